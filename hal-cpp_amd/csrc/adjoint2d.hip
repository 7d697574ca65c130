// adjoint2d.hip -- per-cell-diffusivity ADI, its discrete adjoint and the
// gradient of a field misfit with respect to every cell's diffusivity and to
// the decay rate, with checkpoint-and-recompute.  No reference counterpart:
// the forward model generalises Diffusion2DADI (include/Hal/Pde/ADIequations.h:306-337)
// to a per-cell rate with the face convention of the per-cell explicit
// stencil (src/Pde/PDEequations.cpp:214-219; the face between cells i and i+1
// carries D_i), zero-flux boundaries, and appends MulAll(-k) + Update.
//
// Tick (N = X*Y cells, u in HAL's x*Y+y layout):
//   s1 = (I + L_y/2) u;  s2 = (I - L_x/2)^{-1} s1;  s3 = (I + L_x/2) s2;
//   s4 = (I - L_y/2)^{-1} s3;  u' = s4 - k u.
// Every factor is symmetric, so the adjoint tick applies them transposed in
// reverse order:  m1 = (I - L_y/2)^{-1} lam;  m2 = (I - L_x/2)^{-1} (I + L_x/2) m1;
//   lam' = (I + L_y/2) m2 - k lam,
// and the per-tick gradient contributions are
//   dJ/dD_c -= 1/2 (dy m1 . dy s4 + dx m1 . dx s2 + dx m2 . dx s2 + dy m2 . dy u)
//   dJ/dk   -= <lam, u>
// with dx p = p(x+1,y) - p(x,y) (0 on the last row), dy likewise.
//
// Kernels.  The Thomas factors of a zero-flux line with face rates D are
// fw_j = (D_{j-1}/2) ib_{j-1} and bw_j = (D_j/2) ib_j, with ib = 1/b' the
// reciprocal pivots, so two tables per line direction hold them: ib and
// h = (D/2) ib (fw_j = h_{j-1}, bw_j = h_j), formed once per diffusivity
// field on the host in FP64 and kept in the field's own layout.  Lines along y (contiguous rows) are solved by one workgroup of
// four wavefronts per row: thread-local recurrences composed by wavefront
// scans of affine maps and LDS-exchanged wavefront totals.  Lines along x
// (stride Y) are solved without a transpose: a workgroup stages C adjacent
// columns over all X rows in LDS with row-major loads, one wavefront solves
// each column with E elements per lane, and the result goes back row-major.
// A fused kernel forms the gradient contributions and the adjoint update.
// A forward tick is two launches and an adjoint tick three.  Checkpoints
// every K ticks; the backward pass recomputes each segment from its
// checkpoint, keeping u, s2, s4 of the segment.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

using namespace halhip;

namespace
{
    template <int CTRL, int ROW_MASK = 0xF>
    __device__ __forceinline__ float DppF(float old, float v)
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                      __builtin_bit_cast(int, v), CTRL, ROW_MASK,
                                                                      0xF, false));
    }

    // One step of the wavefront scan of affine maps (A, B) and, with TWO, a
    // second offset B2 under the same multipliers: the element CTRL brings in
    // composes before this one.
    template <int CTRL, int ROW_MASK, bool TWO>
    __device__ __forceinline__ void ScanStepUp(float &A, float &B, float &B2)
    {
        const float tb = DppF<CTRL, ROW_MASK>(0.0f, B);
        if constexpr (TWO)
        {
            const float tb2 = DppF<CTRL, ROW_MASK>(0.0f, B2);
            B2 = fmaf(A, tb2, B2);
        }
        B = fmaf(A, tb, B);
        A *= DppF<CTRL, ROW_MASK>(1.0f, A);
    }

    struct LineArgs
    {
        int32_t nLines;
        int32_t L;
        const float *in;   // [nLines][L]
        const float *rate; // [nLines][L] face rates (face j..j+1 of a line, and line l..l+1)
        const float *ib;   // [nLines][L] reciprocal Thomas pivots of the lines
        const float *h;    // [nLines][L] (D_j / 2) ib_j: bw_j, and fw_{j+1}
        float *out;        // [nLines][L]
        const float *prev; // decay: out = x - k * prev
        float negK;
        float *stash;      // optional copy of x before decay
        const float4 *lineK; // periodic lines: (gamma, alpha, vEnd, 0) per line
        // Dirichlet ghost values, [lo across (L) | hi across (L) | lo along
        // (nLines) | hi along (nLines)]: for rows, bc(-1, y), bc(X, y),
        // bc(x, -1), bc(x, Y).
        const float *edges;
        int32_t wrapAcross; // RHS 0: the neighbour lines wrap
        int32_t dirAcross;  // RHS 0: Dirichlet ghosts beyond lines 0 and nLines - 1
        int32_t dirAlong;   // RHS 0: the implicit Dirichlet source at the line's ends
    };

    // Line solve: a workgroup of kBlockLine threads per line (four
    // wavefronts, EPL adjacent elements per thread, L <= kBlockLine * EPL).
    // With VEC (EPL and L multiples of 4) every array moves as 16-byte
    // vectors.  Each recurrence composes per-thread folds, a wavefront scan
    // of affine maps (DPP) and the wavefront totals exchanged through LDS.
    constexpr int kBlockLine = 256;
    constexpr unsigned kLossBlocks = 1024; // blocks (and partial sums) of the loss seed

    template <int EPL, bool VEC>
    __device__ __forceinline__ void LoadLine(const float *__restrict__ src, size_t base, int j0, int L, float pad,
                                             float (&v)[EPL])
    {
        if constexpr (VEC)
        {
            // L is a multiple of 4: each group of 4 lies wholly inside or
            // wholly past the line
#pragma unroll
            for (int q = 0; q < EPL / 4; q++)
            {
                float4 t = make_float4(pad, pad, pad, pad);
                if (j0 + 4 * q < L)
                {
                    t = *reinterpret_cast<const float4 *>(src + base + j0 + 4 * q);
                }
                v[4 * q] = t.x;
                v[4 * q + 1] = t.y;
                v[4 * q + 2] = t.z;
                v[4 * q + 3] = t.w;
            }
        }
        else
        {
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                v[e] = (j0 + e < L) ? src[base + j0 + e] : pad;
            }
        }
    }

    template <int EPL, bool VEC>
    __device__ __forceinline__ void StoreLine(float *__restrict__ dst, size_t base, int j0, int L, const float (&v)[EPL])
    {
        if constexpr (VEC)
        {
#pragma unroll
            for (int q = 0; q < EPL / 4; q++)
            {
                if (j0 + 4 * q < L)
                {
                    *reinterpret_cast<float4 *>(dst + base + j0 + 4 * q) =
                        make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
                }
            }
        }
        else
        {
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                if (j0 + e < L)
                {
                    dst[base + j0 + e] = v[e];
                }
            }
        }
    }

    // RHS 0: d = c + 1/2 (rate[l] (in[l+1] - c) + rate[l-1] (in[l-1] - c)) across
    //        lines (wrapped, or with Dirichlet ghosts 2 rate[l] (bc - c) beyond
    //        the first and last line), plus rate * bc at a Dirichlet line's ends.
    // RHS 2: d = in.
    // DECAY: out = x - k * prev; STASH: keeps x before the decay.
    // PER: the line is cyclic.  The tables then hold the Thomas factors of
    // the corner-adjusted matrix and the solve runs a second right-hand side
    // (gamma at the first element, alpha at the last) through the same
    // recurrences; the two combine by Sherman-Morrison
    // (include/Hal/Pde/TdmaSolver.h:68-120).
    template <int EPL, int RHS, bool DECAY, bool STASH, bool VEC, bool PER>
    __global__ __launch_bounds__(kBlockLine) void CellLineSolveBlock(LineArgs a)
    {
        __shared__ float sA[kBlockLine / 64];
        __shared__ float sB[kBlockLine / 64];
        __shared__ float sB2[kBlockLine / 64];
        __shared__ float sEnd[4];
        const int t = threadIdx.x;
        const int lane = t & 63;
        const int wave = t >> 6;
        // Lines are dealt so that each XCD (blocks b, b + 8, ...) takes a
        // contiguous range: the neighbour lines the explicit half-step reads
        // are then in the same L2.
        int line = static_cast<int>(blockIdx.x);
        if (a.nLines % 8 == 0)
        {
            line = (line % 8) * (a.nLines / 8) + line / 8;
        }
        const int L = a.L;
        const size_t base = static_cast<size_t>(line) * L;
        const int j0 = t * EPL;
        float d[EPL];
        float ib[EPL];
        LoadLine<EPL, VEC>(a.in, base, j0, L, 0.0f, d);
        LoadLine<EPL, VEC>(a.ib, base, j0, L, 0.0f, ib);
        float h[EPL];
        LoadLine<EPL, VEC>(a.h, base, j0, L, 0.0f, h);
        if constexpr (RHS == 0)
        {
            // explicit half-step across lines: neighbours in lines l +- 1
            float nb[EPL];
            float s[EPL];
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                s[e] = 0.0f;
            }
            const bool hasNext = line + 1 < a.nLines;
            const bool hasPrev = line > 0;
            float dl[EPL]; // this line's face rates (to line l + 1): every cell's own D
            LoadLine<EPL, VEC>(a.rate, base, j0, L, 0.0f, dl);
            if (hasNext || a.wrapAcross)
            {
                const size_t nbase = hasNext ? base + L : 0; // wrapped: line 0
                LoadLine<EPL, VEC>(a.in, nbase, j0, L, 0.0f, nb);
#pragma unroll
                for (int e = 0; e < EPL; e++)
                {
                    s[e] += dl[e] * (nb[e] - d[e]);
                }
            }
            if (hasPrev || a.wrapAcross)
            {
                const size_t pbase = hasPrev ? base - L : static_cast<size_t>(a.nLines - 1) * L;
                float r[EPL];
                LoadLine<EPL, VEC>(a.rate, pbase, j0, L, 0.0f, r);
                LoadLine<EPL, VEC>(a.in, pbase, j0, L, 0.0f, nb);
#pragma unroll
                for (int e = 0; e < EPL; e++)
                {
                    s[e] += r[e] * (nb[e] - d[e]);
                }
            }
            if (a.dirAcross)
            {
#pragma unroll
                for (int e = 0; e < EPL; e++)
                {
                    const int j = j0 + e;
                    if (j < L)
                    {
                        if (line == 0)
                        {
                            s[e] += 2.0f * dl[e] * (a.edges[j] - d[e]);
                        }
                        if (line == a.nLines - 1)
                        {
                            s[e] += 2.0f * dl[e] * (a.edges[L + j] - d[e]);
                        }
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                d[e] = (j0 + e < L) ? fmaf(0.5f, s[e], d[e]) : 0.0f;
            }
            if (a.dirAlong)
            {
                const float lo = a.edges[2 * L + line];
                const float hi = a.edges[2 * L + a.nLines + line];
#pragma unroll
                for (int e = 0; e < EPL; e++)
                {
                    if (j0 + e == 0)
                    {
                        d[e] = fmaf(dl[e], lo, d[e]);
                    }
                    if (j0 + e == L - 1)
                    {
                        d[e] = fmaf(dl[e], hi, d[e]);
                    }
                }
            }
        }
        // PER: the second right-hand side and its solution
        float z[PER ? EPL : 1];
        float4 K = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if constexpr (PER)
        {
            K = a.lineK[line];
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                z[e] = (j0 + e == 0) ? K.x : ((j0 + e == L - 1) ? K.y : 0.0f);
            }
        }
        // bw_j = h_j (0 at the line's end), fw_j = h_{j-1} (0 at its start;
        // the first element's comes from the previous thread)
        const float hLeft = (j0 > 0 && j0 - 1 < L) ? a.h[base + j0 - 1] : 0.0f;
        float fw[EPL];
#pragma unroll
        for (int e = 0; e < EPL; e++)
        {
            fw[e] = (e == 0) ? hLeft : h[e - 1];
        }

        // Forward recurrence d'_j = fw_j d'_{j-1} + d_j, carry into the line 0.
        float A = 1.0f;
        float B = 0.0f;
        float B2 = 0.0f;
#pragma unroll
        for (int e = 0; e < EPL; e++)
        {
            B = fmaf(fw[e], B, d[e]);
            if constexpr (PER)
            {
                B2 = fmaf(fw[e], B2, z[e]);
            }
            A *= fw[e];
        }
        ScanStepUp<0x111, 0xF, PER>(A, B, B2);
        ScanStepUp<0x112, 0xF, PER>(A, B, B2);
        ScanStepUp<0x114, 0xF, PER>(A, B, B2);
        ScanStepUp<0x118, 0xF, PER>(A, B, B2);
        ScanStepUp<0x142, 0xA, PER>(A, B, B2);
        ScanStepUp<0x143, 0xC, PER>(A, B, B2);
        if (lane == 63)
        {
            sA[wave] = A;
            sB[wave] = B;
            sB2[wave] = B2;
        }
        __syncthreads();
        float cw = 0.0f;
        float cw2 = 0.0f;
        for (int g = 0; g < wave; g++)
        {
            cw = fmaf(sA[g], cw, sB[g]);
            if constexpr (PER)
            {
                cw2 = fmaf(sA[g], cw2, sB2[g]);
            }
        }
        // this lane's carry: the preceding lanes' composition applied to cw
        const float Aprev = DppF<0x138>(1.0f, A);
        float acc = fmaf(Aprev, cw, DppF<0x138>(0.0f, B));
        float acc2 = 0.0f;
        if constexpr (PER)
        {
            acc2 = fmaf(Aprev, cw2, DppF<0x138>(0.0f, B2));
        }
#pragma unroll
        for (int e = 0; e < EPL; e++)
        {
            acc = fmaf(fw[e], acc, d[e]);
            d[e] = acc;
            if constexpr (PER)
            {
                acc2 = fmaf(fw[e], acc2, z[e]);
                z[e] = acc2;
            }
        }

        // Backward recurrence x_j = bw_j x_{j+1} + ib_j d'_j from the right.
        float bw[EPL];
#pragma unroll
        for (int e = 0; e < EPL; e++)
        {
            bw[e] = (j0 + e < L - 1) ? h[e] : 0.0f;
        }
        A = 1.0f;
        B = 0.0f;
        B2 = 0.0f;
#pragma unroll
        for (int e = EPL - 1; e >= 0; e--)
        {
            B = fmaf(bw[e], B, ib[e] * d[e]);
            if constexpr (PER)
            {
                B2 = fmaf(bw[e], B2, ib[e] * z[e]);
            }
            A *= bw[e];
        }
#pragma unroll
        for (int k = 0; k < 6; k++)
        {
            const int o = 1 << k;
            const float Bd = __shfl_down(B, o);
            const float Ad = __shfl_down(A, o);
            float Bd2 = 0.0f;
            if constexpr (PER)
            {
                Bd2 = __shfl_down(B2, o);
            }
            if (lane + o < 64)
            {
                B = fmaf(A, Bd, B);
                if constexpr (PER)
                {
                    B2 = fmaf(A, Bd2, B2);
                }
                A *= Ad;
            }
        }
        __syncthreads(); // the forward totals are read
        if (lane == 0)
        {
            sA[wave] = A;
            sB[wave] = B;
            sB2[wave] = B2;
        }
        __syncthreads();
        cw = 0.0f;
        cw2 = 0.0f;
        for (int g = kBlockLine / 64 - 1; g > wave; g--)
        {
            cw = fmaf(sA[g], cw, sB[g]);
            if constexpr (PER)
            {
                cw2 = fmaf(sA[g], cw2, sB2[g]);
            }
        }
        const float An = __shfl_down(A, 1);
        const float Bn = __shfl_down(B, 1);
        acc = (lane == 63) ? cw : fmaf(An, cw, Bn);
        if constexpr (PER)
        {
            const float Bn2 = __shfl_down(B2, 1);
            acc2 = (lane == 63) ? cw2 : fmaf(An, cw2, Bn2);
        }
#pragma unroll
        for (int e = EPL - 1; e >= 0; e--)
        {
            acc = fmaf(bw[e], acc, ib[e] * d[e]);
            d[e] = acc;
            if constexpr (PER)
            {
                acc2 = fmaf(bw[e], acc2, ib[e] * z[e]);
                z[e] = acc2;
            }
        }
        if constexpr (PER)
        {
            // x = y - f z with f = (y_0 + vEnd y_{L-1}) / (1 + z_0 + vEnd z_{L-1})
            const int tLast = (L - 1) / EPL;
            const int eLast = (L - 1) - tLast * EPL;
            if (t == 0)
            {
                sEnd[0] = d[0];
                sEnd[1] = z[0];
            }
            if (t == tLast)
            {
                float yl = 0.0f;
                float zl = 0.0f;
#pragma unroll
                for (int e = 0; e < EPL; e++)
                {
                    if (e == eLast)
                    {
                        yl = d[e];
                        zl = z[e];
                    }
                }
                sEnd[2] = yl;
                sEnd[3] = zl;
            }
            __syncthreads();
            const float f = fmaf(K.z, sEnd[2], sEnd[0]) / (1.0f + fmaf(K.z, sEnd[3], sEnd[1]));
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                d[e] = fmaf(-f, z[e], d[e]);
            }
        }
        if constexpr (STASH)
        {
            StoreLine<EPL, VEC>(a.stash, base, j0, L, d);
        }
        if constexpr (DECAY)
        {
            float pv[EPL];
            LoadLine<EPL, VEC>(a.prev, base, j0, L, 0.0f, pv);
#pragma unroll
            for (int e = 0; e < EPL; e++)
            {
                d[e] = fmaf(a.negK, pv[e], d[e]);
            }
        }
        StoreLine<EPL, VEC>(a.out, base, j0, L, d);
    }

    struct ColArgs
    {
        int32_t X;
        int32_t Y;
        int32_t blocks;    // column blocks, ceil(Y / C)
        const float *in;   // [X][Y]
        const float *D;    // [X][Y] face rates
        const float *Dt;   // [Y][X] the same, column-major
        const float *ib;   // [Y][X] reciprocal pivots of the x-lines, column-major
        const float *h;    // [Y][X] (D_x / 2) ib_x of the x-lines, column-major
        float *out;        // [X][Y]
        const float4 *lineK; // periodic x-lines: (gamma, alpha, vEnd, 0) per column
        const float *edges;  // Dirichlet ghosts [xlo(Y) | xhi(Y) | ylo(X) | yhi(X)]
        int32_t wrapX;       // MODE 1: the column wraps
        int32_t wrapY;       // MODE 0: the neighbour columns wrap
        int32_t dirX;        // Dirichlet x ends (MODE 0: implicit source; MODE 1: ghost faces at 0)
        int32_t dirY;        // MODE 0: Dirichlet ghosts beyond columns 0 and Y - 1
    };

    // C consecutive floats of row x from column y0 (those past Y read 0).
    template <int C, bool VEC>
    __device__ __forceinline__ void LoadRow(const float *__restrict__ src, size_t at, int y0, int Y, float (&v)[C])
    {
        if constexpr (VEC && C == 4)
        {
            const float4 t = *reinterpret_cast<const float4 *>(src + at);
            v[0] = t.x;
            v[1] = t.y;
            v[2] = t.z;
            v[3] = t.w;
        }
        else if constexpr (VEC && C == 2)
        {
            const float2 t = *reinterpret_cast<const float2 *>(src + at);
            v[0] = t.x;
            v[1] = t.y;
        }
        else
        {
#pragma unroll
            for (int c = 0; c < C; c++)
            {
                v[c] = (y0 + c < Y) ? src[at + c] : 0.0f;
            }
        }
    }

    // Lines along x (one per column, stride Y) solved without a transpose.  A
    // workgroup holds C adjacent columns of the field over all X rows in LDS:
    // rows are loaded and stored row-major (C consecutive floats each).  The
    // line tables (ib, h and, for MODE 1, D) are fixed for a run and kept
    // column-major, so each lane reads its own rows of them directly, as
    // contiguous vectors: staging them too would pull every 128-byte line of
    // them into several CUs for 16 bytes each.  Wavefront c solves column c
    // with lane l holding rows [lE, lE + E) (LDS pitch E + 1 per lane, so the
    // lanes' element reads hit distinct banks): per-lane folds and a
    // wavefront scan of affine maps.  (Sharing a column among four
    // wavefronts measured no faster and moved the heavily cancelling decay
    // gradient of the 1024^2 workload by 2e-4.)  Rows past X are identity
    // padding (d = D = ib = 0).
    //   MODE 0: d = (I + L_y/2) in, the explicit half-step across columns
    //           (reads columns y0 - 1 and y0 + C too; wrapped, or with
    //           Dirichlet ghosts beyond the first and last column), plus
    //           rate * bc at a Dirichlet column's ends.
    //   MODE 1: d = (I + A_x/2) in, the linear explicit half-step along the
    //           line (ghost values 0).
    // PER: the column is cyclic (second right-hand side and Sherman-Morrison,
    // as in CellLineSolveBlock).
    // Column blocks are dealt so that each XCD (blocks b, b + 8, ...) takes
    // a contiguous range: blocks that share cache lines share an L2.
    template <int C, int E, int MODE, bool VEC, bool PER>
    __global__ __launch_bounds__(64 * C) void CellColumnSolve(ColArgs a)
    {
        constexpr int kThreads = 64 * C;
        constexpr int kPitch = 64 * (E + 1); // LDS floats per column
        constexpr int kRows = 64 * E;
        __shared__ float sD[C * kPitch];
        const int X = a.X;
        const int Y = a.Y;
        int b = static_cast<int>(blockIdx.x);
        if (a.blocks % 8 == 0)
        {
            b = (b % 8) * (a.blocks / 8) + b / 8;
        }
        const int y0 = b * C;
        const int t = threadIdx.x;
        const int lane = t & 63;
        const int c = t >> 6; // the column this wavefront solves
        const int x0 = lane * E;
        const int col = c * kPitch;
        const int me = lane * (E + 1);
        // This lane's rows of the column's tables (a column past Y reads the
        // last one; its results are not stored), requested first: they do
        // not wait for the staging.
        const size_t colBase = static_cast<size_t>(min(y0 + c, Y - 1)) * X;
        const bool vecT = (E % 4 == 0) && (X % 4 == 0);
        float h[E];
        float ib[E];
        float rr[MODE == 1 ? E : 1];
        if (vecT)
        {
            LoadLine<E, (E % 4 == 0)>(a.ib, colBase, x0, X, 0.0f, ib);
            LoadLine<E, (E % 4 == 0)>(a.h, colBase, x0, X, 0.0f, h);
            if constexpr (MODE == 1)
            {
                LoadLine<E, (E % 4 == 0)>(a.Dt, colBase, x0, X, 0.0f, rr);
            }
        }
        else
        {
            LoadLine<E, false>(a.ib, colBase, x0, X, 0.0f, ib);
            LoadLine<E, false>(a.h, colBase, x0, X, 0.0f, h);
            if constexpr (MODE == 1)
            {
                LoadLine<E, false>(a.Dt, colBase, x0, X, 0.0f, rr);
            }
        }
        const float hLeft = (x0 > 0 && x0 - 1 < X) ? a.h[colBase + x0 - 1] : 0.0f;
        // Every thread holds kMine rows: all their loads are issued before
        // the first LDS store, so one memory latency covers the tile.
        constexpr int kMine = (kRows + kThreads - 1) / kThreads;
        float u[kMine][C];
        float r[kMine][C];
        float uL[kMine];
        float rL[kMine];
        float uR[kMine];
#pragma unroll
        for (int i = 0; i < kMine; i++)
        {
            const int x = t + i * kThreads;
            const size_t row = static_cast<size_t>(x) * Y + y0;
            uL[i] = 0.0f;
            rL[i] = 0.0f;
            uR[i] = 0.0f;
            if (x < X)
            {
                LoadRow<C, VEC>(a.in, row, y0, Y, u[i]);
                if constexpr (MODE == 0)
                {
                    LoadRow<C, VEC>(a.D, row, y0, Y, r[i]);
                    // wrapped: column Y - 1 before column 0, column 0 after Y - 1
                    const size_t rowStart = static_cast<size_t>(x) * Y;
                    uL[i] = (y0 > 0) ? a.in[row - 1] : (a.wrapY ? a.in[rowStart + Y - 1] : 0.0f);
                    rL[i] = (y0 > 0) ? a.D[row - 1] : (a.wrapY ? a.D[rowStart + Y - 1] : 0.0f);
                    uR[i] = (y0 + C < Y) ? a.in[row + C] : (a.wrapY ? a.in[rowStart] : 0.0f);
                }
            }
            else
            {
#pragma unroll
                for (int c = 0; c < C; c++)
                {
                    u[i][c] = 0.0f;
                    r[i][c] = 0.0f;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kMine; i++)
        {
            const int x = t + i * kThreads;
            const int at = x + x / E;
            if constexpr (MODE == 0)
            {
                float dv[C];
#pragma unroll
                for (int c = 0; c < C; c++)
                {
                    const int y = y0 + c;
                    // past column Y - 1 (wrapped) the neighbour is uR, column 0
                    const float up = (c + 1 < C && y + 1 < Y) ? u[i][c + 1] : uR[i];
                    const float um = (c > 0) ? u[i][c - 1] : uL[i];
                    const float rm = (c > 0) ? r[i][c - 1] : rL[i];
                    float sum = 0.0f;
                    sum += (y + 1 < Y || a.wrapY) ? r[i][c] * (up - u[i][c]) : 0.0f;
                    sum += (y > 0 || a.wrapY) ? rm * (um - u[i][c]) : 0.0f;
                    if (a.dirY && x < X && y < Y)
                    {
                        if (y == 0)
                        {
                            sum += 2.0f * r[i][c] * (a.edges[2 * Y + x] - u[i][c]);
                        }
                        if (y == Y - 1)
                        {
                            sum += 2.0f * r[i][c] * (a.edges[2 * Y + X + x] - u[i][c]);
                        }
                    }
                    dv[c] = (y < Y && x < X) ? fmaf(0.5f, sum, u[i][c]) : 0.0f;
                    if (a.dirX && y < Y)
                    {
                        if (x == 0)
                        {
                            dv[c] = fmaf(r[i][c], a.edges[y], dv[c]);
                        }
                        if (x == X - 1)
                        {
                            dv[c] = fmaf(r[i][c], a.edges[Y + y], dv[c]);
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < C; c++)
                {
                    u[i][c] = dv[c];
                }
            }
            if (x < kRows)
            {
#pragma unroll
                for (int c = 0; c < C; c++)
                {
                    sD[c * kPitch + at] = u[i][c];
                }
            }
        }
        __syncthreads();

        float d[E];
#pragma unroll
        for (int e = 0; e < E; e++)
        {
            d[e] = sD[col + me + e];
        }
        if constexpr (MODE == 1)
        {
            // (I + L_x/2) along the column: neighbours x +- 1 from LDS
            // (rows past X have D = 0, so no face reaches them)
            const int xl = (x0 > 0) ? x0 - 1 : X - 1; // wrapped: row X - 1 before row 0
            const int xr = x0 + E;
            const bool left = (x0 > 0 && x0 - 1 < X) || (x0 == 0 && a.wrapX);
            const float dL = left ? sD[col + xl + xl / E] : 0.0f;
            const float rLv = left ? a.Dt[colBase + xl] : 0.0f;
            const float dR = (xr < kRows) ? sD[col + xr + xr / E] : 0.0f;
            const float dW = sD[col]; // row 0, after row X - 1 when wrapped
            float out[E];
#pragma unroll
            for (int e = 0; e < E; e++)
            {
                const int x = x0 + e;
                const float cv = d[e];
                const float dp = (x + 1 < X) ? ((e + 1 < E) ? d[e + 1] : dR) : dW;
                const float dm = (e > 0) ? d[e - 1] : dL;
                const float rm = (e > 0) ? rr[e - 1] : rLv;
                float sum = 0.0f;
                sum += (x + 1 < X || a.wrapX) ? rr[e] * (dp - cv) : 0.0f;
                sum += (x > 0 || a.wrapX) ? rm * (dm - cv) : 0.0f;
                if (a.dirX)
                {
                    // ghost faces 2 D (0 - c) at both ends
                    sum -= (x == 0) ? 2.0f * rr[e] * cv : 0.0f;
                    sum -= (x == X - 1) ? 2.0f * rr[e] * cv : 0.0f;
                }
                out[e] = (x < X) ? fmaf(0.5f, sum, cv) : 0.0f;
            }
#pragma unroll
            for (int e = 0; e < E; e++)
            {
                d[e] = out[e];
            }
        }

        // PER: the second right-hand side (gamma at row 0, alpha at X - 1)
        float z[PER ? E : 1];
        float4 K = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if constexpr (PER)
        {
            K = a.lineK[min(y0 + c, Y - 1)];
#pragma unroll
            for (int e = 0; e < E; e++)
            {
                z[e] = (x0 + e == 0) ? K.x : ((x0 + e == X - 1) ? K.y : 0.0f);
            }
        }

        // forward d'_x = fw_x d'_{x-1} + d_x with fw_x = h_{x-1}
        float A = 1.0f;
        float B = 0.0f;
        float B2 = 0.0f;
#pragma unroll
        for (int e = 0; e < E; e++)
        {
            const float fw = (e == 0) ? hLeft : h[e - 1];
            B = fmaf(fw, B, d[e]);
            if constexpr (PER)
            {
                B2 = fmaf(fw, B2, z[e]);
            }
            A *= fw;
        }
        ScanStepUp<0x111, 0xF, PER>(A, B, B2);
        ScanStepUp<0x112, 0xF, PER>(A, B, B2);
        ScanStepUp<0x114, 0xF, PER>(A, B, B2);
        ScanStepUp<0x118, 0xF, PER>(A, B, B2);
        ScanStepUp<0x142, 0xA, PER>(A, B, B2);
        ScanStepUp<0x143, 0xC, PER>(A, B, B2);
        // carry into this lane: the preceding lanes' inclusive value
        float acc = DppF<0x138>(0.0f, B);
        float acc2 = 0.0f;
        if constexpr (PER)
        {
            acc2 = DppF<0x138>(0.0f, B2);
        }
#pragma unroll
        for (int e = 0; e < E; e++)
        {
            const float fw = (e == 0) ? hLeft : h[e - 1];
            acc = fmaf(fw, acc, d[e]);
            d[e] = acc;
            if constexpr (PER)
            {
                acc2 = fmaf(fw, acc2, z[e]);
                z[e] = acc2;
            }
        }

        // backward x_x = bw_x x_{x+1} + ib_x d'_x with bw_x = h_x (0 at X - 1)
        A = 1.0f;
        B = 0.0f;
        B2 = 0.0f;
#pragma unroll
        for (int e = E - 1; e >= 0; e--)
        {
            const float bw = (x0 + e < X - 1) ? h[e] : 0.0f;
            B = fmaf(bw, B, ib[e] * d[e]);
            if constexpr (PER)
            {
                B2 = fmaf(bw, B2, ib[e] * z[e]);
            }
            A *= bw;
        }
#pragma unroll
        for (int kk = 0; kk < 6; kk++)
        {
            const int o = 1 << kk;
            const float Bd = __shfl_down(B, o);
            const float Ad = __shfl_down(A, o);
            float Bd2 = 0.0f;
            if constexpr (PER)
            {
                Bd2 = __shfl_down(B2, o);
            }
            if (lane + o < 64)
            {
                B = fmaf(A, Bd, B);
                if constexpr (PER)
                {
                    B2 = fmaf(A, Bd2, B2);
                }
                A *= Ad;
            }
        }
        const float Bn = __shfl_down(B, 1);
        acc = (lane == 63) ? 0.0f : Bn;
        if constexpr (PER)
        {
            const float Bn2 = __shfl_down(B2, 1);
            acc2 = (lane == 63) ? 0.0f : Bn2;
        }
#pragma unroll
        for (int e = E - 1; e >= 0; e--)
        {
            const float bw = (x0 + e < X - 1) ? h[e] : 0.0f;
            acc = fmaf(bw, acc, ib[e] * d[e]);
            d[e] = acc;
            if constexpr (PER)
            {
                acc2 = fmaf(bw, acc2, ib[e] * z[e]);
                z[e] = acc2;
            }
        }
        if constexpr (PER)
        {
            // x = y - f z with f = (y_0 + vEnd y_{X-1}) / (1 + z_0 + vEnd z_{X-1}):
            // row 0 is lane 0's first element, row X - 1 lane (X-1)/E's
            const int lLast = (X - 1) / E;
            const int eLast = (X - 1) - lLast * E;
            float yl = 0.0f;
            float zl = 0.0f;
#pragma unroll
            for (int e = 0; e < E; e++)
            {
                if (e == eLast)
                {
                    yl = d[e];
                    zl = z[e];
                }
            }
            const float y0v = __shfl(d[0], 0);
            const float z0v = __shfl(z[0], 0);
            yl = __shfl(yl, lLast);
            zl = __shfl(zl, lLast);
            const float f = fmaf(K.z, yl, y0v) / (1.0f + fmaf(K.z, zl, z0v));
#pragma unroll
            for (int e = 0; e < E; e++)
            {
                d[e] = fmaf(-f, z[e], d[e]);
            }
        }
#pragma unroll
        for (int e = 0; e < E; e++)
        {
            sD[col + me + e] = d[e];
        }
        __syncthreads();
        for (int x = t; x < X; x += kThreads)
        {
            const int at = x + x / E;
            const size_t row = static_cast<size_t>(x) * Y + y0;
            if constexpr (VEC && C == 4)
            {
                *reinterpret_cast<float4 *>(a.out + row) =
                    make_float4(sD[at], sD[kPitch + at], sD[2 * kPitch + at], sD[3 * kPitch + at]);
            }
            else if constexpr (VEC && C == 2)
            {
                *reinterpret_cast<float2 *>(a.out + row) = make_float2(sD[at], sD[kPitch + at]);
            }
            else
            {
#pragma unroll
                for (int cc = 0; cc < C; cc++)
                {
                    if (y0 + cc < Y)
                    {
                        a.out[row + cc] = sD[cc * kPitch + at];
                    }
                }
            }
        }
    }

    struct GradArgs
    {
        int32_t X;
        int32_t Y;
        const float *D;   // [X][Y]
        const float *m1;  // [X][Y]
        const float *m2;  // [X][Y]
        const float *s2;  // [X][Y]
        const float *s4;  // [X][Y]
        const float *u;   // [X][Y]
        const float *lam; // [X][Y]
        float *lamNext;   // [X][Y]
        double *gradD;    // [X][Y], accumulated over the ticks in FP64
        double *partialK; // one per block
        float negK;
        const float *edges; // Dirichlet ghosts [xlo(Y) | xhi(Y) | ylo(X) | yhi(X)]
        int32_t wrapX;
        int32_t wrapY;
        int32_t dirX;
        int32_t dirY;
    };

    // Per-tick gradient contributions and lam' = (I + A_y/2) m2 - k lam.  The
    // decay gradient -<lam, u> cancels heavily (lam has near-zero mean), so
    // its products and sums are formed in FP64.  A cell's interior or wrap
    // face adds -1/2 dq dp per (q, p) pair; each of its Dirichlet faces adds
    // q (bc - p).
    __global__ __launch_bounds__(256) void CellGradient(GradArgs a)
    {
        __shared__ double red[4];
        const int y = blockIdx.x * 64 + (threadIdx.x & 63);
        const int x = blockIdx.y * 4 + (threadIdx.x >> 6);
        const int X = a.X;
        const int Y = a.Y;
        double kPart = 0.0;
        if (x < X && y < Y)
        {
            const size_t i = static_cast<size_t>(x) * Y + y;
            const bool hx = x + 1 < X;
            const bool hy = y + 1 < Y;
            // the cell's own face leads to the next cell, or wraps to the first
            const size_t ix = hx ? i + Y : (a.wrapX ? i - static_cast<size_t>(X - 1) * Y : i);
            const size_t iy = hy ? i + 1 : (a.wrapY ? i - static_cast<size_t>(Y - 1) : i);
            const float dyM1 = a.m1[iy] - a.m1[i];
            const float dyS4 = a.s4[iy] - a.s4[i];
            const float dxM1 = a.m1[ix] - a.m1[i];
            const float dxS2 = a.s2[ix] - a.s2[i];
            const float dxM2 = a.m2[ix] - a.m2[i];
            const float dyM2 = a.m2[iy] - a.m2[i];
            const float dyU = a.u[iy] - a.u[i];
            const float g = dyM1 * dyS4 + (dxM1 + dxM2) * dxS2 + dyM2 * dyU;
            const float c = a.m2[i];
            float gb = 0.0f;
            if (a.dirY && (y == 0 || y == Y - 1))
            {
                const float m1 = a.m1[i];
                const float s4 = a.s4[i];
                const float u = a.u[i];
                if (y == 0)
                {
                    const float bc = a.edges[2 * Y + x];
                    gb += m1 * (bc - s4) + c * (bc - u);
                }
                if (y == Y - 1)
                {
                    const float bc = a.edges[2 * Y + X + x];
                    gb += m1 * (bc - s4) + c * (bc - u);
                }
            }
            if (a.dirX && (x == 0 || x == X - 1))
            {
                const float m12 = a.m1[i] + c;
                const float s2 = a.s2[i];
                if (x == 0)
                {
                    gb += m12 * (a.edges[y] - s2);
                }
                if (x == X - 1)
                {
                    gb += m12 * (a.edges[Y + y] - s2);
                }
            }
            a.gradD[i] += static_cast<double>(gb) - 0.5 * static_cast<double>(g);
            const float lam = a.lam[i];
            kPart = -static_cast<double>(lam) * static_cast<double>(a.u[i]);
            // (I + A_y/2) m2 along the row
            float s = 0.0f;
            if (hy || a.wrapY)
            {
                s += a.D[i] * (a.m2[iy] - c);
            }
            if (y > 0)
            {
                s += a.D[i - 1] * (a.m2[i - 1] - c);
            }
            else if (a.wrapY)
            {
                s += a.D[i + Y - 1] * (a.m2[i + Y - 1] - c);
            }
            if (a.dirY)
            {
                s -= (y == 0) ? 2.0f * a.D[i] * c : 0.0f;
                s -= (y == Y - 1) ? 2.0f * a.D[i] * c : 0.0f;
            }
            a.lamNext[i] = fmaf(a.negK, lam, fmaf(0.5f, s, c));
        }
        // Block sum of the decay-gradient part (deterministic: fixed tree).
        for (int o = 32; o > 0; o >>= 1)
        {
            kPart += __shfl_down(kPart, o);
        }
        if ((threadIdx.x & 63) == 0)
        {
            red[threadIdx.x >> 6] = kPart;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            a.partialK[blockIdx.y * gridDim.x + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
        }
    }

    // loss partials: 1/2 (u - target)^2 per block, and lam = u - target.
    __global__ __launch_bounds__(256) void LossSeed(const float *__restrict__ u, const float *__restrict__ target,
                                                    float *__restrict__ lam, int64_t n, double *__restrict__ partial)
    {
        __shared__ double red[4];
        double s = 0.0;
        for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        {
            const float r = u[i] - target[i];
            lam[i] = r;
            s += 0.5 * static_cast<double>(r) * static_cast<double>(r);
        }
        for (int o = 32; o > 0; o >>= 1)
        {
            s += __shfl_down(s, o);
        }
        if ((threadIdx.x & 63) == 0)
        {
            red[threadIdx.x >> 6] = s;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
        }
    }

    // dst[k] = float(src[k]).
    __global__ __launch_bounds__(256) void ToFloat(const double *__restrict__ src, float *__restrict__ dst, int64_t n)
    {
        for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        {
            dst[i] = static_cast<float>(src[i]);
        }
    }

    // dst[0, n) = 0 (a kernel, not a memset node, so graph replays clear it).
    __global__ __launch_bounds__(256) void FillZero(float *__restrict__ dst, int64_t n)
    {
        for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        {
            dst[i] = 0.0f;
        }
    }

    // Reciprocal Thomas pivots ib = 1/b' of the lines with face rates D:
    // b_j = 1 + (D_{j-1} + D_j)/2 (absent faces dropped), b'_0 = b_0,
    // b'_j = b_j - (D_{j-1}/2)^2 / b'_{j-1}, and h = (D/2) ib.  A Dirichlet
    // line adds D_j to each end's diagonal (its ghost face 2 D_j, halved).  A
    // periodic line (L > 1) adds the wrap face D_{L-1} to both ends and takes
    // the corner-adjusted diagonal of include/Hal/Pde/TdmaSolver.h:68-120
    // (gamma = -b_0, alpha = beta = -D_{L-1}/2: b_0 - gamma, b_{L-1} -
    // alpha beta / gamma); lineK gets (gamma, alpha, beta / gamma, 0) per line.
    // Lines along x (alongX: one per column, element j at j*Y + y) or along y
    // (one per row); ib and h are written in the [X][Y] layout either way.
    void LinePivots(int32_t X, int32_t Y, const std::vector<double> &D, bool alongX, bool wrap, bool dirichlet,
                    std::vector<float> &ib, std::vector<float> &h, std::vector<float> &lineK)
    {
        ib.assign(static_cast<size_t>(X) * Y, 0.0f);
        h.assign(static_cast<size_t>(X) * Y, 0.0f);
        const int32_t nLines = alongX ? Y : X;
        const int32_t L = alongX ? X : Y;
        const bool per = wrap && L > 1;
        lineK.assign(4 * static_cast<size_t>(nLines), 0.0f);
        std::vector<double> b(static_cast<size_t>(L));
        for (int32_t l = 0; l < nLines; l++)
        {
            auto at = [&](int32_t j) -> size_t
            {
                return alongX ? static_cast<size_t>(j) * Y + l : static_cast<size_t>(l) * Y + j;
            };
            for (int32_t j = 0; j < L; j++)
            {
                const double dm = (j > 0) ? D[at(j - 1)] : 0.0;
                const double dp = (j < L - 1) ? D[at(j)] : 0.0;
                b[j] = 1.0 + 0.5 * (dm + dp);
            }
            if (per)
            {
                const double dw = D[at(L - 1)];
                b[0] += 0.5 * dw;
                b[L - 1] += 0.5 * dw;
                const double gamma = -b[0];
                const double alpha = -0.5 * dw;
                b[0] -= gamma;
                b[L - 1] -= alpha * alpha / gamma;
                float *k = lineK.data() + 4 * static_cast<size_t>(l);
                k[0] = static_cast<float>(gamma);
                k[1] = static_cast<float>(alpha);
                k[2] = static_cast<float>(alpha / gamma);
            }
            else if (dirichlet)
            {
                b[0] += D[at(0)];
                b[L - 1] += D[at(L - 1)];
            }
            double bpPrev = 0.0;
            for (int32_t j = 0; j < L; j++)
            {
                double bp = b[j];
                if (j > 0)
                {
                    const double dm = 0.5 * D[at(j - 1)];
                    bp -= dm * dm / bpPrev;
                }
                ib[at(j)] = static_cast<float>(1.0 / bp);
                h[at(j)] = static_cast<float>((j < L - 1) ? 0.5 * D[at(j)] / bp : 0.0);
                bpPrev = bp;
            }
        }
    }

}

struct halhip_adjoint2d
{
    halhip_context *ctx = nullptr;
    int32_t X = 0;
    int32_t Y = 0;
    int32_t ticks = 0;
    int32_t every = 1;
    float decay = 0.0f;
    bool ready = false;
    DeviceArray<float> D;                  // face rates, [X][Y]
    DeviceArray<float> Dt;                 // face rates, column-major [Y][X]
    DeviceArray<float> ibXt, hXt;          // x-line pivots and (D/2) ib, column-major [Y][X]
    DeviceArray<float> ibY, hY;            // y-line pivots and (D/2) ib, [X][Y]
    // Boundaries: wrap flags, Dirichlet ghosts on the non-wrapped axes
    // (edges [xlo(Y) | xhi(Y) | ylo(X) | yhi(X)]), and the periodic lines'
    // Sherman-Morrison constants (lineKX per column, lineKY per row).
    bool wrapX = false;
    bool wrapY = false;
    bool dirX = false;
    bool dirY = false;
    DeviceArray<float> edges;
    DeviceArray<float> lineKX, lineKY;
    DeviceArray<float> u, s2, work;        // field work buffers
    DeviceArray<float> ckpt;               // (ticks/every + 1) checkpoints
    DeviceArray<float> stashU, stashS2, stashS4; // one segment
    DeviceArray<float> lam, lamNext, m1, m2;
    DeviceArray<double> grad;              // dJ/dD, accumulated in FP64
    DeviceArray<double> partialK;
    DeviceArray<double> partialJ;
    StreamClock clock;
    Profiler prof;
    GraphCache graphs; // the forward and gradient passes, per problem
};

namespace
{
    size_t Cells(const halhip_adjoint2d *p)
    {
        return static_cast<size_t>(p->X) * p->Y;
    }

    void RequireAdj(const halhip_adjoint2d *p, const char *caller)
    {
        Require(p != nullptr, std::string(caller) + ": null plan");
        Bind(p->ctx);
    }

    // Periodic lines along y (rows) / along x (columns): a wrapped axis of
    // more than one cell.
    bool PeriodicRows(const halhip_adjoint2d *p)
    {
        return p->wrapY && p->Y > 1;
    }

    bool PeriodicColumns(const halhip_adjoint2d *p)
    {
        return p->wrapX && p->X > 1;
    }

    template <int EPL, int RHS, bool DECAY, bool STASH, bool PER>
    void LaunchBlockLinePer(halhip_adjoint2d *p, const LineArgs &a)
    {
        const bool vec = (EPL % 4 == 0) && (a.L % 4 == 0);
        if (vec)
        {
            hipLaunchKernelGGL((CellLineSolveBlock<EPL, RHS, DECAY, STASH, (EPL % 4 == 0), PER>),
                               dim3(static_cast<unsigned>(a.nLines)), dim3(kBlockLine), 0, p->clock.stream, a);
        }
        else
        {
            hipLaunchKernelGGL((CellLineSolveBlock<EPL, RHS, DECAY, STASH, false, PER>),
                               dim3(static_cast<unsigned>(a.nLines)), dim3(kBlockLine), 0, p->clock.stream, a);
        }
    }

    template <int EPL, int RHS, bool DECAY, bool STASH>
    void LaunchBlockLine(halhip_adjoint2d *p, const LineArgs &a)
    {
        if (PeriodicRows(p))
        {
            LaunchBlockLinePer<EPL, RHS, DECAY, STASH, true>(p, a);
        }
        else
        {
            LaunchBlockLinePer<EPL, RHS, DECAY, STASH, false>(p, a);
        }
    }

    // One line solve per line of the family: a workgroup of four
    // wavefronts per line, EPL = the smallest power of two with
    // 256 * EPL >= L (sides are at most 4096).
    template <int RHS, bool DECAY, bool STASH>
    void LaunchLine(halhip_adjoint2d *p, const LineArgs &a, const char *name)
    {
        p->prof.Begin(p->clock.stream, name);
        int epl = 1;
        while (kBlockLine * epl < a.L)
        {
            epl *= 2;
        }
        switch (epl)
        {
            case 1:
                LaunchBlockLine<1, RHS, DECAY, STASH>(p, a);
                break;
            case 2:
                LaunchBlockLine<2, RHS, DECAY, STASH>(p, a);
                break;
            case 4:
                LaunchBlockLine<4, RHS, DECAY, STASH>(p, a);
                break;
            case 8:
                LaunchBlockLine<8, RHS, DECAY, STASH>(p, a);
                break;
            default:
                LaunchBlockLine<16, RHS, DECAY, STASH>(p, a);
                break;
        }
        Check(hipGetLastError(), name);
        p->prof.End(p->clock.stream);
    }

    template <int C, int E, int MODE, bool PER>
    void LaunchColumnPer(halhip_adjoint2d *p, const ColArgs &a)
    {
        const dim3 grid(static_cast<unsigned>(a.blocks));
        const dim3 block(static_cast<unsigned>(64 * C));
        if (a.Y % C == 0)
        {
            hipLaunchKernelGGL((CellColumnSolve<C, E, MODE, true, PER>), grid, block, 0, p->clock.stream, a);
        }
        else
        {
            hipLaunchKernelGGL((CellColumnSolve<C, E, MODE, false, PER>), grid, block, 0, p->clock.stream, a);
        }
    }

    template <int C, int E, int MODE>
    void LaunchColumnE(halhip_adjoint2d *p, const ColArgs &a)
    {
        if (PeriodicColumns(p))
        {
            LaunchColumnPer<C, E, MODE, true>(p, a);
        }
        else
        {
            LaunchColumnPer<C, E, MODE, false>(p, a);
        }
    }

    // One x-line solve per column: E = the smallest power of two with
    // 64 E >= X, four columns per workgroup (two at E = 64).
    template <int MODE>
    void LaunchColumn(halhip_adjoint2d *p, const float *in, float *out, const char *name)
    {
        ColArgs a{};
        a.X = p->X;
        a.Y = p->Y;
        a.in = in;
        a.D = p->D.get();
        a.Dt = p->Dt.get();
        a.ib = p->ibXt.get();
        a.h = p->hXt.get();
        a.out = out;
        a.lineK = reinterpret_cast<const float4 *>(p->lineKX.get());
        a.edges = p->edges.get();
        a.wrapX = p->wrapX ? 1 : 0;
        a.wrapY = p->wrapY ? 1 : 0;
        a.dirX = p->dirX ? 1 : 0;
        a.dirY = p->dirY ? 1 : 0;
        const int X = a.X;
        const int C = (X > 2048) ? 2 : 4;
        a.blocks = (a.Y + C - 1) / C;
        p->prof.Begin(p->clock.stream, name);
        if (X <= 64)
        {
            LaunchColumnE<4, 1, MODE>(p, a);
        }
        else if (X <= 128)
        {
            LaunchColumnE<4, 2, MODE>(p, a);
        }
        else if (X <= 256)
        {
            LaunchColumnE<4, 4, MODE>(p, a);
        }
        else if (X <= 512)
        {
            LaunchColumnE<4, 8, MODE>(p, a);
        }
        else if (X <= 1024)
        {
            LaunchColumnE<4, 16, MODE>(p, a);
        }
        else if (X <= 2048)
        {
            LaunchColumnE<4, 32, MODE>(p, a);
        }
        else
        {
            LaunchColumnE<2, 64, MODE>(p, a);
        }
        Check(hipGetLastError(), name);
        p->prof.End(p->clock.stream);
    }

    // Lines along y: one per row of the [X][Y] field.
    LineArgs Rows(halhip_adjoint2d *p)
    {
        LineArgs a{};
        a.nLines = p->X;
        a.L = p->Y;
        a.rate = p->D.get();
        a.ib = p->ibY.get();
        a.h = p->hY.get();
        a.lineK = reinterpret_cast<const float4 *>(p->lineKY.get());
        a.edges = p->edges.get();
        a.wrapAcross = p->wrapX ? 1 : 0;
        a.dirAcross = p->dirX ? 1 : 0;
        a.dirAlong = p->dirY ? 1 : 0;
        return a;
    }

    // One forward tick from `in` to `out` (both [X][Y]); with stash pointers,
    // keeps s2 and s4.
    void ForwardTick(halhip_adjoint2d *p, const float *in, float *out, float *s2, float *s4)
    {
        float *s2dst = (s2 != nullptr) ? s2 : p->s2.get();
        LaunchColumn<0>(p, in, s2dst, "adjoint_fwd_solve_x");
        LineArgs b = Rows(p);
        b.in = s2dst;
        b.out = out;
        b.prev = in;
        b.negK = -p->decay;
        if (s4 != nullptr)
        {
            b.stash = s4;
            LaunchLine<0, true, true>(p, b, "adjoint_fwd_solve_y");
        }
        else
        {
            LaunchLine<0, true, false>(p, b, "adjoint_fwd_solve_y");
        }
    }

    // Enqueues the forward model from u0 in p->u, storing the checkpoints;
    // returns the buffer holding u_T (p->u or p->lamNext: tick t reads one
    // and writes the other, so no pointer changes hands and the pass can be
    // replayed as a graph).
    float *ForwardResult(halhip_adjoint2d *p)
    {
        return (p->ticks % 2 == 0) ? p->u.get() : p->lamNext.get();
    }

    float *EnqueueForward(halhip_adjoint2d *p)
    {
        const size_t n = Cells(p);
        float *buf[2] = {p->u.get(), p->lamNext.get()};
        Check(hipMemcpyAsync(p->ckpt.get(), buf[0], n * sizeof(float), hipMemcpyDeviceToDevice, p->clock.stream),
              "hipMemcpyAsync");
        for (int32_t t = 0; t < p->ticks; t++)
        {
            ForwardTick(p, buf[t % 2], buf[(t + 1) % 2], nullptr, nullptr);
            if ((t + 1) % p->every == 0 && t + 1 < p->ticks)
            {
                const size_t slot = static_cast<size_t>((t + 1) / p->every);
                Check(hipMemcpyAsync(p->ckpt.get() + slot * n, buf[(t + 1) % 2], n * sizeof(float),
                                     hipMemcpyDeviceToDevice, p->clock.stream),
                      "hipMemcpyAsync");
            }
        }
        return ForwardResult(p);
    }

    // Enqueues the whole gradient pass from u0 in p->u and the target in
    // p->work: the forward with checkpoints, the loss seed, then per segment
    // (last first) the recompute with stashes and the adjoint ticks.  Loss
    // partials go to partialJ, decay-gradient partials to partialK.
    void EnqueueGradient(halhip_adjoint2d *p)
    {
        const int32_t X = p->X;
        const int32_t Y = p->Y;
        const size_t n = Cells(p);
        hipStream_t s = p->clock.stream;
        float *uT = EnqueueForward(p);
        hipLaunchKernelGGL(LossSeed, dim3(kLossBlocks), dim3(256), 0, s, uT, p->work.get(), p->lam.get(),
                           static_cast<int64_t>(n), p->partialJ.get());
        Check(hipGetLastError(), "LossSeed launch");
        hipLaunchKernelGGL(FillZero, dim3(1024), dim3(256), 0, s, reinterpret_cast<float *>(p->grad.get()),
                           static_cast<int64_t>(2 * n));
        Check(hipGetLastError(), "FillZero launch");
        const dim3 gGrid(static_cast<unsigned>((Y + 63) / 64), static_cast<unsigned>((X + 3) / 4));
        const size_t gBlocks = static_cast<size_t>(gGrid.x) * gGrid.y;
        float *lam[2] = {p->lam.get(), p->lamNext.get()};
        int32_t j = 0; // adjoint ticks done: lam[j % 2] holds the current one
        const int32_t nSeg = (p->ticks + p->every - 1) / p->every;
        for (int32_t seg = nSeg - 1; seg >= 0; seg--)
        {
            const int32_t t0 = seg * p->every;
            const int32_t t1 = std::min(p->ticks, t0 + p->every);
            // Recompute the segment from its checkpoint, keeping u, s2, s4.
            Check(hipMemcpyAsync(p->stashU.get(), p->ckpt.get() + static_cast<size_t>(seg) * n, n * sizeof(float),
                                 hipMemcpyDeviceToDevice, s),
                  "hipMemcpyAsync");
            for (int32_t t = t0; t < t1; t++)
            {
                const size_t k = static_cast<size_t>(t - t0);
                float *next = (t + 1 < t1) ? p->stashU.get() + (k + 1) * n : p->s2.get();
                ForwardTick(p, p->stashU.get() + k * n, next, p->stashS2.get() + k * n, p->stashS4.get() + k * n);
            }
            // Adjoint ticks backwards through the segment.
            for (int32_t t = t1 - 1; t >= t0; t--, j++)
            {
                const size_t k = static_cast<size_t>(t - t0);
                LineArgs b = Rows(p);
                b.in = lam[j % 2];
                b.out = p->m1.get();
                LaunchLine<2, false, false>(p, b, "adjoint_solve_y");
                LaunchColumn<1>(p, p->m1.get(), p->m2.get(), "adjoint_solve_x");
                GradArgs g{};
                g.X = X;
                g.Y = Y;
                g.D = p->D.get();
                g.m1 = p->m1.get();
                g.m2 = p->m2.get();
                g.s2 = p->stashS2.get() + k * n;
                g.s4 = p->stashS4.get() + k * n;
                g.u = p->stashU.get() + k * n;
                g.lam = lam[j % 2];
                g.lamNext = lam[(j + 1) % 2];
                g.gradD = p->grad.get();
                g.partialK = p->partialK.get() + static_cast<size_t>(t) * gBlocks;
                g.negK = -p->decay;
                g.edges = p->edges.get();
                g.wrapX = p->wrapX ? 1 : 0;
                g.wrapY = p->wrapY ? 1 : 0;
                g.dirX = p->dirX ? 1 : 0;
                g.dirY = p->dirY ? 1 : 0;
                p->prof.Begin(s, "adjoint_gradient");
                hipLaunchKernelGGL(CellGradient, gGrid, dim3(256), 0, s, g);
                Check(hipGetLastError(), "CellGradient launch");
                p->prof.End(s);
            }
        }
    }

    // Runs `body` (a whole pass) on the plan's stream: replayed from a graph
    // captured on first use, or launched directly while profiling (so each
    // kernel gets its own events).
    template <class F>
    void RunPass(halhip_adjoint2d *p, int32_t pass, F body)
    {
        if (p->prof.On())
        {
            body();
            return;
        }
        std::string key;
        AppendKey(key, pass);
        hipGraphExec_t exec = FindOrCapture(p->graphs, key, p->clock.stream, body);
        Check(hipGraphLaunch(exec, p->clock.stream), "hipGraphLaunch");
    }
}

extern "C"
{
    int halhip_adjoint2d_create(halhip_context *ctx, int32_t xDim, int32_t yDim, halhip_adjoint2d **out)
    {
        halhip_adjoint2d *p = nullptr;
        int rc = Guard(
            [&]
            {
                Require(out != nullptr, "halhip_adjoint2d_create: null out");
                *out = nullptr;
                Require(ctx != nullptr, "halhip_adjoint2d_create: null context");
                Require(xDim > 0 && yDim > 0, "halhip_adjoint2d_create: dimensions must be positive");
                if (xDim > 4096 || yDim > 4096)
                {
                    Fail(HALHIP_ERR_UNSUPPORTED, "halhip_adjoint2d_create: sides are limited to 4096");
                }
                Bind(ctx);
                p = new halhip_adjoint2d();
                p->ctx = ctx;
                p->X = xDim;
                p->Y = yDim;
                const size_t n = Cells(p);
                for (DeviceArray<float> *b : {&p->D, &p->Dt, &p->ibXt, &p->ibY, &p->hXt, &p->hY, &p->u, &p->s2, &p->work,
                                              &p->lam, &p->lamNext,
                                              &p->m1, &p->m2})
                {
                    b->Allocate(ctx, n);
                }
                p->grad.Allocate(ctx, n);
                p->partialJ.Allocate(ctx, 1024);
                p->edges.Allocate(ctx, 2 * static_cast<size_t>(xDim) + 2 * static_cast<size_t>(yDim));
                p->lineKX.Allocate(ctx, 4 * static_cast<size_t>(yDim));
                p->lineKY.Allocate(ctx, 4 * static_cast<size_t>(xDim));
                p->clock.Create();
                *out = p;
            });
        if (rc != HALHIP_OK && p != nullptr)
        {
            p->clock.Destroy();
            delete p;
        }
        return rc;
    }

    int halhip_adjoint2d_destroy(halhip_adjoint2d *p)
    {
        return Guard(
            [&]
            {
                if (p == nullptr)
                {
                    return;
                }
                Bind(p->ctx);
                p->clock.Destroy();
                delete p;
            });
    }

    int halhip_adjoint2d_set_problem_bc(halhip_adjoint2d *p, const double *diffusivity, double decay, int32_t ticks,
                                        int32_t checkpointEvery, int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequireAdj(p, "halhip_adjoint2d_set_problem");
                Require(diffusivity != nullptr, "halhip_adjoint2d_set_problem: null diffusivity");
                Require(ticks >= 1 && checkpointEvery >= 1, "halhip_adjoint2d_set_problem: ticks and checkpoint >= 1");
                Require(std::isfinite(static_cast<float>(decay)), "halhip_adjoint2d_set_problem: bad decay");
                const int32_t X = p->X;
                const int32_t Y = p->Y;
                const size_t n = Cells(p);
                EdgeValues ev;
                ev.From(bc, X, Y);
                std::vector<double> Dxy(diffusivity, diffusivity + n);
                for (double v : Dxy)
                {
                    // finite in FP32, which the device computes in
                    Require(std::isfinite(static_cast<float>(v)) && v >= 0.0,
                            "halhip_adjoint2d_set_problem: diffusivity must be finite and non-negative");
                }
                const bool wx = wrapX != 0;
                const bool wy = wrapY != 0;
                // The device computes with D in FP32; the pivots are formed
                // in FP64 from the same FP32 values.
                std::vector<float> Df(Dxy.begin(), Dxy.end());
                std::vector<double> Dr(Df.begin(), Df.end());
                std::vector<float> ibx, iby, hx, hy, kx, ky;
                LinePivots(X, Y, Dr, true, wx, ev.dirichlet && !wx, ibx, hx, kx);
                LinePivots(X, Y, Dr, false, wy, ev.dirichlet && !wy, iby, hy, ky);
                std::vector<float> edges;
                edges.reserve(2 * static_cast<size_t>(X) + 2 * static_cast<size_t>(Y));
                for (const std::vector<float> *e : {&ev.xlo, &ev.xhi, &ev.ylo, &ev.yhi})
                {
                    edges.insert(edges.end(), e->begin(), e->end());
                }
                auto up = [&](DeviceArray<float> &dst, const std::vector<float> &src)
                {
                    Check(hipMemcpy(dst.get(), src.data(), src.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
                };
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                // the x-line tables and D again, column-major, for the
                // column solve
                auto columnMajor = [&](const std::vector<float> &src)
                {
                    std::vector<float> t(src.size());
                    for (int32_t x = 0; x < X; x++)
                    {
                        for (int32_t y = 0; y < Y; y++)
                        {
                            t[static_cast<size_t>(y) * X + x] = src[static_cast<size_t>(x) * Y + y];
                        }
                    }
                    return t;
                };
                up(p->D, Df);
                up(p->Dt, columnMajor(Df));
                up(p->ibXt, columnMajor(ibx));
                up(p->ibY, iby);
                up(p->hXt, columnMajor(hx));
                up(p->hY, hy);
                up(p->lineKX, kx);
                up(p->lineKY, ky);
                up(p->edges, edges);
                p->wrapX = wx;
                p->wrapY = wy;
                p->dirX = ev.dirichlet && !wx;
                p->dirY = ev.dirichlet && !wy;
                p->graphs.Clear(); // they hold the old checkpoint and stash buffers
                p->decay = static_cast<float>(decay);
                p->ticks = ticks;
                p->every = checkpointEvery;
                const size_t nCk = static_cast<size_t>((ticks + checkpointEvery - 1) / checkpointEvery);
                p->ckpt.Allocate(p->ctx, nCk * n);
                const size_t seg = static_cast<size_t>(checkpointEvery);
                p->stashU.Allocate(p->ctx, seg * n);
                p->stashS2.Allocate(p->ctx, seg * n);
                p->stashS4.Allocate(p->ctx, seg * n);
                const size_t gblocks = static_cast<size_t>((Y + 63) / 64) * static_cast<size_t>((X + 3) / 4);
                p->partialK.Allocate(p->ctx, gblocks * static_cast<size_t>(ticks));
                p->ready = true;
            });
    }

    int halhip_adjoint2d_set_problem(halhip_adjoint2d *p, const double *diffusivity, double decay, int32_t ticks,
                                     int32_t checkpointEvery)
    {
        return halhip_adjoint2d_set_problem_bc(p, diffusivity, decay, ticks, checkpointEvery, 0, 0, nullptr);
    }

    int halhip_adjoint2d_forward(halhip_adjoint2d *p, const float *u0, float *uT)
    {
        return Guard(
            [&]
            {
                RequireAdj(p, "halhip_adjoint2d_forward");
                Require(p->ready, "halhip_adjoint2d_forward: call halhip_adjoint2d_set_problem first");
                Require(u0 != nullptr && uT != nullptr, "halhip_adjoint2d_forward: null array");
                const size_t n = Cells(p);
                Check(hipMemcpyAsync(p->u.get(), u0, n * sizeof(float), hipMemcpyHostToDevice, p->clock.stream),
                      "hipMemcpyAsync");
                RunPass(p, 0, [&] { EnqueueForward(p); });
                Check(hipMemcpyAsync(uT, ForwardResult(p), n * sizeof(float), hipMemcpyDeviceToHost, p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                p->prof.Collect();
            });
    }

    int halhip_adjoint2d_gradient(halhip_adjoint2d *p, const float *u0, const float *target, double *loss,
                                  float *gradDiffusivity, double *gradDecay)
    {
        return Guard(
            [&]
            {
                RequireAdj(p, "halhip_adjoint2d_gradient");
                Require(p->ready, "halhip_adjoint2d_gradient: call halhip_adjoint2d_set_problem first");
                Require(u0 != nullptr && target != nullptr && loss != nullptr && gradDiffusivity != nullptr &&
                            gradDecay != nullptr,
                        "halhip_adjoint2d_gradient: null array");
                const int32_t X = p->X;
                const int32_t Y = p->Y;
                const size_t n = Cells(p);
                hipStream_t s = p->clock.stream;
                Check(hipMemcpyAsync(p->u.get(), u0, n * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
                // The target goes through the work buffer.
                Check(hipMemcpyAsync(p->work.get(), target, n * sizeof(float), hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync");
                RunPass(p, 1, [&] { EnqueueGradient(p); });
                const dim3 gGrid(static_cast<unsigned>((Y + 63) / 64), static_cast<unsigned>((X + 3) / 4));
                const size_t gBlocks = static_cast<size_t>(gGrid.x) * gGrid.y;
                std::vector<double> hostK(gBlocks * static_cast<size_t>(p->ticks));
                double gk = 0.0;
                const unsigned lossBlocks = kLossBlocks;
                std::vector<double> hostJ(lossBlocks);
                Check(hipMemcpyAsync(hostJ.data(), p->partialJ.get(), lossBlocks * sizeof(double),
                                     hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync");
                hipLaunchKernelGGL(ToFloat, dim3(1024), dim3(256), 0, s, static_cast<const double *>(p->grad.get()),
                                   p->work.get(), static_cast<int64_t>(n));
                Check(hipGetLastError(), "ToFloat launch");
                Check(hipMemcpyAsync(gradDiffusivity, p->work.get(), n * sizeof(float), hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync");
                Check(hipMemcpyAsync(hostK.data(), p->partialK.get(), hostK.size() * sizeof(double),
                                     hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(s), "hipStreamSynchronize");
                double J = 0.0;
                for (double v : hostJ)
                {
                    J += v;
                }
                for (double v : hostK)
                {
                    gk += v;
                }
                *loss = J;
                *gradDecay = gk;
                p->prof.Collect();
            });
    }

    int halhip_adjoint2d_set_profiling(halhip_adjoint2d *p, int32_t on)
    {
        return Guard(
            [&]
            {
                RequireAdj(p, "halhip_adjoint2d_set_profiling");
                p->prof.Reset();
                p->prof.SetOn(on != 0);
            });
    }

    int halhip_adjoint2d_profile(halhip_adjoint2d *p, int32_t maxKernels, int32_t *nKernels, char *names,
                                 int32_t namesLen, double *totalMs, int64_t *launches)
    {
        return Guard(
            [&]
            {
                RequireAdj(p, "halhip_adjoint2d_profile");
                p->prof.Export(maxKernels, nKernels, names, namesLen, totalMs, launches);
            });
    }
}
