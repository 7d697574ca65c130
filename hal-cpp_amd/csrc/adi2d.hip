// adi2d.hip -- Peaceman-Rachford ADI diffusion for gfx950.
//
// Reference semantics: Internal::Diffusion2DADI (include/Hal/Pde/ADIequations.h:306-378),
// ADISolveRow coefficients (ADIequations.h:225-270), explicit half-steps
// (src/Pde/ADIequations.cpp:19-47, ADIequations.h:54-94), Dirichlet implicit
// source (ADIequations.h:198-220) and the Thomas / Sherman-Morrison solves
// (include/Hal/Pde/TdmaSolver.h:27-120).  A tick is sweep A (implicit along x,
// explicit along y) into u*, then sweep B (implicit along y, explicit along
// x) into u'.
//
// Method.  For HAL's uniform-rate lines the Thomas elimination factors
// (modified diagonal b'_i and multipliers) do not depend on the right-hand
// side, so they are computed once on the host in FP64.  The solve is then two
// first-order linear recurrences:
//     forward   d'_i = fw_i * d'_{i-1} + d_i
//     backward  x_i  = bw_i * x_{i+1} + ib_i * d'_i
// (fw = -a_i/b'_{i-1}, bw = -c_i/b'_i, ib = 1/b'_i), and a periodic line adds
// the Sherman-Morrison correction x -= factor * z with z precomputed.  A first
// order recurrence composes associatively, so it runs as a parallel scan.
//
// Layout.  The grid is cut into strips of S full rows (x) by all yDim columns.
// One workgroup owns one strip and keeps it in registers: thread t holds CPT
// adjacent columns of every row of the strip, so each row load is one
// coalesced 16-byte-per-lane access along the contiguous axis.
//   - Sweep B lines (one per row) lie entirely inside the workgroup: an
//     in-thread serial pass over CPT values, a wavefront scan with shuffles and
//     a cross-wavefront fold through LDS.
//   - Sweep A lines (one per column) cross all strips.  Each strip reduces its
//     part of every column to two numbers (forward end value B with zero
//     carry-in, backward start value D with zero carries); a small kernel scans
//     those across strips (AdiStripScan) and hands each strip its carry-in
//     values, after which the strip finishes sweep A, runs sweep B and writes
//     u' in place.  The reduction for the next tick is taken from u' while it
//     is still in registers, so a plain tick reads and writes the field once
//     (8 B/cell) plus ~1 B/cell of strip carries.
// The boundary rows of u* that sweep B's explicit x-term needs from the
// neighbouring strips are recovered from the carries: u*(e+1) is the backward
// carry-in and u*(s-1) = bw_{s-1} u*(s) + ib_{s-1} d'_{s-1}.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <string>

using namespace halhip;

namespace
{
    constexpr int kMaxWaves = 16;

    struct AdiArgs
    {
        int32_t X;
        int32_t Y;
        int32_t P;         // strips
        float hr;          // rate/2, explicit half-step weight
        float r;           // rate, weight of a Dirichlet value in the implicit RHS
        float negDecay;
        int32_t wrapX;
        int32_t wrapY;
        int32_t dirichlet; // ghost-cell Dirichlet on non-wrapped axes
        int32_t vecOk;     // yDim % CPT == 0: rows load as vectors
        // sweep A tables, padded to P*S rows (padding: fw=1, ib=bw=z=0)
        const float *xfw;
        const float *xib;
        const float *xbw;
        const float *xz;
        // sweep B tables, padded to T*CPT columns (same padding)
        const float *yfw;
        const float *yib;
        const float *ybw;
        const float *yz;
        float yvEnd;
        float yInv;
        // Dirichlet edge values
        const float *bcXlo;
        const float *bcXhi;
        const float *bcYlo;
        const float *bcYhi;
        // strip reduction and carries, [2][P][Y] each
        float *red;
        float *carry;
        float *colInfo; // [3][Y]: SM factor, u*(0,y), u*(X-1,y)
        const float *stripA;
        const float *stripC;
        const float *stripE;
        float xvEnd;
        float xInv;
        float xibLast;
        float xz0;
        float xzLast;
    };

    // ---- register-tile helpers ------------------------------------------------

    template <int CPT>
    __device__ __forceinline__ void LoadRow(const float *__restrict__ row, int y0, int Y, bool vec, float (&d)[CPT])
    {
        if (vec && y0 + CPT <= Y)
        {
            if constexpr (CPT % 4 == 0)
            {
#pragma unroll
                for (int q = 0; q < CPT / 4; q++)
                {
                    float4 t = *reinterpret_cast<const float4 *>(row + y0 + 4 * q);
                    d[4 * q + 0] = t.x;
                    d[4 * q + 1] = t.y;
                    d[4 * q + 2] = t.z;
                    d[4 * q + 3] = t.w;
                }
                return;
            }
            else if constexpr (CPT == 2)
            {
                float2 t = *reinterpret_cast<const float2 *>(row + y0);
                d[0] = t.x;
                d[1] = t.y;
                return;
            }
        }
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            d[c] = (y0 + c < Y) ? row[y0 + c] : 0.0f;
        }
    }

    template <int CPT>
    __device__ __forceinline__ void StoreRow(float *__restrict__ row, int y0, int Y, bool vec, const float (&d)[CPT])
    {
        if (vec && y0 + CPT <= Y)
        {
            if constexpr (CPT % 4 == 0)
            {
#pragma unroll
                for (int q = 0; q < CPT / 4; q++)
                {
                    *reinterpret_cast<float4 *>(row + y0 + 4 * q) =
                        make_float4(d[4 * q + 0], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3]);
                }
                return;
            }
            else if constexpr (CPT == 2)
            {
                *reinterpret_cast<float2 *>(row + y0) = make_float2(d[0], d[1]);
                return;
            }
        }
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            if (y0 + c < Y)
            {
                row[y0 + c] = d[c];
            }
        }
    }

    struct LdsEdges
    {
        float *left;  // [kMaxWaves][S] first column of each wavefront
        float *right; // [kMaxWaves][S] last column of each wavefront
        float *first; // [S] column y = 0
        float *last;  // [S] column y = Y-1
    };

    // Sweep A right-hand side in place: v <- v + hr*(dv/dy+ + dv/dy-) plus the
    // Dirichlet implicit term r*bc on rows 0 and X-1.  Rows past X and columns
    // past Y become 0.  ExplicitDiffusionY2ADI (src/Pde/ADIequations.cpp:34-47,
    // ADIequations.h:79-94) + ImplicitBoundarySource (ADIequations.h:198-220).
    template <int CPT, int S>
    __device__ __forceinline__ void SweepARhs(const AdiArgs &a, int s0, int y0, int tid, float (&v)[S][CPT],
                                              const LdsEdges &e)
    {
        const int lane = tid & 63;
        const int wave = tid >> 6;
        const int nWaves = blockDim.x >> 6;
        const int lastT = (a.Y - 1) / CPT;
        const int lastC = (a.Y - 1) % CPT;
        if (lane == 0)
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                e.left[wave * S + r] = v[r][0];
            }
        }
        if (lane == 63)
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                e.right[wave * S + r] = v[r][CPT - 1];
            }
        }
        if (tid == 0)
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                e.first[r] = v[r][0];
            }
        }
        if (tid == lastT)
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    if (c == lastC)
                    {
                        e.last[r] = v[r][c];
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const int x = s0 + r;
            float left = __shfl_up(v[r][CPT - 1], 1);
            float right = __shfl_down(v[r][0], 1);
            if (x >= a.X)
            {
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    v[r][c] = 0.0f;
                }
                continue;
            }
            if (lane == 0 && wave > 0)
            {
                left = e.right[(wave - 1) * S + r];
            }
            if (lane == 63 && wave + 1 < nWaves)
            {
                right = e.left[(wave + 1) * S + r];
            }
            const float firstV = e.first[r];
            const float lastV = e.last[r];
            float bcLo = 0.0f;
            float bcHi = 0.0f;
            if (a.dirichlet && !a.wrapY)
            {
                bcLo = a.bcYlo[x];
                bcHi = a.bcYhi[x];
            }
            const bool srcRowLo = a.dirichlet && !a.wrapX && x == 0;
            const bool srcRowHi = a.dirichlet && !a.wrapX && x == a.X - 1;
            float prev = left;
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                const int y = y0 + c;
                const float cur = v[r][c];
                const float nxt = (c + 1 < CPT) ? v[r][c + 1] : right;
                float rhs = 0.0f;
                if (y < a.Y)
                {
                    float dp;
                    float dm;
                    if (y + 1 < a.Y)
                    {
                        dp = nxt - cur;
                    }
                    else
                    {
                        dp = a.wrapY ? (firstV - cur) : (a.dirichlet ? (bcHi - cur) * 2.0f : 0.0f);
                    }
                    if (y > 0)
                    {
                        dm = prev - cur;
                    }
                    else
                    {
                        dm = a.wrapY ? (lastV - cur) : (a.dirichlet ? (bcLo - cur) * 2.0f : 0.0f);
                    }
                    rhs = cur + (dp + dm) * a.hr;
                    if (srcRowLo || srcRowHi)
                    {
                        float src = 0.0f;
                        if (srcRowLo)
                        {
                            src += a.bcXlo[y];
                        }
                        if (srcRowHi)
                        {
                            src += a.bcXhi[y];
                        }
                        rhs += src * a.r;
                    }
                }
                prev = cur;
                v[r][c] = rhs;
            }
        }
    }

    // Zero-carry local solve of sweep A on the strip (v holds the RHS and is
    // consumed) and write of the strip's two reduction values per column.
    template <int CPT, int S>
    __device__ __forceinline__ void ReduceA(const AdiArgs &a, int p, int s0, int y0, float (&v)[S][CPT])
    {
        float outB[CPT];
        float outD[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            float acc = 0.0f;
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                acc = a.xfw[s0 + r] * acc + v[r][c];
                v[r][c] = acc;
            }
            outB[c] = acc;
            acc = 0.0f;
#pragma unroll
            for (int r = S - 1; r >= 0; r--)
            {
                acc = a.xbw[s0 + r] * acc + a.xib[s0 + r] * v[r][c];
            }
            outD[c] = acc;
        }
        const size_t base = static_cast<size_t>(p) * a.Y;
        const size_t plane = static_cast<size_t>(a.P) * a.Y;
        StoreRow<CPT>(a.red + base, y0, a.Y, a.vecOk != 0, outB);
        StoreRow<CPT>(a.red + plane + base, y0, a.Y, a.vecOk != 0, outD);
    }

    // Finishes sweep A on the strip from the carries (v: RHS in, u* out) and
    // returns u* of the rows just above (s0-1) and below (s0+S) the strip.
    template <int CPT, int S>
    __device__ __forceinline__ void FinishA(const AdiArgs &a, int p, int s0, int y0, float (&v)[S][CPT],
                                            float (&up)[CPT], float (&dn)[CPT])
    {
        const size_t base = static_cast<size_t>(p) * a.Y;
        const size_t plane = static_cast<size_t>(a.P) * a.Y;
        const bool vec = a.vecOk != 0;
        float cf[CPT];
        float xb[CPT];
        float factor[CPT];
        LoadRow<CPT>(a.carry + base, y0, a.Y, vec, cf);
        LoadRow<CPT>(a.carry + plane + base, y0, a.Y, vec, xb);
        LoadRow<CPT>(a.colInfo, y0, a.Y, vec, factor);
        const bool hasUp = s0 > 0;
        const bool hasDn = s0 + S < a.X;
        const float zUp = hasUp ? a.xz[s0 - 1] : 0.0f;
        const float zDn = hasDn ? a.xz[s0 + S] : 0.0f;
        const float bwUp = hasUp ? a.xbw[s0 - 1] : 0.0f;
        const float ibUp = hasUp ? a.xib[s0 - 1] : 0.0f;
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            float acc = cf[c];
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                acc = a.xfw[s0 + r] * acc + v[r][c];
                v[r][c] = acc;
            }
            acc = xb[c];
#pragma unroll
            for (int r = S - 1; r >= 0; r--)
            {
                acc = a.xbw[s0 + r] * acc + a.xib[s0 + r] * v[r][c];
                v[r][c] = acc;
            }
            up[c] = bwUp * v[0][c] + ibUp * cf[c] - factor[c] * zUp;
            dn[c] = xb[c] - factor[c] * zDn;
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                v[r][c] -= factor[c] * a.xz[s0 + r];
            }
        }
    }

    // Sweep B right-hand side in place from u*: v <- v + hr*(dv/dx+ + dv/dx-)
    // plus the Dirichlet implicit term on columns 0 and Y-1.
    // ExplicitDiffusionX2ADI (src/Pde/ADIequations.cpp:19-32, ADIequations.h:54-69).
    template <int CPT, int S>
    __device__ __forceinline__ void SweepBRhs(const AdiArgs &a, int s0, int y0, float (&v)[S][CPT],
                                              const float (&up)[CPT], const float (&dn)[CPT])
    {
        const bool vec = a.vecOk != 0;
        float lo[CPT];
        float hi[CPT];
        // Off-grid x neighbours: the wrapped row of u* or the Dirichlet value.
        if (a.wrapX)
        {
            LoadRow<CPT>(a.colInfo + 2 * static_cast<size_t>(a.Y), y0, a.Y, vec, lo);
            LoadRow<CPT>(a.colInfo + static_cast<size_t>(a.Y), y0, a.Y, vec, hi);
        }
        else if (a.dirichlet)
        {
            LoadRow<CPT>(a.bcXlo, y0, a.Y, vec, lo);
            LoadRow<CPT>(a.bcXhi, y0, a.Y, vec, hi);
        }
        else
        {
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                lo[c] = 0.0f;
                hi[c] = 0.0f;
            }
        }
        float prevRow[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            prevRow[c] = up[c];
        }
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const int x = s0 + r;
            if (x >= a.X)
            {
                break;
            }
            float srcLo = 0.0f;
            float srcHi = 0.0f;
            const bool ySrc = a.dirichlet && !a.wrapY;
            if (ySrc)
            {
                srcLo = a.bcYlo[x];
                srcHi = a.bcYhi[x];
            }
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                const int y = y0 + c;
                const float cur = v[r][c];
                float dDown;
                float dUp;
                if (x + 1 < a.X)
                {
                    dDown = ((r + 1 < S) ? v[r + 1][c] : dn[c]) - cur;
                }
                else
                {
                    dDown = a.wrapX ? (hi[c] - cur) : (a.dirichlet ? (hi[c] - cur) * 2.0f : 0.0f);
                }
                if (x > 0)
                {
                    dUp = prevRow[c] - cur;
                }
                else
                {
                    dUp = a.wrapX ? (lo[c] - cur) : (a.dirichlet ? (lo[c] - cur) * 2.0f : 0.0f);
                }
                float rhs = cur + (dDown + dUp) * a.hr;
                if (ySrc)
                {
                    float src = 0.0f;
                    if (y == 0)
                    {
                        src += srcLo;
                    }
                    if (y == a.Y - 1)
                    {
                        src += srcHi;
                    }
                    rhs += src * a.r;
                }
                prevRow[c] = cur;
                v[r][c] = (y < a.Y) ? rhs : 0.0f;
            }
        }
    }

    // Sweep B solve of every row of the strip across the workgroup.
    template <int CPT, int S>
    __device__ __forceinline__ void SolveB(const AdiArgs &a, int y0, int tid, float (&v)[S][CPT], float *sTot,
                                           float *sAw, float *sEnds)
    {
        const int lane = tid & 63;
        const int wave = tid >> 6;
        const int nWaves = blockDim.x >> 6;
        float fw[CPT];
        float bw[CPT];
        float ib[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            fw[c] = a.yfw[y0 + c];
            bw[c] = a.ybw[y0 + c];
            ib[c] = a.yib[y0 + c];
        }

        // Forward recurrence d'_y = fw_y d'_{y-1} + d_y.
        float aSeg = 1.0f;
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            aSeg *= fw[c];
        }
        float aStage[6];
        float aInc = aSeg;
#pragma unroll
        for (int k = 0; k < 6; k++)
        {
            const int o = 1 << k;
            aStage[k] = aInc;
            float aUp = __shfl_up(aInc, o);
            if (lane >= o)
            {
                aInc = aUp * aInc;
            }
        }
        float aEx = __shfl_up(aInc, 1);
        if (lane == 0)
        {
            aEx = 1.0f;
        }
        if (lane == 63)
        {
            sAw[wave] = aInc;
        }
        float bEx[S];
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            float b = 0.0f;
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                b = fw[c] * b + v[r][c];
            }
#pragma unroll
            for (int k = 0; k < 6; k++)
            {
                const int o = 1 << k;
                float bUp = __shfl_up(b, o);
                if (lane >= o)
                {
                    b = aStage[k] * bUp + b;
                }
            }
            float be = __shfl_up(b, 1);
            bEx[r] = (lane == 0) ? 0.0f : be;
            if (lane == 63)
            {
                sTot[wave * S + r] = b;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            float cw = 0.0f;
            for (int w = 0; w < wave; w++)
            {
                cw = sAw[w] * cw + sTot[w * S + r];
            }
            float acc = aEx * cw + bEx[r];
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                acc = fw[c] * acc + v[r][c];
                v[r][c] = acc;
            }
        }
        __syncthreads();

        // Backward recurrence x_y = bw_y x_{y+1} + ib_y d'_y, scanned from the right.
        aSeg = 1.0f;
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            aSeg *= bw[c];
        }
        aInc = aSeg;
#pragma unroll
        for (int k = 0; k < 6; k++)
        {
            const int o = 1 << k;
            aStage[k] = aInc;
            float aDn = __shfl_down(aInc, o);
            if (lane + o < 64)
            {
                aInc = aDn * aInc;
            }
        }
        aEx = __shfl_down(aInc, 1);
        if (lane == 63)
        {
            aEx = 1.0f;
        }
        if (lane == 0)
        {
            sAw[wave] = aInc;
        }
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            float b = 0.0f;
#pragma unroll
            for (int c = CPT - 1; c >= 0; c--)
            {
                b = bw[c] * b + ib[c] * v[r][c];
            }
#pragma unroll
            for (int k = 0; k < 6; k++)
            {
                const int o = 1 << k;
                float bDn = __shfl_down(b, o);
                if (lane + o < 64)
                {
                    b = aStage[k] * bDn + b;
                }
            }
            float be = __shfl_down(b, 1);
            bEx[r] = (lane == 63) ? 0.0f : be;
            if (lane == 0)
            {
                sTot[wave * S + r] = b;
            }
        }
        __syncthreads();
        const int lastT = (a.Y - 1) / CPT;
        const int lastC = (a.Y - 1) % CPT;
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            float cw = 0.0f;
            for (int w = nWaves - 1; w > wave; w--)
            {
                cw = sAw[w] * cw + sTot[w * S + r];
            }
            float acc = aEx * cw + bEx[r];
#pragma unroll
            for (int c = CPT - 1; c >= 0; c--)
            {
                acc = bw[c] * acc + ib[c] * v[r][c];
                v[r][c] = acc;
            }
            if (a.wrapY)
            {
                if (tid == 0)
                {
                    sEnds[r] = v[r][0];
                }
                if (tid == lastT)
                {
#pragma unroll
                    for (int c = 0; c < CPT; c++)
                    {
                        if (c == lastC)
                        {
                            sEnds[S + r] = v[r][c];
                        }
                    }
                }
            }
        }
        if (a.wrapY)
        {
            // Sherman-Morrison combination (TdmaSolver.h:112-119).
            __syncthreads();
            float z[CPT];
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                z[c] = a.yz[y0 + c];
            }
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                const float factor = (sEnds[r] + sEnds[S + r] * a.yvEnd) * a.yInv;
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    v[r][c] -= factor * z[c];
                }
            }
        }
    }

    // MODE 0: reduction of the current field only (first tick of a run).
    // MODE 1: plain tick in place, then the reduction of u' for the next tick.
    // MODE 2: tick with decay: u' = ADI(u) - k*u (no reduction; the next tick
    //         starts with a MODE 0 pass after point sources are added).
    template <int CPT, int S, int MODE>
    __global__ __launch_bounds__(1024) void AdiStrip(AdiArgs a, float *__restrict__ u)
    {
        __shared__ float sLeft[kMaxWaves * S];
        __shared__ float sRight[kMaxWaves * S];
        __shared__ float sFirst[S];
        __shared__ float sLast[S];
        __shared__ float sTot[kMaxWaves * S];
        __shared__ float sAw[kMaxWaves];
        __shared__ float sEnds[2 * S];
        const LdsEdges edges{sLeft, sRight, sFirst, sLast};

        const int p = blockIdx.x;
        const int s0 = p * S;
        const int tid = threadIdx.x;
        const int y0 = tid * CPT;
        const bool vec = a.vecOk != 0;

        float v[S][CPT];
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            if (s0 + r < a.X)
            {
                LoadRow<CPT>(u + static_cast<size_t>(s0 + r) * a.Y, y0, a.Y, vec, v[r]);
            }
            else
            {
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    v[r][c] = 0.0f;
                }
            }
        }
        SweepARhs<CPT, S>(a, s0, y0, tid, v, edges);
        if constexpr (MODE == 0)
        {
            ReduceA<CPT, S>(a, p, s0, y0, v);
            return;
        }
        float up[CPT];
        float dn[CPT];
        FinishA<CPT, S>(a, p, s0, y0, v, up, dn);
        SweepBRhs<CPT, S>(a, s0, y0, v, up, dn);
        SolveB<CPT, S>(a, y0, tid, v, sTot, sAw, sEnds);
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            if (s0 + r < a.X)
            {
                float *row = u + static_cast<size_t>(s0 + r) * a.Y;
                if constexpr (MODE == 2)
                {
                    float old[CPT];
                    LoadRow<CPT>(row, y0, a.Y, vec, old);
#pragma unroll
                    for (int c = 0; c < CPT; c++)
                    {
                        v[r][c] += old[c] * a.negDecay;
                    }
                }
                StoreRow<CPT>(row, y0, a.Y, vec, v[r]);
            }
        }
        if constexpr (MODE == 1)
        {
            __syncthreads();
            SweepARhs<CPT, S>(a, s0, y0, tid, v, edges);
            ReduceA<CPT, S>(a, p, s0, y0, v);
        }
    }

    // Scan of the strip reductions along each column (sweep A across strips).
    // Block: 64 columns (lanes) x G wavefronts; wavefront g owns strips
    // [g*Q, (g+1)*Q).  Writes the carry-in values of every strip and, per
    // column, the periodic Sherman-Morrison factor and the wrapped rows of u*.
    __global__ __launch_bounds__(1024) void AdiStripScan(AdiArgs a, int32_t Q)
    {
        __shared__ float sB[kMaxWaves][64];
        __shared__ float sA[kMaxWaves];
        __shared__ float sB2[kMaxWaves][64];
        __shared__ float sA2[kMaxWaves];
        __shared__ float sCf[64];
        __shared__ float sX0[64];
        const int lane = threadIdx.x & 63;
        const int g = threadIdx.x >> 6;
        const int G = blockDim.x >> 6;
        const int y = blockIdx.x * 64 + lane;
        const bool active = y < a.Y;
        const int p0 = g * Q;
        const int p1 = min(a.P, p0 + Q);
        const size_t plane = static_cast<size_t>(a.P) * a.Y;
        const float *B = a.red;
        const float *D = a.red + plane;
        float *cfIn = a.carry;
        float *xbIn = a.carry + plane;

        float acc = 0.0f;
        float aSeg = 1.0f;
        for (int p = p0; p < p1; p++)
        {
            const float A = a.stripA[p];
            acc = A * acc + (active ? B[static_cast<size_t>(p) * a.Y + y] : 0.0f);
            aSeg *= A;
        }
        sB[g][lane] = acc;
        if (lane == 0)
        {
            sA[g] = aSeg;
        }
        __syncthreads();
        float cf = 0.0f;
        for (int w = 0; w < g; w++)
        {
            cf = sA[w] * cf + sB[w][lane];
        }
        for (int p = p0; p < p1; p++)
        {
            const size_t k = static_cast<size_t>(p) * a.Y + y;
            const float bv = active ? B[k] : 0.0f;
            if (active)
            {
                cfIn[k] = cf;
            }
            cf = a.stripA[p] * cf + bv;
        }
        if (p0 < p1 && p1 == a.P)
        {
            sCf[lane] = cf;
        }

        acc = 0.0f;
        aSeg = 1.0f;
        for (int p = p1 - 1; p >= p0; p--)
        {
            const size_t k = static_cast<size_t>(p) * a.Y + y;
            const float C = a.stripC[p];
            const float off = active ? (D[k] + a.stripE[p] * cfIn[k]) : 0.0f;
            acc = C * acc + off;
            aSeg *= C;
        }
        sB2[g][lane] = acc;
        if (lane == 0)
        {
            sA2[g] = aSeg;
        }
        __syncthreads();
        float xb = 0.0f;
        for (int w = G - 1; w > g; w--)
        {
            xb = sA2[w] * xb + sB2[w][lane];
        }
        for (int p = p1 - 1; p >= p0; p--)
        {
            const size_t k = static_cast<size_t>(p) * a.Y + y;
            float off = 0.0f;
            if (active)
            {
                xbIn[k] = xb;
                off = D[k] + a.stripE[p] * cfIn[k];
            }
            xb = a.stripC[p] * xb + off;
        }
        if (g == 0)
        {
            sX0[lane] = xb;
        }
        __syncthreads();
        if (g == 0 && active)
        {
            const float x0 = sX0[lane];
            const float xLast = a.xibLast * sCf[lane];
            const float factor = (x0 + xLast * a.xvEnd) * a.xInv;
            a.colInfo[y] = factor;
            a.colInfo[a.Y + y] = x0 - factor * a.xz0;
            a.colInfo[2 * static_cast<size_t>(a.Y) + y] = xLast - factor * a.xzLast;
        }
    }

    // Point sources, applied serially so repeated indices add in list order.
    __global__ void AdiAddSources(float *__restrict__ u, int32_t n, const int64_t *__restrict__ idx,
                                  const float *__restrict__ val)
    {
        if (threadIdx.x == 0 && blockIdx.x == 0)
        {
            for (int k = 0; k < n; k++)
            {
                u[idx[k]] += val[k];
            }
        }
    }

    // ---- host: line factorisation (FP64) ---------------------------------------

    // Thomas factors of one ADI line with the reference's coefficients
    // (ADISolveRow, ADIequations.h:225-270; TDMA / TDMAperiodic,
    // TdmaSolver.h:27-120), padded to `padded` entries with identity rows.
    struct AxisFactors
    {
        std::vector<double> fw, ib, bw, z;
        double vEnd = 0.0;
        double inv = 0.0;
    };

    AxisFactors FactorLine(int32_t n, int32_t padded, double rate, bool wrap, bool dirichlet)
    {
        AxisFactors f;
        f.fw.assign(static_cast<size_t>(padded), 1.0);
        f.ib.assign(static_cast<size_t>(padded), 0.0);
        f.bw.assign(static_cast<size_t>(padded), 0.0);
        f.z.assign(static_cast<size_t>(padded), 0.0);
        const double off = -rate / 2;
        std::vector<double> b(static_cast<size_t>(n));
        double gamma = 0.0;
        if (wrap)
        {
            for (int32_t i = 0; i < n; i++)
            {
                b[i] = 1 - 2 * off;
            }
            gamma = -b[0];
            const double alpha = off;
            const double beta = off;
            b[0] = b[0] - gamma;
            if (n > 1)
            {
                b[n - 1] = b[n - 1] - (alpha * beta) / gamma;
            }
        }
        else
        {
            const double endAdjust = dirichlet ? 1.0 : -1.0;
            for (int32_t i = 0; i < n; i++)
            {
                double weight = 2.0;
                if (i == 0)
                {
                    weight += endAdjust;
                }
                if (i == n - 1)
                {
                    weight += endAdjust;
                }
                b[i] = 1 - weight * off;
            }
        }
        std::vector<double> bp(static_cast<size_t>(n));
        bp[0] = b[0];
        f.fw[0] = 0.0;
        for (int32_t i = 1; i < n; i++)
        {
            const double w = off / bp[i - 1];
            bp[i] = b[i] - w * off;
            f.fw[i] = -w;
        }
        for (int32_t i = 0; i < n; i++)
        {
            f.ib[i] = 1.0 / bp[i];
            f.bw[i] = (i < n - 1) ? -off / bp[i] : 0.0;
        }
        if (wrap)
        {
            // z solves the adjusted system for u = (gamma, 0, ..., 0, alpha).
            std::vector<double> u(static_cast<size_t>(n), 0.0);
            u[0] = gamma;
            if (n > 1)
            {
                u[n - 1] = off;
            }
            std::vector<double> dp(static_cast<size_t>(n));
            double acc = 0.0;
            for (int32_t i = 0; i < n; i++)
            {
                acc = ((i == 0) ? 0.0 : f.fw[i]) * acc + u[i];
                dp[i] = acc;
            }
            acc = 0.0;
            for (int32_t i = n - 1; i >= 0; i--)
            {
                acc = f.bw[i] * acc + f.ib[i] * dp[i];
                f.z[i] = acc;
            }
            f.vEnd = off / gamma;
            const double vz = f.z[0] + f.z[n - 1] * f.vEnd;
            f.inv = 1.0 / (1 + vz);
        }
        return f;
    }

    struct Shape
    {
        int32_t cpt;
        int32_t s;
        int32_t threads;
    };

    Shape ChooseShape(int32_t X, int32_t Y)
    {
        Shape sh{};
        sh.cpt = 1;
        while (sh.cpt < 8 && (Y + sh.cpt - 1) / sh.cpt > 1024)
        {
            sh.cpt *= 2;
        }
        int32_t lanes = (Y + sh.cpt - 1) / sh.cpt;
        sh.threads = ((lanes + 63) / 64) * 64;
        // Largest strip height that keeps CPT*S <= 64 register values and still
        // gives at least one strip per CU; else the smallest.
        const int32_t maxS = std::min(16, 64 / sh.cpt);
        sh.s = 4;
        for (int32_t s = maxS; s >= 4; s /= 2)
        {
            if ((X + s - 1) / s >= 256)
            {
                sh.s = s;
                break;
            }
        }
        return sh;
    }

    using StripKernel = void (*)(AdiArgs, float *);

    template <int CPT, int S>
    StripKernel PickMode(int mode)
    {
        switch (mode)
        {
            case 0:
                return AdiStrip<CPT, S, 0>;
            case 1:
                return AdiStrip<CPT, S, 1>;
            default:
                return AdiStrip<CPT, S, 2>;
        }
    }

    StripKernel PickStrip(int cpt, int s, int mode)
    {
        switch (cpt * 100 + s)
        {
            case 104:
                return PickMode<1, 4>(mode);
            case 108:
                return PickMode<1, 8>(mode);
            case 116:
                return PickMode<1, 16>(mode);
            case 204:
                return PickMode<2, 4>(mode);
            case 208:
                return PickMode<2, 8>(mode);
            case 216:
                return PickMode<2, 16>(mode);
            case 404:
                return PickMode<4, 4>(mode);
            case 408:
                return PickMode<4, 8>(mode);
            case 416:
                return PickMode<4, 16>(mode);
            case 804:
                return PickMode<8, 4>(mode);
            case 808:
                return PickMode<8, 8>(mode);
            default:
                Fail(HALHIP_ERR_UNSUPPORTED, "adi2d: no kernel for the chosen strip shape");
        }
    }
}

struct halhip_adi2d
{
    halhip_context *ctx = nullptr;
    int32_t X = 0;
    int32_t Y = 0;
    Shape shape{};
    int32_t P = 0;
    DeviceArray<float> field;
    DeviceArray<float> tables;
    DeviceArray<float> red;
    DeviceArray<float> carry;
    DeviceArray<float> colInfo;
    DeviceArray<int64_t> srcIdx;
    DeviceArray<float> srcVal;
    int32_t nSources = 0;
    std::string tableKey;
    bool reducedValid = false; // red[] holds the reduction of the current field for tableKey
    AdiArgs args{};
    StreamClock clock;
    Profiler prof;
    GraphCache graphs;
    std::vector<float> hostBuf;
};

namespace
{
    size_t Cells(const halhip_adi2d *p)
    {
        return static_cast<size_t>(p->X) * static_cast<size_t>(p->Y);
    }

    void RequirePlan(const halhip_adi2d *p, const char *caller)
    {
        Require(p != nullptr, std::string(caller) + ": null plan");
        Bind(p->ctx);
    }

    // Builds and uploads the line factors and strip constants for a rate and
    // boundary set, when they differ from the resident ones.
    void PrepareTables(halhip_adi2d *p, double rate, int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        Require(std::isfinite(rate) && rate >= 0.0, "adi2d: diffusion rate must be finite and non-negative");
        EdgeValues e;
        e.From(bc, p->X, p->Y);
        std::string key;
        AppendKey(key, rate);
        AppendKey(key, wrapX);
        AppendKey(key, wrapY);
        e.AppendTo(key);
        if (key == p->tableKey)
        {
            return;
        }
        const int32_t S = p->shape.s;
        const int32_t cpt = p->shape.cpt;
        const int32_t T = p->shape.threads;
        const int32_t xPad = p->P * S;
        const int32_t yPad = T * cpt;
        const bool dir = e.dirichlet;
        AxisFactors fx = FactorLine(p->X, xPad, rate, wrapX != 0, dir);
        AxisFactors fy = FactorLine(p->Y, yPad, rate, wrapY != 0, dir);

        std::vector<float> host;
        auto put = [&](const std::vector<double> &v) -> size_t
        {
            size_t at = host.size();
            for (double d : v)
            {
                host.push_back(static_cast<float>(d));
            }
            return at;
        };
        auto putf = [&](const std::vector<float> &v) -> size_t
        {
            size_t at = host.size();
            host.insert(host.end(), v.begin(), v.end());
            return at;
        };
        std::vector<double> sA(static_cast<size_t>(p->P)), sC(static_cast<size_t>(p->P)), sE(static_cast<size_t>(p->P));
        for (int32_t q = 0; q < p->P; q++)
        {
            double A = 1.0;
            double C = 1.0;
            std::vector<double> dp(static_cast<size_t>(S));
            double acc = 1.0;
            for (int32_t r = 0; r < S; r++)
            {
                const double fw = fx.fw[static_cast<size_t>(q * S + r)];
                A *= fw;
                C *= fx.bw[static_cast<size_t>(q * S + r)];
                acc = fw * acc;
                dp[r] = acc;
            }
            acc = 0.0;
            for (int32_t r = S - 1; r >= 0; r--)
            {
                acc = fx.bw[static_cast<size_t>(q * S + r)] * acc + fx.ib[static_cast<size_t>(q * S + r)] * dp[r];
            }
            sA[q] = A;
            sC[q] = C;
            sE[q] = acc;
        }
        size_t oXfw = put(fx.fw), oXib = put(fx.ib), oXbw = put(fx.bw), oXz = put(fx.z);
        size_t oYfw = put(fy.fw), oYib = put(fy.ib), oYbw = put(fy.bw), oYz = put(fy.z);
        size_t oA = put(sA), oC = put(sC), oE = put(sE);
        size_t oXlo = putf(e.xlo), oXhi = putf(e.xhi), oYlo = putf(e.ylo), oYhi = putf(e.yhi);

        if (p->tables.size() < host.size())
        {
            p->tables.Allocate(p->ctx, host.size());
        }
        Check(hipMemcpyAsync(p->tables.get(), host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice,
                             p->clock.stream),
              "hipMemcpyAsync");
        Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");

        float *t = p->tables.get();
        AdiArgs &a = p->args;
        a.X = p->X;
        a.Y = p->Y;
        a.P = p->P;
        a.hr = static_cast<float>(rate / 2);
        a.r = static_cast<float>(rate);
        a.wrapX = wrapX ? 1 : 0;
        a.wrapY = wrapY ? 1 : 0;
        a.dirichlet = dir ? 1 : 0;
        a.vecOk = (p->Y % cpt == 0) ? 1 : 0;
        a.xfw = t + oXfw;
        a.xib = t + oXib;
        a.xbw = t + oXbw;
        a.xz = t + oXz;
        a.yfw = t + oYfw;
        a.yib = t + oYib;
        a.ybw = t + oYbw;
        a.yz = t + oYz;
        a.yvEnd = static_cast<float>(fy.vEnd);
        a.yInv = static_cast<float>(fy.inv);
        a.bcXlo = t + oXlo;
        a.bcXhi = t + oXhi;
        a.bcYlo = t + oYlo;
        a.bcYhi = t + oYhi;
        a.red = p->red.get();
        a.carry = p->carry.get();
        a.colInfo = p->colInfo.get();
        a.stripA = t + oA;
        a.stripC = t + oC;
        a.stripE = t + oE;
        a.xvEnd = static_cast<float>(fx.vEnd);
        a.xInv = static_cast<float>(fx.inv);
        a.xibLast = static_cast<float>(fx.ib[static_cast<size_t>(p->X - 1)]);
        a.xz0 = static_cast<float>(fx.z[0]);
        a.xzLast = static_cast<float>(fx.z[static_cast<size_t>(p->X - 1)]);
        p->tableKey = key;
        p->reducedValid = false;
        p->graphs.Clear();
    }

    void LaunchStrip(halhip_adi2d *p, int mode, const char *name)
    {
        StripKernel k = PickStrip(p->shape.cpt, p->shape.s, mode);
        p->prof.Begin(p->clock.stream, name);
        hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(p->P)), dim3(static_cast<unsigned>(p->shape.threads)), 0,
                           p->clock.stream, p->args, p->field.get());
        Check(hipGetLastError(), "AdiStrip launch");
        p->prof.End(p->clock.stream);
    }

    void LaunchScan(halhip_adi2d *p)
    {
        const int32_t G = std::min<int32_t>(kMaxWaves, p->P);
        const int32_t Q = (p->P + G - 1) / G;
        p->prof.Begin(p->clock.stream, "adi_strip_scan");
        hipLaunchKernelGGL(AdiStripScan, dim3(static_cast<unsigned>((p->Y + 63) / 64)), dim3(static_cast<unsigned>(64 * G)),
                           0, p->clock.stream, p->args, Q);
        Check(hipGetLastError(), "AdiStripScan launch");
        p->prof.End(p->clock.stream);
    }

    void LaunchSources(halhip_adi2d *p)
    {
        if (p->nSources == 0)
        {
            return;
        }
        p->prof.Begin(p->clock.stream, "adi_sources");
        hipLaunchKernelGGL(AdiAddSources, dim3(1), dim3(64), 0, p->clock.stream, p->field.get(), p->nSources,
                           p->srcIdx.get(), p->srcVal.get());
        Check(hipGetLastError(), "AdiAddSources launch");
        p->prof.End(p->clock.stream);
    }

    // One plain tick's launches (the reduction of the field is resident).
    void PlainTick(halhip_adi2d *p)
    {
        LaunchScan(p);
        LaunchStrip(p, 1, "adi_strip_tick");
    }

    void EnqueueTicks(halhip_adi2d *p, int32_t steps, double rate, double decay, int32_t wrapX, int32_t wrapY,
                      const halhip_bc *bc, bool withSources)
    {
        Require(steps >= 0, "halhip_adi2d_run_steps: negative steps");
        Require(std::isfinite(decay), "halhip_adi2d_run_steps: decay must be finite");
        PrepareTables(p, rate, wrapX, wrapY, bc);
        if (steps == 0)
        {
            return;
        }
        const bool plain = (decay == 0.0) && (!withSources || p->nSources == 0);
        if (plain)
        {
            if (!p->reducedValid)
            {
                LaunchStrip(p, 0, "adi_strip_reduce");
            }
            constexpr int32_t kChunk = 20;
            int32_t done = 0;
            if (!p->prof.On() && steps >= kChunk)
            {
                std::string key = p->tableKey;
                AppendKey(key, kChunk);
                hipGraphExec_t exec = p->graphs.Find(key);
                if (exec == nullptr)
                {
                    hipGraph_t graph;
                    Check(hipStreamBeginCapture(p->clock.stream, hipStreamCaptureModeThreadLocal),
                          "hipStreamBeginCapture");
                    for (int32_t k = 0; k < kChunk; k++)
                    {
                        PlainTick(p);
                    }
                    Check(hipStreamEndCapture(p->clock.stream, &graph), "hipStreamEndCapture");
                    Check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
                    (void)hipGraphDestroy(graph);
                    p->graphs.Insert(key, exec);
                }
                for (; done + kChunk <= steps; done += kChunk)
                {
                    Check(hipGraphLaunch(exec, p->clock.stream), "hipGraphLaunch");
                }
            }
            for (; done < steps; done++)
            {
                PlainTick(p);
            }
            p->reducedValid = true;
            return;
        }
        p->args.negDecay = static_cast<float>(-decay);
        for (int32_t t = 0; t < steps; t++)
        {
            LaunchStrip(p, 0, "adi_strip_reduce");
            LaunchScan(p);
            LaunchStrip(p, 2, "adi_strip_tick_decay");
            if (withSources)
            {
                LaunchSources(p);
            }
        }
        p->reducedValid = false;
    }
}

extern "C"
{
    int halhip_adi2d_create(halhip_context *ctx, int32_t xDim, int32_t yDim, halhip_adi2d **out)
    {
        halhip_adi2d *p = nullptr;
        int rc = Guard(
            [&]
            {
                Require(out != nullptr, "halhip_adi2d_create: null out");
                *out = nullptr;
                Require(ctx != nullptr, "halhip_adi2d_create: null context");
                Require(xDim > 0 && yDim > 0, "halhip_adi2d_create: dimensions must be positive");
                if (yDim > HALHIP_ADI_MAX_Y)
                {
                    Fail(HALHIP_ERR_UNSUPPORTED, "halhip_adi2d_create: yDim exceeds HALHIP_ADI_MAX_Y (8192)");
                }
                Require(static_cast<int64_t>(xDim) * yDim <= (int64_t(1) << 31),
                        "halhip_adi2d_create: grid larger than 2^31 cells");
                Bind(ctx);
                p = new halhip_adi2d();
                p->ctx = ctx;
                p->X = xDim;
                p->Y = yDim;
                p->shape = ChooseShape(xDim, yDim);
                p->P = (xDim + p->shape.s - 1) / p->shape.s;
                const size_t n = Cells(p);
                const size_t plane = static_cast<size_t>(p->P) * static_cast<size_t>(yDim);
                p->field.Allocate(ctx, n);
                p->red.Allocate(ctx, 2 * plane);
                p->carry.Allocate(ctx, 2 * plane);
                p->colInfo.Allocate(ctx, 3 * static_cast<size_t>(yDim));
                p->clock.Create();
                Check(hipMemsetAsync(p->field.get(), 0, n * sizeof(float), p->clock.stream), "hipMemset");
                Check(hipMemsetAsync(p->colInfo.get(), 0, 3 * static_cast<size_t>(yDim) * sizeof(float), p->clock.stream),
                      "hipMemset");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                *out = p;
            });
        if (rc != HALHIP_OK && p != nullptr)
        {
            p->clock.Destroy();
            delete p;
        }
        return rc;
    }

    int halhip_adi2d_destroy(halhip_adi2d *p)
    {
        return Guard(
            [&]
            {
                if (p == nullptr)
                {
                    return;
                }
                Bind(p->ctx);
                p->graphs.Clear();
                p->clock.Destroy();
                delete p;
            });
    }

    int halhip_adi2d_upload(halhip_adi2d *p, const float *field, int64_t n)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_upload");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_adi2d_upload");
                Require(field != nullptr, "halhip_adi2d_upload: null field");
                Check(hipMemcpyAsync(p->field.get(), field, Cells(p) * sizeof(float), hipMemcpyHostToDevice,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                p->reducedValid = false;
            });
    }

    int halhip_adi2d_download(halhip_adi2d *p, float *field, int64_t n)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_download");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_adi2d_download");
                Require(field != nullptr, "halhip_adi2d_download: null field");
                Check(hipMemcpyAsync(field, p->field.get(), Cells(p) * sizeof(float), hipMemcpyDeviceToHost,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
            });
    }

    int halhip_adi2d_set_sources(halhip_adi2d *p, int32_t count, const int64_t *index, const float *value)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_set_sources");
                Require(count >= 0 && (count == 0 || (index != nullptr && value != nullptr)),
                        "halhip_adi2d_set_sources: bad source arrays");
                const int64_t cells = static_cast<int64_t>(Cells(p));
                for (int32_t k = 0; k < count; k++)
                {
                    Require(index[k] >= 0 && index[k] < cells, "halhip_adi2d_set_sources: index out of range");
                }
                p->srcIdx.Release();
                p->srcVal.Release();
                p->nSources = count;
                if (count == 0)
                {
                    return;
                }
                p->srcIdx.Allocate(p->ctx, static_cast<size_t>(count));
                p->srcVal.Allocate(p->ctx, static_cast<size_t>(count));
                Check(hipMemcpy(p->srcIdx.get(), index, sizeof(int64_t) * count, hipMemcpyHostToDevice), "hipMemcpy");
                Check(hipMemcpy(p->srcVal.get(), value, sizeof(float) * count, hipMemcpyHostToDevice), "hipMemcpy");
            });
    }

    int halhip_adi2d_diffuse_f64(halhip_adi2d *p, const double *field, double *deltas, int64_t n, double diffRate,
                                 int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_diffuse_f64");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_adi2d_diffuse_f64");
                Require(field != nullptr && deltas != nullptr, "halhip_adi2d_diffuse_f64: null array");
                const size_t cells = Cells(p);
                p->hostBuf.resize(cells);
                for (size_t k = 0; k < cells; k++)
                {
                    p->hostBuf[k] = static_cast<float>(field[k]);
                }
                Check(hipMemcpyAsync(p->field.get(), p->hostBuf.data(), cells * sizeof(float), hipMemcpyHostToDevice,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                p->reducedValid = false;
                EnqueueTicks(p, 1, diffRate, 0.0, wrapX, wrapY, bc, false);
                Check(hipMemcpyAsync(p->hostBuf.data(), p->field.get(), cells * sizeof(float), hipMemcpyDeviceToHost,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                p->prof.Collect();
                for (size_t k = 0; k < cells; k++)
                {
                    deltas[k] += static_cast<double>(p->hostBuf[k]) - field[k];
                }
            });
    }

    int halhip_adi2d_run_steps_async(halhip_adi2d *p, int32_t steps, double diffRate, double decay, int32_t wrapX,
                                     int32_t wrapY, const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_run_steps");
                p->clock.Mark();
                EnqueueTicks(p, steps, diffRate, decay, wrapX, wrapY, bc, true);
            });
    }

    int halhip_adi2d_sync(halhip_adi2d *p, float *elapsedMs)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_sync");
                p->clock.Sync(elapsedMs);
                p->prof.Collect();
            });
    }

    int halhip_adi2d_run_steps(halhip_adi2d *p, int32_t steps, double diffRate, double decay, int32_t wrapX,
                               int32_t wrapY, const halhip_bc *bc)
    {
        int rc = halhip_adi2d_run_steps_async(p, steps, diffRate, decay, wrapX, wrapY, bc);
        if (rc != HALHIP_OK)
        {
            return rc;
        }
        return halhip_adi2d_sync(p, nullptr);
    }

    int halhip_adi2d_set_profiling(halhip_adi2d *p, int32_t on)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_set_profiling");
                p->prof.Reset();
                p->prof.SetOn(on != 0);
            });
    }

    int halhip_adi2d_profile(halhip_adi2d *p, int32_t maxKernels, int32_t *nKernels, char *names, int32_t namesLen,
                             double *totalMs, int64_t *launches)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_profile");
                p->prof.Export(maxKernels, nKernels, names, namesLen, totalMs, launches);
            });
    }
}
