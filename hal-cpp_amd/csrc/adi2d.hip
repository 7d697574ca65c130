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

#include "adi_device.hpp"

#include <algorithm>
#include <cmath>
#include <string>

using namespace halhip;

namespace
{
    struct Shape
    {
        int32_t cpt;
        int32_t s;
        int32_t threads;
        bool vec;
    };

    // Scan of the strip reductions for any strip count (reloads from memory).
    // Block: 64 columns (lanes) x G wavefronts; wavefront g owns strips
    // [g*Q, (g+1)*Q).  Writes the carry-in values of every strip and, per
    // column, the periodic Sherman-Morrison factor and the wrapped rows of u*.
    __global__ __launch_bounds__(1024) void AdiStripScanGeneric(AdiArgs a, int32_t Q)
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

    // d += v - u
    __global__ __launch_bounds__(256) void AdiAccumulateDelta(float *__restrict__ d, const float *__restrict__ v,
                                                              const float *__restrict__ u, int64_t n)
    {
        for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        {
            d[i] += v[i] - u[i];
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

    // ---- rows longer than one workgroup --------------------------------------------
    //
    // A row of yDim > HALHIP_ADI_MAX_Y columns is cut into G segments of W =
    // 4 x threads columns, and strip p's segment g is one workgroup.  Sweep A
    // is unchanged (its lines are columns).  Sweep B is linear in its two
    // carries, so each segment solves with zero carries (x0) and the segments
    // are joined exactly afterwards:
    //     d'_j = d'0_j + P_j cf,   x_j = x0_j + Q_j xb + R_j cf   (yS <= j < yE)
    // with cf = d' of column yS-1, xb = x of column yE, and the host tables
    //     P_j = prod_{yS<=k<=j} fw_k,  Q_j = prod_{j<=k<yE} bw_k,
    //     R_j = bw_j R_{j+1} + ib_j P_j  (R_yE = 0).
    // The carries follow from two numbers per row and segment: x0 at the
    // first column and x0 at the last (d'0 there is x0 * b').  A periodic y
    // axis applies its Sherman-Morrison closure to the joined row.
    // A tick: AdiSegReduce (sweep A strip reductions), the cross-strip scan,
    // AdiSegSolve (sweep A finish, sweep B right-hand side and zero-carry
    // solve into x0), AdiSegScan (carries per row) and AdiSegFix (x0 plus the
    // carry responses, with decay, written over the field).
    struct SegArgs
    {
        int32_t G;         // segments per row
        int32_t W;         // columns per segment
        float *x0;         // [X][Y] zero-carry sweep-B solution
        float *first;      // [G][X] x0 at each segment's first column
        float *last;       // [G][X] x0 at each segment's last column
        float *cf;         // [G][X] forward carry into each segment
        float *xb;         // [G][X] backward carry into each segment
        float *factor;     // [X] Sherman-Morrison factor of a periodic y axis
        const float *q;    // [G*W] Q_j
        const float *rr;   // [G*W] R_j
        const float *pEnd; // [G] P at each segment's last column
        const float *bEnd; // [G] b' (1/ib) at each segment's last column
        float rLast;       // R at column Y-1
    };

    constexpr int kSegCpt = 4;

    // Sweep A right-hand side of the segment [yS, yE) of every strip row, as
    // SweepARhs forms it for a whole row.  The neighbours of the segment's
    // first and last columns are read from the field (the adjacent segments)
    // or, at the grid's ends, are the y boundary's ghost values.
    template <int CPT, int S, bool VEC>
    __device__ __forceinline__ void SegSweepARhs(const AdiArgs &a, const float *__restrict__ u, int s0, int yS,
                                                 int yE, int tid, float (&v)[S][CPT], const StripLds &L)
    {
        const int lane = tid & 63;
        const int wave = tid >> 6;
        const int nWaves = blockDim.x >> 6;
        const int y0 = yS + tid * CPT;
        const int lastT = (yE - 1 - yS) / CPT;
        const int lastC = (yE - 1 - yS) % CPT;
        if (lane == 0)
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                L.left[wave * S + r] = v[r][0];
            }
        }
        if (lane == 63)
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                L.right[wave * S + r] = v[r][CPT - 1];
            }
        }
        __syncthreads();
        const bool dirY = a.dirichlet && !a.wrapY;
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const int x = s0 + r;
            float left = Dpp<0x138>(0.0f, v[r][CPT - 1]);
            float right = Dpp<0x130>(0.0f, v[r][0]);
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
                left = L.right[(wave - 1) * S + r];
            }
            if (lane == 63 && wave + 1 < nWaves)
            {
                right = L.left[(wave + 1) * S + r];
            }
            const float *row = u + static_cast<size_t>(x) * a.Y;
            if (tid == 0)
            {
                const float c0 = v[r][0];
                left = (yS > 0) ? row[yS - 1] : (a.wrapY ? row[a.Y - 1] : (dirY ? 2.0f * a.bcYlo[x] - c0 : c0));
            }
            float ghostHi = 0.0f;
            if (tid == lastT)
            {
                float cl = 0.0f;
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    cl = (c == lastC) ? v[r][c] : cl;
                }
                ghostHi = (yE < a.Y) ? row[yE] : (a.wrapY ? row[0] : (dirY ? 2.0f * a.bcYhi[x] - cl : cl));
            }
            const bool srcRow = a.dirichlet && !a.wrapX && (x == 0 || x == a.X - 1);
            float prev = left;
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                const float cur = v[r][c];
                float nxt = (c + 1 < CPT) ? v[r][c + 1] : right;
                if (tid == lastT && c == lastC)
                {
                    nxt = ghostHi;
                }
                float rhs = fmaf((nxt - cur) + (prev - cur), a.hr, cur);
                if (srcRow)
                {
                    float src = 0.0f;
                    if (x == 0)
                    {
                        src += a.bcXlo[min(y0 + c, a.Y - 1)];
                    }
                    if (x == a.X - 1)
                    {
                        src += a.bcXhi[min(y0 + c, a.Y - 1)];
                    }
                    rhs = fmaf(src, a.r, rhs);
                }
                prev = cur;
                v[r][c] = (y0 + c < a.Y) ? rhs : 0.0f;
            }
        }
    }

    // Workgroup b runs strip b / G, segment b % G (neighbouring workgroups
    // share rows).
    struct SegTile
    {
        int p, g, s0, yS, yE, y0;
    };

    template <int S>
    __device__ __forceinline__ SegTile SegTileOf(const AdiArgs &a, const SegArgs &sg)
    {
        SegTile t;
        t.p = static_cast<int>(blockIdx.x) / sg.G;
        t.g = static_cast<int>(blockIdx.x) - t.p * sg.G;
        t.s0 = t.p * S;
        t.yS = t.g * sg.W;
        t.yE = min(a.Y, t.yS + sg.W);
        t.y0 = t.yS + static_cast<int>(threadIdx.x) * kSegCpt;
        return t;
    }

    // Sweep A reductions of the current field (the strip kernel's MODE 0).
    template <int S, bool VEC>
    __global__ __launch_bounds__(1024) void AdiSegReduce(AdiArgs a, SegArgs sg, const float *__restrict__ u)
    {
        __shared__ float sLeft[kMaxWaves * S];
        __shared__ float sRight[kMaxWaves * S];
        const StripLds L{sLeft, sRight, nullptr, nullptr, nullptr, nullptr, nullptr};
        const SegTile t = SegTileOf<S>(a, sg);
        float v[S][kSegCpt];
        LoadStripRows<kSegCpt, S, VEC>(a, u, t.s0, t.y0, v);
        SegSweepARhs<kSegCpt, S, VEC>(a, u, t.s0, t.yS, t.yE, static_cast<int>(threadIdx.x), v, L);
        ReduceA<kSegCpt, S, VEC>(a, t.p, t.s0, t.y0, v);
    }

    // Sweep A finished from the scan's carries, sweep B's right-hand side and
    // its zero-carry solve on the segment, written to x0 with the segment's
    // first and last values.
    template <int S, bool VEC>
    __global__ __launch_bounds__(1024) void AdiSegSolve(AdiArgs a, SegArgs sg, const float *__restrict__ u)
    {
        __shared__ float sLeft[kMaxWaves * S];
        __shared__ float sRight[kMaxWaves * S];
        __shared__ float sTot[kMaxWaves * S];
        __shared__ float sAw[kMaxWaves];
        const StripLds L{sLeft, sRight, nullptr, nullptr, sTot, sAw, nullptr};
        const int tid = static_cast<int>(threadIdx.x);
        const SegTile t = SegTileOf<S>(a, sg);
        float v[S][kSegCpt];
        LoadStripRows<kSegCpt, S, VEC>(a, u, t.s0, t.y0, v);
        SegSweepARhs<kSegCpt, S, VEC>(a, u, t.s0, t.yS, t.yE, tid, v, L);
        float up[kSegCpt];
        float dn[kSegCpt];
        FinishA<kSegCpt, S, VEC>(a, t.p, t.s0, t.y0, v, up, dn);
        // SweepBRhs takes the grid-wide thread index of the columns (its
        // Dirichlet terms sit on the grid's first and last columns).
        SweepBRhs<kSegCpt, S, VEC>(a, t.s0, t.y0, t.y0 / kSegCpt, v, up, dn);
        SolveB<kSegCpt, S, true>(a, t.y0, tid, v, L);
        const int lastT = (t.yE - 1 - t.yS) / kSegCpt;
        const int lastC = (t.yE - 1 - t.yS) % kSegCpt;
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const int x = t.s0 + r;
            if (x >= a.X)
            {
                break;
            }
            StoreRow<kSegCpt, VEC>(sg.x0 + static_cast<size_t>(x) * a.Y, t.y0, a.Y, v[r]);
            const size_t k = static_cast<size_t>(t.g) * a.X + x;
            if (tid == 0)
            {
                sg.first[k] = v[r][0];
            }
            if (tid == lastT)
            {
                float cl = 0.0f;
#pragma unroll
                for (int c = 0; c < kSegCpt; c++)
                {
                    cl = (c == lastC) ? v[r][c] : cl;
                }
                sg.last[k] = cl;
            }
        }
    }

    // Carries of every segment of row x: forward from the left (d' of the
    // column before each segment), backward from the right (x of the column
    // after it), and the periodic closure's factor from the row's end values.
    __global__ __launch_bounds__(256) void AdiSegScan(AdiArgs a, SegArgs sg)
    {
        const int x = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
        if (x >= a.X)
        {
            return;
        }
        float cf = 0.0f;
        float cfLast = 0.0f;
        for (int g = 0; g < sg.G; g++)
        {
            const size_t k = static_cast<size_t>(g) * a.X + x;
            sg.cf[k] = cf;
            cfLast = cf;
            cf = fmaf(sg.pEnd[g], cf, sg.last[k] * sg.bEnd[g]);
        }
        float xb = 0.0f;
        for (int g = sg.G - 1; g >= 0; g--)
        {
            const size_t k = static_cast<size_t>(g) * a.X + x;
            const size_t j = static_cast<size_t>(g) * sg.W;
            sg.xb[k] = xb;
            xb = fmaf(sg.q[j], xb, fmaf(sg.rr[j], sg.cf[k], sg.first[k]));
        }
        if (a.wrapY)
        {
            // xb is now x of column 0; column Y-1 has no backward carry.
            const float xLast = fmaf(sg.rLast, cfLast, sg.last[static_cast<size_t>(sg.G - 1) * a.X + x]);
            sg.factor[x] = (xb + xLast * a.yvEnd) * a.yInv;
        }
    }

    // u' = x0 + Q xb + R cf (- factor z on a periodic y axis) (- k u with
    // decay), in place over the field.  Block row blockIdx.x is grid row x;
    // each thread takes V consecutive cells (V = 4 when yDim is a multiple
    // of 4: a group then never straddles a segment), blocks of 256 threads
    // stride along the row.
    template <int V, bool DECAY>
    __global__ __launch_bounds__(256) void AdiSegFix(AdiArgs a, SegArgs sg, float *__restrict__ u)
    {
        const int x = static_cast<int>(blockIdx.x);
        const size_t rowBase = static_cast<size_t>(x) * a.Y;
        const float fac = a.wrapY ? sg.factor[x] : 0.0f;
        for (int y = (static_cast<int>(blockIdx.y) * 256 + static_cast<int>(threadIdx.x)) * V; y < a.Y;
             y += static_cast<int>(gridDim.y) * 256 * V)
        {
            const size_t k = static_cast<size_t>(y / sg.W) * a.X + x;
            const float xb = sg.xb[k];
            const float cf = sg.cf[k];
            const size_t i = rowBase + static_cast<size_t>(y);
            float v[V];
            float old[V];
            // The field and x0 move as vectors; the tables (at any float
            // offset of the plan's table buffer) load per element.
            if constexpr (V == 4)
            {
                const float4 tv = *reinterpret_cast<const float4 *>(sg.x0 + i);
                v[0] = tv.x, v[1] = tv.y, v[2] = tv.z, v[3] = tv.w;
                if constexpr (DECAY)
                {
                    const float4 to = *reinterpret_cast<const float4 *>(u + i);
                    old[0] = to.x, old[1] = to.y, old[2] = to.z, old[3] = to.w;
                }
            }
            else
            {
                v[0] = sg.x0[i];
                if constexpr (DECAY)
                {
                    old[0] = u[i];
                }
            }
#pragma unroll
            for (int c = 0; c < V; c++)
            {
                float r = fmaf(sg.q[y + c], xb, fmaf(sg.rr[y + c], cf, v[c]));
                r = fmaf(-fac, a.yz[y + c], r);
                if constexpr (DECAY)
                {
                    r = fmaf(old[c], a.negDecay, r);
                }
                v[c] = r;
            }
            if constexpr (V == 4)
            {
                *reinterpret_cast<float4 *>(u + i) = make_float4(v[0], v[1], v[2], v[3]);
            }
            else
            {
                u[i] = v[0];
            }
        }
    }

    bool ShapeCompiled(int32_t cpt, int32_t s)
    {
        return (cpt == 1 || cpt == 2 || cpt == 4) ? (s == 4 || s == 8 || s == 16) : (cpt == 8 && (s == 4 || s == 8));
    }

    // Strip shape with at most maxLanes threads per workgroup.
    Shape ChooseShape(int32_t X, int32_t Y, int32_t maxLanes = 1024)
    {
        Shape sh{};
        sh.cpt = 1;
        while (sh.cpt < 8 && (Y + sh.cpt - 1) / sh.cpt > maxLanes)
        {
            sh.cpt *= 2;
        }
        int32_t lanes = (Y + sh.cpt - 1) / sh.cpt;
        sh.threads = ((lanes + 63) / 64) * 64;
        sh.vec = (Y % sh.cpt) == 0;
        // Largest strip height (at most 8: measured fastest at 4096^2, where
        // 16-row strips spill registers) that keeps CPT*S <= 64 register values
        // and still gives at least one strip per CU; else the smallest.
        const int32_t maxS = std::min(8, 64 / sh.cpt);
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

    template <int CPT, int S, bool VEC>
    StripKernel PickMode(int mode)
    {
        switch (mode)
        {
            case 0:
                return AdiStrip<CPT, S, 0, VEC>;
            case 1:
                return AdiStrip<CPT, S, 1, VEC>;
            default:
                return AdiStrip<CPT, S, 2, VEC>;
        }
    }

    template <int CPT, int S>
    StripKernel PickVec(bool vec, int mode)
    {
        return vec ? PickMode<CPT, S, true>(mode) : PickMode<CPT, S, false>(mode);
    }

    StripKernel PickStrip(int cpt, int s, bool vec, int mode)
    {
        switch (cpt * 100 + s)
        {
            case 104:
                return PickVec<1, 4>(vec, mode);
            case 108:
                return PickVec<1, 8>(vec, mode);
            case 116:
                return PickVec<1, 16>(vec, mode);
            case 204:
                return PickVec<2, 4>(vec, mode);
            case 208:
                return PickVec<2, 8>(vec, mode);
            case 216:
                return PickVec<2, 16>(vec, mode);
            case 404:
                return PickVec<4, 4>(vec, mode);
            case 408:
                return PickVec<4, 8>(vec, mode);
            case 416:
                return PickVec<4, 16>(vec, mode);
            case 804:
                return PickVec<8, 4>(vec, mode);
            case 808:
                return PickVec<8, 8>(vec, mode);
            default:
                Fail(HALHIP_ERR_UNSUPPORTED, "adi2d: no kernel for the chosen strip shape");
        }
    }
}

namespace
{
    using LdsKernel = void (*)(AdiArgs, float *, int32_t, RowSolve);

    template <int CPT, int S, int E, int T, bool DECAY>
    LdsKernel PickLdsWindow(int32_t K, bool px)
    {
        if (K == 0)
        {
            return AdiStripLds<CPT, S, E, 0, T, DECAY>;
        }
        if (px)
        {
            return (K <= 4) ? AdiStripLds<CPT, S, E, 4, T, DECAY, true>
                            : AdiStripLds<CPT, S, E, kMaxWindow, T, DECAY, true>;
        }
        return (K <= 4) ? AdiStripLds<CPT, S, E, 4, T, DECAY> : AdiStripLds<CPT, S, E, kMaxWindow, T, DECAY>;
    }

    template <int CPT, int S, int E>
    LdsKernel PickLdsThreads(int threads, int32_t K, bool decay, bool px)
    {
        if (decay)
        {
            return threads == 1024 ? PickLdsWindow<CPT, S, E, 1024, true>(K, px)
                                   : (threads == 512 ? PickLdsWindow<CPT, S, E, 512, true>(K, px) : nullptr);
        }
        return threads == 1024 ? PickLdsWindow<CPT, S, E, 1024, false>(K, px)
                               : (threads == 512 ? PickLdsWindow<CPT, S, E, 512, false>(K, px) : nullptr);
    }

    // Kernel of the LDS row solve for a strip shape, E and workgroup size
    // (512 or 1024) with windowed carries of size K (0: carries from the scan
    // kernel), plain or with decay and sources, for a periodic x axis (px,
    // windowed only) or not, or null when that combination is not compiled.
    LdsKernel PickLds(int cpt, int s, int e, int32_t window, int threads, bool decay = false, bool px = false)
    {
        switch (cpt * 10000 + s * 100 + e)
        {
            case 40832:
                return PickLdsThreads<4, 8, 32>(threads, window, decay, px);
            case 80432:
                return PickLdsThreads<8, 4, 32>(threads, window, decay, px);
            case 20816:
                return PickLdsThreads<2, 8, 16>(threads, window, decay, px);
            case 40416:
                return PickLdsThreads<4, 4, 16>(threads, window, decay, px);
            case 10808:
                return PickLdsThreads<1, 8, 8>(threads, window, decay, px);
            case 20408:
                return PickLdsThreads<2, 4, 8>(threads, window, decay, px);
            case 10404:
                return PickLdsThreads<1, 4, 4>(threads, window, decay, px);
            default:
                return nullptr;
        }
    }

    // Elements per lane E of the LDS row solve for this shape (0: the shape
    // does not fit it).  Needs whole 16-byte rows, yDim = threads * CPT with a
    // whole number W of wavefronts per row, and a compiled (CPT, S, E).
    int32_t LdsElements(const Shape &sh, int32_t Y)
    {
        if (!sh.vec || Y % 64 != 0 || sh.threads * sh.cpt != Y)
        {
            return 0;
        }
        const int32_t waves = sh.threads / 64;
        if (waves % sh.s != 0)
        {
            return 0;
        }
        const int32_t W = waves / sh.s;
        const int32_t e = Y / (64 * W);
        if (64 * W * e != Y || PickLds(sh.cpt, sh.s, e, 0, sh.threads) == nullptr)
        {
            return 0;
        }
        return e;
    }

    // The automatic strip shape.  Up to yDim 2048, 512-thread workgroups
    // with more columns per thread are faster when the LDS row solve covers
    // them (measured 1024^2: 2 x 4 at 10.4 us/tick vs 1 x 4 at 11.4;
    // 2048^2: 4 x 8 at 17.7 vs 2 x 8 at 18.6); otherwise the widest
    // workgroup (fewest columns per thread) is used.
    Shape AutoShape(int32_t X, int32_t Y)
    {
        if (Y <= 2048)
        {
            const Shape narrow = ChooseShape(X, Y, 512);
            if (LdsElements(narrow, Y) > 0)
            {
                return narrow;
            }
        }
        return ChooseShape(X, Y);
    }

    // Constants of the LDS row solve (RowSolve) for lines of n = 64 W E
    // elements, computed in FP64.  End diagonals follow ADISolveRow
    // (ADIequations.h:225-270): 1 + r/2 (zero flux) or 1 + 3r/2 (Dirichlet).
    struct RowTables
    {
        RowSolve solve{};
        std::vector<double> data;
        size_t oPw = 0, oLane = 0, oChunk = 0, oPowE = 0;
    };

    // periodic: the cyclic line (TDMAperiodic's matrix, include/Hal/Pde/TdmaSolver.h:68-120):
    // diagonal 1 + r everywhere and corners -r/2, i.e. T + U C U^T with
    // C = [[1 + r - beta, -r/2], [-r/2, 0]]; otherwise C is diagonal with the
    // end coefficients of ADISolveRow.
    RowTables BuildRowTables(int32_t n, int32_t E, int32_t W, double rate, bool dirichlet, bool periodic)
    {
        RowTables t;
        const double hr = rate / 2;
        const double beta = ((1 + rate) + std::sqrt(1 + 2 * rate)) / 2;
        const double phi = hr / beta;
        const double ib = 1 / beta;
        const double bEnd = dirichlet ? 1 + 1.5 * rate : 1 + 0.5 * rate;
        double c00 = bEnd - beta;
        double c01 = 0.0;
        double c11 = bEnd - (1 + rate);
        if (periodic)
        {
            c00 = (1 + rate) - beta;
            c01 = -hr;
            c11 = 0.0;
        }
        const double c10 = c01;
        auto pw = [&](double e) { return std::pow(phi, e); };
        // G = U^T T^-1 U (symmetric), Kmat = (I + C G)^-1 C
        const double g00 = ib * (1 - pw(2.0 * n)) / (1 - phi * phi);
        const double g01 = ib * pw(n - 1.0);
        const double g11 = ib;
        const double m00 = 1 + c00 * g00 + c01 * g01, m01 = c00 * g01 + c01 * g11;
        const double m10 = c10 * g00 + c11 * g01, m11 = 1 + c10 * g01 + c11 * g11;
        const double det = m00 * m11 - m01 * m10;
        const double k00 = (m11 * c00 - m01 * c10) / det, k01 = (m11 * c01 - m01 * c11) / det;
        const double k10 = (m00 * c10 - m10 * c00) / det, k11 = (m00 * c11 - m10 * c01) / det;
        // (alpha, gamma) = Kmat ib (X_0, X_{n-1}); h = alpha / (1 - phi^2);
        // t = gamma - h phi^(n+1)
        const double q00 = k00 * ib / (1 - phi * phi), q01 = k01 * ib / (1 - phi * phi);
        const double q10 = k10 * ib - q00 * pw(n + 1.0), q11 = k11 * ib - q01 * pw(n + 1.0);
        RowSolve &r = t.solve;
        r.phi = static_cast<float>(phi);
        r.ib = static_cast<float>(ib);
        r.phiWave = static_cast<float>(pw(64.0 * E));
        r.q00 = static_cast<float>(q00);
        r.q01 = static_cast<float>(q01);
        r.q10 = static_cast<float>(q10);
        r.q11 = static_cast<float>(q11);
        r.W = W;
        t.oPw = t.data.size();
        for (int k = 0; k < 6; k++)
        {
            t.data.push_back(pw(static_cast<double>(E) * (1 << k)));
        }
        t.oLane = t.data.size();
        for (int l = 0; l < 64; l++)
        {
            t.data.push_back(pw(static_cast<double>(E) * (l % 16 + 1)));
        }
        for (int l = 0; l < 64; l++)
        {
            t.data.push_back(pw(static_cast<double>(E) * (l % 32 + 1)));
        }
        for (int l = 0; l < 64; l++)
        {
            t.data.push_back(pw(static_cast<double>(E) * l));
        }
        for (int l = 0; l < 64; l++)
        {
            t.data.push_back(pw(static_cast<double>(E) * (63 - l)));
        }
        for (int l = 0; l < 64; l++)
        {
            t.data.push_back(pw(static_cast<double>(E) * (16 - l % 16)));
        }
        t.oChunk = t.data.size();
        for (int k = 0; k < 64 * W; k++)
        {
            t.data.push_back(pw(static_cast<double>(E) * k));
        }
        for (int k = 0; k < 64 * W; k++)
        {
            t.data.push_back(pw(static_cast<double>(n) - static_cast<double>(E) * k - E));
        }
        t.oPowE = t.data.size();
        for (int j = 0; j < E; j++)
        {
            t.data.push_back(pw(j));
        }
        return t;
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
    // The same sources sorted by strip for the LDS decay kernel (stable, so
    // list order within a strip): offsets [P+1], (x, y) pairs and values,
    // built for strip height srcStripS (0: not built).
    std::vector<int64_t> srcHostIdx;
    std::vector<float> srcHostVal;
    DeviceArray<int32_t> srcStrip;
    DeviceArray<int32_t> srcXY;
    DeviceArray<float> srcStripVal;
    int32_t srcStripS = 0;
    std::string tableKey;
    bool reducedValid = false; // red half redCur holds the reduction of the current field for tableKey
    int32_t redCur = 0;        // which half of red[] holds the current reductions
    int32_t wantRow = HALHIP_ADI_AUTO;
    int32_t wantCarry = HALHIP_ADI_AUTO;
    int32_t useRow = HALHIP_ADI_ROWS_WORKGROUP_SCAN;
    int32_t useCarry = HALHIP_ADI_CARRIES_SCAN;
    int32_t window = 0;        // K of the windowed carries
    int32_t ldsE = 0;          // elements per lane of the LDS row solve
    RowSolve rs{};
    AdiArgs args{};
    StreamClock clock;
    Profiler prof;
    GraphCache graphs;
    bool prepareOnly = false;   // inside halhip_adi2d_prepare_steps
    // Rows cut into segments across workgroups (segG > 0): yDim above
    // HALHIP_ADI_MAX_Y, or forced with HALHIP_ADI_ROWS_SEGMENTED.
    int32_t segG = 0;
    DeviceArray<float> segX0;
    DeviceArray<float> segBuf;
    SegArgs seg{};
    HostStaging<float> hostBuf; // pinned staging of the per-call FP64 paths
    ChunkedTransfer chunks;
    // The per-call DiffuseADI's field (allocated on its first call): the
    // call swaps it in for the tick, so the resident field survives.
    DeviceArray<float> callField;
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

    // Strips whose reductions a periodic x axis's column values need: the
    // smallest k such that the terms dropped from u*(0, y) (weight
    // C_0 .. C_{k-1}) and from the forward carry out of the last strip
    // (weight A_{P-k} .. A_{P-1}) are at most 2^-28; P when there is none.
    // Unlike WindowSize, these products start at the grid's ends, where the
    // strip multipliers are not zero.
    int32_t PeriodicWindowSize(const std::vector<double> &A, const std::vector<double> &C)
    {
        const double tau = std::ldexp(1.0, -28);
        const int32_t P = static_cast<int32_t>(A.size());
        double prodC = 1.0;
        double prodA = 1.0;
        for (int32_t k = 1; k < P; k++)
        {
            prodC *= std::fabs(C[static_cast<size_t>(k - 1)]);
            prodA *= std::fabs(A[static_cast<size_t>(P - k)]);
            if (prodC <= tau && prodA <= tau)
            {
                return k;
            }
        }
        return P;
    }

    // (Re)allocates the strip reduction / carry buffers for the plan's shape.
    void AllocateStrips(halhip_adi2d *p)
    {
        p->P = (p->X + p->shape.s - 1) / p->shape.s;
        const size_t plane = static_cast<size_t>(p->P) * static_cast<size_t>(p->Y);
        p->red.Allocate(p->ctx, 4 * plane); // two halves of (B, D) for windowed ticks
        p->redCur = 0;
        p->carry.Allocate(p->ctx, 2 * plane);
        p->tableKey.clear();
        p->reducedValid = false;
        p->graphs.Clear();
    }

    // Cuts the rows into segments of 4 x T columns (T = 1024 threads above
    // HALHIP_ADI_MAX_Y; T = 64 when forced on shorter rows, so that they
    // still span several segments), or restores whole-row strips (on =
    // false).  The caller reallocates the strip buffers.
    void ConfigureSegments(halhip_adi2d *p, bool on)
    {
        if (!on)
        {
            p->segG = 0;
            p->segX0.Release();
            p->segBuf.Release();
            p->seg = SegArgs{};
            p->shape = AutoShape(p->X, p->Y);
            p->useRow = HALHIP_ADI_ROWS_WORKGROUP_SCAN;
            return;
        }
        const int32_t T = (p->Y > HALHIP_ADI_MAX_Y) ? 1024 : 64;
        const int32_t W = T * kSegCpt;
        const int32_t G = (p->Y + W - 1) / W;
        p->segG = G;
        p->shape.cpt = kSegCpt;
        p->shape.threads = T;
        p->shape.vec = (p->Y % kSegCpt) == 0;
        p->shape.s = (static_cast<int64_t>((p->X + 7) / 8) * G >= 512) ? 8 : 4;
        const size_t X = static_cast<size_t>(p->X);
        p->segX0.Release();
        p->segBuf.Release();
        p->segX0.Allocate(p->ctx, Cells(p));
        p->segBuf.Allocate(p->ctx, 4 * static_cast<size_t>(G) * X + X);
        SegArgs &g = p->seg;
        g.G = G;
        g.W = W;
        g.x0 = p->segX0.get();
        g.first = p->segBuf.get();
        g.last = g.first + static_cast<size_t>(G) * X;
        g.cf = g.last + static_cast<size_t>(G) * X;
        g.xb = g.cf + static_cast<size_t>(G) * X;
        g.factor = g.xb + static_cast<size_t>(G) * X;
        p->useRow = HALHIP_ADI_ROWS_SEGMENTED;
        p->useCarry = HALHIP_ADI_CARRIES_SCAN;
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
        AppendKey(key, p->wantRow);
        AppendKey(key, p->wantCarry);
        e.AppendTo(key);
        if (key == p->tableKey)
        {
            return;
        }
        const int32_t S = p->shape.s;
        const int32_t cpt = p->shape.cpt;
        const int32_t T = p->shape.threads;
        const int32_t xPad = p->P * S;
        const int32_t yPad = (p->segG > 0) ? p->segG * T * cpt : T * cpt;
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
        // Reduction weights: B = sum_j wB_j rhs_j with wB_j = prod_{m>j} fw_m;
        // D = sum_j wD_j rhs_j with wD_j = sum_{i>=j} (prod_{m<i} bw_m) ib_i
        // (prod_{j<m<=i} fw_m), indices within the strip.
        std::vector<double> redW(static_cast<size_t>(p->P) * 2 * S);
        for (int32_t q = 0; q < p->P; q++)
        {
            const size_t o = static_cast<size_t>(q) * S;
            for (int32_t j = 0; j < S; j++)
            {
                double wb = 1.0;
                for (int32_t m = j + 1; m < S; m++)
                {
                    wb *= fx.fw[o + m];
                }
                double wd = 0.0;
                double bwProd = 1.0; // prod_{m<i} bw_m
                for (int32_t i = 0; i < S; i++)
                {
                    if (i >= j)
                    {
                        double fwProd = 1.0;
                        for (int32_t m = j + 1; m <= i; m++)
                        {
                            fwProd *= fx.fw[o + m];
                        }
                        wd += bwProd * fx.ib[o + i] * fwProd;
                    }
                    bwProd *= fx.bw[o + i];
                }
                redW[2 * o + j] = wb;
                redW[2 * o + S + j] = wd;
            }
        }
        size_t oXfw = put(fx.fw), oXib = put(fx.ib), oXbw = put(fx.bw), oXz = put(fx.z);
        size_t oYfw = put(fy.fw), oYib = put(fy.ib), oYbw = put(fy.bw), oYz = put(fy.z);
        size_t oA = put(sA), oC = put(sC), oE = put(sE), oW = put(redW);
        size_t oXlo = putf(e.xlo), oXhi = putf(e.xhi), oYlo = putf(e.ylo), oYhi = putf(e.yhi);

        // Segment responses to the carries (SegArgs), in FP64 from the
        // sweep-B factors.
        size_t oSq = 0, oSr = 0, oSp = 0, oSb = 0;
        double segRLast = 0.0;
        if (p->segG > 0)
        {
            const int32_t W = T * cpt;
            std::vector<double> q(static_cast<size_t>(yPad), 0.0), rr(static_cast<size_t>(yPad), 0.0);
            std::vector<double> pEnd(static_cast<size_t>(p->segG)), bEnd(static_cast<size_t>(p->segG));
            std::vector<double> prod(static_cast<size_t>(W));
            for (int32_t g = 0; g < p->segG; g++)
            {
                const int32_t yS = g * W;
                const int32_t yE = std::min(p->Y, yS + W);
                double acc = 1.0;
                for (int32_t j = yS; j < yE; j++)
                {
                    acc *= fy.fw[static_cast<size_t>(j)];
                    prod[static_cast<size_t>(j - yS)] = acc;
                }
                pEnd[static_cast<size_t>(g)] = acc;
                bEnd[static_cast<size_t>(g)] = 1.0 / fy.ib[static_cast<size_t>(yE - 1)];
                double qn = 1.0;
                double rn = 0.0;
                for (int32_t j = yE - 1; j >= yS; j--)
                {
                    const size_t k = static_cast<size_t>(j);
                    q[k] = fy.bw[k] * qn;
                    rr[k] = fy.bw[k] * rn + fy.ib[k] * prod[static_cast<size_t>(j - yS)];
                    qn = q[k];
                    rn = rr[k];
                }
            }
            segRLast = rr[static_cast<size_t>(p->Y - 1)];
            oSq = put(q);
            oSr = put(rr);
            oSp = put(pEnd);
            oSb = put(bEnd);
        }

        // Row solver and carries for plain ticks.
        const int32_t ldsE = (p->segG > 0 || p->wantRow == HALHIP_ADI_ROWS_WORKGROUP_SCAN) ? 0 : LdsElements(p->shape, p->Y);
        int32_t K = (ldsE == 0 || p->wantCarry == HALHIP_ADI_CARRIES_SCAN) ? 0 : WindowSize(sA, sC);
        if (K > 0 && PickLds(p->shape.cpt, p->shape.s, ldsE, K, p->shape.threads) == nullptr)
        {
            K = 0;
        }
        // A periodic x axis also folds the first and last colWindow strips in
        // every strip.  With strip multipliers of magnitude <= 1 that window
        // is at most K + 1; anything longer takes the scan kernel instead.
        const int32_t colWindow = wrapX ? PeriodicWindowSize(sA, sC) : 0;
        if (K > 0 && colWindow > K + 1)
        {
            K = 0;
        }
        size_t oRow = 0;
        RowTables rt;
        if (ldsE > 0)
        {
            rt = BuildRowTables(p->Y, ldsE, p->shape.threads / 64 / p->shape.s, rate, dir && !wrapY, wrapY != 0);
            oRow = put(rt.data);
        }

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
        a.vecOk = p->shape.vec ? 1 : 0;
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
        a.redW = t + oW;
        a.xvEnd = static_cast<float>(fx.vEnd);
        a.xInv = static_cast<float>(fx.inv);
        a.xibLast = static_cast<float>(fx.ib[static_cast<size_t>(p->X - 1)]);
        a.xz0 = static_cast<float>(fx.z[0]);
        a.xzLast = static_cast<float>(fx.z[static_cast<size_t>(p->X - 1)]);
        a.window = K;
        a.colWindow = colWindow;
        p->ldsE = ldsE;
        p->window = K;
        p->useRow = (p->segG > 0) ? HALHIP_ADI_ROWS_SEGMENTED
                                  : ((ldsE > 0) ? HALHIP_ADI_ROWS_LDS : HALHIP_ADI_ROWS_WORKGROUP_SCAN);
        p->useCarry = (K > 0) ? HALHIP_ADI_CARRIES_WINDOW : HALHIP_ADI_CARRIES_SCAN;
        if (p->segG > 0)
        {
            p->seg.q = t + oSq;
            p->seg.rr = t + oSr;
            p->seg.pEnd = t + oSp;
            p->seg.bEnd = t + oSb;
            p->seg.rLast = static_cast<float>(segRLast);
        }
        if (ldsE > 0)
        {
            p->rs = rt.solve;
            p->rs.pw = t + oRow + rt.oPw;
            p->rs.lane = t + oRow + rt.oLane;
            p->rs.chunk = t + oRow + rt.oChunk;
            p->rs.powE = t + oRow + rt.oPowE;
        }
        p->tableKey = key;
        p->reducedValid = false;
        p->graphs.Clear();
    }

    void LaunchStrip(halhip_adi2d *p, int mode, const char *name)
    {
        StripKernel k = PickStrip(p->shape.cpt, p->shape.s, p->args.vecOk != 0, mode);
        p->prof.Begin(p->clock.stream, name);
        hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(p->P)), dim3(static_cast<unsigned>(p->shape.threads)), 0,
                           p->clock.stream, p->args, p->field.get());
        Check(hipGetLastError(), "AdiStrip launch");
        p->prof.End(p->clock.stream);
    }

    float *RedHalf(halhip_adi2d *p, int32_t half)
    {
        return p->red.get() + static_cast<size_t>(half) * 2 * static_cast<size_t>(p->P) * static_cast<size_t>(p->Y);
    }

    void LaunchLds(halhip_adi2d *p, bool decay)
    {
        const int32_t window = (p->useCarry == HALHIP_ADI_CARRIES_WINDOW) ? p->window : 0;
        LdsKernel k = PickLds(p->shape.cpt, p->shape.s, p->ldsE, window, p->shape.threads, decay,
                              window > 0 && p->args.wrapX != 0);
        Require(k != nullptr, "adi2d: no LDS kernel for the strip shape");
        p->prof.Begin(p->clock.stream, decay ? "adi_strip_tick_decay" : "adi_strip_tick");
        hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(p->P)), dim3(static_cast<unsigned>(p->shape.threads)), 0,
                           p->clock.stream, p->args, p->field.get(), p->window, p->rs);
        Check(hipGetLastError(), "AdiStripLds launch");
        p->prof.End(p->clock.stream);
    }

    void LaunchScan(halhip_adi2d *p)
    {
        p->prof.Begin(p->clock.stream, "adi_strip_scan");
        // up to 16 strips per thread in registers (1024 strips); more reload
        const int32_t Qs = (p->P + kScanSegsWide - 1) / kScanSegsWide;
        if (Qs <= 16)
        {
            auto kernel = (Qs <= 4) ? AdiStripScan<4> : ((Qs <= 8) ? AdiStripScan<8> : AdiStripScan<16>);
            hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>((p->Y + kScanCols - 1) / kScanCols)),
                               dim3(kScanCols * kScanSegsWide), 0, p->clock.stream, p->args, Qs);
        }
        else
        {
            const int32_t G = std::min<int32_t>(kMaxWaves, p->P);
            const int32_t Q = (p->P + G - 1) / G;
            hipLaunchKernelGGL(AdiStripScanGeneric, dim3(static_cast<unsigned>((p->Y + 63) / 64)),
                               dim3(static_cast<unsigned>(64 * G)), 0, p->clock.stream, p->args, Q);
        }
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

    // One tick's launches (the reduction of the field is resident in red
    // half redCur).  Windowed ticks read that half and write the other.
    // decay: the LDS kernel with decay and sources (needs the LDS row solve).
    void PlainTick(halhip_adi2d *p, bool decay = false)
    {
        if (p->useCarry == HALHIP_ADI_CARRIES_WINDOW)
        {
            p->args.redIn = RedHalf(p, p->redCur);
            p->args.red = RedHalf(p, 1 - p->redCur);
            LaunchLds(p, decay);
            p->redCur = 1 - p->redCur;
            return;
        }
        p->args.red = RedHalf(p, p->redCur);
        LaunchScan(p);
        if (p->useRow == HALHIP_ADI_ROWS_LDS)
        {
            LaunchLds(p, decay);
        }
        else
        {
            LaunchStrip(p, 1, "adi_strip_tick");
        }
    }

    using SegKernel = void (*)(AdiArgs, SegArgs, const float *);

    template <int S>
    SegKernel PickSeg(bool solve, bool vec)
    {
        if (solve)
        {
            return vec ? AdiSegSolve<S, true> : AdiSegSolve<S, false>;
        }
        return vec ? AdiSegReduce<S, true> : AdiSegReduce<S, false>;
    }

    void LaunchSeg(halhip_adi2d *p, bool solve)
    {
        const SegKernel k = (p->shape.s == 8) ? PickSeg<8>(solve, p->shape.vec) : PickSeg<4>(solve, p->shape.vec);
        p->prof.Begin(p->clock.stream, solve ? "adi_seg_solve" : "adi_seg_reduce");
        hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(p->P) * static_cast<unsigned>(p->segG)),
                           dim3(static_cast<unsigned>(p->shape.threads)), 0, p->clock.stream, p->args, p->seg,
                           static_cast<const float *>(p->field.get()));
        Check(hipGetLastError(), "AdiSeg launch");
        p->prof.End(p->clock.stream);
    }

    // One tick of a segmented plan (see SegArgs): every tick starts from the
    // reductions of the current field, so point sources added after the
    // previous tick need nothing more.
    void SegTick(halhip_adi2d *p, bool decay, bool sources)
    {
        p->args.red = RedHalf(p, p->redCur);
        LaunchSeg(p, false);
        LaunchScan(p);
        LaunchSeg(p, true);
        p->prof.Begin(p->clock.stream, "adi_seg_scan");
        hipLaunchKernelGGL(AdiSegScan, dim3(static_cast<unsigned>((p->X + 255) / 256)), dim3(256), 0,
                           p->clock.stream, p->args, p->seg);
        Check(hipGetLastError(), "AdiSegScan launch");
        p->prof.End(p->clock.stream);
        const int32_t perBlock = 256 * (p->shape.vec ? 4 : 1);
        const unsigned chunks = static_cast<unsigned>(std::min<int32_t>((p->Y + perBlock - 1) / perBlock, 65535));
        using FixKernel = void (*)(AdiArgs, SegArgs, float *);
        const FixKernel fix = p->shape.vec ? (decay ? AdiSegFix<4, true> : AdiSegFix<4, false>)
                                           : (decay ? AdiSegFix<1, true> : AdiSegFix<1, false>);
        p->prof.Begin(p->clock.stream, "adi_seg_fix");
        hipLaunchKernelGGL(fix, dim3(static_cast<unsigned>(p->X), chunks), dim3(256), 0, p->clock.stream, p->args,
                           p->seg, p->field.get());
        Check(hipGetLastError(), "AdiSegFix launch");
        p->prof.End(p->clock.stream);
        if (sources)
        {
            LaunchSources(p);
        }
    }

    // Strip-sorted copy of the point sources for strip height S.
    void BuildStripSources(halhip_adi2d *p)
    {
        const int32_t S = p->shape.s;
        if (p->srcStripS == S)
        {
            return;
        }
        const int32_t n = p->nSources;
        std::vector<int32_t> offsets(static_cast<size_t>(p->P) + 1, 0);
        for (int32_t k = 0; k < n; k++)
        {
            offsets[static_cast<size_t>(p->srcHostIdx[k] / p->Y / S) + 1]++;
        }
        for (int32_t q = 0; q < p->P; q++)
        {
            offsets[q + 1] += offsets[q];
        }
        std::vector<int32_t> cursor(offsets.begin(), offsets.end() - 1);
        std::vector<int32_t> xy(2 * static_cast<size_t>(n));
        std::vector<float> val(static_cast<size_t>(n));
        for (int32_t k = 0; k < n; k++)
        {
            const int64_t i = p->srcHostIdx[k];
            const int32_t at = cursor[static_cast<size_t>(i / p->Y / S)]++;
            xy[2 * static_cast<size_t>(at)] = static_cast<int32_t>(i / p->Y);
            xy[2 * static_cast<size_t>(at) + 1] = static_cast<int32_t>(i % p->Y);
            val[static_cast<size_t>(at)] = p->srcHostVal[k];
        }
        p->srcStrip.Release();
        p->srcXY.Release();
        p->srcStripVal.Release();
        p->srcStrip.Allocate(p->ctx, offsets.size());
        p->srcXY.Allocate(p->ctx, xy.size());
        p->srcStripVal.Allocate(p->ctx, val.size());
        Check(hipMemcpy(p->srcStrip.get(), offsets.data(), sizeof(int32_t) * offsets.size(), hipMemcpyHostToDevice),
              "hipMemcpy");
        Check(hipMemcpy(p->srcXY.get(), xy.data(), sizeof(int32_t) * xy.size(), hipMemcpyHostToDevice), "hipMemcpy");
        Check(hipMemcpy(p->srcStripVal.get(), val.data(), sizeof(float) * val.size(), hipMemcpyHostToDevice),
              "hipMemcpy");
        p->srcStripS = S;
        p->graphs.Clear();
    }

    // `steps` ticks of the loop body `tick`, replayed as 20-tick graphs
    // keyed by `key` and the reduction half the chunk starts from (even: a
    // windowed chunk leaves redCur as it found it).  In a prepare call the
    // chunk is captured for both halves, so a later run of this length
    // replays a graph whatever ticks ran in between, and nothing is
    // enqueued.
    template <class F>
    void RunChunked(halhip_adi2d *p, int32_t steps, std::string key, F tick)
    {
        constexpr int32_t kChunk = 20;
        int32_t done = 0;
        if (!p->prof.On() && steps >= kChunk)
        {
            AppendKey(key, kChunk);
            auto chunk = [&](int32_t half)
            {
                std::string k = key;
                AppendKey(k, half);
                const int32_t saved = p->redCur;
                p->redCur = half;
                hipGraphExec_t exec = FindOrCapture(p->graphs, k, p->clock.stream,
                                                    [&]
                                                    {
                                                        for (int32_t t = 0; t < kChunk; t++)
                                                        {
                                                            tick();
                                                        }
                                                    });
                p->redCur = saved;
                return exec;
            };
            if (p->prepareOnly)
            {
                chunk(p->redCur);
                if (p->useCarry == HALHIP_ADI_CARRIES_WINDOW)
                {
                    chunk(1 - p->redCur);
                }
                return;
            }
            hipGraphExec_t exec = chunk(p->redCur);
            for (; done + kChunk <= steps; done += kChunk)
            {
                Check(hipGraphLaunch(exec, p->clock.stream), "hipGraphLaunch");
            }
        }
        if (p->prepareOnly)
        {
            return;
        }
        for (; done < steps; done++)
        {
            tick();
        }
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
        p->args.red = RedHalf(p, p->redCur);
        if (p->segG > 0)
        {
            const bool src = withSources && p->nSources > 0;
            const bool dec = decay != 0.0;
            p->args.negDecay = static_cast<float>(-decay);
            std::string key = p->tableKey;
            AppendKey(key, -2); // segmented ticks
            AppendKey(key, p->args.negDecay);
            AppendKey(key, src ? 1 : 0);
            RunChunked(p, steps, key, [p, dec, src] { SegTick(p, dec, src); });
            p->reducedValid = false;
            return;
        }
        if (plain)
        {
            if (!p->reducedValid && !p->prepareOnly)
            {
                LaunchStrip(p, 0, "adi_strip_reduce");
            }
            RunChunked(p, steps, p->tableKey, [p] { PlainTick(p); });
            p->reducedValid = p->reducedValid || !p->prepareOnly;
            return;
        }
        p->args.negDecay = static_cast<float>(-decay);
        if (p->useRow == HALHIP_ADI_ROWS_LDS)
        {
            // Decay and sources inside the LDS tick kernel, which also leaves
            // the reduction of u' for the next tick.
            const bool src = withSources && p->nSources > 0;
            if (src)
            {
                BuildStripSources(p);
            }
            p->args.srcStrip = src ? p->srcStrip.get() : nullptr;
            p->args.srcXY = src ? p->srcXY.get() : nullptr;
            p->args.srcVal = src ? p->srcStripVal.get() : nullptr;
            if (!p->reducedValid && !p->prepareOnly)
            {
                LaunchStrip(p, 0, "adi_strip_reduce");
            }
            std::string key = p->tableKey;
            AppendKey(key, -1); // decay ticks
            AppendKey(key, p->args.negDecay);
            AppendKey(key, src ? 1 : 0);
            RunChunked(p, steps, key, [p] { PlainTick(p, true); });
            p->reducedValid = p->reducedValid || !p->prepareOnly;
            return;
        }
        if (p->prepareOnly)
        {
            return; // the register path's model ticks launch directly
        }
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

namespace halhip
{
    void AdiDeltaOf(halhip_adi2d *p, const float *u, float *d, int32_t X, int32_t Y, double rate, int32_t wrapX,
                    int32_t wrapY, const halhip_bc *bc, hipStream_t caller)
    {
        Require(p != nullptr, "delta_adi: null ADI plan");
        Require(p->X == X && p->Y == Y, "delta_adi: the ADI plan's dimensions differ from the grid's");
        Bind(p->ctx);
        const size_t n = Cells(p);
        Check(hipStreamSynchronize(caller), "hipStreamSynchronize");
        Check(hipMemcpyAsync(p->field.get(), u, n * sizeof(float), hipMemcpyDeviceToDevice, p->clock.stream),
              "hipMemcpyAsync");
        p->reducedValid = false;
        EnqueueTicks(p, 1, rate, 0.0, wrapX, wrapY, bc, false);
        hipLaunchKernelGGL(AdiAccumulateDelta, dim3(1024), dim3(256), 0, p->clock.stream, d,
                           static_cast<const float *>(p->field.get()), u, static_cast<int64_t>(n));
        Check(hipGetLastError(), "AdiAccumulateDelta launch");
        Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
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
                Require(static_cast<int64_t>(xDim) * yDim <= (int64_t(1) << 31),
                        "halhip_adi2d_create: grid larger than 2^31 cells");
                Bind(ctx);
                p = new halhip_adi2d();
                p->ctx = ctx;
                p->X = xDim;
                p->Y = yDim;
                const size_t n = Cells(p);
                p->field.Allocate(ctx, n);
                if (yDim > HALHIP_ADI_MAX_Y)
                {
                    ConfigureSegments(p, true);
                }
                else
                {
                    p->shape = AutoShape(xDim, yDim);
                }
                AllocateStrips(p);
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

    int halhip_adi2d_save_grid(halhip_adi2d *p, const char *path, int32_t wrapX, int32_t wrapY,
                                  int32_t updateCt)
    {
        std::vector<float> host;
        int rc = Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_save_grid");
                host.resize(Cells(p));
            });
        if (rc != HALHIP_OK)
        {
            return rc;
        }
        rc = halhip_adi2d_download(p, host.data(), static_cast<int64_t>(host.size()));
        if (rc != HALHIP_OK)
        {
            return rc;
        }
        return Guard([&] { SavePdeGrid2D(path, p->X, p->Y, wrapX != 0, wrapY != 0, updateCt, host); });
    }

    int halhip_adi2d_load_grid(halhip_adi2d *p, const char *path, int32_t *wrapX, int32_t *wrapY,
                                  int32_t *updateCt)
    {
        std::vector<float> host;
        int rc = Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_load_grid");
                bool wx = false;
                bool wy = false;
                int32_t ct = 0;
                LoadPdeGrid2D(path, p->X, p->Y, wx, wy, ct, host);
                if (wrapX != nullptr)
                {
                    *wrapX = wx ? 1 : 0;
                }
                if (wrapY != nullptr)
                {
                    *wrapY = wy ? 1 : 0;
                }
                if (updateCt != nullptr)
                {
                    *updateCt = ct;
                }
            });
        if (rc != HALHIP_OK)
        {
            return rc;
        }
        return halhip_adi2d_upload(p, host.data(), static_cast<int64_t>(host.size()));
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
                p->srcHostIdx.assign(index, index + count);
                p->srcHostVal.assign(value, value + count);
                p->srcStripS = 0;
                p->graphs.Clear();
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
                if (p->callField.get() == nullptr)
                {
                    p->callField.Allocate(p->ctx, cells);
                }
                // One ungraphed tick on the call's own field; the swap back
                // runs on every exit, and the reductions left behind belong
                // to the call's field, so the resident field's are redone.
                struct SwapBack
                {
                    halhip_adi2d *plan;
                    ~SwapBack()
                    {
                        plan->field.SwapWith(plan->callField);
                        plan->reducedValid = false;
                    }
                };
                p->field.SwapWith(p->callField);
                p->reducedValid = false;
                SwapBack restore{p};
                p->chunks.Upload(field, cells, p->hostBuf.data(), p->field.get(), p->clock.stream);
                EnqueueTicks(p, 1, diffRate, 0.0, wrapX, wrapY, bc, false);
                p->chunks.DownloadAdd(p->field.get(), cells, p->hostBuf.data(), deltas, field, p->clock.stream);
                p->prof.Collect();
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

    int halhip_adi2d_prepare_steps(halhip_adi2d *p, int32_t steps, double diffRate, double decay, int32_t wrapX,
                                   int32_t wrapY, const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_prepare_steps");
                PrepareScope scope(p->prepareOnly);
                EnqueueTicks(p, steps, diffRate, decay, wrapX, wrapY, bc, true);
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

    int halhip_adi2d_set_strip_shape(halhip_adi2d *p, int32_t columnsPerThread, int32_t stripRows, int32_t vector)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_set_strip_shape");
                if (p->segG > 0)
                {
                    // Segmented rows: 4 columns per thread; the strip height is free.
                    Require(columnsPerThread == kSegCpt && (stripRows == 4 || stripRows == 8),
                            "halhip_adi2d_set_strip_shape: segmented rows take (4, 4) or (4, 8)");
                    Require(!vector || p->Y % kSegCpt == 0,
                            "halhip_adi2d_set_strip_shape: vector rows need yDim to be a multiple of 4");
                    Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                    p->shape.s = stripRows;
                    p->shape.vec = vector != 0;
                    AllocateStrips(p);
                    return;
                }
                Require(ShapeCompiled(columnsPerThread, stripRows),
                        "halhip_adi2d_set_strip_shape: (columnsPerThread, stripRows) must be one of "
                        "{1,2,4}x{4,8,16} or 8x{4,8}");
                const int32_t lanes = (p->Y + columnsPerThread - 1) / columnsPerThread;
                Require(lanes <= 1024, "halhip_adi2d_set_strip_shape: yDim needs more than 1024 threads at this width");
                Require(!vector || p->Y % columnsPerThread == 0,
                        "halhip_adi2d_set_strip_shape: vector rows need yDim to be a multiple of columnsPerThread");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                p->shape.cpt = columnsPerThread;
                p->shape.s = stripRows;
                p->shape.threads = ((lanes + 63) / 64) * 64;
                p->shape.vec = vector != 0;
                AllocateStrips(p);
            });
    }

    int halhip_adi2d_set_methods(halhip_adi2d *p, int32_t rowSolver, int32_t carries)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_set_methods");
                Require(rowSolver >= HALHIP_ADI_AUTO && rowSolver <= HALHIP_ADI_ROWS_SEGMENTED,
                        "halhip_adi2d_set_methods: unknown row solver");
                Require(carries >= HALHIP_ADI_AUTO && carries <= HALHIP_ADI_CARRIES_WINDOW,
                        "halhip_adi2d_set_methods: unknown carry method");
                const bool longRows = p->Y > HALHIP_ADI_MAX_Y;
                if (longRows && rowSolver != HALHIP_ADI_AUTO && rowSolver != HALHIP_ADI_ROWS_SEGMENTED)
                {
                    Fail(HALHIP_ERR_UNSUPPORTED,
                         "halhip_adi2d_set_methods: rows longer than HALHIP_ADI_MAX_Y are always segmented");
                }
                if (longRows && carries == HALHIP_ADI_CARRIES_WINDOW)
                {
                    Fail(HALHIP_ERR_UNSUPPORTED, "halhip_adi2d_set_methods: segmented rows take scan carries");
                }
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                const bool seg = longRows || rowSolver == HALHIP_ADI_ROWS_SEGMENTED;
                if (seg != (p->segG > 0))
                {
                    ConfigureSegments(p, seg);
                    AllocateStrips(p);
                }
                p->wantRow = rowSolver;
                p->wantCarry = carries;
                p->tableKey.clear();
                p->reducedValid = false;
                p->graphs.Clear();
            });
    }

    int halhip_adi2d_get_methods(halhip_adi2d *p, int32_t *rowSolver, int32_t *carries, int32_t *window)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_get_methods");
                Require(rowSolver != nullptr && carries != nullptr && window != nullptr,
                        "halhip_adi2d_get_methods: null output");
                *rowSolver = p->useRow;
                *carries = p->useCarry;
                *window = p->window;
            });
    }

    int halhip_adi2d_get_strip_shape(halhip_adi2d *p, int32_t *columnsPerThread, int32_t *stripRows, int32_t *vector)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_adi2d_get_strip_shape");
                Require(columnsPerThread != nullptr && stripRows != nullptr && vector != nullptr,
                        "halhip_adi2d_get_strip_shape: null output");
                *columnsPerThread = p->shape.cpt;
                *stripRows = p->shape.s;
                *vector = p->shape.vec ? 1 : 0;
            });
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
