// adi_device.hpp -- device side of the ADI strip solver, shared by the
// single-grid plan (adi2d.hip) and the batched-wells plan (batch2d.hip).
// See adi2d.hip for the method.
#pragma once

#include "common.hpp"

#include <cmath>
#include <vector>

namespace
{
    using namespace halhip;

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
        // strip reduction and carries, [2][P][Y] each (red: written by this
        // launch; redIn: the current tick's reductions, read by windowed ticks)
        float *red;
        const float *redIn;
        float *carry;
        float *colInfo; // [3][Y]: SM factor, u*(0,y), u*(X-1,y)
        const float *stripA;
        const float *stripC;
        const float *stripE;
        const float *redW; // [P][2][S]: per-row weights of the strip reductions B and D
        float xvEnd;
        float xInv;
        float xibLast;
        float xz0;
        float xzLast;
        // point sources applied by MODE 3 (batched wells): pairs (x, y) and
        // values, entries [srcBegin, srcEnd) belong to the current strip
        const int32_t *srcXY;
        const float *srcVal;
        int32_t srcBegin;
        int32_t srcEnd;
        // single grid with decay/sources on the LDS path: per-strip offsets
        // [P+1] into srcXY / srcVal (null: no sources)
        const int32_t *srcStrip;
        int32_t window; // K of windowed carries (0: carries from the scan kernel)
        // periodic x with windowed carries: the first and last colWindow
        // strips' reductions give the column values (PeriodicWindowSize)
        int32_t colWindow;
    };

    // ---- wavefront primitives ------------------------------------------------------

    // DPP lane moves (GFX9 encodings): row_shl:n = 0x100+n, row_shr:n =
    // 0x110+n, wave_shl:1 = 0x130, wave_shr:1 = 0x138, row_bcast:15 = 0x142,
    // row_bcast:31 = 0x143.  Lanes without a source (or
    // in rows outside rowMask) receive `old`.
    template <int CTRL, int ROW_MASK = 0xF>
    __device__ __forceinline__ float Dpp(float old, float v)
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                      __builtin_bit_cast(int, v), CTRL, ROW_MASK,
                                                                      0xF, false));
    }

    // Inclusive left-to-right scan of the affine maps (a_l, b_l) over the 64
    // lanes of a wavefront, composing x -> a*x + b.  The six stages of the
    // multiplier scan do not depend on the row, so they are computed once
    // (ScanStagesForward) and each row only scans its offsets (ScanRowForward).
    __device__ __forceinline__ float ScanStagesForward(float a, float (&stage)[6])
    {
        stage[0] = a;
        a = Dpp<0x111>(1.0f, a) * a;
        stage[1] = a;
        a = Dpp<0x112>(1.0f, a) * a;
        stage[2] = a;
        a = Dpp<0x114>(1.0f, a) * a;
        stage[3] = a;
        a = Dpp<0x118>(1.0f, a) * a;
        stage[4] = a;
        a = Dpp<0x142, 0xA>(1.0f, a) * a;
        stage[5] = a;
        a = Dpp<0x143, 0xC>(1.0f, a) * a;
        return a;
    }

    __device__ __forceinline__ float ScanRowForward(float b, const float (&stage)[6])
    {
        b = fmaf(stage[0], Dpp<0x111>(0.0f, b), b);
        b = fmaf(stage[1], Dpp<0x112>(0.0f, b), b);
        b = fmaf(stage[2], Dpp<0x114>(0.0f, b), b);
        b = fmaf(stage[3], Dpp<0x118>(0.0f, b), b);
        b = fmaf(stage[4], Dpp<0x142, 0xA>(0.0f, b), b);
        b = fmaf(stage[5], Dpp<0x143, 0xC>(0.0f, b), b);
        return b;
    }

    // Right-to-left counterparts (Kogge-Stone with shuffles down).
    __device__ __forceinline__ float ScanStagesBackward(float a, int lane, float (&stage)[6])
    {
#pragma unroll
        for (int k = 0; k < 6; k++)
        {
            const int o = 1 << k;
            stage[k] = a;
            const float d = __shfl_down(a, o);
            a = (lane + o < 64) ? d * a : a;
        }
        return a;
    }

    __device__ __forceinline__ float ScanRowBackward(float b, int lane, const float (&stage)[6])
    {
#pragma unroll
        for (int k = 0; k < 6; k++)
        {
            const int o = 1 << k;
            const float d = __shfl_down(b, o);
            b = (lane + o < 64) ? fmaf(stage[k], d, b) : b;
        }
        return b;
    }

    // ---- register-tile helpers ------------------------------------------------------

    template <int CPT, bool VEC>
    __device__ __forceinline__ void LoadRow(const float *__restrict__ row, int y0, int Y, float (&d)[CPT])
    {
        if constexpr (VEC)
        {
            if (y0 + CPT <= Y)
            {
                if constexpr (CPT % 4 == 0)
                {
#pragma unroll
                    for (int q = 0; q < CPT / 4; q++)
                    {
                        const float4 t = *reinterpret_cast<const float4 *>(row + y0 + 4 * q);
                        d[4 * q + 0] = t.x;
                        d[4 * q + 1] = t.y;
                        d[4 * q + 2] = t.z;
                        d[4 * q + 3] = t.w;
                    }
                }
                else if constexpr (CPT == 2)
                {
                    const float2 t = *reinterpret_cast<const float2 *>(row + y0);
                    d[0] = t.x;
                    d[1] = t.y;
                }
                else
                {
                    d[0] = row[y0];
                }
            }
            else
            {
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    d[c] = 0.0f;
                }
            }
        }
        else
        {
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                d[c] = (y0 + c < Y) ? row[y0 + c] : 0.0f;
            }
        }
    }

    template <int CPT, bool VEC>
    __device__ __forceinline__ void StoreRow(float *__restrict__ row, int y0, int Y, const float (&d)[CPT])
    {
        if constexpr (VEC)
        {
            if (y0 + CPT <= Y)
            {
                if constexpr (CPT % 4 == 0)
                {
#pragma unroll
                    for (int q = 0; q < CPT / 4; q++)
                    {
                        *reinterpret_cast<float4 *>(row + y0 + 4 * q) =
                            make_float4(d[4 * q + 0], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3]);
                    }
                }
                else if constexpr (CPT == 2)
                {
                    *reinterpret_cast<float2 *>(row + y0) = make_float2(d[0], d[1]);
                }
                else
                {
                    row[y0] = d[0];
                }
            }
        }
        else
        {
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                if (y0 + c < Y)
                {
                    row[y0 + c] = d[c];
                }
            }
        }
    }

    // A thread's CPT columns as vectors with no column check: the caller
    // guarantees y0 + CPT <= Y and y0 a multiple of CPT.
    template <int CPT>
    __device__ __forceinline__ void LoadRowFull(const float *__restrict__ row, int y0, float (&d)[CPT])
    {
        if constexpr (CPT % 4 == 0)
        {
#pragma unroll
            for (int q = 0; q < CPT / 4; q++)
            {
                const float4 t = *reinterpret_cast<const float4 *>(row + y0 + 4 * q);
                d[4 * q + 0] = t.x;
                d[4 * q + 1] = t.y;
                d[4 * q + 2] = t.z;
                d[4 * q + 3] = t.w;
            }
        }
        else if constexpr (CPT == 2)
        {
            const float2 t = *reinterpret_cast<const float2 *>(row + y0);
            d[0] = t.x;
            d[1] = t.y;
        }
        else
        {
            d[0] = row[y0];
        }
    }

    struct StripLds
    {
        float *left;  // [kMaxWaves][S] first column of each wavefront
        float *right; // [kMaxWaves][S] last column of each wavefront
        float *first; // [S] column y = 0
        float *last;  // [S] column y = Y-1
        float *tot;   // [kMaxWaves][S] wavefront scan totals
        float *aw;    // [kMaxWaves] wavefront multiplier products
        float *ends;  // [2][S] first / last solution values of each row (periodic y)
    };

    // Sweep A right-hand side in place: v <- v + hr*(dv/dy+ + dv/dy-) plus the
    // Dirichlet implicit term r*bc on rows 0 and X-1.  Rows past X and columns
    // past Y become 0.  An off-grid neighbour is replaced by a ghost value g
    // with g - c equal to the reference's difference: the wrapped cell
    // (periodic), 2*bc - c (Dirichlet, whose difference is (bc - c)*2) or c
    // (zero flux).  ExplicitDiffusionY2ADI (src/Pde/ADIequations.cpp:34-47,
    // ADIequations.h:79-94) + ImplicitBoundarySource (ADIequations.h:198-220).
    template <int CPT, int S, bool VEC>
    __device__ __forceinline__ void SweepARhs(const AdiArgs &a, int s0, int y0, int tid, float (&v)[S][CPT],
                                              const StripLds &L)
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
        if (tid == 0)
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                L.first[r] = v[r][0];
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
                        L.last[r] = v[r][c];
                    }
                }
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
            if (tid == 0)
            {
                const float c0 = v[r][0];
                left = a.wrapY ? L.last[r] : (dirY ? 2.0f * a.bcYlo[x] - c0 : c0);
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
                ghostHi = a.wrapY ? L.first[r] : (dirY ? 2.0f * a.bcYhi[x] - cl : cl);
                if constexpr (VEC)
                {
                    right = ghostHi;
                }
            }
            const bool srcRow = a.dirichlet && !a.wrapX && (x == 0 || x == a.X - 1);
            float prev = left;
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                const float cur = v[r][c];
                float nxt = (c + 1 < CPT) ? v[r][c + 1] : right;
                if constexpr (!VEC)
                {
                    if (tid == lastT && c == lastC)
                    {
                        nxt = ghostHi;
                    }
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

    // Zero-carry local solve of sweep A on the strip (v holds the RHS and is
    // consumed) and write of the strip's two reduction values per column.
    template <int CPT, int S, bool VEC>
    __device__ __forceinline__ void ReduceA(const AdiArgs &a, int p, int s0, int y0, float (&v)[S][CPT])
    {
        float acc[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            acc[c] = 0.0f;
        }
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const float f = a.xfw[s0 + r];
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                acc[c] = fmaf(f, acc[c], v[r][c]);
                v[r][c] = acc[c];
            }
        }
        const size_t base = static_cast<size_t>(p) * a.Y;
        const size_t plane = static_cast<size_t>(a.P) * a.Y;
        StoreRow<CPT, VEC>(a.red + base, y0, a.Y, acc);
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            acc[c] = 0.0f;
        }
#pragma unroll
        for (int r = S - 1; r >= 0; r--)
        {
            const float bwr = a.xbw[s0 + r];
            const float ibr = a.xib[s0 + r];
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                acc[c] = fmaf(bwr, acc[c], ibr * v[r][c]);
            }
        }
        StoreRow<CPT, VEC>(a.red + plane + base, y0, a.Y, acc);
    }

    // Per-column values of a periodic x axis (Sherman-Morrison on the cyclic
    // sweep-A lines): the correction factor, u*(0, y) and u*(X-1, y).  The
    // scan kernel writes them to colInfo; with windowed carries each strip
    // computes them itself (WindowPeriodicCols).
    template <int CPT>
    struct PeriodicCols
    {
        float factor[CPT];
        float top[CPT];
        float bottom[CPT];
    };

    // Finishes sweep A on the strip from its carries (v: RHS in, u* out): cf
    // is d' of row s0-1 (forward carry-in), dn on entry u* of row s0+S
    // (backward carry-in).  Returns u* of the rows just above (s0-1, in up)
    // and below (s0+S, in dn) the strip.  On a periodic x axis the factor
    // comes from pc when PCREG, else from colInfo.
    template <int CPT, int S, bool VEC, bool PCREG = false>
    __device__ __forceinline__ void FinishACore(const AdiArgs &a, int s0, int y0, float (&v)[S][CPT],
                                                const float (&cf)[CPT], float (&up)[CPT], float (&dn)[CPT],
                                                const PeriodicCols<CPT> *pc = nullptr)
    {
        float acc[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            acc[c] = cf[c];
        }
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const float f = a.xfw[s0 + r];
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                acc[c] = fmaf(f, acc[c], v[r][c]);
                v[r][c] = acc[c];
            }
        }
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            acc[c] = dn[c];
        }
#pragma unroll
        for (int r = S - 1; r >= 0; r--)
        {
            const float bwr = a.xbw[s0 + r];
            const float ibr = a.xib[s0 + r];
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                acc[c] = fmaf(bwr, acc[c], ibr * v[r][c]);
                v[r][c] = acc[c];
            }
        }
        if (s0 > 0)
        {
            const float bwUp = a.xbw[s0 - 1];
            const float ibUp = a.xib[s0 - 1];
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                up[c] = fmaf(bwUp, v[0][c], ibUp * cf[c]);
            }
        }
        else
        {
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                up[c] = 0.0f;
            }
        }
        if (a.wrapX)
        {
            float factor[CPT];
            if constexpr (PCREG)
            {
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    factor[c] = pc->factor[c];
                }
            }
            else
            {
                LoadRow<CPT, VEC>(a.colInfo, y0, a.Y, factor);
            }
            const float zUp = (s0 > 0) ? a.xz[s0 - 1] : 0.0f;
            const float zDn = (s0 + S < a.X) ? a.xz[s0 + S] : 0.0f;
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                const float z = a.xz[s0 + r];
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    v[r][c] = fmaf(-factor[c], z, v[r][c]);
                }
            }
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                up[c] = fmaf(-factor[c], zUp, up[c]);
                dn[c] = fmaf(-factor[c], zDn, dn[c]);
            }
        }
    }

    // FinishACore with the carries written by the cross-strip scan kernel.
    template <int CPT, int S, bool VEC>
    __device__ __forceinline__ void FinishA(const AdiArgs &a, int p, int s0, int y0, float (&v)[S][CPT],
                                            float (&up)[CPT], float (&dn)[CPT])
    {
        const size_t base = static_cast<size_t>(p) * a.Y;
        const size_t plane = static_cast<size_t>(a.P) * a.Y;
        float cf[CPT];
        LoadRow<CPT, VEC>(a.carry + base, y0, a.Y, cf);
        LoadRow<CPT, VEC>(a.carry + plane + base, y0, a.Y, dn);
        FinishACore<CPT, S, VEC>(a, s0, y0, v, cf, up, dn);
    }

    // Sweep B right-hand side in place from u*: v <- v + hr*(dv/dx+ + dv/dx-)
    // plus the Dirichlet implicit term on columns 0 and Y-1, with the same
    // ghost-value treatment of off-grid neighbours as sweep A.
    // ExplicitDiffusionX2ADI (src/Pde/ADIequations.cpp:19-32, ADIequations.h:54-69).
    template <int CPT, int S, bool VEC, bool PCREG = false>
    __device__ __forceinline__ void SweepBRhs(const AdiArgs &a, int s0, int y0, int tid, float (&v)[S][CPT],
                                              const float (&up)[CPT], const float (&dn)[CPT],
                                              const PeriodicCols<CPT> *pc = nullptr)
    {
        const int lastT = (a.Y - 1) / CPT;
        const int lastC = (a.Y - 1) % CPT;
        const bool dirX = a.dirichlet && !a.wrapX;
        const bool dirY = a.dirichlet && !a.wrapY;
        float prevRow[CPT];
        if (s0 == 0)
        {
            // Ghost row above row 0.
            float g[CPT];
            if (a.wrapX)
            {
                if constexpr (PCREG)
                {
#pragma unroll
                    for (int c = 0; c < CPT; c++)
                    {
                        g[c] = pc->bottom[c];
                    }
                }
                else
                {
                    LoadRow<CPT, VEC>(a.colInfo + 2 * static_cast<size_t>(a.Y), y0, a.Y, g);
                }
            }
            else if (dirX)
            {
                LoadRow<CPT, VEC>(a.bcXlo, y0, a.Y, g);
            }
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                prevRow[c] = a.wrapX ? g[c] : (dirX ? fmaf(2.0f, g[c], -v[0][c]) : v[0][c]);
            }
        }
        else
        {
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                prevRow[c] = up[c];
            }
        }
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const int x = s0 + r;
            if (x >= a.X)
            {
                break;
            }
            float next[CPT];
            if (x + 1 < a.X)
            {
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    next[c] = (r + 1 < S) ? v[r + 1 < S ? r + 1 : r][c] : dn[c];
                }
            }
            else
            {
                float g[CPT];
                if (a.wrapX)
                {
                    if constexpr (PCREG)
                    {
#pragma unroll
                        for (int c = 0; c < CPT; c++)
                        {
                            g[c] = pc->top[c];
                        }
                    }
                    else
                    {
                        LoadRow<CPT, VEC>(a.colInfo + static_cast<size_t>(a.Y), y0, a.Y, g);
                    }
                }
                else if (dirX)
                {
                    LoadRow<CPT, VEC>(a.bcXhi, y0, a.Y, g);
                }
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    next[c] = a.wrapX ? g[c] : (dirX ? fmaf(2.0f, g[c], -v[r][c]) : v[r][c]);
                }
            }
            float srcLo = 0.0f;
            float srcHi = 0.0f;
            if (dirY)
            {
                srcLo = (tid == 0) ? a.bcYlo[x] * a.r : 0.0f;
                srcHi = (tid == lastT) ? a.bcYhi[x] * a.r : 0.0f;
            }
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                const float cur = v[r][c];
                float rhs = fmaf((next[c] - cur) + (prevRow[c] - cur), a.hr, cur);
                if (dirY)
                {
                    rhs += (c == 0) ? srcLo : 0.0f;
                    rhs += (c == lastC) ? srcHi : 0.0f;
                }
                prevRow[c] = cur;
                v[r][c] = (y0 + c < a.Y) ? rhs : 0.0f;
            }
        }
    }

    // Sweep B solve of every row of the strip across the workgroup.  SEG:
    // the workgroup holds one segment of longer rows, solved with zero
    // carries at both of its ends and without the periodic closure (the
    // segments are joined afterwards).
    template <int CPT, int S, bool SEG = false>
    __device__ __forceinline__ void SolveB(const AdiArgs &a, int y0, int tid, float (&v)[S][CPT], const StripLds &L)
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
        float stage[6];
        const float aInc = ScanStagesForward(aSeg, stage);
        const float aEx = Dpp<0x138>(1.0f, aInc);
        if (lane == 63)
        {
            L.aw[wave] = aInc;
        }
        float bEx[S];
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            float b = 0.0f;
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                b = fmaf(fw[c], b, v[r][c]);
            }
            b = ScanRowForward(b, stage);
            bEx[r] = Dpp<0x138>(0.0f, b);
            if (lane == 63)
            {
                L.tot[wave * S + r] = b;
            }
        }
        __syncthreads();
        // Lane r folds the totals of the wavefronts before this one for row r.
        float cwLane = 0.0f;
        if (lane < S)
        {
            for (int w = 0; w < wave; w++)
            {
                cwLane = fmaf(L.aw[w], cwLane, L.tot[w * S + lane]);
            }
        }
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const float cw = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, cwLane), r));
            float acc = fmaf(aEx, cw, bEx[r]);
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                acc = fmaf(fw[c], acc, v[r][c]);
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
        const float aIncB = ScanStagesBackward(aSeg, lane, stage);
        float aExB = __shfl_down(aIncB, 1);
        aExB = (lane == 63) ? 1.0f : aExB;
        if (lane == 0)
        {
            L.aw[wave] = aIncB;
        }
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            float b = 0.0f;
#pragma unroll
            for (int c = CPT - 1; c >= 0; c--)
            {
                b = fmaf(bw[c], b, ib[c] * v[r][c]);
            }
            b = ScanRowBackward(b, lane, stage);
            const float be = __shfl_down(b, 1);
            bEx[r] = (lane == 63) ? 0.0f : be;
            if (lane == 0)
            {
                L.tot[wave * S + r] = b;
            }
        }
        __syncthreads();
        cwLane = 0.0f;
        if (lane < S)
        {
            for (int w = nWaves - 1; w > wave; w--)
            {
                cwLane = fmaf(L.aw[w], cwLane, L.tot[w * S + lane]);
            }
        }
        const int lastT = (a.Y - 1) / CPT;
        const int lastC = (a.Y - 1) % CPT;
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const float cw = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, cwLane), r));
            float acc = fmaf(aExB, cw, bEx[r]);
#pragma unroll
            for (int c = CPT - 1; c >= 0; c--)
            {
                acc = fmaf(bw[c], acc, ib[c] * v[r][c]);
                v[r][c] = acc;
            }
            if (!SEG && a.wrapY)
            {
                if (tid == 0)
                {
                    L.ends[r] = v[r][0];
                }
                if (tid == lastT)
                {
#pragma unroll
                    for (int c = 0; c < CPT; c++)
                    {
                        if (c == lastC)
                        {
                            L.ends[S + r] = v[r][c];
                        }
                    }
                }
            }
        }
        if (!SEG && a.wrapY)
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
                const float factor = (L.ends[r] + L.ends[S + r] * a.yvEnd) * a.yInv;
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    v[r][c] = fmaf(-factor, z[c], v[r][c]);
                }
            }
        }
    }

    // ---- windowed carries ------------------------------------------------------------
    //
    // Sweep A's lines are diagonally dominant, so the carry a strip receives
    // from strip q decays like the product of the strip multipliers in
    // between.  When the host has found K with every product of K consecutive
    // strip multipliers A (forward) and C (backward) below 2^-28, each strip
    // composes its carries from the K strips on either side and the
    // cross-strip scan kernel is not needed:
    //   F_{q+1} = A_q F_q + B_q            (d' of the row before strip q+1)
    //   X_q     = D_q + E_q F_q + C_q X_{q+1}   (u* of the first row of strip q)
    // starting from F = 0 K strips up (exact at the top of the grid) and
    // X = 0 K strips down (exact at the bottom).
    constexpr int kMaxWindow = 8;

    //
    // The window's rows are requested in batches of at most 32 values before
    // any of the batch is used, so a window costs a few memory round trips
    // rather than one per strip; the fold order, and so every bit, is the
    // strip-by-strip one.
    template <int CPT, bool VEC, int KW>
    __device__ __forceinline__ void WindowCarries(const AdiArgs &a, int p, int y0, int K, float (&cf)[CPT],
                                                  float (&xb)[CPT])
    {
        const size_t plane = static_cast<size_t>(a.P) * a.Y;
        const float *B = a.redIn;
        const float *D = a.redIn + plane;
        const int qLo = max(0, p - K);
        const int nUp = p - qLo;                 // strips above folded into cf
        const int nDn = min(a.P - 1, p + K) - p; // strips below folded into xb
        constexpr int kUp = (32 / CPT < KW) ? (32 / CPT > 0 ? 32 / CPT : 1) : KW;        // B rows per batch
        constexpr int kDn = (16 / CPT < KW) ? (16 / CPT > 0 ? 16 / CPT : 1) : KW;        // (B, D) pairs per batch
        float F[CPT];
        float X[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            F[c] = 0.0f;
            X[c] = 0.0f;
        }
#pragma unroll
        for (int i0 = 0; i0 < KW; i0 += kUp)
        {
            if (i0 < nUp)
            {
                float bq[kUp][CPT] = {};
#pragma unroll
                for (int j = 0; j < kUp; j++)
                {
                    if (i0 + j < nUp && i0 + j < KW)
                    {
                        LoadRow<CPT, VEC>(B + static_cast<size_t>(qLo + i0 + j) * a.Y, y0, a.Y, bq[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < kUp; j++)
                {
                    if (i0 + j < nUp && i0 + j < KW)
                    {
                        const float A = a.stripA[qLo + i0 + j];
#pragma unroll
                        for (int c = 0; c < CPT; c++)
                        {
                            F[c] = fmaf(A, F[c], bq[j][c]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            cf[c] = F[c];
        }
        // X_{p+1} = sum_q (C_{p+1} ... C_{q-1}) (D_q + E_q F_q), q = p+1 .. p+nDn,
        // with F_q = A_{q-1} F_{q-1} + B_{q-1}.
        float coef = 1.0f;
#pragma unroll
        for (int i0 = 0; i0 < KW; i0 += kDn)
        {
            if (i0 < nDn)
            {
                float bq[kDn][CPT] = {};
                float dq[kDn][CPT] = {};
#pragma unroll
                for (int j = 0; j < kDn; j++)
                {
                    if (i0 + j < nDn && i0 + j < KW)
                    {
                        const int q = p + 1 + i0 + j;
                        LoadRow<CPT, VEC>(B + static_cast<size_t>(q - 1) * a.Y, y0, a.Y, bq[j]);
                        LoadRow<CPT, VEC>(D + static_cast<size_t>(q) * a.Y, y0, a.Y, dq[j]);
                    }
                }
#pragma unroll
                for (int j = 0; j < kDn; j++)
                {
                    if (i0 + j < nDn && i0 + j < KW)
                    {
                        const int q = p + 1 + i0 + j;
                        const float A = a.stripA[q - 1];
                        const float Eq = a.stripE[q];
#pragma unroll
                        for (int c = 0; c < CPT; c++)
                        {
                            F[c] = fmaf(A, F[c], bq[j][c]);
                            X[c] = fmaf(coef, fmaf(Eq, F[c], dq[j][c]), X[c]);
                        }
                        coef *= a.stripC[q];
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            xb[c] = X[c];
        }
    }

    // Windowed carries of a strip with K <= KW full strips on both sides
    // (K <= p <= P-1-K): every load is issued before any is used
    // (WindowLoads), then WindowFold composes them in the order
    // WindowCarries does, so the results are the same bits.
    // up: B rows p-K .. p-1; bd: B rows p .. p+K-1; wd: D rows p+1 .. p+K.
    template <int CPT, bool VEC, int KW>
    __device__ __forceinline__ void WindowLoads(const AdiArgs &a, int p, int y0, int K, float (&up)[KW][CPT],
                                                float (&bd)[KW][CPT], float (&wd)[KW][CPT])
    {
        const size_t plane = static_cast<size_t>(a.P) * a.Y;
#pragma unroll
        for (int i = 0; i < KW; i++)
        {
            if (i < K)
            {
                LoadRow<CPT, VEC>(a.redIn + static_cast<size_t>(p - K + i) * a.Y, y0, a.Y, up[i]);
                LoadRow<CPT, VEC>(a.redIn + static_cast<size_t>(p + i) * a.Y, y0, a.Y, bd[i]);
                LoadRow<CPT, VEC>(a.redIn + plane + static_cast<size_t>(p + 1 + i) * a.Y, y0, a.Y, wd[i]);
            }
        }
    }

    template <int CPT, int KW>
    __device__ __forceinline__ void WindowFold(const AdiArgs &a, int p, int K, const float (&up)[KW][CPT],
                                               const float (&bd)[KW][CPT], const float (&wd)[KW][CPT],
                                               float (&cf)[CPT], float (&xb)[CPT])
    {
        float F[CPT];
        float X[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            F[c] = 0.0f;
            X[c] = 0.0f;
        }
#pragma unroll
        for (int i = 0; i < KW; i++)
        {
            if (i < K)
            {
                const float A = a.stripA[p - K + i];
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    F[c] = fmaf(A, F[c], up[i][c]);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            cf[c] = F[c];
        }
        float coef = 1.0f;
#pragma unroll
        for (int i = 0; i < KW; i++)
        {
            if (i < K)
            {
                const int q = p + 1 + i;
                const float A = a.stripA[q - 1];
                const float Eq = a.stripE[q];
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    F[c] = fmaf(A, F[c], bd[i][c]);
                    X[c] = fmaf(coef, fmaf(Eq, F[c], wd[i][c]), X[c]);
                }
                coef *= a.stripC[q];
            }
        }
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            xb[c] = X[c];
        }
    }

    // The periodic-x column values from the first and last K' = a.colWindow
    // strips' reductions, as the scan kernel's full folds truncated where
    // the dropped terms are below 2^-28: the forward carry out of the last
    // strip (strips P-K' .. P-1, from 0) and the backward value at the top
    // (strips 0 .. K'-1, forward carries from 0 at the top).  KW only sizes
    // the batches of rows in flight.
    template <int CPT, bool VEC, int KW>
    __device__ __forceinline__ void WindowPeriodicCols(const AdiArgs &a, int y0, PeriodicCols<CPT> &pc)
    {
        const size_t plane = static_cast<size_t>(a.P) * a.Y;
        const float *B = a.redIn;
        const float *D = a.redIn + plane;
        const int kk = a.colWindow;
        // rows are requested in batches of KR strips (at most 48 values)
        // before they are folded
        constexpr int KR = (3 * KW * CPT <= 48) ? KW : ((3 * 4 * CPT <= 48) ? 4 : 1);
        float cfEnd[CPT];
        float F[CPT];
        float X[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            cfEnd[c] = 0.0f;
            F[c] = 0.0f;
            X[c] = 0.0f;
        }
        float coef = 1.0f;
        for (int i0 = 0; i0 < kk; i0 += KR)
        {
            float be[KR][CPT]; // B rows P-kk+i (the end)
            float dt[KR][CPT]; // D rows i (the top)
            float bt[KR][CPT]; // B rows i-1 (the top)
#pragma unroll
            for (int j = 0; j < KR; j++)
            {
                const int i = i0 + j;
                if (i < kk)
                {
                    LoadRow<CPT, VEC>(B + static_cast<size_t>(a.P - kk + i) * a.Y, y0, a.Y, be[j]);
                    LoadRow<CPT, VEC>(D + static_cast<size_t>(i) * a.Y, y0, a.Y, dt[j]);
                    if (i > 0)
                    {
                        LoadRow<CPT, VEC>(B + static_cast<size_t>(i - 1) * a.Y, y0, a.Y, bt[j]);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < KR; j++)
            {
                const int i = i0 + j;
                if (i < kk)
                {
                    const float Ae = a.stripA[a.P - kk + i];
#pragma unroll
                    for (int c = 0; c < CPT; c++)
                    {
                        cfEnd[c] = fmaf(Ae, cfEnd[c], be[j][c]);
                    }
                    if (i > 0)
                    {
                        const float At = a.stripA[i - 1];
#pragma unroll
                        for (int c = 0; c < CPT; c++)
                        {
                            F[c] = fmaf(At, F[c], bt[j][c]);
                        }
                    }
                    const float Eq = a.stripE[i];
#pragma unroll
                    for (int c = 0; c < CPT; c++)
                    {
                        X[c] = fmaf(coef, fmaf(Eq, F[c], dt[j][c]), X[c]);
                    }
                    coef *= a.stripC[i];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            const float x0 = X[c];
            const float xLast = a.xibLast * cfEnd[c];
            const float factor = (x0 + xLast * a.xvEnd) * a.xInv;
            pc.factor[c] = factor;
            pc.top[c] = x0 - factor * a.xz0;
            pc.bottom[c] = xLast - factor * a.xzLast;
        }
    }

    // A full window: K (<= KW) strips on both sides of strip p.
    template <int KW>
    __device__ __forceinline__ bool FullWindow(const AdiArgs &a, int p, int K)
    {
        return KW > 0 && K <= KW && p >= K && p + K <= a.P - 1;
    }

    // Whether a full window's 3 KW rows of CPT values fit in registers
    // beside the strip (larger windows fold strip by strip).
    template <int CPT, int KW>
    constexpr bool kWindowInRegisters = KW > 0 && 3 * KW * CPT <= 48;

    // The strip's rows s0 .. s0+S-1 (rows past the grid read as 0).
    template <int CPT, int S, bool VEC>
    __device__ __forceinline__ void LoadStripRows(const AdiArgs &a, const float *__restrict__ u, int s0, int y0,
                                                  float (&v)[S][CPT])
    {
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            if (s0 + r < a.X)
            {
                LoadRow<CPT, VEC>(u + static_cast<size_t>(s0 + r) * a.Y, y0, a.Y, v[r]);
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
    }

    template <int CPT, int KW>
    __device__ __forceinline__ void WindowFold(const AdiArgs &a, int p, const float (&wb)[2 * KW][CPT],
                                               const float (&wd)[KW][CPT], float (&cf)[CPT], float (&xb)[CPT])
    {
        float F[CPT];
        float X[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            F[c] = 0.0f;
            X[c] = 0.0f;
        }
#pragma unroll
        for (int i = 0; i < KW; i++)
        {
            const float A = a.stripA[p - KW + i];
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                F[c] = fmaf(A, F[c], wb[i][c]);
            }
        }
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            cf[c] = F[c];
        }
        float coef = 1.0f;
#pragma unroll
        for (int i = 0; i < KW; i++)
        {
            const int q = p + 1 + i;
            const float A = a.stripA[q - 1];
            const float Eq = a.stripE[q];
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                F[c] = fmaf(A, F[c], wb[KW + i][c]);
                X[c] = fmaf(coef, fmaf(Eq, F[c], wd[i][c]), X[c]);
            }
            coef *= a.stripC[q];
        }
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            xb[c] = X[c];
        }
    }

    // MODE 0: reduction of the current field only (first tick of a run).
    // MODE 1: plain tick in place, then the reduction of u' for the next tick.
    // MODE 2: tick with decay: u' = ADI(u) - k*u (no reduction; the next tick
    //         starts with a MODE 0 pass after point sources are added).
    // MODE 3: tick with decay and point sources kept in registers:
    //         u' = ADI(u) - k*u + sources, then the reduction of u' for the next
    //         tick (batched wells).
    // VEC: yDim is a multiple of CPT, so rows move as 16-byte vectors.
    // One workgroup runs strip p of the grid described by `a`; u is the grid.
    // KW > 0: the carries are composed from the neighbouring strips'
    // reductions in a.redIn (windowed carries, K = a.window <= KW) instead of
    // read from the scan kernel's carry arrays.
    template <int CPT, int S, int MODE, bool VEC, int KW = 0>
    __device__ __forceinline__ void StripBody(const AdiArgs &a, float *__restrict__ u, int p)
    {
        __shared__ float sLeft[kMaxWaves * S];
        __shared__ float sRight[kMaxWaves * S];
        __shared__ float sFirst[S];
        __shared__ float sLast[S];
        __shared__ float sTot[kMaxWaves * S];
        __shared__ float sAw[kMaxWaves];
        __shared__ float sEnds[2 * S];
        const StripLds L{sLeft, sRight, sFirst, sLast, sTot, sAw, sEnds};

        const int s0 = p * S;
        const int tid = threadIdx.x;
        const int y0 = tid * CPT;

        // Windowed carries first: their loads are in flight while the
        // strip's rows load.  With a full window every load is issued before
        // the fold waits for its own (loads complete in order); near the
        // grid's ends the window is folded strip by strip.
        float cf[CPT];
        float dn[CPT];
        float v[S][CPT];
        if constexpr (KW > 0 && MODE != 0 && kWindowInRegisters<CPT, KW>)
        {
            if (FullWindow<KW>(a, p, a.window))
            {
                float wu[KW][CPT];
                float wb[KW][CPT];
                float wd[KW][CPT];
                WindowLoads<CPT, VEC, KW>(a, p, y0, a.window, wu, wb, wd);
                LoadStripRows<CPT, S, VEC>(a, u, s0, y0, v);
                WindowFold<CPT, KW>(a, p, a.window, wu, wb, wd, cf, dn);
            }
            else
            {
                WindowCarries<CPT, VEC, KW>(a, p, y0, a.window, cf, dn);
                LoadStripRows<CPT, S, VEC>(a, u, s0, y0, v);
            }
        }
        else
        {
            if constexpr (KW > 0 && MODE != 0)
            {
                WindowCarries<CPT, VEC, KW>(a, p, y0, a.window, cf, dn);
            }
            LoadStripRows<CPT, S, VEC>(a, u, s0, y0, v);
        }
        float old[MODE == 3 ? S : 1][CPT];
        if constexpr (MODE == 3)
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    old[r][c] = v[r][c];
                }
            }
        }
        SweepARhs<CPT, S, VEC>(a, s0, y0, tid, v, L);
        if constexpr (MODE == 0)
        {
            ReduceA<CPT, S, VEC>(a, p, s0, y0, v);
            return;
        }
        else
        {
            float up[CPT];
            if constexpr (KW > 0)
            {
                FinishACore<CPT, S, VEC>(a, s0, y0, v, cf, up, dn);
            }
            else
            {
                FinishA<CPT, S, VEC>(a, p, s0, y0, v, up, dn);
            }
            SweepBRhs<CPT, S, VEC>(a, s0, y0, tid, v, up, dn);
            SolveB<CPT, S>(a, y0, tid, v, L);
            if constexpr (MODE == 3)
            {
#pragma unroll
                for (int r = 0; r < S; r++)
                {
#pragma unroll
                    for (int c = 0; c < CPT; c++)
                    {
                        v[r][c] = fmaf(old[r][c], a.negDecay, v[r][c]);
                    }
                }
                // Point sources of this strip (PDEGrid2D::Add), in list order.
                for (int k = a.srcBegin; k < a.srcEnd; k++)
                {
                    const int sx = a.srcXY[2 * k] - s0;
                    const int sy = a.srcXY[2 * k + 1] - y0;
                    const float sv = a.srcVal[k];
#pragma unroll
                    for (int r = 0; r < S; r++)
                    {
#pragma unroll
                        for (int c = 0; c < CPT; c++)
                        {
                            v[r][c] += (sx == r && sy == c) ? sv : 0.0f;
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                if (s0 + r < a.X)
                {
                    float *row = u + static_cast<size_t>(s0 + r) * a.Y;
                    if constexpr (MODE == 2)
                    {
                        float old[CPT];
                        LoadRow<CPT, VEC>(row, y0, a.Y, old);
#pragma unroll
                        for (int c = 0; c < CPT; c++)
                        {
                            v[r][c] = fmaf(old[c], a.negDecay, v[r][c]);
                        }
                    }
                    StoreRow<CPT, VEC>(row, y0, a.Y, v[r]);
                }
            }
            if constexpr (MODE == 1 || MODE == 3)
            {
                __syncthreads();
                SweepARhs<CPT, S, VEC>(a, s0, y0, tid, v, L);
                ReduceA<CPT, S, VEC>(a, p, s0, y0, v);
            }
        }
    }

    template <int CPT, int S, int MODE, bool VEC>
    __global__ __launch_bounds__(1024) void AdiStrip(AdiArgs a, float *__restrict__ u)
    {
        StripBody<CPT, S, MODE, VEC>(a, u, blockIdx.x);
    }

    // ---- sweep B through LDS with constant coefficients -----------------------------
    //
    // A non-periodic sweep-B line is M x = d with off-diagonals -r/2,
    // interior diagonal 1+r and end diagonals b_0, b_{n-1}.  Let T be M with
    // its first diagonal replaced by beta = ((1+r) + sqrt(1+2r))/2, the fixed
    // point of Thomas's b'_i = (1+r) - (r/2)^2 / b'_{i-1}, and its last by
    // 1+r: T's Thomas factors are constant (b' = beta, multipliers
    // phi = (r/2)/beta), and M = T + U diag(b_0 - beta, b_{n-1} - 1 - r) U^T
    // with U = [e_0, e_{n-1}].  By the Woodbury identity
    //     x = w - Z K U^T w,   w = T^-1 d,   Z = T^-1 U,
    // where in the scaled variable X = beta * w
    //     X_i = d'_i + phi X_{i+1},   d'_i = d_i + phi d'_{i-1}
    // and Z's columns are closed-form geometric sequences, so
    //     x_i = (X_i - h phi^i - t phi^{n-1-i}) / beta
    // with (h, t) = Q (X_0, X_{n-1}) for a host-computed 2x2 Q.
    //
    // Layout: the strip's S rows sit in LDS; W wavefronts share a row, lane l
    // of wavefront h owns the E consecutive elements of chunk k = 64h + l,
    // stored with a one-float pad per chunk (stride E+1) so that lanes reading
    // element j of their chunks hit distinct banks.
    struct RowSolve
    {
        float phi;        // multiplier of both recurrences
        float ib;         // 1 / beta
        float phiWave;    // phi^(64 E): one wavefront's product
        float q00, q01, q10, q11;
        int32_t W;        // wavefronts per row
        const float *pw;  // [6]: phi^(E 2^k), uniform scan stages
        const float *lane; // [5][64]: phi^(E(l%16+1)), phi^(E(l%32+1)), phi^(E l), phi^(E(63-l)),
                           //        phi^(E(16-l%16))
        const float *chunk; // [2][64 W]: phi^(E k), phi^(n - E k - E)
        const float *powE;  // [E]: phi^j
    };

    template <int E>
    __device__ __forceinline__ int ChunkAddr(int k, int j)
    {
        return k * (E + 1) + j;
    }

    template <int E>
    __device__ __forceinline__ int RowAddr(int y)
    {
        return ChunkAddr<E>(y / E, y % E);
    }

    // Sweep A right-hand side of the strip (v: u in, RHS out) with the row
    // neighbours read from the copy of the same rows staged in LDS.  yDim is
    // non-periodic here: the ghost value at y = -1 / Y is the cell itself
    // (zero flux) or 2 bc - c (Dirichlet).  Only the first and last strips
    // (edge) add the Dirichlet implicit term of rows 0 and X-1 and zero the
    // rows past X.  Same arithmetic as SweepARhs.
    //
    // RhsACore takes the y neighbours of the thread's edge columns, per row:
    // Ln (column y0-1) and Rn (column y0+CPT), with the row's far end on a
    // periodic y axis; at a non-periodic edge they are replaced by ghosts.
    template <int CPT, int S, bool EDGE>
    __device__ __forceinline__ void RhsACore(const AdiArgs &a, int s0, int y0, bool first, bool last,
                                             float (&v)[S][CPT], const float (&Ln)[S], const float (&Rn)[S])
    {
        const bool ghostL = first && !a.wrapY;
        const bool ghostR = last && !a.wrapY;
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const int x = min(s0 + r, a.X - 1);
            float L = Ln[r];
            float R = Rn[r];
            const float c0 = v[r][0];
            const float cl = v[r][CPT - 1];
            const float gl = a.dirichlet ? fmaf(2.0f, a.bcYlo[x], -c0) : c0;
            const float gr = a.dirichlet ? fmaf(2.0f, a.bcYhi[x], -cl) : cl;
            L = ghostL ? gl : L;
            R = ghostR ? gr : R;
            // (nxt - cur) + (prev - cur) = e_c - e_{c-1} with e_c = nxt - cur
            float eLeft = v[r][0] - L;
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                const float cur = v[r][c];
                const float nxt = (c + 1 < CPT) ? v[r][c + 1 < CPT ? c + 1 : c] : R;
                const float e = nxt - cur;
                v[r][c] = fmaf(e - eLeft, a.hr, cur);
                eLeft = e;
            }
        }
        if constexpr (EDGE)
        {
            const bool dirX = a.dirichlet && !a.wrapX;
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                const int x = s0 + r;
                if (x >= a.X)
                {
#pragma unroll
                    for (int c = 0; c < CPT; c++)
                    {
                        v[r][c] = 0.0f;
                    }
                }
                else if (dirX && (x == 0 || x == a.X - 1))
                {
#pragma unroll
                    for (int c = 0; c < CPT; c++)
                    {
                        float src = 0.0f;
                        if (x == 0)
                        {
                            src += a.bcXlo[y0 + c];
                        }
                        if (x == a.X - 1)
                        {
                            src += a.bcXhi[y0 + c];
                        }
                        v[r][c] = fmaf(src, a.r, v[r][c]);
                    }
                }
            }
        }
    }

    // RhsACore with the neighbours read from the rows in LDS.
    template <int CPT, int S, int E, bool EDGE>
    __device__ __forceinline__ void RhsAFromLds(const AdiArgs &a, const float *rows, int rowStride, int s0, int y0,
                                                bool first, bool last, float (&v)[S][CPT])
    {
        const int yl = (y0 > 0) ? y0 - 1 : (a.wrapY ? a.Y - 1 : 0);
        const int yr = (y0 + CPT < a.Y) ? y0 + CPT : (a.wrapY ? 0 : a.Y - 1);
        const int al = RowAddr<E>(yl);
        const int ar = RowAddr<E>(yr);
        float Ln[S];
        float Rn[S];
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            Ln[r] = rows[r * rowStride + al];
            Rn[r] = rows[r * rowStride + ar];
        }
        RhsACore<CPT, S, EDGE>(a, s0, y0, first, last, v, Ln, Rn);
    }

    // RhsACore with the neighbours taken from the adjacent lanes (DPP wave
    // shifts) and, for a wavefront's first and last lane, from the
    // neighbouring wavefronts' edge columns in sFirst / sLast ([wave][S],
    // written before a barrier).  The workgroup holds whole rows, so the
    // first wavefront's left neighbour is the last wavefront's last column
    // (the row's far end), and the other way round.
    template <int CPT, int S, bool EDGE>
    __device__ __forceinline__ void RhsAFromLanes(const AdiArgs &a, const float *sFirst, const float *sLast, int s0,
                                                  int y0, bool first, bool last, float (&v)[S][CPT])
    {
        const int wave = threadIdx.x >> 6;
        const int nW = blockDim.x >> 6;
        const int wl = (wave > 0) ? wave - 1 : nW - 1;
        const int wr = (wave + 1 < nW) ? wave + 1 : 0;
        float Ln[S];
        float Rn[S];
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const float fromLeft = sLast[wl * S + r];
            const float fromRight = sFirst[wr * S + r];
            Ln[r] = Dpp<0x138>(fromLeft, v[r][CPT - 1]);  // lane 0 keeps fromLeft
            Rn[r] = Dpp<0x130>(fromRight, v[r][0]);       // lane 63 keeps fromRight
        }
        RhsACore<CPT, S, EDGE>(a, s0, y0, first, last, v, Ln, Rn);
    }

    // Sweep B right-hand side for a strip away from the x boundaries (rows
    // s0-1 and s0+S exist and are up / dn).  Same arithmetic as SweepBRhs.
    template <int CPT, int S>
    __device__ __forceinline__ void RhsBInterior(const AdiArgs &a, int s0, bool first, bool last, float (&v)[S][CPT],
                                                 const float (&up)[CPT], const float (&dn)[CPT])
    {
        // (nxt - cur) + (prev - cur) = e_r - e_{r-1} with e_r = nxt - cur
        float eUp[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            eUp[c] = v[0][c] - up[c];
        }
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const int x = s0 + r;
            const float srcLo = (a.dirichlet && !a.wrapY && first) ? a.bcYlo[x] * a.r : 0.0f;
            const float srcHi = (a.dirichlet && !a.wrapY && last) ? a.bcYhi[x] * a.r : 0.0f;
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                const float cur = v[r][c];
                const float nxt = (r + 1 < S) ? v[r + 1 < S ? r + 1 : r][c] : dn[c];
                const float e = nxt - cur;
                float rhs = fmaf(e - eUp[c], a.hr, cur);
                if (c == 0)
                {
                    rhs += srcLo;
                }
                if (c == CPT - 1)
                {
                    rhs += srcHi;
                }
                eUp[c] = e;
                v[r][c] = rhs;
            }
        }
    }

    // The strip's two reduction values as dot products of the sweep-A
    // right-hand side with per-row weights (host FP64): B = sum_r wB_r rhs_r
    // (forward end value, zero carry-in) and D = sum_r wD_r rhs_r (backward
    // start value, zero carries), the same linear maps as ReduceA.
    template <int CPT, int S>
    __device__ __forceinline__ void ReduceADot(const AdiArgs &a, int p, int y0, const float (&v)[S][CPT])
    {
        float bAcc[CPT];
        float dAcc[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++)
        {
            bAcc[c] = 0.0f;
            dAcc[c] = 0.0f;
        }
        const float *wB = a.redW + static_cast<size_t>(p) * 2 * S;
#pragma unroll
        for (int r = 0; r < S; r++)
        {
            const float b = wB[r];
            const float d = wB[S + r];
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                bAcc[c] = fmaf(b, v[r][c], bAcc[c]);
                dAcc[c] = fmaf(d, v[r][c], dAcc[c]);
            }
        }
        const size_t base = static_cast<size_t>(p) * a.Y;
        const size_t plane = static_cast<size_t>(a.P) * a.Y;
        StoreRow<CPT, true>(a.red + base, y0, a.Y, bAcc);
        StoreRow<CPT, true>(a.red + plane + base, y0, a.Y, dAcc);
    }

    // KW: 0 takes the carries from the scan kernel, else windowed carries
    // with K <= KW.  EDGE: the strip holds row 0 or row X-1 (or rows past X).
    // PX (with KW > 0): periodic x axis; the strip forms the Sherman-Morrison
    // column values from the first and last strips' reductions.
    // DECAY: the tick is u' = ADI(u) - k*u + sources (a.negDecay; the strip's
    // sources are entries [a.srcStrip[p], a.srcStrip[p+1]) of a.srcXY /
    // a.srcVal, in list order, or none when a.srcStrip is null), and the
    // reduction for the next tick is taken of that u'.
    // rows: blockDim.x * (E+1) floats (S rows x 64W chunks x (E+1)).
    // sFirst / sLast: [waves][S] floats for the edge columns of each
    // wavefront.
    template <int CPT, int S, int E, int KW, bool EDGE, bool DECAY = false, bool PX = false>
    __device__ __forceinline__ void StripLdsBody(const AdiArgs &a, float *__restrict__ u, int p, int K,
                                                 const RowSolve &rs, float *rows, float *sTotF, float *sTotB,
                                                 float *sFirst, float *sLast)
    {

        const int s0 = p * S;
        const int tid = threadIdx.x;
        const int y0 = tid * CPT;
        const int lane = tid & 63;
        const int wave = tid >> 6;
        const int W = rs.W;
        const int rowStride = 64 * W * (E + 1);
        const bool first = tid == 0;
        const bool last = tid == static_cast<int>(blockDim.x) - 1;

        // Periodic x with windowed carries: the column values first, while
        // nothing else is held in registers.
        PeriodicCols<CPT> pc;
        if constexpr (PX)
        {
            WindowPeriodicCols<CPT, true, KW>(a, y0, pc);
        }

        // Interior strips with a full window request the carries' rows and
        // then the strip's rows before waiting for any of them (loads
        // complete in order, so the window fold waits only for its own).
        // Rows are whole (yDim = threads * CPT); interior strips lie inside
        // the grid.
        float cf[CPT];
        float up[CPT];
        float dn[CPT];
        float v[S][CPT];
        const bool fullWindow = kWindowInRegisters<CPT, KW> && FullWindow<KW>(a, p, K);
        if constexpr (KW > 0 && !EDGE && kWindowInRegisters<CPT, KW>)
        {
            if (fullWindow)
            {
                float wu[KW][CPT];
                float wb[KW][CPT];
                float wd[KW][CPT];
                WindowLoads<CPT, true, KW>(a, p, y0, K, wu, wb, wd);
#pragma unroll
                for (int r = 0; r < S; r++)
                {
                    LoadRowFull<CPT>(u + static_cast<size_t>(s0 + r) * a.Y, y0, v[r]);
                }
                WindowFold<CPT, KW>(a, p, K, wu, wb, wd, cf, dn);
            }
        }
        if (!(KW > 0 && !EDGE && kWindowInRegisters<CPT, KW> && fullWindow))
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                if (!EDGE || s0 + r < a.X)
                {
                    LoadRowFull<CPT>(u + static_cast<size_t>(s0 + r) * a.Y, y0, v[r]);
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
        }
        if constexpr (KW > 0)
        {
            if (EDGE || !fullWindow)
            {
                WindowCarries<CPT, true, KW>(a, p, y0, K, cf, dn);
            }
        }
        else
        {
            const size_t base = static_cast<size_t>(p) * a.Y;
            const size_t plane = static_cast<size_t>(a.P) * a.Y;
            LoadRow<CPT, true>(a.carry + base, y0, a.Y, cf);
            LoadRow<CPT, true>(a.carry + plane + base, y0, a.Y, dn);
        }
        // The wavefronts' edge columns for RhsAFromLanes.  They live apart
        // from the rows buffer, so the right-hand sides can go to LDS below
        // without a second barrier.
        if (lane == 0)
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                sFirst[wave * S + r] = v[r][0];
            }
        }
        if (lane == 63)
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                sLast[wave * S + r] = v[r][CPT - 1];
            }
        }
        __syncthreads();
        RhsAFromLanes<CPT, S, EDGE>(a, sFirst, sLast, s0, y0, first, last, v);
        if constexpr (PX)
        {
            // windowed carries on a periodic x axis: no scan kernel wrote
            // colInfo, so the column values come from pc
            FinishACore<CPT, S, true, true>(a, s0, y0, v, cf, up, dn, &pc);
            if constexpr (EDGE)
            {
                SweepBRhs<CPT, S, true, true>(a, s0, y0, tid, v, up, dn, &pc);
            }
            else
            {
                RhsBInterior<CPT, S>(a, s0, first, last, v, up, dn);
            }
        }
        else
        {
            FinishACore<CPT, S, true>(a, s0, y0, v, cf, up, dn);
            if constexpr (EDGE)
            {
                SweepBRhs<CPT, S, true>(a, s0, y0, tid, v, up, dn);
            }
            else
            {
                RhsBInterior<CPT, S>(a, s0, first, last, v, up, dn);
            }
        }

        // Right-hand sides to LDS, row-major by chunk.
#pragma unroll
        for (int r = 0; r < S; r++)
        {
#pragma unroll
            for (int c = 0; c < CPT; c++)
            {
                rows[r * rowStride + RowAddr<E>(y0 + c)] = v[r][c];
            }
        }
        __syncthreads();

        // Row phase: wavefront `wave` solves part h of row `row`.
        const int row = wave / W;
        const int h = wave - row * W;
        const int k = h * 64 + lane;
        const bool active = row < S && s0 + row < a.X;
        float *line = rows + row * rowStride;
        const float phi = rs.phi;
        float d[E];
        float b = 0.0f;
        if (active)
        {
#pragma unroll
            for (int j = 0; j < E; j++)
            {
                d[j] = line[ChunkAddr<E>(k, j)];
                b = fmaf(phi, b, d[j]);
            }
            // inclusive scan over the wavefront's lanes
            b = fmaf(rs.pw[0], Dpp<0x111>(0.0f, b), b);
            b = fmaf(rs.pw[1], Dpp<0x112>(0.0f, b), b);
            b = fmaf(rs.pw[2], Dpp<0x114>(0.0f, b), b);
            b = fmaf(rs.pw[3], Dpp<0x118>(0.0f, b), b);
            b = fmaf(rs.lane[lane], Dpp<0x142, 0xA>(0.0f, b), b);
            b = fmaf(rs.lane[64 + lane], Dpp<0x143, 0xC>(0.0f, b), b);
            if (lane == 63)
            {
                sTotF[wave] = b;
            }
        }
        __syncthreads();
        float xLast = 0.0f; // d'_{n-1} = X_{n-1}
        if (active)
        {
            float cw = 0.0f;
            for (int g = 0; g < W; g++)
            {
                const float t = sTotF[row * W + g];
                cw = (g < h) ? fmaf(rs.phiWave, cw, t) : cw;
                xLast = fmaf(rs.phiWave, xLast, t);
            }
            float acc = fmaf(rs.lane[128 + lane], cw, Dpp<0x138>(0.0f, b));
#pragma unroll
            for (int j = 0; j < E; j++)
            {
                acc = fmaf(phi, acc, d[j]);
                d[j] = acc;
            }
            b = 0.0f;
#pragma unroll
            for (int j = E - 1; j >= 0; j--)
            {
                b = fmaf(phi, b, d[j]);
            }
            // inclusive suffix scan over the wavefront's lanes: within each
            // row of 16 lanes by DPP row_shl, then the row totals (lanes 16,
            // 32, 48) folded in from the rows below
            b = fmaf(rs.pw[0], Dpp<0x101>(0.0f, b), b);
            b = fmaf(rs.pw[1], Dpp<0x102>(0.0f, b), b);
            b = fmaf(rs.pw[2], Dpp<0x104>(0.0f, b), b);
            b = fmaf(rs.pw[3], Dpp<0x108>(0.0f, b), b);
            {
                const float t1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, b), 16));
                const float t2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, b), 32));
                const float t3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, b), 48));
                const float s2 = t3;
                const float s1 = fmaf(rs.pw[4], s2, t2);
                const float s0 = fmaf(rs.pw[4], s1, t1);
                const int rw = lane >> 4;
                const float below = (rw == 0) ? s0 : ((rw == 1) ? s1 : ((rw == 2) ? s2 : 0.0f));
                b = fmaf(rs.lane[256 + lane], below, b);
            }
            if (lane == 0)
            {
                sTotB[wave] = b;
            }
        }
        __syncthreads();
        if (active)
        {
            float cw = 0.0f;
            float x0 = 0.0f;
            for (int g = W - 1; g >= 0; g--)
            {
                const float t = sTotB[row * W + g];
                cw = (g > h) ? fmaf(rs.phiWave, cw, t) : cw;
                x0 = fmaf(rs.phiWave, x0, t);
            }
            const float hc = fmaf(rs.q00, x0, rs.q01 * xLast);
            const float tc = fmaf(rs.q10, x0, rs.q11 * xLast);
            const float hl = -rs.ib * hc * rs.chunk[k];
            const float tl = -rs.ib * tc * rs.chunk[64 * W + k];
            float acc = fmaf(rs.lane[192 + lane], cw, Dpp<0x130>(0.0f, b));
#pragma unroll
            for (int j = E - 1; j >= 0; j--)
            {
                acc = fmaf(phi, acc, d[j]);
                line[ChunkAddr<E>(k, j)] = fmaf(hl, rs.powE[j], fmaf(tl, rs.powE[E - 1 - j], rs.ib * acc));
            }
        }
        __syncthreads();

        // Column phase: u' from LDS, store, and next tick's strip reduction.
        if constexpr (DECAY)
        {
            // u' = ADI(u) - k*u from the rows still in HBM, then the strip's
            // point sources (PDEGrid2D::MulAll + Add), as the reference sums
            // its deltas
            float old[S][CPT];
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                if (s0 + r < a.X)
                {
                    LoadRow<CPT, true>(u + static_cast<size_t>(s0 + r) * a.Y, y0, a.Y, old[r]);
                }
                else
                {
#pragma unroll
                    for (int c = 0; c < CPT; c++)
                    {
                        old[r][c] = 0.0f;
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < S; r++)
            {
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    v[r][c] = (s0 + r < a.X) ? fmaf(old[r][c], a.negDecay, rows[r * rowStride + RowAddr<E>(y0 + c)])
                                             : 0.0f;
                }
            }
            if (a.srcStrip != nullptr)
            {
                const int kEnd = a.srcStrip[p + 1];
                for (int k = a.srcStrip[p]; k < kEnd; k++)
                {
                    const int sx = a.srcXY[2 * k] - s0;
                    const int sy = a.srcXY[2 * k + 1] - y0;
                    const float sv = a.srcVal[k];
#pragma unroll
                    for (int r = 0; r < S; r++)
                    {
#pragma unroll
                        for (int c = 0; c < CPT; c++)
                        {
                            v[r][c] += (sx == r && sy == c) ? sv : 0.0f;
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                if (s0 + r < a.X)
                {
                    StoreRow<CPT, true>(u + static_cast<size_t>(s0 + r) * a.Y, y0, a.Y, v[r]);
                }
            }
            // RhsAFromLds reads the neighbours' edge columns of u' from LDS;
            // each thread rewrites its own edge columns (the rest of the row
            // in LDS is not read again)
#pragma unroll
            for (int r = 0; r < S; r++)
            {
                rows[r * rowStride + RowAddr<E>(y0)] = v[r][0];
                if constexpr (CPT > 1)
                {
                    rows[r * rowStride + RowAddr<E>(y0 + CPT - 1)] = v[r][CPT - 1];
                }
            }
            __syncthreads();
        }
        else
        {
#pragma unroll
            for (int r = 0; r < S; r++)
            {
#pragma unroll
                for (int c = 0; c < CPT; c++)
                {
                    v[r][c] = (s0 + r < a.X) ? rows[r * rowStride + RowAddr<E>(y0 + c)] : 0.0f;
                }
                if (s0 + r < a.X)
                {
                    StoreRow<CPT, true>(u + static_cast<size_t>(s0 + r) * a.Y, y0, a.Y, v[r]);
                }
            }
        }
        RhsAFromLds<CPT, S, E, EDGE>(a, rows, rowStride, s0, y0, first, last, v);
        ReduceADot<CPT, S>(a, p, y0, v);
    }

    // Strip index of workgroup b: with windowed carries the strips of one
    // contiguous range go to one XCD (workgroups are dealt to the 8 XCDs in
    // turn), so the reductions a strip reads from its neighbours are mostly
    // in that XCD's L2.
    __device__ __forceinline__ int StripOfBlock(int b, int P, bool xcdMajor)
    {
        if (!xcdMajor)
        {
            return b;
        }
        const int per = P / 8;
        return (b % 8) * per + b / 8;
    }

    // T: threads per workgroup (= yDim / CPT); the LDS row buffer holds
    // T * (E+1) floats.
    template <int CPT, int S, int E, int KW, int T, bool DECAY = false, bool PX = false>
    __global__ __launch_bounds__(T) void AdiStripLds(AdiArgs a, float *__restrict__ u, int32_t K, RowSolve rs)
    {
        static_assert(T * (E + 1) * 4 <= 140 * 1024, "LDS row buffer too large");
        __shared__ float rows[T * (E + 1)];
        __shared__ float sTotF[kMaxWaves];
        __shared__ float sTotB[kMaxWaves];
        __shared__ float sFirst[kMaxWaves * S];
        __shared__ float sLast[kMaxWaves * S];
        const int p = StripOfBlock(blockIdx.x, a.P, KW > 0 && (a.P % 8 == 0));
        if (p == 0 || (p + 1) * S >= a.X)
        {
            StripLdsBody<CPT, S, E, KW, true, DECAY, PX>(a, u, p, K, rs, rows, sTotF, sTotB, sFirst, sLast);
        }
        else
        {
            StripLdsBody<CPT, S, E, KW, false, DECAY, PX>(a, u, p, K, rs, rows, sTotF, sTotB, sFirst, sLast);
        }
    }

    // Scan of the strip reductions along each column (sweep A across strips).
    // Block: COLS columns x SEGS segments; thread t takes column
    // y0 + (t % COLS) and segment g = t / COLS, which owns strips
    // [g*Q, (g+1)*Q) with Q <= QMAX.  A thread first loads its segment's
    // reduction values into registers (all loads in flight together), folds
    // them, exchanges segment totals through LDS, then runs the recurrences
    // from registers.  Writes the carry-in values of every strip and, per
    // column, the periodic Sherman-Morrison factor and the wrapped rows of
    // u*.
    constexpr int kScanCols = 16;
    constexpr int kScanSegs = 16;
    // the single grid's scan: 16 columns x 64 segments (1024 threads), so a
    // 4096-wide grid fills 256 workgroups and each thread walks P/64 strips
    constexpr int kScanSegsWide = 64;

    template <int QMAX, int COLS = kScanCols, int SEGS = kScanSegs>
    __device__ __forceinline__ void ScanBody(const AdiArgs &a, int yBase, int32_t Q)
    {
        __shared__ float sB[SEGS][COLS];
        __shared__ float sA[SEGS];
        __shared__ float sB2[SEGS][COLS];
        __shared__ float sA2[SEGS];
        __shared__ float sCf[COLS];
        __shared__ float sX0[COLS];
        const int col = threadIdx.x & (COLS - 1);
        const int g = threadIdx.x / COLS;
        const int y = yBase + col;
        const bool active = y < a.Y;
        const int p0 = g * Q;
        const int cnt = max(0, min(a.P, p0 + Q) - p0);
        const size_t plane = static_cast<size_t>(a.P) * a.Y;
        const float *B = a.red;
        const float *D = a.red + plane;
        float *cfIn = a.carry;
        float *xbIn = a.carry + plane;

        float bv[QMAX];
        float dv[QMAX];
        float cv[QMAX];
#pragma unroll
        for (int q = 0; q < QMAX; q++)
        {
            const bool ok = active && q < cnt;
            const size_t k = ok ? static_cast<size_t>(p0 + q) * a.Y + y : 0;
            bv[q] = ok ? B[k] : 0.0f;
            dv[q] = ok ? D[k] : 0.0f;
        }
        float acc = 0.0f;
        float aSeg = 1.0f;
#pragma unroll
        for (int q = 0; q < QMAX; q++)
        {
            const float A = (q < cnt) ? a.stripA[p0 + q] : 1.0f;
            acc = fmaf(A, acc, bv[q]);
            aSeg *= A;
        }
        sB[g][col] = acc;
        if (col == 0)
        {
            sA[g] = aSeg;
        }
        __syncthreads();
        float cf = 0.0f;
        for (int w = 0; w < g; w++)
        {
            cf = fmaf(sA[w], cf, sB[w][col]);
        }
#pragma unroll
        for (int q = 0; q < QMAX; q++)
        {
            cv[q] = cf;
            const float A = (q < cnt) ? a.stripA[p0 + q] : 1.0f;
            cf = fmaf(A, cf, bv[q]);
        }
        if (cnt > 0 && p0 + cnt == a.P)
        {
            sCf[col] = cf;
        }
        acc = 0.0f;
        aSeg = 1.0f;
#pragma unroll
        for (int q = QMAX - 1; q >= 0; q--)
        {
            const bool in = q < cnt;
            const float C = in ? a.stripC[p0 + q] : 1.0f;
            const float off = in ? fmaf(a.stripE[p0 + q], cv[q], dv[q]) : 0.0f;
            acc = fmaf(C, acc, off);
            aSeg *= C;
        }
        sB2[g][col] = acc;
        if (col == 0)
        {
            sA2[g] = aSeg;
        }
        __syncthreads();
        float xb = 0.0f;
        for (int w = SEGS - 1; w > g; w--)
        {
            xb = fmaf(sA2[w], xb, sB2[w][col]);
        }
#pragma unroll
        for (int q = QMAX - 1; q >= 0; q--)
        {
            const bool in = q < cnt;
            if (in && active)
            {
                const size_t k = static_cast<size_t>(p0 + q) * a.Y + y;
                xbIn[k] = xb;
                cfIn[k] = cv[q];
            }
            const float C = in ? a.stripC[p0 + q] : 1.0f;
            const float off = in ? fmaf(a.stripE[p0 + q], cv[q], dv[q]) : 0.0f;
            xb = fmaf(C, xb, off);
        }
        if (g == 0)
        {
            sX0[col] = xb;
        }
        __syncthreads();
        if (g == 0 && active)
        {
            const float x0 = sX0[col];
            const float xLast = a.xibLast * sCf[col];
            const float factor = (x0 + xLast * a.xvEnd) * a.xInv;
            a.colInfo[y] = factor;
            a.colInfo[a.Y + y] = x0 - factor * a.xz0;
            a.colInfo[2 * static_cast<size_t>(a.Y) + y] = xLast - factor * a.xzLast;
        }
    }

    template <int QMAX>
    __global__ __launch_bounds__(kScanCols *kScanSegsWide) void AdiStripScan(AdiArgs a, int32_t Q)
    {
        ScanBody<QMAX, kScanCols, kScanSegsWide>(a, blockIdx.x * kScanCols, Q);
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

    // Smallest K <= kMaxWindow such that every product of K consecutive
    // strip multipliers, forward (A) and backward (C), that a windowed carry
    // drops is at most 2^-28; 0 when there is none.  Windows reaching past the
    // first or last strip drop nothing.
    int32_t WindowSize(const std::vector<double> &A, const std::vector<double> &C)
    {
        const double tau = std::ldexp(1.0, -28);
        const int32_t P = static_cast<int32_t>(A.size());
        for (int32_t k = 1; k <= kMaxWindow; k++)
        {
            bool ok = true;
            for (int32_t p = 0; p < P && ok; p++)
            {
                if (p - k >= 0)
                {
                    double prod = 1.0;
                    for (int32_t q = p - k; q < p; q++)
                    {
                        prod *= std::fabs(A[static_cast<size_t>(q)]);
                    }
                    ok = prod <= tau;
                }
                if (ok && p + k <= P - 1)
                {
                    double prod = 1.0;
                    for (int32_t q = p + 1; q <= p + k; q++)
                    {
                        prod *= std::fabs(C[static_cast<size_t>(q)]);
                    }
                    ok = prod <= tau;
                }
            }
            if (ok)
            {
                return k;
            }
        }
        return 0;
    }

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

}
