// diffusion2d.hip -- explicit 5-point diffusion with fused decay, point
// sources and commit, for gfx950.
//
// Replaces the reference's two-dispatch Vulkan step (Diffusion2D.comp +
// DiffusionUpdate2D.comp, src/Gpu/DiffusionKernel2D.cpp:244-304) with one
// pass per tick: each lane reads its cell and four neighbours and writes the
// committed value, so a tick moves 8 B/cell of HBM traffic (read u, write u')
// instead of ~28.  Lanes run along y, the contiguous axis of HAL's x*yDim+y
// layout, so every wavefront row access is a coalesced 256-byte segment; the
// x-neighbour rows are re-read from L2.
//
// Semantics follow src/Pde/PDEequations.cpp:140-161 (zero flux / periodic),
// include/Hal/Pde/PDEequations.h:50-98, 264-287 (Dirichlet ghost
// (bc - c)*2), src/Pde/PDEequations.cpp:199-228 (per-cell face rates) and the
// PDEGrid2D tick Diffusion -> Advection -> MulAll(-k) -> Add -> Update
// (include/Hal/Grids/PDEGrid2D.h:103-218, 289-375).  First-order upwind
// advection: uniform velocity src/Pde/PDEequations.cpp:380-413 and
// include/Hal/Pde/PDEequations.h:413-447; per-cell velocities in flux form
// src/Pde/PDEequations.cpp:535-573 and include/Hal/Pde/PDEequations.h:721-783.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <string>

using namespace halhip;

namespace
{
    constexpr int kBlockY = 64;
    constexpr int kBlockX = 4;
    constexpr int kInlineSources = 64; // sources scanned per lane; above this a dense map is used

    struct ExplicitArgs
    {
        int32_t X;
        int32_t Y;
        float rate;
        float xVel; // uniform advection velocity (0: none)
        float yVel;
        float negDecay;
        int32_t wrapX;
        int32_t wrapY;
        int32_t dirichlet;
        const float *bcXlo; // [Y] value at x = -1
        const float *bcXhi; // [Y] value at x = X
        const float *bcYlo; // [X] value at y = -1
        const float *bcYhi; // [X] value at y = Y
        int32_t nSources;   // inline list (<= kInlineSources) ...
        const int64_t *srcIdx;
        const float *srcVal;
        const float *srcDense; // ... or a dense per-cell map (null if unused)
    };

    // Difference to the neighbour at (nx, ny) of the cell (x, y) with value
    // c, reading along one axis; off-grid handling per the axis mode.
    __device__ __forceinline__ float NeighbourDiffX(const float *__restrict__ u, const ExplicitArgs &a, int nx, int y,
                                                    float c)
    {
        if (nx >= 0 && nx < a.X)
        {
            return u[static_cast<size_t>(nx) * a.Y + y] - c;
        }
        if (a.wrapX)
        {
            int w = (nx < 0) ? nx + a.X : nx - a.X;
            return u[static_cast<size_t>(w) * a.Y + y] - c;
        }
        if (a.dirichlet)
        {
            float bc = (nx < 0) ? a.bcXlo[y] : a.bcXhi[y];
            return (bc - c) * 2.0f;
        }
        return 0.0f;
    }

    __device__ __forceinline__ float NeighbourDiffY(const float *__restrict__ row, const ExplicitArgs &a, int x,
                                                    int ny, float c)
    {
        if (ny >= 0 && ny < a.Y)
        {
            return row[ny] - c;
        }
        if (a.wrapY)
        {
            int w = (ny < 0) ? ny + a.Y : ny - a.Y;
            return row[w] - c;
        }
        if (a.dirichlet)
        {
            float bc = (ny < 0) ? a.bcYlo[x] : a.bcYhi[x];
            return (bc - c) * 2.0f;
        }
        return 0.0f;
    }

    // Upwind advection term from the four neighbour differences (d+x, d-x,
    // d+y, d-y) of a cell: x axis then y axis, as Advection2 adds them.
    __device__ __forceinline__ float AdvectTerm(float delta, float xVel, float yVel, float dxp, float dxm, float dyp,
                                                float dym)
    {
        if (xVel > 0.0f)
        {
            delta += dxm * xVel;
        }
        else if (xVel < 0.0f)
        {
            delta -= dxp * xVel;
        }
        if (yVel > 0.0f)
        {
            delta += dym * yVel;
        }
        else if (yVel < 0.0f)
        {
            delta -= dyp * yVel;
        }
        return delta;
    }

    // One tick.  TICK: out = u + (rate*lap + advection + negDecay*u + src);
    // otherwise out = rate*lap + advection (the interface deltas).
    template <bool TICK>
    __global__ __launch_bounds__(kBlockY *kBlockX) void ExplicitStep(ExplicitArgs a, const float *__restrict__ u,
                                                                      float *__restrict__ out)
    {
        const int y = blockIdx.x * kBlockY + threadIdx.x;
        const int x = blockIdx.y * kBlockX + threadIdx.y;
        if (x >= a.X || y >= a.Y)
        {
            return;
        }
        const size_t i = static_cast<size_t>(x) * a.Y + y;
        const float *row = u + static_cast<size_t>(x) * a.Y;
        const float c = row[y];
        const float dxp = NeighbourDiffX(u, a, x + 1, y, c);
        const float dxm = NeighbourDiffX(u, a, x - 1, y, c);
        const float dyp = NeighbourDiffY(row, a, x, y + 1, c);
        const float dym = NeighbourDiffY(row, a, x, y - 1, c);
        float s = dxp;
        s += dxm;
        s += dyp;
        s += dym;
        float delta = s * a.rate;
        delta = AdvectTerm(delta, a.xVel, a.yVel, dxp, dxm, dyp, dym);
        if (!TICK)
        {
            out[i] = delta;
            return;
        }
        delta += c * a.negDecay;
        if (a.srcDense != nullptr)
        {
            delta += a.srcDense[i];
        }
        else
        {
            for (int k = 0; k < a.nSources; k++)
            {
                if (a.srcIdx[k] == static_cast<int64_t>(i))
                {
                    delta += a.srcVal[k];
                }
            }
        }
        out[i] = c + delta;
    }

    // ---- several ticks per launch (temporal blocking) --------------------------
    //
    // Workgroup g owns rows [x0, x0+B), x0 = g*B, and holds rows
    // [x0-T, x0+B+T) of the field in LDS (two copies, ping-pong); rows wrap
    // on a periodic x axis and do not exist past a non-wrapped edge.  Tick t
    // updates local rows [t, B+2T-t): every row it reads was updated at tick
    // t-1 or is past the grid edge, where the boundary rule replaces it.
    // After T ticks rows [T, T+B) are exact and are stored.  The arithmetic
    // is ExplicitStep's, term for term, so results are bit-identical to T
    // single-tick launches.
    constexpr int kMultiThreads = 1024;
    constexpr int kMultiRowsPerThread = 8; // rows of the held B + 2T one thread covers

    // Threads: thread `tid` works on column y = tid % Y of rows
    // tid / Y, tid / Y + R, ... (R = 1024 / Y rows per pass; Y <= 1024), so
    // its y neighbours and y-edge rules are fixed for the whole launch.
    template <bool WX, bool WY, bool DIR>
    __global__ __launch_bounds__(kMultiThreads) void ExplicitMulti(ExplicitArgs a, int32_t B, int32_t T,
                                                                   const float *__restrict__ in,
                                                                   float *__restrict__ out)
    {
        extern __shared__ float lds[];
        __shared__ int sRow[64]; // global row of each local row, -1 past a non-wrapped edge
        const int rowsHeld = B + 2 * T;
        float *buf[2] = {lds, lds + static_cast<size_t>(rowsHeld) * a.Y};
        const int x0 = blockIdx.x * B;
        if (threadIdx.x < rowsHeld)
        {
            int gx = x0 - T + static_cast<int>(threadIdx.x);
            if (WX)
            {
                gx %= a.X;
                gx = gx < 0 ? gx + a.X : gx;
            }
            else if (gx < 0 || gx >= a.X)
            {
                gx = -1;
            }
            sRow[threadIdx.x] = gx;
        }
        __syncthreads();
        const int rowsPerPass = kMultiThreads / a.Y;
        const int y = static_cast<int>(threadIdx.x) % a.Y;
        const int r0 = static_cast<int>(threadIdx.x) / a.Y;
        const bool lane = r0 < rowsPerPass;
        // y neighbours: in-row index, or an edge rule (zero flux: difference
        // 0; Dirichlet: (bc - c) * 2 with bc at (x, -1) / (x, Y))
        const bool hiIn = WY || y + 1 < a.Y;
        const bool loIn = WY || y > 0;
        const int yHi = (y + 1 < a.Y) ? y + 1 : 0;
        const int yLo = (y > 0) ? y - 1 : a.Y - 1;
        const float bcXlo = DIR && !WX ? a.bcXlo[y] : 0.0f;
        const float bcXhi = DIR && !WX ? a.bcXhi[y] : 0.0f;
        // Source term of each of this thread's rows (list order), gathered
        // once per launch: row j of the thread is local row r0 + j*R.
        float srcAdd[kMultiRowsPerThread];
        if (lane)
        {
#pragma unroll
            for (int j = 0; j < kMultiRowsPerThread; j++)
            {
                const int l = r0 + j * rowsPerPass;
                srcAdd[j] = 0.0f;
                const int gx = (l < rowsHeld) ? sRow[l] : -1;
                if (gx >= 0)
                {
                    const size_t i = static_cast<size_t>(gx) * a.Y + y;
                    if (a.srcDense != nullptr)
                    {
                        srcAdd[j] = a.srcDense[i];
                    }
                    else
                    {
                        for (int k = 0; k < a.nSources; k++)
                        {
                            if (a.srcIdx[k] == static_cast<int64_t>(i))
                            {
                                srcAdd[j] += a.srcVal[k];
                            }
                        }
                    }
                }
            }
            for (int l = r0; l < rowsHeld; l += rowsPerPass)
            {
                const int gx = sRow[l];
                buf[0][l * a.Y + y] = (gx >= 0) ? in[static_cast<size_t>(gx) * a.Y + y] : 0.0f;
            }
        }
        __syncthreads();
        int cur = 0;
        for (int t = 1; t <= T; t++)
        {
            const float *src = buf[cur];
            float *dst = buf[1 - cur];
            if (lane)
            {
#pragma unroll
                for (int j = 0; j < kMultiRowsPerThread; j++)
                {
                    const int l = r0 + j * rowsPerPass;
                    if (l < t || l >= rowsHeld - t)
                    {
                        continue;
                    }
                    const int gx = sRow[l];
                    if (gx < 0)
                    {
                        continue;
                    }
                    const float *row = src + l * a.Y;
                    const float c = row[y];
                    // x+1, x-1, y+1, y-1 in ExplicitStep's order
                    const float dxp = (WX || gx + 1 < a.X) ? row[a.Y + y] - c : (DIR ? (bcXhi - c) * 2.0f : 0.0f);
                    const float dxm = (WX || gx > 0) ? row[y - a.Y] - c : (DIR ? (bcXlo - c) * 2.0f : 0.0f);
                    const float dyp = hiIn ? row[yHi] - c : (DIR ? (a.bcYhi[gx] - c) * 2.0f : 0.0f);
                    const float dym = loIn ? row[yLo] - c : (DIR ? (a.bcYlo[gx] - c) * 2.0f : 0.0f);
                    float s = dxp;
                    s += dxm;
                    s += dyp;
                    s += dym;
                    float delta = s * a.rate;
                    delta = AdvectTerm(delta, a.xVel, a.yVel, dxp, dxm, dyp, dym);
                    delta += c * a.negDecay;
                    delta += srcAdd[j];
                    dst[l * a.Y + y] = c + delta;
                }
            }
            __syncthreads();
            cur = 1 - cur;
        }
        if (lane)
        {
            for (int l = T + r0; l < T + B; l += rowsPerPass)
            {
                const int gx = x0 + (l - T);
                if (gx < a.X)
                {
                    out[static_cast<size_t>(gx) * a.Y + y] = buf[cur][l * a.Y + y];
                }
            }
        }
    }

    // Per-cell-velocity upwind advection delta (flux form).  Off-grid field
    // values and velocities are 0 without edge arrays (bc / vbc null), the
    // edge values otherwise; vbc's x edges hold the x velocity at x = -1 / X,
    // its y edges the y velocity at y = -1 / Y.
    struct AdvectCellsArgs
    {
        int32_t X;
        int32_t Y;
        int32_t wrapX;
        int32_t wrapY;
        const float *bcXlo, *bcXhi, *bcYlo, *bcYhi;
        const float *vXlo, *vXhi, *vYlo, *vYhi;
    };

    __device__ __forceinline__ float ShiftX(const float *v, const AdvectCellsArgs &a, int x, int y, const float *lo,
                                            const float *hi)
    {
        if (x >= 0 && x < a.X)
        {
            return v[static_cast<size_t>(x) * a.Y + y];
        }
        if (a.wrapX)
        {
            return v[static_cast<size_t>(x < 0 ? x + a.X : x - a.X) * a.Y + y];
        }
        return lo != nullptr ? (x < 0 ? lo[y] : hi[y]) : 0.0f;
    }

    __device__ __forceinline__ float ShiftY(const float *v, const AdvectCellsArgs &a, int x, int y, const float *lo,
                                            const float *hi)
    {
        if (y >= 0 && y < a.Y)
        {
            return v[static_cast<size_t>(x) * a.Y + y];
        }
        if (a.wrapY)
        {
            return v[static_cast<size_t>(x) * a.Y + (y < 0 ? y + a.Y : y - a.Y)];
        }
        return lo != nullptr ? (y < 0 ? lo[x] : hi[x]) : 0.0f;
    }

    __global__ __launch_bounds__(kBlockY *kBlockX) void AdvectCellsDelta(AdvectCellsArgs a,
                                                                          const float *__restrict__ u,
                                                                          const float *__restrict__ xv,
                                                                          const float *__restrict__ yv,
                                                                          float *__restrict__ out)
    {
        const int y = blockIdx.x * kBlockY + threadIdx.x;
        const int x = blockIdx.y * kBlockX + threadIdx.y;
        if (x >= a.X || y >= a.Y)
        {
            return;
        }
        const size_t i = static_cast<size_t>(x) * a.Y + y;
        const float xVelP = (x == a.X - 1 && !a.wrapX) ? ShiftX(xv, a, x + 1, y, a.vXlo, a.vXhi) : xv[i];
        const float yVelP = (y == a.Y - 1 && !a.wrapY) ? ShiftY(yv, a, x, y + 1, a.vYlo, a.vYhi) : yv[i];
        const float xVelM = ShiftX(xv, a, x - 1, y, a.vXlo, a.vXhi);
        const float yVelM = ShiftY(yv, a, x, y - 1, a.vYlo, a.vYhi);
        const float c = u[i];
        float xFluxP = 0.0f;
        float xFluxM = 0.0f;
        float yFluxP = 0.0f;
        float yFluxM = 0.0f;
        if (xVelP > 0.0f)
        {
            xFluxP = xVelP * c;
        }
        else if (xVelP < 0.0f)
        {
            xFluxP = xVelP * ShiftX(u, a, x + 1, y, a.bcXlo, a.bcXhi);
        }
        if (xVelM > 0.0f)
        {
            xFluxM = xVelM * ShiftX(u, a, x - 1, y, a.bcXlo, a.bcXhi);
        }
        else if (xVelM < 0.0f)
        {
            xFluxM = xVelM * c;
        }
        if (yVelP > 0.0f)
        {
            yFluxP = yVelP * c;
        }
        else if (yVelP < 0.0f)
        {
            yFluxP = yVelP * ShiftY(u, a, x, y + 1, a.bcYlo, a.bcYhi);
        }
        if (yVelM > 0.0f)
        {
            yFluxM = yVelM * ShiftY(u, a, x, y - 1, a.bcYlo, a.bcYhi);
        }
        else if (yVelM < 0.0f)
        {
            yFluxM = yVelM * c;
        }
        out[i] = -xFluxP + xFluxM - yFluxP + yFluxM;
    }

    // Per-cell face rates: +x/+y face uses this cell's rate (0 past a
    // non-wrapped edge, where the neighbour difference is 0 anyway), -x/-y
    // face the neighbour's.  src/Pde/PDEequations.cpp:199-228.
    __global__ __launch_bounds__(kBlockY *kBlockX) void ExplicitRatesDelta(int32_t X, int32_t Y, int32_t wrapX,
                                                                            int32_t wrapY,
                                                                            const float *__restrict__ u,
                                                                            const float *__restrict__ rx,
                                                                            const float *__restrict__ ry,
                                                                            float *__restrict__ out)
    {
        const int y = blockIdx.x * kBlockY + threadIdx.x;
        const int x = blockIdx.y * kBlockX + threadIdx.y;
        if (x >= X || y >= Y)
        {
            return;
        }
        const size_t i = static_cast<size_t>(x) * Y + y;
        const float c = u[i];
        auto at = [&](const float *f, int xx, int yy, bool &ok) -> float
        {
            ok = true;
            if (xx < 0 || xx >= X)
            {
                if (!wrapX)
                {
                    ok = false;
                    return 0.0f;
                }
                xx = (xx < 0) ? xx + X : xx - X;
            }
            if (yy < 0 || yy >= Y)
            {
                if (!wrapY)
                {
                    ok = false;
                    return 0.0f;
                }
                yy = (yy < 0) ? yy + Y : yy - Y;
            }
            return f[static_cast<size_t>(xx) * Y + yy];
        };
        bool ok;
        float s = 0.0f;
        float v = at(u, x + 1, y, ok);
        s += ok ? (v - c) * rx[i] : 0.0f;
        v = at(u, x - 1, y, ok);
        float r = at(rx, x - 1, y, ok);
        s += ok ? (v - c) * r : 0.0f;
        v = at(u, x, y + 1, ok);
        s += ok ? (v - c) * ry[i] : 0.0f;
        v = at(u, x, y - 1, ok);
        r = at(ry, x, y - 1, ok);
        s += ok ? (v - c) * r : 0.0f;
        out[i] = s;
    }
}

namespace
{
    // ---- PDEGrid2D delta operations on the resident field ------------------------
    //
    // include/Hal/Grids/PDEGrid2D.h:89-245: every call reads the committed
    // field u and changes the deltas d; Update commits u += d, d = 0.

    __global__ __launch_bounds__(256) void DeltaAddInto(float *__restrict__ d, const float *__restrict__ s, int64_t n)
    {
        for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        {
            d[i] += s[i];
        }
    }

    // Field-wide ops: 0 SetAll (d = v - u), 1 AddAll (d += v), 2 MulAll
    // (d += u v), 3 SetAll from values (d = vals - u), 4 Update (u += d,
    // d = 0), 5 SetNonNegative (d = -u where u + d < 0), 6 clear (d = 0).
    __global__ __launch_bounds__(256) void DeltaFill(int32_t op, float v, float *__restrict__ u, float *__restrict__ d,
                                                     const float *__restrict__ vals, int64_t n)
    {
        for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        {
            switch (op)
            {
                case 0:
                    d[i] = v - u[i];
                    break;
                case 1:
                    d[i] += v;
                    break;
                case 2:
                    d[i] += u[i] * v;
                    break;
                case 3:
                    d[i] = vals[i] - u[i];
                    break;
                case 4:
                    u[i] += d[i];
                    d[i] = 0.0f;
                    break;
                case 5:
                    if (u[i] + d[i] < 0.0f)
                    {
                        d[i] = 0.0f - u[i];
                    }
                    break;
                default:
                    d[i] = 0.0f;
                    break;
            }
        }
    }

    // Point ops in list order (one lane, so repeated cells apply in order):
    // 0 Set (d = v - u), 1 Add (d += v), 2 Mul (d += u v).
    __global__ void DeltaPointOps(int32_t count, const int32_t *__restrict__ op, const int64_t *__restrict__ idx,
                                  const float *__restrict__ val, const float *__restrict__ u, float *__restrict__ d)
    {
        if (threadIdx.x != 0 || blockIdx.x != 0)
        {
            return;
        }
        for (int32_t k = 0; k < count; k++)
        {
            const int64_t i = idx[k];
            switch (op[k])
            {
                case 0:
                    d[i] = val[k] - u[i];
                    break;
                case 1:
                    d[i] += val[k];
                    break;
                default:
                    d[i] += u[i] * val[k];
                    break;
            }
        }
    }

    // Per-block partials: 0 max |d|, 1 max |d| / (|u| + offset), 2 max u,
    // 3 min u, 4 sum of u (FP64).
    __global__ __launch_bounds__(256) void DeltaReduce(int32_t mode, float offset, const float *__restrict__ u,
                                                       const float *__restrict__ d, int64_t n,
                                                       double *__restrict__ partial)
    {
        __shared__ double red[256];
        double acc = (mode == 2) ? -INFINITY : ((mode == 3) ? INFINITY : 0.0);
        for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        {
            switch (mode)
            {
                case 0:
                    acc = fmax(acc, static_cast<double>(fabsf(d[i])));
                    break;
                case 1:
                    acc = fmax(acc, static_cast<double>(fabsf(d[i] / (fabsf(u[i]) + offset))));
                    break;
                case 2:
                    acc = fmax(acc, static_cast<double>(u[i]));
                    break;
                case 3:
                    acc = fmin(acc, static_cast<double>(u[i]));
                    break;
                default:
                    acc += static_cast<double>(u[i]);
                    break;
            }
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1)
        {
            if (static_cast<int>(threadIdx.x) < o)
            {
                const double b = red[threadIdx.x + o];
                const double a = red[threadIdx.x];
                red[threadIdx.x] = (mode == 4) ? a + b : ((mode == 3) ? fmin(a, b) : fmax(a, b));
            }
            __syncthreads();
        }
        if (threadIdx.x == 0)
        {
            partial[blockIdx.x] = red[0];
        }
    }

    constexpr unsigned kDeltaBlocks = 1024; // grid of the field-wide delta kernels
}

struct halhip_diffusion2d
{
    halhip_context *ctx = nullptr;
    int32_t X = 0;
    int32_t Y = 0;
    DeviceArray<float> field[2]; // ping-pong resident field
    int cur = 0;
    DeviceArray<float> aux[2];   // interface scratch (field copy, delta)
    DeviceArray<float> cellA;    // per-cell inputs (rates, velocities, values), first use
    DeviceArray<float> cellB;
    DeviceArray<float> cellEdges; // field and velocity edges of per-cell advection
    DeviceArray<float> deltas;   // PDEGrid2D deltas of the resident field, first use
    DeviceArray<int32_t> pointOp; // point-op lists
    DeviceArray<int64_t> pointIdx;
    DeviceArray<float> pointVal;
    DeviceArray<double> partial; // reduction partials
    DeviceArray<float> bc;       // 2Y + 2X Dirichlet edges
    DeviceArray<int64_t> srcIdx;
    DeviceArray<float> srcVal;
    DeviceArray<float> srcDense;
    int32_t nSources = 0;
    double xVel = 0.0; // uniform advection of the resident ticks
    double yVel = 0.0;
    int32_t multiB = 0; // pinned multi-tick shape (0: automatic)
    int32_t multiT = 0;
    EdgeValues edges;
    std::string edgeKey;
    StreamClock clock;
    Profiler prof;
    GraphCache graphs;
    bool prepareOnly = false; // inside halhip_diffusion2d_prepare_steps
    HostStaging<float> hostBuf; // pinned staging of the per-call FP64 paths
    ChunkedTransfer chunks;
};

namespace
{
    size_t Cells(const halhip_diffusion2d *p)
    {
        return static_cast<size_t>(p->X) * static_cast<size_t>(p->Y);
    }

    void RequirePlan(const halhip_diffusion2d *p, const char *caller)
    {
        Require(p != nullptr, std::string(caller) + ": null plan");
        Bind(p->ctx);
    }

    // Uploads Dirichlet edges when they changed and fills the kernel args.
    ExplicitArgs MakeArgs(halhip_diffusion2d *p, double rate, double decay, int32_t wrapX, int32_t wrapY,
                          const halhip_bc *bc, double xVel = 0.0, double yVel = 0.0)
    {
        EdgeValues e;
        e.From(bc, p->X, p->Y);
        std::string key;
        e.AppendTo(key);
        if (e.dirichlet && key != p->edgeKey)
        {
            std::vector<float> packed;
            packed.insert(packed.end(), e.xlo.begin(), e.xlo.end());
            packed.insert(packed.end(), e.xhi.begin(), e.xhi.end());
            packed.insert(packed.end(), e.ylo.begin(), e.ylo.end());
            packed.insert(packed.end(), e.yhi.begin(), e.yhi.end());
            Check(hipMemcpyAsync(p->bc.get(), packed.data(), packed.size() * sizeof(float), hipMemcpyHostToDevice,
                                 p->clock.stream),
                  "hipMemcpyAsync");
            Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
            p->edgeKey = key;
        }
        ExplicitArgs a{};
        a.X = p->X;
        a.Y = p->Y;
        a.rate = static_cast<float>(rate);
        a.xVel = static_cast<float>(xVel);
        a.yVel = static_cast<float>(yVel);
        a.negDecay = static_cast<float>(-decay);
        a.wrapX = wrapX ? 1 : 0;
        a.wrapY = wrapY ? 1 : 0;
        a.dirichlet = e.dirichlet ? 1 : 0;
        a.bcXlo = p->bc.get();
        a.bcXhi = p->bc.get() + p->Y;
        a.bcYlo = p->bc.get() + 2 * p->Y;
        a.bcYhi = p->bc.get() + 2 * p->Y + p->X;
        a.nSources = (p->srcDense.get() != nullptr) ? 0 : p->nSources;
        a.srcIdx = p->srcIdx.get();
        a.srcVal = p->srcVal.get();
        a.srcDense = p->srcDense.get();
        return a;
    }

    dim3 Grid(const halhip_diffusion2d *p)
    {
        return dim3(static_cast<unsigned>((p->Y + kBlockY - 1) / kBlockY),
                    static_cast<unsigned>((p->X + kBlockX - 1) / kBlockX));
    }

    // Device buffer b with at least n elements, allocated on first use.
    template <class T>
    T *Lazy(halhip_diffusion2d *p, DeviceArray<T> &b, size_t n)
    {
        if (b.get() == nullptr || b.size() < n)
        {
            b.Allocate(p->ctx, n);
        }
        return b.get();
    }

    // Pinned staging of at least n floats, once the plan's stream is idle
    // (earlier calls' copies out of it have finished).
    float *Stage(halhip_diffusion2d *p, size_t n)
    {
        Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
        p->hostBuf.resize(n);
        return p->hostBuf.data();
    }

    // dst[0, n) = float(src[0, n)) through staging[0, n), enqueued.
    void Upload64(halhip_diffusion2d *p, const double *src, size_t n, float *staging, float *dst)
    {
        for (size_t k = 0; k < n; k++)
        {
            staging[k] = static_cast<float>(src[k]);
        }
        Check(hipMemcpyAsync(dst, staging, n * sizeof(float), hipMemcpyHostToDevice, p->clock.stream),
              "hipMemcpyAsync");
    }

    // out = per-cell-rate diffusion delta of u (rates already on the device).
    void LaunchRates(halhip_diffusion2d *p, const float *u, const float *rx, const float *ry, float *out,
                     int32_t wrapX, int32_t wrapY)
    {
        p->prof.Begin(p->clock.stream, "explicit_rates_delta");
        hipLaunchKernelGGL(ExplicitRatesDelta, Grid(p), dim3(kBlockY, kBlockX), 0, p->clock.stream, p->X, p->Y,
                           wrapX ? 1 : 0, wrapY ? 1 : 0, u, rx, ry, out);
        Check(hipGetLastError(), "ExplicitRatesDelta launch");
        p->prof.End(p->clock.stream);
    }

    // Edge arrays of per-cell advection uploaded to cellEdges (field edges,
    // then velocity edges) and the kernel's arguments.  Enqueued copies read
    // `staging` (2 * (2X + 2Y) floats).
    AdvectCellsArgs CellAdvectArgs(halhip_diffusion2d *p, int32_t wrapX, int32_t wrapY, const halhip_bc *bc,
                                   const halhip_bc *velBc, float *staging)
    {
        EdgeValues e;
        e.From(bc, p->X, p->Y);
        EdgeValues v;
        v.From(velBc, p->X, p->Y);
        Require(!v.dirichlet || e.dirichlet, "advect_cells: velocity edges need field edges (Dirichlet bc)");
        const size_t edgeLen = 2 * static_cast<size_t>(p->X) + 2 * static_cast<size_t>(p->Y);
        AdvectCellsArgs a{};
        a.X = p->X;
        a.Y = p->Y;
        a.wrapX = wrapX ? 1 : 0;
        a.wrapY = wrapY ? 1 : 0;
        if (!e.dirichlet)
        {
            return a;
        }
        float *eb = Lazy(p, p->cellEdges, 2 * edgeLen);
        size_t at = 0;
        for (const EdgeValues *ev : {&e, &v})
        {
            for (const std::vector<float> *edge : {&ev->xlo, &ev->xhi, &ev->ylo, &ev->yhi})
            {
                std::copy(edge->begin(), edge->end(), staging + at);
                at += edge->size();
            }
        }
        Check(hipMemcpyAsync(eb, staging, at * sizeof(float), hipMemcpyHostToDevice, p->clock.stream),
              "hipMemcpyAsync");
        a.bcXlo = eb;
        a.bcXhi = eb + p->Y;
        a.bcYlo = eb + 2 * p->Y;
        a.bcYhi = eb + 2 * p->Y + p->X;
        if (v.dirichlet)
        {
            float *vb = eb + edgeLen;
            a.vXlo = vb;
            a.vXhi = vb + p->Y;
            a.vYlo = vb + 2 * p->Y;
            a.vYhi = vb + 2 * p->Y + p->X;
        }
        return a;
    }

    void LaunchAdvectCells(halhip_diffusion2d *p, const AdvectCellsArgs &a, const float *u, const float *xv,
                           const float *yv, float *out)
    {
        p->prof.Begin(p->clock.stream, "advect_cells_delta");
        hipLaunchKernelGGL(AdvectCellsDelta, Grid(p), dim3(kBlockY, kBlockX), 0, p->clock.stream, a, u, xv, yv, out);
        Check(hipGetLastError(), "AdvectCellsDelta launch");
        p->prof.End(p->clock.stream);
    }

    // The deltas of the resident field (zero on first use).
    float *Deltas(halhip_diffusion2d *p)
    {
        if (p->deltas.get() == nullptr)
        {
            p->deltas.Allocate(p->ctx, Cells(p));
            Check(hipMemsetAsync(p->deltas.get(), 0, Cells(p) * sizeof(float), p->clock.stream), "hipMemsetAsync");
        }
        return p->deltas.get();
    }

    void LaunchFill(halhip_diffusion2d *p, int32_t op, float v, const float *vals)
    {
        hipLaunchKernelGGL(DeltaFill, dim3(kDeltaBlocks), dim3(256), 0, p->clock.stream, op, v,
                           p->field[p->cur].get(), Deltas(p), vals, static_cast<int64_t>(Cells(p)));
        Check(hipGetLastError(), "DeltaFill launch");
    }

    // deltas += scratch (a delta computed into aux[1])
    void AddScratch(halhip_diffusion2d *p)
    {
        hipLaunchKernelGGL(DeltaAddInto, dim3(kDeltaBlocks), dim3(256), 0, p->clock.stream, Deltas(p),
                           static_cast<const float *>(p->aux[1].get()), static_cast<int64_t>(Cells(p)));
        Check(hipGetLastError(), "DeltaAddInto launch");
    }

    // Rows per workgroup B and ticks per launch T of the multi-tick kernel
    // (both 0 when yDim is too long for its LDS rows).
    struct MultiShape
    {
        int32_t B = 0;
        int32_t T = 0;
    };

    constexpr size_t kMultiLds = 64 * 1024; // dynamic LDS per workgroup

    // Whether (B, T) fits the multi-tick kernel for rows of Y cells: the
    // B + 2T held rows, twice, in kMultiLds bytes, at most
    // kMultiRowsPerThread rows per thread and 64 held rows.
    bool MultiFits(int32_t Y, const MultiShape &c)
    {
        if (Y > kMultiThreads || c.B < 1 || c.T < 1)
        {
            return false;
        }
        const int32_t rowsPerPass = kMultiThreads / Y;
        const int32_t held = c.B + 2 * c.T;
        const size_t bytes = 2 * static_cast<size_t>(held) * static_cast<size_t>(Y) * sizeof(float);
        return held <= 64 && bytes <= kMultiLds && (held + rowsPerPass - 1) / rowsPerPass <= kMultiRowsPerThread;
    }

    MultiShape ChooseMulti(const halhip_diffusion2d *p)
    {
        if (p->multiB > 0)
        {
            return MultiShape{p->multiB, p->multiT};
        }
        // 8 ticks per launch over the tallest workgroups (B <= 4 rows) that
        // still give one workgroup per CU: redundant halo work is (B + T) / B,
        // and at 256^2 spreading over all 256 CUs wins (1 x 8: 1.85 us/tick,
        // 4 x 8: 1.97); then fewer ticks where the rows do not fit.
        const int32_t b = std::max(1, std::min(4, p->X / 256));
        for (const MultiShape c : {MultiShape{b, 8}, MultiShape{4, 8}, MultiShape{4, 4}, MultiShape{2, 2}})
        {
            if (MultiFits(p->Y, c))
            {
                return c;
            }
        }
        return MultiShape{};
    }

    void LaunchMulti(halhip_diffusion2d *p, const ExplicitArgs &a, const MultiShape &m, int32_t ticks,
                     const float *in, float *out)
    {
        const size_t lds = 2 * static_cast<size_t>(m.B + 2 * ticks) * static_cast<size_t>(p->Y) * sizeof(float);
        auto kernel = a.wrapX ? (a.wrapY ? (a.dirichlet ? ExplicitMulti<true, true, true> : ExplicitMulti<true, true, false>)
                                         : (a.dirichlet ? ExplicitMulti<true, false, true> : ExplicitMulti<true, false, false>))
                              : (a.wrapY ? (a.dirichlet ? ExplicitMulti<false, true, true> : ExplicitMulti<false, true, false>)
                                         : (a.dirichlet ? ExplicitMulti<false, false, true> : ExplicitMulti<false, false, false>));
        p->prof.Begin(p->clock.stream, "explicit_multi");
        hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>((p->X + m.B - 1) / m.B)), dim3(kMultiThreads), lds,
                           p->clock.stream, a, m.B, ticks, in, out);
        Check(hipGetLastError(), "ExplicitMulti launch");
        p->prof.End(p->clock.stream);
    }

    void LaunchTick(halhip_diffusion2d *p, const ExplicitArgs &a, const float *in, float *out)
    {
        p->prof.Begin(p->clock.stream, "explicit_tick");
        hipLaunchKernelGGL(ExplicitStep<true>, Grid(p), dim3(kBlockY, kBlockX), 0, p->clock.stream, a, in, out);
        Check(hipGetLastError(), "ExplicitStep launch");
        p->prof.End(p->clock.stream);
    }

    // Enqueues `steps` ticks; full chunks of ticks replay a cached graph
    // (an even launch count keeps the ping-pong parity, which is part of
    // the key), the rest launch directly.  With profiling on, every tick is
    // launched directly so each gets its own event pair.  In a prepare call
    // the chunk is captured for both parities and nothing is enqueued.
    void EnqueueTicks(halhip_diffusion2d *p, int32_t steps, double rate, double decay, int32_t wrapX, int32_t wrapY,
                      const halhip_bc *bc)
    {
        Require(steps >= 0, "halhip_diffusion2d_run_steps: negative steps");
        ExplicitArgs a = MakeArgs(p, rate, decay, wrapX, wrapY, bc, p->xVel, p->yVel);
        const MultiShape m = ChooseMulti(p);
        // Graph chunks of six multi-tick launches of T ticks each, or of 50
        // single-tick launches.
        const int32_t launches = m.T > 0 ? 6 : 50;
        const int32_t chunk = m.T > 0 ? launches * m.T : launches;
        auto launchAll = [&](int c0, int32_t count)
        {
            int c = c0;
            for (int32_t k = 0; k < count; k++)
            {
                if (m.T > 0)
                {
                    LaunchMulti(p, a, m, m.T, p->field[c].get(), p->field[1 - c].get());
                }
                else
                {
                    LaunchTick(p, a, p->field[c].get(), p->field[1 - c].get());
                }
                c = 1 - c;
            }
        };
        auto chunkGraph = [&](int c0)
        {
            std::string key;
            AppendKey(key, a);
            AppendKey(key, m);
            AppendKey(key, c0);
            return FindOrCapture(p->graphs, key, p->clock.stream, [&] { launchAll(c0, launches); });
        };
        int32_t done = 0;
        if (!p->prof.On() && steps >= chunk)
        {
            if (p->prepareOnly)
            {
                chunkGraph(p->cur);
                chunkGraph(1 - p->cur);
                return;
            }
            hipGraphExec_t exec = chunkGraph(p->cur);
            for (; done + chunk <= steps; done += chunk)
            {
                Check(hipGraphLaunch(exec, p->clock.stream), "hipGraphLaunch");
            }
        }
        if (p->prepareOnly)
        {
            return;
        }
        while (done < steps)
        {
            const int32_t t = m.T > 0 ? std::min(m.T, steps - done) : 1;
            if (m.T > 0)
            {
                LaunchMulti(p, a, m, t, p->field[p->cur].get(), p->field[1 - p->cur].get());
            }
            else
            {
                LaunchTick(p, a, p->field[p->cur].get(), p->field[1 - p->cur].get());
            }
            p->cur = 1 - p->cur;
            done += t;
        }
    }
}

extern "C"
{
    int halhip_diffusion2d_create(halhip_context *ctx, int32_t xDim, int32_t yDim, halhip_diffusion2d **out)
    {
        halhip_diffusion2d *p = nullptr;
        int rc = Guard(
            [&]
            {
                Require(out != nullptr, "halhip_diffusion2d_create: null out");
                *out = nullptr;
                Require(ctx != nullptr, "halhip_diffusion2d_create: null context");
                Require(xDim > 0 && yDim > 0, "halhip_diffusion2d_create: dimensions must be positive");
                Require(static_cast<int64_t>(xDim) * yDim <= (int64_t(1) << 31),
                        "halhip_diffusion2d_create: grid larger than 2^31 cells");
                Bind(ctx);
                p = new halhip_diffusion2d();
                p->ctx = ctx;
                p->X = xDim;
                p->Y = yDim;
                size_t n = Cells(p);
                p->field[0].Allocate(ctx, n);
                p->field[1].Allocate(ctx, n);
                p->aux[0].Allocate(ctx, n);
                p->bc.Allocate(ctx, 2 * static_cast<size_t>(xDim) + 2 * static_cast<size_t>(yDim));
                p->clock.Create();
                Check(hipMemsetAsync(p->field[0].get(), 0, n * sizeof(float), p->clock.stream), "hipMemset");
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

    int halhip_diffusion2d_destroy(halhip_diffusion2d *p)
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

    int halhip_diffusion2d_diffuse_f64(halhip_diffusion2d *p, const double *field, double *deltas, int64_t n,
                                       double diffRate, int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_diffuse_f64");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_diffusion2d_diffuse_f64");
                Require(field != nullptr && deltas != nullptr, "halhip_diffusion2d_diffuse_f64: null array");
                size_t cells = Cells(p);
                p->hostBuf.resize(cells);
                ExplicitArgs a = MakeArgs(p, diffRate, 0.0, wrapX, wrapY, bc);
                if (p->aux[1].get() == nullptr)
                {
                    p->aux[1].Allocate(p->ctx, cells);
                }
                p->chunks.Upload(field, cells, p->hostBuf.data(), p->aux[0].get(), p->clock.stream);
                p->prof.Begin(p->clock.stream, "explicit_delta");
                hipLaunchKernelGGL(ExplicitStep<false>, Grid(p), dim3(kBlockY, kBlockX), 0, p->clock.stream, a,
                                   p->aux[0].get(), p->aux[1].get());
                Check(hipGetLastError(), "ExplicitStep launch");
                p->prof.End(p->clock.stream);
                p->chunks.DownloadAdd(p->aux[1].get(), cells, p->hostBuf.data(), deltas, nullptr, p->clock.stream);
            });
    }

    int halhip_diffusion2d_diffuse_rates_f64(halhip_diffusion2d *p, const double *field, double *deltas,
                                             const double *ratesX, const double *ratesY, int64_t n, int32_t wrapX,
                                             int32_t wrapY)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_diffuse_rates_f64");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_diffusion2d_diffuse_rates_f64");
                Require(field != nullptr && deltas != nullptr && ratesX != nullptr && ratesY != nullptr,
                        "halhip_diffusion2d_diffuse_rates_f64: null array");
                const size_t cells = Cells(p);
                // field -> aux[0], rates -> cellA / cellB (persistent: the
                // resident field must survive), delta -> aux[1]
                float *stage = Stage(p, 3 * cells);
                float *rx = Lazy(p, p->cellA, cells);
                float *ry = Lazy(p, p->cellB, cells);
                Lazy(p, p->aux[1], cells);
                Upload64(p, field, cells, stage, p->aux[0].get());
                Upload64(p, ratesX, cells, stage + cells, rx);
                Upload64(p, ratesY, cells, stage + 2 * cells, ry);
                LaunchRates(p, p->aux[0].get(), rx, ry, p->aux[1].get(), wrapX, wrapY);
                p->chunks.DownloadAdd(p->aux[1].get(), cells, stage, deltas, nullptr, p->clock.stream);
            });
    }

    int halhip_diffusion2d_advect_f64(halhip_diffusion2d *p, const double *field, double *deltas, int64_t n,
                                      double xVel, double yVel, int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_advect_f64");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_diffusion2d_advect_f64");
                Require(field != nullptr && deltas != nullptr, "halhip_diffusion2d_advect_f64: null array");
                Require(std::isfinite(xVel) && std::isfinite(yVel), "halhip_diffusion2d_advect_f64: bad velocity");
                const size_t cells = Cells(p);
                p->hostBuf.resize(cells);
                for (size_t k = 0; k < cells; k++)
                {
                    p->hostBuf[k] = static_cast<float>(field[k]);
                }
                Check(hipMemcpyAsync(p->aux[0].get(), p->hostBuf.data(), cells * sizeof(float),
                                     hipMemcpyHostToDevice, p->clock.stream),
                      "hipMemcpyAsync");
                // rate 0: the kernel's delta is the advection term alone
                ExplicitArgs a = MakeArgs(p, 0.0, 0.0, wrapX, wrapY, bc, xVel, yVel);
                if (p->aux[1].get() == nullptr)
                {
                    p->aux[1].Allocate(p->ctx, cells);
                }
                p->prof.Begin(p->clock.stream, "advect_delta");
                hipLaunchKernelGGL(ExplicitStep<false>, Grid(p), dim3(kBlockY, kBlockX), 0, p->clock.stream, a,
                                   p->aux[0].get(), p->aux[1].get());
                Check(hipGetLastError(), "ExplicitStep launch");
                p->prof.End(p->clock.stream);
                Check(hipMemcpyAsync(p->hostBuf.data(), p->aux[1].get(), cells * sizeof(float),
                                     hipMemcpyDeviceToHost, p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                for (size_t k = 0; k < cells; k++)
                {
                    deltas[k] += static_cast<double>(p->hostBuf[k]);
                }
            });
    }

    int halhip_diffusion2d_advect_cells_f64(halhip_diffusion2d *p, const double *field, double *deltas,
                                            const double *xVels, const double *yVels, int64_t n, int32_t wrapX,
                                            int32_t wrapY, const halhip_bc *bc, const halhip_bc *velBc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_advect_cells_f64");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_diffusion2d_advect_cells_f64");
                Require(field != nullptr && deltas != nullptr && xVels != nullptr && yVels != nullptr,
                        "halhip_diffusion2d_advect_cells_f64: null array");
                const size_t cells = Cells(p);
                const size_t edges = 4 * (static_cast<size_t>(p->X) + static_cast<size_t>(p->Y));
                float *stage = Stage(p, 3 * cells + edges);
                float *xv = Lazy(p, p->cellA, cells);
                float *yv = Lazy(p, p->cellB, cells);
                Lazy(p, p->aux[1], cells);
                const AdvectCellsArgs a = CellAdvectArgs(p, wrapX, wrapY, bc, velBc, stage + 3 * cells);
                Upload64(p, field, cells, stage, p->aux[0].get());
                Upload64(p, xVels, cells, stage + cells, xv);
                Upload64(p, yVels, cells, stage + 2 * cells, yv);
                LaunchAdvectCells(p, a, p->aux[0].get(), xv, yv, p->aux[1].get());
                p->chunks.DownloadAdd(p->aux[1].get(), cells, stage, deltas, nullptr, p->clock.stream);
            });
    }

    // ---- PDEGrid2D deltas of the resident field ------------------------------------

    int halhip_diffusion2d_delta_point_ops(halhip_diffusion2d *p, int32_t count, const int32_t *ops,
                                           const int64_t *index, const double *values)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_delta_point_ops");
                Require(count >= 0, "halhip_diffusion2d_delta_point_ops: negative count");
                if (count == 0)
                {
                    return;
                }
                Require(ops != nullptr && index != nullptr && values != nullptr,
                        "halhip_diffusion2d_delta_point_ops: null array");
                const int64_t cells = static_cast<int64_t>(Cells(p));
                std::vector<int32_t> op(static_cast<size_t>(count));
                std::vector<float> val(static_cast<size_t>(count));
                for (int32_t k = 0; k < count; k++)
                {
                    Require(index[k] >= 0 && index[k] < cells, "halhip_diffusion2d_delta_point_ops: index out of range");
                    Require(ops[k] >= HALHIP_DELTA_SET && ops[k] <= HALHIP_DELTA_SCALE,
                            "halhip_diffusion2d_delta_point_ops: unknown op");
                    // Scale(v) adds u (v - 1): a Mul by v - 1, formed in FP64
                    op[k] = (ops[k] == HALHIP_DELTA_SCALE) ? HALHIP_DELTA_MUL : ops[k];
                    val[k] = static_cast<float>((ops[k] == HALHIP_DELTA_SCALE) ? values[k] - 1.0 : values[k]);
                }
                Deltas(p);
                // the lists' buffers may still be read by the previous call's kernel
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                int32_t *dOp = Lazy(p, p->pointOp, static_cast<size_t>(count));
                int64_t *dIdx = Lazy(p, p->pointIdx, static_cast<size_t>(count));
                float *dVal = Lazy(p, p->pointVal, static_cast<size_t>(count));
                Check(hipMemcpy(dOp, op.data(), op.size() * sizeof(int32_t), hipMemcpyHostToDevice), "hipMemcpy");
                Check(hipMemcpy(dIdx, index, static_cast<size_t>(count) * sizeof(int64_t), hipMemcpyHostToDevice),
                      "hipMemcpy");
                Check(hipMemcpy(dVal, val.data(), val.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
                hipLaunchKernelGGL(DeltaPointOps, dim3(1), dim3(64), 0, p->clock.stream, count,
                                   static_cast<const int32_t *>(dOp), static_cast<const int64_t *>(dIdx),
                                   static_cast<const float *>(dVal), static_cast<const float *>(p->field[p->cur].get()),
                                   p->deltas.get());
                Check(hipGetLastError(), "DeltaPointOps launch");
            });
    }

    int halhip_diffusion2d_delta_all(halhip_diffusion2d *p, int32_t op, double value)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_delta_all");
                Require(std::isfinite(value), "halhip_diffusion2d_delta_all: value must be finite");
                switch (op)
                {
                    case HALHIP_DELTA_SET:
                        LaunchFill(p, 0, static_cast<float>(value), nullptr);
                        break;
                    case HALHIP_DELTA_ADD:
                        LaunchFill(p, 1, static_cast<float>(value), nullptr);
                        break;
                    case HALHIP_DELTA_MUL:
                        LaunchFill(p, 2, static_cast<float>(value), nullptr);
                        break;
                    case HALHIP_DELTA_SCALE:
                        LaunchFill(p, 2, static_cast<float>(value - 1.0), nullptr);
                        break;
                    case HALHIP_DELTA_SET_NON_NEGATIVE:
                        LaunchFill(p, 5, 0.0f, nullptr);
                        break;
                    default:
                        Fail(HALHIP_ERR_INVALID_ARGUMENT, "halhip_diffusion2d_delta_all: unknown op");
                }
            });
    }

    int halhip_diffusion2d_delta_set_values(halhip_diffusion2d *p, const double *values, int64_t n)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_delta_set_values");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_diffusion2d_delta_set_values");
                Require(values != nullptr, "halhip_diffusion2d_delta_set_values: null array");
                const size_t cells = Cells(p);
                float *stage = Stage(p, cells);
                float *v = Lazy(p, p->cellA, cells);
                Upload64(p, values, cells, stage, v);
                LaunchFill(p, 3, 0.0f, v);
            });
    }

    int halhip_diffusion2d_delta_diffuse(halhip_diffusion2d *p, double diffRate, int32_t wrapX, int32_t wrapY,
                                         const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_delta_diffuse");
                Require(std::isfinite(diffRate), "halhip_diffusion2d_delta_diffuse: rate must be finite");
                ExplicitArgs a = MakeArgs(p, diffRate, 0.0, wrapX, wrapY, bc);
                Lazy(p, p->aux[1], Cells(p));
                p->prof.Begin(p->clock.stream, "explicit_delta");
                hipLaunchKernelGGL(ExplicitStep<false>, Grid(p), dim3(kBlockY, kBlockX), 0, p->clock.stream, a,
                                   static_cast<const float *>(p->field[p->cur].get()), p->aux[1].get());
                Check(hipGetLastError(), "ExplicitStep launch");
                p->prof.End(p->clock.stream);
                AddScratch(p);
            });
    }

    int halhip_diffusion2d_delta_advect(halhip_diffusion2d *p, double xVel, double yVel, int32_t wrapX, int32_t wrapY,
                                        const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_delta_advect");
                Require(std::isfinite(xVel) && std::isfinite(yVel), "halhip_diffusion2d_delta_advect: bad velocity");
                // rate 0: the kernel's delta is the advection term alone
                ExplicitArgs a = MakeArgs(p, 0.0, 0.0, wrapX, wrapY, bc, xVel, yVel);
                Lazy(p, p->aux[1], Cells(p));
                p->prof.Begin(p->clock.stream, "advect_delta");
                hipLaunchKernelGGL(ExplicitStep<false>, Grid(p), dim3(kBlockY, kBlockX), 0, p->clock.stream, a,
                                   static_cast<const float *>(p->field[p->cur].get()), p->aux[1].get());
                Check(hipGetLastError(), "ExplicitStep launch");
                p->prof.End(p->clock.stream);
                AddScratch(p);
            });
    }

    int halhip_diffusion2d_delta_diffuse_rates(halhip_diffusion2d *p, const double *ratesX, const double *ratesY,
                                               int64_t n, int32_t wrapX, int32_t wrapY)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_delta_diffuse_rates");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_diffusion2d_delta_diffuse_rates");
                Require(ratesX != nullptr && ratesY != nullptr, "halhip_diffusion2d_delta_diffuse_rates: null array");
                const size_t cells = Cells(p);
                float *stage = Stage(p, 2 * cells);
                float *rx = Lazy(p, p->cellA, cells);
                float *ry = Lazy(p, p->cellB, cells);
                Lazy(p, p->aux[1], cells);
                Upload64(p, ratesX, cells, stage, rx);
                Upload64(p, ratesY, cells, stage + cells, ry);
                LaunchRates(p, p->field[p->cur].get(), rx, ry, p->aux[1].get(), wrapX, wrapY);
                AddScratch(p);
            });
    }

    int halhip_diffusion2d_delta_advect_cells(halhip_diffusion2d *p, const double *xVels, const double *yVels,
                                              int64_t n, int32_t wrapX, int32_t wrapY, const halhip_bc *bc,
                                              const halhip_bc *velBc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_delta_advect_cells");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_diffusion2d_delta_advect_cells");
                Require(xVels != nullptr && yVels != nullptr, "halhip_diffusion2d_delta_advect_cells: null array");
                const size_t cells = Cells(p);
                const size_t edges = 4 * (static_cast<size_t>(p->X) + static_cast<size_t>(p->Y));
                float *stage = Stage(p, 2 * cells + edges);
                float *xv = Lazy(p, p->cellA, cells);
                float *yv = Lazy(p, p->cellB, cells);
                Lazy(p, p->aux[1], cells);
                const AdvectCellsArgs a = CellAdvectArgs(p, wrapX, wrapY, bc, velBc, stage + 2 * cells);
                Upload64(p, xVels, cells, stage, xv);
                Upload64(p, yVels, cells, stage + cells, yv);
                LaunchAdvectCells(p, a, p->field[p->cur].get(), xv, yv, p->aux[1].get());
                AddScratch(p);
            });
    }

    int halhip_diffusion2d_delta_adi(halhip_diffusion2d *p, halhip_adi2d *adi, double diffRate, int32_t wrapX,
                                     int32_t wrapY, const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_delta_adi");
                AdiDeltaOf(adi, p->field[p->cur].get(), Deltas(p), p->X, p->Y, diffRate, wrapX, wrapY, bc,
                           p->clock.stream);
            });
    }

    int halhip_diffusion2d_delta_reduce(halhip_diffusion2d *p, int32_t what, double offset, double *out)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_delta_reduce");
                Require(out != nullptr, "halhip_diffusion2d_delta_reduce: null output");
                Require(what >= HALHIP_REDUCE_MAX_DELTA && what <= HALHIP_REDUCE_FIELD_SUM,
                        "halhip_diffusion2d_delta_reduce: unknown reduction");
                Require(std::isfinite(offset), "halhip_diffusion2d_delta_reduce: offset must be finite");
                double *part = Lazy(p, p->partial, kDeltaBlocks);
                hipLaunchKernelGGL(DeltaReduce, dim3(kDeltaBlocks), dim3(256), 0, p->clock.stream, what,
                                   static_cast<float>(offset), static_cast<const float *>(p->field[p->cur].get()),
                                   static_cast<const float *>(Deltas(p)), static_cast<int64_t>(Cells(p)), part);
                Check(hipGetLastError(), "DeltaReduce launch");
                std::vector<double> host(kDeltaBlocks);
                Check(hipMemcpyAsync(host.data(), part, kDeltaBlocks * sizeof(double), hipMemcpyDeviceToHost,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                double r = host[0];
                for (size_t k = 1; k < host.size(); k++)
                {
                    r = (what == HALHIP_REDUCE_FIELD_SUM) ? r + host[k]
                                                          : ((what == HALHIP_REDUCE_FIELD_MIN) ? std::min(r, host[k])
                                                                                               : std::max(r, host[k]));
                }
                *out = r;
            });
    }

    int halhip_diffusion2d_delta_update(halhip_diffusion2d *p)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_delta_update");
                LaunchFill(p, 4, 0.0f, nullptr);
            });
    }

    int halhip_diffusion2d_download_deltas(halhip_diffusion2d *p, float *deltas, int64_t n)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_download_deltas");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_diffusion2d_download_deltas");
                Require(deltas != nullptr, "halhip_diffusion2d_download_deltas: null array");
                Check(hipMemcpyAsync(deltas, Deltas(p), Cells(p) * sizeof(float), hipMemcpyDeviceToHost,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
            });
    }

    int halhip_diffusion2d_set_multi_shape(halhip_diffusion2d *p, int32_t rowsPerGroup, int32_t ticksPerLaunch)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_set_multi_shape");
                const bool automatic = rowsPerGroup == 0 && ticksPerLaunch == 0;
                Require(automatic || MultiFits(p->Y, MultiShape{rowsPerGroup, ticksPerLaunch}),
                        "halhip_diffusion2d_set_multi_shape: the shape does not fit this grid's rows");
                p->multiB = rowsPerGroup;
                p->multiT = ticksPerLaunch;
                p->graphs.Clear();
            });
    }

    int halhip_diffusion2d_get_multi_shape(halhip_diffusion2d *p, int32_t *rowsPerGroup, int32_t *ticksPerLaunch)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_get_multi_shape");
                Require(rowsPerGroup != nullptr && ticksPerLaunch != nullptr,
                        "halhip_diffusion2d_get_multi_shape: null output");
                const MultiShape m = ChooseMulti(p);
                *rowsPerGroup = m.B;
                *ticksPerLaunch = m.T;
            });
    }

    int halhip_diffusion2d_set_advection(halhip_diffusion2d *p, double xVel, double yVel)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_set_advection");
                Require(std::isfinite(xVel) && std::isfinite(yVel), "halhip_diffusion2d_set_advection: bad velocity");
                // PDEGrid2D::CheckAdvectionStable (include/Hal/Grids/PDEGrid2D.h:444-451)
                Require(std::fabs(xVel) + std::fabs(yVel) <= 0.5,
                        "halhip_diffusion2d_set_advection: advection rate component sum above stable maximum value "
                        "of 0.5");
                p->xVel = xVel;
                p->yVel = yVel;
            });
    }

    int halhip_diffusion2d_upload(halhip_diffusion2d *p, const float *field, int64_t n)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_upload");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_diffusion2d_upload");
                Require(field != nullptr, "halhip_diffusion2d_upload: null field");
                Check(hipMemcpyAsync(p->field[p->cur].get(), field, Cells(p) * sizeof(float), hipMemcpyHostToDevice,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                if (p->deltas.get() != nullptr)
                {
                    LaunchFill(p, 6, 0.0f, nullptr); // a new committed field has no pending deltas
                }
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
            });
    }

    int halhip_diffusion2d_save_grid(halhip_diffusion2d *p, const char *path, int32_t wrapX, int32_t wrapY,
                                  int32_t updateCt)
    {
        std::vector<float> host;
        int rc = Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_save_grid");
                host.resize(Cells(p));
            });
        if (rc != HALHIP_OK)
        {
            return rc;
        }
        rc = halhip_diffusion2d_download(p, host.data(), static_cast<int64_t>(host.size()));
        if (rc != HALHIP_OK)
        {
            return rc;
        }
        return Guard([&] { SavePdeGrid2D(path, p->X, p->Y, wrapX != 0, wrapY != 0, updateCt, host); });
    }

    int halhip_diffusion2d_load_grid(halhip_diffusion2d *p, const char *path, int32_t *wrapX, int32_t *wrapY,
                                  int32_t *updateCt)
    {
        std::vector<float> host;
        int rc = Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_load_grid");
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
        return halhip_diffusion2d_upload(p, host.data(), static_cast<int64_t>(host.size()));
    }

    int halhip_diffusion2d_download(halhip_diffusion2d *p, float *field, int64_t n)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_download");
                RequireLength(n, static_cast<int64_t>(Cells(p)), "halhip_diffusion2d_download");
                Require(field != nullptr, "halhip_diffusion2d_download: null field");
                Check(hipMemcpyAsync(field, p->field[p->cur].get(), Cells(p) * sizeof(float), hipMemcpyDeviceToHost,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
            });
    }

    int halhip_diffusion2d_set_sources(halhip_diffusion2d *p, int32_t count, const int64_t *index,
                                       const float *value)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_set_sources");
                Require(count >= 0 && (count == 0 || (index != nullptr && value != nullptr)),
                        "halhip_diffusion2d_set_sources: bad source arrays");
                int64_t cells = static_cast<int64_t>(Cells(p));
                for (int32_t k = 0; k < count; k++)
                {
                    Require(index[k] >= 0 && index[k] < cells, "halhip_diffusion2d_set_sources: index out of range");
                }
                p->srcDense.Release();
                p->srcIdx.Release();
                p->srcVal.Release();
                p->nSources = count;
                p->graphs.Clear();
                if (count == 0)
                {
                    return;
                }
                if (count <= kInlineSources)
                {
                    p->srcIdx.Allocate(p->ctx, static_cast<size_t>(count));
                    p->srcVal.Allocate(p->ctx, static_cast<size_t>(count));
                    Check(hipMemcpy(p->srcIdx.get(), index, sizeof(int64_t) * count, hipMemcpyHostToDevice),
                          "hipMemcpy");
                    Check(hipMemcpy(p->srcVal.get(), value, sizeof(float) * count, hipMemcpyHostToDevice),
                          "hipMemcpy");
                    return;
                }
                std::vector<float> dense(static_cast<size_t>(cells), 0.0f);
                for (int32_t k = 0; k < count; k++)
                {
                    dense[static_cast<size_t>(index[k])] += value[k];
                }
                p->srcDense.Allocate(p->ctx, dense.size());
                Check(hipMemcpy(p->srcDense.get(), dense.data(), dense.size() * sizeof(float), hipMemcpyHostToDevice),
                      "hipMemcpy");
            });
    }

    int halhip_diffusion2d_run_steps_async(halhip_diffusion2d *p, int32_t steps, double diffRate, double decay,
                                           int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_run_steps");
                p->clock.Mark();
                EnqueueTicks(p, steps, diffRate, decay, wrapX, wrapY, bc);
            });
    }

    int halhip_diffusion2d_sync(halhip_diffusion2d *p, float *elapsedMs)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_sync");
                p->clock.Sync(elapsedMs);
                p->prof.Collect();
            });
    }

    int halhip_diffusion2d_prepare_steps(halhip_diffusion2d *p, int32_t steps, double diffRate, double decay,
                                         int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_prepare_steps");
                PrepareScope scope(p->prepareOnly);
                EnqueueTicks(p, steps, diffRate, decay, wrapX, wrapY, bc);
            });
    }

    int halhip_diffusion2d_run_steps(halhip_diffusion2d *p, int32_t steps, double diffRate, double decay,
                                     int32_t wrapX, int32_t wrapY, const halhip_bc *bc)
    {
        int rc = halhip_diffusion2d_run_steps_async(p, steps, diffRate, decay, wrapX, wrapY, bc);
        if (rc != HALHIP_OK)
        {
            return rc;
        }
        return halhip_diffusion2d_sync(p, nullptr);
    }

    int halhip_diffusion2d_set_profiling(halhip_diffusion2d *p, int32_t on)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_set_profiling");
                p->prof.Reset();
                p->prof.SetOn(on != 0);
            });
    }

    int halhip_diffusion2d_profile(halhip_diffusion2d *p, int32_t maxKernels, int32_t *nKernels, char *names,
                                   int32_t namesLen, double *totalMs, int64_t *launches)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_diffusion2d_profile");
                p->prof.Export(maxKernels, nKernels, names, namesLen, totalMs, launches);
            });
    }
}
