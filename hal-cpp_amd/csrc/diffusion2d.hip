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

struct halhip_diffusion2d
{
    halhip_context *ctx = nullptr;
    int32_t X = 0;
    int32_t Y = 0;
    DeviceArray<float> field[2]; // ping-pong resident field
    int cur = 0;
    DeviceArray<float> aux[2];   // interface scratch (delta, rates)
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
                size_t cells = Cells(p);
                if (p->aux[1].get() == nullptr)
                {
                    p->aux[1].Allocate(p->ctx, cells);
                }
                // field -> aux0, ratesX -> aux1, ratesY -> field[1-cur] scratch is not
                // used: the resident field must survive, so rates go through a
                // dedicated pair of buffers allocated on first use.
                static_assert(sizeof(float) == 4, "");
                std::vector<float> rx(cells), ry(cells);
                p->hostBuf.resize(cells);
                for (size_t k = 0; k < cells; k++)
                {
                    p->hostBuf[k] = static_cast<float>(field[k]);
                    rx[k] = static_cast<float>(ratesX[k]);
                    ry[k] = static_cast<float>(ratesY[k]);
                }
                DeviceArray<float> drx, dry;
                drx.Allocate(p->ctx, cells);
                dry.Allocate(p->ctx, cells);
                Check(hipMemcpyAsync(p->aux[0].get(), p->hostBuf.data(), cells * 4, hipMemcpyHostToDevice,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipMemcpyAsync(drx.get(), rx.data(), cells * 4, hipMemcpyHostToDevice, p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipMemcpyAsync(dry.get(), ry.data(), cells * 4, hipMemcpyHostToDevice, p->clock.stream),
                      "hipMemcpyAsync");
                p->prof.Begin(p->clock.stream, "explicit_rates_delta");
                hipLaunchKernelGGL(ExplicitRatesDelta, Grid(p), dim3(kBlockY, kBlockX), 0, p->clock.stream, p->X,
                                   p->Y, wrapX ? 1 : 0, wrapY ? 1 : 0, p->aux[0].get(), drx.get(), dry.get(),
                                   p->aux[1].get());
                Check(hipGetLastError(), "ExplicitRatesDelta launch");
                p->prof.End(p->clock.stream);
                Check(hipMemcpyAsync(p->hostBuf.data(), p->aux[1].get(), cells * 4, hipMemcpyDeviceToHost,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                for (size_t k = 0; k < cells; k++)
                {
                    deltas[k] += static_cast<double>(p->hostBuf[k]);
                }
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
                EdgeValues e;
                e.From(bc, p->X, p->Y);
                EdgeValues v;
                v.From(velBc, p->X, p->Y);
                Require(!v.dirichlet || e.dirichlet,
                        "halhip_diffusion2d_advect_cells_f64: velocity edges need field edges (Dirichlet bc)");
                const size_t cells = Cells(p);
                // field, x and y velocities, then the edge arrays, in one upload
                std::vector<float> host(3 * cells);
                for (size_t k = 0; k < cells; k++)
                {
                    host[k] = static_cast<float>(field[k]);
                    host[cells + k] = static_cast<float>(xVels[k]);
                    host[2 * cells + k] = static_cast<float>(yVels[k]);
                }
                const size_t edgeLen = 2 * static_cast<size_t>(p->X) + 2 * static_cast<size_t>(p->Y);
                if (e.dirichlet)
                {
                    for (const EdgeValues *ev : {&e, &v})
                    {
                        if (ev->dirichlet)
                        {
                            host.insert(host.end(), ev->xlo.begin(), ev->xlo.end());
                            host.insert(host.end(), ev->xhi.begin(), ev->xhi.end());
                            host.insert(host.end(), ev->ylo.begin(), ev->ylo.end());
                            host.insert(host.end(), ev->yhi.begin(), ev->yhi.end());
                        }
                    }
                }
                DeviceArray<float> dev;
                dev.Allocate(p->ctx, host.size() + cells);
                Check(hipMemcpyAsync(dev.get(), host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                float *d = dev.get();
                AdvectCellsArgs a{};
                a.X = p->X;
                a.Y = p->Y;
                a.wrapX = wrapX ? 1 : 0;
                a.wrapY = wrapY ? 1 : 0;
                if (e.dirichlet)
                {
                    float *eb = d + 3 * cells;
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
                }
                float *out = d + host.size();
                p->prof.Begin(p->clock.stream, "advect_cells_delta");
                hipLaunchKernelGGL(AdvectCellsDelta, Grid(p), dim3(kBlockY, kBlockX), 0, p->clock.stream, a,
                                   static_cast<const float *>(d), static_cast<const float *>(d + cells),
                                   static_cast<const float *>(d + 2 * cells), out);
                Check(hipGetLastError(), "AdvectCellsDelta launch");
                p->prof.End(p->clock.stream);
                p->hostBuf.resize(cells);
                Check(hipMemcpyAsync(p->hostBuf.data(), out, cells * sizeof(float), hipMemcpyDeviceToHost,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                for (size_t k = 0; k < cells; k++)
                {
                    deltas[k] += static_cast<double>(p->hostBuf[k]);
                }
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
