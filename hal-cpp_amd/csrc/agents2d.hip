// agents2d.hip -- cell-bucket agent forces and move for gfx950.
//
// Reference semantics: CpuAgentForceBackend2D::AccumulateForces
// (src/Agents/CpuAgentForceBackend2D.cpp:66-176): clamped floor cell per
// agent (:14-27), exclusive per-cell offsets (:85-97), bucket order by
// ascending agent index within a cell (:99-109), then per agent the cells
// [floor(p-R), ceil(p+R)) in x-major order, bucket order inside a cell,
// linear spring k*((ri+rj)-d) for 0 < d < R with minimum-image displacement
// (:111-175).  The move applies AgentPT2D::MoveSafePT
// (include/Hal/Agents/AgentPT2D.h:157-186) to p + dt*v.
//
// A tick is ten launches, all captured in one hipGraph for multi-tick runs:
//   fill     counts <- 0
//   count    cell_i = clamped floor cell; ticket_i = atomicAdd(count[cell_i], 1)
//   scan     offsets = exclusive prefix of counts (three passes: per-block
//            sums, scan of the block sums, block-local scan + block prefix;
//            wavefront shuffles + LDS inside a block)
//   scatter  bucket[offsets[cell_i] + ticket_i] = i   (a permutation)
//   order    each cell's slice sorted by agent index, which makes the bucket
//            identical to the reference's stable scatter whatever order the
//            atomics handed out tickets in
//   gather   bucket-ordered copies of positions and radii, so the force pass
//            reads its neighbours from a few contiguous slices
//   force    one lane per bucket slot; neighbour cells from the offsets
//   move     p += dt*v with the MoveSafePT rules
// The bins (offsets, bucket) are therefore bit-identical to the reference's
// for the same FP32 positions; forces are FP32 sums in the reference's order.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <string>

using namespace halhip;

namespace
{
    constexpr int kScanThreads = 256;
    constexpr int kScanItems = 16;                              // per thread
    constexpr int kScanTile = kScanThreads * kScanItems;        // per block
    constexpr int kThreads = 256;

    __device__ __forceinline__ int ClampCell(float v, int dim)
    {
        int c = static_cast<int>(floorf(v));
        return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
    }

    __global__ __launch_bounds__(kThreads) void FillZero(int32_t *__restrict__ p, int64_t n)
    {
        int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
        for (; i < n; i += stride)
        {
            p[i] = 0;
        }
    }

    __global__ __launch_bounds__(kThreads) void AgentCount(int32_t N, int32_t X, int32_t Y,
                                                             const float2 *__restrict__ pos,
                                                             int32_t *__restrict__ counts, int32_t *__restrict__ cellOf,
                                                             int32_t *__restrict__ ticketOf)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= N)
        {
            return;
        }
        const float2 p = pos[i];
        const int cell = ClampCell(p.x, X) * Y + ClampCell(p.y, Y);
        cellOf[i] = cell;
        ticketOf[i] = atomicAdd(&counts[cell], 1);
    }

    __device__ __forceinline__ int WaveInclusiveSum(int v, int lane)
    {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            int u = __shfl_up(v, o);
            if (lane >= o)
            {
                v += u;
            }
        }
        return v;
    }

    // Exclusive block-wide sum of one value per thread; also returns the total.
    __device__ __forceinline__ int BlockExclusiveSum(int v, int *sWave, int &total)
    {
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
        const int nWaves = blockDim.x >> 6;
        int inc = WaveInclusiveSum(v, lane);
        if (lane == 63)
        {
            sWave[wave] = inc;
        }
        __syncthreads();
        int before = 0;
        total = 0;
        for (int w = 0; w < nWaves; w++)
        {
            int s = sWave[w];
            before += (w < wave) ? s : 0;
            total += s;
        }
        __syncthreads();
        return before + inc - v;
    }

    // Pass 1: per-tile totals of counts[0..C).
    __global__ __launch_bounds__(kScanThreads) void ScanTileSums(const int32_t *__restrict__ counts, int64_t C,
                                                                 int32_t *__restrict__ tileSums)
    {
        __shared__ int sWave[kScanThreads / 64];
        const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
        int s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; k++)
        {
            s += (base + k < C) ? counts[base + k] : 0;
        }
        int total;
        BlockExclusiveSum(s, sWave, total);
        if (threadIdx.x == 0)
        {
            tileSums[blockIdx.x] = total;
        }
    }

    // Pass 2: exclusive scan of the tile totals in place (one block, any count).
    __global__ __launch_bounds__(1024) void ScanTileOffsets(int32_t *__restrict__ tileSums, int32_t nTiles)
    {
        __shared__ int sWave[16];
        int carry = 0;
        for (int base = 0; base < nTiles; base += blockDim.x)
        {
            const int i = base + threadIdx.x;
            const int v = (i < nTiles) ? tileSums[i] : 0;
            int total;
            const int ex = BlockExclusiveSum(v, sWave, total);
            if (i < nTiles)
            {
                tileSums[i] = carry + ex;
            }
            carry += total;
        }
    }

    // Pass 3: offsets[c] = tile prefix + in-tile exclusive prefix; offsets[C] = N.
    __global__ __launch_bounds__(kScanThreads) void ScanApply(const int32_t *__restrict__ counts, int64_t C,
                                                              const int32_t *__restrict__ tileOffsets,
                                                              int32_t *__restrict__ offsets, int32_t N)
    {
        __shared__ int sWave[kScanThreads / 64];
        const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
        int v[kScanItems];
        int s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; k++)
        {
            v[k] = (base + k < C) ? counts[base + k] : 0;
            s += v[k];
        }
        int total;
        int run = tileOffsets[blockIdx.x] + BlockExclusiveSum(s, sWave, total);
#pragma unroll
        for (int k = 0; k < kScanItems; k++)
        {
            if (base + k < C)
            {
                offsets[base + k] = run;
            }
            run += v[k];
        }
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            offsets[C] = N;
        }
    }

    __global__ __launch_bounds__(kThreads) void AgentScatter(int32_t N, const int32_t *__restrict__ cellOf,
                                                             const int32_t *__restrict__ ticketOf,
                                                             const int32_t *__restrict__ offsets,
                                                             int32_t *__restrict__ bucket)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= N)
        {
            return;
        }
        bucket[offsets[cellOf[i]] + ticketOf[i]] = i;
    }

    // Sorts every cell's slice of the bucket by agent index (insertion sort:
    // slices hold a handful of agents).
    __global__ __launch_bounds__(kThreads) void AgentOrderCells(int64_t C, const int32_t *__restrict__ offsets,
                                                                int32_t *__restrict__ bucket)
    {
        const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (c >= C)
        {
            return;
        }
        const int s = offsets[c];
        const int e = offsets[c + 1];
        for (int a = s + 1; a < e; a++)
        {
            const int key = bucket[a];
            int b = a - 1;
            while (b >= s && bucket[b] > key)
            {
                bucket[b + 1] = bucket[b];
                b--;
            }
            bucket[b + 1] = key;
        }
    }

    __global__ __launch_bounds__(kThreads) void AgentGather(int32_t N, const int32_t *__restrict__ bucket,
                                                            const float2 *__restrict__ pos,
                                                            const float *__restrict__ rad, float2 *__restrict__ spos,
                                                            float *__restrict__ srad)
    {
        const int k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= N)
        {
            return;
        }
        const int j = bucket[k];
        spos[k] = pos[j];
        srad[k] = rad[j];
    }

    // Minimum-image displacement b - a on one axis (CpuAgentForceBackend2D.cpp:30-45).
    __device__ __forceinline__ float Disp(float a, float b, float dim, bool wrap)
    {
        float d = b - a;
        if (wrap && fabsf(d) > dim * 0.5f)
        {
            d = (a > b) ? (d + dim) : (d - dim);
        }
        return d;
    }

    __device__ __forceinline__ int WrapIndex(int v, int dim)
    {
        int m = v % dim;
        return m < 0 ? m + dim : m;
    }

    // One lane per bucket slot k; writes the velocity delta of agent bucket[k].
    __global__ __launch_bounds__(kThreads) void AgentForces(int32_t N, int32_t X, int32_t Y, float R, float kSpring,
                                                            int32_t wrapX, int32_t wrapY,
                                                            const int32_t *__restrict__ offsets,
                                                            const int32_t *__restrict__ bucket,
                                                            const float2 *__restrict__ spos,
                                                            const float *__restrict__ srad, float2 *__restrict__ vel)
    {
        const int k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= N)
        {
            return;
        }
        const int i = bucket[k];
        const float2 p = spos[k];
        const float ri = srad[k];
        const float fX = static_cast<float>(X);
        const float fY = static_cast<float>(Y);
        const int x0 = static_cast<int>(floorf(p.x - R));
        const int x1 = static_cast<int>(ceilf(p.x + R));
        const int y0 = static_cast<int>(floorf(p.y - R));
        const int y1 = static_cast<int>(ceilf(p.y + R));
        float vx = 0.0f;
        float vy = 0.0f;
        for (int xs = x0; xs < x1; xs++)
        {
            int cx = xs;
            if (cx < 0 || cx >= X)
            {
                if (!wrapX)
                {
                    continue;
                }
                cx = WrapIndex(cx, X);
            }
            for (int ys = y0; ys < y1; ys++)
            {
                int cy = ys;
                if (cy < 0 || cy >= Y)
                {
                    if (!wrapY)
                    {
                        continue;
                    }
                    cy = WrapIndex(cy, Y);
                }
                const int64_t cell = static_cast<int64_t>(cx) * Y + cy;
                const int s = offsets[cell];
                const int e = offsets[cell + 1];
                for (int q = s; q < e; q++)
                {
                    if (q == k)
                    {
                        continue;
                    }
                    const float2 o = spos[q];
                    const float dx = Disp(p.x, o.x, fX, wrapX != 0);
                    const float dy = Disp(p.y, o.y, fY, wrapY != 0);
                    const float dist = sqrtf(dx * dx + dy * dy);
                    if (dist < R && dist > 0.0f)
                    {
                        const float force = kSpring * ((ri + srad[q]) - dist);
                        vx -= (dx / dist) * force;
                        vy -= (dy / dist) * force;
                    }
                }
            }
        }
        vel[i] = make_float2(vx, vy);
    }

    // p += dt*v with MoveSafePT: off-grid target -> wrap a wrapped axis, keep
    // the old coordinate on a non-wrapped one.
    __global__ __launch_bounds__(kThreads) void AgentMove(int32_t N, int32_t X, int32_t Y, float dt, int32_t wrapX,
                                                          int32_t wrapY, float2 *__restrict__ pos,
                                                          const float2 *__restrict__ vel)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= N)
        {
            return;
        }
        const float2 o = pos[i];
        const float2 v = vel[i];
        float nx = o.x + dt * v.x;
        float ny = o.y + dt * v.y;
        const float fX = static_cast<float>(X);
        const float fY = static_cast<float>(Y);
        const int gx = static_cast<int>(floorf(nx));
        const int gy = static_cast<int>(floorf(ny));
        const bool in = gx >= 0 && gx < X && gy >= 0 && gy < Y;
        if (!in)
        {
            if (wrapX)
            {
                float m = fmodf(nx, fX);
                nx = (m < 0.0f) ? (fX + m) : m;
            }
            else if (!(nx >= 0.0f && nx < fX))
            {
                nx = o.x;
            }
            if (wrapY)
            {
                float m = fmodf(ny, fY);
                ny = (m < 0.0f) ? (fY + m) : m;
            }
            else if (!(ny >= 0.0f && ny < fY))
            {
                ny = o.y;
            }
        }
        pos[i] = make_float2(nx, ny);
    }

    unsigned Blocks(int64_t n, int threads)
    {
        return static_cast<unsigned>(std::max<int64_t>(1, (n + threads - 1) / threads));
    }
}

struct halhip_agents2d
{
    halhip_context *ctx = nullptr;
    int32_t X = 0;
    int32_t Y = 0;
    int64_t C = 0;
    int32_t capacity = 0;
    int32_t N = 0;
    int32_t nTiles = 0;
    DeviceArray<float2> pos;
    DeviceArray<float> rad;
    DeviceArray<float2> vel;
    DeviceArray<float2> spos;
    DeviceArray<float> srad;
    DeviceArray<int32_t> counts;
    DeviceArray<int32_t> offsets;
    DeviceArray<int32_t> tileSums;
    DeviceArray<int32_t> cellOf;
    DeviceArray<int32_t> ticketOf;
    DeviceArray<int32_t> bucket;
    bool binsValid = false;
    StreamClock clock;
    Profiler prof;
    GraphCache graphs;
    std::vector<float> hostBuf;
};

namespace
{
    void RequirePlan(const halhip_agents2d *p, const char *caller)
    {
        Require(p != nullptr, std::string(caller) + ": null plan");
        Bind(p->ctx);
    }

    template <class K, class... A>
    void Launch(halhip_agents2d *p, const char *name, K kernel, unsigned blocks, unsigned threads, A... args)
    {
        p->prof.Begin(p->clock.stream, name);
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), 0, p->clock.stream, args...);
        Check(hipGetLastError(), name);
        p->prof.End(p->clock.stream);
    }

    // Bins of the current positions: counts, offsets, stable bucket, and the
    // bucket-ordered position / radius copies.
    void EnqueueBins(halhip_agents2d *p)
    {
        const int32_t N = p->N;
        Launch(p, "agents_fill", FillZero, Blocks(p->C, kThreads * 8), kThreads, p->counts.get(), p->C);
        if (N > 0)
        {
            Launch(p, "agents_count", AgentCount, Blocks(N, kThreads), kThreads, N, p->X, p->Y, p->pos.get(),
                   p->counts.get(), p->cellOf.get(), p->ticketOf.get());
        }
        Launch(p, "agents_scan_tiles", ScanTileSums, static_cast<unsigned>(p->nTiles), kScanThreads,
               static_cast<const int32_t *>(p->counts.get()), p->C, p->tileSums.get());
        Launch(p, "agents_scan_tile_offsets", ScanTileOffsets, 1u, 1024u, p->tileSums.get(), p->nTiles);
        Launch(p, "agents_scan_apply", ScanApply, static_cast<unsigned>(p->nTiles), kScanThreads,
               static_cast<const int32_t *>(p->counts.get()), p->C, static_cast<const int32_t *>(p->tileSums.get()),
               p->offsets.get(), N);
        if (N > 0)
        {
            Launch(p, "agents_scatter", AgentScatter, Blocks(N, kThreads), kThreads, N,
                   static_cast<const int32_t *>(p->cellOf.get()), static_cast<const int32_t *>(p->ticketOf.get()),
                   static_cast<const int32_t *>(p->offsets.get()), p->bucket.get());
            Launch(p, "agents_order", AgentOrderCells, Blocks(p->C, kThreads), kThreads, p->C,
                   static_cast<const int32_t *>(p->offsets.get()), p->bucket.get());
            Launch(p, "agents_gather", AgentGather, Blocks(N, kThreads), kThreads, N,
                   static_cast<const int32_t *>(p->bucket.get()), static_cast<const float2 *>(p->pos.get()),
                   static_cast<const float *>(p->rad.get()), p->spos.get(), p->srad.get());
        }
    }

    void EnqueueForces(halhip_agents2d *p, float R, float k, int32_t wrapX, int32_t wrapY)
    {
        if (p->N == 0)
        {
            return;
        }
        Launch(p, "agents_force", AgentForces, Blocks(p->N, kThreads), kThreads, p->N, p->X, p->Y, R, k,
               wrapX ? 1 : 0, wrapY ? 1 : 0, static_cast<const int32_t *>(p->offsets.get()),
               static_cast<const int32_t *>(p->bucket.get()), static_cast<const float2 *>(p->spos.get()),
               static_cast<const float *>(p->srad.get()), p->vel.get());
    }

    void EnqueueTick(halhip_agents2d *p, float R, float k, float dt, int32_t wrapX, int32_t wrapY)
    {
        EnqueueBins(p);
        EnqueueForces(p, R, k, wrapX, wrapY);
        if (p->N > 0)
        {
            Launch(p, "agents_move", AgentMove, Blocks(p->N, kThreads), kThreads, p->N, p->X, p->Y, dt,
                   wrapX ? 1 : 0, wrapY ? 1 : 0, p->pos.get(), static_cast<const float2 *>(p->vel.get()));
        }
    }

    void CheckParams(const halhip_agents2d *p, double R, double k, double dt)
    {
        Require(std::isfinite(R) && R >= 0.0, "agents2d: interaction radius must be finite and non-negative");
        Require(R <= 2.0 * std::max(p->X, p->Y), "agents2d: interaction radius larger than twice the grid");
        Require(std::isfinite(k), "agents2d: spring constant must be finite");
        Require(std::isfinite(dt), "agents2d: dt must be finite");
    }

    void EnqueueTicks(halhip_agents2d *p, int32_t ticks, double R, double k, double dt, int32_t wrapX, int32_t wrapY)
    {
        Require(ticks >= 0, "halhip_agents2d_run_ticks: negative ticks");
        CheckParams(p, R, k, dt);
        const float fR = static_cast<float>(R);
        const float fk = static_cast<float>(k);
        const float fdt = static_cast<float>(dt);
        if (ticks == 0)
        {
            EnqueueBins(p);
            EnqueueForces(p, fR, fk, wrapX, wrapY);
            p->binsValid = true;
            return;
        }
        constexpr int32_t kChunk = 10;
        int32_t done = 0;
        if (!p->prof.On() && ticks >= kChunk)
        {
            std::string key;
            AppendKey(key, p->N);
            AppendKey(key, fR);
            AppendKey(key, fk);
            AppendKey(key, fdt);
            AppendKey(key, wrapX);
            AppendKey(key, wrapY);
            hipGraphExec_t exec = p->graphs.Find(key);
            if (exec == nullptr)
            {
                hipGraph_t graph;
                Check(hipStreamBeginCapture(p->clock.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                for (int32_t t = 0; t < kChunk; t++)
                {
                    EnqueueTick(p, fR, fk, fdt, wrapX, wrapY);
                }
                Check(hipStreamEndCapture(p->clock.stream, &graph), "hipStreamEndCapture");
                Check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
                (void)hipGraphDestroy(graph);
                p->graphs.Insert(key, exec);
            }
            for (; done + kChunk <= ticks; done += kChunk)
            {
                Check(hipGraphLaunch(exec, p->clock.stream), "hipGraphLaunch");
            }
        }
        for (; done < ticks; done++)
        {
            EnqueueTick(p, fR, fk, fdt, wrapX, wrapY);
        }
        // The bins describe the positions before the last move.
        p->binsValid = true;
    }
}

extern "C"
{
    int halhip_agents2d_create(halhip_context *ctx, int32_t xDim, int32_t yDim, int32_t maxAgents,
                               halhip_agents2d **out)
    {
        halhip_agents2d *p = nullptr;
        int rc = Guard(
            [&]
            {
                Require(out != nullptr, "halhip_agents2d_create: null out");
                *out = nullptr;
                Require(ctx != nullptr, "halhip_agents2d_create: null context");
                Require(xDim > 0 && yDim > 0, "halhip_agents2d_create: dimensions must be positive");
                Require(maxAgents >= 0, "halhip_agents2d_create: negative capacity");
                Require(static_cast<int64_t>(xDim) * yDim < (int64_t(1) << 31),
                        "halhip_agents2d_create: more than 2^31-1 bins");
                Bind(ctx);
                p = new halhip_agents2d();
                p->ctx = ctx;
                p->X = xDim;
                p->Y = yDim;
                p->C = static_cast<int64_t>(xDim) * yDim;
                p->capacity = maxAgents;
                p->nTiles = static_cast<int32_t>((p->C + kScanTile - 1) / kScanTile);
                const size_t cap = static_cast<size_t>(std::max(1, maxAgents));
                p->pos.Allocate(ctx, cap);
                p->rad.Allocate(ctx, cap);
                p->vel.Allocate(ctx, cap);
                p->spos.Allocate(ctx, cap);
                p->srad.Allocate(ctx, cap);
                p->cellOf.Allocate(ctx, cap);
                p->ticketOf.Allocate(ctx, cap);
                p->bucket.Allocate(ctx, cap);
                p->counts.Allocate(ctx, static_cast<size_t>(p->C));
                p->offsets.Allocate(ctx, static_cast<size_t>(p->C) + 1);
                p->tileSums.Allocate(ctx, static_cast<size_t>(p->nTiles));
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

    int halhip_agents2d_destroy(halhip_agents2d *p)
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

    int halhip_agents2d_upload(halhip_agents2d *p, const float *positions, const float *radii, int32_t nAgents)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_upload");
                Require(nAgents >= 0 && nAgents <= p->capacity,
                        "halhip_agents2d_upload: agent count outside [0, construction-time capacity]");
                Require(nAgents == 0 || (positions != nullptr && radii != nullptr),
                        "halhip_agents2d_upload: null array");
                if (nAgents != p->N)
                {
                    p->graphs.Clear();
                }
                p->N = nAgents;
                p->binsValid = false;
                if (nAgents == 0)
                {
                    return;
                }
                Check(hipMemcpyAsync(p->pos.get(), positions, sizeof(float2) * nAgents, hipMemcpyHostToDevice,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipMemcpyAsync(p->rad.get(), radii, sizeof(float) * nAgents, hipMemcpyHostToDevice,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
            });
    }

    int halhip_agents2d_download_positions(halhip_agents2d *p, float *positions, int32_t nAgents)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_download_positions");
                Require(nAgents == p->N, "halhip_agents2d_download_positions: count differs from the uploaded count");
                Require(nAgents == 0 || positions != nullptr, "halhip_agents2d_download_positions: null array");
                if (nAgents == 0)
                {
                    return;
                }
                Check(hipMemcpyAsync(positions, p->pos.get(), sizeof(float2) * nAgents, hipMemcpyDeviceToHost,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
            });
    }

    int halhip_agents2d_download_velocities(halhip_agents2d *p, float *velocityDeltas, int32_t nAgents)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_download_velocities");
                Require(nAgents == p->N, "halhip_agents2d_download_velocities: count differs from the uploaded count");
                Require(p->binsValid, "halhip_agents2d_download_velocities: no forces computed since the upload");
                Require(nAgents == 0 || velocityDeltas != nullptr, "halhip_agents2d_download_velocities: null array");
                if (nAgents == 0)
                {
                    return;
                }
                Check(hipMemcpyAsync(velocityDeltas, p->vel.get(), sizeof(float2) * nAgents, hipMemcpyDeviceToHost,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
            });
    }

    int halhip_agents2d_download_bins(halhip_agents2d *p, int32_t *offsets, int32_t *bucket)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_download_bins");
                Require(p->binsValid, "halhip_agents2d_download_bins: no bins computed since the upload");
                Require(offsets != nullptr && (p->N == 0 || bucket != nullptr), "halhip_agents2d_download_bins: null array");
                Check(hipMemcpyAsync(offsets, p->offsets.get(), sizeof(int32_t) * (p->C + 1), hipMemcpyDeviceToHost,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                if (p->N > 0)
                {
                    Check(hipMemcpyAsync(bucket, p->bucket.get(), sizeof(int32_t) * p->N, hipMemcpyDeviceToHost,
                                         p->clock.stream),
                          "hipMemcpyAsync");
                }
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
            });
    }

    int halhip_agents2d_forces_f64(halhip_agents2d *p, const double *positions, const double *radii,
                                   int32_t nAgents, double *velocityDeltas, double interactionRad,
                                   double springConstant, int32_t wrapX, int32_t wrapY)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_forces_f64");
                Require(nAgents >= 0 && nAgents <= p->capacity,
                        "halhip_agents2d_forces_f64: more agents than the construction-time capacity");
                Require(nAgents == 0 || (positions != nullptr && radii != nullptr && velocityDeltas != nullptr),
                        "halhip_agents2d_forces_f64: null array");
                CheckParams(p, interactionRad, springConstant, 0.0);
                const size_t n = static_cast<size_t>(nAgents);
                p->hostBuf.resize(3 * n);
                for (size_t i = 0; i < 2 * n; i++)
                {
                    p->hostBuf[i] = static_cast<float>(positions[i]);
                }
                for (size_t i = 0; i < n; i++)
                {
                    p->hostBuf[2 * n + i] = static_cast<float>(radii[i]);
                }
                if (nAgents != p->N)
                {
                    p->graphs.Clear();
                }
                p->N = nAgents;
                if (n > 0)
                {
                    Check(hipMemcpyAsync(p->pos.get(), p->hostBuf.data(), sizeof(float2) * n, hipMemcpyHostToDevice,
                                         p->clock.stream),
                          "hipMemcpyAsync");
                    Check(hipMemcpyAsync(p->rad.get(), p->hostBuf.data() + 2 * n, sizeof(float) * n,
                                         hipMemcpyHostToDevice, p->clock.stream),
                          "hipMemcpyAsync");
                }
                EnqueueBins(p);
                EnqueueForces(p, static_cast<float>(interactionRad), static_cast<float>(springConstant), wrapX, wrapY);
                p->binsValid = true;
                if (n > 0)
                {
                    Check(hipMemcpyAsync(p->hostBuf.data(), p->vel.get(), sizeof(float2) * n, hipMemcpyDeviceToHost,
                                         p->clock.stream),
                          "hipMemcpyAsync");
                }
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                p->prof.Collect();
                for (size_t i = 0; i < 2 * n; i++)
                {
                    velocityDeltas[i] = static_cast<double>(p->hostBuf[i]);
                }
            });
    }

    int halhip_agents2d_run_ticks_async(halhip_agents2d *p, int32_t ticks, double interactionRad,
                                        double springConstant, double dt, int32_t wrapX, int32_t wrapY)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_run_ticks");
                p->clock.Mark();
                EnqueueTicks(p, ticks, interactionRad, springConstant, dt, wrapX, wrapY);
            });
    }

    int halhip_agents2d_sync(halhip_agents2d *p, float *elapsedMs)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_sync");
                p->clock.Sync(elapsedMs);
                p->prof.Collect();
            });
    }

    int halhip_agents2d_run_ticks(halhip_agents2d *p, int32_t ticks, double interactionRad, double springConstant,
                                  double dt, int32_t wrapX, int32_t wrapY)
    {
        int rc = halhip_agents2d_run_ticks_async(p, ticks, interactionRad, springConstant, dt, wrapX, wrapY);
        if (rc != HALHIP_OK)
        {
            return rc;
        }
        return halhip_agents2d_sync(p, nullptr);
    }

    int halhip_agents2d_set_profiling(halhip_agents2d *p, int32_t on)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_set_profiling");
                p->prof.Reset();
                p->prof.SetOn(on != 0);
            });
    }

    int halhip_agents2d_profile(halhip_agents2d *p, int32_t maxKernels, int32_t *nKernels, char *names,
                                int32_t namesLen, double *totalMs, int64_t *launches)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_profile");
                p->prof.Export(maxKernels, nKernels, names, namesLen, totalMs, launches);
            });
    }
}
