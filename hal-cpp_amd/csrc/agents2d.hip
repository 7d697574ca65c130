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
// Agents stay resident in slot order: sorted by (cell, agent index) as of
// the previous binning, with id[k] the agent index held in slot k (after an
// upload, slot k holds agent k).  Agents move little per tick, so slot order
// stays close to spatial order and every pass below touches memory almost in
// order.  A tick is four launches (the count is part of the previous tick's
// force pass after the first), captured in hipGraphs for multi-tick runs:
//   count    per slot: cell = clamped floor cell; ticket = atomicAdd(count[cell], 1)
//   scan     offsets = exclusive prefix of counts in one pass: tiles claimed
//            in order, each publishing its total and summing its
//            predecessors' published totals (a bounded wait); the pass also
//            resets the counts to 0 for the next binning
//   scatter  (id, position, radius) of each slot to offsets[cell] + ticket
//            in the other buffer set (a permutation)
//   order    each cell's slice sorted by agent index, which makes the ids in
//            slot order identical to the reference's bucket (stable scatter)
//            whatever order the atomics handed out tickets in
//   force    one lane per slot; neighbour cells from the offsets, their
//            agents read from contiguous slices
//   move     p += dt*v with the MoveSafePT rules, in slot order
// The bins (offsets, bucket = ids in slot order) are therefore bit-identical
// to the reference's for the same FP32 positions; forces are FP32 sums in
// the reference's order.  Downloads by agent index scatter through id.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <type_traits>

using namespace halhip;

namespace
{
    constexpr int kScanThreads = 256;
    constexpr int kScanRounds = 4;                              // int4 loads per thread
    constexpr int kScanTile = kScanThreads * 4 * kScanRounds;   // counts per block
    constexpr int kThreads = 256;

    __device__ __forceinline__ int ClampCell(float v, int dim)
    {
        int c = static_cast<int>(floorf(v));
        return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
    }

    // Counts slot i's agent into its clamped cell and takes a ticket there
    // (the scatter computes the cell again from the same position).
    __global__ __launch_bounds__(kThreads) void AgentCount(int32_t N, int32_t X, int32_t Y,
                                                             const float2 *__restrict__ pos,
                                                             int32_t *__restrict__ counts,
                                                             int32_t *__restrict__ ticketOf)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= N)
        {
            return;
        }
        const float2 p = pos[i];
        const int cell = ClampCell(p.x, X) * Y + ClampCell(p.y, Y);
        ticketOf[i] = atomicAdd(&counts[cell], 1);
    }

    // Inclusive sum over the wavefront's lanes by DPP: shifts within each
    // 16-lane row, then each row's total broadcast into the rows above it
    // (row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2 and 3).
    // Lanes a shift leaves without a source add 0.
    __device__ __forceinline__ int WaveInclusiveSum(int v)
    {
        v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false); // row_shr:1
        v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false); // row_shr:2
        v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false); // row_shr:4
        v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false); // row_shr:8
        v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false); // row_bcast:15
        v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false); // row_bcast:31
        return v;
    }

    // Exclusive block-wide sums of R sequences at once (value r of every
    // thread is one sequence); total[r] gets each sequence's sum.  sWave
    // holds R * (waves per block) ints.
    template <int R>
    __device__ __forceinline__ void BlockExclusiveSums(const int (&v)[R], int *sWave, int (&ex)[R], int (&total)[R])
    {
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
        const int nWaves = blockDim.x >> 6;
        int inc[R];
#pragma unroll
        for (int r = 0; r < R; r++)
        {
            inc[r] = WaveInclusiveSum(v[r]);
        }
        if (lane == 63)
        {
#pragma unroll
            for (int r = 0; r < R; r++)
            {
                sWave[r * nWaves + wave] = inc[r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; r++)
        {
            int before = 0;
            int sum = 0;
            for (int w = 0; w < nWaves; w++)
            {
                const int t = sWave[r * nWaves + w];
                before += (w < wave) ? t : 0;
                sum += t;
            }
            ex[r] = before + inc[r] - v[r];
            total[r] = sum;
        }
        __syncthreads();
    }

    // Exclusive block-wide sum of one value per thread; also returns the total.
    __device__ __forceinline__ int BlockExclusiveSum(int v, int *sWave, int &total)
    {
        const int in[1] = {v};
        int ex[1];
        int tot[1];
        BlockExclusiveSums<1>(in, sWave, ex, tot);
        total = tot[0];
        return ex[0];
    }

    // The counts and offsets arrays are padded to whole tiles (padding counts
    // stay 0), so every access below is a whole, aligned int4.  In round r
    // thread t covers counts [4(256 r + t), 4(256 r + t) + 4) of its tile.

    // Pass 1: per-tile totals of the counts.
    __global__ __launch_bounds__(kScanThreads) void ScanTileSums(const int4 *__restrict__ counts,
                                                                 int32_t *__restrict__ tileSums)
    {
        __shared__ int sWave[kScanThreads / 64];
        const int4 *tile = counts + static_cast<int64_t>(blockIdx.x) * (kScanTile / 4);
        int s = 0;
#pragma unroll
        for (int r = 0; r < kScanRounds; r++)
        {
            const int4 v = tile[r * kScanThreads + threadIdx.x];
            s += (v.x + v.y) + (v.z + v.w);
        }
        int total;
        BlockExclusiveSum(s, sWave, total);
        if (threadIdx.x == 0)
        {
            tileSums[blockIdx.x] = total;
        }
    }

    // Pass 2: offsets = tile prefix (summed here from the tile totals) +
    // in-tile exclusive prefix (offsets past the last real cell read N), and
    // the counts reset to 0 for the next binning.
    __global__ __launch_bounds__(kScanThreads) void ScanApply(int4 *__restrict__ counts,
                                                              const int32_t *__restrict__ tileSums,
                                                              int4 *__restrict__ offsets, int32_t *__restrict__ offsetsTail,
                                                              int32_t N)
    {
        __shared__ int sWave[kScanThreads / 64];
        const int64_t tileBase = static_cast<int64_t>(blockIdx.x) * (kScanTile / 4);
        int run;
        {
            int part = 0;
            for (int t = threadIdx.x; t < static_cast<int>(blockIdx.x); t += kScanThreads)
            {
                part += tileSums[t];
            }
            BlockExclusiveSum(part, sWave, run);
        }
#pragma unroll
        for (int r = 0; r < kScanRounds; r++)
        {
            const int64_t i = tileBase + r * kScanThreads + threadIdx.x;
            const int4 v = counts[i];
            int total;
            const int o = run + BlockExclusiveSum((v.x + v.y) + (v.z + v.w), sWave, total);
            offsets[i] = make_int4(o, o + v.x, o + v.x + v.y, o + v.x + v.y + v.z);
            counts[i] = make_int4(0, 0, 0, 0);
            run += total;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            *offsetsTail = N;
        }
    }

    // One-pass scan.  Workgroup t takes tile t.  Each tile publishes its
    // total as soon as its counts are summed, then sums the published totals
    // of all its predecessors (one round trip: every thread polls a few of
    // them) for its prefix.  A published word is (epoch << 32) | total, so a
    // word from an earlier scan is never taken for this one and the words
    // need no clearing.  The last tile has seen every other tile's word, so
    // every workgroup has read the epoch by then: it advances the epoch.
    // Waits are on lower tiles only; workgroups are dispatched in index
    // order, so the lowest unfinished tile never waits on one not running.
    // The wait is bounded all the same: a thread that has polled spinLimit
    // times without seeing a total stops waiting and raises *fault (sticky);
    // the workgroup finishes, the kernels that read the offsets do nothing
    // while the fault is up, and the host reports it and clears the scan
    // state.  No counter is shared by all workgroups (such atomics serialize
    // at the coherence point).  A published word carries its epoch and value
    // together and no other data is read on its strength, so relaxed
    // agent-scope atomics suffice: they reach the coherence point shared by
    // the XCDs without the whole-L2 write-back and invalidation that release
    // and acquire would add.
    struct ScanState
    {
        unsigned long long *tiles; // [nTiles] published totals
        int32_t *words;            // [2] epoch, [3] fault ([0], [1] unused)
        int32_t spinLimit;
    };

    __global__ __launch_bounds__(kScanThreads) void ScanOnePass(const int4 *__restrict__ counts,
                                                                int4 *__restrict__ offsets,
                                                                int32_t *__restrict__ offsetsTail, int32_t N,
                                                                int32_t nTiles, ScanState st)
    {
        __shared__ int sWave[kScanRounds * (kScanThreads / 64)];
        __shared__ uint32_t sEpoch;
        if (threadIdx.x == 0)
        {
            sEpoch = static_cast<uint32_t>(__hip_atomic_load(&st.words[2], __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT));
        }
        __syncthreads();
        const int tile = static_cast<int>(blockIdx.x);
        const uint32_t epoch = sEpoch;
        const int64_t tileBase = static_cast<int64_t>(tile) * (kScanTile / 4);
        int4 v[kScanRounds];
#pragma unroll
        for (int r = 0; r < kScanRounds; r++)
        {
            v[r] = counts[tileBase + r * kScanThreads + threadIdx.x];
        }
        // the rounds' block sums together, then chained round after round
        int sums[kScanRounds];
#pragma unroll
        for (int r = 0; r < kScanRounds; r++)
        {
            sums[r] = (v[r].x + v[r].y) + (v[r].z + v[r].w);
        }
        int ex[kScanRounds];
        int roundTotal[kScanRounds];
        BlockExclusiveSums<kScanRounds>(sums, sWave, ex, roundTotal);
        int tileTotal = 0;
#pragma unroll
        for (int r = 0; r < kScanRounds; r++)
        {
            ex[r] += tileTotal;
            tileTotal += roundTotal[r];
        }
        if (threadIdx.x == 0)
        {
            __hip_atomic_store(&st.tiles[tile],
                               (static_cast<unsigned long long>(epoch) << 32) | static_cast<uint32_t>(tileTotal),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // the prefix: every predecessor's published total
        int part = 0;
        int spins = 0;
        bool gaveUp = false;
        for (int j = threadIdx.x; j < tile && !gaveUp; j += kScanThreads)
        {
            for (;;)
            {
                const unsigned long long w =
                    __hip_atomic_load(&st.tiles[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (static_cast<uint32_t>(w >> 32) == epoch)
                {
                    part += static_cast<int>(static_cast<uint32_t>(w));
                    break;
                }
                if (++spins > st.spinLimit)
                {
                    gaveUp = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (gaveUp)
        {
            atomicExch(&st.words[3], 1);
        }
        int run;
        BlockExclusiveSum(part, sWave, run); // (its barriers: every poll above is done)
        if (tile == nTiles - 1 && threadIdx.x == 0)
        {
            // every other tile published, so every workgroup has read the
            // epoch: the next scan's words are told apart from these
            // (epoch 0 is the cleared words' own, never used)
            const uint32_t next = (epoch + 1u == 0u) ? 1u : epoch + 1u;
            __hip_atomic_store(&st.words[2], static_cast<int32_t>(next), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
#pragma unroll
        for (int r = 0; r < kScanRounds; r++)
        {
            const int64_t i = tileBase + r * kScanThreads + threadIdx.x;
            const int o = run + ex[r];
            offsets[i] = make_int4(o, o + v[r].x, o + v[r].x + v[r].y, o + v[r].x + v[r].y + v[r].z);
        }
        if (tile == 0 && threadIdx.x == 0)
        {
            *offsetsTail = N;
        }
    }

    // Moves slot k's (id, position, radius) to slot offsets[cell] + ticket,
    // with the clamped cell of its position as the count computed it (no
    // radius with rad null: every agent has the same radius).
    __global__ __launch_bounds__(kThreads) void AgentScatter(int32_t N, int32_t X, int32_t Y,
                                                             const int32_t *__restrict__ ticketOf,
                                                             const int32_t *__restrict__ offsets,
                                                             const int32_t *__restrict__ id,
                                                             const float2 *__restrict__ pos,
                                                             const float *__restrict__ rad,
                                                             const float2 *__restrict__ carry,
                                                             int32_t *__restrict__ nId, float2 *__restrict__ nPos,
                                                             float *__restrict__ nRad, float2 *__restrict__ nCarry,
                                                             const int32_t *__restrict__ scanFault)
    {
        const int k = XcdContiguousBlock(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
        if (k >= N || *scanFault != 0)
        {
            return; // (a scan that gave up left offsets that must not be used)
        }
        const float2 q = pos[k];
        const int dst = offsets[ClampCell(q.x, X) * Y + ClampCell(q.y, Y)] + ticketOf[k];
        nId[dst] = id[k];
        nPos[dst] = q;
        if (rad != nullptr)
        {
            nRad[dst] = rad[k];
        }
        if (carry != nullptr)
        {
            nCarry[dst] = carry[k];
        }
    }

    // Sorts one cell's slice [s, e) of slots by agent index, moving position,
    // radius (unless rad is null) and (when kept) carried velocity along
    // (insertion sort: slices hold a handful of agents).
    __device__ __forceinline__ void OrderSlice(int s, int e, int32_t *__restrict__ id, float2 *__restrict__ pos,
                                               float *__restrict__ rad, float2 *__restrict__ carry)
    {
        for (int a = s + 1; a < e; a++)
        {
            const int key = id[a];
            if (id[a - 1] <= key)
            {
                continue;
            }
            const float2 kp = pos[a];
            const float kr = (rad != nullptr) ? rad[a] : 0.0f;
            const float2 kc = (carry != nullptr) ? carry[a] : make_float2(0.0f, 0.0f);
            int b = a - 1;
            while (b >= s && id[b] > key)
            {
                id[b + 1] = id[b];
                pos[b + 1] = pos[b];
                if (rad != nullptr)
                {
                    rad[b + 1] = rad[b];
                }
                if (carry != nullptr)
                {
                    carry[b + 1] = carry[b];
                }
                b--;
            }
            id[b + 1] = key;
            pos[b + 1] = kp;
            if (rad != nullptr)
            {
                rad[b + 1] = kr;
            }
            if (carry != nullptr)
            {
                carry[b + 1] = kc;
            }
        }
    }

    // Sorts every cell's slice, four consecutive cells per thread: their five
    // offsets are one int4 and one int (offsets holds whole scan tiles plus
    // 4 entries, and c0 is a multiple of 4, so both loads are in bounds).
    // It also resets its four cells' counts for the next binning (the scan
    // has read them); counts holds whole tiles, so the int4 is in bounds.
    __global__ __launch_bounds__(kThreads) void AgentOrderCells(int64_t C, const int32_t *__restrict__ offsets,
                                                                int32_t *__restrict__ id, float2 *__restrict__ pos,
                                                                float *__restrict__ rad, float2 *__restrict__ carry,
                                                                int32_t *__restrict__ counts,
                                                                const int32_t *__restrict__ scanFault)
    {
        const int64_t c0 =
            4 * (static_cast<int64_t>(XcdContiguousBlock(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x);
        // Offsets from a scan that gave up are not monotonic: one cell's
        // range could span most slots, which no insertion sort may walk
        // (the host then clears the counts).
        if (c0 >= C || *scanFault != 0)
        {
            return;
        }
        *reinterpret_cast<int4 *>(counts + c0) = make_int4(0, 0, 0, 0);
        const int4 o = *reinterpret_cast<const int4 *>(offsets + c0);
        const int o4 = offsets[c0 + 4];
        const int b[5] = {o.x, o.y, o.z, o.w, o4};
        // Cells of two agents (most of the cells that need ordering at about
        // one agent per cell): both ids of all four cells are requested
        // before any is compared, and a pair out of order is swapped in
        // straight-line code.  Larger cells take the insertion sort.
        int first[4];
        int second[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
        {
            const bool pair = c0 + j < C && b[j + 1] - b[j] == 2;
            first[j] = pair ? id[b[j]] : 0;
            second[j] = pair ? id[b[j] + 1] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
        {
            if (c0 + j >= C)
            {
                continue;
            }
            const int n = b[j + 1] - b[j];
            if (n == 2)
            {
                if (first[j] > second[j])
                {
                    const int s = b[j];
                    id[s] = second[j];
                    id[s + 1] = first[j];
                    const float2 p0 = pos[s];
                    const float2 p1 = pos[s + 1];
                    pos[s] = p1;
                    pos[s + 1] = p0;
                    if (rad != nullptr)
                    {
                        const float r0 = rad[s];
                        const float r1 = rad[s + 1];
                        rad[s] = r1;
                        rad[s + 1] = r0;
                    }
                    if (carry != nullptr)
                    {
                        const float2 k0 = carry[s];
                        const float2 k1 = carry[s + 1];
                        carry[s] = k1;
                        carry[s + 1] = k0;
                    }
                }
            }
            else if (n > 2)
            {
                OrderSlice(b[j], b[j + 1], id, pos, rad, carry);
            }
        }
    }

    // out[k] = v[id[k]]: agent-index order to slot order.
    __global__ __launch_bounds__(kThreads) void AgentGather(int32_t N, const int32_t *__restrict__ id,
                                                            const float2 *__restrict__ v, float2 *__restrict__ out)
    {
        const int k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k < N)
        {
            out[k] = v[id[k]];
        }
    }

    // out[id[k]] = v[k]: slot order to agent-index order.
    __global__ __launch_bounds__(kThreads) void AgentUnscatter(int32_t N, const int32_t *__restrict__ id,
                                                               const float2 *__restrict__ v, float2 *__restrict__ out)
    {
        const int k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= N)
        {
            return;
        }
        out[id[k]] = v[k];
    }

    // Slot k holds agent k (after an upload).
    __global__ __launch_bounds__(kThreads) void AgentIota(int32_t N, int32_t *__restrict__ id)
    {
        const int k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k < N)
        {
            id[k] = k;
        }
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

    // Interaction of the agent in slot k (position p, radius ri) with the
    // agents of slots [s, e), accumulated into (vx, vy) in slot order.
    // One candidate q (position o, radius rq) acting on slot k: hardware
    // square root and reciprocal (within 1 ulp); the FP32 sums stay within
    // 1e-5 of the reference's FP64 ones.
    __device__ __forceinline__ void AccumulateOne(int q, float2 o, float rq, int k, float2 p, float ri, float R,
                                                  float kSpring, float fX, float fY, bool wrapX, bool wrapY,
                                                  float &vx, float &vy)
    {
        const float dx = Disp(p.x, o.x, fX, wrapX);
        const float dy = Disp(p.y, o.y, fY, wrapY);
        const float dist = __builtin_amdgcn_sqrtf(dx * dx + dy * dy);
        if (q != k && dist < R && dist > 0.0f)
        {
            const float f = kSpring * ((ri + rq) - dist) * __builtin_amdgcn_rcpf(dist);
            vx -= dx * f;
            vy -= dy * f;
        }
    }

    // UR: every agent's radius is ri (srad is not read).
    template <bool UR>
    __device__ __forceinline__ void AccumulateRange(int s, int e, int k, float2 p, float ri, float R, float kSpring,
                                                    float fX, float fY, bool wrapX, bool wrapY,
                                                    const float2 *__restrict__ spos, const float *__restrict__ srad,
                                                    float &vx, float &vy)
    {
        auto one = [&](int q, float2 o, float rq)
        { AccumulateOne(q, o, rq, k, p, ri, R, kSpring, fX, fY, wrapX, wrapY, vx, vy); };
        // candidates in pairs, both loads in flight together, summed in slot
        // order
        int q = s;
        for (; q + 1 < e; q += 2)
        {
            const float2 o0 = spos[q];
            const float2 o1 = spos[q + 1];
            const float r0 = UR ? ri : srad[q];
            const float r1 = UR ? ri : srad[q + 1];
            one(q, o0, r0);
            one(q + 1, o1, r1);
        }
        if (q < e)
        {
            one(q, spos[q], UR ? ri : srad[q]);
        }
    }

    // What the move does to the agent in slot i: p += dt*v with MoveSafePT
    // (an off-grid target wraps a wrapped axis and keeps the old coordinate
    // on a non-wrapped one).  With a carried velocity (friction or a speed
    // cap set), the velocity is carry + this tick's force delta v, as
    // SphericalAgent2D accumulates SumForces into xVel/yVel
    // (include/Hal/Agents/SphericalAgent2D.h:100-128); its norm is capped at
    // maxVel when maxVel > 0 (CapVelocity, :139-150), the move uses it
    // (ForceMove, :153-156), and friction * velocity is carried to the next
    // tick (ApplyFriction, :133-137).
    struct MoveArgs
    {
        int32_t on;      // 0: forces only
        float dt;
        float2 *posNext; // moved positions, slot order
        float2 *carry;   // carried velocities (null: none)
        float friction;
        float maxVel;
        // the next binning's count of the moved positions (AgentCount's
        // work): a ticket in the clamped cell
        int32_t *counts;
        int32_t *ticketOf;
    };

    __device__ __forceinline__ float2 MovedPosition(const MoveArgs &m, int i, float2 o, float2 v, int32_t X,
                                                    int32_t Y, bool wrapX, bool wrapY)
    {
        if (m.carry != nullptr)
        {
            const float2 c = m.carry[i];
            v.x += c.x;
            v.y += c.y;
            if (m.maxVel > 0.0f)
            {
                const float normSq = v.x * v.x + v.y * v.y;
                if (normSq > m.maxVel * m.maxVel)
                {
                    const float f = m.maxVel / sqrtf(normSq);
                    v.x *= f;
                    v.y *= f;
                }
            }
            m.carry[i] = make_float2(m.friction * v.x, m.friction * v.y);
        }
        float nx = o.x + m.dt * v.x;
        float ny = o.y + m.dt * v.y;
        const float fX = static_cast<float>(X);
        const float fY = static_cast<float>(Y);
        const int gx = static_cast<int>(floorf(nx));
        const int gy = static_cast<int>(floorf(ny));
        const bool in = gx >= 0 && gx < X && gy >= 0 && gy < Y;
        if (!in)
        {
            if (wrapX)
            {
                float md = fmodf(nx, fX);
                nx = (md < 0.0f) ? (fX + md) : md;
            }
            else if (!(nx >= 0.0f && nx < fX))
            {
                nx = o.x;
            }
            if (wrapY)
            {
                float md = fmodf(ny, fY);
                ny = (md < 0.0f) ? (fY + md) : md;
            }
            else if (!(ny >= 0.0f && ny < fY))
            {
                ny = o.y;
            }
        }
        return make_float2(nx, ny);
    }

    // Moves slot k's agent and counts it into the next binning.
    __device__ __forceinline__ void MoveAndCount(const MoveArgs &m, int k, float2 o, float2 v, int32_t X, int32_t Y,
                                                 bool wrapX, bool wrapY)
    {
        const float2 q = MovedPosition(m, k, o, v, X, Y, wrapX, wrapY);
        m.posNext[k] = q;
        const int cell = ClampCell(q.x, X) * Y + ClampCell(q.y, Y);
        m.ticketOf[k] = atomicAdd(&m.counts[cell], 1);
    }

    // One lane per slot k; writes the velocity delta of slot k's agent to
    // vel[k] (slot order) and, with move.on, its moved position to
    // move.posNext[k] and its count for the next binning.  Slots are sorted by cell x*Y + y, so the cells
    // (cx, ya .. yb-1) of one neighbour row are the contiguous slot range
    // [offsets[cx*Y+ya], offsets[cx*Y+yb]); visiting the rows x-major and the
    // wrapped y segments in increasing ys keeps the reference's order.
    // Candidates per neighbour row loaded up front by the force pass's
    // common case (about 2.9 agents per 3-cell row at one agent per cell).
    constexpr int kHead = 4;

    // UR: every agent has radius r0, and no radius is loaded.
    template <bool WX, bool WY, bool UR>
    __global__ __launch_bounds__(kThreads) void AgentForces(int32_t N, int32_t X, int32_t Y, float R, float kSpring,
                                                            const int32_t *__restrict__ offsets,
                                                            const float2 *__restrict__ spos,
                                                            const float *__restrict__ srad, float r0,
                                                            float2 *__restrict__ vel, MoveArgs move,
                                                            const int32_t *__restrict__ scanFault)
    {
        // Each XCD takes a contiguous range of slots: the neighbour rows a
        // slot reads (about Y slots away) then sit in the same L2.
        const int k = XcdContiguousBlock(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
        if (k >= N || *scanFault != 0)
        {
            return; // (no neighbour ranges from a scan that gave up)
        }
        const float2 p = spos[k];
        const float ri = UR ? r0 : srad[k];
        const float fX = static_cast<float>(X);
        const float fY = static_cast<float>(Y);
        const int x0 = static_cast<int>(floorf(p.x - R));
        const int x1 = static_cast<int>(ceilf(p.x + R));
        const int y0 = static_cast<int>(floorf(p.y - R));
        const int y1 = static_cast<int>(ceilf(p.y + R));
        float vx = 0.0f;
        float vy = 0.0f;
        if (!WX && !WY && x1 - x0 <= 3)
        {
            // Up to three neighbour rows: all six slot bounds are requested
            // together, then the rows are summed x-major as below.
            const int ya = max(y0, 0);
            const int yb = min(y1, Y);
            int rs[3];
            int re[3];
#pragma unroll
            for (int r = 0; r < 3; r++)
            {
                const int xs = x0 + r;
                const bool ok = xs < x1 && xs >= 0 && xs < X && ya < yb;
                const int64_t rowBase = static_cast<int64_t>(ok ? xs : 0) * Y;
                rs[r] = ok ? offsets[rowBase + ya] : 0;
                re[r] = ok ? offsets[rowBase + yb] : 0;
            }
            // The first kHead candidates of every row are requested before
            // any is used (one memory round trip for most agents instead of
            // one per pair); the rest of a row follows in pairs.  The sum
            // keeps slot order within each row and rows x-major.
            float2 o[3][kHead];
            float rq[3][kHead];
#pragma unroll
            for (int r = 0; r < 3; r++)
            {
#pragma unroll
                for (int j = 0; j < kHead; j++)
                {
                    const int q = rs[r] + j;
                    o[r][j] = make_float2(0.0f, 0.0f);
                    rq[r][j] = 0.0f;
                    if (q < re[r])
                    {
                        o[r][j] = spos[q];
                        rq[r][j] = UR ? ri : srad[q];
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < 3; r++)
            {
#pragma unroll
                for (int j = 0; j < kHead; j++)
                {
                    if (rs[r] + j < re[r])
                    {
                        AccumulateOne(rs[r] + j, o[r][j], rq[r][j], k, p, ri, R, kSpring, fX, fY, false, false, vx,
                                      vy);
                    }
                }
                AccumulateRange<UR>(rs[r] + kHead, re[r], k, p, ri, R, kSpring, fX, fY, false, false, spos, srad,
                                    vx, vy);
            }
            vel[k] = make_float2(vx, vy);
            if (move.on)
            {
                MoveAndCount(move, k, p, make_float2(vx, vy), X, Y, false, false);
            }
            return;
        }
        for (int xs = x0; xs < x1; xs++)
        {
            int cx = xs;
            if (cx < 0 || cx >= X)
            {
                if (!WX)
                {
                    continue;
                }
                cx = WrapIndex(cx, X);
            }
            const int64_t rowBase = static_cast<int64_t>(cx) * Y;
            if (!WY)
            {
                const int ya = max(y0, 0);
                const int yb = min(y1, Y);
                if (ya < yb)
                {
                    AccumulateRange<UR>(offsets[rowBase + ya], offsets[rowBase + yb], k, p, ri, R, kSpring, fX, fY,
                                        WX, WY, spos, srad, vx, vy);
                }
                continue;
            }
            for (int ys = y0; ys < y1;)
            {
                const int cy = WrapIndex(ys, Y);
                const int len = min(y1 - ys, Y - cy);
                AccumulateRange<UR>(offsets[rowBase + cy], offsets[rowBase + cy + len], k, p, ri, R, kSpring, fX,
                                    fY, WX, WY, spos, srad, vx, vy);
                ys += len;
            }
        }
        vel[k] = make_float2(vx, vy);
        if (move.on)
        {
            MoveAndCount(move, k, p, make_float2(vx, vy), X, Y, WX, WY);
        }
    }

    unsigned Blocks(int64_t n, int threads)
    {
        return static_cast<unsigned>(std::max<int64_t>(1, (n + threads - 1) / threads));
    }
}

// ---- bins without allocation (halhip_agents2d_bins_noalloc) ----------------
//
// For callers that cannot allocate device memory (code under graph capture,
// serving loops): the binning's working memory is one file-scope device
// array per GPU, `scratch`, of fixed size.  Host code reaches it only through
// the address hipGetSymbolAddress returns on that device (kept per device by
// InitBinsScratch, called when a context is created); the array's name in
// host code is not a device address.  Layout in int32 words:
//   counts  [kNaMaxBins]         per-bin counts (whole scan tiles)
//   offsets [kNaMaxBins + 64]    exclusive prefix, padded like the plan's
//   ticket  [kNaMaxAgents]       per-agent ticket from the count atomics
//   tiles   [kNaMaxBins / tile]  per-tile totals
namespace
{
    constexpr int64_t kNaMaxBins = int64_t(1) << 20;
    constexpr int64_t kNaMaxAgents = int64_t(1) << 20;
    constexpr int64_t kNaCounts = 0;
    constexpr int64_t kNaOffsets = kNaCounts + kNaMaxBins;
    constexpr int64_t kNaTicket = kNaOffsets + kNaMaxBins + 64;
    constexpr int64_t kNaTiles = kNaTicket + kNaMaxAgents;
    constexpr int64_t kNaWords = kNaTiles + kNaMaxBins / kScanTile;
    static_assert(kNaMaxBins % kScanTile == 0, "bins must fill whole scan tiles");
    static_assert(kNaMaxBins / kScanTile <= 4096, "the scan's tile prefix must fit one pass");

    static __device__ __align__(256) int32_t scratch[kNaWords];

    constexpr int kNaMaxDevices = 64;
    std::mutex g_naMutex;
    int32_t *g_naScratch[kNaMaxDevices] = {};
    hipError_t g_naLookupError[kNaMaxDevices] = {};

    __global__ __launch_bounds__(kThreads) void FillZero4(int4 *__restrict__ p, int64_t n4)
    {
        const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
        if (i < n4)
        {
            p[i] = make_int4(0, 0, 0, 0);
        }
    }

    // Per agent: ticket = atomicAdd(counts[cell], 1).
    __global__ __launch_bounds__(kThreads) void AgentTickets(int32_t N, int32_t X, int32_t Y,
                                                             const float2 *__restrict__ pos,
                                                             int32_t *__restrict__ counts, int32_t *__restrict__ ticket)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= N)
        {
            return;
        }
        const float2 p = pos[i];
        ticket[i] = atomicAdd(&counts[ClampCell(p.x, X) * Y + ClampCell(p.y, Y)], 1);
    }

    // bucket[offsets[cell] + ticket] = agent index (a permutation of 0..N-1).
    __global__ __launch_bounds__(kThreads) void AgentScatterIndex(int32_t N, int32_t X, int32_t Y,
                                                                  const float2 *__restrict__ pos,
                                                                  const int32_t *__restrict__ ticket,
                                                                  const int32_t *__restrict__ offsets,
                                                                  int32_t *__restrict__ bucket)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= N)
        {
            return;
        }
        const float2 p = pos[i];
        bucket[offsets[ClampCell(p.x, X) * Y + ClampCell(p.y, Y)] + ticket[i]] = i;
    }

    // Each bin's slice of bucket sorted ascending (insertion sort: slices
    // hold a handful of agents), one bin per thread.
    __global__ __launch_bounds__(kThreads) void AgentOrderIndex(int64_t C, const int32_t *__restrict__ offsets,
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

    // The device address of `scratch` on ctx's device (set at context creation).
    int32_t *BinsScratch(const halhip_context *ctx)
    {
        std::lock_guard<std::mutex> lock(g_naMutex);
        int32_t *ptr = (ctx->device >= 0 && ctx->device < kNaMaxDevices) ? g_naScratch[ctx->device] : nullptr;
        if (ptr == nullptr)
        {
            const hipError_t e = (ctx->device >= 0 && ctx->device < kNaMaxDevices) ? g_naLookupError[ctx->device]
                                                                                    : hipSuccess;
            Fail(HALHIP_ERR_HIP, std::string("halhip_agents2d_bins_noalloc: the device address of the scratch array "
                                             "is not known: ") +
                                     (e != hipSuccess ? hipGetErrorString(e) : "no context was created on this device"));
        }
        return ptr;
    }
}

namespace halhip
{
    void InitBinsScratch(int32_t device)
    {
        std::lock_guard<std::mutex> lock(g_naMutex);
        if (device < 0 || device >= kNaMaxDevices || g_naScratch[device] != nullptr)
        {
            return;
        }
        void *ptr = nullptr;
        // The device is current (halhip_context_create bound it).  A failed
        // lookup leaves the address unknown, which only the no-allocation
        // entry point reports; every other entry point stays usable.
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
        {
            (void)hipGetLastError();
            g_naLookupError[device] = e;
            return;
        }
        g_naScratch[device] = static_cast<int32_t *>(ptr);
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
    // Two buffer sets of (id, position, radius) in slot order; `cur` holds
    // the resident agents, the other receives the next binning's scatter.
    DeviceArray<int32_t> id[2];
    DeviceArray<float2> pos[2];
    DeviceArray<float> rad[2];
    // Radii change only on upload.  When they are all equal (radius0), the
    // force pass takes that value instead of loading each candidate's, and
    // the binning moves no radii (rad[] then holds the upload's set only).
    bool uniformRadius = false;
    float radius0 = 0.0f;
    int32_t cur = 0;
    DeviceArray<float2> vel;    // slot order of set `cur`
    // Moved positions, slot order of set `cur`: after a tick the resident
    // positions are here (moved == true), already counted into
    // counts / ticketOf, until the next binning scatters them into
    // the other set.
    DeviceArray<float2> posNext;
    bool moved = false;
    // Carried velocity (slot order, per buffer set), kept while friction or
    // a speed cap is set; allocated on first use.
    DeviceArray<float2> carry[2];
    bool motion = false;
    float friction = 0.0f;
    float maxVel = 0.0f;
    DeviceArray<float2> outBuf; // agent-index order, for downloads
    DeviceArray<int32_t> counts;
    DeviceArray<int32_t> offsets;
    DeviceArray<int32_t> tileSums;
    // one-pass scan: published tile words and [ticket, finished, epoch, fault]
    DeviceArray<unsigned long long> scanTiles;
    DeviceArray<int32_t> scanWords;
    int32_t scanSpinLimit = 1 << 20;
    DeviceArray<int32_t> ticketOf;
    bool binsValid = false;
    StreamClock clock;
    Profiler prof;
    GraphCache graphs;
    bool prepareOnly = false;   // inside halhip_agents2d_prepare_ticks
    HostStaging<float> hostBuf; // pinned staging of the per-call FP64 paths
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

    // Published words cleared, epoch 1, no fault (synchronous).
    void ResetScanState(halhip_agents2d *p)
    {
        Check(hipMemsetAsync(p->scanTiles.get(), 0, static_cast<size_t>(p->nTiles) * sizeof(unsigned long long),
                             p->clock.stream),
              "hipMemsetAsync");
        const int32_t words[4] = {0, 0, 1, 0};
        Check(hipMemcpyAsync(p->scanWords.get(), words, sizeof(words), hipMemcpyHostToDevice, p->clock.stream),
              "hipMemcpyAsync");
        // the order pass, which resets the counts, stopped on the fault
        Check(hipMemsetAsync(p->counts.get(), 0, static_cast<size_t>(p->nTiles) * kScanTile * sizeof(int32_t),
                             p->clock.stream),
              "hipMemsetAsync");
        Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
    }

    // Reports (and clears) a bin scan whose wait gave up since the last
    // check; the bins it produced are not to be trusted.
    void CheckScanFault(halhip_agents2d *p, const char *caller)
    {
        int32_t fault = 0;
        Check(hipMemcpyAsync(&fault, p->scanWords.get() + 3, sizeof(fault), hipMemcpyDeviceToHost, p->clock.stream),
              "hipMemcpyAsync");
        Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
        if (fault != 0)
        {
            // a scan that gave up may have left words a later scan could
            // take for its own: clear them all and restart the epochs
            ResetScanState(p);
            p->binsValid = false;
            Fail(HALHIP_ERR_HIP, std::string(caller) + ": the bin scan's look-back wait gave up "
                                                       "(spin limit reached); bins and positions since the last "
                                                       "check are invalid");
        }
    }

    // Resident positions in slot order of set `cur`.
    float2 *CurPositions(halhip_agents2d *p)
    {
        return p->moved ? p->posNext.get() : p->pos[p->cur].get();
    }

    // Bins of the resident positions: counts, offsets, and the agents moved
    // to the other buffer set in bucket order, which becomes the resident set.
    void EnqueueBins(halhip_agents2d *p)
    {
        const int32_t N = p->N;
        const int c = p->cur;
        const int n = 1 - c;
        const float2 *positions = CurPositions(p);
        if (N > 0 && !p->moved)
        {
            // (moved positions were counted by the move itself)
            Launch(p, "agents_count", AgentCount, Blocks(N, kThreads), kThreads, N, p->X, p->Y, positions,
                   p->counts.get(), p->ticketOf.get());
        }
        int4 *counts4 = reinterpret_cast<int4 *>(p->counts.get());
        const int64_t cPad = static_cast<int64_t>(p->nTiles) * kScanTile;
        ScanState st{p->scanTiles.get(), p->scanWords.get(), p->scanSpinLimit};
        Launch(p, "agents_scan", ScanOnePass, static_cast<unsigned>(p->nTiles), kScanThreads, counts4,
               reinterpret_cast<int4 *>(p->offsets.get()), p->offsets.get() + cPad, N, p->nTiles, st);
        if (N > 0)
        {
            Launch(p, "agents_scatter", AgentScatter, Blocks(N, kThreads), kThreads, N,
                   p->X, p->Y, static_cast<const int32_t *>(p->ticketOf.get()),
                   static_cast<const int32_t *>(p->offsets.get()), static_cast<const int32_t *>(p->id[c].get()),
                   positions, static_cast<const float *>(p->uniformRadius ? nullptr : p->rad[c].get()),
                   static_cast<const float2 *>(p->motion ? p->carry[c].get() : nullptr), p->id[n].get(),
                   p->pos[n].get(), p->rad[n].get(), p->motion ? p->carry[n].get() : nullptr,
                   static_cast<const int32_t *>(p->scanWords.get() + 3));
            Launch(p, "agents_order", AgentOrderCells, Blocks((p->C + 3) / 4, kThreads), kThreads, p->C,
                   static_cast<const int32_t *>(p->offsets.get()), p->id[n].get(), p->pos[n].get(),
                   p->uniformRadius ? nullptr : p->rad[n].get(), p->motion ? p->carry[n].get() : nullptr,
                   p->counts.get(), static_cast<const int32_t *>(p->scanWords.get() + 3));
        }
        p->cur = n;
        p->moved = false;
    }

    // Forces on the freshly binned set `cur`; with move, the moved positions
    // go to posNext (the force pass reads every neighbour's old position).
    void EnqueueForces(halhip_agents2d *p, float R, float k, int32_t wrapX, int32_t wrapY, bool move, float dt)
    {
        if (p->N == 0)
        {
            return;
        }
        const int c = p->cur;
        MoveArgs m{};
        m.on = move ? 1 : 0;
        m.dt = dt;
        m.posNext = p->posNext.get();
        m.carry = p->motion ? p->carry[c].get() : nullptr;
        m.friction = p->friction;
        m.maxVel = p->maxVel;
        m.counts = p->counts.get();
        m.ticketOf = p->ticketOf.get();
        auto pick = [&](auto ur)
        {
            constexpr bool UR = decltype(ur)::value;
            return wrapX ? (wrapY ? AgentForces<true, true, UR> : AgentForces<true, false, UR>)
                         : (wrapY ? AgentForces<false, true, UR> : AgentForces<false, false, UR>);
        };
        auto kernel = p->uniformRadius ? pick(std::true_type{}) : pick(std::false_type{});
        Launch(p, "agents_force", kernel, Blocks(p->N, kThreads), kThreads, p->N, p->X, p->Y, R, k,
               static_cast<const int32_t *>(p->offsets.get()), static_cast<const float2 *>(p->pos[c].get()),
               static_cast<const float *>(p->rad[c].get()), p->radius0, p->vel.get(), m,
               static_cast<const int32_t *>(p->scanWords.get() + 3));
        p->moved = move;
    }

    void EnqueueTick(halhip_agents2d *p, float R, float k, float dt, int32_t wrapX, int32_t wrapY)
    {
        EnqueueBins(p);
        EnqueueForces(p, R, k, wrapX, wrapY, true, dt);
    }

    // Whether every radius equals the first (no early exit, so the loop
    // vectorizes).
    bool UniformRadii(const float *radii, size_t n)
    {
        bool same = true;
        for (size_t i = 1; i < n; i++)
        {
            same &= radii[i] == radii[0];
        }
        return same;
    }

    // Copies the resident agents in and resets slot k to agent k.  uniform:
    // every radius equals radii[0].
    void UploadAgents(halhip_agents2d *p, const float *positions, const float *radii, int32_t nAgents, bool uniform)
    {
        if (nAgents != p->N)
        {
            p->graphs.Clear();
        }
        if (p->moved)
        {
            // the last move counted its positions for a binning that will
            // not happen: start the counts from zero again
            const size_t cPad = static_cast<size_t>(p->nTiles) * kScanTile;
            Check(hipMemsetAsync(p->counts.get(), 0, cPad * sizeof(int32_t), p->clock.stream), "hipMemsetAsync");
        }
        p->N = nAgents;
        p->binsValid = false;
        p->moved = false;
        p->uniformRadius = false;
        if (nAgents == 0)
        {
            return;
        }
        p->uniformRadius = uniform;
        p->radius0 = radii[0];
        const int c = p->cur;
        Check(hipMemcpyAsync(p->pos[c].get(), positions, sizeof(float2) * nAgents, hipMemcpyHostToDevice,
                             p->clock.stream),
              "hipMemcpyAsync");
        Check(hipMemcpyAsync(p->rad[c].get(), radii, sizeof(float) * nAgents, hipMemcpyHostToDevice, p->clock.stream),
              "hipMemcpyAsync");
        hipLaunchKernelGGL(AgentIota, dim3(Blocks(nAgents, kThreads)), dim3(kThreads), 0, p->clock.stream, nAgents,
                           p->id[c].get());
        Check(hipGetLastError(), "AgentIota launch");
        if (p->motion)
        {
            Check(hipMemsetAsync(p->carry[c].get(), 0, sizeof(float2) * nAgents, p->clock.stream), "hipMemsetAsync");
        }
    }

    // Slot-order float2 values of the resident set to host memory in
    // agent-index order.
    void DownloadByIndex(halhip_agents2d *p, const float2 *slotValues, float *out)
    {
        hipLaunchKernelGGL(AgentUnscatter, dim3(Blocks(p->N, kThreads)), dim3(kThreads), 0, p->clock.stream, p->N,
                           static_cast<const int32_t *>(p->id[p->cur].get()), slotValues, p->outBuf.get());
        Check(hipGetLastError(), "AgentUnscatter launch");
        Check(hipMemcpyAsync(out, p->outBuf.get(), sizeof(float2) * p->N, hipMemcpyDeviceToHost, p->clock.stream),
              "hipMemcpyAsync");
        Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
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
            if (p->prepareOnly)
            {
                return;
            }
            EnqueueBins(p);
            EnqueueForces(p, fR, fk, wrapX, wrapY, false, 0.0f);
            p->binsValid = true;
            return;
        }
        // 10-tick graph chunks.  A chunk leaves `cur` as it found it (even
        // length) and the positions moved; the first chunk after an upload or
        // a forces-only pass starts from the binned set instead, so `moved`
        // is part of the key and each chunk is looked up on its own.  A
        // prepare call captures the chunk from both buffer sets, moved and
        // (if it is not) as it is now, and enqueues nothing.
        constexpr int32_t kChunk = 10;
        auto chunkGraph = [&](int32_t c0, bool moved0)
        {
            std::string key;
            AppendKey(key, p->N);
            AppendKey(key, c0);
            AppendKey(key, fR);
            AppendKey(key, fk);
            AppendKey(key, fdt);
            AppendKey(key, wrapX);
            AppendKey(key, wrapY);
            AppendKey(key, p->motion ? 1 : 0);
            AppendKey(key, p->friction);
            AppendKey(key, p->maxVel);
            AppendKey(key, moved0 ? 1 : 0);
            AppendKey(key, p->uniformRadius ? 1 : 0);
            AppendKey(key, p->radius0);
            const int32_t savedCur = p->cur;
            const bool savedMoved = p->moved;
            p->cur = c0;
            p->moved = moved0;
            // capturing runs the ticks' host-side bookkeeping too, which
            // leaves cur as it was and moved set
            hipGraphExec_t exec = FindOrCapture(p->graphs, key, p->clock.stream,
                                                [&]
                                                {
                                                    for (int32_t t = 0; t < kChunk; t++)
                                                    {
                                                        EnqueueTick(p, fR, fk, fdt, wrapX, wrapY);
                                                    }
                                                });
            p->cur = savedCur;
            p->moved = savedMoved;
            return exec;
        };
        int32_t done = 0;
        if (p->prepareOnly)
        {
            if (!p->prof.On() && ticks >= kChunk)
            {
                for (int32_t c0 = 0; c0 < 2; c0++)
                {
                    chunkGraph(c0, true);
                    if (!p->moved)
                    {
                        chunkGraph(c0, false);
                    }
                }
            }
            return;
        }
        while (!p->prof.On() && ticks - done >= kChunk)
        {
            hipGraphExec_t exec = chunkGraph(p->cur, p->moved);
            Check(hipGraphLaunch(exec, p->clock.stream), "hipGraphLaunch");
            p->moved = true;
            done += kChunk;
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
                for (int b = 0; b < 2; b++)
                {
                    p->id[b].Allocate(ctx, cap);
                    p->pos[b].Allocate(ctx, cap);
                    p->rad[b].Allocate(ctx, cap);
                }
                p->vel.Allocate(ctx, cap);
                p->posNext.Allocate(ctx, cap);
                p->outBuf.Allocate(ctx, cap);
                p->ticketOf.Allocate(ctx, cap);
                // counts and offsets padded to whole scan tiles; counts start
                // at 0 and the order pass resets them after every binning
                const size_t cPad = static_cast<size_t>(p->nTiles) * kScanTile;
                p->counts.Allocate(ctx, cPad);
                p->offsets.Allocate(ctx, cPad + 4);
                p->tileSums.Allocate(ctx, static_cast<size_t>(p->nTiles));
                p->scanTiles.Allocate(ctx, static_cast<size_t>(p->nTiles));
                p->scanWords.Allocate(ctx, 4);
                p->clock.Create();
                Check(hipMemsetAsync(p->counts.get(), 0, cPad * sizeof(int32_t), p->clock.stream), "hipMemsetAsync");
                // published words of epoch 0 never match the first scan's epoch 1
                ResetScanState(p);
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
                UploadAgents(p, positions, radii, nAgents, UniformRadii(radii, static_cast<size_t>(nAgents)));
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
                DownloadByIndex(p, CurPositions(p), positions);
                CheckScanFault(p, "halhip_agents2d_download_positions");
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
                DownloadByIndex(p, p->vel.get(), velocityDeltas);
                CheckScanFault(p, "halhip_agents2d_download_velocities");
            });
    }

    int halhip_agents2d_set_motion(halhip_agents2d *p, double friction, double maxVel)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_set_motion");
                Require(std::isfinite(friction) && friction >= 0.0 && friction <= 1.0,
                        "halhip_agents2d_set_motion: friction must lie in [0, 1]");
                Require(std::isfinite(maxVel) && maxVel >= 0.0,
                        "halhip_agents2d_set_motion: maxVel must be finite and non-negative");
                const bool on = friction > 0.0 || maxVel > 0.0;
                if (on && !p->motion)
                {
                    const size_t cap = static_cast<size_t>(std::max(1, p->capacity));
                    for (int b = 0; b < 2; b++)
                    {
                        if (p->carry[b].get() == nullptr)
                        {
                            p->carry[b].Allocate(p->ctx, cap);
                        }
                    }
                    // velocities start at rest
                    Check(hipMemsetAsync(p->carry[p->cur].get(), 0, sizeof(float2) * cap, p->clock.stream),
                          "hipMemsetAsync");
                    Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                }
                p->motion = on;
                p->friction = static_cast<float>(friction);
                p->maxVel = static_cast<float>(maxVel);
            });
    }

    int halhip_agents2d_upload_motion(halhip_agents2d *p, const float *velocities, int32_t nAgents)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_upload_motion");
                Require(p->motion, "halhip_agents2d_upload_motion: no friction or speed cap set");
                Require(nAgents == p->N, "halhip_agents2d_upload_motion: count differs from the uploaded count");
                Require(nAgents == 0 || velocities != nullptr, "halhip_agents2d_upload_motion: null array");
                if (nAgents == 0)
                {
                    return;
                }
                // agent-index order in; slot order on the device
                Check(hipMemcpyAsync(p->outBuf.get(), velocities, sizeof(float2) * nAgents, hipMemcpyHostToDevice,
                                     p->clock.stream),
                      "hipMemcpyAsync");
                hipLaunchKernelGGL(AgentGather, dim3(Blocks(nAgents, kThreads)), dim3(kThreads), 0, p->clock.stream,
                                   nAgents, static_cast<const int32_t *>(p->id[p->cur].get()),
                                   static_cast<const float2 *>(p->outBuf.get()), p->carry[p->cur].get());
                Check(hipGetLastError(), "AgentGather launch");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
            });
    }

    int halhip_agents2d_download_motion(halhip_agents2d *p, float *velocities, int32_t nAgents)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_download_motion");
                Require(p->motion, "halhip_agents2d_download_motion: no friction or speed cap set");
                Require(nAgents == p->N, "halhip_agents2d_download_motion: count differs from the uploaded count");
                Require(nAgents == 0 || velocities != nullptr, "halhip_agents2d_download_motion: null array");
                if (nAgents == 0)
                {
                    return;
                }
                DownloadByIndex(p, p->carry[p->cur].get(), velocities);
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
                    Check(hipMemcpyAsync(bucket, p->id[p->cur].get(), sizeof(int32_t) * p->N, hipMemcpyDeviceToHost,
                                         p->clock.stream),
                          "hipMemcpyAsync");
                }
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                CheckScanFault(p, "halhip_agents2d_download_bins");
            });
    }

    int halhip_agents2d_set_scan_spin_limit(halhip_agents2d *p, int32_t limit)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_set_scan_spin_limit");
                Require(limit >= 0, "halhip_agents2d_set_scan_spin_limit: negative limit");
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                p->scanSpinLimit = limit;
                p->graphs.Clear(); // they hold the old limit
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
                bool uniform = true;
                for (size_t i = 0; i < n; i++)
                {
                    p->hostBuf[2 * n + i] = static_cast<float>(radii[i]);
                    uniform &= p->hostBuf[2 * n + i] == p->hostBuf[2 * n];
                }
                UploadAgents(p, p->hostBuf.data(), p->hostBuf.data() + 2 * n, nAgents, uniform);
                EnqueueBins(p);
                EnqueueForces(p, static_cast<float>(interactionRad), static_cast<float>(springConstant), wrapX, wrapY,
                              false, 0.0f);
                p->binsValid = true;
                if (n > 0)
                {
                    DownloadByIndex(p, p->vel.get(), p->hostBuf.data());
                }
                Check(hipStreamSynchronize(p->clock.stream), "hipStreamSynchronize");
                p->prof.Collect();
                CheckScanFault(p, "halhip_agents2d_forces_f64");
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

    int halhip_agents2d_prepare_ticks(halhip_agents2d *p, int32_t ticks, double interactionRad,
                                      double springConstant, double dt, int32_t wrapX, int32_t wrapY)
    {
        return Guard(
            [&]
            {
                RequirePlan(p, "halhip_agents2d_prepare_ticks");
                PrepareScope scope(p->prepareOnly);
                EnqueueTicks(p, ticks, interactionRad, springConstant, dt, wrapX, wrapY);
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

    int halhip_agents2d_bins_noalloc(halhip_context *ctx, const float *positions, int32_t nAgents, int32_t xDim,
                                     int32_t yDim, int32_t *offsets, int32_t *bucket, void *stream)
    {
        return Guard(
            [&]
            {
                Require(ctx != nullptr, "halhip_agents2d_bins_noalloc: null context");
                Require(xDim > 0 && yDim > 0, "halhip_agents2d_bins_noalloc: dimensions must be positive");
                Require(nAgents >= 0, "halhip_agents2d_bins_noalloc: negative agent count");
                const int64_t C = static_cast<int64_t>(xDim) * yDim;
                Require(C <= kNaMaxBins, "halhip_agents2d_bins_noalloc: xDim*yDim exceeds "
                                         "HALHIP_AGENTS2D_NOALLOC_MAX_BINS (1048576)");
                Require(nAgents <= kNaMaxAgents, "halhip_agents2d_bins_noalloc: nAgents exceeds "
                                                 "HALHIP_AGENTS2D_NOALLOC_MAX_AGENTS (1048576)");
                Require(offsets != nullptr, "halhip_agents2d_bins_noalloc: null offsets");
                Require(nAgents == 0 || (positions != nullptr && bucket != nullptr),
                        "halhip_agents2d_bins_noalloc: null positions or bucket");
                Bind(ctx);
                int32_t *base = BinsScratch(ctx);
                hipStream_t s = static_cast<hipStream_t>(stream);
                int32_t *counts = base + kNaCounts;
                int32_t *offs = base + kNaOffsets;
                int32_t *ticket = base + kNaTicket;
                int32_t *tiles = base + kNaTiles;
                const int32_t nTiles = static_cast<int32_t>((C + kScanTile - 1) / kScanTile);
                const int64_t cPad = static_cast<int64_t>(nTiles) * kScanTile;
                // The previous call's counts may not have been reset (a call
                // can stop part-way), so they are cleared here, by a kernel:
                // a memset node would not replay reliably inside a graph.
                const int64_t n4 = cPad / 4;
                hipLaunchKernelGGL(FillZero4, dim3(Blocks(n4, kThreads)), dim3(kThreads), 0, s,
                                   reinterpret_cast<int4 *>(counts), n4);
                Check(hipGetLastError(), "FillZero4 launch");
                const float2 *pos = reinterpret_cast<const float2 *>(positions);
                if (nAgents > 0)
                {
                    hipLaunchKernelGGL(AgentTickets, dim3(Blocks(nAgents, kThreads)), dim3(kThreads), 0, s, nAgents,
                                       xDim, yDim, pos, counts, ticket);
                    Check(hipGetLastError(), "AgentTickets launch");
                }
                hipLaunchKernelGGL(ScanTileSums, dim3(static_cast<unsigned>(nTiles)), dim3(kScanThreads), 0, s,
                                   reinterpret_cast<const int4 *>(counts), tiles);
                Check(hipGetLastError(), "ScanTileSums launch");
                hipLaunchKernelGGL(ScanApply, dim3(static_cast<unsigned>(nTiles)), dim3(kScanThreads), 0, s,
                                   reinterpret_cast<int4 *>(counts), static_cast<const int32_t *>(tiles),
                                   reinterpret_cast<int4 *>(offs), offs + cPad, nAgents);
                Check(hipGetLastError(), "ScanApply launch");
                if (nAgents > 0)
                {
                    hipLaunchKernelGGL(AgentScatterIndex, dim3(Blocks(nAgents, kThreads)), dim3(kThreads), 0, s,
                                       nAgents, xDim, yDim, pos, static_cast<const int32_t *>(ticket),
                                       static_cast<const int32_t *>(offs), bucket);
                    Check(hipGetLastError(), "AgentScatterIndex launch");
                    hipLaunchKernelGGL(AgentOrderIndex, dim3(Blocks(C, kThreads)), dim3(kThreads), 0, s, C,
                                       static_cast<const int32_t *>(offs), bucket);
                    Check(hipGetLastError(), "AgentOrderIndex launch");
                }
                Check(hipMemcpyAsync(offsets, offs, sizeof(int32_t) * static_cast<size_t>(C + 1),
                                     hipMemcpyDeviceToDevice, s),
                      "hipMemcpyAsync");
            });
    }
}
