// common.hpp -- runtime plumbing shared by every plan in libhalhip.so:
// status/error handling across the C ABI, the device context, owned device
// buffers, HIP-event kernel profiling and cached hipGraph replays.
#pragma once

#include "halhip.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <exception>
#include <map>
#include <new>
#include <string>
#include <utility>
#include <vector>

struct halhip_context
{
    int32_t device = 0;
    int32_t failNextAllocs = 0; // test hook (halhip_context_fail_next_allocs)
};

namespace halhip
{
    // Workgroup b of a launch of n goes to XCD b % 8 (the dispatcher deals
    // workgroups to the 8 XCDs in turn).  This maps b to a block index such
    // that each XCD runs one contiguous range of [0, n) (a bijection for any
    // n), so work items that share data (neighbouring slots, rows or cells)
    // meet in one XCD's L2.
    __device__ __forceinline__ int XcdContiguousBlock(int b, int n)
    {
        const int j = b & 7;
        const int q = n >> 3;
        const int rem = n & 7;
        return j * q + (j < rem ? j : rem) + (b >> 3);
    }

    // Error carried from the implementation to the C ABI, where guard()
    // turns it into a status code plus the thread's last-error string.
    struct Error : std::exception
    {
        int code;
        std::string message;
        Error(int c, std::string m) : code(c), message(std::move(m)) {}
        const char *what() const noexcept override { return message.c_str(); }
    };

    void SetLastError(const std::string &message);

    [[noreturn]] inline void Fail(int code, const std::string &message)
    {
        throw Error(code, message);
    }

    inline void Require(bool ok, const std::string &message)
    {
        if (!ok)
        {
            Fail(HALHIP_ERR_INVALID_ARGUMENT, message);
        }
    }

    inline void Check(hipError_t e, const char *what)
    {
        if (e != hipSuccess)
        {
            int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? HALHIP_ERR_OUT_OF_MEMORY
                                                                                     : HALHIP_ERR_HIP;
            Fail(code, std::string(what) + ": " + hipGetErrorString(e));
        }
    }

    // Runs f, mapping any exception to a status code and clearing or setting
    // the calling thread's last error.
    template <class F>
    int Guard(F &&f)
    {
        try
        {
            f();
            SetLastError("");
            return HALHIP_OK;
        }
        catch (const Error &e)
        {
            SetLastError(e.message);
            return e.code;
        }
        catch (const std::bad_alloc &)
        {
            SetLastError("host allocation failed");
            return HALHIP_ERR_OUT_OF_MEMORY;
        }
        catch (const std::exception &e)
        {
            SetLastError(e.what());
            return HALHIP_ERR_HIP;
        }
    }

    // Looks up, once per device and with that device current, the device
    // address of the no-allocation binning's scratch array (agents2d.hip).
    void InitBinsScratch(int32_t device);

    // d[0, X*Y) += ADI(u) - u: one ADI tick (no decay or sources) of the
    // device field u, run by the ADI plan on its own stream after the work
    // queued on `caller` (adi2d.hip); `caller` then waits for it.  The plan's
    // resident field is replaced.
    void AdiDeltaOf(halhip_adi2d *adi, const float *u, float *d, int32_t X, int32_t Y, double rate, int32_t wrapX,
                    int32_t wrapY, const halhip_bc *bc, hipStream_t caller);

    // Makes ctx's device current for the calling thread.
    inline void Bind(const halhip_context *ctx)
    {
        Check(hipSetDevice(ctx->device), "hipSetDevice");
    }

    // Pinned (page-locked) host staging for the per-call FP64 interfaces:
    // copies between it and the device run at full link speed, which
    // pageable memory does not.  Grows only; the contents are not kept
    // across a resize (every user writes before it reads).
    template <typename T>
    class HostStaging
    {
    public:
        HostStaging() = default;
        HostStaging(const HostStaging &) = delete;
        HostStaging &operator=(const HostStaging &) = delete;
        ~HostStaging() { Release(); }

        void resize(size_t count)
        {
            if (count <= m_count)
            {
                return;
            }
            Release();
            Check(hipHostMalloc(reinterpret_cast<void **>(&m_ptr), count * sizeof(T), hipHostMallocDefault),
                  "hipHostMalloc");
            m_count = count;
        }
        void Release()
        {
            if (m_ptr != nullptr)
            {
                (void)hipHostFree(m_ptr);
                m_ptr = nullptr;
                m_count = 0;
            }
        }
        T *data() { return m_ptr; }
        T &operator[](size_t k) { return m_ptr[k]; }
        size_t size() const { return m_count; }

    private:
        T *m_ptr = nullptr;
        size_t m_count = 0;
    };

    // The per-call FP64 paths' copies in pieces of at least kMinPiece values
    // (at most kChunks) through pinned staging: the host converts piece c+1
    // while piece c is copied up, and adds piece c into the caller's array
    // while piece c+1 comes down.  Smaller arrays move in one piece, where a
    // copy's fixed cost would outweigh the overlap.
    class ChunkedTransfer
    {
    public:
        static constexpr int kChunks = 8;
        static constexpr size_t kMinPiece = 256 * 1024;
        ChunkedTransfer() = default;
        ChunkedTransfer(const ChunkedTransfer &) = delete;
        ChunkedTransfer &operator=(const ChunkedTransfer &) = delete;
        ~ChunkedTransfer()
        {
            for (hipEvent_t e : m_ev)
            {
                if (e != nullptr)
                {
                    (void)hipEventDestroy(e);
                }
            }
        }

        // dev[0, n) = float(src[0, n)), enqueued on s.
        void Upload(const double *src, size_t n, float *staging, float *dev, hipStream_t s)
        {
            const int pieces = Pieces(n);
            for (int c = 0; c < pieces; c++)
            {
                const size_t b = Begin(n, c, pieces);
                const size_t e = Begin(n, c + 1, pieces);
                for (size_t k = b; k < e; k++)
                {
                    staging[k] = static_cast<float>(src[k]);
                }
                if (e > b)
                {
                    Check(hipMemcpyAsync(dev + b, staging + b, (e - b) * sizeof(float), hipMemcpyHostToDevice, s),
                          "hipMemcpyAsync");
                }
            }
        }

        // dst[k] += double(dev[k]) (minus base[k] when base is given) once
        // the work queued on s before this call has finished; returns when
        // every piece has been added.
        void DownloadAdd(const float *dev, size_t n, float *staging, double *dst, const double *base, hipStream_t s)
        {
            if (m_ev[0] == nullptr)
            {
                for (hipEvent_t &e : m_ev)
                {
                    Check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreateWithFlags");
                }
            }
            const int pieces = Pieces(n);
            for (int c = 0; c < pieces; c++)
            {
                const size_t b = Begin(n, c, pieces);
                const size_t e = Begin(n, c + 1, pieces);
                if (e > b)
                {
                    Check(hipMemcpyAsync(staging + b, dev + b, (e - b) * sizeof(float), hipMemcpyDeviceToHost, s),
                          "hipMemcpyAsync");
                }
                if (pieces > 1)
                {
                    Check(hipEventRecord(m_ev[c], s), "hipEventRecord");
                }
            }
            for (int c = 0; c < pieces; c++)
            {
                if (pieces > 1)
                {
                    Check(hipEventSynchronize(m_ev[c]), "hipEventSynchronize");
                }
                else
                {
                    Check(hipStreamSynchronize(s), "hipStreamSynchronize");
                }
                const size_t b = Begin(n, c, pieces);
                const size_t e = Begin(n, c + 1, pieces);
                if (base != nullptr)
                {
                    for (size_t k = b; k < e; k++)
                    {
                        dst[k] += static_cast<double>(staging[k]) - base[k];
                    }
                }
                else
                {
                    for (size_t k = b; k < e; k++)
                    {
                        dst[k] += static_cast<double>(staging[k]);
                    }
                }
            }
        }

    private:
        static int Pieces(size_t n)
        {
            const size_t p = n / kMinPiece;
            return static_cast<int>(p < 1 ? 1 : (p > static_cast<size_t>(kChunks) ? kChunks : p));
        }
        static size_t Begin(size_t n, int c, int pieces) { return n * static_cast<size_t>(c) / pieces; }
        hipEvent_t m_ev[kChunks] = {};
    };

    // Device allocation owned by a plan; freed on destruction.
    template <class T>
    class DeviceArray
    {
    public:
        DeviceArray() = default;
        DeviceArray(const DeviceArray &) = delete;
        DeviceArray &operator=(const DeviceArray &) = delete;
        ~DeviceArray() { Release(); }

        void Allocate(halhip_context *ctx, size_t count)
        {
            Release();
            if (ctx->failNextAllocs > 0)
            {
                ctx->failNextAllocs--;
                Fail(HALHIP_ERR_OUT_OF_MEMORY, "device allocation failed (forced by halhip_context_fail_next_allocs)");
            }
            size_t bytes = (count == 0 ? 1 : count) * sizeof(T);
            Check(hipMalloc(reinterpret_cast<void **>(&m_ptr), bytes), "hipMalloc");
            m_count = count;
        }
        void Release()
        {
            if (m_ptr != nullptr)
            {
                (void)hipFree(m_ptr);
                m_ptr = nullptr;
                m_count = 0;
            }
        }
        T *get() const { return m_ptr; }
        size_t size() const { return m_count; }
        void SwapWith(DeviceArray &other)
        {
            std::swap(m_ptr, other.m_ptr);
            std::swap(m_count, other.m_count);
        }

    private:
        T *m_ptr = nullptr;
        size_t m_count = 0;
    };

    // Per-kernel device time accumulated with HIP events bracketing each
    // launch on the plan's stream.  Off by default; when off Begin/End are
    // no-ops and launches are not serialised by event records.
    class Profiler
    {
    public:
        ~Profiler();
        void SetOn(bool on) { m_on = on; }
        bool On() const { return m_on; }
        void Begin(hipStream_t s, const char *name);
        void End(hipStream_t s);
        // Waits for the recorded events and folds them into the totals.
        void Collect();
        void Reset();
        int Export(int32_t maxKernels, int32_t *nKernels, char *names, int32_t namesLen, double *totalMs,
                   int64_t *launches);
        // Total milliseconds collected under `name` (0 if none).
        double TotalMs(const std::string &name) const
        {
            for (const auto &t : m_totals)
            {
                if (t.first == name)
                {
                    return t.second.first;
                }
            }
            return 0.0;
        }

    private:
        struct Pending
        {
            std::string name;
            hipEvent_t start;
            hipEvent_t stop;
        };
        hipEvent_t Take();
        bool m_on = false;
        std::vector<Pending> m_pending;
        std::vector<hipEvent_t> m_free;
        std::vector<std::pair<std::string, std::pair<double, int64_t>>> m_totals;
    };

    // A stream plus elapsed-time events for the async entry points.
    struct StreamClock
    {
        hipStream_t stream = nullptr;
        hipEvent_t start = nullptr;
        hipEvent_t stop = nullptr;
        bool pending = false;
        void Create();
        void Destroy();
        void Mark() { Check(hipEventRecord(start, stream), "hipEventRecord"); pending = true; }
        void Sync(float *elapsedMs);
    };

    // Instantiated graphs replaying a fixed launch sequence, keyed by a byte
    // string describing the launch parameters.
    class GraphCache
    {
    public:
        ~GraphCache() { Clear(); }
        hipGraphExec_t Find(const std::string &key) const;
        void Insert(const std::string &key, hipGraphExec_t exec);
        void Clear();

    private:
        std::map<std::string, hipGraphExec_t> m_execs;
    };

    template <class T>
    inline void AppendKey(std::string &key, const T &v)
    {
        key.append(reinterpret_cast<const char *>(&v), sizeof(T));
    }

    // The cached graph for `key`, or a new one holding the launches body()
    // enqueues on s (captured, instantiated and cached).  Host-side state
    // that body() changes while it is captured changes as if it had run.
    template <class F>
    hipGraphExec_t FindOrCapture(GraphCache &cache, const std::string &key, hipStream_t s, F body)
    {
        hipGraphExec_t exec = cache.Find(key);
        if (exec != nullptr)
        {
            return exec;
        }
        hipGraph_t graph;
        Check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try
        {
            body();
        }
        catch (...)
        {
            if (hipStreamEndCapture(s, &graph) == hipSuccess)
            {
                (void)hipGraphDestroy(graph);
            }
            throw;
        }
        Check(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        Check(e, "hipGraphInstantiate");
        cache.Insert(key, exec);
        return exec;
    }

    // Sets a plan's "build graphs only" flag for one prepare call: the
    // run path then captures every graph a run of the same arguments would
    // replay and enqueues nothing.
    class PrepareScope
    {
    public:
        explicit PrepareScope(bool &flag) : m_flag(flag) { m_flag = true; }
        ~PrepareScope() { m_flag = false; }
        PrepareScope(const PrepareScope &) = delete;
        PrepareScope &operator=(const PrepareScope &) = delete;

    private:
        bool &m_flag;
    };

    // Host copy of a Dirichlet description expanded to four edge arrays.
    struct EdgeValues
    {
        bool dirichlet = false;
        std::vector<float> xlo, xhi, ylo, yhi; // [Y], [Y], [X], [X]
        void From(const halhip_bc *bc, int32_t X, int32_t Y);
        void AppendTo(std::string &key) const;
    };

    // HAL's PDEGrid2D file layout (src/Grids/PDEGridSerialization.cpp:33-61,
    // raw little-endian as BinaryWriter writes it): int32 xDim, int32 yDim,
    // bool wrapX, bool wrapY, int32 updateCt, xDim*yDim FP64 field values,
    // xDim*yDim FP64 deltas.  Resident plans hold committed state, so they
    // save zero deltas and refuse files whose deltas are not all zero.
    void SavePdeGrid2D(const char *path, int32_t X, int32_t Y, bool wrapX, bool wrapY, int32_t updateCt,
                       const std::vector<float> &field);
    void LoadPdeGrid2D(const char *path, int32_t X, int32_t Y, bool &wrapX, bool &wrapY, int32_t &updateCt,
                       std::vector<float> &field);

    // Number of bytes in a flat field, checked against the plan's shape.
    inline void RequireLength(int64_t n, int64_t expect, const char *caller)
    {
        Require(n == expect, std::string(caller) + ": field and deltas must each hold xDim*yDim values");
    }
}
