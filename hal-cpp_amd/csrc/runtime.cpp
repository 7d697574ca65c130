// runtime.cpp -- context, error state, profiling and graph cache of
// libhalhip.so, and the host-side seeded input generator.
#include "common.hpp"

#include <cstdint>
#include <cstdio>

namespace halhip
{
    namespace
    {
        thread_local std::string t_lastError;
    }

    void SetLastError(const std::string &message)
    {
        t_lastError = message;
    }

    void SavePdeGrid2D(const char *path, int32_t X, int32_t Y, bool wrapX, bool wrapY, int32_t updateCt,
                       const std::vector<float> &field)
    {
        Require(path != nullptr, "save_grid: null path");
        std::FILE *f = std::fopen(path, "wb");
        Require(f != nullptr, std::string("save_grid: cannot open ") + path + " for writing");
        const unsigned char wx = wrapX ? 1 : 0;
        const unsigned char wy = wrapY ? 1 : 0;
        bool ok = std::fwrite(&X, sizeof(X), 1, f) == 1 && std::fwrite(&Y, sizeof(Y), 1, f) == 1 &&
                  std::fwrite(&wx, 1, 1, f) == 1 && std::fwrite(&wy, 1, 1, f) == 1 &&
                  std::fwrite(&updateCt, sizeof(updateCt), 1, f) == 1;
        std::vector<double> row(static_cast<size_t>(Y));
        const size_t n = static_cast<size_t>(X) * static_cast<size_t>(Y);
        for (size_t base = 0; ok && base < n; base += row.size())
        {
            for (size_t k = 0; k < row.size(); k++)
            {
                row[k] = static_cast<double>(field[base + k]);
            }
            ok = std::fwrite(row.data(), sizeof(double), row.size(), f) == row.size();
        }
        std::fill(row.begin(), row.end(), 0.0);
        for (size_t base = 0; ok && base < n; base += row.size())
        {
            ok = std::fwrite(row.data(), sizeof(double), row.size(), f) == row.size();
        }
        ok = (std::fclose(f) == 0) && ok;
        Require(ok, std::string("save_grid: write failed: ") + path);
    }

    void LoadPdeGrid2D(const char *path, int32_t X, int32_t Y, bool &wrapX, bool &wrapY, int32_t &updateCt,
                       std::vector<float> &field)
    {
        Require(path != nullptr, "load_grid: null path");
        std::FILE *f = std::fopen(path, "rb");
        Require(f != nullptr, std::string("load_grid: cannot open ") + path);
        int32_t fx = 0;
        int32_t fy = 0;
        unsigned char wx = 0;
        unsigned char wy = 0;
        int32_t ct = 0;
        bool ok = std::fread(&fx, sizeof(fx), 1, f) == 1 && std::fread(&fy, sizeof(fy), 1, f) == 1 &&
                  std::fread(&wx, 1, 1, f) == 1 && std::fread(&wy, 1, 1, f) == 1 &&
                  std::fread(&ct, sizeof(ct), 1, f) == 1;
        // Internal::CheckLoadedDims (src/Grids/GridSerialization.cpp:14-32)
        ok = ok && fx > 0 && fy > 0 && static_cast<int64_t>(fx) * fy <= INT32_MAX;
        if (!ok)
        {
            std::fclose(f);
            Fail(HALHIP_ERR_INVALID_ARGUMENT, "load_grid: failed read or invalid dimensions in saved data");
        }
        if (fx != X || fy != Y)
        {
            std::fclose(f);
            Fail(HALHIP_ERR_INVALID_ARGUMENT, "load_grid: saved grid is " + std::to_string(fx) + "x" +
                                                  std::to_string(fy) + ", the plan " + std::to_string(X) + "x" +
                                                  std::to_string(Y));
        }
        const size_t n = static_cast<size_t>(X) * static_cast<size_t>(Y);
        field.resize(n);
        std::vector<double> row(static_cast<size_t>(Y));
        for (size_t base = 0; ok && base < n; base += row.size())
        {
            ok = std::fread(row.data(), sizeof(double), row.size(), f) == row.size();
            for (size_t k = 0; ok && k < row.size(); k++)
            {
                field[base + k] = static_cast<float>(row[k]);
            }
        }
        bool zeroDeltas = true;
        for (size_t base = 0; ok && base < n; base += row.size())
        {
            ok = std::fread(row.data(), sizeof(double), row.size(), f) == row.size();
            for (size_t k = 0; ok && k < row.size(); k++)
            {
                zeroDeltas = zeroDeltas && row[k] == 0.0;
            }
        }
        std::fclose(f);
        Require(ok, "load_grid: failed read or invalid dimensions in saved data");
        Require(zeroDeltas, "load_grid: the saved grid has pending deltas (saved between a stencil call and "
                            "Update); resident plans hold committed state only");
        wrapX = wx != 0;
        wrapY = wy != 0;
        updateCt = ct;
    }

    Profiler::~Profiler()
    {
        for (auto &p : m_pending)
        {
            (void)hipEventDestroy(p.start);
            (void)hipEventDestroy(p.stop);
        }
        for (auto e : m_free)
        {
            (void)hipEventDestroy(e);
        }
    }

    hipEvent_t Profiler::Take()
    {
        if (!m_free.empty())
        {
            hipEvent_t e = m_free.back();
            m_free.pop_back();
            return e;
        }
        hipEvent_t e;
        Check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }

    void Profiler::Begin(hipStream_t s, const char *name)
    {
        if (!m_on)
        {
            return;
        }
        Pending p{name, Take(), Take()};
        Check(hipEventRecord(p.start, s), "hipEventRecord");
        m_pending.push_back(p);
    }

    void Profiler::End(hipStream_t s)
    {
        if (!m_on || m_pending.empty())
        {
            return;
        }
        Check(hipEventRecord(m_pending.back().stop, s), "hipEventRecord");
    }

    void Profiler::Collect()
    {
        for (auto &p : m_pending)
        {
            Check(hipEventSynchronize(p.stop), "hipEventSynchronize");
            float ms = 0.0f;
            Check(hipEventElapsedTime(&ms, p.start, p.stop), "hipEventElapsedTime");
            bool found = false;
            for (auto &t : m_totals)
            {
                if (t.first == p.name)
                {
                    t.second.first += ms;
                    t.second.second += 1;
                    found = true;
                    break;
                }
            }
            if (!found)
            {
                m_totals.push_back({p.name, {static_cast<double>(ms), 1}});
            }
            m_free.push_back(p.start);
            m_free.push_back(p.stop);
        }
        m_pending.clear();
    }

    void Profiler::Reset()
    {
        Collect();
        m_totals.clear();
    }

    int Profiler::Export(int32_t maxKernels, int32_t *nKernels, char *names, int32_t namesLen, double *totalMs,
                         int64_t *launches)
    {
        Collect();
        std::string joined;
        int32_t n = 0;
        for (auto &t : m_totals)
        {
            if (n >= maxKernels)
            {
                break;
            }
            if (!joined.empty())
            {
                joined += ",";
            }
            joined += t.first;
            totalMs[n] = t.second.first;
            launches[n] = t.second.second;
            n++;
        }
        *nKernels = n;
        if (namesLen > 0)
        {
            std::snprintf(names, static_cast<size_t>(namesLen), "%s", joined.c_str());
        }
        return HALHIP_OK;
    }

    void StreamClock::Create()
    {
        Check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        Check(hipEventCreate(&start), "hipEventCreate");
        Check(hipEventCreate(&stop), "hipEventCreate");
    }

    void StreamClock::Destroy()
    {
        if (stream != nullptr)
        {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
            stream = nullptr;
        }
        if (start != nullptr)
        {
            (void)hipEventDestroy(start);
            start = nullptr;
        }
        if (stop != nullptr)
        {
            (void)hipEventDestroy(stop);
            stop = nullptr;
        }
    }

    void StreamClock::Sync(float *elapsedMs)
    {
        if (pending)
        {
            Check(hipEventRecord(stop, stream), "hipEventRecord");
        }
        Check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (elapsedMs != nullptr)
        {
            *elapsedMs = 0.0f;
            if (pending)
            {
                Check(hipEventElapsedTime(elapsedMs, start, stop), "hipEventElapsedTime");
            }
        }
        pending = false;
    }

    hipGraphExec_t GraphCache::Find(const std::string &key) const
    {
        auto it = m_execs.find(key);
        return (it == m_execs.end()) ? nullptr : it->second;
    }

    void GraphCache::Insert(const std::string &key, hipGraphExec_t exec)
    {
        if (m_execs.size() >= 16)
        {
            Clear();
        }
        m_execs[key] = exec;
    }

    void GraphCache::Clear()
    {
        for (auto &kv : m_execs)
        {
            (void)hipGraphExecDestroy(kv.second);
        }
        m_execs.clear();
    }

    void EdgeValues::From(const halhip_bc *bc, int32_t X, int32_t Y)
    {
        dirichlet = (bc != nullptr) && (bc->kind == HALHIP_BC_DIRICHLET);
        Require(bc == nullptr || bc->kind == HALHIP_BC_NONE || bc->kind == HALHIP_BC_DIRICHLET,
                "halhip_bc.kind must be HALHIP_BC_NONE or HALHIP_BC_DIRICHLET");
        xlo.assign(static_cast<size_t>(Y), 0.0f);
        xhi.assign(static_cast<size_t>(Y), 0.0f);
        ylo.assign(static_cast<size_t>(X), 0.0f);
        yhi.assign(static_cast<size_t>(X), 0.0f);
        if (!dirichlet)
        {
            return;
        }
        auto fill = [&](std::vector<float> &dst, const double *src)
        {
            for (size_t i = 0; i < dst.size(); i++)
            {
                dst[i] = static_cast<float>(src != nullptr ? src[i] : bc->value);
            }
        };
        fill(xlo, bc->xlo);
        fill(xhi, bc->xhi);
        fill(ylo, bc->ylo);
        fill(yhi, bc->yhi);
    }

    void EdgeValues::AppendTo(std::string &key) const
    {
        AppendKey(key, dirichlet);
        if (dirichlet)
        {
            key.append(reinterpret_cast<const char *>(xlo.data()), xlo.size() * sizeof(float));
            key.append(reinterpret_cast<const char *>(xhi.data()), xhi.size() * sizeof(float));
            key.append(reinterpret_cast<const char *>(ylo.data()), ylo.size() * sizeof(float));
            key.append(reinterpret_cast<const char *>(yhi.data()), yhi.size() * sizeof(float));
        }
    }
}

using namespace halhip;

extern "C"
{
    const char *halhip_last_error(void)
    {
        return t_lastError.c_str();
    }

    int halhip_device_count(int32_t *count)
    {
        return Guard(
            [&]
            {
                Require(count != nullptr, "halhip_device_count: null count");
                int n = 0;
                if (hipGetDeviceCount(&n) != hipSuccess)
                {
                    n = 0;
                }
                *count = n;
            });
    }

    int halhip_context_create(int32_t device, halhip_context **out)
    {
        return Guard(
            [&]
            {
                Require(out != nullptr, "halhip_context_create: null out");
                *out = nullptr;
                int n = 0;
                if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
                {
                    Fail(HALHIP_ERR_NO_DEVICE, "halhip_context_create: no GPU device is available");
                }
                Require(device >= 0 && device < n, "halhip_context_create: device index out of range");
                Check(hipSetDevice(device), "hipSetDevice");
                hipDeviceProp_t prop;
                Check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
                if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
                {
                    Fail(HALHIP_ERR_NO_DEVICE, std::string("halhip_context_create: device is ") + prop.gcnArchName +
                                                   ", this library is built for gfx950 only");
                }
                InitBinsScratch(device);
                auto *ctx = new halhip_context();
                ctx->device = device;
                *out = ctx;
            });
    }

    int halhip_context_destroy(halhip_context *ctx)
    {
        return Guard([&] { delete ctx; });
    }

    int halhip_context_synchronize(halhip_context *ctx)
    {
        return Guard(
            [&]
            {
                Require(ctx != nullptr, "halhip_context_synchronize: null context");
                Bind(ctx);
                Check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            });
    }

    int halhip_context_fail_next_allocs(halhip_context *ctx, int32_t count)
    {
        return Guard(
            [&]
            {
                Require(ctx != nullptr, "halhip_context_fail_next_allocs: null context");
                Require(count >= 0, "halhip_context_fail_next_allocs: negative count");
                ctx->failNextAllocs = count;
            });
    }

    // Hal::Rng(seed).NextDouble() stream: splitmix64 seeding and xoshiro256**
    // (Blackman & Vigna), top 53 bits scaled by 2^-53; src/Rng.cpp:13-101.
    int halhip_rng_seed_state(uint64_t seed, uint64_t *state)
    {
        return Guard(
            [&]
            {
                Require(state != nullptr, "halhip_rng_seed_state: null state");
                // splitmix64 over the seed fills the four state words
                uint64_t st = seed;
                for (int w = 0; w < 4; w++)
                {
                    st += 0x9E3779B97F4A7C15ULL;
                    uint64_t z = st;
                    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
                    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
                    state[w] = z ^ (z >> 31);
                }
            });
    }

    int halhip_rng_next_doubles(uint64_t *state, int64_t n, double *out)
    {
        return Guard(
            [&]
            {
                Require(state != nullptr, "halhip_rng_next_doubles: null state");
                Require(n >= 0 && (n == 0 || out != nullptr), "halhip_rng_next_doubles: bad output");
                uint64_t s[4] = {state[0], state[1], state[2], state[3]};
                auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
                // xoshiro256** step; the top 53 bits scaled to [0, 1)
                for (int64_t i = 0; i < n; i++)
                {
                    uint64_t r = rotl(s[1] * 5, 7) * 9;
                    uint64_t t = s[1] << 17;
                    s[2] ^= s[0];
                    s[3] ^= s[1];
                    s[1] ^= s[2];
                    s[0] ^= s[3];
                    s[2] ^= t;
                    s[3] = rotl(s[3], 45);
                    out[i] = static_cast<double>(r >> 11) * (1.0 / 9007199254740992.0);
                }
                for (int w = 0; w < 4; w++)
                {
                    state[w] = s[w];
                }
            });
    }

    int halhip_rng_doubles(uint64_t seed, int64_t n, double *out)
    {
        uint64_t state[4];
        const int rc = halhip_rng_seed_state(seed, state);
        return rc != HALHIP_OK ? rc : halhip_rng_next_doubles(state, n, out);
    }
}
