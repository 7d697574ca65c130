// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" shim over the *compiled HAL-CPP reference* CPU backends
// (built from the reference's own sources by oracle/Makefile into
// oracle/_ref/libhalref.so).  It is used to write golden vectors and to
// cross-check the FP64 restatement in hal_oracle.cpp; nothing in the product
// links it.  Each entry point calls the reference API it names.
#include "Hal/Grids/PDEGridSerialization.h"
#include "Hal/Agents/CpuAgentForceBackend2D.h"
#include "Hal/IO/BinaryReader.h"
#include "Hal/IO/BinaryWriter.h"
#include "Hal/Grids/PDEGrid2D.h"
#include "Hal/Pde/ADIequations.h"
#include "Hal/Pde/CpuAdiDiffusionBackend2D.h"
#include "Hal/Pde/CpuDiffusionBackend2D.h"
#include "Hal/Pde/PDEequations.h"
#include "Hal/Rng.h"

#include <chrono>
#include <cstdint>
#include <vector>

namespace
{
    struct EdgeBc
    {
        int32_t X;
        int32_t Y;
        const double *xlo;
        const double *xhi;
        const double *ylo;
        const double *yhi;
        double operator()(int32_t x, int32_t y) const
        {
            if (x < 0)
            {
                return xlo[y];
            }
            if (x >= X)
            {
                return xhi[y];
            }
            if (y < 0)
            {
                return ylo[x];
            }
            return yhi[x];
        }
    };
}

extern "C"
{
    // Hal::Rng(seed).NextDouble() stream.
    void ref_rng_doubles(uint64_t seed, int64_t n, double *out)
    {
        Hal::Rng rng(seed);
        for (int64_t i = 0; i < n; i++)
        {
            out[i] = rng.NextDouble();
        }
    }

    void ref_rng_u64(uint64_t seed, int64_t n, uint64_t *out)
    {
        Hal::Rng rng(seed);
        for (int64_t i = 0; i < n; i++)
        {
            out[i] = rng.NextUint64();
        }
    }

    // Hal::Rng(seed) after `skip` NextDouble draws, saved with Rng::Save
    // (src/Rng.cpp:122-125) to `path`.
    int32_t ref_rng_save(uint64_t seed, int64_t skip, const char *path)
    {
        Hal::Rng rng(seed);
        for (int64_t i = 0; i < skip; i++)
        {
            (void)rng.NextDouble();
        }
        Hal::BinaryWriter writer(path);
        rng.Save(writer);
        return writer.IsGood() ? 0 : 1;
    }

    // A Hal::Rng restored with Rng::Load (src/Rng.cpp:127-130) from `path`,
    // then n NextDouble draws.
    void ref_rng_load_doubles(const char *path, int64_t n, double *out)
    {
        Hal::Rng rng(0);
        Hal::BinaryReader reader(path);
        rng.Load(reader);
        for (int64_t i = 0; i < n; i++)
        {
            out[i] = rng.NextDouble();
        }
    }

    // Hal::CpuDiffusionBackend2D::Diffuse, or Internal::Diffusion2<BCFn> when
    // edge arrays are given.
    void ref_diffuse(const double *field, double *deltas, double rate, int32_t X, int32_t Y, int32_t wrapX,
                     int32_t wrapY, const double *bcXlo, const double *bcXhi, const double *bcYlo,
                     const double *bcYhi)
    {
        size_t n = static_cast<size_t>(X) * Y;
        if (bcXlo == nullptr)
        {
            std::vector<double> f(field, field + n);
            std::vector<double> d(deltas, deltas + n);
            Hal::CpuDiffusionBackend2D backend;
            backend.Diffuse(f, d, rate, X, Y, wrapX != 0, wrapY != 0);
            std::copy(d.begin(), d.end(), deltas);
            return;
        }
        EdgeBc bc{X, Y, bcXlo, bcXhi, bcYlo, bcYhi};
        Hal::Internal::Diffusion2(field, deltas, rate, X, Y, wrapX != 0, wrapY != 0, bc);
    }

    void ref_diffuse_rates(const double *field, double *deltas, const double *ratesX, const double *ratesY,
                           int32_t X, int32_t Y, int32_t wrapX, int32_t wrapY)
    {
        Hal::Internal::Diffusion2(field, deltas, ratesX, ratesY, X, Y, wrapX != 0, wrapY != 0);
    }

    // Hal::CpuAdiDiffusionBackend2D::DiffuseADI, or Internal::Diffusion2DADI<BCFn>.
    void ref_adi(const double *field, double *deltas, double rate, int32_t X, int32_t Y, int32_t wrapX,
                 int32_t wrapY, const double *bcXlo, const double *bcXhi, const double *bcYlo, const double *bcYhi)
    {
        size_t n = static_cast<size_t>(X) * Y;
        if (bcXlo == nullptr)
        {
            std::vector<double> f(field, field + n);
            std::vector<double> d(deltas, deltas + n);
            Hal::CpuAdiDiffusionBackend2D backend;
            backend.DiffuseADI(f, d, rate, X, Y, wrapX != 0, wrapY != 0);
            std::copy(d.begin(), d.end(), deltas);
            return;
        }
        EdgeBc bc{X, Y, bcXlo, bcXhi, bcYlo, bcYhi};
        std::vector<double> scratch(n, 0.0);
        Hal::TdmaSolver tdma(X > Y ? X : Y);
        Hal::Internal::Diffusion2DADI(field, scratch.data(), deltas, rate, X, Y, wrapX != 0, wrapY != 0, bc, tdma);
    }

    // Full model ticks through Hal::PDEGrid2D: Diffusion or DiffusionADI
    // (adi != 0; constant Dirichlet value when hasBc), MulAll(-decay),
    // Add(src), Update().
    void ref_grid_ticks(double *field, int32_t ticks, int32_t adi, double rate, double decay, int32_t X, int32_t Y,
                        int32_t wrapX, int32_t wrapY, int32_t hasBc, double bcValue, int32_t nSrc,
                        const int64_t *srcIdx, const double *srcVal)
    {
        Hal::PDEGrid2D grid(X, Y, wrapX != 0, wrapY != 0);
        int32_t n = X * Y;
        for (int32_t i = 0; i < n; i++)
        {
            grid.Set(i, field[i]);
        }
        grid.Update();
        for (int32_t t = 0; t < ticks; t++)
        {
            if (adi != 0)
            {
                if (hasBc != 0)
                {
                    grid.DiffusionADI(rate, bcValue);
                }
                else
                {
                    grid.DiffusionADI(rate);
                }
            }
            else
            {
                grid.Diffusion(rate);
            }
            if (decay != 0.0)
            {
                grid.MulAll(-decay);
            }
            for (int32_t s = 0; s < nSrc; s++)
            {
                grid.Add(static_cast<int32_t>(srcIdx[s]), srcVal[s]);
            }
            grid.Update();
        }
        for (int32_t i = 0; i < n; i++)
        {
            field[i] = grid.Get(i);
        }
    }

    // CPU baseline timing: `warmup` untimed then `ticks` timed ticks of
    // PDEGrid2D::DiffusionADI(rate, bcValue) + Update on one thread, clocked
    // with std::chrono::steady_clock as examples/BackendBenchmark.cpp:50-54
    // does.  Returns the seconds of the timed ticks; field is updated.
    double ref_adi_ticks_timed(double *field, int32_t warmup, int32_t ticks, double rate, int32_t X, int32_t Y,
                               double bcValue)
    {
        Hal::PDEGrid2D grid(X, Y, false, false);
        int32_t n = X * Y;
        for (int32_t i = 0; i < n; i++)
        {
            grid.Set(i, field[i]);
        }
        grid.Update();
        for (int32_t t = 0; t < warmup; t++)
        {
            grid.DiffusionADI(rate, bcValue);
            grid.Update();
        }
        auto t0 = std::chrono::steady_clock::now();
        for (int32_t t = 0; t < ticks; t++)
        {
            grid.DiffusionADI(rate, bcValue);
            grid.Update();
        }
        auto t1 = std::chrono::steady_clock::now();
        for (int32_t i = 0; i < n; i++)
        {
            field[i] = grid.Get(i);
        }
        return std::chrono::duration<double>(t1 - t0).count();
    }

    // CPU baseline of the per-call interfaces: one untimed then `calls` timed
    // CpuDiffusionBackend2D::Diffuse (adi == 0) or
    // CpuAdiDiffusionBackend2D::DiffuseADI calls on the same caller-owned
    // vectors, as examples/BackendBenchmark.cpp:60-140 times a backend.
    // Returns the seconds of the timed calls.
    double ref_diffuse_calls_timed(const double *field, double rate, int32_t X, int32_t Y, int32_t wrapX,
                                   int32_t wrapY, int32_t adi, int32_t calls)
    {
        size_t n = static_cast<size_t>(X) * Y;
        std::vector<double> f(field, field + n);
        std::vector<double> d(n, 0.0);
        Hal::CpuDiffusionBackend2D explicitBackend;
        Hal::CpuAdiDiffusionBackend2D adiBackend;
        auto call = [&]()
        {
            if (adi != 0)
            {
                adiBackend.DiffuseADI(f, d, rate, X, Y, wrapX != 0, wrapY != 0);
            }
            else
            {
                explicitBackend.Diffuse(f, d, rate, X, Y, wrapX != 0, wrapY != 0);
            }
        };
        call();
        auto t0 = std::chrono::steady_clock::now();
        for (int32_t c = 0; c < calls; c++)
        {
            call();
        }
        auto t1 = std::chrono::steady_clock::now();
        return std::chrono::duration<double>(t1 - t0).count();
    }

    // Hal::Internal::Advection2 with a uniform velocity (zero flux / periodic,
    // or Dirichlet edge values when xlo is non-null).
    void ref_advect(const double *field, double *deltas, double xVel, double yVel, int32_t X, int32_t Y,
                    int32_t wrapX, int32_t wrapY, const double *bcXlo, const double *bcXhi, const double *bcYlo,
                    const double *bcYhi)
    {
        if (bcXlo == nullptr)
        {
            Hal::Internal::Advection2(field, deltas, xVel, yVel, X, Y, wrapX != 0, wrapY != 0);
            return;
        }
        EdgeBc bc{X, Y, bcXlo, bcXhi, bcYlo, bcYhi};
        Hal::Internal::Advection2(field, deltas, xVel, yVel, X, Y, wrapX != 0, wrapY != 0, bc);
    }

    // Hal::Internal::Advection2 with per-cell velocities (zero flux /
    // periodic, or field and velocity edge values when bcXlo is non-null).
    void ref_advect_cells(const double *field, double *deltas, const double *xVels, const double *yVels, int32_t X,
                          int32_t Y, int32_t wrapX, int32_t wrapY, const double *bcXlo, const double *bcXhi,
                          const double *bcYlo, const double *bcYhi, const double *vXlo, const double *vXhi,
                          const double *vYlo, const double *vYhi)
    {
        if (bcXlo == nullptr)
        {
            Hal::Internal::Advection2(field, deltas, xVels, yVels, X, Y, wrapX != 0, wrapY != 0);
            return;
        }
        EdgeBc bc{X, Y, bcXlo, bcXhi, bcYlo, bcYhi};
        EdgeBc vbc{X, Y, vXlo, vXhi, vYlo, vYhi};
        Hal::Internal::Advection2(field, deltas, xVels, yVels, X, Y, wrapX != 0, wrapY != 0, bc, vbc, vbc);
    }

    // Model ticks through Hal::PDEGrid2D: Diffusion + Advection (with a
    // constant boundary value when hasBc) + MulAll(-decay) + Add + Update.
    void ref_model_ticks(double *field, int32_t ticks, double rate, double xVel, double yVel, double decay, int32_t X,
                         int32_t Y, int32_t wrapX, int32_t wrapY, int32_t hasBc, double bcValue, int32_t nSrc,
                         const int64_t *srcIdx, const double *srcVal)
    {
        Hal::PDEGrid2D grid(X, Y, wrapX != 0, wrapY != 0);
        int32_t n = X * Y;
        for (int32_t i = 0; i < n; i++)
        {
            grid.Set(i, field[i]);
        }
        grid.Update();
        for (int32_t t = 0; t < ticks; t++)
        {
            if (hasBc != 0)
            {
                grid.Diffusion(rate, bcValue);
                grid.Advection(xVel, yVel, bcValue);
            }
            else
            {
                grid.Diffusion(rate);
                grid.Advection(xVel, yVel);
            }
            if (decay != 0.0)
            {
                grid.MulAll(-decay);
            }
            for (int32_t s = 0; s < nSrc; s++)
            {
                grid.Add(static_cast<int32_t>(srcIdx[s]), srcVal[s]);
            }
            grid.Update();
        }
        for (int32_t i = 0; i < n; i++)
        {
            field[i] = grid.Get(i);
        }
    }

    // CPU baseline timing of whole model ticks through Hal::PDEGrid2D
    // (as ref_grid_ticks): `warmup` untimed then `ticks` timed ticks, one
    // thread, std::chrono::steady_clock.  Returns the timed seconds.
    double ref_grid_ticks_timed(double *field, int32_t warmup, int32_t ticks, int32_t adi, double rate, double decay,
                                int32_t X, int32_t Y, int32_t wrapX, int32_t wrapY, int32_t hasBc, double bcValue,
                                int32_t nSrc, const int64_t *srcIdx, const double *srcVal)
    {
        Hal::PDEGrid2D grid(X, Y, wrapX != 0, wrapY != 0);
        int32_t n = X * Y;
        for (int32_t i = 0; i < n; i++)
        {
            grid.Set(i, field[i]);
        }
        grid.Update();
        auto tick = [&]()
        {
            if (adi != 0)
            {
                if (hasBc != 0)
                {
                    grid.DiffusionADI(rate, bcValue);
                }
                else
                {
                    grid.DiffusionADI(rate);
                }
            }
            else
            {
                grid.Diffusion(rate);
            }
            if (decay != 0.0)
            {
                grid.MulAll(-decay);
            }
            for (int32_t s = 0; s < nSrc; s++)
            {
                grid.Add(static_cast<int32_t>(srcIdx[s]), srcVal[s]);
            }
            grid.Update();
        };
        for (int32_t t = 0; t < warmup; t++)
        {
            tick();
        }
        auto t0 = std::chrono::steady_clock::now();
        for (int32_t t = 0; t < ticks; t++)
        {
            tick();
        }
        auto t1 = std::chrono::steady_clock::now();
        for (int32_t i = 0; i < n; i++)
        {
            field[i] = grid.Get(i);
        }
        return std::chrono::duration<double>(t1 - t0).count();
    }

    // CPU baseline timing of `calls` Hal::CpuAgentForceBackend2D::AccumulateForces
    // calls (after one untimed call, as examples/BackendBenchmark.cpp:75-77).
    double ref_agent_forces_timed(const double *pos, const double *radii, int32_t N, int32_t calls, double R,
                                  double k, int32_t X, int32_t Y, int32_t wrapX, int32_t wrapY)
    {
        std::vector<double> p(pos, pos + 2 * static_cast<size_t>(N));
        std::vector<double> r(radii, radii + N);
        std::vector<double> v;
        Hal::CpuAgentForceBackend2D backend;
        backend.AccumulateForces(p, r, v, R, k, X, Y, wrapX != 0, wrapY != 0);
        auto t0 = std::chrono::steady_clock::now();
        for (int32_t c = 0; c < calls; c++)
        {
            backend.AccumulateForces(p, r, v, R, k, X, Y, wrapX != 0, wrapY != 0);
        }
        auto t1 = std::chrono::steady_clock::now();
        return std::chrono::duration<double>(t1 - t0).count();
    }

    // Hal::SavePDEGrid2D of a grid holding `field` and `deltas`
    // (src/Grids/PDEGridSerialization.cpp:33-43) to `path`.
    int32_t ref_save_pdegrid2d(const char *path, const double *field, const double *deltas, int32_t X, int32_t Y,
                               int32_t wrapX, int32_t wrapY, int32_t updateCt)
    {
        Hal::PDEGrid2D grid(X, Y, wrapX != 0, wrapY != 0);
        const size_t n = static_cast<size_t>(X) * Y;
        grid.SetState(std::vector<double>(field, field + n), std::vector<double>(deltas, deltas + n), updateCt);
        Hal::BinaryWriter writer(path);
        Hal::SavePDEGrid2D(writer, grid);
        return writer.IsGood() ? 0 : 1;
    }

    // Hal::LoadPDEGrid2D (src/Grids/PDEGridSerialization.cpp:45-61): fills
    // dims, flags, update count, field and deltas (n = capacity); returns 0,
    // or 1 when the reference throws.
    int32_t ref_load_pdegrid2d(const char *path, int32_t *dims, double *field, double *deltas, int64_t n)
    {
        try
        {
            Hal::BinaryReader reader(path);
            Hal::PDEGrid2D grid = Hal::LoadPDEGrid2D(reader);
            dims[0] = grid.Xdim();
            dims[1] = grid.Ydim();
            dims[2] = grid.IsWrapX() ? 1 : 0;
            dims[3] = grid.IsWrapY() ? 1 : 0;
            dims[4] = grid.UpdateCt();
            const size_t len = static_cast<size_t>(grid.Xdim()) * grid.Ydim();
            if (static_cast<int64_t>(len) > n)
            {
                return 1;
            }
            std::copy(grid.GetField().begin(), grid.GetField().end(), field);
            std::copy(grid.GetDeltas().begin(), grid.GetDeltas().end(), deltas);
            return 0;
        }
        catch (const std::exception &)
        {
            return 1;
        }
    }

    // Hal::CpuAgentForceBackend2D::AccumulateForces.
    void ref_agent_forces(const double *pos, const double *radii, int32_t N, double *vel, double R, double k,
                          int32_t X, int32_t Y, int32_t wrapX, int32_t wrapY)
    {
        std::vector<double> p(pos, pos + 2 * static_cast<size_t>(N));
        std::vector<double> r(radii, radii + N);
        std::vector<double> v;
        Hal::CpuAgentForceBackend2D backend;
        backend.AccumulateForces(p, r, v, R, k, X, Y, wrapX != 0, wrapY != 0);
        std::copy(v.begin(), v.end(), vel);
    }
}
