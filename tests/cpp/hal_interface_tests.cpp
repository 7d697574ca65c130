// hal_interface_tests.cpp -- model-side C++ code written against HAL's three
// strategy interfaces, run on the Hal::Hip backends.  It restates the
// reference's interface tests (tests/IDiffusionBackend2DTests.cpp,
// tests/IAgentForceBackend2DTests.cpp) with closed-form and brute-force known
// answers, and a model-owned tick loop (Diffuse, decay, commit as in
// PDEGrid2D::Update), and the same model code on the device-resident
// Hal::Hip::HipPDEGrid2D.  Exit status 0 when every check passes; 77 when no GPU
// is available (skip, like tests/GpuTestSupport.cpp:9-20).
#include "Hal/Hip/HipBackends.h"

#include <algorithm>
#include <cmath>
#include <memory>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
    int g_failures = 0;
    int g_checks = 0;

    void Check(bool ok, const std::string &what)
    {
        g_checks++;
        if (!ok)
        {
            g_failures++;
            std::printf("FAIL: %s\n", what.c_str());
        }
    }

    bool Near(double a, double b, double tol)
    {
        return std::fabs(a - b) <= tol * std::max(1.0, std::fabs(b));
    }

    std::vector<double> Eigen(int nx, int ny)
    {
        std::vector<double> f(static_cast<size_t>(nx * ny));
        for (int x = 0; x < nx; x++)
        {
            for (int y = 0; y < ny; y++)
            {
                f[static_cast<size_t>(x * ny + y)] = std::sin(2.0 * M_PI * x / nx) * std::sin(2.0 * M_PI * y / ny);
            }
        }
        return f;
    }

    // tests/IDiffusionBackend2DTests.cpp:23-55
    void ExplicitEigenmode(Hal::IDiffusionBackend2D &backend)
    {
        const int nx = 16;
        const int ny = 20;
        const double r = 0.05;
        std::vector<double> f = Eigen(nx, ny);
        std::vector<double> d(f.size(), 0.0);
        backend.Diffuse(f, d, r, nx, ny, true, true);
        const double lam = 1.0 - 4.0 * r * (std::pow(std::sin(M_PI / nx), 2) + std::pow(std::sin(M_PI / ny), 2));
        bool ok = true;
        for (size_t i = 0; i < f.size(); i++)
        {
            ok = ok && Near(f[i] + d[i], lam * f[i], 1e-4);
        }
        Check(ok, "explicit eigenmode through IDiffusionBackend2D");
    }

    // tests/IDiffusionBackend2DTests.cpp:76-144
    void AdiEigenmode(Hal::IAdiDiffusionBackend2D &backend)
    {
        const int nx = 12;
        const int ny = 18;
        const double r = 0.3;
        std::vector<double> f = Eigen(nx, ny);
        auto cn = [&](int n)
        {
            const double s = std::pow(std::sin(M_PI / n), 2);
            return (1.0 - 2.0 * r * s) / (1.0 + 2.0 * r * s);
        };
        const double lam = cn(nx) * cn(ny);
        for (int call = 0; call < 2; call++)
        {
            std::vector<double> d(f.size(), 0.0);
            backend.DiffuseADI(f, d, r, nx, ny, true, true);
            bool ok = true;
            for (size_t i = 0; i < f.size(); i++)
            {
                ok = ok && Near(f[i] + d[i], lam * f[i], 1e-4);
            }
            Check(ok, "ADI eigenmode through IAdiDiffusionBackend2D, call " + std::to_string(call));
        }
    }

    template <class F>
    bool Throws(F &&f)
    {
        try
        {
            f();
        }
        catch (const std::invalid_argument &)
        {
            return true;
        }
        return false;
    }

    // tests/IAgentForceBackend2DTests.cpp:25-121 with Hal::Rng(1234) draws.
    void AgentBruteForce(Hal::IAgentForceBackend2D &backend)
    {
        const int dim = 16;
        const int n = 60;
        std::vector<double> u(3 * n);
        halhip_rng_doubles(1234, 3 * n, u.data());
        std::vector<double> pos;
        std::vector<double> rad;
        for (int i = 0; i < n; i++)
        {
            pos.push_back(u[3 * i] * dim);
            pos.push_back(u[3 * i + 1] * dim);
            rad.push_back(0.4 + u[3 * i + 2] * 0.4);
        }
        auto wrapDisp = [&](double a, double b)
        {
            double d = b - a;
            if (std::fabs(d) > dim * 0.5)
            {
                d = (a > b) ? d + dim : d - dim;
            }
            return d;
        };
        std::vector<double> expect(pos.size(), 0.0);
        for (int i = 0; i < n; i++)
        {
            for (int j = 0; j < n; j++)
            {
                if (i == j)
                {
                    continue;
                }
                const double dx = wrapDisp(pos[2 * i], pos[2 * j]);
                const double dy = wrapDisp(pos[2 * i + 1], pos[2 * j + 1]);
                const double dist = std::sqrt(dx * dx + dy * dy);
                if (dist < 3.0 && dist > 0.0)
                {
                    const double force = 0.4 * ((rad[i] + rad[j]) - dist);
                    expect[2 * i] -= (dx / dist) * force;
                    expect[2 * i + 1] -= (dy / dist) * force;
                }
            }
        }
        std::vector<double> got;
        backend.AccumulateForces(pos, rad, got, 3.0, 0.4, dim, dim, true, true);
        bool ok = got.size() == expect.size();
        for (size_t k = 0; ok && k < expect.size(); k++)
        {
            ok = Near(got[k], expect[k], 1e-3);
        }
        Check(ok, "agent forces vs brute force through IAgentForceBackend2D");
        std::vector<double> d;
        Check(Throws([&] { backend.AccumulateForces(pos, rad, d, 1.0, 1.0, 8, 8, true, true); }),
              "agent dims differing from construction throw");
        pos.pop_back();
        Check(Throws([&] { backend.AccumulateForces(pos, rad, d, 1.0, 1.0, dim, dim, true, true); }),
              "agent positions/radii size mismatch throws");
    }

    // A model that owns main() and calls HAL each tick: explicit diffusion,
    // decay and a point source committed like PDEGrid2D::Update, compared with
    // the same loop in FP64 on the host.
    void ModelTickLoop(Hal::IDiffusionBackend2D &backend)
    {
        const int nx = 32;
        const int ny = 24;
        std::vector<double> field(static_cast<size_t>(nx * ny));
        halhip_rng_doubles(42, nx * ny, field.data());
        std::vector<double> host = field;
        std::vector<double> deltas(field.size(), 0.0);
        const int src = 10 * ny + 7;
        for (int tick = 0; tick < 50; tick++)
        {
            backend.Diffuse(field, deltas, 0.2, nx, ny, false, true);
            for (size_t i = 0; i < field.size(); i++)
            {
                deltas[i] += field[i] * -1e-3;
            }
            deltas[src] += 1.0;
            for (size_t i = 0; i < field.size(); i++)
            {
                field[i] += deltas[i];
                deltas[i] = 0.0;
            }
            std::vector<double> hd(host.size(), 0.0);
            for (int x = 0; x < nx; x++)
            {
                for (int y = 0; y < ny; y++)
                {
                    const size_t i = static_cast<size_t>(x * ny + y);
                    const double c = host[i];
                    double s = 0.0;
                    s += (x + 1 < nx) ? host[i + ny] - c : 0.0;
                    s += (x > 0) ? host[i - ny] - c : 0.0;
                    s += host[static_cast<size_t>(x * ny + (y + 1) % ny)] - c;
                    s += host[static_cast<size_t>(x * ny + (y + ny - 1) % ny)] - c;
                    hd[i] = s * 0.2 - c * 1e-3;
                }
            }
            hd[src] += 1.0;
            for (size_t i = 0; i < host.size(); i++)
            {
                host[i] += hd[i];
            }
        }
        double num = 0.0;
        double den = 0.0;
        for (size_t i = 0; i < host.size(); i++)
        {
            num += (field[i] - host[i]) * (field[i] - host[i]);
            den += host[i] * host[i];
        }
        Check(std::sqrt(num / den) < 1e-5, "model-owned tick loop matches FP64 host loop (rel L2 < 1e-5)");
    }

    // PDEGrid2D::Advection through the adapter: upwind differences on a
    // periodic grid, and the flux form with a uniform velocity field equal
    // to the uniform form (src/Pde/PDEequations.cpp:380-413, 535-573).
    void Advection(Hal::Hip::HipDiffusionBackend2D &backend)
    {
        const int nx = 9;
        const int ny = 7;
        std::vector<double> f(static_cast<size_t>(nx * ny));
        for (size_t i = 0; i < f.size(); i++)
        {
            f[i] = std::sin(0.7 * static_cast<double>(i)) + 2.0;
        }
        const double vx = 0.25;
        const double vy = -0.2;
        std::vector<double> d(f.size(), 0.0);
        backend.Advect(f, d, vx, vy, nx, ny, true, true);
        bool ok = true;
        for (int x = 0; x < nx; x++)
        {
            for (int y = 0; y < ny; y++)
            {
                const size_t i = static_cast<size_t>(x * ny + y);
                const double c = f[i];
                const double up = f[static_cast<size_t>(((x + nx - 1) % nx) * ny + y)] - c;
                const double side = f[static_cast<size_t>(x * ny + (y + 1) % ny)] - c;
                const double want = up * vx - side * vy;
                ok = ok && std::fabs(d[i] - want) < 1e-6;
            }
        }
        Check(ok, "uniform upwind advection on a periodic grid");
        std::vector<double> dc(f.size(), 0.0);
        backend.AdvectCells(f, dc, std::vector<double>(f.size(), vx), std::vector<double>(f.size(), vy), nx, ny, true,
                            true);
        ok = true;
        for (size_t i = 0; i < f.size(); i++)
        {
            ok = ok && std::fabs(dc[i] - d[i]) < 1e-6;
        }
        Check(ok, "flux-form advection with a uniform velocity field equals the uniform form");
    }

    // One explicit model tick on the host in FP64, periodic on both axes:
    // Diffusion(r) + Advection(vx > 0, vy < 0) + MulAll(m) + Add(src, a) + Update.
    void HostExplicitTick(std::vector<double> &u, int nx, int ny, double r, double vx, double vy, double m, int src,
                          double a)
    {
        std::vector<double> d(u.size(), 0.0);
        for (int x = 0; x < nx; x++)
        {
            for (int y = 0; y < ny; y++)
            {
                const size_t i = static_cast<size_t>(x * ny + y);
                const double c = u[i];
                const double xp = u[static_cast<size_t>(((x + 1) % nx) * ny + y)];
                const double xm = u[static_cast<size_t>(((x + nx - 1) % nx) * ny + y)];
                const double yp = u[static_cast<size_t>(x * ny + (y + 1) % ny)];
                const double ym = u[static_cast<size_t>(x * ny + (y + ny - 1) % ny)];
                d[i] = r * (xp + xm + yp + ym - 4.0 * c) + (xm - c) * vx - (yp - c) * vy + m * c;
            }
        }
        d[static_cast<size_t>(src)] += a;
        for (size_t i = 0; i < u.size(); i++)
        {
            u[i] += d[i];
        }
    }

    double RelL2(const std::vector<double> &a, const std::vector<double> &b)
    {
        double num = 0.0;
        double den = 0.0;
        for (size_t i = 0; i < a.size(); i++)
        {
            num += (a[i] - b[i]) * (a[i] - b[i]);
            den += b[i] * b[i];
        }
        return std::sqrt(num / den);
    }

    // Model code written against PDEGrid2D's tick calls, run on the
    // device-resident grid: explicit ticks with advection, decay and a
    // source against the FP64 host loop; ADI ticks (and switching between
    // the stencils mid-run) against the array adapters; misuse throws; and
    // a save/load round trip in HAL's PDEGrid2D file layout.
    void ResidentGrid(Hal::Hip::HipContext &ctx, const std::string &scratchDir)
    {
        const int nx = 24;
        const int ny = 20;
        const int src = 5 * ny + 9;
        std::vector<double> u0(static_cast<size_t>(nx * ny));
        halhip_rng_doubles(7, nx * ny, u0.data());
        Hal::Hip::HipPDEGrid2D grid(ctx, nx, ny, true, true);
        grid.SetField(u0);
        std::vector<double> host = u0;
        for (int tick = 0; tick < 40; tick++)
        {
            grid.Diffusion(0.15);
            grid.Advection(0.2, -0.1);
            grid.MulAll(-2e-3);
            grid.Add(src, 0.5);
            grid.Update();
            HostExplicitTick(host, nx, ny, 0.15, 0.2, -0.1, -2e-3, src, 0.5);
        }
        Check(grid.UpdateCt() == 40, "resident grid counts its ticks");
        Check(RelL2(grid.GetField(), host) < 1e-5,
              "resident explicit + advection + decay + source ticks match FP64 host loop (rel L2 < 1e-5)");
        Check(Near(grid.Get(3, 4), grid.GetField()[static_cast<size_t>(3 * ny + 4)], 0.0), "Get(x, y) = x*yDim + y");

        // ADI, explicit and ADI ticks in a row with a Dirichlet boundary.
        Hal::Hip::HipPDEGrid2D walls(ctx, nx, ny, false, false);
        walls.SetField(u0);
        Hal::Hip::HipAdiDiffusionBackend2D adi(ctx);
        Hal::Hip::HipDiffusionBackend2D expl(ctx);
        Hal::Hip::DirichletBoundary bc;
        bc.value = 0.25;
        std::vector<double> ref = u0;
        for (int tick = 0; tick < 6; tick++)
        {
            std::vector<double> d(ref.size(), 0.0);
            if (tick == 2 || tick == 3)
            {
                walls.Diffusion(0.2, 0.25);
                expl.Diffuse(ref, d, 0.2, nx, ny, false, false, bc);
            }
            else
            {
                walls.DiffusionADI(3.0, 0.25);
                adi.DiffuseADI(ref, d, 3.0, nx, ny, false, false, bc);
            }
            walls.MulAll(-0.01);
            walls.Add(src, 1.0);
            walls.Update();
            for (size_t i = 0; i < ref.size(); i++)
            {
                ref[i] += d[i] - 0.01 * ref[i];
            }
            ref[static_cast<size_t>(src)] += 1.0;
        }
        Check(RelL2(walls.GetField(), ref) < 1e-5, "resident ADI / explicit ticks match the array adapters");

        // The same ADI model ticks on a grid whose rows fill whole
        // wavefronts (yDim a multiple of 64): decay and sources then run
        // inside the LDS tick kernel.
        {
            const int wx = 96;
            const int wy = 128;
            std::vector<double> w0(static_cast<size_t>(wx * wy));
            halhip_rng_doubles(11, wx * wy, w0.data());
            Hal::Hip::HipPDEGrid2D wide(ctx, wx, wy, false, false);
            wide.SetField(w0);
            std::vector<double> wref = w0;
            const int wsrc[3] = {0, 40 * wy + 77, wx * wy - 1};
            for (int tick = 0; tick < 25; tick++)
            {
                wide.DiffusionADI(5.0, 0.0);
                wide.MulAll(-2e-3);
                for (int k = 0; k < 3; k++)
                {
                    wide.Add(wsrc[k], 0.5 * (k + 1));
                }
                wide.Update();
                std::vector<double> d(wref.size(), 0.0);
                Hal::Hip::DirichletBoundary zero;
                zero.value = 0.0;
                adi.DiffuseADI(wref, d, 5.0, wx, wy, false, false, zero);
                for (size_t i = 0; i < wref.size(); i++)
                {
                    wref[i] += d[i] - 2e-3 * wref[i];
                }
                for (int k = 0; k < 3; k++)
                {
                    wref[static_cast<size_t>(wsrc[k])] += 0.5 * (k + 1);
                }
            }
            Check(RelL2(wide.GetField(), wref) < 1e-5,
                  "resident ADI model ticks on 64-multiple rows match the array adapter loop");
        }

        // A tick with no calls commits nothing but counts.
        const std::vector<double> before = walls.GetField();
        walls.Update();
        Check(walls.GetField() == before && walls.UpdateCt() == 7, "empty tick keeps the field and counts");

        Check(Throws([&] { walls.Diffusion(0.3); }), "unstable explicit rate throws");
        Check(Throws([&] { walls.Advection(0.4, 0.2); }), "unstable advection velocity throws");
        // Ticks the fused kernels do not cover run call by call on device
        // deltas: two diffusion calls, differing boundary values, advection
        // with ADI.  Each matches the same calls through the array adapters.
        {
            std::vector<double> r0 = walls.GetField();
            std::vector<double> d(r0.size(), 0.0);
            walls.Diffusion(0.1);
            walls.DiffusionADI(1.0);
            walls.Advection(0.1, 0.1, 2.0);
            Check(Throws([&] { walls.Save(scratchDir + "/pending.bin"); }), "saving with pending calls throws");
            walls.Update();
            expl.Diffuse(r0, d, 0.1, nx, ny, false, false);
            adi.DiffuseADI(r0, d, 1.0, nx, ny, false, false);
            Hal::Hip::DirichletBoundary two;
            two.value = 2.0;
            expl.Advect(r0, d, 0.1, 0.1, nx, ny, false, false, two);
            for (size_t i = 0; i < r0.size(); i++)
            {
                r0[i] += d[i];
            }
            Check(RelL2(walls.GetField(), r0) < 1e-5,
                  "a tick of explicit + ADI diffusion + advection with its own boundary matches the adapters");
        }
        Check(Throws([&] { walls.Add(nx * ny, 1.0); }), "out-of-range Add throws");
        Check(Throws([&] { walls.Set(-1, 1.0); }), "out-of-range Set throws");
        Check(Throws([&] { walls.Diffusion(std::vector<double>(3, 0.1), std::vector<double>(3, 0.1)); }),
              "per-cell rates of the wrong length throw");

        const std::string path = scratchDir + "/resident_grid.bin";
        walls.Save(path);
        Hal::Hip::HipPDEGrid2D back(ctx, nx, ny, false, false);
        back.Load(path);
        Check(back.UpdateCt() == walls.UpdateCt() && back.GetField() == walls.GetField(),
              "resident grid save/load round trip is exact");
        Hal::Hip::HipPDEGrid2D wrapped(ctx, nx, ny, true, false);
        Check(Throws([&] { wrapped.Load(path); }), "loading a file with other wrap flags throws");
        Hal::Hip::HipPDEGrid2D other(ctx, nx, ny + 1, false, false);
        Check(Throws([&] { other.Load(path); }), "loading a file with other dimensions throws");
    }
}

namespace
{
    void WriteDoubles(const std::string &path, const std::vector<double> &v)
    {
        FILE *f = std::fopen(path.c_str(), "wb");
        Check(f != nullptr, "open " + path);
        if (f != nullptr)
        {
            Check(std::fwrite(v.data(), sizeof(double), v.size(), f) == v.size(), "write " + path);
            std::fclose(f);
        }
    }

    std::vector<double> Draws(uint64_t seed, int n, double scale, double shift)
    {
        std::vector<double> v(static_cast<size_t>(n));
        halhip_rng_doubles(seed, n, v.data());
        for (double &x : v)
        {
            x = scale * (x + shift);
        }
        return v;
    }

    // Model code over the whole PDEGrid2D tick surface on the device grid.
    // The results go to files that the Python side compares with the FP64
    // oracle running the same calls (tests/test_cpp_interfaces.py).
    void ResidentGridScenarios(Hal::Hip::HipContext &ctx, const std::string &dir)
    {
        // 1: fused ADI model ticks (DiffusionADI with a boundary value,
        // MulAll, Add) on rows that fill whole wavefronts
        {
            const int X = 96;
            const int Y = 128;
            Hal::Hip::HipPDEGrid2D g(ctx, X, Y, false, false);
            g.SetField(Draws(11, X * Y, 1.0, 0.0));
            const int src[3] = {0, 40 * Y + 77, X * Y - 1};
            for (int t = 0; t < 25; t++)
            {
                g.DiffusionADI(5.0, 0.0);
                g.MulAll(-2e-3);
                for (int k = 0; k < 3; k++)
                {
                    g.Add(src[k], 0.5 * (k + 1));
                }
                g.Update();
            }
            WriteDoubles(dir + "/facade_adi.bin", g.GetField());
        }
        // 2: a tick of every kind of call: boundary callables, per-cell
        // velocities with velocity callables, a second (ADI) diffusion,
        // Set / Scale / Mul / Add, AddAll / MulAll, SetNonNegative, MaxDelta
        {
            const int X = 40;
            const int Y = 36;
            Hal::Hip::HipPDEGrid2D g(ctx, X, Y, false, false);
            g.SetField(Draws(21, X * Y, 1.0, -0.2));
            const std::vector<double> xv = Draws(22, X * Y, 0.1, -0.5);
            const std::vector<double> yv = Draws(23, X * Y, 0.1, -0.5);
            auto bc = [](int32_t x, int32_t y) { return 0.5 + 0.01 * x - 0.02 * y; };
            auto xVelBC = [](int32_t, int32_t) { return 0.05; };
            auto yVelBC = [](int32_t, int32_t) { return -0.03; };
            std::vector<double> stats;
            for (int t = 0; t < 8; t++)
            {
                g.Diffusion(0.1, bc);
                g.Advection(xv, yv, bc, xVelBC, yVelBC);
                g.DiffusionADI(2.0);
                g.Set(7, 5, 1.5);
                g.Scale(100, 1.2);
                g.Mul(200, -0.1);
                g.Add(300, 0.25);
                g.AddAll(1e-3);
                g.MulAll(-5e-3);
                if (t % 2 == 1)
                {
                    g.SetNonNegative();
                }
                stats.push_back(g.MaxDelta());
                stats.push_back(g.MaxDeltaScaled(0.1));
                g.Update();
            }
            stats.push_back(g.GetMax());
            stats.push_back(g.GetMin());
            stats.push_back(g.GetAvg());
            WriteDoubles(dir + "/facade_general.bin", g.GetField());
            WriteDoubles(dir + "/facade_general_stats.bin", stats);
        }
        // 3: per-cell rates with uniform advection, on a grid periodic in x;
        // then SetAll from values and a fused tick from the same state
        {
            const int X = 32;
            const int Y = 40;
            Hal::Hip::HipPDEGrid2D g(ctx, X, Y, true, false);
            g.SetField(Draws(31, X * Y, 1.0, 0.0));
            const std::vector<double> rx = Draws(32, X * Y, 0.2, 0.0);
            const std::vector<double> ry = Draws(33, X * Y, 0.2, 0.0);
            for (int t = 0; t < 12; t++)
            {
                g.Diffusion(rx, ry);
                g.Advection(0.1, 0.05);
                g.MulAll(-1e-3);
                g.Add(5, 0.1);
                g.Update();
            }
            const std::vector<double> vals = Draws(34, X * Y, 1.0, 0.0);
            g.SetAll(vals.data());
            g.Diffusion(0.2);
            g.Update();
            g.Diffusion(0.2);
            g.MulAll(-1e-3);
            g.Update();
            WriteDoubles(dir + "/facade_rates.bin", g.GetField());
        }
    }
}

int main(int argc, char **argv)
{
    // argv[1]: a writable folder for the file round trip (default /tmp)
    const std::string scratchDir = argc > 1 ? argv[1] : "/tmp";
    std::unique_ptr<Hal::Hip::HipContext> ctx = Hal::Hip::HipContext::TryCreate(0);
    if (!ctx)
    {
        std::printf("SKIP: no GPU device\n");
        return 77;
    }
    {
        Hal::Hip::HipDiffusionBackend2D fixed(*ctx, 16, 20);
        ExplicitEigenmode(fixed);
        Hal::Hip::HipDiffusionBackend2D any(*ctx);
        ExplicitEigenmode(any);
        ModelTickLoop(any);
        Advection(any);
        Check(Throws([&] { fixed.SetAdvection(0.3, -0.3); }), "unstable advection velocity throws");
        std::vector<double> field(12, 0.0);
        std::vector<double> deltas(11, 0.0);
        Check(Throws([&] { any.Diffuse(field, deltas, 0.1, 3, 4, false, false); }), "explicit size mismatch throws");
        std::vector<double> f12(12, 0.0);
        std::vector<double> d12(12, 0.0);
        Check(Throws([&] { fixed.Diffuse(f12, d12, 0.1, 3, 4, false, false); }),
              "explicit dims differing from construction throw");
    }
    {
        Hal::Hip::HipAdiDiffusionBackend2D fixed(*ctx, 12, 18);
        AdiEigenmode(fixed);
        Hal::Hip::HipAdiDiffusionBackend2D any(*ctx);
        AdiEigenmode(any);
        std::vector<double> field(12, 0.0);
        std::vector<double> deltas(11, 0.0);
        Check(Throws([&] { any.DiffuseADI(field, deltas, 0.1, 3, 4, false, false); }), "ADI size mismatch throws");
        deltas.resize(12);
        bool threw = false;
        try
        {
            any.DiffuseADI(field, deltas, 0.1, 3, 4, false, false);
        }
        catch (...)
        {
            threw = true;
        }
        Check(!threw, "ADI on matching sizes does not throw");
        // Dirichlet fixed point (tests/PDEADITests.cpp:219-235).
        std::vector<double> flat(35, 4.0);
        std::vector<double> dd(35, 0.0);
        Hal::Hip::DirichletBoundary bc;
        bc.value = 4.0;
        any.DiffuseADI(flat, dd, 0.9, 7, 5, false, false, bc);
        bool ok = true;
        for (double v : dd)
        {
            ok = ok && std::fabs(v) < 1e-5;
        }
        Check(ok, "ADI Dirichlet fixed point");
    }
    {
        Hal::Hip::HipAgentForceBackend2D agents(*ctx, 16, 16, 60);
        AgentBruteForce(agents);
        std::vector<double> pos = {4.0, 4.0, 4.0, 4.0};
        std::vector<double> rad = {0.5, 0.5};
        std::vector<double> out;
        Hal::Hip::HipAgentForceBackend2D small(*ctx, 8, 8, 4);
        small.AccumulateForces(pos, rad, out, 2.0, 1.0, 8, 8, false, false);
        Check(out.size() == 4 && out[0] == 0.0 && out[1] == 0.0 && out[2] == 0.0 && out[3] == 0.0,
              "coincident agents give zero force");
    }
    ResidentGrid(*ctx, scratchDir);
    ResidentGridScenarios(*ctx, scratchDir);
    std::printf("%d checks, %d failures\n", g_checks, g_failures);
    return g_failures == 0 ? 0 : 1;
}
