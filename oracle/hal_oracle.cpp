// hal_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of HAL-CPP's CPU hot path, used as the checker for the
// HIP backend. Only tests, the smoke entry point and the CPU-baseline leg of
// the benchmark may load this library; the product path never links it.
//
// Every function names the reference lines whose behaviour it restates
// (paths relative to the HAL-CPP source tree). Floating-point operation order
// follows the reference so that this restatement and the compiled reference
// agree to the last bit on the golden vectors.
//
// Pinning (what checks each function; tests/test_oracle.py):
//   oracle_rng_*            golden stream from the compiled reference
//   oracle_diffuse*         golden vectors; periodic eigenmode closed form;
//                           mass invariance; per-cell == uniform rates
//   oracle_adi*             golden vectors (zero-flux, periodic, Dirichlet);
//                           Crank-Nicolson eigenmode product; Dirichlet fixed
//                           point and linear ramp; scipy banded solve
//   oracle_agent_*          golden vectors; brute-force all-pairs sum; sorted
//                           (cell, index) order for bins; carried-velocity
//                           ticks: the reference's friction/cap unit-test
//                           values, geometric-series closed form, force +
//                           carry sum, and "off == plain ticks"
//   oracle_advect*          golden vectors (uniform and per-cell velocity,
//                           zero flux, periodic, Dirichlet); mass invariance
//                           of the flux form; uniform velocity fields give
//                           the uniform-velocity form in the interior
//   oracle_model_ticks      golden tick compositions
//   oracle_adi_cell_*       per-cell-diffusivity ADI and its adjoint: no
//                           reference counterpart.  Pinned, for every boundary
//                           kind, by "uniform D == oracle_adi", the explicit
//                           half-step against the per-cell-rate stencil
//                           oracle_diffuse_rates, the dot-product test
//                           <Ax,y>=<x,A^T y> and FP64 central finite
//                           differences; at full size by the committed
//                           central-difference fixture.
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace
{
    // True modulo; src/MathUtil.cpp:41-46.
    inline int32_t WrapI(int32_t v, int32_t n)
    {
        int32_t m = v % n;
        return (m < 0) ? (n + m) : m;
    }

    // Neighbour difference along one axis of a flat x*Y+y field.  Off-grid
    // neighbours give 0 (zero flux), the wrapped cell (periodic) or the
    // Dirichlet ghost (bc - c)*2 when pBc is given.
    // src/Pde/PDEequations.cpp:24-60; include/Hal/Pde/PDEequations.h:50-98.
    inline double DeltaX(const double *f, double c, int32_t x, int32_t y, int32_t X, int32_t Y, bool wrap,
                         const double *bcLo, const double *bcHi)
    {
        if (x >= 0 && x < X)
        {
            return f[static_cast<size_t>(x) * Y + y] - c;
        }
        if (wrap)
        {
            return f[static_cast<size_t>(WrapI(x, X)) * Y + y] - c;
        }
        if (bcLo != nullptr)
        {
            double bc = (x < 0) ? bcLo[y] : bcHi[y];
            return (bc - c) * 2;
        }
        return 0;
    }

    inline double DeltaY(const double *f, double c, int32_t x, int32_t y, int32_t X, int32_t Y, bool wrap,
                         const double *bcLo, const double *bcHi)
    {
        (void)X;
        if (y >= 0 && y < Y)
        {
            return f[static_cast<size_t>(x) * Y + y] - c;
        }
        if (wrap)
        {
            return f[static_cast<size_t>(x) * Y + WrapI(y, Y)] - c;
        }
        if (bcLo != nullptr)
        {
            double bc = (y < 0) ? bcLo[x] : bcHi[x];
            return (bc - c) * 2;
        }
        return 0;
    }

    // Value of a per-cell array (field or velocity) at (x, y) along x: the
    // cell, the wrapped cell, or past a non-wrapped edge the boundary value
    // (lo at x = -1, hi at x = X), 0 without one.
    // src/Pde/PDEequations.cpp:163-198; include/Hal/Pde/PDEequations.h:579-615.
    inline double ShiftX(const double *v, int32_t x, int32_t y, int32_t X, int32_t Y, bool wrap, const double *lo,
                         const double *hi)
    {
        if (x >= 0 && x < X)
        {
            return v[static_cast<size_t>(x) * Y + y];
        }
        if (wrap)
        {
            return v[static_cast<size_t>(WrapI(x, X)) * Y + y];
        }
        if (lo != nullptr)
        {
            return (x < 0) ? lo[y] : hi[y];
        }
        return 0.0;
    }

    inline double ShiftY(const double *v, int32_t x, int32_t y, int32_t X, int32_t Y, bool wrap, const double *lo,
                         const double *hi)
    {
        (void)X;
        if (y >= 0 && y < Y)
        {
            return v[static_cast<size_t>(x) * Y + y];
        }
        if (wrap)
        {
            return v[static_cast<size_t>(x) * Y + WrapI(y, Y)];
        }
        if (lo != nullptr)
        {
            return (y < 0) ? lo[x] : hi[x];
        }
        return 0.0;
    }

    // Thomas algorithm on explicit coefficient arrays: a sub-diagonal (a[0]
    // unused), b diagonal, c super-diagonal (c[n-1] unused).
    // include/Hal/Pde/TdmaSolver.h:27-66.
    void Thomas(int32_t n, const double *a, const double *b, const double *c, const double *d, double *x,
                std::vector<double> &work)
    {
        work.resize(static_cast<size_t>(n) * 2);
        double *bp = work.data();
        double *dp = work.data() + n;
        bp[0] = b[0];
        dp[0] = d[0];
        for (int32_t i = 1; i < n; i++)
        {
            double w = a[i] / bp[i - 1];
            bp[i] = b[i] - w * c[i - 1];
            dp[i] = d[i] - w * dp[i - 1];
        }
        int32_t i = n - 1;
        double v = dp[i] / bp[i];
        x[i] = v;
        for (i -= 1; i >= 0; i--)
        {
            v = (dp[i] - c[i] * v) / bp[i];
            x[i] = v;
        }
    }

    // Cyclic tridiagonal solve by Sherman-Morrison: two Thomas solves on the
    // corner-adjusted matrix, then a rank-1 combination.
    // include/Hal/Pde/TdmaSolver.h:68-120.
    void ThomasPeriodic(int32_t n, const double *a, const double *b, const double *c, const double *d, double *x,
                        std::vector<double> &work)
    {
        const double gamma = -b[0];
        const double alpha = a[0];
        const double beta = c[n - 1];
        std::vector<double> bAdj(b, b + n);
        bAdj[0] = b[0] - gamma;
        if (n - 1 != 0)
        {
            bAdj[n - 1] = b[n - 1] - (alpha * beta) / gamma;
        }
        std::vector<double> u(static_cast<size_t>(n), 0.0);
        u[0] = gamma;
        if (n - 1 != 0)
        {
            u[n - 1] = alpha;
        }
        std::vector<double> ys(static_cast<size_t>(n));
        std::vector<double> zs(static_cast<size_t>(n));
        Thomas(n, a, bAdj.data(), c, d, ys.data(), work);
        Thomas(n, a, bAdj.data(), c, u.data(), zs.data(), work);
        double vEnd = beta / gamma;
        double vy = ys[0] + ys[n - 1] * vEnd;
        double vz = zs[0] + zs[n - 1] * vEnd;
        double factor = vy / (1 + vz);
        for (int32_t i = 0; i < n; i++)
        {
            x[i] = ys[i] - factor * zs[i];
        }
    }

    // Worker threads for the line loops of the ADI sweeps and the per-agent
    // force sums (1 by default; oracle_set_threads).  Lines and agents are
    // independent, so the result does not depend on the count.
    int g_threads = 1;

    // Runs body(begin, end) over [0, n) split into g_threads contiguous
    // ranges, one thread each.
    template <class F>
    void ParallelRanges(int32_t n, F body)
    {
        const int32_t T = std::max(1, std::min<int32_t>(g_threads, n));
        if (T == 1)
        {
            body(0, n);
            return;
        }
        std::vector<std::thread> pool;
        for (int32_t t = 0; t < T; t++)
        {
            const int32_t b = static_cast<int32_t>(static_cast<int64_t>(n) * t / T);
            const int32_t e = static_cast<int32_t>(static_cast<int64_t>(n) * (t + 1) / T);
            pool.emplace_back([=] { body(b, e); });
        }
        for (auto &th : pool)
        {
            th.join();
        }
    }

    // Coefficients of one ADI line.  off = -rate/2; periodic lines use
    // b = 1 - 2*off; otherwise each in-line neighbour weighs 1 and past an
    // end a Dirichlet ghost weighs 2, a zero-flux edge 0.
    // include/Hal/Pde/ADIequations.h:225-270.
    struct LineCoefs
    {
        int32_t n = 0;
        bool wrap = false;
        std::vector<double> a, b, c;
        LineCoefs(int32_t len, double rate, bool wrapLine, bool hasBc) : n(len), wrap(wrapLine)
        {
            double off = -rate / 2;
            a.assign(static_cast<size_t>(n), off);
            c.assign(static_cast<size_t>(n), off);
            b.resize(static_cast<size_t>(n));
            if (wrap)
            {
                for (int32_t i = 0; i < n; i++)
                {
                    b[i] = 1 - 2 * off;
                }
                return;
            }
            double endAdjust = hasBc ? 1.0 : -1.0;
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
        // TDMA / TDMAperiodic (include/Hal/Pde/TdmaSolver.h:27-120).
        void Solve(const double *d, double *x, std::vector<double> &work) const
        {
            if (wrap)
            {
                ThomasPeriodic(n, a.data(), b.data(), c.data(), d, x, work);
            }
            else
            {
                Thomas(n, a.data(), b.data(), c.data(), d, x, work);
            }
        }
    };

    void SolveLine(int32_t n, double rate, bool wrap, bool hasBc, const double *d, double *x,
                   std::vector<double> &work)
    {
        LineCoefs(n, rate, wrap, hasBc).Solve(d, x, work);
    }

    // Peaceman-Rachford ADI: x-implicit sweep (explicit in y) into scratch,
    // then y-implicit sweep (explicit in x) into deltas += v - field.
    // include/Hal/Pde/ADIequations.h:306-378; src/Pde/ADIequations.cpp:19-47;
    // ImplicitBoundarySource include/Hal/Pde/ADIequations.h:198-220.
    // The x-sweep handles lines in groups of kGroup adjacent columns, so its
    // reads and writes run along rows; every line's arithmetic is unchanged.
    void Adi2D(const double *field, double *deltas, double rate, int32_t X, int32_t Y, bool wrapX, bool wrapY,
               const double *bcXlo, const double *bcXhi, const double *bcYlo, const double *bcYhi)
    {
        constexpr int32_t kGroup = 16;
        bool hasBc = (bcXlo != nullptr);
        std::vector<double> scratch(static_cast<size_t>(X) * Y);
        const LineCoefs lineX(X, rate, wrapX, hasBc);
        const LineCoefs lineY(Y, rate, wrapY, hasBc);
        const int32_t groups = (Y + kGroup - 1) / kGroup;
        ParallelRanges(groups, [&](int32_t g0, int32_t g1)
        {
            std::vector<double> work;
            std::vector<double> d(static_cast<size_t>(X) * kGroup);
            std::vector<double> v(static_cast<size_t>(X));
            std::vector<double> out(static_cast<size_t>(X) * kGroup);
            for (int32_t g = g0; g < g1; g++)
            {
                const int32_t y0 = g * kGroup;
                const int32_t ny = std::min(kGroup, Y - y0);
                for (int32_t i = 0; i < X; i++)
                {
                    for (int32_t j = 0; j < ny; j++)
                    {
                        const int32_t y = y0 + j;
                        double c = field[static_cast<size_t>(i) * Y + y];
                        double s = DeltaY(field, c, i, y + 1, X, Y, wrapY, bcYlo, bcYhi);
                        s += DeltaY(field, c, i, y - 1, X, Y, wrapY, bcYlo, bcYhi);
                        double rhs = c + (s * rate) / 2;
                        if (hasBc && !wrapX)
                        {
                            double src = 0.0;
                            if (i == 0)
                            {
                                src += bcXlo[y];
                            }
                            if (i == X - 1)
                            {
                                src += bcXhi[y];
                            }
                            rhs += src * 2.0 * rate / 2;
                        }
                        d[static_cast<size_t>(j) * X + i] = rhs;
                    }
                }
                for (int32_t j = 0; j < ny; j++)
                {
                    lineX.Solve(d.data() + static_cast<size_t>(j) * X, v.data(), work);
                    for (int32_t i = 0; i < X; i++)
                    {
                        out[static_cast<size_t>(i) * kGroup + j] = v[i];
                    }
                }
                for (int32_t i = 0; i < X; i++)
                {
                    for (int32_t j = 0; j < ny; j++)
                    {
                        scratch[static_cast<size_t>(i) * Y + y0 + j] = out[static_cast<size_t>(i) * kGroup + j];
                    }
                }
            }
        });
        ParallelRanges(X, [&](int32_t x0, int32_t x1)
        {
            std::vector<double> work;
            std::vector<double> d(static_cast<size_t>(Y));
            std::vector<double> v(static_cast<size_t>(Y));
            for (int32_t x = x0; x < x1; x++)
            {
                for (int32_t i = 0; i < Y; i++)
                {
                    double c = scratch[static_cast<size_t>(x) * Y + i];
                    double s = DeltaX(scratch.data(), c, x + 1, i, X, Y, wrapX, bcXlo, bcXhi);
                    s += DeltaX(scratch.data(), c, x - 1, i, X, Y, wrapX, bcXlo, bcXhi);
                    double rhs = c + (s * rate) / 2;
                    if (hasBc && !wrapY)
                    {
                        double src = 0.0;
                        if (i == 0)
                        {
                            src += bcYlo[x];
                        }
                        if (i == Y - 1)
                        {
                            src += bcYhi[x];
                        }
                        rhs += src * 2.0 * rate / 2;
                    }
                    d[i] = rhs;
                }
                lineY.Solve(d.data(), v.data(), work);
                for (int32_t i = 0; i < Y; i++)
                {
                    size_t k = static_cast<size_t>(x) * Y + i;
                    deltas[k] += v[i] - field[k];
                }
            }
        });
    }

    // xoshiro256** seeded through splitmix64; src/Rng.cpp:13-20, 37-45, 56-69,
    // 91-96.
    struct Xoshiro
    {
        uint64_t s[4];
        static uint64_t SplitMix(uint64_t &st)
        {
            st += 0x9E3779B97F4A7C15ULL;
            uint64_t z = st;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            return z ^ (z >> 31);
        }
        static uint64_t Rotl(uint64_t x, int k)
        {
            return (x << k) | (x >> (64 - k));
        }
        explicit Xoshiro(uint64_t seed)
        {
            uint64_t st = seed;
            for (int i = 0; i < 4; i++)
            {
                s[i] = SplitMix(st);
            }
        }
        uint64_t Next()
        {
            uint64_t r = Rotl(s[1] * 5, 7) * 9;
            uint64_t t = s[1] << 17;
            s[2] ^= s[0];
            s[3] ^= s[1];
            s[1] ^= s[2];
            s[0] ^= s[3];
            s[2] ^= t;
            s[3] = Rotl(s[3], 45);
            return r;
        }
        double NextDouble()
        {
            return static_cast<double>(Next() >> 11) * (1.0 / 9007199254740992.0);
        }
    };

    // Clamped floor cell of a coordinate; src/Agents/CpuAgentForceBackend2D.cpp:14-27.
    inline int32_t CellCoord(double v, int32_t dim)
    {
        int32_t c = static_cast<int32_t>(std::floor(v));
        if (c < 0)
        {
            c = 0;
        }
        else if (c >= dim)
        {
            c = dim - 1;
        }
        return c;
    }

    // b - a along the shorter way round a wrapped axis;
    // src/Agents/CpuAgentForceBackend2D.cpp:30-45.
    inline double Disp(double a, double b, int32_t dim, bool wrap)
    {
        double d = b - a;
        if (wrap)
        {
            double dimF = static_cast<double>(dim);
            if (std::fabs(d) > dimF * 0.5)
            {
                d = (a > b) ? (d + dimF) : (d - dimF);
            }
        }
        return d;
    }
}

extern "C"
{
    // ---- RNG -----------------------------------------------------------------

    // n draws of Rng(seed).NextDouble(); src/Rng.cpp:91-96.
    void oracle_rng_doubles(uint64_t seed, int64_t n, double *out)
    {
        Xoshiro r(seed);
        for (int64_t i = 0; i < n; i++)
        {
            out[i] = r.NextDouble();
        }
    }

    // n raw NextUint64 outputs; src/Rng.cpp:56-69.
    void oracle_rng_u64(uint64_t seed, int64_t n, uint64_t *out)
    {
        Xoshiro r(seed);
        for (int64_t i = 0; i < n; i++)
        {
            out[i] = r.Next();
        }
    }

    // ---- explicit diffusion ----------------------------------------------------

    // deltas += rate * sum of four neighbour differences;
    // src/Pde/PDEequations.cpp:140-161 (bc arrays null) and
    // include/Hal/Pde/PDEequations.h:264-287 (Dirichlet ghost from edge arrays:
    // bcXlo[y] = bc(-1,y), bcXhi[y] = bc(X,y), bcYlo[x] = bc(x,-1), bcYhi[x] = bc(x,Y)).
    void oracle_diffuse(const double *field, double *deltas, double rate, int32_t X, int32_t Y, int32_t wrapX,
                        int32_t wrapY, const double *bcXlo, const double *bcXhi, const double *bcYlo,
                        const double *bcYhi)
    {
        for (int32_t x = 0; x < X; x++)
        {
            for (int32_t y = 0; y < Y; y++)
            {
                size_t i = static_cast<size_t>(x) * Y + y;
                double c = field[i];
                double s = DeltaX(field, c, x + 1, y, X, Y, wrapX != 0, bcXlo, bcXhi);
                s += DeltaX(field, c, x - 1, y, X, Y, wrapX != 0, bcXlo, bcXhi);
                s += DeltaY(field, c, x, y + 1, X, Y, wrapY != 0, bcYlo, bcYhi);
                s += DeltaY(field, c, x, y - 1, X, Y, wrapY != 0, bcYlo, bcYhi);
                deltas[i] += s * rate;
            }
        }
    }

    // Per-cell-rate explicit diffusion: the +x/+y face uses this cell's rate
    // (0 past a non-wrapped edge), the -x/-y face the neighbour's;
    // src/Pde/PDEequations.cpp:163-228.
    void oracle_diffuse_rates(const double *field, double *deltas, const double *ratesX, const double *ratesY,
                              int32_t X, int32_t Y, int32_t wrapX, int32_t wrapY)
    {
        auto dispX = [&](const double *v, int32_t x, int32_t y) -> double
        {
            if (x >= 0 && x < X)
            {
                return v[static_cast<size_t>(x) * Y + y];
            }
            if (wrapX)
            {
                return v[static_cast<size_t>(WrapI(x, X)) * Y + y];
            }
            return 0.0;
        };
        auto dispY = [&](const double *v, int32_t x, int32_t y) -> double
        {
            if (y >= 0 && y < Y)
            {
                return v[static_cast<size_t>(x) * Y + y];
            }
            if (wrapY)
            {
                return v[static_cast<size_t>(x) * Y + WrapI(y, Y)];
            }
            return 0.0;
        };
        for (int32_t x = 0; x < X; x++)
        {
            for (int32_t y = 0; y < Y; y++)
            {
                size_t i = static_cast<size_t>(x) * Y + y;
                double c = field[i];
                double dp1x = (x == X - 1 && !wrapX) ? dispX(ratesX, x + 1, y) : ratesX[i];
                double dm1x = dispX(ratesX, x - 1, y);
                double dp1y = (y == Y - 1 && !wrapY) ? dispY(ratesY, x, y + 1) : ratesY[i];
                double dm1y = dispY(ratesY, x, y - 1);
                double s = DeltaX(field, c, x + 1, y, X, Y, wrapX != 0, nullptr, nullptr) * dp1x;
                s += DeltaX(field, c, x - 1, y, X, Y, wrapX != 0, nullptr, nullptr) * dm1x;
                s += DeltaY(field, c, x, y + 1, X, Y, wrapY != 0, nullptr, nullptr) * dp1y;
                s += DeltaY(field, c, x, y - 1, X, Y, wrapY != 0, nullptr, nullptr) * dm1y;
                deltas[i] += s;
            }
        }
    }

    // `ticks` model ticks of explicit diffusion + decay + point sources, each
    // committed like PDEGrid2D::Update: deltas = Diffuse(u) ; deltas += -k*u
    // (MulAll(-k)) ; deltas[src] += s (Add) ; u += deltas.
    // include/Hal/Grids/PDEGrid2D.h:103-145, 178-186, 212-218.
    void oracle_explicit_ticks(double *field, int32_t ticks, double rate, double decay, int32_t X, int32_t Y,
                               int32_t wrapX, int32_t wrapY, int32_t nSrc, const int64_t *srcIdx,
                               const double *srcVal)
    {
        size_t n = static_cast<size_t>(X) * Y;
        std::vector<double> deltas(n, 0.0);
        for (int32_t t = 0; t < ticks; t++)
        {
            oracle_diffuse(field, deltas.data(), rate, X, Y, wrapX, wrapY, nullptr, nullptr, nullptr, nullptr);
            if (decay != 0.0)
            {
                for (size_t i = 0; i < n; i++)
                {
                    deltas[i] += field[i] * (-decay);
                }
            }
            for (int32_t s = 0; s < nSrc; s++)
            {
                deltas[static_cast<size_t>(srcIdx[s])] += srcVal[s];
            }
            for (size_t i = 0; i < n; i++)
            {
                field[i] += deltas[i];
                deltas[i] = 0.0;
            }
        }
    }

    // deltas += first-order upwind advection with a uniform velocity: the
    // upwind neighbour difference times the velocity along each axis, with
    // the same off-grid rules as diffusion (Dirichlet ghost (bc - c)*2 when
    // edges are given).  src/Pde/PDEequations.cpp:380-413 (zero flux /
    // periodic), include/Hal/Pde/PDEequations.h:413-447 (boundary function).
    void oracle_advect(const double *field, double *deltas, double xVel, double yVel, int32_t X, int32_t Y,
                       int32_t wrapX, int32_t wrapY, const double *bcXlo, const double *bcXhi, const double *bcYlo,
                       const double *bcYhi)
    {
        for (int32_t x = 0; x < X; x++)
        {
            for (int32_t y = 0; y < Y; y++)
            {
                const size_t i = static_cast<size_t>(x) * Y + y;
                const double c = field[i];
                if (xVel > 0)
                {
                    deltas[i] += DeltaX(field, c, x - 1, y, X, Y, wrapX != 0, bcXlo, bcXhi) * xVel;
                }
                else if (xVel < 0)
                {
                    deltas[i] -= DeltaX(field, c, x + 1, y, X, Y, wrapX != 0, bcXlo, bcXhi) * xVel;
                }
                if (yVel > 0)
                {
                    deltas[i] += DeltaY(field, c, x, y - 1, X, Y, wrapY != 0, bcYlo, bcYhi) * yVel;
                }
                else if (yVel < 0)
                {
                    deltas[i] -= DeltaY(field, c, x, y + 1, X, Y, wrapY != 0, bcYlo, bcYhi) * yVel;
                }
            }
        }
    }

    // deltas += conservative first-order upwind advection with per-cell
    // velocities: the flux through each face takes the velocity of the cell
    // on its low side (the cell itself on the +x/+y face, except at a
    // non-wrapped high edge where the off-grid velocity is used) and the
    // upwind cell's value.  Without boundary arrays off-grid values and
    // velocities are 0 (nothing crosses the walls); with them (bc for the
    // field, vbc for the velocity components: xlo/xhi the x velocity at
    // x = -1 / X, ylo/yhi the y velocity at y = -1 / Y) they are the given
    // values.  src/Pde/PDEequations.cpp:535-573, include/Hal/Pde/PDEequations.h:721-783.
    void oracle_advect_cells(const double *field, double *deltas, const double *xVels, const double *yVels, int32_t X,
                             int32_t Y, int32_t wrapX, int32_t wrapY, const double *bcXlo, const double *bcXhi,
                             const double *bcYlo, const double *bcYhi, const double *vXlo, const double *vXhi,
                             const double *vYlo, const double *vYhi)
    {
        const bool wx = wrapX != 0;
        const bool wy = wrapY != 0;
        for (int32_t x = 0; x < X; x++)
        {
            for (int32_t y = 0; y < Y; y++)
            {
                const size_t i = static_cast<size_t>(x) * Y + y;
                const double xVelP = (x == X - 1 && !wx) ? ShiftX(xVels, x + 1, y, X, Y, wx, vXlo, vXhi) : xVels[i];
                const double yVelP = (y == Y - 1 && !wy) ? ShiftY(yVels, x, y + 1, X, Y, wy, vYlo, vYhi) : yVels[i];
                const double xVelM = ShiftX(xVels, x - 1, y, X, Y, wx, vXlo, vXhi);
                const double yVelM = ShiftY(yVels, x, y - 1, X, Y, wy, vYlo, vYhi);
                const double c = field[i];
                double xFluxP = 0.0;
                double xFluxM = 0.0;
                double yFluxP = 0.0;
                double yFluxM = 0.0;
                if (xVelP > 0)
                {
                    xFluxP = xVelP * c;
                }
                else if (xVelP < 0)
                {
                    xFluxP = xVelP * ShiftX(field, x + 1, y, X, Y, wx, bcXlo, bcXhi);
                }
                if (xVelM > 0)
                {
                    xFluxM = xVelM * ShiftX(field, x - 1, y, X, Y, wx, bcXlo, bcXhi);
                }
                else if (xVelM < 0)
                {
                    xFluxM = xVelM * c;
                }
                if (yVelP > 0)
                {
                    yFluxP = yVelP * c;
                }
                else if (yVelP < 0)
                {
                    yFluxP = yVelP * ShiftY(field, x, y + 1, X, Y, wy, bcYlo, bcYhi);
                }
                if (yVelM > 0)
                {
                    yFluxM = yVelM * ShiftY(field, x, y - 1, X, Y, wy, bcYlo, bcYhi);
                }
                else if (yVelM < 0)
                {
                    yFluxM = yVelM * c;
                }
                deltas[i] += -xFluxP + xFluxM - yFluxP + yFluxM;
            }
        }
    }

    // `ticks` model ticks Diffusion(rate) + Advection(xVel, yVel) +
    // MulAll(-decay) + Add(sources) + Update, with Dirichlet edges for both
    // stencils when given (PDEGrid2D::Diffusion / Advection with a boundary,
    // include/Hal/Grids/PDEGrid2D.h:289-375).
    void oracle_model_ticks(double *field, int32_t ticks, double rate, double xVel, double yVel, double decay,
                            int32_t X, int32_t Y, int32_t wrapX, int32_t wrapY, const double *bcXlo,
                            const double *bcXhi, const double *bcYlo, const double *bcYhi, int32_t nSrc,
                            const int64_t *srcIdx, const double *srcVal)
    {
        const size_t n = static_cast<size_t>(X) * Y;
        std::vector<double> deltas(n, 0.0);
        for (int32_t t = 0; t < ticks; t++)
        {
            oracle_diffuse(field, deltas.data(), rate, X, Y, wrapX, wrapY, bcXlo, bcXhi, bcYlo, bcYhi);
            oracle_advect(field, deltas.data(), xVel, yVel, X, Y, wrapX, wrapY, bcXlo, bcXhi, bcYlo, bcYhi);
            if (decay != 0.0)
            {
                for (size_t i = 0; i < n; i++)
                {
                    deltas[i] += field[i] * (-decay);
                }
            }
            for (int32_t s = 0; s < nSrc; s++)
            {
                deltas[static_cast<size_t>(srcIdx[s])] += srcVal[s];
            }
            for (size_t i = 0; i < n; i++)
            {
                field[i] += deltas[i];
                deltas[i] = 0.0;
            }
        }
    }

    // ---- tridiagonal and ADI ------------------------------------------------

    void oracle_thomas(int32_t n, const double *a, const double *b, const double *c, const double *d, double *x)
    {
        std::vector<double> work;
        Thomas(n, a, b, c, d, x, work);
    }

    void oracle_thomas_periodic(int32_t n, const double *a, const double *b, const double *c, const double *d,
                                double *x)
    {
        std::vector<double> work;
        ThomasPeriodic(n, a, b, c, d, x, work);
    }

    // One ADI line solve with the reference's coefficients (ADISolveRow).
    // Worker threads of the ADI line loops and the per-agent force sums
    // (results do not depend on the count).
    void oracle_set_threads(int32_t threads)
    {
        g_threads = threads < 1 ? 1 : threads;
    }

    void oracle_adi_line(int32_t n, double rate, int32_t wrap, int32_t hasBc, const double *d, double *x)
    {
        std::vector<double> work;
        SolveLine(n, rate, wrap != 0, hasBc != 0, d, x, work);
    }

    // deltas += ADI(u) - u.  Null bc arrays: zero flux / periodic
    // (CpuAdiDiffusionBackend2D::DiffuseADI, src/Pde/CpuAdiDiffusionBackend2D.cpp:17-32);
    // non-null: Dirichlet ghost values as in oracle_diffuse.
    void oracle_adi(const double *field, double *deltas, double rate, int32_t X, int32_t Y, int32_t wrapX,
                    int32_t wrapY, const double *bcXlo, const double *bcXhi, const double *bcYlo,
                    const double *bcYhi)
    {
        Adi2D(field, deltas, rate, X, Y, wrapX != 0, wrapY != 0, bcXlo, bcXhi, bcYlo, bcYhi);
    }

    // `ticks` model ticks of ADI + decay + point sources, committed each tick
    // (PDEGrid2D::DiffusionADI / MulAll / Add / Update).
    void oracle_adi_ticks(double *field, int32_t ticks, double rate, double decay, int32_t X, int32_t Y,
                          int32_t wrapX, int32_t wrapY, const double *bcXlo, const double *bcXhi,
                          const double *bcYlo, const double *bcYhi, int32_t nSrc, const int64_t *srcIdx,
                          const double *srcVal)
    {
        size_t n = static_cast<size_t>(X) * Y;
        std::vector<double> deltas(n, 0.0);
        for (int32_t t = 0; t < ticks; t++)
        {
            Adi2D(field, deltas.data(), rate, X, Y, wrapX != 0, wrapY != 0, bcXlo, bcXhi, bcYlo, bcYhi);
            if (decay != 0.0)
            {
                for (size_t i = 0; i < n; i++)
                {
                    deltas[i] += field[i] * (-decay);
                }
            }
            for (int32_t s = 0; s < nSrc; s++)
            {
                deltas[static_cast<size_t>(srcIdx[s])] += srcVal[s];
            }
            for (size_t i = 0; i < n; i++)
            {
                field[i] += deltas[i];
                deltas[i] = 0.0;
            }
        }
    }

    // ---- agents -------------------------------------------------------------

    // Passes 1-3 of CpuAgentForceBackend2D::AccumulateForces: exclusive cell
    // offsets[C+1] and the stable (ascending agent index) bucket order.
    // src/Agents/CpuAgentForceBackend2D.cpp:85-109.
    void oracle_agent_bins(const double *pos, int32_t N, int32_t X, int32_t Y, int32_t *offsets, int32_t *bucket)
    {
        int64_t C = static_cast<int64_t>(X) * Y;
        std::memset(offsets, 0, sizeof(int32_t) * static_cast<size_t>(C + 1));
        for (int32_t i = 0; i < N; i++)
        {
            int32_t cx = CellCoord(pos[2 * static_cast<size_t>(i)], X);
            int32_t cy = CellCoord(pos[2 * static_cast<size_t>(i) + 1], Y);
            offsets[static_cast<size_t>(cx) * Y + cy + 1]++;
        }
        for (int64_t c = 0; c < C; c++)
        {
            offsets[c + 1] += offsets[c];
        }
        std::vector<int32_t> cursor(offsets, offsets + C);
        for (int32_t i = 0; i < N; i++)
        {
            int32_t cx = CellCoord(pos[2 * static_cast<size_t>(i)], X);
            int32_t cy = CellCoord(pos[2 * static_cast<size_t>(i) + 1], Y);
            size_t cell = static_cast<size_t>(cx) * Y + cy;
            bucket[cursor[cell]++] = i;
        }
    }

    // Pass 4: per agent, cells [floor(p-R), ceil(p+R)) per axis in bucket
    // order, linear spring k*((ri+rj)-d) for 0 < d < R, v -= (disp/d)*F.
    // Overwrites vel (2N).  src/Agents/CpuAgentForceBackend2D.cpp:111-175.
    void oracle_agent_forces(const double *pos, const double *radii, int32_t N, double *vel, double R, double k,
                             int32_t X, int32_t Y, int32_t wrapX, int32_t wrapY)
    {
        int64_t C = static_cast<int64_t>(X) * Y;
        std::vector<int32_t> offsets(static_cast<size_t>(C + 1));
        std::vector<int32_t> bucket(static_cast<size_t>(N > 0 ? N : 1));
        oracle_agent_bins(pos, N, X, Y, offsets.data(), bucket.data());
        // each agent's sum is independent: split over the worker threads
        ParallelRanges(N, [&](int32_t i0, int32_t i1) { for (int32_t i = i0; i < i1; i++)
        {
            double px = pos[2 * static_cast<size_t>(i)];
            double py = pos[2 * static_cast<size_t>(i) + 1];
            double rad = radii[i];
            int32_t x0 = static_cast<int32_t>(std::floor(px - R));
            int32_t x1 = static_cast<int32_t>(std::ceil(px + R));
            int32_t y0 = static_cast<int32_t>(std::floor(py - R));
            int32_t y1 = static_cast<int32_t>(std::ceil(py + R));
            double vx = 0.0;
            double vy = 0.0;
            for (int32_t xs = x0; xs < x1; xs++)
            {
                int32_t cx = xs;
                bool inX = (cx >= 0) && (cx < X);
                if (!inX)
                {
                    if (!wrapX)
                    {
                        continue;
                    }
                    cx = WrapI(cx, X);
                }
                for (int32_t ys = y0; ys < y1; ys++)
                {
                    int32_t cy = ys;
                    bool inY = (cy >= 0) && (cy < Y);
                    if (!inY)
                    {
                        if (!wrapY)
                        {
                            continue;
                        }
                        cy = WrapI(cy, Y);
                    }
                    size_t cell = static_cast<size_t>(cx) * Y + cy;
                    for (int32_t q = offsets[cell]; q < offsets[cell + 1]; q++)
                    {
                        int32_t j = bucket[q];
                        if (j == i)
                        {
                            continue;
                        }
                        double dx = Disp(px, pos[2 * static_cast<size_t>(j)], X, wrapX != 0);
                        double dy = Disp(py, pos[2 * static_cast<size_t>(j) + 1], Y, wrapY != 0);
                        double dist = std::sqrt((dx * dx) + (dy * dy));
                        if ((dist < R) && (dist > 0.0))
                        {
                            double force = k * ((rad + radii[j]) - dist);
                            vx -= (dx / dist) * force;
                            vy -= (dy / dist) * force;
                        }
                    }
                }
            }
            vel[2 * static_cast<size_t>(i)] = vx;
            vel[2 * static_cast<size_t>(i) + 1] = vy;
        } });
    }

    // Move with AgentPT2D::MoveSafePT semantics (include/Hal/Agents/AgentPT2D.h:157-186):
    // target p + dt*v; if the target cell is off the grid, a wrapped axis wraps
    // (fmod, src/MathUtil.cpp:48-53) and a non-wrapped axis keeps its old
    // coordinate.
    void oracle_agent_move(double *pos, const double *vel, int32_t N, double dt, int32_t X, int32_t Y,
                           int32_t wrapX, int32_t wrapY)
    {
        for (int32_t i = 0; i < N; i++)
        {
            double ox = pos[2 * static_cast<size_t>(i)];
            double oy = pos[2 * static_cast<size_t>(i) + 1];
            double nx = ox + dt * vel[2 * static_cast<size_t>(i)];
            double ny = oy + dt * vel[2 * static_cast<size_t>(i) + 1];
            int32_t fx = static_cast<int32_t>(std::floor(nx));
            int32_t fy = static_cast<int32_t>(std::floor(ny));
            bool in = (fx >= 0) && (fx < X) && (fy >= 0) && (fy < Y);
            if (!in)
            {
                if (wrapX)
                {
                    double m = std::fmod(nx, static_cast<double>(X));
                    nx = (m < 0) ? (X + m) : m;
                }
                else if (!((nx >= 0) && (nx < X)))
                {
                    nx = ox;
                }
                if (wrapY)
                {
                    double m = std::fmod(ny, static_cast<double>(Y));
                    ny = (m < 0) ? (Y + m) : m;
                }
                else if (!((ny >= 0) && (ny < Y)))
                {
                    ny = oy;
                }
            }
            pos[2 * static_cast<size_t>(i)] = nx;
            pos[2 * static_cast<size_t>(i) + 1] = ny;
        }
    }

    // `ticks` ticks of forces + move with a carried velocity `carry` (2N,
    // updated in place), as SphericalAgent2D keeps xVel/yVel
    // (include/Hal/Agents/SphericalAgent2D.h:100-156): v = carry + force
    // delta (SumForces accumulates into the velocity), |v| capped at maxVel
    // when maxVel > 0 (CapVelocity, :139-150), move by dt*v (ForceMove with
    // MoveSafePT), carry = friction*v (ApplyFriction, :133-137).
    void oracle_agent_ticks_motion(double *pos, double *carry, const double *radii, int32_t N, int32_t ticks,
                                   double R, double k, double dt, double friction, double maxVel, int32_t X,
                                   int32_t Y, int32_t wrapX, int32_t wrapY)
    {
        std::vector<double> vel(2 * static_cast<size_t>(N > 0 ? N : 1));
        for (int32_t t = 0; t < ticks; t++)
        {
            oracle_agent_forces(pos, radii, N, vel.data(), R, k, X, Y, wrapX, wrapY);
            for (int32_t i = 0; i < N; i++)
            {
                double vx = vel[2 * static_cast<size_t>(i)] + carry[2 * static_cast<size_t>(i)];
                double vy = vel[2 * static_cast<size_t>(i) + 1] + carry[2 * static_cast<size_t>(i) + 1];
                const double normSq = vx * vx + vy * vy;
                if (maxVel > 0.0 && normSq > maxVel * maxVel)
                {
                    const double f = maxVel / std::sqrt(normSq);
                    vx *= f;
                    vy *= f;
                }
                vel[2 * static_cast<size_t>(i)] = vx;
                vel[2 * static_cast<size_t>(i) + 1] = vy;
                carry[2 * static_cast<size_t>(i)] = friction * vx;
                carry[2 * static_cast<size_t>(i) + 1] = friction * vy;
            }
            oracle_agent_move(pos, vel.data(), N, dt, X, Y, wrapX, wrapY);
        }
    }

    // `ticks` ticks of forces + move.
    void oracle_agent_ticks(double *pos, const double *radii, int32_t N, int32_t ticks, double R, double k,
                            double dt, int32_t X, int32_t Y, int32_t wrapX, int32_t wrapY)
    {
        std::vector<double> vel(2 * static_cast<size_t>(N > 0 ? N : 1));
        for (int32_t t = 0; t < ticks; t++)
        {
            oracle_agent_forces(pos, radii, N, vel.data(), R, k, X, Y, wrapX, wrapY);
            oracle_agent_move(pos, vel.data(), N, dt, X, Y, wrapX, wrapY);
        }
    }
}

// ---- per-cell-diffusivity ADI and its adjoint (no reference counterpart) ----
//
// Forward model, generalising Diffusion2DADI (include/Hal/Pde/ADIequations.h:306-378)
// to a per-cell rate D with the face convention of the per-cell explicit
// stencil (src/Pde/PDEequations.cpp:199-228): the face between cells i and
// i+1 along an axis carries D_i, and on a wrapped axis the face between the
// last and the first cell carries D_{n-1}.  Boundaries per axis, as in the
// reference's ADI: periodic on a wrapped axis; otherwise zero flux, or
// Dirichlet with ghost values bc (the reference's DeltaX2D ghost rule
// (bc - c)*2, include/Hal/Pde/PDEequations.h:50-69), whose face carries the
// boundary cell's own D.  With
//   (L_x u)_i = D_i (u_{i+1} - u_i) + D_{i-1} (u_{i-1} - u_i)
//               + [Dirichlet end] 2 D_i (bc - u_i)        (absent faces dropped)
// (affine: L u = A u + g) a tick is
//   s1 = (I + L_y/2) u ;  s2 = (I - A_x/2)^{-1} (s1 + g_x/2) ;
//   s3 = (I + L_x/2) s2 ; s4 = (I - A_y/2)^{-1} (s3 + g_y/2) ;  u' = s4 - k u,
// which for uniform D = r is exactly the reference's ADI tick (diagonal
// 1 + 3r/2 and right-hand side r*bc at implicit Dirichlet ends, the cyclic
// line with corners -r/2 on a wrapped axis) plus MulAll(-k).  The loss is
// J = 1/2 |u_T - target|^2.
namespace
{
    // Boundary description of the per-cell model: wrap flags and, for a
    // Dirichlet grid, the ghost values as four edge arrays
    // [xlo(Y) | xhi(Y) | ylo(X) | yhi(X)] (bc(-1,y), bc(X,y), bc(x,-1), bc(x,Y));
    // edges == nullptr is zero flux on the non-wrapped axes.
    struct CellBounds
    {
        bool wrapX = false;
        bool wrapY = false;
        const double *edges = nullptr;
    };

    // Ghost value beyond cell (x, y) on the low (hi = false) or high end of axis.
    inline double Ghost(const CellBounds &b, int32_t X, int32_t Y, int axis, bool hi, int32_t x, int32_t y)
    {
        if (axis == 0)
        {
            return b.edges[(hi ? Y : 0) + y];
        }
        return b.edges[2 * Y + (hi ? X : 0) + x];
    }

    // Line solve (I - A/2) x = d along one line of a strided field: element j
    // of the line is at base + j*stride; the line's face rates are D at the
    // same addressing.  A periodic line is cyclic (corners -D_{n-1}/2,
    // include/Hal/Pde/TdmaSolver.h:68-120); a Dirichlet line adds D_j to the
    // diagonal of each end (its ghost face 2 D_j, halved).
    void CellLineSolve(int32_t n, const double *D, size_t base, size_t stride, const double *d, double *x,
                       bool wrap, bool dirichlet, std::vector<double> &work)
    {
        if (wrap && n == 1)
        {
            // the wrap face of a one-cell line joins the cell to itself
            x[base] = d[base];
            return;
        }
        std::vector<double> a(static_cast<size_t>(n)), b(static_cast<size_t>(n)), c(static_cast<size_t>(n));
        for (int32_t j = 0; j < n; j++)
        {
            double dm = (j > 0) ? D[base + (j - 1) * stride] : 0.0;
            double dp = (j < n - 1) ? D[base + j * stride] : 0.0;
            double diag = 0.5 * (dm + dp);
            if (wrap)
            {
                // the wrap face (n-1 .. 0) carries D_{n-1}
                const double dw = D[base + (n - 1) * stride];
                if (j == 0)
                {
                    diag += 0.5 * dw;
                    dm = dw;
                }
                if (j == n - 1)
                {
                    diag += 0.5 * dw;
                    dp = dw;
                }
            }
            else if (dirichlet)
            {
                const double own = D[base + j * stride];
                diag += (j == 0 ? own : 0.0) + (j == n - 1 ? own : 0.0);
            }
            a[j] = -0.5 * dm;
            c[j] = -0.5 * dp;
            b[j] = 1.0 + diag;
        }
        std::vector<double> dd(static_cast<size_t>(n)), xx(static_cast<size_t>(n));
        for (int32_t j = 0; j < n; j++)
        {
            dd[j] = d[base + j * stride];
        }
        if (wrap)
        {
            ThomasPeriodic(n, a.data(), b.data(), c.data(), dd.data(), xx.data(), work);
        }
        else
        {
            Thomas(n, a.data(), b.data(), c.data(), dd.data(), xx.data(), work);
        }
        for (int32_t j = 0; j < n; j++)
        {
            x[base + j * stride] = xx[j];
        }
    }

    // out = in + 1/2 L in along x (axis 0) or y (axis 1); withBc = false
    // applies the linear part A only (ghost values 0).
    void CellExplicit(int32_t X, int32_t Y, const double *D, const double *in, double *out, int axis,
                      const CellBounds &bd, bool withBc)
    {
        const bool wrap = (axis == 0) ? bd.wrapX : bd.wrapY;
        const bool dir = !wrap && bd.edges != nullptr;
        const int32_t n = (axis == 0) ? X : Y;
        ParallelRanges(X, [&](int32_t x0, int32_t x1) { for (int32_t x = x0; x < x1; x++)
        {
            for (int32_t y = 0; y < Y; y++)
            {
                const size_t i = static_cast<size_t>(x) * Y + y;
                const int32_t j = (axis == 0) ? x : y;
                const size_t step = (axis == 0) ? static_cast<size_t>(Y) : 1;
                const double c = in[i];
                double s = 0.0;
                if (j + 1 < n)
                {
                    s += D[i] * (in[i + step] - c);
                }
                else if (wrap && n > 1)
                {
                    s += D[i] * (in[i - (n - 1) * step] - c);
                }
                if (j > 0)
                {
                    s += D[i - step] * (in[i - step] - c);
                }
                else if (wrap && n > 1)
                {
                    s += D[i + (n - 1) * step] * (in[i + (n - 1) * step] - c);
                }
                if (dir)
                {
                    if (j == 0)
                    {
                        s += 2.0 * D[i] * ((withBc ? Ghost(bd, X, Y, axis, false, x, y) : 0.0) - c);
                    }
                    if (j == n - 1)
                    {
                        s += 2.0 * D[i] * ((withBc ? Ghost(bd, X, Y, axis, true, x, y) : 0.0) - c);
                    }
                }
                out[i] = c + 0.5 * s;
            }
        } });
    }

    // d += g/2 along the axis: D_i * bc at each Dirichlet end cell.
    void CellImplicitSource(int32_t X, int32_t Y, const double *D, double *d, int axis, const CellBounds &bd)
    {
        const bool wrap = (axis == 0) ? bd.wrapX : bd.wrapY;
        if (wrap || bd.edges == nullptr)
        {
            return;
        }
        for (int32_t x = 0; x < X; x++)
        {
            for (int32_t y = 0; y < Y; y++)
            {
                const int32_t j = (axis == 0) ? x : y;
                const int32_t n = (axis == 0) ? X : Y;
                const size_t i = static_cast<size_t>(x) * Y + y;
                if (j == 0)
                {
                    d[i] += D[i] * Ghost(bd, X, Y, axis, false, x, y);
                }
                if (j == n - 1)
                {
                    d[i] += D[i] * Ghost(bd, X, Y, axis, true, x, y);
                }
            }
        }
    }

    // x-implicit (axis 0: lines along x, one per y) or y-implicit solve.
    void CellSolve(int32_t X, int32_t Y, const double *D, const double *d, double *x, int axis, const CellBounds &bd)
    {
        if (axis == 0)
        {
            const bool dir = !bd.wrapX && bd.edges != nullptr;
            ParallelRanges(Y, [&](int32_t y0, int32_t y1)
            {
                std::vector<double> work;
                for (int32_t y = y0; y < y1; y++)
                {
                    CellLineSolve(X, D, static_cast<size_t>(y), static_cast<size_t>(Y), d, x, bd.wrapX, dir, work);
                }
            });
        }
        else
        {
            const bool dir = !bd.wrapY && bd.edges != nullptr;
            ParallelRanges(X, [&](int32_t x0, int32_t x1)
            {
                std::vector<double> work;
                for (int32_t xx = x0; xx < x1; xx++)
                {
                    CellLineSolve(Y, D, static_cast<size_t>(xx) * Y, 1, d, x, bd.wrapY, dir, work);
                }
            });
        }
    }

    // One forward tick; s2 and s4 receive the intermediate states if non-null.
    void CellTick(int32_t X, int32_t Y, const double *D, double k, const double *u, double *uNext, double *s2,
                  double *s4, const CellBounds &bd)
    {
        const size_t n = static_cast<size_t>(X) * Y;
        std::vector<double> t1(n), t2(n), t3(n), t4(n);
        CellExplicit(X, Y, D, u, t1.data(), 1, bd, true);
        CellImplicitSource(X, Y, D, t1.data(), 0, bd);
        CellSolve(X, Y, D, t1.data(), t2.data(), 0, bd);
        CellExplicit(X, Y, D, t2.data(), t3.data(), 0, bd, true);
        CellImplicitSource(X, Y, D, t3.data(), 1, bd);
        CellSolve(X, Y, D, t3.data(), t4.data(), 1, bd);
        for (size_t i = 0; i < n; i++)
        {
            uNext[i] = t4[i] - k * u[i];
        }
        if (s2 != nullptr)
        {
            std::copy(t2.begin(), t2.end(), s2);
        }
        if (s4 != nullptr)
        {
            std::copy(t4.begin(), t4.end(), s4);
        }
    }

    CellBounds Bounds(int32_t wrapX, int32_t wrapY, const double *edges)
    {
        CellBounds b;
        b.wrapX = wrapX != 0;
        b.wrapY = wrapY != 0;
        b.edges = edges;
        return b;
    }

    void CellForward(double *u, const double *D, double k, int32_t ticks, int32_t X, int32_t Y, const CellBounds &bd)
    {
        const size_t n = static_cast<size_t>(X) * Y;
        std::vector<double> next(n);
        for (int32_t t = 0; t < ticks; t++)
        {
            CellTick(X, Y, D, k, u, next.data(), nullptr, nullptr, bd);
            std::copy(next.begin(), next.end(), u);
        }
    }

    // Transposed tick: lambdaOut = M^T lambdaIn with M the linear part of the
    // tick at fixed D (every factor is symmetric, so M^T applies them in
    // reverse order: (I - A_y/2)^{-1}, (I + A_x/2), (I - A_x/2)^{-1},
    // (I + A_y/2), - k).
    void CellTickTranspose(const double *lambdaIn, double *lambdaOut, const double *D, double k, int32_t X,
                           int32_t Y, const CellBounds &bd)
    {
        const size_t n = static_cast<size_t>(X) * Y;
        std::vector<double> m1(n), w(n), m2(n);
        CellSolve(X, Y, D, lambdaIn, m1.data(), 1, bd);
        CellExplicit(X, Y, D, m1.data(), w.data(), 0, bd, false);
        CellSolve(X, Y, D, w.data(), m2.data(), 0, bd);
        CellExplicit(X, Y, D, m2.data(), lambdaOut, 1, bd, false);
        for (size_t i = 0; i < n; i++)
        {
            lambdaOut[i] -= k * lambdaIn[i];
        }
    }

    // J and its gradients by the discrete adjoint with every intermediate
    // stored.  Per tick (q, p pairs (m1, s4), (m2, u) on y faces and
    // (m1, s2), (m2, s2) on x faces, m1 = (I - A_y/2)^{-1} lam,
    // m2 = (I - A_x/2)^{-1} (I + A_x/2) m1):
    //   dJ/dD_c += 1/2 sum q^T (dL/dD_c) p
    // which is -1/2 dq dp on the interior or wrap face c owns
    // (dp = p(next) - p(c)) and q_c (bc - p_c) on each of c's Dirichlet faces;
    //   dJ/dk  -= <lam_{t+1}, u_t>.
    double CellGradient(const double *u0, const double *D, double k, int32_t ticks, int32_t X, int32_t Y,
                        const double *target, double *gradD, double *gradK, const CellBounds &bd)
    {
        const size_t n = static_cast<size_t>(X) * Y;
        std::vector<std::vector<double>> u(static_cast<size_t>(ticks) + 1, std::vector<double>(n));
        std::vector<std::vector<double>> s2(static_cast<size_t>(ticks), std::vector<double>(n));
        std::vector<std::vector<double>> s4(static_cast<size_t>(ticks), std::vector<double>(n));
        std::copy(u0, u0 + n, u[0].begin());
        for (int32_t t = 0; t < ticks; t++)
        {
            CellTick(X, Y, D, k, u[t].data(), u[t + 1].data(), s2[t].data(), s4[t].data(), bd);
        }
        double J = 0.0;
        std::vector<double> lam(n), m1(n), w(n), m2(n), next(n);
        for (size_t i = 0; i < n; i++)
        {
            const double r = u[ticks][i] - target[i];
            J += 0.5 * r * r;
            lam[i] = r;
        }
        std::fill(gradD, gradD + n, 0.0);
        double gk = 0.0;
        std::vector<double> gkRow(static_cast<size_t>(X));
        const bool dirX = !bd.wrapX && bd.edges != nullptr;
        const bool dirY = !bd.wrapY && bd.edges != nullptr;
        // forward differences over the face a cell owns (0 where it owns none)
        auto dx = [&](const std::vector<double> &p, size_t i, int32_t x)
        {
            if (x + 1 < X)
            {
                return p[i + Y] - p[i];
            }
            return (bd.wrapX && X > 1) ? p[i - static_cast<size_t>(X - 1) * Y] - p[i] : 0.0;
        };
        auto dy = [&](const std::vector<double> &p, size_t i, int32_t y)
        {
            if (y + 1 < Y)
            {
                return p[i + 1] - p[i];
            }
            return (bd.wrapY && Y > 1) ? p[i - static_cast<size_t>(Y - 1)] - p[i] : 0.0;
        };
        for (int32_t t = ticks - 1; t >= 0; t--)
        {
            CellSolve(X, Y, D, lam.data(), m1.data(), 1, bd);
            CellExplicit(X, Y, D, m1.data(), w.data(), 0, bd, false);
            CellSolve(X, Y, D, w.data(), m2.data(), 0, bd);
            // per-row partial sums of dJ/dk, added in row order: the same
            // bits for any thread count
            ParallelRanges(X, [&](int32_t x0, int32_t x1)
            {
                for (int32_t x = x0; x < x1; x++)
                {
                    double row = 0.0;
                    for (int32_t y = 0; y < Y; y++)
                    {
                        const size_t i = static_cast<size_t>(x) * Y + y;
                        const double g = dy(m1, i, y) * dy(s4[t], i, y) + dx(m1, i, x) * dx(s2[t], i, x) +
                                         dx(m2, i, x) * dx(s2[t], i, x) + dy(m2, i, y) * dy(u[t], i, y);
                        double gb = 0.0;
                        for (int hi = 0; hi < 2; hi++)
                        {
                            if (dirY && y == (hi ? Y - 1 : 0))
                            {
                                const double bc = Ghost(bd, X, Y, 1, hi != 0, x, y);
                                gb += m1[i] * (bc - s4[t][i]) + m2[i] * (bc - u[t][i]);
                            }
                            if (dirX && x == (hi ? X - 1 : 0))
                            {
                                const double bc = Ghost(bd, X, Y, 0, hi != 0, x, y);
                                gb += (m1[i] + m2[i]) * (bc - s2[t][i]);
                            }
                        }
                        gradD[i] += gb - 0.5 * g;
                        row -= lam[i] * u[t][i];
                    }
                    gkRow[static_cast<size_t>(x)] = row;
                }
            });
            for (int32_t x = 0; x < X; x++)
            {
                gk += gkRow[static_cast<size_t>(x)];
            }
            CellExplicit(X, Y, D, m2.data(), next.data(), 1, bd, false);
            for (size_t i = 0; i < n; i++)
            {
                next[i] -= k * lam[i];
            }
            lam.swap(next);
        }
        *gradK = gk;
        return J;
    }
}

extern "C"
{
    // `ticks` forward ticks of the per-cell-diffusivity model with zero-flux
    // boundaries; u updated in place.
    void oracle_adi_cell_forward(double *u, const double *D, double k, int32_t ticks, int32_t X, int32_t Y)
    {
        CellForward(u, D, k, ticks, X, Y, CellBounds{});
    }

    // The same with wrap flags and Dirichlet ghost values (edges: null for
    // zero flux, else [xlo(Y) | xhi(Y) | ylo(X) | yhi(X)]).
    void oracle_adi_cell_forward_bc(double *u, const double *D, double k, int32_t ticks, int32_t X, int32_t Y,
                                    int32_t wrapX, int32_t wrapY, const double *edges)
    {
        CellForward(u, D, k, ticks, X, Y, Bounds(wrapX, wrapY, edges));
    }

    // out = in + 1/2 L in along one axis (0: x, 1: y) with the ghost values
    // (withBc = 0: the linear part only).
    void oracle_adi_cell_explicit(const double *in, double *out, const double *D, int32_t X, int32_t Y, int32_t axis,
                                  int32_t wrapX, int32_t wrapY, const double *edges, int32_t withBc)
    {
        CellExplicit(X, Y, D, in, out, axis, Bounds(wrapX, wrapY, edges), withBc != 0);
    }

    void oracle_adi_cell_tick_transpose(const double *lambdaIn, double *lambdaOut, const double *D, double k,
                                        int32_t X, int32_t Y)
    {
        CellTickTranspose(lambdaIn, lambdaOut, D, k, X, Y, CellBounds{});
    }

    void oracle_adi_cell_tick_transpose_bc(const double *lambdaIn, double *lambdaOut, const double *D, double k,
                                           int32_t X, int32_t Y, int32_t wrapX, int32_t wrapY, const double *edges)
    {
        CellTickTranspose(lambdaIn, lambdaOut, D, k, X, Y, Bounds(wrapX, wrapY, edges));
    }

    // Loss J = 1/2 |u_T - target|^2 after `ticks` ticks from u0, and its
    // gradient with respect to every cell's D (gradD, X*Y) and to k (gradK).
    // Returns J.
    double oracle_adi_cell_gradient(const double *u0, const double *D, double k, int32_t ticks, int32_t X,
                                    int32_t Y, const double *target, double *gradD, double *gradK)
    {
        return CellGradient(u0, D, k, ticks, X, Y, target, gradD, gradK, CellBounds{});
    }

    double oracle_adi_cell_gradient_bc(const double *u0, const double *D, double k, int32_t ticks, int32_t X,
                                       int32_t Y, const double *target, double *gradD, double *gradK, int32_t wrapX,
                                       int32_t wrapY, const double *edges)
    {
        return CellGradient(u0, D, k, ticks, X, Y, target, gradD, gradK, Bounds(wrapX, wrapY, edges));
    }
}
