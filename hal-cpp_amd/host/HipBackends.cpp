// HipBackends.cpp -- Hal::Hip adapters of HAL-CPP's strategy interfaces over
// the C ABI.  Host code only: compiled with g++ and linked against
// libhalhip.so.  This file defines only Hal::Hip symbols, so it compiles into
// HAL's own library next to HAL's Logger, interface bases and
// Internal::CheckFieldSizes; a build without HAL links
// standalone/HalInterfaces.cpp for those instead.
#include "Hal/Hip/HipBackends.h"

#include <cmath>

#include "Hal/Log/Logger.h"

#include <iostream>
#include <stdexcept>
#include <string>

namespace Hal
{
    namespace Hip
    {
        namespace
        {
            [[noreturn]] void Throw(bool invalid, const std::string &message)
            {
                Logger::Instance().Log(LogSeverity::Error, message);
                if (invalid)
                {
                    throw std::invalid_argument(message);
                }
                throw std::runtime_error(message);
            }

            void RequireDims(bool fixed, int32_t xDim, int32_t yDim, int32_t planX, int32_t planY, const char *caller)
            {
                if (fixed && (xDim != planX || yDim != planY))
                {
                    Throw(true, std::string(caller) + " called with dimensions differing from construction");
                }
            }

            // halhip_bc view of a DirichletBoundary (arrays must be empty or
            // hold the edge lengths).
            halhip_bc ToBc(const DirichletBoundary &b, int32_t xDim, int32_t yDim, const char *caller)
            {
                auto check = [&](const std::vector<double> &v, int32_t n)
                {
                    if (!v.empty() && v.size() != static_cast<size_t>(n))
                    {
                        Throw(true, std::string(caller) + ": Dirichlet edge array has the wrong length");
                    }
                    return v.empty() ? nullptr : v.data();
                };
                halhip_bc bc{};
                bc.kind = HALHIP_BC_DIRICHLET;
                bc.value = b.value;
                bc.xlo = check(b.xlo, yDim);
                bc.xhi = check(b.xhi, yDim);
                bc.ylo = check(b.ylo, xDim);
                bc.yhi = check(b.yhi, xDim);
                return bc;
            }
        }

        void ThrowIfFailed(int status, const char *caller)
        {
            if (status == HALHIP_OK)
            {
                return;
            }
            Throw(status == HALHIP_ERR_INVALID_ARGUMENT, std::string(caller) + ": " + halhip_last_error());
        }

        // ---- context ---------------------------------------------------------------------

        HipContext::HipContext(int32_t device) : m_ctx(nullptr)
        {
            ThrowIfFailed(halhip_context_create(device, &m_ctx), "Hal::Hip::HipContext");
        }

        HipContext::~HipContext()
        {
            halhip_context_destroy(m_ctx);
        }

        std::unique_ptr<HipContext> HipContext::TryCreate(int32_t device)
        {
            halhip_context *probe = nullptr;
            int rc = halhip_context_create(device, &probe);
            if (rc == HALHIP_ERR_NO_DEVICE)
            {
                return nullptr;
            }
            ThrowIfFailed(rc, "Hal::Hip::HipContext::TryCreate");
            halhip_context_destroy(probe);
            return std::make_unique<HipContext>(device);
        }

        // ---- explicit diffusion ----------------------------------------------------------

        HipDiffusionBackend2D::HipDiffusionBackend2D(HipContext &context)
            : m_context(context), m_fixed(false), m_xDim(0), m_yDim(0), m_plan(nullptr)
        {
        }

        HipDiffusionBackend2D::HipDiffusionBackend2D(HipContext &context, int32_t xDim, int32_t yDim)
            : m_context(context), m_fixed(true), m_xDim(xDim), m_yDim(yDim), m_plan(nullptr)
        {
            ThrowIfFailed(halhip_diffusion2d_create(context.Handle(), xDim, yDim, &m_plan),
                          "HipDiffusionBackend2D");
        }

        HipDiffusionBackend2D::~HipDiffusionBackend2D()
        {
            halhip_diffusion2d_destroy(m_plan);
        }

        halhip_diffusion2d *HipDiffusionBackend2D::Plan(int32_t xDim, int32_t yDim, const char *caller)
        {
            RequireDims(m_fixed, xDim, yDim, m_xDim, m_yDim, caller);
            if (m_plan == nullptr || xDim != m_xDim || yDim != m_yDim)
            {
                halhip_diffusion2d_destroy(m_plan);
                m_plan = nullptr;
                ThrowIfFailed(halhip_diffusion2d_create(m_context.Handle(), xDim, yDim, &m_plan), caller);
                m_xDim = xDim;
                m_yDim = yDim;
            }
            return m_plan;
        }

        void HipDiffusionBackend2D::Diffuse(const std::vector<double> &field,
                                            std::vector<double>       &deltas,
                                            double                     diffRate,
                                            int32_t                    xDim,
                                            int32_t                    yDim,
                                            bool                       wrapX,
                                            bool                       wrapY)
        {
            const char *caller = "HipDiffusionBackend2D::Diffuse";
            Internal::CheckFieldSizes(field, deltas, xDim, yDim, caller);
            halhip_diffusion2d *plan = Plan(xDim, yDim, caller);
            ThrowIfFailed(halhip_diffusion2d_diffuse_f64(plan, field.data(), deltas.data(),
                                                         static_cast<int64_t>(field.size()), diffRate, wrapX, wrapY,
                                                         nullptr),
                          caller);
        }

        void HipDiffusionBackend2D::Diffuse(const std::vector<double> &field,
                                            std::vector<double>       &deltas,
                                            double                     diffRate,
                                            int32_t                    xDim,
                                            int32_t                    yDim,
                                            bool                       wrapX,
                                            bool                       wrapY,
                                            const DirichletBoundary   &bc)
        {
            const char *caller = "HipDiffusionBackend2D::Diffuse (Dirichlet)";
            Internal::CheckFieldSizes(field, deltas, xDim, yDim, caller);
            halhip_diffusion2d *plan = Plan(xDim, yDim, caller);
            halhip_bc b = ToBc(bc, xDim, yDim, caller);
            ThrowIfFailed(halhip_diffusion2d_diffuse_f64(plan, field.data(), deltas.data(),
                                                         static_cast<int64_t>(field.size()), diffRate, wrapX, wrapY,
                                                         &b),
                          caller);
        }

        void HipDiffusionBackend2D::DiffuseRates(const std::vector<double> &field,
                                                 std::vector<double>       &deltas,
                                                 const std::vector<double> &ratesX,
                                                 const std::vector<double> &ratesY,
                                                 int32_t                    xDim,
                                                 int32_t                    yDim,
                                                 bool                       wrapX,
                                                 bool                       wrapY)
        {
            const char *caller = "HipDiffusionBackend2D::DiffuseRates";
            Internal::CheckFieldSizes(field, deltas, xDim, yDim, caller);
            if (ratesX.size() != field.size() || ratesY.size() != field.size())
            {
                Throw(true, std::string(caller) + ": rate arrays must hold xDim*yDim values");
            }
            halhip_diffusion2d *plan = Plan(xDim, yDim, caller);
            ThrowIfFailed(halhip_diffusion2d_diffuse_rates_f64(plan, field.data(), deltas.data(), ratesX.data(),
                                                               ratesY.data(), static_cast<int64_t>(field.size()),
                                                               wrapX, wrapY),
                          caller);
        }

        void HipDiffusionBackend2D::RunSteps(std::vector<float> &field,
                                             int32_t             steps,
                                             float               diffRate,
                                             bool                wrapX,
                                             bool                wrapY,
                                             float               decay)
        {
            const char *caller = "HipDiffusionBackend2D::RunSteps";
            if (!m_fixed)
            {
                Throw(true, std::string(caller) + " needs a backend constructed with dimensions");
            }
            halhip_diffusion2d *plan = Plan(m_xDim, m_yDim, caller);
            ThrowIfFailed(halhip_diffusion2d_upload(plan, field.data(), static_cast<int64_t>(field.size())), caller);
            ThrowIfFailed(halhip_diffusion2d_run_steps(plan, steps, diffRate, decay, wrapX, wrapY, nullptr), caller);
            ThrowIfFailed(halhip_diffusion2d_download(plan, field.data(), static_cast<int64_t>(field.size())), caller);
        }

        void HipDiffusionBackend2D::Advect(const std::vector<double> &field,
                                           std::vector<double>       &deltas,
                                           double                     xVel,
                                           double                     yVel,
                                           int32_t                    xDim,
                                           int32_t                    yDim,
                                           bool                       wrapX,
                                           bool                       wrapY)
        {
            const char *caller = "HipDiffusionBackend2D::Advect";
            Internal::CheckFieldSizes(field, deltas, xDim, yDim, caller);
            halhip_diffusion2d *plan = Plan(xDim, yDim, caller);
            ThrowIfFailed(halhip_diffusion2d_advect_f64(plan, field.data(), deltas.data(),
                                                        static_cast<int64_t>(field.size()), xVel, yVel, wrapX, wrapY,
                                                        nullptr),
                          caller);
        }

        void HipDiffusionBackend2D::Advect(const std::vector<double> &field,
                                           std::vector<double>       &deltas,
                                           double                     xVel,
                                           double                     yVel,
                                           int32_t                    xDim,
                                           int32_t                    yDim,
                                           bool                       wrapX,
                                           bool                       wrapY,
                                           const DirichletBoundary   &bc)
        {
            const char *caller = "HipDiffusionBackend2D::Advect (Dirichlet)";
            Internal::CheckFieldSizes(field, deltas, xDim, yDim, caller);
            halhip_diffusion2d *plan = Plan(xDim, yDim, caller);
            halhip_bc b = ToBc(bc, xDim, yDim, caller);
            ThrowIfFailed(halhip_diffusion2d_advect_f64(plan, field.data(), deltas.data(),
                                                        static_cast<int64_t>(field.size()), xVel, yVel, wrapX, wrapY,
                                                        &b),
                          caller);
        }

        void HipDiffusionBackend2D::AdvectCells(const std::vector<double> &field,
                                                std::vector<double>       &deltas,
                                                const std::vector<double> &xVels,
                                                const std::vector<double> &yVels,
                                                int32_t                    xDim,
                                                int32_t                    yDim,
                                                bool                       wrapX,
                                                bool                       wrapY)
        {
            const char *caller = "HipDiffusionBackend2D::AdvectCells";
            Internal::CheckFieldSizes(field, deltas, xDim, yDim, caller);
            if (xVels.size() != field.size() || yVels.size() != field.size())
            {
                Throw(true, std::string(caller) + ": velocities must hold xDim*yDim values");
            }
            halhip_diffusion2d *plan = Plan(xDim, yDim, caller);
            ThrowIfFailed(halhip_diffusion2d_advect_cells_f64(plan, field.data(), deltas.data(), xVels.data(),
                                                              yVels.data(), static_cast<int64_t>(field.size()), wrapX,
                                                              wrapY, nullptr, nullptr),
                          caller);
        }

        void HipDiffusionBackend2D::SetAdvection(double xVel, double yVel)
        {
            const char *caller = "HipDiffusionBackend2D::SetAdvection";
            if (!m_fixed)
            {
                Throw(true, std::string(caller) + " needs a backend constructed with dimensions");
            }
            ThrowIfFailed(halhip_diffusion2d_set_advection(Plan(m_xDim, m_yDim, caller), xVel, yVel), caller);
        }

        // ---- device-resident PDEGrid2D ----------------------------------------------------

        HipPDEGrid2D::HipPDEGrid2D(HipContext &context, int32_t xDim, int32_t yDim, bool wrapX, bool wrapY)
            : m_context(context), m_xDim(xDim), m_yDim(yDim), m_wrapX(wrapX), m_wrapY(wrapY), m_updateCt(0),
              m_explicit(nullptr), m_adi(nullptr), m_resident(Stencil::Explicit), m_stencil(Stencil::None),
              m_rate(0.0), m_advect(false), m_xVel(0.0), m_yVel(0.0), m_hasBc(false), m_bc(0.0), m_bcSet(false),
              m_mulAll(0.0), m_explicitVel{0.0, 0.0}
        {
            ThrowIfFailed(halhip_diffusion2d_create(context.Handle(), xDim, yDim, &m_explicit), "HipPDEGrid2D");
        }

        HipPDEGrid2D::~HipPDEGrid2D()
        {
            halhip_adi2d_destroy(m_adi);
            halhip_diffusion2d_destroy(m_explicit);
        }

        void HipPDEGrid2D::MoveFieldTo(Stencil s)
        {
            if (s == m_resident)
            {
                return;
            }
            if (s == Stencil::Adi && m_adi == nullptr)
            {
                ThrowIfFailed(halhip_adi2d_create(m_context.Handle(), m_xDim, m_yDim, &m_adi), "HipPDEGrid2D::DiffusionADI");
            }
            const int64_t n = static_cast<int64_t>(m_xDim) * m_yDim;
            std::vector<float> host(static_cast<size_t>(n));
            if (m_resident == Stencil::Adi)
            {
                ThrowIfFailed(halhip_adi2d_download(m_adi, host.data(), n), "HipPDEGrid2D");
                ThrowIfFailed(halhip_diffusion2d_upload(m_explicit, host.data(), n), "HipPDEGrid2D");
            }
            else
            {
                ThrowIfFailed(halhip_diffusion2d_download(m_explicit, host.data(), n), "HipPDEGrid2D");
                ThrowIfFailed(halhip_adi2d_upload(m_adi, host.data(), n), "HipPDEGrid2D");
            }
            m_resident = s;
        }

        void HipPDEGrid2D::SetField(const std::vector<double> &field)
        {
            const char *caller = "HipPDEGrid2D::SetField";
            const int64_t n = static_cast<int64_t>(m_xDim) * m_yDim;
            if (static_cast<int64_t>(field.size()) != n)
            {
                Throw(true, std::string(caller) + ": field must hold xDim*yDim values");
            }
            std::vector<float> host(field.begin(), field.end());
            if (m_resident == Stencil::Adi)
            {
                ThrowIfFailed(halhip_adi2d_upload(m_adi, host.data(), n), caller);
            }
            else
            {
                ThrowIfFailed(halhip_diffusion2d_upload(m_explicit, host.data(), n), caller);
            }
        }

        std::vector<double> HipPDEGrid2D::GetField()
        {
            const char *caller = "HipPDEGrid2D::GetField";
            const int64_t n = static_cast<int64_t>(m_xDim) * m_yDim;
            std::vector<float> host(static_cast<size_t>(n));
            if (m_resident == Stencil::Adi)
            {
                ThrowIfFailed(halhip_adi2d_download(m_adi, host.data(), n), caller);
            }
            else
            {
                ThrowIfFailed(halhip_diffusion2d_download(m_explicit, host.data(), n), caller);
            }
            return std::vector<double>(host.begin(), host.end());
        }

        double HipPDEGrid2D::Get(int32_t i)
        {
            if (i < 0 || static_cast<int64_t>(i) >= static_cast<int64_t>(m_xDim) * m_yDim)
            {
                Throw(true, "HipPDEGrid2D::Get: index out of range");
            }
            return GetField()[static_cast<size_t>(i)];
        }

        void HipPDEGrid2D::SetBoundary(bool has, double value, const char *caller)
        {
            if (m_bcSet && (has != m_hasBc || (has && value != m_bc)))
            {
                Throw(true, std::string(caller) + ": the calls of one tick must share their boundary value");
            }
            m_bcSet = true;
            m_hasBc = has;
            m_bc = value;
        }

        void HipPDEGrid2D::SetStencil(Stencil s, double rate, const char *caller)
        {
            if (m_stencil != Stencil::None)
            {
                Throw(true, std::string(caller) + ": one diffusion call per tick on the device grid");
            }
            if (s == Stencil::Adi && m_advect)
            {
                Throw(true, std::string(caller) + ": advection is fused with explicit diffusion only");
            }
            m_stencil = s;
            m_rate = rate;
        }

        void HipPDEGrid2D::Diffusion(double diffCoef)
        {
            // PDEGrid2D::CheckDiffusionStable (include/Hal/Grids/PDEGrid2D.h:436-442)
            if (diffCoef > 0.25)
            {
                Throw(true, "HipPDEGrid2D::Diffusion: Diffusion rate above stable maximum value of 0.25");
            }
            SetBoundary(false, 0.0, "HipPDEGrid2D::Diffusion");
            SetStencil(Stencil::Explicit, diffCoef, "HipPDEGrid2D::Diffusion");
        }

        void HipPDEGrid2D::Diffusion(double diffCoef, double boundaryValue)
        {
            if (diffCoef > 0.25)
            {
                Throw(true, "HipPDEGrid2D::Diffusion: Diffusion rate above stable maximum value of 0.25");
            }
            SetBoundary(true, boundaryValue, "HipPDEGrid2D::Diffusion");
            SetStencil(Stencil::Explicit, diffCoef, "HipPDEGrid2D::Diffusion");
        }

        void HipPDEGrid2D::DiffusionADI(double diffRate)
        {
            SetBoundary(false, 0.0, "HipPDEGrid2D::DiffusionADI");
            SetStencil(Stencil::Adi, diffRate, "HipPDEGrid2D::DiffusionADI");
        }

        void HipPDEGrid2D::DiffusionADI(double diffRate, double boundaryValue)
        {
            SetBoundary(true, boundaryValue, "HipPDEGrid2D::DiffusionADI");
            SetStencil(Stencil::Adi, diffRate, "HipPDEGrid2D::DiffusionADI");
        }

        void HipPDEGrid2D::Advection(double xVel, double yVel)
        {
            const char *caller = "HipPDEGrid2D::Advection";
            if (std::fabs(xVel) + std::fabs(yVel) > 0.5)
            {
                Throw(true, std::string(caller) + ": Advection rate component sum above stable maximum value of 0.5");
            }
            if (m_stencil == Stencil::Adi || m_advect)
            {
                Throw(true, std::string(caller) + ": one advection call per tick, fused with explicit diffusion only");
            }
            SetBoundary(false, 0.0, caller);
            m_advect = true;
            m_xVel = xVel;
            m_yVel = yVel;
        }

        void HipPDEGrid2D::Advection(double xVel, double yVel, double boundaryValue)
        {
            const char *caller = "HipPDEGrid2D::Advection";
            if (std::fabs(xVel) + std::fabs(yVel) > 0.5)
            {
                Throw(true, std::string(caller) + ": Advection rate component sum above stable maximum value of 0.5");
            }
            if (m_stencil == Stencil::Adi || m_advect)
            {
                Throw(true, std::string(caller) + ": one advection call per tick, fused with explicit diffusion only");
            }
            SetBoundary(true, boundaryValue, caller);
            m_advect = true;
            m_xVel = xVel;
            m_yVel = yVel;
        }

        void HipPDEGrid2D::MulAll(double val)
        {
            m_mulAll += val;
        }

        void HipPDEGrid2D::Add(int32_t i, double val)
        {
            if (i < 0 || static_cast<int64_t>(i) >= static_cast<int64_t>(m_xDim) * m_yDim)
            {
                Throw(true, "HipPDEGrid2D::Add: index out of range");
            }
            m_adds.emplace_back(static_cast<int64_t>(i), static_cast<float>(val));
        }

        void HipPDEGrid2D::Update()
        {
            const char *caller = "HipPDEGrid2D::Update";
            const bool pending = m_stencil != Stencil::None || m_advect || m_mulAll != 0.0 || !m_adds.empty();
            if (pending)
            {
                halhip_bc bc{};
                bc.kind = HALHIP_BC_DIRICHLET;
                bc.value = m_bc;
                const halhip_bc *pbc = m_hasBc ? &bc : nullptr;
                std::vector<int64_t> idx;
                std::vector<float> val;
                for (const auto &a : m_adds)
                {
                    idx.push_back(a.first);
                    val.push_back(a.second);
                }
                // MulAll(v) adds v*u to the deltas: the tick's decay is -sum(v).
                const double decay = -m_mulAll;
                if (m_stencil == Stencil::Adi)
                {
                    MoveFieldTo(Stencil::Adi);
                    if (m_adds != m_adiSources)
                    {
                        ThrowIfFailed(halhip_adi2d_set_sources(m_adi, static_cast<int32_t>(idx.size()), idx.data(),
                                                               val.data()),
                                      caller);
                        m_adiSources = m_adds;
                    }
                    ThrowIfFailed(halhip_adi2d_run_steps_async(m_adi, 1, m_rate, decay, m_wrapX, m_wrapY, pbc), caller);
                }
                else
                {
                    // explicit diffusion (rate 0 when the tick has none)
                    MoveFieldTo(Stencil::Explicit);
                    if (m_adds != m_explicitSources)
                    {
                        ThrowIfFailed(halhip_diffusion2d_set_sources(m_explicit, static_cast<int32_t>(idx.size()),
                                                                     idx.data(), val.data()),
                                      caller);
                        m_explicitSources = m_adds;
                    }
                    const double vx = m_advect ? m_xVel : 0.0;
                    const double vy = m_advect ? m_yVel : 0.0;
                    if (vx != m_explicitVel[0] || vy != m_explicitVel[1])
                    {
                        ThrowIfFailed(halhip_diffusion2d_set_advection(m_explicit, vx, vy), caller);
                        m_explicitVel[0] = vx;
                        m_explicitVel[1] = vy;
                    }
                    const double rate = (m_stencil == Stencil::Explicit) ? m_rate : 0.0;
                    ThrowIfFailed(halhip_diffusion2d_run_steps_async(m_explicit, 1, rate, decay, m_wrapX, m_wrapY, pbc),
                                  caller);
                }
            }
            m_stencil = Stencil::None;
            m_rate = 0.0;
            m_advect = false;
            m_hasBc = false;
            m_bcSet = false;
            m_mulAll = 0.0;
            m_adds.clear();
            m_updateCt++;
        }

        void HipPDEGrid2D::Save(const std::string &path)
        {
            const char *caller = "HipPDEGrid2D::Save";
            if (m_stencil != Stencil::None || m_advect || m_mulAll != 0.0 || !m_adds.empty())
            {
                Throw(true, std::string(caller) + ": the grid has pending calls; Update before saving");
            }
            if (m_resident == Stencil::Adi)
            {
                ThrowIfFailed(halhip_adi2d_save_grid(m_adi, path.c_str(), m_wrapX, m_wrapY, m_updateCt), caller);
            }
            else
            {
                ThrowIfFailed(halhip_diffusion2d_save_grid(m_explicit, path.c_str(), m_wrapX, m_wrapY, m_updateCt),
                              caller);
            }
        }

        void HipPDEGrid2D::Load(const std::string &path)
        {
            const char *caller = "HipPDEGrid2D::Load";
            int32_t wx = 0;
            int32_t wy = 0;
            int32_t ct = 0;
            if (m_stencil != Stencil::None || m_advect || m_mulAll != 0.0 || !m_adds.empty())
            {
                Throw(true, std::string(caller) + ": the grid has pending calls; Update before loading");
            }
            MoveFieldTo(Stencil::Explicit);
            ThrowIfFailed(halhip_diffusion2d_load_grid(m_explicit, path.c_str(), &wx, &wy, &ct), caller);
            if ((wx != 0) != m_wrapX || (wy != 0) != m_wrapY)
            {
                Throw(true, std::string(caller) + ": the saved grid's wrap flags differ from this grid's");
            }
            m_updateCt = ct;
        }

        // ---- ADI diffusion ---------------------------------------------------------------

        HipAdiDiffusionBackend2D::HipAdiDiffusionBackend2D(HipContext &context)
            : m_context(context), m_fixed(false), m_xDim(0), m_yDim(0), m_plan(nullptr)
        {
        }

        HipAdiDiffusionBackend2D::HipAdiDiffusionBackend2D(HipContext &context, int32_t xDim, int32_t yDim)
            : m_context(context), m_fixed(true), m_xDim(xDim), m_yDim(yDim), m_plan(nullptr)
        {
            ThrowIfFailed(halhip_adi2d_create(context.Handle(), xDim, yDim, &m_plan), "HipAdiDiffusionBackend2D");
        }

        HipAdiDiffusionBackend2D::~HipAdiDiffusionBackend2D()
        {
            halhip_adi2d_destroy(m_plan);
        }

        halhip_adi2d *HipAdiDiffusionBackend2D::Plan(int32_t xDim, int32_t yDim, const char *caller)
        {
            RequireDims(m_fixed, xDim, yDim, m_xDim, m_yDim, caller);
            if (m_plan == nullptr || xDim != m_xDim || yDim != m_yDim)
            {
                halhip_adi2d_destroy(m_plan);
                m_plan = nullptr;
                ThrowIfFailed(halhip_adi2d_create(m_context.Handle(), xDim, yDim, &m_plan), caller);
                m_xDim = xDim;
                m_yDim = yDim;
            }
            return m_plan;
        }

        void HipAdiDiffusionBackend2D::DiffuseADI(const std::vector<double> &field,
                                                  std::vector<double>       &deltas,
                                                  double                     diffRate,
                                                  int32_t                    xDim,
                                                  int32_t                    yDim,
                                                  bool                       wrapX,
                                                  bool                       wrapY)
        {
            const char *caller = "HipAdiDiffusionBackend2D::DiffuseADI";
            Internal::CheckFieldSizes(field, deltas, xDim, yDim, caller);
            halhip_adi2d *plan = Plan(xDim, yDim, caller);
            ThrowIfFailed(halhip_adi2d_diffuse_f64(plan, field.data(), deltas.data(), static_cast<int64_t>(field.size()),
                                                   diffRate, wrapX, wrapY, nullptr),
                          caller);
        }

        void HipAdiDiffusionBackend2D::DiffuseADI(const std::vector<double> &field,
                                                  std::vector<double>       &deltas,
                                                  double                     diffRate,
                                                  int32_t                    xDim,
                                                  int32_t                    yDim,
                                                  bool                       wrapX,
                                                  bool                       wrapY,
                                                  const DirichletBoundary   &bc)
        {
            const char *caller = "HipAdiDiffusionBackend2D::DiffuseADI (Dirichlet)";
            Internal::CheckFieldSizes(field, deltas, xDim, yDim, caller);
            halhip_adi2d *plan = Plan(xDim, yDim, caller);
            halhip_bc b = ToBc(bc, xDim, yDim, caller);
            ThrowIfFailed(halhip_adi2d_diffuse_f64(plan, field.data(), deltas.data(), static_cast<int64_t>(field.size()),
                                                   diffRate, wrapX, wrapY, &b),
                          caller);
        }

        void HipAdiDiffusionBackend2D::RunSteps(std::vector<float> &field,
                                                int32_t             steps,
                                                float               diffRate,
                                                bool                wrapX,
                                                bool                wrapY)
        {
            const char *caller = "HipAdiDiffusionBackend2D::RunSteps";
            if (!m_fixed)
            {
                Throw(true, std::string(caller) + " needs a backend constructed with dimensions");
            }
            halhip_adi2d *plan = Plan(m_xDim, m_yDim, caller);
            ThrowIfFailed(halhip_adi2d_upload(plan, field.data(), static_cast<int64_t>(field.size())), caller);
            ThrowIfFailed(halhip_adi2d_run_steps(plan, steps, diffRate, 0.0, wrapX, wrapY, nullptr), caller);
            ThrowIfFailed(halhip_adi2d_download(plan, field.data(), static_cast<int64_t>(field.size())), caller);
        }

        void HipAdiDiffusionBackend2D::RunStepsDirichlet(std::vector<float> &field,
                                                         int32_t             steps,
                                                         float               diffRate,
                                                         bool                wrapX,
                                                         bool                wrapY,
                                                         double              boundaryValue)
        {
            const char *caller = "HipAdiDiffusionBackend2D::RunStepsDirichlet";
            if (!m_fixed)
            {
                Throw(true, std::string(caller) + " needs a backend constructed with dimensions");
            }
            halhip_adi2d *plan = Plan(m_xDim, m_yDim, caller);
            halhip_bc b{};
            b.kind = HALHIP_BC_DIRICHLET;
            b.value = boundaryValue;
            ThrowIfFailed(halhip_adi2d_upload(plan, field.data(), static_cast<int64_t>(field.size())), caller);
            ThrowIfFailed(halhip_adi2d_run_steps(plan, steps, diffRate, 0.0, wrapX, wrapY, &b), caller);
            ThrowIfFailed(halhip_adi2d_download(plan, field.data(), static_cast<int64_t>(field.size())), caller);
        }

        // ---- agents ----------------------------------------------------------------------

        HipAgentForceBackend2D::HipAgentForceBackend2D(HipContext &context, int32_t xDim, int32_t yDim,
                                                       int32_t maxAgents)
            : m_xDim(xDim), m_yDim(yDim), m_maxAgents(maxAgents), m_plan(nullptr)
        {
            ThrowIfFailed(halhip_agents2d_create(context.Handle(), xDim, yDim, maxAgents, &m_plan),
                          "HipAgentForceBackend2D");
        }

        HipAgentForceBackend2D::~HipAgentForceBackend2D()
        {
            halhip_agents2d_destroy(m_plan);
        }

        void HipAgentForceBackend2D::AccumulateForces(const std::vector<double> &positions,
                                                      const std::vector<double> &radii,
                                                      std::vector<double>       &velocityDeltas,
                                                      double                     interactionRad,
                                                      double                     springConstant,
                                                      int32_t                    xDim,
                                                      int32_t                    yDim,
                                                      bool                       wrapX,
                                                      bool                       wrapY)
        {
            const char *caller = "HipAgentForceBackend2D::AccumulateForces";
            RequireDims(true, xDim, yDim, m_xDim, m_yDim, caller);
            if (positions.size() != radii.size() * 2)
            {
                Throw(true, std::string(caller) + ": positions must hold two values per radius");
            }
            if (radii.size() > static_cast<size_t>(m_maxAgents))
            {
                Throw(true, std::string(caller) + " called with more agents than the construction-time capacity");
            }
            velocityDeltas.assign(positions.size(), 0.0);
            ThrowIfFailed(halhip_agents2d_forces_f64(m_plan, positions.data(), radii.data(),
                                                     static_cast<int32_t>(radii.size()), velocityDeltas.data(),
                                                     interactionRad, springConstant, wrapX, wrapY),
                          caller);
        }

        void HipAgentForceBackend2D::RunTicks(std::vector<float>       &positions,
                                              const std::vector<float> &radii,
                                              int32_t                   ticks,
                                              float                     interactionRad,
                                              float                     springConstant,
                                              float                     dt,
                                              bool                      wrapX,
                                              bool                      wrapY)
        {
            const char *caller = "HipAgentForceBackend2D::RunTicks";
            if (positions.size() != radii.size() * 2)
            {
                Throw(true, std::string(caller) + ": positions must hold two values per radius");
            }
            const int32_t n = static_cast<int32_t>(radii.size());
            ThrowIfFailed(halhip_agents2d_upload(m_plan, positions.data(), radii.data(), n), caller);
            ThrowIfFailed(halhip_agents2d_run_ticks(m_plan, ticks, interactionRad, springConstant, dt, wrapX, wrapY),
                          caller);
            ThrowIfFailed(halhip_agents2d_download_positions(m_plan, positions.data(), n), caller);
        }
    }
}
