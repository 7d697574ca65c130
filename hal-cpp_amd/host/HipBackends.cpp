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

        namespace
        {
            // halhip_bc over a recorded boundary (constant or edge arrays).
            struct BcView
            {
                halhip_bc bc{};
                const halhip_bc *Ptr() const { return bc.kind == HALHIP_BC_DIRICHLET ? &bc : nullptr; }
            };

            BcView ViewOf(bool has, bool constant, double value, const DirichletBoundary *edges)
            {
                BcView v;
                if (!has)
                {
                    return v;
                }
                v.bc.kind = HALHIP_BC_DIRICHLET;
                v.bc.value = value;
                if (!constant && edges != nullptr)
                {
                    v.bc.xlo = edges->xlo.data();
                    v.bc.xhi = edges->xhi.data();
                    v.bc.ylo = edges->ylo.data();
                    v.bc.yhi = edges->yhi.data();
                }
                return v;
            }

            // PDEGrid2D::CheckDiffusionStable / CheckAdvectionStable
            // (include/Hal/Grids/PDEGrid2D.h:436-451)
            void CheckDiffusionRate(double diffCoef, const char *caller)
            {
                if (diffCoef > 0.25)
                {
                    Throw(true, std::string(caller) + ": Diffusion rate above stable maximum value of 0.25");
                }
            }

            void CheckAdvectionRate(double xVel, double yVel, const char *caller)
            {
                if (std::fabs(xVel) + std::fabs(yVel) > 0.5)
                {
                    Throw(true, std::string(caller) + ": Advection rate component sum above stable maximum value of 0.5");
                }
            }
        }

        HipPDEGrid2D::HipPDEGrid2D(HipContext &context, int32_t xDim, int32_t yDim, bool wrapX, bool wrapY)
            : m_context(context), m_xDim(xDim), m_yDim(yDim), m_wrapX(wrapX), m_wrapY(wrapY), m_updateCt(0),
              m_explicit(nullptr), m_adi(nullptr), m_resident(Stencil::Explicit), m_general(false),
              m_explicitVel{0.0, 0.0}
        {
            ThrowIfFailed(halhip_diffusion2d_create(context.Handle(), xDim, yDim, &m_explicit), "HipPDEGrid2D");
        }

        HipPDEGrid2D::~HipPDEGrid2D()
        {
            halhip_adi2d_destroy(m_adi);
            halhip_diffusion2d_destroy(m_explicit);
        }

        halhip_adi2d *HipPDEGrid2D::Adi()
        {
            if (m_adi == nullptr)
            {
                ThrowIfFailed(halhip_adi2d_create(m_context.Handle(), m_xDim, m_yDim, &m_adi), "HipPDEGrid2D::DiffusionADI");
            }
            return m_adi;
        }

        void HipPDEGrid2D::MoveFieldTo(Stencil s)
        {
            if (s == m_resident)
            {
                return;
            }
            Adi();
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

        void HipPDEGrid2D::ClearTick()
        {
            m_ops.clear();
            m_pointOps.clear();
            m_pointIdx.clear();
            m_pointVal.clear();
            m_general = false;
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
                ThrowIfFailed(halhip_diffusion2d_upload(m_explicit, host.data(), n), caller); // clears the deltas
            }
            ClearTick();
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
            CheckIndex(i, "HipPDEGrid2D::Get");
            return GetField()[static_cast<size_t>(i)];
        }

        void HipPDEGrid2D::CheckIndex(int32_t i, const char *caller) const
        {
            if (i < 0 || static_cast<int64_t>(i) >= static_cast<int64_t>(m_xDim) * m_yDim)
            {
                Throw(true, std::string(caller) + ": index out of range");
            }
        }

        // The committed field's reductions and the pending deltas' need both
        // on the explicit plan; the deltas are materialised first.
        double HipPDEGrid2D::Reduce(int32_t what, double offset, const char *caller)
        {
            if (what == HALHIP_REDUCE_MAX_DELTA || what == HALHIP_REDUCE_MAX_DELTA_SCALED)
            {
                Materialize();
                FlushPoints();
            }
            else if (m_resident != Stencil::Explicit)
            {
                // the pending calls were recorded against the committed field,
                // which is the same wherever it lives
                MoveFieldTo(Stencil::Explicit);
            }
            double out = 0.0;
            ThrowIfFailed(halhip_diffusion2d_delta_reduce(m_explicit, what, offset, &out), caller);
            return out;
        }

        double HipPDEGrid2D::GetMax()
        {
            return Reduce(HALHIP_REDUCE_FIELD_MAX, 0.0, "HipPDEGrid2D::GetMax");
        }

        double HipPDEGrid2D::GetMin()
        {
            return Reduce(HALHIP_REDUCE_FIELD_MIN, 0.0, "HipPDEGrid2D::GetMin");
        }

        double HipPDEGrid2D::GetAvg()
        {
            return Reduce(HALHIP_REDUCE_FIELD_SUM, 0.0, "HipPDEGrid2D::GetAvg") / Length();
        }

        double HipPDEGrid2D::MaxDelta()
        {
            return Reduce(HALHIP_REDUCE_MAX_DELTA, 0.0, "HipPDEGrid2D::MaxDelta");
        }

        double HipPDEGrid2D::MaxDeltaScaled(double denomOffset)
        {
            return Reduce(HALHIP_REDUCE_MAX_DELTA_SCALED, denomOffset, "HipPDEGrid2D::MaxDeltaScaled");
        }

        void HipPDEGrid2D::Record(Op op)
        {
            if (m_general)
            {
                Issue(op);
                return;
            }
            m_ops.push_back(std::move(op));
        }

        void HipPDEGrid2D::Set(int32_t i, double val)
        {
            CheckIndex(i, "HipPDEGrid2D::Set");
            Op op(Op::Kind::Point);
            op.code = HALHIP_DELTA_SET;
            op.index = i;
            op.a = val;
            Record(std::move(op));
        }

        void HipPDEGrid2D::Add(int32_t i, double val)
        {
            CheckIndex(i, "HipPDEGrid2D::Add");
            Op op(Op::Kind::Point);
            op.code = HALHIP_DELTA_ADD;
            op.index = i;
            op.a = val;
            Record(std::move(op));
        }

        void HipPDEGrid2D::Mul(int32_t i, double val)
        {
            CheckIndex(i, "HipPDEGrid2D::Mul");
            Op op(Op::Kind::Point);
            op.code = HALHIP_DELTA_MUL;
            op.index = i;
            op.a = val;
            Record(std::move(op));
        }

        void HipPDEGrid2D::Scale(int32_t i, double val)
        {
            CheckIndex(i, "HipPDEGrid2D::Scale");
            Op op(Op::Kind::Point);
            op.code = HALHIP_DELTA_SCALE;
            op.index = i;
            op.a = val;
            Record(std::move(op));
        }

        void HipPDEGrid2D::SetAll(double val)
        {
            Op op(Op::Kind::All);
            op.code = HALHIP_DELTA_SET;
            op.a = val;
            Record(std::move(op));
        }

        void HipPDEGrid2D::SetAll(const double *pVals)
        {
            if (pVals == nullptr)
            {
                Throw(true, "HipPDEGrid2D::SetAll: null values");
            }
            Op op(Op::Kind::SetValues);
            op.u = std::make_shared<const std::vector<double>>(pVals, pVals + Length());
            Record(std::move(op));
        }

        void HipPDEGrid2D::AddAll(double val)
        {
            Op op(Op::Kind::All);
            op.code = HALHIP_DELTA_ADD;
            op.a = val;
            Record(std::move(op));
        }

        void HipPDEGrid2D::MulAll(double val)
        {
            Op op(Op::Kind::All);
            op.code = HALHIP_DELTA_MUL;
            op.a = val;
            Record(std::move(op));
        }

        void HipPDEGrid2D::SetNonNegative()
        {
            Op op(Op::Kind::All);
            op.code = HALHIP_DELTA_SET_NON_NEGATIVE;
            Record(std::move(op));
        }

        void HipPDEGrid2D::Diffusion(double diffCoef)
        {
            CheckDiffusionRate(diffCoef, "HipPDEGrid2D::Diffusion");
            Op op(Op::Kind::Diffusion);
            op.a = diffCoef;
            Record(std::move(op));
        }

        void HipPDEGrid2D::Diffusion(double diffCoef, double boundaryValue)
        {
            CheckDiffusionRate(diffCoef, "HipPDEGrid2D::Diffusion");
            Op op(Op::Kind::Diffusion);
            op.a = diffCoef;
            op.hasBc = true;
            op.constBc = true;
            op.bcValue = boundaryValue;
            Record(std::move(op));
        }

        void HipPDEGrid2D::DiffusionEdges(double diffCoef, const DirichletBoundary &bc)
        {
            CheckDiffusionRate(diffCoef, "HipPDEGrid2D::Diffusion");
            Op op(Op::Kind::Diffusion);
            op.a = diffCoef;
            op.hasBc = true;
            op.bc = std::make_shared<const DirichletBoundary>(bc);
            Record(std::move(op));
        }

        void HipPDEGrid2D::Diffusion(const std::vector<double> &diffRatesX, const std::vector<double> &diffRatesY)
        {
            if (static_cast<int64_t>(diffRatesX.size()) != Length() || static_cast<int64_t>(diffRatesY.size()) != Length())
            {
                Throw(true, "HipPDEGrid2D::Diffusion: rates must hold xDim*yDim values");
            }
            Op op(Op::Kind::DiffusionRates);
            op.u = std::make_shared<const std::vector<double>>(diffRatesX);
            op.v = std::make_shared<const std::vector<double>>(diffRatesY);
            Record(std::move(op));
        }

        void HipPDEGrid2D::DiffusionADI(double diffRate)
        {
            Op op(Op::Kind::Adi);
            op.a = diffRate;
            Record(std::move(op));
        }

        void HipPDEGrid2D::DiffusionADI(double diffRate, double boundaryValue)
        {
            Op op(Op::Kind::Adi);
            op.a = diffRate;
            op.hasBc = true;
            op.constBc = true;
            op.bcValue = boundaryValue;
            Record(std::move(op));
        }

        void HipPDEGrid2D::DiffusionADIEdges(double diffRate, const DirichletBoundary &bc)
        {
            Op op(Op::Kind::Adi);
            op.a = diffRate;
            op.hasBc = true;
            op.bc = std::make_shared<const DirichletBoundary>(bc);
            Record(std::move(op));
        }

        void HipPDEGrid2D::Advection(double xVel, double yVel)
        {
            CheckAdvectionRate(xVel, yVel, "HipPDEGrid2D::Advection");
            Op op(Op::Kind::Advection);
            op.a = xVel;
            op.b = yVel;
            Record(std::move(op));
        }

        void HipPDEGrid2D::Advection(double xVel, double yVel, double boundaryValue)
        {
            CheckAdvectionRate(xVel, yVel, "HipPDEGrid2D::Advection");
            Op op(Op::Kind::Advection);
            op.a = xVel;
            op.b = yVel;
            op.hasBc = true;
            op.constBc = true;
            op.bcValue = boundaryValue;
            Record(std::move(op));
        }

        void HipPDEGrid2D::AdvectionEdges(double xVel, double yVel, const DirichletBoundary &bc)
        {
            CheckAdvectionRate(xVel, yVel, "HipPDEGrid2D::Advection");
            Op op(Op::Kind::Advection);
            op.a = xVel;
            op.b = yVel;
            op.hasBc = true;
            op.bc = std::make_shared<const DirichletBoundary>(bc);
            Record(std::move(op));
        }

        void HipPDEGrid2D::Advection(const std::vector<double> &xVels, const std::vector<double> &yVels)
        {
            if (static_cast<int64_t>(xVels.size()) != Length() || static_cast<int64_t>(yVels.size()) != Length())
            {
                Throw(true, "HipPDEGrid2D::Advection: velocities must hold xDim*yDim values");
            }
            Op op(Op::Kind::AdvectionCells);
            op.u = std::make_shared<const std::vector<double>>(xVels);
            op.v = std::make_shared<const std::vector<double>>(yVels);
            Record(std::move(op));
        }

        void HipPDEGrid2D::AdvectionCellsEdges(const std::vector<double> &xVels, const std::vector<double> &yVels,
                                               const DirichletBoundary &bc, const DirichletBoundary &velBc)
        {
            if (static_cast<int64_t>(xVels.size()) != Length() || static_cast<int64_t>(yVels.size()) != Length())
            {
                Throw(true, "HipPDEGrid2D::Advection: velocities must hold xDim*yDim values");
            }
            Op op(Op::Kind::AdvectionCells);
            op.u = std::make_shared<const std::vector<double>>(xVels);
            op.v = std::make_shared<const std::vector<double>>(yVels);
            op.hasBc = true;
            op.bc = std::make_shared<const DirichletBoundary>(bc);
            op.velBc = std::make_shared<const DirichletBoundary>(velBc);
            Record(std::move(op));
        }

        // The fused tick takes one uniform stencil (explicit or ADI), at most
        // one uniform advection (explicit ticks only), MulAll and Add, with
        // no boundary or one shared constant boundary value.
        bool HipPDEGrid2D::Fusable() const
        {
            int stencils = 0;
            int advections = 0;
            bool adi = false;
            bool bcSeen = false;
            bool hasBc = false;
            double bcValue = 0.0;
            for (const Op &op : m_ops)
            {
                switch (op.kind)
                {
                    case Op::Kind::Diffusion:
                    case Op::Kind::Adi:
                    case Op::Kind::Advection:
                        if (op.hasBc && !op.constBc)
                        {
                            return false;
                        }
                        if (bcSeen && (op.hasBc != hasBc || (op.hasBc && op.bcValue != bcValue)))
                        {
                            return false;
                        }
                        bcSeen = true;
                        hasBc = op.hasBc;
                        bcValue = op.bcValue;
                        if (op.kind == Op::Kind::Advection)
                        {
                            advections++;
                        }
                        else
                        {
                            stencils++;
                            adi = adi || op.kind == Op::Kind::Adi;
                        }
                        break;
                    case Op::Kind::Point:
                        if (op.code != HALHIP_DELTA_ADD)
                        {
                            return false;
                        }
                        break;
                    case Op::Kind::All:
                        if (op.code != HALHIP_DELTA_MUL)
                        {
                            return false;
                        }
                        break;
                    default:
                        return false;
                }
            }
            return stencils <= 1 && advections <= 1 && !(adi && advections > 0);
        }

        void HipPDEGrid2D::RunFused()
        {
            const char *caller = "HipPDEGrid2D::Update";
            Stencil stencil = Stencil::None;
            double rate = 0.0;
            double vx = 0.0;
            double vy = 0.0;
            bool hasBc = false;
            double bcValue = 0.0;
            double mulAll = 0.0;
            std::vector<std::pair<int64_t, float>> adds;
            for (const Op &op : m_ops)
            {
                switch (op.kind)
                {
                    case Op::Kind::Diffusion:
                    case Op::Kind::Adi:
                        stencil = (op.kind == Op::Kind::Adi) ? Stencil::Adi : Stencil::Explicit;
                        rate = op.a;
                        hasBc = hasBc || op.hasBc;
                        bcValue = op.hasBc ? op.bcValue : bcValue;
                        break;
                    case Op::Kind::Advection:
                        vx = op.a;
                        vy = op.b;
                        hasBc = hasBc || op.hasBc;
                        bcValue = op.hasBc ? op.bcValue : bcValue;
                        break;
                    case Op::Kind::All:
                        mulAll += op.a;
                        break;
                    default:
                        adds.emplace_back(op.index, static_cast<float>(op.a));
                        break;
                }
            }
            halhip_bc bc{};
            bc.kind = HALHIP_BC_DIRICHLET;
            bc.value = bcValue;
            const halhip_bc *pbc = hasBc ? &bc : nullptr;
            std::vector<int64_t> idx;
            std::vector<float> val;
            for (const auto &a : adds)
            {
                idx.push_back(a.first);
                val.push_back(a.second);
            }
            // MulAll(v) adds v*u to the deltas: the tick's decay is -sum(v).
            const double decay = -mulAll;
            if (stencil == Stencil::Adi)
            {
                MoveFieldTo(Stencil::Adi);
                if (adds != m_adiSources)
                {
                    ThrowIfFailed(halhip_adi2d_set_sources(m_adi, static_cast<int32_t>(idx.size()), idx.data(), val.data()),
                                  caller);
                    m_adiSources = adds;
                }
                ThrowIfFailed(halhip_adi2d_run_steps_async(m_adi, 1, rate, decay, m_wrapX, m_wrapY, pbc), caller);
                return;
            }
            // explicit diffusion (rate 0 when the tick has none)
            MoveFieldTo(Stencil::Explicit);
            if (adds != m_explicitSources)
            {
                ThrowIfFailed(halhip_diffusion2d_set_sources(m_explicit, static_cast<int32_t>(idx.size()), idx.data(),
                                                             val.data()),
                              caller);
                m_explicitSources = adds;
            }
            if (vx != m_explicitVel[0] || vy != m_explicitVel[1])
            {
                ThrowIfFailed(halhip_diffusion2d_set_advection(m_explicit, vx, vy), caller);
                m_explicitVel[0] = vx;
                m_explicitVel[1] = vy;
            }
            const double r = (stencil == Stencil::Explicit) ? rate : 0.0;
            ThrowIfFailed(halhip_diffusion2d_run_steps_async(m_explicit, 1, r, decay, m_wrapX, m_wrapY, pbc), caller);
        }

        // Switches the pending tick to device deltas on the explicit plan:
        // the committed field moves there and the recorded calls run in order.
        void HipPDEGrid2D::Materialize()
        {
            if (m_general)
            {
                return;
            }
            MoveFieldTo(Stencil::Explicit);
            m_general = true;
            std::vector<Op> ops;
            ops.swap(m_ops);
            for (const Op &op : ops)
            {
                Issue(op);
            }
        }

        void HipPDEGrid2D::FlushPoints()
        {
            if (m_pointOps.empty())
            {
                return;
            }
            ThrowIfFailed(halhip_diffusion2d_delta_point_ops(m_explicit, static_cast<int32_t>(m_pointOps.size()),
                                                             m_pointOps.data(), m_pointIdx.data(), m_pointVal.data()),
                          "HipPDEGrid2D");
            m_pointOps.clear();
            m_pointIdx.clear();
            m_pointVal.clear();
        }

        void HipPDEGrid2D::Issue(const Op &op)
        {
            const char *caller = "HipPDEGrid2D";
            if (op.kind == Op::Kind::Point)
            {
                m_pointOps.push_back(op.code);
                m_pointIdx.push_back(op.index);
                m_pointVal.push_back(op.a);
                return;
            }
            FlushPoints();
            const BcView bc = ViewOf(op.hasBc, op.constBc, op.bcValue, op.bc.get());
            const int64_t n = Length();
            switch (op.kind)
            {
                case Op::Kind::Diffusion:
                    ThrowIfFailed(halhip_diffusion2d_delta_diffuse(m_explicit, op.a, m_wrapX, m_wrapY, bc.Ptr()), caller);
                    break;
                case Op::Kind::DiffusionRates:
                    ThrowIfFailed(halhip_diffusion2d_delta_diffuse_rates(m_explicit, op.u->data(), op.v->data(), n,
                                                                         m_wrapX, m_wrapY),
                                  caller);
                    break;
                case Op::Kind::Adi:
                    ThrowIfFailed(halhip_diffusion2d_delta_adi(m_explicit, Adi(), op.a, m_wrapX, m_wrapY, bc.Ptr()), caller);
                    break;
                case Op::Kind::Advection:
                    ThrowIfFailed(halhip_diffusion2d_delta_advect(m_explicit, op.a, op.b, m_wrapX, m_wrapY, bc.Ptr()),
                                  caller);
                    break;
                case Op::Kind::AdvectionCells:
                {
                    const BcView vel = ViewOf(op.velBc != nullptr, false, 0.0, op.velBc.get());
                    ThrowIfFailed(halhip_diffusion2d_delta_advect_cells(m_explicit, op.u->data(), op.v->data(), n,
                                                                        m_wrapX, m_wrapY, bc.Ptr(), vel.Ptr()),
                                  caller);
                    break;
                }
                case Op::Kind::All:
                    ThrowIfFailed(halhip_diffusion2d_delta_all(m_explicit, op.code, op.a), caller);
                    break;
                case Op::Kind::SetValues:
                    ThrowIfFailed(halhip_diffusion2d_delta_set_values(m_explicit, op.u->data(), n), caller);
                    break;
                default:
                    break;
            }
        }

        void HipPDEGrid2D::Update()
        {
            const char *caller = "HipPDEGrid2D::Update";
            if (!m_general && Fusable())
            {
                if (!m_ops.empty())
                {
                    RunFused();
                }
            }
            else
            {
                Materialize();
                FlushPoints();
                ThrowIfFailed(halhip_diffusion2d_delta_update(m_explicit), caller);
            }
            ClearTick();
            m_updateCt++;
        }

        void HipPDEGrid2D::Save(const std::string &path)
        {
            const char *caller = "HipPDEGrid2D::Save";
            if (!m_ops.empty() || m_general)
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
            if (!m_ops.empty() || m_general)
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
