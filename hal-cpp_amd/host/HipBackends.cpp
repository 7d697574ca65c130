// HipBackends.cpp -- Hal::Hip adapters of HAL-CPP's strategy interfaces over
// the C ABI, plus the interface base classes, CheckFieldSizes and Logger.
// Host code only: compiled with g++ and linked against libhalhip.so.
#include "Hal/Hip/HipBackends.h"

#include "Hal/Log/Logger.h"

#include <iostream>
#include <stdexcept>
#include <string>

namespace Hal
{
    // ---- Logger (src/Log/Logger.cpp:16-37 behaviour) ------------------------------------

    Logger::Logger() : m_minSeverity(LogSeverity::Info) {}

    Logger::~Logger() {}

    Logger &Logger::Instance()
    {
        static Logger instance;
        return instance;
    }

    void Logger::SetMinSeverity(LogSeverity severity)
    {
        m_minSeverity = severity;
    }

    void Logger::Log(LogSeverity severity, const std::string &message)
    {
        if (severity < m_minSeverity)
        {
            return;
        }
        std::ostream &out = (severity == LogSeverity::Error) ? std::cerr : std::cout;
        out << "[" << ToString(severity) << "] " << message << std::endl;
    }

    const char *ToString(LogSeverity severity)
    {
        switch (severity)
        {
            case LogSeverity::Debug:
                return "DEBUG";
            case LogSeverity::Info:
                return "INFO";
            case LogSeverity::Warning:
                return "WARNING";
            case LogSeverity::Error:
                return "ERROR";
        }
        return "UNKNOWN";
    }

    // ---- interface bases -------------------------------------------------------------

    IDiffusionBackend2D::IDiffusionBackend2D() {}
    IDiffusionBackend2D::~IDiffusionBackend2D() {}
    IAdiDiffusionBackend2D::IAdiDiffusionBackend2D() {}
    IAdiDiffusionBackend2D::~IAdiDiffusionBackend2D() {}
    IAgentForceBackend2D::IAgentForceBackend2D() {}
    IAgentForceBackend2D::~IAgentForceBackend2D() {}

    namespace Internal
    {
        void CheckFieldSizes(const std::vector<double> &field,
                             const std::vector<double> &deltas,
                             int32_t                    xDim,
                             int32_t                    yDim,
                             const char                *pCaller)
        {
            const size_t n = static_cast<size_t>(xDim) * static_cast<size_t>(yDim);
            if (xDim <= 0 || yDim <= 0 || field.size() != n || deltas.size() != n)
            {
                const std::string message = std::string(pCaller) + ": field and deltas must each hold xDim*yDim values";
                Logger::Instance().Log(LogSeverity::Error, message);
                throw std::invalid_argument(message);
            }
        }
    }

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
