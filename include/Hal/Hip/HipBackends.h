#pragma once

// MI355X (gfx950) implementations of HAL-CPP's three strategy interfaces,
// built on the C ABI in halhip.h.  They replace the Vulkan adapters
// Hal::Gpu::GpuDiffusionBackend2D, GpuAdiDiffusionBackend2D and
// GpuAgentForceBackend2D (src/Gpu/Gpu*Backend2D.cpp) with the same contract:
// FP64 host arrays in and out, FP32 on the device, synchronous calls, errors
// logged through Hal::Logger and thrown as std::invalid_argument (misuse) or
// std::runtime_error (device failure).
//
// A backend constructed with dimensions is sized like the reference's
// (other dimensions throw std::invalid_argument); one constructed without
// dimensions re-plans on the first call with new dimensions.  One thread per
// backend object.

#include "Hal/Agents/IAgentForceBackend2D.h"
#include "Hal/Pde/IAdiDiffusionBackend2D.h"
#include "Hal/Pde/IDiffusionBackend2D.h"
#include "halhip.h"

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace Hal
{
    namespace Hip
    {
        // Logs and throws for a failed halhip status (no-op for HALHIP_OK).
        void ThrowIfFailed(int status, const char *caller);

        // Replaces Hal::Gpu::ComputeContext (include/Hal/Gpu/ComputeContext.h:18-49).
        class HipContext
        {
        public:
            explicit HipContext(int32_t device = 0);
            ~HipContext();

            HipContext(const HipContext &other) = delete;
            HipContext &operator=(const HipContext &other) = delete;

            halhip_context *Handle() const { return m_ctx; }

            // nullptr when no usable GPU exists, mirroring
            // TestHelpers::TryCreateComputeContext (tests/GpuTestSupport.cpp:9-20).
            static std::unique_ptr<HipContext> TryCreate(int32_t device = 0);

        private:
            halhip_context *m_ctx;
        };

        // Dirichlet edge values, as halhip_bc: a constant or four edge arrays
        // (bc(-1,y), bc(X,y), bc(x,-1), bc(x,Y)).
        struct DirichletBoundary
        {
            double value = 0.0;
            std::vector<double> xlo, xhi, ylo, yhi;
        };

        class HipDiffusionBackend2D : public IDiffusionBackend2D
        {
        public:
            explicit HipDiffusionBackend2D(HipContext &context);
            HipDiffusionBackend2D(HipContext &context, int32_t xDim, int32_t yDim);
            ~HipDiffusionBackend2D();

            void Diffuse(const std::vector<double> &field,
                         std::vector<double>       &deltas,
                         double                     diffRate,
                         int32_t                    xDim,
                         int32_t                    yDim,
                         bool                       wrapX,
                         bool                       wrapY) override;

            // PDEGrid2D::Diffusion(diffCoef, boundaryValue / BCFn) (include/Hal/Grids/PDEGrid2D.h:289-310).
            void Diffuse(const std::vector<double> &field,
                         std::vector<double>       &deltas,
                         double                     diffRate,
                         int32_t                    xDim,
                         int32_t                    yDim,
                         bool                       wrapX,
                         bool                       wrapY,
                         const DirichletBoundary   &bc);

            // Per-cell face rates (src/Pde/PDEequations.cpp:199-228).
            void DiffuseRates(const std::vector<double> &field,
                              std::vector<double>       &deltas,
                              const std::vector<double> &ratesX,
                              const std::vector<double> &ratesY,
                              int32_t                    xDim,
                              int32_t                    yDim,
                              bool                       wrapX,
                              bool                       wrapY);

            // PDEGrid2D::Advection(xVel, yVel[, boundary]) on the interface's
            // terms: deltas += first-order upwind advection
            // (src/Pde/PDEequations.cpp:380-413, include/Hal/Pde/PDEequations.h:413-447).
            void Advect(const std::vector<double> &field,
                        std::vector<double>       &deltas,
                        double                     xVel,
                        double                     yVel,
                        int32_t                    xDim,
                        int32_t                    yDim,
                        bool                       wrapX,
                        bool                       wrapY);
            void Advect(const std::vector<double> &field,
                        std::vector<double>       &deltas,
                        double                     xVel,
                        double                     yVel,
                        int32_t                    xDim,
                        int32_t                    yDim,
                        bool                       wrapX,
                        bool                       wrapY,
                        const DirichletBoundary   &bc);

            // PDEGrid2D::Advection(xVels, yVels): per-cell velocities in flux
            // form with zero-flux walls (src/Pde/PDEequations.cpp:535-573).
            void AdvectCells(const std::vector<double> &field,
                             std::vector<double>       &deltas,
                             const std::vector<double> &xVels,
                             const std::vector<double> &yVels,
                             int32_t                    xDim,
                             int32_t                    yDim,
                             bool                       wrapX,
                             bool                       wrapY);

            // Uniform advection added to every RunSteps tick (Diffusion +
            // Advection + MulAll + Update); 0, 0 turns it off.
            void SetAdvection(double xVel, double yVel);

            // DiffusionKernel2D::RunSteps (include/Hal/Gpu/DiffusionKernel2D.h:34-38):
            // `steps` full ticks on the device, one upload and one readback;
            // decay and point sources are fused into each tick.
            void RunSteps(std::vector<float> &field,
                          int32_t             steps,
                          float               diffRate,
                          bool                wrapX,
                          bool                wrapY,
                          float               decay = 0.0f);

        private:
            halhip_diffusion2d *Plan(int32_t xDim, int32_t yDim, const char *caller);

            HipContext &m_context;
            bool m_fixed;
            int32_t m_xDim;
            int32_t m_yDim;
            halhip_diffusion2d *m_plan;
        };

        class HipAdiDiffusionBackend2D : public IAdiDiffusionBackend2D
        {
        public:
            explicit HipAdiDiffusionBackend2D(HipContext &context);
            HipAdiDiffusionBackend2D(HipContext &context, int32_t xDim, int32_t yDim);
            ~HipAdiDiffusionBackend2D();

            void DiffuseADI(const std::vector<double> &field,
                            std::vector<double>       &deltas,
                            double                     diffRate,
                            int32_t                    xDim,
                            int32_t                    yDim,
                            bool                       wrapX,
                            bool                       wrapY) override;

            // PDEGrid2D::DiffusionADI(diffRate, boundaryValue / BCFn)
            // (include/Hal/Grids/PDEGrid2D.h:328-347).
            void DiffuseADI(const std::vector<double> &field,
                            std::vector<double>       &deltas,
                            double                     diffRate,
                            int32_t                    xDim,
                            int32_t                    yDim,
                            bool                       wrapX,
                            bool                       wrapY,
                            const DirichletBoundary   &bc);

            // AdiKernel2D::RunSteps (include/Hal/Gpu/AdiKernel2D.h:33-39), with
            // an optional constant Dirichlet value.
            void RunSteps(std::vector<float> &field,
                          int32_t             steps,
                          float               diffRate,
                          bool                wrapX,
                          bool                wrapY);
            void RunStepsDirichlet(std::vector<float> &field,
                                   int32_t             steps,
                                   float               diffRate,
                                   bool                wrapX,
                                   bool                wrapY,
                                   double              boundaryValue);

        private:
            halhip_adi2d *Plan(int32_t xDim, int32_t yDim, const char *caller);

            HipContext &m_context;
            bool m_fixed;
            int32_t m_xDim;
            int32_t m_yDim;
            halhip_adi2d *m_plan;
        };

        // A PDEGrid2D whose field stays on the GPU, for model code written
        // against Hal::PDEGrid2D's tick calls (include/Hal/Grids/PDEGrid2D.h):
        // Diffusion / DiffusionADI / Advection / MulAll / Add accumulate a
        // pending tick, and Update runs it as one fused device tick and
        // commits, like PDEGrid2D::Update (include/Hal/Grids/PDEGrid2D.h:178-186).
        // Ticks are queued asynchronously; GetField, Get and Save wait for them.
        // A tick holds at most one diffusion call (explicit or ADI); Advection
        // combines with explicit diffusion (or stands alone), and the
        // boundary value, when given, must be the same for every call of a
        // tick.  Other combinations throw std::invalid_argument.  Get(i)
        // downloads the whole field.  One thread per grid object.
        class HipPDEGrid2D
        {
        public:
            HipPDEGrid2D(HipContext &context, int32_t xDim, int32_t yDim, bool wrapX = false, bool wrapY = false);
            ~HipPDEGrid2D();

            HipPDEGrid2D(const HipPDEGrid2D &other) = delete;
            HipPDEGrid2D &operator=(const HipPDEGrid2D &other) = delete;

            int32_t Xdim() const { return m_xDim; }
            int32_t Ydim() const { return m_yDim; }
            bool IsWrapX() const { return m_wrapX; }
            bool IsWrapY() const { return m_wrapY; }
            int32_t UpdateCt() const { return m_updateCt; }

            // Committed field (PDEGrid2D::SetState with zero deltas / GetField).
            void SetField(const std::vector<double> &field);
            std::vector<double> GetField();
            double Get(int32_t i);
            double Get(int32_t x, int32_t y) { return Get(x * m_yDim + y); }

            // PDEGrid2D::Diffusion (include/Hal/Grids/PDEGrid2D.h:289-310; rate <= 0.25)
            void Diffusion(double diffCoef);
            void Diffusion(double diffCoef, double boundaryValue);
            // PDEGrid2D::DiffusionADI (include/Hal/Grids/PDEGrid2D.h:328-347)
            void DiffusionADI(double diffRate);
            void DiffusionADI(double diffRate, double boundaryValue);
            // PDEGrid2D::Advection (include/Hal/Grids/PDEGrid2D.h:349-375; |xVel|+|yVel| <= 0.5)
            void Advection(double xVel, double yVel);
            void Advection(double xVel, double yVel, double boundaryValue);
            // PDEGrid2D::MulAll / Add (include/Hal/Grids/PDEGrid2D.h:103-130, 212-218)
            void MulAll(double val);
            void Add(int32_t i, double val);
            void Add(int32_t x, int32_t y, double val) { Add(x * m_yDim + y, val); }
            // PDEGrid2D::Update: runs and commits the pending tick.
            void Update();

            // SavePDEGrid2D / LoadPDEGrid2D file layout (src/Grids/PDEGridSerialization.cpp:33-61).
            void Save(const std::string &path);
            void Load(const std::string &path);

        private:
            enum class Stencil
            {
                None,
                Explicit,
                Adi
            };
            void SetBoundary(bool has, double value, const char *caller);
            void SetStencil(Stencil s, double rate, const char *caller);
            void MoveFieldTo(Stencil s);

            HipContext &m_context;
            int32_t m_xDim;
            int32_t m_yDim;
            bool m_wrapX;
            bool m_wrapY;
            int32_t m_updateCt;
            halhip_diffusion2d *m_explicit;
            halhip_adi2d *m_adi;
            Stencil m_resident; // plan holding the committed field
            // pending tick
            Stencil m_stencil;
            double m_rate;
            bool m_advect;
            double m_xVel;
            double m_yVel;
            bool m_hasBc;
            double m_bc;
            bool m_bcSet;
            double m_mulAll;
            std::vector<std::pair<int64_t, float>> m_adds;
            // sources last handed to each plan
            std::vector<std::pair<int64_t, float>> m_explicitSources;
            std::vector<std::pair<int64_t, float>> m_adiSources;
            double m_explicitVel[2];
        };

        class HipAgentForceBackend2D : public IAgentForceBackend2D
        {
        public:
            HipAgentForceBackend2D(HipContext &context, int32_t xDim, int32_t yDim, int32_t maxAgents);
            ~HipAgentForceBackend2D();

            void AccumulateForces(const std::vector<double> &positions,
                                  const std::vector<double> &radii,
                                  std::vector<double>       &velocityDeltas,
                                  double                     interactionRad,
                                  double                     springConstant,
                                  int32_t                    xDim,
                                  int32_t                    yDim,
                                  bool                       wrapX,
                                  bool                       wrapY) override;

            // Resident ticks (bin -> scan -> stable scatter -> force -> move
            // with MoveSafePT rules, include/Hal/Agents/AgentPT2D.h:157-186).
            void RunTicks(std::vector<float>       &positions,
                          const std::vector<float> &radii,
                          int32_t                   ticks,
                          float                     interactionRad,
                          float                     springConstant,
                          float                     dt,
                          bool                      wrapX,
                          bool                      wrapY);

        private:
            int32_t m_xDim;
            int32_t m_yDim;
            int32_t m_maxAgents;
            halhip_agents2d *m_plan;
        };
    }
}
