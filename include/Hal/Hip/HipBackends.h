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
        // every call adds to the pending tick's deltas of the committed field
        // and Update commits, like PDEGrid2D::Update (:178-186).  Calls are
        // recorded; a tick of one uniform-rate diffusion (explicit or ADI,
        // optionally with a constant boundary value), uniform advection
        // (explicit ticks), MulAll and Add runs on Update as one fused device
        // tick.  Any other tick -- per-cell rates or velocities, boundary
        // callables, Set / Mul / Scale / SetAll / AddAll, SetNonNegative, or
        // several diffusion calls -- and any query of the pending deltas
        // (MaxDelta) runs its calls in order on device-resident deltas.
        // Ticks are queued asynchronously; GetField, Get and Save wait for
        // them.  Get(i) downloads the whole field.  One thread per grid
        // object.
        class HipPDEGrid2D
        {
        public:
            HipPDEGrid2D(HipContext &context, int32_t xDim, int32_t yDim, bool wrapX = false, bool wrapY = false);
            ~HipPDEGrid2D();

            HipPDEGrid2D(const HipPDEGrid2D &other) = delete;
            HipPDEGrid2D &operator=(const HipPDEGrid2D &other) = delete;

            int32_t Xdim() const { return m_xDim; }
            int32_t Ydim() const { return m_yDim; }
            int32_t Length() const { return m_xDim * m_yDim; }
            bool IsWrapX() const { return m_wrapX; }
            bool IsWrapY() const { return m_wrapY; }
            int32_t UpdateCt() const { return m_updateCt; }

            // Committed field (PDEGrid2D::SetState with zero deltas / GetField);
            // SetField drops the pending tick.
            void SetField(const std::vector<double> &field);
            std::vector<double> GetField();
            double Get(int32_t i);
            double Get(int32_t x, int32_t y) { return Get(x * m_yDim + y); }
            // PDEGrid2D::GetMax / GetMin / GetAvg of the committed field (:133-162)
            double GetMax();
            double GetMin();
            double GetAvg();

            // PDEGrid2D::Set / Add / Mul / Scale (include/Hal/Grids/PDEGrid2D.h:89-145)
            void Set(int32_t i, double val);
            void Set(int32_t x, int32_t y, double val) { Set(x * m_yDim + y, val); }
            void Add(int32_t i, double val);
            void Add(int32_t x, int32_t y, double val) { Add(x * m_yDim + y, val); }
            void Mul(int32_t i, double val);
            void Mul(int32_t x, int32_t y, double val) { Mul(x * m_yDim + y, val); }
            void Scale(int32_t i, double val);
            void Scale(int32_t x, int32_t y, double val) { Scale(x * m_yDim + y, val); }
            // PDEGrid2D::SetAll / AddAll / MulAll (:177-218); SetAll(pVals) reads Length() values
            void SetAll(double val);
            void SetAll(const double *pVals);
            void AddAll(double val);
            void MulAll(double val);
            // PDEGrid2D::MaxDelta / MaxDeltaScaled / SetNonNegative (:220-259)
            double MaxDelta();
            double MaxDeltaScaled(double denomOffset);
            void SetNonNegative();

            // PDEGrid2D::Diffusion (include/Hal/Grids/PDEGrid2D.h:289-327; uniform rate <= 0.25)
            void Diffusion(double diffCoef);
            void Diffusion(double diffCoef, double boundaryValue);
            template <typename BCFn>
            void Diffusion(double diffCoef, BCFn bc)
            {
                DiffusionEdges(diffCoef, EdgesOf(bc));
            }
            void Diffusion(const std::vector<double> &diffRatesX, const std::vector<double> &diffRatesY);
            // PDEGrid2D::DiffusionADI (include/Hal/Grids/PDEGrid2D.h:328-347)
            void DiffusionADI(double diffRate);
            void DiffusionADI(double diffRate, double boundaryValue);
            template <typename BCFn>
            void DiffusionADI(double diffRate, BCFn bc)
            {
                DiffusionADIEdges(diffRate, EdgesOf(bc));
            }
            // PDEGrid2D::Advection (include/Hal/Grids/PDEGrid2D.h:349-399; uniform |xVel|+|yVel| <= 0.5)
            void Advection(double xVel, double yVel);
            void Advection(double xVel, double yVel, double boundaryValue);
            template <typename BCFn>
            void Advection(double xVel, double yVel, BCFn bc)
            {
                AdvectionEdges(xVel, yVel, EdgesOf(bc));
            }
            void Advection(const std::vector<double> &xVels, const std::vector<double> &yVels);
            // Off-grid field values from bc; x velocities at x = -1 / xDim from
            // xVelBC, y velocities at y = -1 / yDim from yVelBC.
            template <typename BCFn, typename XVelBCFn, typename YVelBCFn>
            void Advection(const std::vector<double> &xVels, const std::vector<double> &yVels, BCFn bc,
                           XVelBCFn xVelBC, YVelBCFn yVelBC)
            {
                DirichletBoundary vel;
                vel.xlo.resize(static_cast<size_t>(m_yDim));
                vel.xhi.resize(static_cast<size_t>(m_yDim));
                vel.ylo.resize(static_cast<size_t>(m_xDim));
                vel.yhi.resize(static_cast<size_t>(m_xDim));
                for (int32_t y = 0; y < m_yDim; y++)
                {
                    vel.xlo[static_cast<size_t>(y)] = xVelBC(-1, y);
                    vel.xhi[static_cast<size_t>(y)] = xVelBC(m_xDim, y);
                }
                for (int32_t x = 0; x < m_xDim; x++)
                {
                    vel.ylo[static_cast<size_t>(x)] = yVelBC(x, -1);
                    vel.yhi[static_cast<size_t>(x)] = yVelBC(x, m_yDim);
                }
                AdvectionCellsEdges(xVels, yVels, EdgesOf(bc), vel);
            }
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
            // One recorded call of the pending tick.
            struct Op
            {
                enum class Kind
                {
                    Diffusion,
                    DiffusionRates,
                    Adi,
                    Advection,
                    AdvectionCells,
                    Point,
                    All,
                    SetValues
                };
                explicit Op(Kind k) : kind(k) {}
                Kind kind;
                double a = 0.0; // rate, x velocity or value
                double b = 0.0; // y velocity
                int32_t code = 0; // point / field-wide op (HALHIP_DELTA_*)
                int64_t index = 0;
                bool hasBc = false;
                bool constBc = false; // a constant boundary value (bcValue) rather than edge arrays
                double bcValue = 0.0;
                std::shared_ptr<const DirichletBoundary> bc;    // edge arrays (constBc false)
                std::shared_ptr<const DirichletBoundary> velBc; // per-cell advection velocity edges
                std::shared_ptr<const std::vector<double>> u;  // rates / velocities / values
                std::shared_ptr<const std::vector<double>> v;
            };

            // Ghost-cell values of a boundary callable bc(x, y).
            template <typename BCFn>
            DirichletBoundary EdgesOf(BCFn bc) const
            {
                DirichletBoundary e;
                e.xlo.resize(static_cast<size_t>(m_yDim));
                e.xhi.resize(static_cast<size_t>(m_yDim));
                e.ylo.resize(static_cast<size_t>(m_xDim));
                e.yhi.resize(static_cast<size_t>(m_xDim));
                for (int32_t y = 0; y < m_yDim; y++)
                {
                    e.xlo[static_cast<size_t>(y)] = bc(-1, y);
                    e.xhi[static_cast<size_t>(y)] = bc(m_xDim, y);
                }
                for (int32_t x = 0; x < m_xDim; x++)
                {
                    e.ylo[static_cast<size_t>(x)] = bc(x, -1);
                    e.yhi[static_cast<size_t>(x)] = bc(x, m_yDim);
                }
                return e;
            }
            void DiffusionEdges(double diffCoef, const DirichletBoundary &bc);
            void DiffusionADIEdges(double diffRate, const DirichletBoundary &bc);
            void AdvectionEdges(double xVel, double yVel, const DirichletBoundary &bc);
            void AdvectionCellsEdges(const std::vector<double> &xVels, const std::vector<double> &yVels,
                                     const DirichletBoundary &bc, const DirichletBoundary &velBc);

            void Record(Op op);
            void CheckIndex(int32_t i, const char *caller) const;
            bool Fusable() const;
            void RunFused();
            void Materialize();
            void Issue(const Op &op);
            void FlushPoints();
            void ClearTick();
            double Reduce(int32_t what, double offset, const char *caller);
            void MoveFieldTo(Stencil s);
            halhip_adi2d *Adi();

            HipContext &m_context;
            int32_t m_xDim;
            int32_t m_yDim;
            bool m_wrapX;
            bool m_wrapY;
            int32_t m_updateCt;
            halhip_diffusion2d *m_explicit;
            halhip_adi2d *m_adi;
            Stencil m_resident; // plan holding the committed field
            // the pending tick: recorded calls, or (m_general) calls already
            // applied to the explicit plan's device deltas, with point ops
            // waiting in m_points to go in one batch
            std::vector<Op> m_ops;
            bool m_general;
            std::vector<int32_t> m_pointOps;
            std::vector<int64_t> m_pointIdx;
            std::vector<double> m_pointVal;
            // sources and velocity last handed to each plan's fused tick
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
