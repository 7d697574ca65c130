#pragma once

// Strategy interface for cell-bucket agent force accumulation, with the
// signature and contract of HAL-CPP's Hal::IAgentForceBackend2D
// (include/Hal/Agents/IAgentForceBackend2D.h:12-37): per agent, the linear
// spring springConstant * ((rad_i + rad_j) - dist) summed over every other
// agent closer than interactionRad (and not at the same position).

#include <cstdint>
#include <vector>

namespace Hal
{
    class IAgentForceBackend2D
    {
    public:
        IAgentForceBackend2D();
        virtual ~IAgentForceBackend2D();

        IAgentForceBackend2D(const IAgentForceBackend2D &other) = delete;
        IAgentForceBackend2D &operator=(const IAgentForceBackend2D &other) = delete;

        // positions: flat (x0, y0, x1, y1, ...); radii: one per agent.
        // velocityDeltas is resized to positions.size() and overwritten with
        // each agent's (vx, vy).  Throws std::invalid_argument unless
        // positions.size() == 2 * radii.size().
        virtual void AccumulateForces(const std::vector<double> &positions,
                                      const std::vector<double> &radii,
                                      std::vector<double>       &velocityDeltas,
                                      double                     interactionRad,
                                      double                     springConstant,
                                      int32_t                    xDim,
                                      int32_t                    yDim,
                                      bool                       wrapX,
                                      bool                       wrapY) = 0;
    };
}
