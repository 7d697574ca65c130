#pragma once

// Strategy interface for one explicit 2D diffusion step.  Declared with the
// same signature and contract as HAL-CPP's Hal::IDiffusionBackend2D
// (include/Hal/Pde/IDiffusionBackend2D.h:13-33) so that model code written
// against it runs unchanged on Hal::Hip::HipDiffusionBackend2D.
//
// Fields are flat, caller-owned arrays indexed x*yDim + y.

#include <cstdint>
#include <vector>

namespace Hal
{
    class IDiffusionBackend2D
    {
    public:
        IDiffusionBackend2D();
        virtual ~IDiffusionBackend2D();

        IDiffusionBackend2D(const IDiffusionBackend2D &other) = delete;
        IDiffusionBackend2D &operator=(const IDiffusionBackend2D &other) = delete;

        // deltas[i] += this step's change of field[i] (accumulates; the caller
        // clears deltas if needed).  Throws std::invalid_argument unless field
        // and deltas both hold xDim*yDim values.
        virtual void Diffuse(const std::vector<double> &field,
                             std::vector<double>       &deltas,
                             double                     diffRate,
                             int32_t                    xDim,
                             int32_t                    yDim,
                             bool                       wrapX,
                             bool                       wrapY) = 0;
    };

    namespace Internal
    {
        // Logs and throws std::invalid_argument unless field and deltas both
        // hold xDim*yDim values (src/Pde/IDiffusionBackend2D.cpp:20-33).
        void CheckFieldSizes(const std::vector<double> &field,
                             const std::vector<double> &deltas,
                             int32_t                    xDim,
                             int32_t                    yDim,
                             const char                *pCaller);
    }
}
