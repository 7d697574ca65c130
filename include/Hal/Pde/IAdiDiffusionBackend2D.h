#pragma once

// Strategy interface for one implicit (ADI) 2D diffusion step, with the
// signature and contract of HAL-CPP's Hal::IAdiDiffusionBackend2D
// (include/Hal/Pde/IAdiDiffusionBackend2D.h:12-32): zero-flux or periodic
// boundaries, same flat layout as IDiffusionBackend2D.

#include <cstdint>
#include <vector>

namespace Hal
{
    class IAdiDiffusionBackend2D
    {
    public:
        IAdiDiffusionBackend2D();
        virtual ~IAdiDiffusionBackend2D();

        IAdiDiffusionBackend2D(const IAdiDiffusionBackend2D &other) = delete;
        IAdiDiffusionBackend2D &operator=(const IAdiDiffusionBackend2D &other) = delete;

        // deltas[i] += ADI(field)[i] - field[i].  Unconditionally stable.
        // Throws std::invalid_argument unless field and deltas both hold
        // xDim*yDim values.
        virtual void DiffuseADI(const std::vector<double> &field,
                                std::vector<double>       &deltas,
                                double                     diffRate,
                                int32_t                    xDim,
                                int32_t                    yDim,
                                bool                       wrapX,
                                bool                       wrapY) = 0;
    };
}
