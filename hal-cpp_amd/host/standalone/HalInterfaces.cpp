// HalInterfaces.cpp -- the few HAL-CPP definitions the Hal::Hip adapters
// need, for a build of the adapters WITHOUT HAL (this repository's own
// libhal_hip_backends.so and its tests).  Inside HAL, HAL's own
// src/Log/Logger.cpp, src/Pde/IDiffusionBackend2D.cpp,
// src/Pde/IAdiDiffusionBackend2D.cpp and src/Agents/IAgentForceBackend2D.cpp
// define these symbols and this file is not compiled.
//
// Behaviour kept identical to HAL's because callers observe it:
//   - Logger: "[SEVERITY] text" lines, errors to stderr, the rest to stdout,
//     below the minimum severity dropped (src/Log/Logger.cpp:16-37);
//   - CheckFieldSizes: logs, then throws std::invalid_argument unless both
//     arrays hold xDim*yDim values (src/Pde/IDiffusionBackend2D.cpp:20-33).
#include "Hal/Agents/IAgentForceBackend2D.h"
#include "Hal/Log/Logger.h"
#include "Hal/Pde/IAdiDiffusionBackend2D.h"
#include "Hal/Pde/IDiffusionBackend2D.h"

#include <iostream>
#include <stdexcept>
#include <string>

namespace Hal
{
    // ---- Logger ----------------------------------------------------------------------

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

}
