#pragma once

// Severity logger with the interface of HAL-CPP's Hal::Logger
// (include/Hal/Log/Logger.h, src/Log/Logger.cpp:16-37): messages at or above
// the minimum severity go to stdout, errors to stderr, as "[SEVERITY] text".
// The HIP backends log every error before throwing, as HAL does.

#include <cstdint>
#include <string>

namespace Hal
{
    enum class LogSeverity : uint32_t
    {
        Debug,
        Info,
        Warning,
        Error
    };

    class Logger
    {
    public:
        Logger();
        ~Logger();

        static Logger &Instance();

        void Log(LogSeverity severity, const std::string &message);
        void SetMinSeverity(LogSeverity severity);

    private:
        LogSeverity m_minSeverity;
    };

    const char *ToString(LogSeverity severity);
}
