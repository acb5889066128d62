// Small host utilities used by the plugin classes and drivers
// (counterparts of Common/Util.h: ensure, Timer).
#pragma once

#include <chrono>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>

// The reference's ensure() prints and exits (Common/Util.h:13-15); a library
// must not exit its host process, so this throws std::logic_error instead.
#define ensure(expression)                                                                              \
    do {                                                                                                \
        if (!(expression))                                                                              \
            throw std::logic_error(std::string("ensure failed: ") + __FILE__ + ":" +                   \
                                   std::to_string(__LINE__) + " " #expression);                        \
    } while (0)

class Timer {
public:
    Timer() : start(std::chrono::steady_clock::now()) {}
    uint64_t Elapsed() const {
        return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - start)
            .count();
    }
    double Seconds() const { return Elapsed() / 1e9; }

private:
    std::chrono::steady_clock::time_point start;
};

// Throws std::runtime_error with ps_last_error() when a C-ABI call fails, as
// the reference's ERR() does for CUDA errors (Puzzles/GPU.cu:5-10).
void PsCheck(int status, const char* what);

// Initialises the library on device $PS_DEVICE (default 0) once per process;
// the GPU plugins and drivers call it before their first device operation.
void EnsurePsRuntime();
