// Timer.hpp -- wall-clock phase timer with the interface of the reference's
// src/Timer.hpp: reset() restarts it, elapsed() returns the seconds since.
#pragma once

#include <chrono>

class Timer {
   public:
    Timer() { reset(); }
    void reset() { t0_ = std::chrono::steady_clock::now(); }
    double elapsed() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

   private:
    std::chrono::steady_clock::time_point t0_;
};
