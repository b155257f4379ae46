// timer.h - accumulating wall-clock timer of the drop-in API (lib/timer.h).
#ifndef TIMER_H
#define TIMER_H

#include <chrono>

class Timer {
public:
    void start() {
        running_ = true;
        t0_ = std::chrono::high_resolution_clock::now();
    }
    void stop() {
        if (!running_) return;
        running_ = false;
        total_ += std::chrono::high_resolution_clock::now() - t0_;
    }
    void reset() { total_ = std::chrono::duration<double>(0); }
    unsigned long long get_time_milli_sec() const {
        return (unsigned long long)(get_time_sec() * 1000.0);
    }
    // while running: time since start(); otherwise the accumulated total
    double get_time_sec() const {
        if (running_) return std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0_).count();
        return total_.count();
    }
    double operator()() const { return get_time_sec(); }

private:
    bool running_ = false;
    std::chrono::high_resolution_clock::time_point t0_{};
    std::chrono::duration<double> total_{0};
};

#endif
