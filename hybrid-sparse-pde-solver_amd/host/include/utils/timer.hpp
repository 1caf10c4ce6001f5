#pragma once
// timer.hpp — wall-clock Timer / ScopedTimer (reference timer.hpp:13-63).
#include <chrono>
#include <iostream>
#include <string>

namespace hsps {

class Timer {
public:
    using Clock = std::chrono::high_resolution_clock;
    using Duration = std::chrono::duration<double>;
    Timer() { reset(); }
    void start() { start_ = Clock::now(); running_ = true; }
    void stop() { running_ = false; accum_ += Clock::now() - start_; }
    void reset() { running_ = false; accum_ = Duration::zero(); }
    double elapsed() const { return running_ ? (accum_ + (Clock::now() - start_)).count() : accum_.count(); }
    void restart() { reset(); start(); }

private:
    Clock::time_point start_;
    Duration accum_ = Duration::zero();
    bool running_ = false;
};

class ScopedTimer {
public:
    explicit ScopedTimer(const std::string& label, std::ostream& os = std::cout) : label_(label), os_(os) { t_.start(); }
    ~ScopedTimer() {
        t_.stop();
        os_ << "[Timer] " << label_ << " : " << t_.elapsed() << " s\n";
    }

private:
    std::string label_;
    std::ostream& os_;
    Timer t_;
};

}  // namespace hsps
