#pragma once
// logger.hpp — level-filtered, mutex-guarded logger (reference logger.hpp:17-69).
#include <iostream>
#include <mutex>
#include <sstream>
#include <string>

namespace hsps {

enum class LogLevel { DEBUG, INFO, WARN, ERROR };

class Logger {
public:
    static Logger& instance() {
        static Logger l;
        return l;
    }
    void set_level(LogLevel l) { level_ = l; }
    void set_stream(std::ostream& os) { out_ = &os; }
    template <typename... Args>
    void log(LogLevel l, Args&&... args) {
        if (l < level_) return;
        std::ostringstream ss;
        (ss << ... << args);
        std::lock_guard<std::mutex> lock(mtx_);
        *out_ << "[" << name(l) << "] " << ss.str() << "\n";
    }

private:
    Logger() = default;
    static const char* name(LogLevel l) {
        switch (l) {
            case LogLevel::DEBUG: return "DEBUG";
            case LogLevel::INFO: return " INFO";
            case LogLevel::WARN: return " WARN";
            case LogLevel::ERROR: return "ERROR";
        }
        return "?????";
    }
    LogLevel level_ = LogLevel::WARN;
    std::ostream* out_ = &std::cerr;
    std::mutex mtx_;
};

}  // namespace hsps

#define HSPS_LOG_DEBUG(...) ::hsps::Logger::instance().log(::hsps::LogLevel::DEBUG, __VA_ARGS__)
#define HSPS_LOG_INFO(...) ::hsps::Logger::instance().log(::hsps::LogLevel::INFO, __VA_ARGS__)
#define HSPS_LOG_WARN(...) ::hsps::Logger::instance().log(::hsps::LogLevel::WARN, __VA_ARGS__)
#define HSPS_LOG_ERROR(...) ::hsps::Logger::instance().log(::hsps::LogLevel::ERROR, __VA_ARGS__)
