// Error vocabulary: recoverable conditions are returned as Status, programmer
// errors are thrown as Error.  Mirrors include/sovereign/status.hpp:25-69 of
// the reference; the enum values are also the C-ABI return codes of
// include/sov_hip.h (Ok = 0 ... NumericalFailure = 7).
#pragma once

#include <stdexcept>
#include <string>

namespace sov {

enum class Status {
    Ok = 0,
    Unsupported,
    NotImplemented,
    InvalidArgument,
    DimensionMismatch,
    OutOfMemory,
    BackendError,
    NumericalFailure,
};

constexpr const char* to_string(Status s) noexcept {
    switch (s) {
        case Status::Ok: return "ok";
        case Status::Unsupported: return "unsupported by this backend";
        case Status::NotImplemented: return "not implemented yet";
        case Status::InvalidArgument: return "invalid argument";
        case Status::DimensionMismatch: return "dimension mismatch";
        case Status::OutOfMemory: return "out of memory";
        case Status::BackendError: return "backend error";
        case Status::NumericalFailure: return "numerical failure";
    }
    return "unknown status";
}

constexpr bool ok(Status s) noexcept { return s == Status::Ok; }

class Error : public std::runtime_error {
public:
    explicit Error(std::string what) : std::runtime_error(std::move(what)) {}
    Error(Status s, const std::string& context)
        : std::runtime_error(context + ": " + to_string(s)), status_(s) {}
    Status status() const noexcept { return status_; }

private:
    Status status_ = Status::BackendError;
};

inline void check(Status s, const char* context) {
    if (s != Status::Ok) throw Error(s, context);
}

}  // namespace sov
