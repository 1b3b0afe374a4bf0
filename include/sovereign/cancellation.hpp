// Cooperative cancellation flag (reference include/sovereign/cancellation.hpp:23-30).
// The device loop polls it between launch intervals.
#pragma once

#include <atomic>

namespace sov {

class CancellationToken {
public:
    void cancel() noexcept { flag_.store(true, std::memory_order_relaxed); }
    bool is_cancelled() const noexcept { return flag_.load(std::memory_order_relaxed); }

private:
    std::atomic<bool> flag_{false};
};

}  // namespace sov
