// host_util.h -- host-side helpers shared by the C entry points: a small
// worker pool for independent chunks of host work, and the guard that keeps
// C++ exceptions from crossing the C ABI.
#pragma once
#include <algorithm>
#include <exception>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/neuro_hip.h"

namespace nh {

int set_error(int code, const std::string &msg);

// Run fn on up to min(16, hardware threads, tasks) threads (the caller's
// included); fn pulls its own work items.  Falls back to the calling thread
// alone when threads cannot be started.
template <class F> void run_workers(F &fn, size_t tasks) {
    const size_t hw = std::max<unsigned>(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(16, hw), std::max<size_t>(tasks, 1));
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; ++t) {
        try {
            pool.emplace_back([&fn] { fn(); });
        } catch (...) {
            break;
        }
    }
    fn();
    for (auto &th : pool) th.join();
}

// Body of a C entry point with host exceptions turned into status codes
// (none may cross the C ABI): out of host memory -> NH_ERR_CAPACITY.
template <class F> int guarded(F &&fn) noexcept {
    try {
        return fn();
    } catch (const std::bad_alloc &) {
        return set_error(NH_ERR_CAPACITY, "out of host memory");
    } catch (const std::exception &e) {
        return set_error(NH_ERR_HIP, std::string("host error: ") + e.what());
    } catch (...) {
        return set_error(NH_ERR_HIP, "host error");
    }
}

}  // namespace nh
