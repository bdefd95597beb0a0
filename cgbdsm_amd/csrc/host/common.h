// Shared host utilities: limits, a persistent thread pool with dynamic
// chunking (stands in for the reference's TBB task arena, main_calig.cpp:57-62),
// and small helpers.
#pragma once
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "cgbdsm.h"

namespace cg {

constexpr uint32_t kMaxQV = CGBDSM_MAX_QV;
constexpr uint32_t kMaxQE = CGBDSM_MAX_QE;
constexpr uint32_t kMaxDir = CGBDSM_MAX_DIR;
constexpr uint8_t kNoEdge = CGBDSM_NO_EDGE;

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Default thread count of the reference: hardware_concurrency / NUMA nodes
// (index/calig.cpp:1092-1095).  NUMA nodes are read from sysfs.
int DefaultThreads();
int NumaNodes();

class ThreadPool {
public:
    explicit ThreadPool(int n);
    ~ThreadPool();
    int size() const { return n_; }
    // fn(begin, end, tid) over [0, n) in chunks of `grain`; the caller is tid 0.
    void For(size_t n, size_t grain, const std::function<void(size_t, size_t, int)>& fn);
    // fn(i, tid) for each i.
    template <class F>
    void Each(size_t n, size_t grain, F&& f) {
        For(n, grain, [&](size_t b, size_t e, int tid) {
            for (size_t i = b; i < e; i++) f(i, tid);
        });
    }

private:
    void Worker(int tid);
    int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    uint64_t gen_ = 0;
    bool stop_ = false;
    int running_ = 0;
    const std::function<void(size_t, size_t, int)>* job_ = nullptr;
    size_t job_n_ = 0, job_grain_ = 1;
    std::atomic<size_t> next_{0};
};

// Concatenate per-thread vectors into one (mergeTasks, index/calig.cpp:91-109).
template <class T>
void Gather(std::vector<std::vector<T>>& local, std::vector<T>& global) {
    size_t total = 0;
    for (auto& l : local) total += l.size();
    global.clear();
    global.reserve(total);
    for (auto& l : local) {
        global.insert(global.end(), l.begin(), l.end());
        l.clear();
    }
}

}  // namespace cg
