#include "common.h"

#include <dirent.h>

#include <algorithm>
#include <cstring>

namespace cg {

int NumaNodes() {
    int n = 0;
    DIR* d = opendir("/sys/devices/system/node");
    if (!d) return 1;
    while (dirent* e = readdir(d)) {
        if (strncmp(e->d_name, "node", 4) == 0 && e->d_name[4] >= '0' && e->d_name[4] <= '9') n++;
    }
    closedir(d);
    return std::max(n, 1);
}

int DefaultThreads() {
    int total = (int)std::thread::hardware_concurrency();
    if (total <= 0) total = 1;
    int t = total / NumaNodes();
    return t > 0 ? t : total;
}

ThreadPool::ThreadPool(int n) : n_(std::max(n, 1)) {
    for (int t = 1; t < n_; t++) th_.emplace_back([this, t] { Worker(t); });
}

ThreadPool::~ThreadPool() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
        gen_++;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
}

void ThreadPool::Worker(int tid) {
    uint64_t seen = 0;
    for (;;) {
        const std::function<void(size_t, size_t, int)>* job;
        size_t n, grain;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            if (stop_) return;
            job = job_;
            n = job_n_;
            grain = job_grain_;
        }
        for (;;) {
            size_t b = next_.fetch_add(grain);
            if (b >= n) break;
            (*job)(b, std::min(n, b + grain), tid);
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (--running_ == 0) done_cv_.notify_all();
        }
    }
}

void ThreadPool::For(size_t n, size_t grain, const std::function<void(size_t, size_t, int)>& fn) {
    if (n == 0) return;
    if (grain == 0) grain = 1;
    if (n_ == 1 || n <= grain) {
        fn(0, n, 0);
        return;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        job_ = &fn;
        job_n_ = n;
        job_grain_ = grain;
        next_.store(0);
        running_ = n_ - 1;
        gen_++;
    }
    cv_.notify_all();
    for (;;) {
        size_t b = next_.fetch_add(grain);
        if (b >= n) break;
        fn(b, std::min(n, b + grain), 0);
    }
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return running_ == 0; });
    job_ = nullptr;
}

}  // namespace cg
