// cgcsm_* command-line drivers: the two-stage continuous matching pipeline.
//
// Follows main_calig.cpp:41-438: the same argv (`<query> <data> <update>
// <gpu_id> [batch_size]`, batch_size counted in ints, two per edge), the
// same Stage-1 thread (updateIndex, ConstructUpdate + mergeDeltas,
// PrepareUpdateAll, PrepareUpdateIndex) handing batches to the main thread
// through a mutex + condition variable, the same Stage-2 sequence per batch
// (TransferUpdateAll, TransferUpdateIndex with the skip of empty batches,
// BuildLocalIndexBitAll, MatchingBitAll, BatchUpdateGlobalIndex) and the
// same result and profiling lines.  Stage 2 runs on the gfx950 engine
// through the cgbdsm_* C ABI only; each stage is timed on the host around a
// call followed by a stream synchronisation, as the reference's synchronous
// calls are.
//
// The executable name selects the variant (main_calig.cpp is compiled into
// one binary per CMake target):
//   cgcsm_merged_matching         merged matching, LDS stacks
//   cgcsm_gamma                   same path (the GAMMA index is the only layout)
//   cgcsm_merged_matching_global  same counts; the final DFS keeps no
//                                 per-wave stacks in global memory, so the
//                                 only difference from the LDS variant in
//                                 the reference (enumeration.cu:1157-1490)
//                                 does not exist here
//   cgcsm_gpma                    the main index kept as a packed-memory
//                                 array (CGBDSM_LAYOUT_GPMA)
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cgbdsm.h"
#include "common.h"
#include "session.h"

namespace {

using Clock = std::chrono::high_resolution_clock;

double Ms(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

void Check(cgbdsm_status st, const char* what) {
    if (st != CGBDSM_OK) {
        std::cerr << "[cgbdsm] " << what << " failed (" << st << "): " << cgbdsm_last_error() << "\n";
        std::exit(1);
    }
}

std::string BaseName(const char* p) {
    std::string s(p);
    size_t k = s.find_last_of('/');
    return k == std::string::npos ? s : s.substr(k + 1);
}

int Usage(const std::string& name) {
    std::cerr << "usage: " << name << " <query> <data.graph | binary CSR prefix> <update> <gpu_id> [batch_size]\n"
              << "  batch_size counts ints of the update stream (two per edge); default: one batch\n";
    return 2;
}

}  // namespace

int main(int argc, char* argv[]) {
    const std::string prog = BaseName(argv[0]);
    if (argc < 5) return Usage(prog);
    const std::string query_path = argv[1], data_path = argv[2], update_path = argv[3];
    const int gpu = std::atoi(argv[4]);
    const uint32_t batch_size = argc > 5 ? (uint32_t)std::strtoul(argv[5], nullptr, 10) : UINT32_MAX;

    const int numa = cg::NumaNodes();
    const int total_cores = (int)std::thread::hardware_concurrency();
    const int cores_per_node = total_cores / std::max(numa, 1);
    std::cout << "[Init] NUMA nodes: " << numa << ", total cores: " << total_cores
              << ", cores_per_node: " << cores_per_node << "\n"
              << "[Pipeline] 2-Stage: S1 → NUMA 0\n";
    if (const char* v = cgbdsm_version()) std::cout << "[Engine] " << v << " (" << prog << ")\n";

    cgbdsm_ctx* ctx = nullptr;
    Check(cgbdsm_ctx_create(gpu, 0, &ctx), "cgbdsm_ctx_create");

    std::unique_ptr<cg::Session> sess;
    std::cout << "-----------Symbi Loading graphs ------------" << std::endl;
    try {
        sess = std::make_unique<cg::Session>(query_path, data_path, update_path, batch_size,
                                             std::max(cores_per_node, 1), 5);
    } catch (const std::exception& e) {
        std::cerr << "[host] " << e.what() << "\n";
        return 1;
    }
    const uint32_t num_updates = sess->num_updates();
    const uint32_t num_batches = sess->num_batches();
    std::cout << "Total updates: " << num_updates << ", batch_size: " << batch_size
              << ", num_batches: " << num_batches << '\n';

    cgbdsm_query* q = nullptr;
    try {
        sess->BuildIndex();
        std::cout << "----------- CPU Load Graph ------------\n";
        std::cout << "----------- GPU Load Graph ------------\n";
        cgbdsm_plan tables;
        cg::FillQueryTables(sess->query(), &sess->plan(), sess->ebits(), &tables);
        Check(cgbdsm_query_create(ctx, &tables, &q), "cgbdsm_query_create");
        if (prog == "cgcsm_gpma") Check(cgbdsm_query_set_layout(q, CGBDSM_LAYOUT_GPMA), "cgbdsm_query_set_layout");
        std::cout << "----------- GPU Memory Allocation ------------\n";
        const std::vector<uint64_t>& keys = sess->InitialKeys();
        uint32_t card[CGBDSM_MAX_DIR] = {};
        float avg[CGBDSM_MAX_DIR] = {};
        Check(cgbdsm_index_init(q, keys.data(), keys.size(), sess->encoder().DV(), card, avg),
              "cgbdsm_index_init");
        std::cout << "----------- Query Plan Generation ------------\n";
        sess->MakePlan(card, avg);
        cgbdsm_plan plan;
        sess->plan().Fill(&plan, sess->ebits());
        Check(cgbdsm_query_set_plan(q, &plan), "cgbdsm_query_set_plan");
        sess->plan().Print();
        std::cout << "--------- Incremental Matching (2-Stage Pipeline) --------\n";
    } catch (const std::exception& e) {
        std::cerr << "[host] " << e.what() << "\n";
        return 1;
    }

    if (num_batches == 0) {
        std::cout << "No updates to process.\n";
        return 0;
    }

    const auto pipe_t0 = Clock::now();
    unsigned long long num_positive_matches = 0ull;
    double s2_apply = 0, s2_uidx = 0, s2_commit = 0, s2_masks = 0, s2_match = 0, s1_wait = 0;

    std::mutex mtx;
    std::condition_variable cv;
    int s1_batch = -1;
    bool s1_end = false;
    std::string s1_error;
    std::vector<cg::KeyBatch> slots(num_batches);
    std::vector<cg::S1Timing> s1_t(num_batches);

    std::thread s1([&] {
        try {
            for (uint32_t b = 0; b < num_batches; b++) {
                cg::KeyBatch kb = sess->RunBatch(b, &s1_t[b]);
                std::lock_guard<std::mutex> lk(mtx);
                slots[b] = std::move(kb);
                s1_batch = (int)b;
                cv.notify_one();
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(mtx);
            s1_error = e.what();
        }
        std::lock_guard<std::mutex> lk(mtx);
        s1_end = true;
        cv.notify_one();
    });

    for (uint32_t b = 0; b < num_batches; b++) {
        cg::KeyBatch kb;
        {
            auto w0 = Clock::now();
            std::unique_lock<std::mutex> lk(mtx);
            cv.wait(lk, [&] { return s1_batch >= (int)b || s1_end; });
            s1_wait += Ms(w0, Clock::now());
            if (s1_batch < (int)b) break;  // Stage 1 failed
            kb = std::move(slots[b]);
        }
        auto t0 = Clock::now();
        Check(cgbdsm_batch_stage(q, kb.ins.data(), kb.ins.size(), kb.del.data(), kb.del.size(), kb.upd.data(),
                                 kb.upd.size(), kb.edge_ok.data(), kb.dv),
              "cgbdsm_batch_stage");
        Check(cgbdsm_transfer_update_all(q), "cgbdsm_transfer_update_all");
        Check(cgbdsm_sync(q), "cgbdsm_sync");
        auto t1 = Clock::now();
        s2_apply += Ms(t0, t1);
        int nonempty = 0;
        Check(cgbdsm_transfer_update_index(q, &nonempty), "cgbdsm_transfer_update_index");
        if (!nonempty) continue;
        Check(cgbdsm_sync(q), "cgbdsm_sync");
        auto t2 = Clock::now();
        s2_uidx += Ms(t1, t2);

        Check(cgbdsm_build_local_index(q), "cgbdsm_build_local_index");
        Check(cgbdsm_sync(q), "cgbdsm_sync");
        auto t3 = Clock::now();
        s2_masks += Ms(t2, t3);
        Check(cgbdsm_matching(q, &num_positive_matches, nullptr), "cgbdsm_matching");
        auto t4 = Clock::now();
        s2_match += Ms(t3, t4);
        Check(cgbdsm_batch_update_global_index(q), "cgbdsm_batch_update_global_index");
        Check(cgbdsm_sync(q), "cgbdsm_sync");
        s2_commit += Ms(t4, Clock::now());
    }
    s1.join();
    if (!s1_error.empty()) {
        std::cerr << "[host] Stage 1: " << s1_error << "\n";
        return 1;
    }
    // The reference prints host wall time and a device event pair bracketing
    // the same region (utils/cuda_helpers.h:84-102); both are wall time here.
    const double pipe_ms = Ms(pipe_t0, Clock::now());
    std::cout << "Incremental Matching (2-Stage Pipeline), time (ms): " << (unsigned long)pipe_ms << "(host), "
              << (unsigned long)pipe_ms << "(kernel)\n";

    double s1_ui = 0, s1_cu = 0, s1_pa = 0, s1_pi = 0;
    for (const auto& t : s1_t) {
        s1_ui += t.update_index_ms;
        s1_cu += t.construct_update_ms;
        s1_pa += t.prepare_all_ms;
        s1_pi += t.prepare_index_ms;
    }
    std::cout << "Num Positive Matches: " << num_positive_matches << '\n';
    std::cout << "\n========== 2-Stage Pipeline Profiling Summary ==========\n";
    std::cout << "Total batches: " << num_batches << "\n";
    std::cout << "\n--- Stage 1 (CPU: updateIndex + ConstructUpdate + PrepareUpdate) ---\n";
    std::cout << "[S1] updateIndex:        " << s1_ui << " ms\n";
    std::cout << "[S1] ConstructUpdate:    " << s1_cu << " ms\n";
    std::cout << "[S1] PrepareUpdateAll:   " << s1_pa << " ms\n";
    std::cout << "[S1] PrepareUpdateIndex: " << s1_pi << " ms\n";
    const double s1_total = s1_ui + s1_cu + s1_pa + s1_pi;
    std::cout << "[S1] TOTAL:              " << s1_total << " ms\n";
    std::cout << "\n--- Stage 2 (GPU: Transfer + Match) ---\n";
    std::cout << "[S2] TransferUpdateAll:  " << s2_apply << " ms\n";
    std::cout << "[S2] TransferUpdateIdx:  " << s2_uidx << " ms\n";
    std::cout << "[S2] UpdateGlobalIndex:  " << s2_commit << " ms\n";
    std::cout << "[S2] BuildLocalIndex:    " << s2_masks << " ms\n";
    std::cout << "[S2] Matching:           " << s2_match << " ms\n";
    const double s2_total = s2_apply + s2_uidx + s2_commit + s2_masks + s2_match;
    std::cout << "[S2] TOTAL:              " << s2_total << " ms\n";
    std::cout << "\n--- Sync Waits ---\n";
    std::cout << "[Sync] S2 wait for S1:   " << s1_wait << " ms\n";
    std::cout << "\n--- Bottleneck = max(S1, S2) ---\n";
    std::cout << "S1 avg/batch: " << s1_total / num_batches << " ms\n";
    std::cout << "S2 avg/batch: " << s2_total / num_batches << " ms\n";
    std::cout << "===================================================\n";

    cgbdsm_query_destroy(q);
    cgbdsm_ctx_destroy(ctx);
    return 0;
}
