// One continuous-matching job on the host: CaLiG Stage 1, the query model,
// the planner and the key encoder, advanced batch by batch.  Shared by the
// cgcsm_* executables and the C ABI in cgbdsm_host.h.
#pragma once
#include <memory>
#include <string>

#include "calig.h"
#include "gen.h"
#include "keys.h"
#include "plan.h"
#include "query.h"

namespace cg {

struct S1Timing {
    double update_index_ms = 0, construct_update_ms = 0, prepare_all_ms = 0, prepare_index_ms = 0;
};

class Session {
public:
    Session(const std::string& query_path, const std::string& data_path, const std::string& update_path,
            uint32_t batch_ints, int threads, uint32_t ebits);
    Session(const SynthGraph& g, const SampledQuery& sq, uint32_t batch_ints, int threads, uint32_t ebits);

    // constructCand + staticFilter (main_calig.cpp:106-107).
    void BuildIndex();
    // Initial keys; fills card/avg (ConvertGlobalIndex host half).
    const std::vector<uint64_t>& InitialKeys();
    void MakePlan(const uint32_t* card, const float* avg);
    // Stage 1 of batch `b` (main_calig.cpp:220-262): updateIndex,
    // ConstructUpdate, mergeDeltas, PrepareUpdateAll, PrepareUpdateIndex.
    const KeyBatch& RunBatch(uint32_t b, S1Timing* t);
    // Stage 1 over stream ints [start, start + count) (batches of any size;
    // callers keep them contiguous and in order).
    const KeyBatch& RunRange(uint32_t start, uint32_t count, S1Timing* t);
    // The same pieces one at a time, for callers that time them separately.
    uint64_t BatchPairs(uint32_t b) const;

    uint32_t num_batches() const { return num_batches_; }
    uint32_t num_updates() const { return calig_->UpdateSize(); }
    uint32_t batch_ints() const { return batch_ints_; }
    uint32_t ebits() const { return ebits_; }
    CaLiG& calig() { return *calig_; }
    const Query& query() const { return query_; }
    Plan& plan() { return *plan_; }
    KeyEncoder& encoder() { return *enc_; }
    const uint32_t* card() const { return card_; }
    const float* avg() const { return avg_; }
    const KeyBatch& current() const { return cur_; }

private:
    void Setup(uint32_t batch_ints, uint32_t ebits);
    std::unique_ptr<CaLiG> calig_;
    Query query_;
    std::unique_ptr<Plan> plan_;
    std::unique_ptr<KeyEncoder> enc_;
    uint32_t batch_ints_ = UINT32_MAX, num_batches_ = 0, ebits_ = 5;
    std::vector<uint64_t> init_keys_;
    uint32_t card_[kMaxDir] = {};
    float avg_[kMaxDir] = {};
    KeyBatch cur_;
};

}  // namespace cg
