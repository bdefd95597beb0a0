// Host half of the Stage-2 helper: compressed vertex ids and 64-bit key
// encoding of the CaLiG deltas (CaLiGHelperGamma, index/calig_helper_gamma.cu).
//
//   buildGlobalVertexMapping / UpdateVertexMappingCPU  (:77-105, :132-181)
//     every data vertex with any lit bit gets the next compressed id; ids are
//     never reused.  The reference assigns new ids in ska::flat_hash_set
//     iteration order; here they are assigned in ascending original id, which
//     only permutes compressed ids (row order), never counts.
//   ConvertGlobalIndex host half                      (:238-277)
//   PrepareUpdateAll                                   (:363-401)
//     ins kept only when the owner is still lit; del kept when both ends mapped
//   PrepareUpdateIndex                                 (:403-445)
//     upd kept when both ends mapped; edge_ok[e] = both directions non-empty
#pragma once
#include <vector>

#include "calig.h"
#include "query.h"

namespace cg {

struct KeyBatch {
    std::vector<uint64_t> ins, del, upd;
    std::vector<uint8_t> edge_ok;  // [qe]
    uint32_t dv = 0;
};

class KeyEncoder {
public:
    KeyEncoder(const CaLiG& calig, const Query& q, uint32_t ebits);
    // Maps every currently lit vertex and returns one key per candidate entry
    // of every lit (vertex, query vertex) pair; fills card[ndir], avg[ndir].
    std::vector<uint64_t> InitialKeys(uint32_t* card, float* avg, ThreadPool& pool);
    KeyBatch Prepare(const BatchDeltas& d, ThreadPool& pool);
    uint32_t DV() const { return (uint32_t)old_of_.size(); }
    uint32_t NewId(uint32_t v) const { return v < new_of_.size() ? new_of_[v] : UINT32_MAX; }
    const std::vector<uint32_t>& OldIds() const { return old_of_; }
    uint32_t ebits() const { return ebits_; }
    uint64_t Encode(uint32_t e, uint32_t u, uint32_t v) const;

private:
    void ExtendMapping();
    const CaLiG& c_;
    const Query& q_;
    uint32_t ebits_;
    std::vector<uint32_t> new_of_;  // original -> compressed
    std::vector<uint32_t> old_of_;  // compressed -> original
};

}  // namespace cg
