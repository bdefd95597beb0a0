// Stage-1 CaLiG candidate index (host C++), restated from index/calig.{h,cpp}.
//
// State per data vertex v (index/calig.h:28-33):
//   label            query label of v, or -1 when the label is not in the query
//   nei              sorted adjacency (the reference's hash set)
//   cand[ui][uj]     for every query vertex ui with v's label and every query
//                    neighbour uj of ui: the candidate neighbours of v for uj
//   G_Li             lit bitmask: bit ui set <=> v is a candidate of ui
//   isChecked[ui]    "checked since the last relevant change" flags
//
// Layout: the cand sets of all vertices live in one flat vector of sorted
// vectors; vertex v owns a block starting at cand_base_[v], and the slot of
// (ui, uj) inside the block is fixed per label (slot_[ui][uj]).  Parallel
// phases that mutate sets are grouped by owner vertex (every owner handled
// by one thread), which replaces the reference's 163 840 striped mutexes
// (index/calig.cpp:24-43) and makes each phase race free.
//
// Per-batch sequence (CaLiG::updateIndex, index/calig.cpp:950-1069):
//   pairs -> turnOnCond -> G.nei inserts -> turnOnCheck -> [turnOffCheck]
//   -> addAndCheck* -> addAndCheckDelete -> delAndCheck* -> li_activated deltas
// then ConstructUpdate (index/calig.cpp:776-804) and mergeDeltas (:1071-1088).
#pragma once
#include <string>
#include <unordered_map>
#include <vector>

#include "common.h"

namespace cg {

struct DeltaOp {
    uint32_t vi, vj, ui, uj;  // owner vi (for query vertex ui) lists vj (for uj)
};

// One batch's deltas; replaces delta_{ins,del,update}_merged_[batch]
// (index/calig.h:86-96) with a value per batch (the reference clears the outer
// vectors and then indexes them, index/calig.cpp:957-959 vs :1073-1075).
struct BatchDeltas {
    std::vector<DeltaOp> ins, del, upd;
};

class CaLiG {
public:
    explicit CaLiG(int threads = 0);
    ~CaLiG();

    // Loaders (index/calig.cpp:111-323, 719-733).
    void InputQ(const std::string& path);
    void InputG(const std::string& path);
    void InputUpdate(const std::string& path, int max_num = 0);
    // Binary CSR prefix (index/calig.cpp:218-323): <prefix>.meta.txt,
    // .vertex.bin (u64 row_ptr[n+1]), .edge.bin (u32 adj[m]).  Labels are
    // drawn with the C library's rand() from its initial state, every tenth
    // edge (by a multiplicative hash of the pair) becomes the update stream,
    // shuffled by std::mt19937(42).
    void InputGBin(const std::string& prefix);
    // In-memory equivalents (labels are the raw integer labels).
    void SetQuery(const std::vector<int>& labels, const std::vector<std::pair<uint32_t, uint32_t>>& edges);
    void SetData(const std::vector<int>& labels, const std::vector<std::pair<uint32_t, uint32_t>>& edges);
    void SetUpdates(const std::vector<std::pair<uint32_t, uint32_t>>& pairs);

    void ConstructCand();    // index/calig.cpp:325-356
    void StaticFilter();     // index/calig.cpp:705-717
    // start/count are in ints (two per edge), as main_calig.cpp:221-222.
    void UpdateIndex(uint32_t start, uint32_t count);          // :950-1069
    void ConstructUpdate(uint32_t start, uint32_t count);      // :776-804
    BatchDeltas MergeDeltas();                                 // :1071-1088

    uint32_t UpdateSize() const { return (uint32_t)update_.size(); }
    const std::vector<int>& UpdateStream() const { return update_; }
    uint32_t NumVertices() const { return (uint32_t)glabel_.size(); }
    int Threads() const { return pool_.size(); }
    ThreadPool& Pool() { return pool_; }

    // Read access for the key encoder and parity exports.
    uint32_t Li(uint32_t v) const { return li_[v]; }
    const std::vector<uint32_t>& LiAll() const { return li_; }
    int Label(uint32_t v) const { return glabel_[v]; }
    uint32_t QueryVertices() const { return qn_; }
    const std::vector<int>& QueryNbrs(uint32_t ui) const { return qnei_[ui]; }
    const std::vector<int>& QueryVerticesWithLabel(int label) const;
    // cand[ui][uj] of v; nullptr if v has no such slot.
    const std::vector<int>* Cand(uint32_t v, uint32_t ui, uint32_t uj) const;
    const std::vector<int>& Nei(uint32_t v) const { return nei_[v]; }

private:
    struct Task { int ui, vi; };
    struct TmpTask { int ui, uj, vi, vj; };
    struct DelTask { int ui, vi, lb; };

    int SlotOf(int ui, int uj) const { return slot_[ui * kMaxQV + uj]; }
    std::vector<int>& CandRef(int v, int ui, int uj) { return cands_[cand_base_[v] + SlotOf(ui, uj)]; }
    bool HasCand(int v, int ui) const {
        return glabel_[v] >= 0 && glabel_[v] == qlabel_[ui];
    }
    uint8_t* CheckedPtr(int v, int ui) { return &checked_[chk_base_[v] + ui_k_[ui]]; }

    void BuildQueryTables();
    void AllocCand();
    bool TryNei(size_t th, int v, int ui, std::vector<int>& used, const std::vector<int>& to_check) const;
    bool CheckNei(int v, int ui) const;
    void LiOr(int v, int ui);
    void LiAnd(int v, int ui);
    bool TestAndSetChecked(int v, int ui);
    void ClearChecked(int v, int ui);

    // Phases.
    void TurnOnCheck(std::vector<TmpTask>& tmp, std::vector<Task>& add_out);
    void AddAndCheck(std::vector<Task>& global, std::vector<Task>& next, std::vector<Task>& add_del);
    void DelAndCheck(std::vector<Task>& global, std::vector<Task>& next);
    void TurnOff(int v, std::vector<Task>& out);
    // Run fn(task, tid) grouped by owner(task) so each owner is single-threaded.
    template <class T, class Owner, class F>
    void ByOwner(const std::vector<T>& tasks, Owner owner, F fn);

    ThreadPool pool_;
    int T_;
    // query (CaLiG's view, raw labels)
    uint32_t qn_ = 0;
    std::vector<int> qlabel_;
    std::vector<std::vector<int>> qnei_;
    std::vector<std::vector<std::pair<int, std::vector<int>>>> rep_nei_;  // ui -> (label, nbrs) if >1
    std::unordered_map<int, std::vector<int>> labels_;                    // label -> query vertices
    std::vector<int> slot_;        // [ui*12+uj] -> slot inside the vertex block
    std::vector<int> ui_k_;        // ui -> rank among query vertices with its label
    std::unordered_map<int, int> block_size_;  // label -> slots per vertex
    // data
    std::vector<int> glabel_;
    std::vector<std::vector<int>> nei_;
    std::vector<uint64_t> cand_base_;
    std::vector<std::vector<int>> cands_;
    std::vector<uint64_t> chk_base_;
    std::vector<uint8_t> checked_;
    std::vector<uint32_t> li_;
    std::vector<int> update_;
    // per-thread delta buffers
    std::vector<std::vector<DeltaOp>> d_ins_, d_del_, d_upd_;
    std::vector<std::vector<Task>> li_activated_;
};

}  // namespace cg
