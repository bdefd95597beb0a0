// Query planner: per-edge indexing orders, the global matching order,
// rebuild flags and cartesian-product tags.
//
// Restates graph/plan.cpp:
//   BuildIndexingOrder            plan.cpp:1180-1275 (BFS from the edge's endpoints)
//   SelectStartingEdge            plan.cpp:428-533   (4-level tie break)
//   GenerateGlobalMatchingOrder   plan.cpp:535-745   (4-level greedy, float avg_degrees)
//   GenerateCartesianProductInfoGlobal plan.cpp:899-952
//   GenerateRebuildFlagsWithOrder plan.cpp:1277-1334 (edges on simple paths from the
//                                 global start vertices to the edge's endpoints)
//   GenerateRebuildVFlagsWithOrder plan.cpp:1336-1370
// Orders change the work done, never the set of counted embeddings.
#pragma once
#include <bitset>

#include "query.h"

namespace cg {

struct Order {
    uint8_t vs[kMaxQV] = {};
    uint8_t bni_offs[kMaxQV + 1] = {};
    uint8_t bni[kMaxDir] = {};
};

struct Plan {
    const Query* q = nullptr;
    Order indexing[kMaxQE];
    Order global;
    uint8_t cp[kMaxQV] = {};
    std::bitset<kMaxQV * kMaxQV> rebuild_flags[kMaxQE];
    std::bitset<kMaxQV> rebuild_v[kMaxQE];

    explicit Plan(const Query& query) : q(&query) {}
    void GenerateIndexingOrders();                                        // plan.cpp:21-30
    void GenerateMatchingOrders(const uint32_t* card, const float* avg);  // plan.cpp:31-54
    // Fill the POD tables handed to the device and the oracle.
    void Fill(cgbdsm_plan* out, uint32_t ebits) const;
    void Print() const;  // Plan::PrintOrders (plan.cpp:954-1008)

private:
    void BuildIndexingOrder(uint8_t u0, uint8_t u1, Order& o);
    std::pair<uint8_t, uint8_t> SelectStartingEdge(const uint32_t* card, const float* avg);
    void GlobalOrder(const uint32_t* card, const float* avg);
    void CartesianInfo(const std::vector<uint8_t>& order);
    void RebuildFlags(uint8_t u0, uint8_t u1, std::initializer_list<uint8_t> starts,
                      std::bitset<kMaxQV * kMaxQV>& flags);
    void RebuildVFlags(uint8_t u0, uint8_t u1, uint32_t index);
};

// Fill only the query-model part of the plan tables (before the matching order
// exists): counts, eidx, dir_to_edge, qe list, indexing orders.
void FillQueryTables(const Query& q, const Plan* p, uint32_t ebits, cgbdsm_plan* out);

}  // namespace cg
