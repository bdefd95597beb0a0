// Query graph model and edge numbering.
//
// Follows CPUGraphLoader (graph/graph.cpp:43-121) for parsing and
// SetQueryMeta (graph/graph.cpp:123-199) for the numbering that defines the
// edge tag and the visibility rule:
//   * directed edge index eidx[u][uu] = position of uu in u's sorted adjacency
//     within the query's CSR (graph.cpp:137-148);
//   * undirected edges qe_list (u < uu, ascending u then uu) and their two
//     directions qe_eidx (graph.cpp:184-199).
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "common.h"

namespace cg {

struct Query {
    uint32_t vcount = 0, ecount = 0;
    std::vector<int> raw_labels;                       // labels as written in the file
    std::vector<uint32_t> vlabels;                     // dense remapped labels
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> nbrs;  // sorted (nbr, elabel)
    uint8_t eidx[kMaxQV * kMaxQV];                     // [u*kMaxQV+uu]
    std::vector<std::pair<uint32_t, uint32_t>> qe_list;
    std::vector<std::pair<uint32_t, uint32_t>> qe_eidx;
    uint8_t dir_to_edge[kMaxDir];

    uint8_t E(uint32_t u, uint32_t uu) const { return eidx[u * kMaxQV + uu]; }
    bool Adjacent(uint32_t u, uint32_t uu) const { return eidx[u * kMaxQV + uu] != kNoEdge; }
    uint32_t Degree(uint32_t u) const { return (uint32_t)nbrs[u].size(); }
};

// Parse a `.graph` query file ("t ...", "v id label", "e u v label").
// Throws cg::Error on malformed input or limit violations
// (MAX_VCOUNT/MAX_ECOUNT, graph.cpp:128-133).
Query LoadQuery(const std::string& path);
Query QueryFromEdges(const std::vector<int>& labels, const std::vector<std::pair<uint32_t, uint32_t>>& edges);
void WriteQuery(const std::string& path, const Query& q);

}  // namespace cg
