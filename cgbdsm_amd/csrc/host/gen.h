// Seeded synthetic workloads: labeled R-MAT data graphs, a shuffled
// initial/stream split, and random-walk query sampling (tree, dense, sparse).
// The reference ships a query sampler (sampler/graph_sampler.h) that grows
// BFS/DFS trees; the workloads here ask for random-walk queries, so the
// sampler is written from those definitions:
//   tree   BFS tree over the vertices of a random walk (labels copied)
//   dense  induced subgraph on the walk's vertices, kept if avg degree >= 3
//   sparse the edges the walk traversed (not induced)
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "common.h"

namespace cg {

struct Rng {  // xoshiro256** seeded by splitmix64
    uint64_t s[4];
    explicit Rng(uint64_t seed);
    uint64_t Next();
    double Uniform() { return (Next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t Below(uint64_t n) { return n ? Next() % n : 0; }
};

struct SynthGraph {
    uint32_t n = 0;
    std::vector<int> labels;
    std::vector<std::pair<uint32_t, uint32_t>> initial, stream;  // undirected, a < b not required
};

struct RmatSpec {
    uint32_t scale = 14;
    double avg_degree = 8;
    double a = 0.57, b = 0.19, c = 0.19;  // d = 1 - a - b - c
    uint32_t num_labels = 8;
    double zipf_s = 0.0;                  // 0: uniform labels
    double initial_frac = 0.9;            // share of edges in the initial graph
    uint64_t stream_edges = 0;            // 0: all remaining edges
    uint64_t seed = 1;
};

SynthGraph GenerateRmat(const RmatSpec& spec, ThreadPool* pool = nullptr);

enum class QueryKind { Tree = 0, Dense = 1, Sparse = 2 };

struct SampledQuery {
    std::vector<int> labels;
    std::vector<std::pair<uint32_t, uint32_t>> edges;
};

// Samples a connected query of `size` vertices from the union of initial and
// stream edges; retries with derived seeds until the kind's rule holds.
SampledQuery SampleQuery(const SynthGraph& g, uint32_t size, QueryKind kind, uint64_t seed);

// The same sampler with the graph's adjacency built once (many queries).
class QuerySampler {
public:
    explicit QuerySampler(const SynthGraph& g);
    SampledQuery Sample(uint32_t size, QueryKind kind, uint64_t seed) const;

private:
    const SynthGraph& g_;
    std::vector<uint32_t> deg_;
    std::vector<uint64_t> off_;
    std::vector<uint32_t> adj_;
};

void WriteDataGraph(const std::string& path, const SynthGraph& g);
void WriteStream(const std::string& path, const SynthGraph& g);
// The whole graph (initial and stream edges) as the binary CSR prefix read
// by CaLiG::InputGBin; num_labels is written as the vertex label classes.
void WriteBinaryCsr(const std::string& prefix, const SynthGraph& g, uint32_t num_labels);
void WriteQueryFile(const std::string& path, const SampledQuery& q);

}  // namespace cg
