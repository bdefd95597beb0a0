#include "gen.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <unordered_set>

namespace cg {

static uint64_t SplitMix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

Rng::Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& w : s) w = SplitMix(x);
}

uint64_t Rng::Next() {
    auto rotl = [](uint64_t v, int k) { return (v << k) | (v >> (64 - k)); };
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
}

SynthGraph GenerateRmat(const RmatSpec& sp, ThreadPool* pool) {
    SynthGraph g;
    g.n = 1u << sp.scale;
    const uint64_t target = (uint64_t)std::llround(g.n * sp.avg_degree / 2.0);
    // Edges are drawn in independent seeded chunks (parallel when a pool is
    // given), then self loops and duplicates are removed; chunks are drawn
    // until `target` distinct edges exist.
    const double ab = sp.a + sp.b, abc = ab + sp.c;
    std::vector<uint64_t> keys;
    keys.reserve(target + target / 8);
    uint64_t chunk_id = 0;
    while (keys.size() < target) {
        uint64_t need = target - keys.size();
        uint64_t draw = need + need / 8 + 1024;
        const uint64_t per = 1u << 20;
        uint64_t nchunks = (draw + per - 1) / per;
        std::vector<std::vector<uint64_t>> parts(nchunks);
        auto work = [&](size_t c) {
            Rng r(sp.seed * 1000003ULL + (chunk_id + c) * 7919ULL + 17);
            uint64_t cnt = std::min(per, draw - c * per);
            parts[c].reserve(cnt);
            for (uint64_t i = 0; i < cnt; i++) {
                uint32_t u = 0, v = 0;
                for (uint32_t bit = 0; bit < sp.scale; bit++) {
                    double p = r.Uniform();
                    uint32_t du = 0, dv = 0;
                    if (p < sp.a) {
                    } else if (p < ab) {
                        dv = 1;
                    } else if (p < abc) {
                        du = 1;
                    } else {
                        du = dv = 1;
                    }
                    u = (u << 1) | du;
                    v = (v << 1) | dv;
                }
                if (u == v) continue;
                uint32_t x = std::min(u, v), y = std::max(u, v);
                parts[c].push_back(((uint64_t)x << 32) | y);
            }
        };
        if (pool)
            pool->Each(nchunks, 1, [&](size_t c, int) { work(c); });
        else
            for (size_t c = 0; c < nchunks; c++) work(c);
        chunk_id += nchunks;
        for (auto& p : parts) keys.insert(keys.end(), p.begin(), p.end());
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    }
    // Keep a seeded random subset of exactly `target` distinct edges.
    Rng r(sp.seed ^ 0x5bd1e995ULL);
    for (uint64_t i = keys.size(); i > 1; i--) std::swap(keys[i - 1], keys[r.Below(i)]);
    keys.resize(target);
    // Labels.
    g.labels.resize(g.n);
    Rng lr(sp.seed * 31 + 7);
    if (sp.zipf_s <= 0.0) {
        for (auto& l : g.labels) l = (int)lr.Below(sp.num_labels);
    } else {
        std::vector<double> cdf(sp.num_labels);
        double s = 0;
        for (uint32_t k = 0; k < sp.num_labels; k++) cdf[k] = (s += 1.0 / std::pow(k + 1.0, sp.zipf_s));
        for (auto& c : cdf) c /= s;
        for (auto& l : g.labels) {
            double p = lr.Uniform();
            l = (int)(std::lower_bound(cdf.begin(), cdf.end(), p) - cdf.begin());
            if (l >= (int)sp.num_labels) l = (int)sp.num_labels - 1;
        }
    }
    // Split: the first initial_frac of the (already shuffled) edges is the
    // initial graph; the stream follows.
    uint64_t n_init = (uint64_t)std::llround(sp.initial_frac * (double)keys.size());
    uint64_t n_stream = keys.size() - n_init;
    if (sp.stream_edges && sp.stream_edges < n_stream) n_stream = sp.stream_edges;
    g.initial.reserve(n_init);
    for (uint64_t i = 0; i < n_init; i++) g.initial.emplace_back((uint32_t)(keys[i] >> 32), (uint32_t)keys[i]);
    g.stream.reserve(n_stream);
    for (uint64_t i = n_init; i < n_init + n_stream; i++)
        g.stream.emplace_back((uint32_t)(keys[i] >> 32), (uint32_t)keys[i]);
    return g;
}

QuerySampler::QuerySampler(const SynthGraph& g) : g_(g) {
    // Adjacency of the final graph.
    deg_.assign(g.n + 1, 0);
    for (auto* set : {&g.initial, &g.stream})
        for (auto& [a, b] : *set) {
            deg_[a]++;
            deg_[b]++;
        }
    off_.assign(g.n + 1, 0);
    for (uint32_t v = 0; v < g.n; v++) off_[v + 1] = off_[v] + deg_[v];
    adj_.resize(off_[g.n]);
    std::vector<uint64_t> pos(off_.begin(), off_.end() - 1);
    for (auto* set : {&g.initial, &g.stream})
        for (auto& [a, b] : *set) {
            adj_[pos[a]++] = b;
            adj_[pos[b]++] = a;
        }
}

SampledQuery SampleQuery(const SynthGraph& g, uint32_t size, QueryKind kind, uint64_t seed) {
    return QuerySampler(g).Sample(size, kind, seed);
}

SampledQuery QuerySampler::Sample(uint32_t size, QueryKind kind, uint64_t seed) const {
    if (size < 2 || size > kMaxQV) throw Error("query size out of range");
    const SynthGraph& g = g_;
    const auto& deg = deg_;
    const auto& off = off_;
    const auto& adj = adj_;
    auto adjacent = [&](uint32_t a, uint32_t b) {
        for (uint64_t i = off[a]; i < off[a + 1]; i++)
            if (adj[i] == b) return true;
        return false;
    };
    for (uint64_t attempt = 0; attempt < 100000; attempt++) {
        Rng r(seed * 2654435761ULL + attempt * 40503ULL + 1);
        uint32_t start = (uint32_t)r.Below(g.n);
        if (deg[start] == 0) continue;
        std::vector<uint32_t> vs{start};
        std::vector<std::pair<uint32_t, uint32_t>> walked;  // in local ids
        uint32_t cur = start;
        for (uint32_t step = 0; step < 64 * size && vs.size() < size; step++) {
            uint32_t nx = adj[off[cur] + r.Below(deg[cur])];
            auto it = std::find(vs.begin(), vs.end(), nx);
            uint32_t li_cur = (uint32_t)(std::find(vs.begin(), vs.end(), cur) - vs.begin());
            uint32_t li_nx;
            if (it == vs.end()) {
                li_nx = (uint32_t)vs.size();
                vs.push_back(nx);
            } else {
                li_nx = (uint32_t)(it - vs.begin());
            }
            if (li_cur != li_nx) {
                auto e = std::make_pair(std::min(li_cur, li_nx), std::max(li_cur, li_nx));
                if (std::find(walked.begin(), walked.end(), e) == walked.end()) walked.push_back(e);
            }
            cur = nx;
        }
        if (vs.size() < size) continue;
        SampledQuery q;
        for (uint32_t v : vs) q.labels.push_back(g.labels[v]);
        if (kind == QueryKind::Sparse) {
            q.edges = walked;
        } else {
            std::vector<std::pair<uint32_t, uint32_t>> induced;
            for (uint32_t i = 0; i < size; i++)
                for (uint32_t j = i + 1; j < size; j++)
                    if (adjacent(vs[i], vs[j])) induced.emplace_back(i, j);
            if (kind == QueryKind::Dense) {
                if (2.0 * induced.size() / size < 3.0) continue;
                if (induced.size() > kMaxQE) continue;
                q.edges = induced;
            } else {
                // BFS tree over the induced subgraph, rooted at the walk's start.
                std::vector<int> parent(size, -1);
                std::vector<uint32_t> queue{0};
                parent[0] = 0;
                for (size_t h = 0; h < queue.size(); h++) {
                    uint32_t x = queue[h];
                    for (auto& [a, b] : induced) {
                        uint32_t y = a == x ? b : (b == x ? a : UINT32_MAX);
                        if (y == UINT32_MAX || parent[y] != -1) continue;
                        parent[y] = (int)x;
                        queue.push_back(y);
                        q.edges.emplace_back(std::min(x, y), std::max(x, y));
                    }
                }
            }
        }
        if (q.edges.size() > kMaxQE) continue;
        return q;
    }
    throw Error("could not sample a query of the requested kind");
}

void WriteDataGraph(const std::string& path, const SynthGraph& g) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) throw Error("cannot write " + path);
    for (uint32_t v = 0; v < g.n; v++) fprintf(f, "v %u %d\n", v, g.labels[v]);
    for (auto& [a, b] : g.initial) fprintf(f, "e %u %u 0\n", a, b);
    fclose(f);
}

void WriteStream(const std::string& path, const SynthGraph& g) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) throw Error("cannot write " + path);
    for (auto& [a, b] : g.stream) fprintf(f, "e %u %u 0\n", a, b);
    fclose(f);
}

void WriteBinaryCsr(const std::string& prefix, const SynthGraph& g, uint32_t num_labels) {
    std::vector<std::vector<uint32_t>> adj(g.n);
    for (const auto* part : {&g.initial, &g.stream})
        for (auto& [a, b] : *part) {
            if (a == b || a >= g.n || b >= g.n) continue;
            adj[a].push_back(b);
            adj[b].push_back(a);
        }
    std::vector<uint64_t> offs((size_t)g.n + 1, 0);
    uint32_t maxdeg = 0;
    for (uint32_t v = 0; v < g.n; v++) {
        auto& s = adj[v];
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
        offs[v + 1] = offs[v] + s.size();
        maxdeg = std::max<uint32_t>(maxdeg, (uint32_t)s.size());
    }
    FILE* f = fopen((prefix + ".meta.txt").c_str(), "w");
    if (!f) throw Error("cannot write " + prefix + ".meta.txt");
    fprintf(f, "%u %llu 4 8 4 0 %u 0 %u 0\n", g.n, (unsigned long long)offs[g.n], maxdeg, num_labels);
    fclose(f);
    f = fopen((prefix + ".vertex.bin").c_str(), "wb");
    if (!f) throw Error("cannot write " + prefix + ".vertex.bin");
    fwrite(offs.data(), sizeof(uint64_t), offs.size(), f);
    fclose(f);
    f = fopen((prefix + ".edge.bin").c_str(), "wb");
    if (!f) throw Error("cannot write " + prefix + ".edge.bin");
    for (uint32_t v = 0; v < g.n; v++) fwrite(adj[v].data(), sizeof(uint32_t), adj[v].size(), f);
    fclose(f);
}

void WriteQueryFile(const std::string& path, const SampledQuery& q) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) throw Error("cannot write " + path);
    for (size_t i = 0; i < q.labels.size(); i++) fprintf(f, "v %zu %d\n", i, q.labels[i]);
    for (auto& [a, b] : q.edges) fprintf(f, "e %u %u 0\n", a, b);
    fclose(f);
}

}  // namespace cg
