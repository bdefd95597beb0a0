#include "query.h"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <unordered_map>

namespace cg {

namespace {

// Edge numbering of SetQueryMeta (graph/graph.cpp:137-148, 184-199).
void Finish(Query& q) {
    if (q.vcount < 2) throw Error("query must have at least 2 vertices");
    if (q.vcount > kMaxQV || q.ecount > kMaxQE)
        throw Error("query has " + std::to_string(q.vcount) + " vertices / " + std::to_string(q.ecount) +
                    " edges; limits are " + std::to_string(kMaxQV) + " / " + std::to_string(kMaxQE));
    memset(q.eidx, kNoEdge, sizeof(q.eidx));
    memset(q.dir_to_edge, kNoEdge, sizeof(q.dir_to_edge));
    uint32_t pos = 0;
    for (uint32_t u = 0; u < q.vcount; u++)
        for (auto& [uu, l] : q.nbrs[u]) q.eidx[u * kMaxQV + uu] = (uint8_t)pos++;
    if (pos != 2 * q.ecount) throw Error("query has duplicate edges or self loops");
    q.qe_list.clear();
    q.qe_eidx.clear();
    for (uint32_t u = 0; u < q.vcount; u++)
        for (auto& [uu, l] : q.nbrs[u]) {
            if (u > uu) continue;
            q.qe_list.emplace_back(u, uu);
            q.qe_eidx.emplace_back(q.E(u, uu), q.E(uu, u));
        }
    for (uint32_t e = 0; e < q.qe_list.size(); e++) {
        q.dir_to_edge[q.qe_eidx[e].first] = (uint8_t)e;
        q.dir_to_edge[q.qe_eidx[e].second] = (uint8_t)e;
    }
}

void AddEdge(Query& q, uint32_t a, uint32_t b, uint32_t lab) {
    if (a == b) throw Error("query self loop");
    auto ins = [&](uint32_t x, uint32_t y) {
        std::pair<uint32_t, uint32_t> p(y, lab);
        auto it = std::lower_bound(q.nbrs[x].begin(), q.nbrs[x].end(), p);
        q.nbrs[x].insert(it, p);
    };
    ins(a, b);
    ins(b, a);
    q.ecount++;
}

}  // namespace

Query LoadQuery(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw Error("Failed to open: " + path);
    Query q;
    std::unordered_map<int, uint32_t> vmap, emap;
    std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> edges;
    char type;
    while (in >> type) {
        if (type == 't') {
            std::string a, b;
            in >> a >> b;
        } else if (type == 'v') {
            long id;
            int lab;
            in >> id >> lab;
            if (id < 0 || id >= (long)kMaxQV * 4) throw Error("bad query vertex id");
            if (!vmap.count(lab)) {
                uint32_t s = (uint32_t)vmap.size();
                vmap[lab] = s;
            }
            if ((size_t)id >= q.raw_labels.size()) {
                q.raw_labels.resize(id + 1, -1);
                q.vlabels.resize(id + 1, UINT32_MAX);
            }
            // CaLiG::inputQ maps label -1 to 0 (index/calig.cpp:124-126).
            q.raw_labels[id] = lab == -1 ? 0 : lab;
            q.vlabels[id] = vmap[lab];
            q.vcount++;
        } else if (type == 'e') {
            long a, b;
            int lab;
            in >> a >> b >> lab;
            if (!emap.count(lab)) {
                uint32_t s = (uint32_t)emap.size();
                emap[lab] = s;
            }
            edges.emplace_back((uint32_t)a, (uint32_t)b, emap[lab]);
        } else {
            std::string rest;
            std::getline(in, rest);
        }
    }
    if (q.vcount != q.raw_labels.size()) throw Error("query vertex ids must be 0..n-1");
    q.nbrs.assign(q.vcount, {});
    for (auto& [a, b, l] : edges) {
        if (a >= q.vcount || b >= q.vcount) throw Error("query edge endpoint out of range");
        AddEdge(q, a, b, l);
    }
    Finish(q);
    return q;
}

Query QueryFromEdges(const std::vector<int>& labels, const std::vector<std::pair<uint32_t, uint32_t>>& edges) {
    Query q;
    q.vcount = (uint32_t)labels.size();
    q.raw_labels = labels;
    std::map<int, uint32_t> vmap;
    q.vlabels.resize(labels.size());
    for (size_t i = 0; i < labels.size(); i++) {
        if (!vmap.count(labels[i])) {
            uint32_t s = (uint32_t)vmap.size();
            vmap[labels[i]] = s;
        }
        q.vlabels[i] = vmap[labels[i]];
    }
    q.nbrs.assign(q.vcount, {});
    for (auto& [a, b] : edges) AddEdge(q, a, b, 0);
    Finish(q);
    return q;
}

void WriteQuery(const std::string& path, const Query& q) {
    std::ofstream out(path);
    if (!out) throw Error("cannot write " + path);
    // No "t" header: CaLiG::inputQ (index/calig.cpp:111-145) reads only v/e lines.
    for (uint32_t u = 0; u < q.vcount; u++) out << "v " << u << " " << q.raw_labels[u] << "\n";
    for (auto& [a, b] : q.qe_list) out << "e " << a << " " << b << " 0\n";
}

}  // namespace cg
