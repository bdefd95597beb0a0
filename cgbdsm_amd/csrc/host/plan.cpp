#include "plan.h"

#include <algorithm>
#include <cstring>
#include <iostream>
#include <limits>

namespace cg {

void Plan::GenerateIndexingOrders() {
    for (uint32_t i = 0; i < q->ecount; i++)
        BuildIndexingOrder((uint8_t)q->qe_list[i].first, (uint8_t)q->qe_list[i].second, indexing[i]);
}

// BFS level order from (u0, u1); backward neighbours are listed by position.
// Position 0 carries the sentinel bni = {1} of plan.cpp:1211-1214.
void Plan::BuildIndexingOrder(uint8_t u0, uint8_t u1, Order& o) {
    std::vector<uint8_t> order{u0, u1}, pre{u0, u1}, cur;
    std::bitset<kMaxQV> seen;
    seen[u0] = seen[u1] = true;
    while (!pre.empty()) {
        for (uint8_t p : pre)
            for (auto& [nb, l] : q->nbrs[p])
                if (!seen[nb]) {
                    seen[nb] = true;
                    cur.push_back((uint8_t)nb);
                    order.push_back((uint8_t)nb);
                }
        pre.swap(cur);
        cur.clear();
    }
    if (order.size() != q->vcount) throw Error("query graph is not connected");
    o.vs[0] = order[0];
    o.bni_offs[0] = 0;
    o.bni[0] = 1;
    uint32_t cum = 1;
    o.bni_offs[1] = 1;
    for (uint32_t j = 1; j < q->vcount; j++) {
        o.vs[j] = order[j];
        for (uint32_t k = 0; k < j; k++)
            if (q->Adjacent(order[j], o.vs[k])) o.bni[cum++] = (uint8_t)k;
        o.bni_offs[j + 1] = (uint8_t)cum;
    }
}

void Plan::GenerateMatchingOrders(const uint32_t* card, const float* avg) {
    GlobalOrder(card, avg);
    for (uint32_t i = 0; i < q->ecount; i++) {
        uint8_t u0 = (uint8_t)q->qe_list[i].first, u1 = (uint8_t)q->qe_list[i].second;
        rebuild_flags[i].reset();
        rebuild_v[i].reset();
        RebuildFlags(u0, u1, {global.vs[0], global.vs[1]}, rebuild_flags[i]);
        RebuildVFlags(u0, u1, i);
    }
}

std::pair<uint8_t, uint8_t> Plan::SelectStartingEdge(const uint32_t* card, const float* avg) {
    using E = std::pair<uint8_t, uint8_t>;
    std::vector<E> sel, next;
    // Level 1: maximum endpoint degree sum.
    uint32_t best = 0;
    for (auto& [a, b] : q->qe_list) {
        uint32_t s = q->Degree(a) + q->Degree(b);
        if (s > best) {
            best = s;
            sel.clear();
            sel.emplace_back(a, b);
        } else if (s == best) {
            sel.emplace_back(a, b);
        }
    }
    // Level 2: most triangles through the edge.
    if (sel.size() > 1) {
        uint32_t bt = 0;
        for (auto& [a, b] : sel) {
            uint32_t t = 0;
            for (uint32_t w = 0; w < q->vcount; w++) {
                bool fa = false, fb = false;
                for (auto& [x, l] : q->nbrs[w]) {
                    if (x == a) fa = true;
                    else if (x == b) fb = true;
                }
                if (fa && fb) t++;
            }
            if (t > bt) {
                bt = t;
                next.clear();
                next.emplace_back(a, b);
            } else if (t == bt) {
                next.emplace_back(a, b);
            }
        }
        sel.swap(next);
        next.clear();
    }
    // Level 3: fewest candidate edges.
    if (sel.size() > 1) {
        uint32_t bc = UINT32_MAX;
        for (auto& [a, b] : sel) {
            uint32_t c = card[q->E(a, b)];
            if (c < bc) {
                bc = c;
                next.clear();
                next.emplace_back(a, b);
            } else if (c == bc) {
                next.emplace_back(a, b);
            }
        }
        sel.swap(next);
        next.clear();
    }
    // Level 4: smallest summed average degree of the other incident edges.
    if (sel.size() > 1) {
        float bs = std::numeric_limits<float>::max();
        for (auto& [a, b] : sel) {
            float s = 0.0f;
            for (auto& [x, l] : q->nbrs[a])
                if (x != b) s += avg[q->E(a, x)];
            for (auto& [x, l] : q->nbrs[b])
                if (x != a) s += avg[q->E(b, x)];
            if (s < bs) {
                bs = s;
                next.clear();
                next.emplace_back(a, b);
            } else if (s == bs) {
                next.emplace_back(a, b);
            }
        }
        sel.swap(next);
        next.clear();
    }
    return sel[0];
}

void Plan::GlobalOrder(const uint32_t* card, const float* avg) {
    const uint32_t V = q->vcount;
    std::vector<uint8_t> order;
    std::bitset<kMaxQV> vis;
    auto [f, s] = SelectStartingEdge(card, avg);
    vis[f] = vis[s] = true;
    order = {f, s};
    for (uint32_t i = 2; i < V; i++) {
        std::vector<uint8_t> sel, next;
        {
            // Extendable vertices with their smallest incoming average degree,
            // truncated to an integer as the reference's pair<uint8_t,uint32_t> does.
            std::vector<std::pair<uint8_t, uint32_t>> cand;
            for (uint32_t u = 0; u < V; u++) {
                if (vis[u]) continue;
                for (auto& [bn, l] : q->nbrs[u]) {
                    if (!vis[bn]) continue;
                    float d = avg[q->E(bn, u)];
                    if (cand.empty() || cand.back().first != u)
                        cand.emplace_back((uint8_t)u, (uint32_t)d);
                    else if (d < (float)cand.back().second)
                        cand.back().second = (uint32_t)d;
                }
            }
            if (cand.empty()) throw Error("query graph is not connected");
            // std::sort with the reference's comparator (plan.cpp:580), so
            // ties fall the same way with the same standard library.
            std::sort(cand.begin(), cand.end(), [](const auto& a, const auto& b) { return a.second < b.second; });
            size_t end = cand.size();
            for (size_t j = 0; j + 1 < cand.size(); j++) {
                if (cand[j + 1].second >= (uint32_t)(cand[j].second * (2u << 9))) {
                    end = j + 1;
                    break;
                }
            }
            for (size_t j = 0; j < end; j++) sel.push_back(cand[j].first);
            std::sort(sel.begin(), sel.end());
        }
        // 1: most backward neighbours.
        if (sel.size() > 1) {
            uint32_t best = 0;
            for (uint8_t u : sel) {
                uint32_t c = 0;
                for (auto& [bn, l] : q->nbrs[u]) c += vis[bn] ? 1 : 0;
                if (c > best) {
                    best = c;
                    next.clear();
                    next.push_back(u);
                } else if (c == best) {
                    next.push_back(u);
                }
            }
            sel.swap(next);
            next.clear();
        }
        // 2: most visited vertices adjacent to unvisited neighbours.
        if (sel.size() > 1) {
            uint32_t best = 0;
            for (uint8_t u : sel) {
                uint32_t c = 0;
                std::vector<bool> tv(V, false);
                for (auto& [fn, l] : q->nbrs[u]) {
                    if (vis[fn]) continue;
                    for (auto& [bn, l2] : q->nbrs[fn])
                        if (vis[bn] && !tv[bn]) {
                            tv[bn] = true;
                            c++;
                        }
                }
                if (c > best) {
                    best = c;
                    next.clear();
                    next.push_back(u);
                } else if (c == best) {
                    next.push_back(u);
                }
            }
            sel.swap(next);
            next.clear();
        }
        // 3: most dangling forward neighbours.
        if (sel.size() > 1) {
            uint32_t best = 0;
            for (uint8_t u : sel) {
                uint32_t c = 0;
                for (auto& [fn, l] : q->nbrs[u]) {
                    if (vis[fn]) continue;
                    bool none = true;
                    for (auto& [bn, l2] : q->nbrs[fn])
                        if (vis[bn]) {
                            none = false;
                            break;
                        }
                    if (none) c++;
                }
                if (c > best) {
                    best = c;
                    next.clear();
                    next.push_back(u);
                } else if (c == best) {
                    next.push_back(u);
                }
            }
            sel.swap(next);
            next.clear();
        }
        // 4: fewest estimated candidates.
        if (sel.size() > 1) {
            float best = std::numeric_limits<float>::max();
            for (uint8_t u : sel) {
                float c = std::numeric_limits<float>::max();
                for (auto& [bn, l] : q->nbrs[u])
                    if (vis[bn]) c = std::min(c, avg[q->E(bn, u)]);
                if (c < best) {
                    best = c;
                    next.clear();
                    next.push_back(u);
                } else if (c == best) {
                    next.push_back(u);
                }
            }
            sel.swap(next);
            next.clear();
        }
        order.push_back(sel[0]);
        vis[sel[0]] = true;
    }
    CartesianInfo(order);
    // Backward neighbours by position, nearest first, then ordered by
    // ascending average degree with std::sort as the reference does
    // (plan.cpp:720-744).
    uint32_t cum = 0;
    global.bni_offs[0] = 0;
    for (uint32_t j = 0; j < V; j++) {
        uint8_t u = order[j];
        global.vs[j] = u;
        for (int k = (int)j - 1; k >= 0; k--)
            if (q->Adjacent(u, global.vs[k])) global.bni[cum++] = (uint8_t)k;
        global.bni_offs[j + 1] = (uint8_t)cum;
        std::sort(global.bni + global.bni_offs[j], global.bni + global.bni_offs[j + 1], [&](uint8_t a, uint8_t b) {
            return avg[q->E(global.vs[a], u)] < avg[q->E(global.vs[b], u)];
        });
    }
}

void Plan::CartesianInfo(const std::vector<uint8_t>& order) {
    const uint32_t V = q->vcount;
    uint8_t type;
    uint8_t last = order[V - 1];
    if (q->Degree(last) == 1) {
        cp[V - 1] = CGBDSM_CP_TREE_SINGLE;
        type = CGBDSM_CP_TREE_PRODUCT;
    } else {
        cp[V - 1] = CGBDSM_CP_NONTREE_SINGLE;
        type = CGBDSM_CP_NONTREE_PRODUCT;
    }
    for (uint32_t i = 1; i < V - 1; i++) {
        if (type == CGBDSM_CP_NONE) {
            cp[V - 1 - i] = CGBDSM_CP_NONE;
            continue;
        }
        uint8_t v = order[V - 1 - i];
        bool valid = true;
        for (uint32_t j = 0; j < i; j++)
            if (q->Adjacent(v, order[V - 1 - j])) {
                valid = false;
                break;
            }
        if (valid) {
            if (q->Degree(v) == 1 && type == CGBDSM_CP_TREE_PRODUCT) {
                cp[V - 1 - i] = CGBDSM_CP_TREE_PRODUCT;
            } else {
                cp[V - 1 - i] = CGBDSM_CP_NONTREE_PRODUCT;
                type = CGBDSM_CP_NONTREE_PRODUCT;
            }
        } else {
            cp[V - 1 - i] = CGBDSM_CP_NONE;
            type = CGBDSM_CP_NONE;
        }
    }
}

void Plan::RebuildFlags(uint8_t u0, uint8_t u1, std::initializer_list<uint8_t> starts,
                        std::bitset<kMaxQV * kMaxQV>& flags) {
    const uint32_t V = q->vcount;
    flags[u0 * kMaxQV + u1] = true;
    flags[u1 * kMaxQV + u0] = true;
    std::vector<uint8_t> path(V + 1, 0xFF);
    for (uint8_t s : starts) {
        std::bitset<kMaxQV> vis;
        vis[s] = true;
        path[0] = s;
        std::vector<uint32_t> ci(V + 1, 0);
        uint32_t depth = 1;
        for (;;) {
            while (ci[depth] < q->nbrs[path[depth - 1]].size()) {
                uint8_t nx = (uint8_t)q->nbrs[path[depth - 1]][ci[depth]].first;
                if (vis[nx]) {
                    ci[depth]++;
                    continue;
                }
                if (nx == u0 || nx == u1) {
                    path[depth] = nx;
                    for (uint32_t i = 0; i < depth; i++) {
                        flags[path[i] * kMaxQV + path[i + 1]] = true;
                        flags[path[i + 1] * kMaxQV + path[i]] = true;
                    }
                    path[depth] = 0xFF;
                    ci[depth]++;
                    continue;
                }
                vis[nx] = true;
                path[depth] = nx;
                depth++;
            }
            if (ci[depth] >= q->nbrs[path[depth - 1]].size()) {
                ci[depth] = 0;
                depth--;
                if (depth == 0) break;
                vis[path[depth]] = false;
                ci[depth]++;
            }
        }
    }
}

void Plan::RebuildVFlags(uint8_t, uint8_t, uint32_t index) {
    const Order& o = indexing[index];
    for (uint32_t i = 0; i < q->vcount; i++) {
        uint8_t u = o.vs[i];
        if (i == 0) {
            rebuild_v[index][u] = true;
            continue;
        }
        uint32_t sum = 0, n = o.bni_offs[i + 1] - o.bni_offs[i];
        for (uint32_t k = o.bni_offs[i]; k < o.bni_offs[i + 1]; k++)
            sum += rebuild_flags[index][o.vs[o.bni[k]] * kMaxQV + u] ? 1 : 0;
        if (sum == 0)
            rebuild_v[index][u] = false;
        else if (sum == n)
            rebuild_v[index][u] = true;
        else
            throw Error("Error in order generation!");
    }
}

void FillQueryTables(const Query& q, const Plan* p, uint32_t ebits, cgbdsm_plan* out) {
    memset(out, 0, sizeof(*out));
    out->qv = q.vcount;
    out->qe = q.ecount;
    out->ndir = 2 * q.ecount;
    out->ebits = ebits;
    memcpy(out->eidx, q.eidx, sizeof(out->eidx));
    memcpy(out->dir_to_edge, q.dir_to_edge, sizeof(out->dir_to_edge));
    for (uint32_t i = 0; i < q.ecount; i++) {
        out->qe_u[i] = (uint8_t)q.qe_list[i].first;
        out->qe_v[i] = (uint8_t)q.qe_list[i].second;
    }
    if (p) {
        for (uint32_t i = 0; i < q.ecount; i++) {
            memcpy(out->io_vs[i], p->indexing[i].vs, kMaxQV);
            memcpy(out->io_bni_offs[i], p->indexing[i].bni_offs, kMaxQV + 1);
            memcpy(out->io_bni[i], p->indexing[i].bni, kMaxDir);
        }
    }
}

void Plan::Fill(cgbdsm_plan* out, uint32_t ebits) const {
    FillQueryTables(*q, this, ebits, out);
    memcpy(out->go_vs, global.vs, kMaxQV);
    memcpy(out->go_bni_offs, global.bni_offs, kMaxQV + 1);
    memcpy(out->go_bni, global.bni, kMaxDir);
    memcpy(out->go_cp, cp, kMaxQV);
    for (uint32_t a = 0; a < q->vcount; a++) {
        out->go_lcls[a] = (uint8_t)a;
        for (uint32_t b = 0; b < a; b++)
            if (q->vlabels[global.vs[b]] == q->vlabels[global.vs[a]]) {
                out->go_lcls[a] = (uint8_t)b;
                break;
            }
    }
    for (uint32_t i = 0; i < q->ecount; i++)
        for (uint32_t u = 0; u < q->vcount; u++) out->rebuild_v[i][u] = rebuild_v[i][u] ? 1 : 0;
}

void Plan::Print() const {
    std::cout << "Orders:\n";
    for (uint32_t i = 0; i < q->ecount; i++) {
        std::cout << '[' << q->qe_list[i].first << ',' << q->qe_list[i].second << "]:\n    indexing order:";
        const Order& o = indexing[i];
        for (uint32_t j = 0; j < q->vcount; j++) {
            std::cout << (uint32_t)o.vs[j] << (rebuild_v[i][o.vs[j]] ? "-build-(" : "-(");
            for (uint32_t k = o.bni_offs[j]; k < o.bni_offs[j + 1]; k++) {
                std::cout << (uint32_t)o.vs[o.bni[k]] << '@' << (uint32_t)o.bni[k];
                if (k != o.bni_offs[j + 1] - 1u) std::cout << ' ';
            }
            std::cout << ") ";
        }
        std::cout << "\n    order: ";
        for (uint32_t j = 0; j < q->vcount; j++) {
            std::cout << (uint32_t)global.vs[j] << "-(";
            for (uint32_t k = global.bni_offs[j]; k < global.bni_offs[j + 1]; k++) {
                std::cout << (uint32_t)global.vs[global.bni[k]] << '@' << (uint32_t)global.bni[k];
                if (k != global.bni_offs[j + 1] - 1u) std::cout << ' ';
            }
            std::cout << ") ";
        }
        std::cout << "\n    cartesian product info: ";
        for (uint32_t j = 0; j < q->vcount; j++) {
            std::cout << (uint32_t)global.vs[j];
            if (j == 0)
                std::cout << ' ';
            else
                std::cout << "-(" << (uint32_t)cp[j] << ") ";
        }
        std::cout << "\n";
    }
}

}  // namespace cg
