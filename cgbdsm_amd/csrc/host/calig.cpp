#include "calig.h"

#include <algorithm>
#include <cstdlib>
#include <random>
#include <cstring>
#include <fstream>

#include "textio.h"

namespace cg {

namespace {
inline bool SortedHas(const std::vector<int>& s, int x) { return std::binary_search(s.begin(), s.end(), x); }
inline bool SortedInsert(std::vector<int>& s, int x) {
    auto it = std::lower_bound(s.begin(), s.end(), x);
    if (it != s.end() && *it == x) return false;
    s.insert(it, x);
    return true;
}
inline bool SortedErase(std::vector<int>& s, int x) {
    auto it = std::lower_bound(s.begin(), s.end(), x);
    if (it == s.end() || *it != x) return false;
    s.erase(it);
    return true;
}
const std::vector<int> kEmpty;
}  // namespace

CaLiG::CaLiG(int threads) : pool_(threads > 0 ? threads : DefaultThreads()), T_(pool_.size()) {
    d_ins_.resize(T_);
    d_del_.resize(T_);
    d_upd_.resize(T_);
    li_activated_.resize(T_);
}

CaLiG::~CaLiG() = default;

// ---------------------------------------------------------------- loaders

void CaLiG::InputQ(const std::string& path) {
    // index/calig.cpp:111-164: "v id label" lines, then "e a b label" lines;
    // label -1 is read as 0.
    TextReader r(path);
    std::vector<int> labels;
    std::vector<std::pair<uint32_t, uint32_t>> edges;
    char c;
    while (r.NextChar(&c)) {
        if (c == 'v') {
            long id = r.Int(), lb = r.Int();
            if (id < 0 || id > 4096) throw Error("bad query vertex id");
            if ((size_t)id >= labels.size()) labels.resize(id + 1, 0);
            labels[id] = lb == -1 ? 0 : (int)lb;
        } else if (c == 'e') {
            long a = r.Int(), b = r.Int();
            r.Int();
            edges.emplace_back((uint32_t)a, (uint32_t)b);
        } else {
            r.SkipLine();
        }
    }
    SetQuery(labels, edges);
}

void CaLiG::SetQuery(const std::vector<int>& labels, const std::vector<std::pair<uint32_t, uint32_t>>& edges) {
    qn_ = (uint32_t)labels.size();
    if (qn_ > kMaxQV) throw Error("query too large");
    qlabel_ = labels;
    qnei_.assign(qn_, {});
    for (auto& [a, b] : edges) {
        if (a >= qn_ || b >= qn_) throw Error("query edge out of range");
        SortedInsert(qnei_[a], (int)b);
        SortedInsert(qnei_[b], (int)a);
    }
    BuildQueryTables();
}

void CaLiG::BuildQueryTables() {
    labels_.clear();
    for (uint32_t u = 0; u < qn_; u++) labels_[qlabel_[u]].push_back((int)u);
    rep_nei_.assign(qn_, {});
    for (uint32_t u = 0; u < qn_; u++) {
        std::vector<std::pair<int, std::vector<int>>> groups;
        for (int uj : qnei_[u]) {
            int lb = qlabel_[uj];
            auto it = std::find_if(groups.begin(), groups.end(), [&](auto& g) { return g.first == lb; });
            if (it == groups.end()) groups.push_back({lb, {uj}});
            else it->second.push_back(uj);
        }
        for (auto& g : groups)
            if (g.second.size() > 1) rep_nei_[u].push_back(g);
    }
    slot_.assign(kMaxQV * kMaxQV, -1);
    ui_k_.assign(qn_, -1);
    block_size_.clear();
    for (auto& [lb, us] : labels_) {
        int off = 0;
        for (size_t k = 0; k < us.size(); k++) {
            int ui = us[k];
            ui_k_[ui] = (int)k;
            for (size_t p = 0; p < qnei_[ui].size(); p++) slot_[ui * kMaxQV + qnei_[ui][p]] = off + (int)p;
            off += (int)qnei_[ui].size();
        }
        block_size_[lb] = off;
    }
}

const std::vector<int>& CaLiG::QueryVerticesWithLabel(int label) const {
    auto it = labels_.find(label);
    return it == labels_.end() ? kEmpty : it->second;
}

void CaLiG::InputG(const std::string& path) {
    // index/calig.cpp:166-216: dense ascending vertex ids (gaps get label -1),
    // labels absent from the query become -1, edges touching them are dropped.
    TextReader r(path);
    std::vector<int> labels;
    std::vector<std::pair<uint32_t, uint32_t>> edges;
    char c;
    while (r.NextChar(&c)) {
        if (c == 'v') {
            long id = r.Int(), lb = r.Int();
            if (id < 0) throw Error("bad data vertex id");
            if ((size_t)id >= labels.size()) labels.resize(id + 1, -1);
            labels[id] = (int)lb;
        } else if (c == 'e') {
            long a = r.Int(), b = r.Int();
            r.Int();
            if (a >= 0 && b >= 0) edges.emplace_back((uint32_t)a, (uint32_t)b);
        } else {
            r.SkipLine();
        }
    }
    SetData(labels, edges);
}

void CaLiG::InputGBin(const std::string& prefix) {
    uint64_t m = 0;
    uint32_t n = 0, vsize = 0, esize = 0, vlbl = 0, elbl = 0, maxdeg = 0, feat = 0, nvcls = 0, necls = 0;
    {
        std::ifstream meta(prefix + ".meta.txt");
        if (!meta) throw Error("cannot open " + prefix + ".meta.txt");
        meta >> n >> m >> vsize >> esize >> vlbl >> elbl >> maxdeg >> feat >> nvcls >> necls;
        if (!meta) throw Error("bad meta file " + prefix + ".meta.txt");
    }
    if (vsize != 4 || esize != 8) throw Error("binary CSR needs 4-byte vertex and 8-byte edge ids");
    if (nvcls == 0) throw Error("binary CSR meta names no vertex label classes");
    std::vector<uint64_t> offs((size_t)n + 1);
    std::vector<uint32_t> adj(m);
    {
        std::ifstream f(prefix + ".vertex.bin", std::ios::binary);
        if (!f.read(reinterpret_cast<char*>(offs.data()), (std::streamsize)(offs.size() * sizeof(uint64_t))))
            throw Error("short read of " + prefix + ".vertex.bin");
    }
    {
        std::ifstream f(prefix + ".edge.bin", std::ios::binary);
        if (m && !f.read(reinterpret_cast<char*>(adj.data()), (std::streamsize)(m * sizeof(uint32_t))))
            throw Error("short read of " + prefix + ".edge.bin");
    }
    if (offs[n] != m) throw Error("binary CSR: row_ptr[n] differs from the edge count");
    for (uint32_t v = 0; v < n; v++)
        if (offs[v] > offs[v + 1]) throw Error("binary CSR: row_ptr not monotone");
    // Labels: one rand() draw per vertex in id order, from the C library's
    // initial state (the reference draws them first thing in its process).
    std::vector<int> labels(n);
    std::srand(1);
    for (uint32_t v = 0; v < n; v++) labels[v] = std::rand() % (int)nvcls;
    // Every tenth undirected edge by a multiplicative hash of (min, max) goes
    // to the update stream as (larger id, smaller id); the rest is the graph.
    auto streamed = [](uint32_t a, uint32_t b) {
        const uint64_t key = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
        return (key * 0x9e3779b97f4a7c15ULL) % 10ULL == 0;
    };
    std::vector<std::pair<uint32_t, uint32_t>> edges, upd;
    edges.reserve(m / 2);
    for (uint32_t v = 0; v < n; v++)
        for (uint64_t i = offs[v]; i < offs[v + 1]; i++) {
            const uint32_t u = adj[i];
            if (u >= n) throw Error("binary CSR: neighbour id out of range");
            if (streamed(v, u)) {
                if (v < u) upd.emplace_back(u, v);
            } else if (v < u) {
                edges.emplace_back(v, u);
            }
        }
    std::vector<size_t> order(upd.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::mt19937 rng(42);
    std::shuffle(order.begin(), order.end(), rng);
    std::vector<std::pair<uint32_t, uint32_t>> shuffled(upd.size());
    for (size_t i = 0; i < order.size(); i++) shuffled[i] = upd[order[i]];
    SetData(labels, edges);
    SetUpdates(shuffled);
}

void CaLiG::SetData(const std::vector<int>& labels, const std::vector<std::pair<uint32_t, uint32_t>>& edges) {
    const size_t n = labels.size();
    glabel_.assign(n, -1);
    for (size_t v = 0; v < n; v++) glabel_[v] = labels_.count(labels[v]) ? labels[v] : -1;
    // Build sorted, de-duplicated adjacency.
    std::vector<uint64_t> deg(n + 1, 0);
    for (auto& [a, b] : edges) {
        if (a >= n || b >= n || a == b) continue;
        if (glabel_[a] == -1 || glabel_[b] == -1) continue;
        deg[a]++;
        deg[b]++;
    }
    nei_.assign(n, {});
    for (size_t v = 0; v < n; v++) nei_[v].reserve(deg[v]);
    for (auto& [a, b] : edges) {
        if (a >= n || b >= n || a == b) continue;
        if (glabel_[a] == -1 || glabel_[b] == -1) continue;
        nei_[a].push_back((int)b);
        nei_[b].push_back((int)a);
    }
    pool_.Each(n, 4096, [&](size_t v, int) {
        auto& s = nei_[v];
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
    });
    li_.assign(n, 0);
}

void CaLiG::InputUpdate(const std::string& path, int max_num) {
    // index/calig.cpp:719-733: "<char> v1 v2 w" per line.
    TextReader r(path);
    update_.clear();
    char c;
    int cnt = 0;
    while (r.NextChar(&c)) {
        long a = r.Int(), b = r.Int();
        r.Int();
        if (max_num != 0 && ++cnt > max_num) break;
        update_.push_back((int)a);
        update_.push_back((int)b);
    }
}

void CaLiG::SetUpdates(const std::vector<std::pair<uint32_t, uint32_t>>& pairs) {
    update_.clear();
    update_.reserve(pairs.size() * 2);
    for (auto& [a, b] : pairs) {
        update_.push_back((int)a);
        update_.push_back((int)b);
    }
}

// ---------------------------------------------------------------- index

void CaLiG::AllocCand() {
    const size_t n = glabel_.size();
    cand_base_.assign(n + 1, 0);
    chk_base_.assign(n + 1, 0);
    for (size_t v = 0; v < n; v++) {
        int lb = glabel_[v];
        uint64_t bs = 0, ks = 0;
        if (lb >= 0) {
            bs = (uint64_t)block_size_[lb];
            ks = labels_[lb].size();
        }
        cand_base_[v + 1] = cand_base_[v] + bs;
        chk_base_[v + 1] = chk_base_[v] + ks;
    }
    cands_.clear();
    cands_.resize(cand_base_[n]);
    checked_.assign(chk_base_[n], 0);
}

const std::vector<int>* CaLiG::Cand(uint32_t v, uint32_t ui, uint32_t uj) const {
    if (v >= glabel_.size() || ui >= qn_ || uj >= qn_) return nullptr;
    if (!HasCand((int)v, (int)ui)) return nullptr;
    int s = SlotOf((int)ui, (int)uj);
    if (s < 0) return nullptr;
    return &cands_[cand_base_[v] + s];
}

void CaLiG::ConstructCand() {
    // index/calig.cpp:325-356: every vertex starts lit for every query vertex
    // with its label; cand[ui][uj] = neighbours labelled like uj.
    AllocCand();
    const size_t n = glabel_.size();
    pool_.Each(n, 1024, [&](size_t v, int) {
        li_[v] = 0;
        int lb = glabel_[v];
        if (lb < 0) return;
        for (int ui : labels_.at(lb)) {
            for (int uj : qnei_[ui]) {
                auto& s = CandRef((int)v, ui, uj);
                s.clear();
                for (int vj : nei_[v])
                    if (glabel_[vj] == qlabel_[uj]) s.push_back(vj);  // nei_ is sorted
            }
            li_[v] |= 1u << ui;
            *CheckedPtr((int)v, ui) = 0;
        }
    });
}

inline void CaLiG::LiOr(int v, int ui) { __atomic_fetch_or(&li_[v], 1u << ui, __ATOMIC_RELAXED); }
inline void CaLiG::LiAnd(int v, int ui) { __atomic_fetch_and(&li_[v], ~(1u << ui), __ATOMIC_RELAXED); }
inline bool CaLiG::TestAndSetChecked(int v, int ui) {
    return __atomic_exchange_n(CheckedPtr(v, ui), (uint8_t)1, __ATOMIC_RELAXED) != 0;
}
inline void CaLiG::ClearChecked(int v, int ui) { __atomic_store_n(CheckedPtr(v, ui), (uint8_t)0, __ATOMIC_RELAXED); }

// Injective assignment of distinct candidates to query neighbours sharing a
// label (tryNei, index/calig.cpp:358-383).
bool CaLiG::TryNei(size_t th, int v, int ui, std::vector<int>& used, const std::vector<int>& to_check) const {
    if (th == to_check.size()) return true;
    int uj = to_check[th];
    const auto& s = cands_[cand_base_[v] + SlotOf(ui, uj)];
    for (int vj : s) {
        if (glabel_[vj] != qlabel_[uj]) continue;
        if (std::find(used.begin(), used.end(), vj) != used.end()) continue;
        used.push_back(vj);
        if (TryNei(th + 1, v, ui, used, to_check)) return true;
        used.pop_back();
    }
    return false;
}

// checkNei (index/calig.cpp:385-408).
bool CaLiG::CheckNei(int v, int ui) const {
    for (int uj : qnei_[ui])
        if (cands_[cand_base_[v] + SlotOf(ui, uj)].empty()) return false;
    std::vector<int> used;
    for (auto& g : rep_nei_[ui]) {
        used.clear();
        if (!TryNei(0, v, ui, used, g.second)) return false;
    }
    return true;
}

void CaLiG::TurnOff(int v, std::vector<Task>& out) {
    // turnOff (index/calig.cpp:499-516)
    int lb = glabel_[v];
    if (lb < 0) return;
    for (int ui : labels_.at(lb)) {
        if (!(__atomic_load_n(&li_[v], __ATOMIC_RELAXED) & (1u << ui))) continue;
        if (!CheckNei(v, ui)) {
            LiAnd(v, ui);
            out.push_back({ui, v});
        }
    }
}

template <class T, class Owner, class F>
void CaLiG::ByOwner(const std::vector<T>& tasks, Owner owner, F fn) {
    const size_t n = tasks.size();
    if (n == 0) return;
    if (T_ == 1 || n < 4096) {
        for (size_t i = 0; i < n; i++) fn(tasks[i], 0);
        return;
    }
    const uint32_t NB = (uint32_t)T_ * 64;
    std::vector<uint32_t> cnt(NB + 1, 0), b(n);
    for (size_t i = 0; i < n; i++) {
        uint32_t x = (uint32_t)owner(tasks[i]);
        x ^= x >> 16;
        x *= 0x45d9f3bu;
        x ^= x >> 16;
        b[i] = x % NB;
        cnt[b[i] + 1]++;
    }
    for (uint32_t k = 0; k < NB; k++) cnt[k + 1] += cnt[k];
    std::vector<uint32_t> order(n);
    {
        std::vector<uint32_t> pos(cnt.begin(), cnt.end() - 1);
        for (size_t i = 0; i < n; i++) order[pos[b[i]]++] = (uint32_t)i;
    }
    pool_.For(NB, 1, [&](size_t bb, size_t be, int tid) {
        for (size_t k = bb; k < be; k++)
            for (uint32_t j = cnt[k]; j < cnt[k + 1]; j++) fn(tasks[order[j]], tid);
    });
}

void CaLiG::DelAndCheck(std::vector<Task>& global, std::vector<Task>& next) {
    // delAndCheck (index/calig.cpp:410-475)
    std::vector<std::vector<TmpTask>> ltmp(T_);
    pool_.For(global.size(), 256, [&](size_t b, size_t e, int tid) {
        for (size_t i = b; i < e; i++) {
            int ui = global[i].ui, vi = global[i].vi;
            for (int uj : qnei_[ui])
                for (int vj : cands_[cand_base_[vi] + SlotOf(ui, uj)]) {
                    if (glabel_[vj] != qlabel_[uj]) continue;
                    if (li_[vj] & (1u << uj)) ltmp[tid].push_back({ui, uj, vi, vj});
                }
        }
    });
    std::vector<TmpTask> tmp;
    Gather(ltmp, tmp);

    std::vector<std::vector<Task>> lnext(T_);
    std::vector<std::vector<DelTask>> lcheck(T_);
    ByOwner(tmp, [](const TmpTask& t) { return t.vj; }, [&](const TmpTask& t, int tid) {
        auto& s = CandRef(t.vj, t.uj, t.ui);
        if (SortedErase(s, t.vi)) d_del_[tid].push_back({(uint32_t)t.vj, (uint32_t)t.vi, (uint32_t)t.uj, (uint32_t)t.ui});
        if (s.empty()) {
            LiAnd(t.vj, t.uj);
            lnext[tid].push_back({t.uj, t.vj});
        } else {
            int lb = qlabel_[t.ui];
            for (auto& g : rep_nei_[t.uj])
                if (g.first == lb) {
                    lcheck[tid].push_back({t.uj, t.vj, lb});
                    ClearChecked(t.vj, t.uj);
                    break;
                }
        }
    });
    std::vector<DelTask> check;
    Gather(lcheck, check);
    // The reference re-checks only the group of the removed candidate's label
    // (tryNei on rep_nei[lb], index/calig.cpp:452-470) but dedups the checks
    // of a pair with one flag per (vertex, query vertex): when candidates of
    // two label groups of one pair go in the same round, whichever check
    // runs first sets the flag and the other group is never re-checked, so
    // the pair can stay lit without an injective support (a TBB-order
    // dependent result).  Here the one check of a pair covers all of its
    // groups (checkNei), which makes the lit bits the greatest fixpoint.
    pool_.For(check.size(), 256, [&](size_t b, size_t e, int tid) {
        for (size_t i = b; i < e; i++) {
            int uj = check[i].ui, vj = check[i].vi;
            if (TestAndSetChecked(vj, uj)) continue;
            if (!(__atomic_load_n(&li_[vj], __ATOMIC_RELAXED) & (1u << uj))) continue;  // emptied in phase 2
            if (!CheckNei(vj, uj)) {
                LiAnd(vj, uj);
                lnext[tid].push_back({uj, vj});
            }
        }
    });
    Gather(lnext, next);
}

void CaLiG::StaticFilter() {
    // staticFilter (index/calig.cpp:705-717): turnOff everything, then the
    // delAndCheck fixpoint.
    std::vector<std::vector<Task>> local(T_);
    pool_.For(glabel_.size(), 1024, [&](size_t b, size_t e, int tid) {
        for (size_t v = b; v < e; v++) TurnOff((int)v, local[tid]);
    });
    std::vector<Task> global, next;
    Gather(local, global);
    while (!global.empty()) {
        DelAndCheck(global, next);
        global.swap(next);
        next.clear();
    }
    // Deletions recorded during the initial filter are not a batch's deltas.
    for (auto& d : d_del_) d.clear();
}

void CaLiG::TurnOnCheck(std::vector<TmpTask>& tmp, std::vector<Task>& add_out) {
    // turnOnCheck (index/calig.cpp:657-693)
    std::vector<std::vector<Task>> ladd(T_), lcheck(T_);
    ByOwner(tmp, [](const TmpTask& t) { return t.vi; }, [&](const TmpTask& t, int tid) {
        SortedInsert(CandRef(t.vi, t.ui, t.uj), t.vj);
        if (li_[t.vi] & (1u << t.ui)) {
            ladd[tid].push_back({t.ui, t.vi});
        } else {
            lcheck[tid].push_back({t.ui, t.vi});
            ClearChecked(t.vi, t.ui);
        }
    });
    std::vector<Task> check;
    Gather(lcheck, check);
    pool_.For(check.size(), 256, [&](size_t b, size_t e, int tid) {
        for (size_t i = b; i < e; i++) {
            int v1 = check[i].vi, ui = check[i].ui;
            if (TestAndSetChecked(v1, ui)) continue;
            if (__atomic_load_n(&li_[v1], __ATOMIC_RELAXED) & (1u << ui)) continue;
            if (CheckNei(v1, ui)) {
                LiOr(v1, ui);
                ladd[tid].push_back({ui, v1});
                li_activated_[tid].push_back({ui, v1});
            }
        }
    });
    Gather(ladd, add_out);
}

void CaLiG::AddAndCheck(std::vector<Task>& global, std::vector<Task>& next, std::vector<Task>& add_del) {
    // addAndCheck (index/calig.cpp:536-638); both reference branches (bucketed
    // for >= 100 000 tasks, direct otherwise) compute the same sets.
    std::vector<std::vector<TmpTask>> ltmp(T_);
    std::vector<std::vector<Task>> lcheck(T_);
    pool_.For(global.size(), 256, [&](size_t b, size_t e, int tid) {
        for (size_t i = b; i < e; i++) {
            int ui = global[i].ui, vi = global[i].vi;
            for (int uj : qnei_[ui])
                for (int vj : cands_[cand_base_[vi] + SlotOf(ui, uj)]) {
                    if (glabel_[vj] != qlabel_[uj]) continue;
                    ltmp[tid].push_back({ui, uj, vi, vj});
                    if (!(li_[vj] & (1u << uj))) {
                        lcheck[tid].push_back({uj, vj});
                        ClearChecked(vj, uj);
                    }
                }
        }
    });
    std::vector<TmpTask> tmp;
    Gather(ltmp, tmp);
    ByOwner(tmp, [](const TmpTask& t) { return t.vj; }, [&](const TmpTask& t, int tid) {
        if (SortedInsert(CandRef(t.vj, t.uj, t.ui), t.vi))
            d_ins_[tid].push_back({(uint32_t)t.vj, (uint32_t)t.vi, (uint32_t)t.uj, (uint32_t)t.ui});
    });
    std::vector<Task> check;
    Gather(lcheck, check);
    std::vector<std::vector<Task>> lnext(T_), ladd_del(T_);
    pool_.For(check.size(), 256, [&](size_t b, size_t e, int tid) {
        for (size_t i = b; i < e; i++) {
            int vj = check[i].vi, uj = check[i].ui;
            if (TestAndSetChecked(vj, uj)) continue;
            if (CheckNei(vj, uj)) {
                LiOr(vj, uj);
                lnext[tid].push_back({uj, vj});
                li_activated_[tid].push_back({uj, vj});
            } else {
                ladd_del[tid].push_back({uj, vj});
            }
        }
    });
    Gather(lnext, next);
    std::vector<Task> ad;
    Gather(ladd_del, ad);
    add_del.insert(add_del.end(), ad.begin(), ad.end());
}

void CaLiG::UpdateIndex(uint32_t start, uint32_t count) {
    for (auto& v : d_ins_) v.clear();
    for (auto& v : d_del_) v.clear();
    for (auto& v : d_upd_) v.clear();
    for (auto& v : li_activated_) v.clear();

    const uint32_t end = (uint32_t)std::min<uint64_t>((uint64_t)start + count, update_.size());
    const uint32_t num_pairs = end > start ? (end - start) / 2 : 0;
    const int n = (int)glabel_.size();

    // Pairs -> G.nei inserts + turnOnCond tasks (index/calig.cpp:975-999, 640-655).
    std::vector<std::vector<std::pair<int, int>>> lupd(T_);
    std::vector<std::vector<TmpTask>> ltmp(T_);
    auto turn_on_cond = [&](int v1, int v2, int tid) {
        for (int ui : labels_.at(glabel_[v1]))
            for (int uj : qnei_[ui])
                if (glabel_[v2] == qlabel_[uj]) ltmp[tid].push_back({ui, uj, v1, v2});
    };
    pool_.For(num_pairs, 512, [&](size_t b, size_t e, int tid) {
        for (size_t idx = b; idx < e; idx++) {
            int v1 = update_[start + 2 * idx], v2 = update_[start + 2 * idx + 1];
            // Negative ids encode deletions in the reference (index/calig.cpp:980-989);
            // the stream formats handled here carry insertions only.
            if (v1 < 0 || v2 < 0 || v1 >= n || v2 >= n || v1 == v2) continue;
            if (glabel_[v1] == -1 || glabel_[v2] == -1) continue;
            if (SortedHas(nei_[v1], v2)) continue;
            lupd[tid].push_back({v1, v2});
            lupd[tid].push_back({v2, v1});
            turn_on_cond(v1, v2, tid);
            turn_on_cond(v2, v1, tid);
        }
    });
    std::vector<std::pair<int, int>> upd;
    Gather(lupd, upd);
    ByOwner(upd, [](const std::pair<int, int>& p) { return p.first; },
            [&](const std::pair<int, int>& p, int) { SortedInsert(nei_[p.first], p.second); });

    std::vector<TmpTask> tmp;
    Gather(ltmp, tmp);
    std::vector<Task> add, next, add_del, del;
    TurnOnCheck(tmp, add);

    auto dedup = [](std::vector<Task>& v) {
        std::sort(v.begin(), v.end(), [](const Task& a, const Task& b) {
            return a.vi != b.vi ? a.vi < b.vi : a.ui < b.ui;
        });
        v.erase(std::unique(v.begin(), v.end(), [](const Task& a, const Task& b) {
                    return a.vi == b.vi && a.ui == b.ui;
                }),
                v.end());
    };
    dedup(add);
    while (!add.empty()) {
        AddAndCheck(add, next, add_del);
        add.swap(next);
        next.clear();
        dedup(add);
    }
    // addAndCheckDelete (index/calig.cpp:695-703)
    for (auto& t : add_del)
        if (!(li_[t.vi] & (1u << t.ui))) del.push_back(t);
    while (!del.empty()) {
        DelAndCheck(del, next);
        del.swap(next);
        next.clear();
    }
    // Newly lit pairs emit all their candidate entries (index/calig.cpp:1048-1067).
    std::vector<Task> act;
    Gather(li_activated_, act);
    pool_.For(act.size(), 256, [&](size_t b, size_t e, int tid) {
        for (size_t i = b; i < e; i++) {
            int vi = act[i].vi, ui = act[i].ui;
            if (!(li_[vi] & (1u << ui))) continue;
            for (int uj : qnei_[ui])
                for (int vj : cands_[cand_base_[vi] + SlotOf(ui, uj)])
                    d_ins_[tid].push_back({(uint32_t)vi, (uint32_t)vj, (uint32_t)ui, (uint32_t)uj});
        }
    });
}

void CaLiG::ConstructUpdate(uint32_t start, uint32_t count) {
    const uint32_t end = (uint32_t)std::min<uint64_t>((uint64_t)start + count, update_.size());
    const uint32_t num_pairs = end > start ? (end - start) / 2 : 0;
    const int n = (int)glabel_.size();
    pool_.For(num_pairs, 512, [&](size_t b, size_t e, int tid) {
        for (size_t idx = b; idx < e; idx++) {
            int v1 = update_[start + 2 * idx], v2 = update_[start + 2 * idx + 1];
            if (v1 < 0 || v2 < 0 || v1 >= n || v2 >= n) continue;
            int lb = glabel_[v1];
            if (lb < 0) continue;
            for (int u1 : labels_.at(lb)) {
                if (!(li_[v1] & (1u << u1))) continue;
                for (int u2 : qnei_[u1]) {
                    if (!SortedHas(cands_[cand_base_[v1] + SlotOf(u1, u2)], v2)) continue;
                    DeltaOp a{(uint32_t)v1, (uint32_t)v2, (uint32_t)u1, (uint32_t)u2};
                    DeltaOp r{(uint32_t)v2, (uint32_t)v1, (uint32_t)u2, (uint32_t)u1};
                    d_upd_[tid].push_back(a);
                    d_del_[tid].push_back(a);
                    d_upd_[tid].push_back(r);
                    d_del_[tid].push_back(r);
                }
            }
        }
    });
}

BatchDeltas CaLiG::MergeDeltas() {
    BatchDeltas out;
    Gather(d_ins_, out.ins);
    Gather(d_del_, out.del);
    Gather(d_upd_, out.upd);
    return out;
}

}  // namespace cg
