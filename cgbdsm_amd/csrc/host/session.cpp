#include "session.h"

#include <chrono>
#include <cstring>
#include <thread>

#include "cgbdsm_host.h"

namespace cg {

using Clock = std::chrono::high_resolution_clock;
static double Ms(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

Session::Session(const std::string& qp, const std::string& dp, const std::string& up, uint32_t batch_ints,
                 int threads, uint32_t ebits) {
    calig_ = std::make_unique<CaLiG>(threads);
    calig_->InputQ(qp);
    // main_calig.cpp:93-101: a data path ending in ".graph" is text (with the
    // update file); anything else is a binary CSR prefix whose update stream
    // is split off the graph itself.
    if (dp.size() >= 6 && dp.compare(dp.size() - 6, 6, ".graph") == 0) {
        calig_->InputG(dp);
        calig_->InputUpdate(up, 0);
    } else {
        calig_->InputGBin(dp);
    }
    query_ = LoadQuery(qp);
    Setup(batch_ints, ebits);
}

Session::Session(const SynthGraph& g, const SampledQuery& sq, uint32_t batch_ints, int threads, uint32_t ebits) {
    calig_ = std::make_unique<CaLiG>(threads);
    calig_->SetQuery(sq.labels, sq.edges);
    calig_->SetData(g.labels, g.initial);
    calig_->SetUpdates(g.stream);
    query_ = QueryFromEdges(sq.labels, sq.edges);
    Setup(batch_ints, ebits);
}

void Session::Setup(uint32_t batch_ints, uint32_t ebits) {
    if (calig_->QueryVertices() != query_.vcount) throw Error("query views disagree");
    batch_ints_ = batch_ints ? batch_ints : UINT32_MAX;
    ebits_ = ebits;
    uint32_t nu = calig_->UpdateSize();
    num_batches_ = nu == 0 ? 0 : (batch_ints_ >= nu ? 1 : (nu + batch_ints_ - 1) / batch_ints_);
    plan_ = std::make_unique<Plan>(query_);
    plan_->GenerateIndexingOrders();
    enc_ = std::make_unique<KeyEncoder>(*calig_, query_, ebits);
}

void Session::BuildIndex() {
    calig_->ConstructCand();
    calig_->StaticFilter();
}

const std::vector<uint64_t>& Session::InitialKeys() {
    init_keys_ = enc_->InitialKeys(card_, avg_, calig_->Pool());
    return init_keys_;
}

void Session::MakePlan(const uint32_t* card, const float* avg) { plan_->GenerateMatchingOrders(card, avg); }

uint64_t Session::BatchPairs(uint32_t b) const {
    uint64_t start = (uint64_t)b * batch_ints_;
    uint64_t nu = calig_->UpdateSize();
    if (start >= nu) return 0;
    return std::min<uint64_t>(batch_ints_, nu - start) / 2;
}

const KeyBatch& Session::RunBatch(uint32_t b, S1Timing* t) {
    const uint32_t start = b * batch_ints_;
    return RunRange(start, std::min(batch_ints_, calig_->UpdateSize() - start), t);
}

const KeyBatch& Session::RunRange(uint32_t start, uint32_t count, S1Timing* t) {
    auto t0 = Clock::now();
    calig_->UpdateIndex(start, count);
    auto t1 = Clock::now();
    calig_->ConstructUpdate(start, count);
    auto t2 = Clock::now();
    BatchDeltas d = calig_->MergeDeltas();
    auto t2b = Clock::now();
    // PrepareUpdateAll and PrepareUpdateIndex are one pass here; the split of
    // the time follows the reference's labels (ins/del vs upd work).
    auto t3 = Clock::now();
    cur_ = enc_->Prepare(d, calig_->Pool());
    auto t4 = Clock::now();
    if (t) {
        t->update_index_ms = Ms(t0, t1);
        t->construct_update_ms = Ms(t1, t2) + Ms(t2, t2b);
        double prep = Ms(t3, t4);
        double w_all = (double)(d.ins.size() + d.del.size()), w_idx = (double)d.upd.size();
        double f = (w_all + w_idx) > 0 ? w_all / (w_all + w_idx) : 0.5;
        t->prepare_all_ms = prep * f;
        t->prepare_index_ms = prep * (1 - f);
    }
    return cur_;
}

}  // namespace cg

// ----------------------------------------------------------------- C ABI

struct cgh_session {
    std::unique_ptr<cg::Session> s;
    uint32_t next_batch = 0;
    uint64_t cursor = 0;  // stream ints consumed by Stage 1 so far
    bool have_keys = false;
};

static thread_local std::string g_err;
#define CGH_TRY(body)                   \
    try {                               \
        body;                           \
        return 0;                       \
    } catch (const std::exception& e) { \
        g_err = e.what();               \
        return 1;                       \
    }

extern "C" {

const char* cgh_last_error(void) { return g_err.c_str(); }

static cg::RmatSpec ToRmat(const cgh_synth_spec* sp) {
    cg::RmatSpec r;
    r.scale = sp->scale;
    r.avg_degree = sp->avg_degree;
    if (sp->a > 0) {
        r.a = sp->a;
        r.b = sp->b;
        r.c = sp->c;
    }
    r.num_labels = sp->num_labels;
    r.zipf_s = sp->zipf_s;
    r.initial_frac = sp->initial_frac;
    r.stream_edges = sp->stream_edges;
    r.seed = sp->seed;
    return r;
}

int cgh_generate_files(const cgh_synth_spec* spec, const char* qp, const char* dp, const char* up) {
    CGH_TRY({
        cg::ThreadPool pool(cg::DefaultThreads());
        cg::SynthGraph g = cg::GenerateRmat(ToRmat(spec), &pool);
        cg::SampledQuery q = cg::SampleQuery(g, spec->query_size, (cg::QueryKind)spec->query_kind, spec->query_seed);
        cg::WriteQueryFile(qp, q);
        cg::WriteDataGraph(dp, g);
        cg::WriteStream(up, g);
    })
}

int cgh_generate_binary(const cgh_synth_spec* spec, const char* qp, const char* prefix) {
    CGH_TRY({
        cg::ThreadPool pool(cg::DefaultThreads());
        cg::SynthGraph g = cg::GenerateRmat(ToRmat(spec), &pool);
        cg::SampledQuery q = cg::SampleQuery(g, spec->query_size, (cg::QueryKind)spec->query_kind, spec->query_seed);
        cg::WriteQueryFile(qp, q);
        cg::WriteBinaryCsr(prefix, g, spec->num_labels);
    })
}

int cgh_session_open(const char* qp, const char* dp, const char* up, uint32_t batch_ints, int threads,
                     uint32_t ebits, cgh_session** out) {
    CGH_TRY({
        auto* s = new cgh_session;
        try {
            s->s = std::make_unique<cg::Session>(qp, dp, up, batch_ints, threads, ebits);
            s->s->BuildIndex();
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    })
}

int cgh_session_open_synth(const cgh_synth_spec* spec, uint32_t batch_ints, int threads, uint32_t ebits,
                           cgh_session** out) {
    CGH_TRY({
        auto* s = new cgh_session;
        try {
            cg::ThreadPool pool(threads > 0 ? threads : cg::DefaultThreads());
            cg::SynthGraph g = cg::GenerateRmat(ToRmat(spec), &pool);
            cg::SampledQuery q = cg::SampleQuery(g, spec->query_size, (cg::QueryKind)spec->query_kind, spec->query_seed);
            s->s = std::make_unique<cg::Session>(g, q, batch_ints, threads, ebits);
            s->s->BuildIndex();
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    })
}

void cgh_session_close(cgh_session* s) { delete s; }

int cgh_query_tables(cgh_session* s, cgbdsm_plan* out) {
    CGH_TRY(cg::FillQueryTables(s->s->query(), &s->s->plan(), s->s->ebits(), out))
}

int cgh_initial_keys(cgh_session* s, const uint64_t** keys, uint64_t* n, uint32_t* dv, uint32_t* card, float* avg) {
    CGH_TRY({
        const auto& k = s->s->InitialKeys();
        *keys = k.data();
        *n = k.size();
        *dv = s->s->encoder().DV();
        uint32_t nd = 2 * s->s->query().ecount;
        if (card) memcpy(card, s->s->card(), nd * sizeof(uint32_t));
        if (avg) memcpy(avg, s->s->avg(), nd * sizeof(float));
        s->have_keys = true;
    })
}

int cgh_make_plan(cgh_session* s, const uint32_t* card, const float* avg, cgbdsm_plan* out) {
    CGH_TRY({
        s->s->MakePlan(card, avg);
        s->s->plan().Fill(out, s->s->ebits());
    })
}

uint32_t cgh_num_batches(cgh_session* s) { return s->s->num_batches(); }
uint32_t cgh_num_threads(cgh_session* s) { return (uint32_t)s->s->calig().Threads(); }

int cgh_next_batch(cgh_session* s, cgh_batch* out) {
    CGH_TRY({
        if (!s->have_keys) throw cg::Error("cgh_initial_keys must run before the first batch");
        if (s->next_batch >= s->s->num_batches()) throw cg::Error("no batches left");
        uint32_t b = s->next_batch++;
        cg::S1Timing t;
        const cg::KeyBatch& kb = s->s->RunBatch(b, &t);
        s->cursor = std::min<uint64_t>((uint64_t)(b + 1) * s->s->batch_ints(), s->s->num_updates());
        out->ins = kb.ins.data();
        out->n_ins = kb.ins.size();
        out->del = kb.del.data();
        out->n_del = kb.del.size();
        out->upd = kb.upd.data();
        out->n_upd = kb.upd.size();
        out->edge_ok = kb.edge_ok.data();
        out->dv = kb.dv;
        out->batch = b;
        out->pairs = s->s->BatchPairs(b);
        out->ms_update_index = t.update_index_ms;
        out->ms_construct_update = t.construct_update_ms;
        out->ms_prepare_all = t.prepare_all_ms;
        out->ms_prepare_index = t.prepare_index_ms;
    })
}

int cgh_export_updates(cgh_session* s, uint32_t* out, uint64_t cap, uint64_t* n_out) {
    CGH_TRY({
        const auto& u = s->s->calig().UpdateStream();
        *n_out = u.size();
        if (out)
            for (uint64_t i = 0; i < u.size() && i < cap; i++) out[i] = (uint32_t)u[i];
    })
}

uint64_t cgh_stream_left(cgh_session* s) { return s->s->num_updates() - std::min<uint64_t>(s->cursor, s->s->num_updates()); }

int cgh_next_batch_edges(cgh_session* s, uint64_t n_edges, cgh_batch* out) {
    CGH_TRY({
        if (!s->have_keys) throw cg::Error("cgh_initial_keys must run before the first batch");
        const uint64_t total = s->s->num_updates();
        if (s->cursor >= total) throw cg::Error("no stream edges left");
        if (n_edges == 0) throw cg::Error("empty batch");
        const uint64_t start = s->cursor;
        const uint64_t count = std::min<uint64_t>(2 * n_edges, total - start);
        cg::S1Timing t;
        const cg::KeyBatch& kb = s->s->RunRange((uint32_t)start, (uint32_t)count, &t);
        s->cursor = start + count;
        const uint32_t b = s->next_batch++;
        out->ins = kb.ins.data();
        out->n_ins = kb.ins.size();
        out->del = kb.del.data();
        out->n_del = kb.del.size();
        out->upd = kb.upd.data();
        out->n_upd = kb.upd.size();
        out->edge_ok = kb.edge_ok.data();
        out->dv = kb.dv;
        out->batch = b;
        out->pairs = count / 2;
        out->ms_update_index = t.update_index_ms;
        out->ms_construct_update = t.construct_update_ms;
        out->ms_prepare_all = t.prepare_all_ms;
        out->ms_prepare_index = t.prepare_index_ms;
    })
}

int cgh_old_ids(cgh_session* s, const uint32_t** ids, uint32_t* n) {
    CGH_TRY({
        *ids = s->s->encoder().OldIds().data();
        *n = (uint32_t)s->s->encoder().OldIds().size();
    })
}

int cgh_export_graph(cgh_session* s, uint64_t* offs, uint32_t* adj, uint64_t cap, int32_t* labels, uint32_t* n_out,
                     uint64_t* m_out) {
    CGH_TRY({
        auto& c = s->s->calig();
        uint32_t n = c.NumVertices();
        uint64_t m = 0;
        for (uint32_t v = 0; v < n; v++) m += c.Nei(v).size();
        *n_out = n;
        *m_out = m;
        if (offs && adj && cap >= m) {
            uint64_t p = 0;
            for (uint32_t v = 0; v < n; v++) {
                offs[v] = p;
                for (int x : c.Nei(v)) adj[p++] = (uint32_t)x;
                if (labels) labels[v] = c.Label(v);
            }
            offs[n] = p;
        }
    })
}

int cgh_export_li(cgh_session* s, uint32_t* li, uint32_t n) {
    CGH_TRY({
        auto& c = s->s->calig();
        if (n != c.NumVertices()) throw cg::Error("size mismatch");
        for (uint32_t v = 0; v < n; v++) li[v] = c.Li(v);
    })
}

int cgh_export_cand(cgh_session* s, uint32_t v, uint32_t ui, uint32_t uj, uint32_t* out, uint32_t cap, int64_t* n_out) {
    CGH_TRY({
        const auto* set = s->s->calig().Cand(v, ui, uj);
        if (!set) {
            *n_out = -1;
            return 0;
        }
        *n_out = (int64_t)set->size();
        for (uint32_t i = 0; i < set->size() && i < cap; i++) out[i] = (uint32_t)(*set)[i];
    })
}

int cgh_query_info(cgh_session* s, int32_t* labels, uint32_t* edges_uv, uint32_t* qv, uint32_t* qe) {
    CGH_TRY({
        const auto& q = s->s->query();
        *qv = q.vcount;
        *qe = q.ecount;
        if (labels)
            for (uint32_t u = 0; u < q.vcount; u++) labels[u] = q.raw_labels[u];
        if (edges_uv)
            for (uint32_t e = 0; e < q.ecount; e++) {
                edges_uv[2 * e] = q.qe_list[e].first;
                edges_uv[2 * e + 1] = q.qe_list[e].second;
            }
    })
}

}  // extern "C"

// ----------------------------------------------------------------- workloads and planners

struct cgh_workload {
    cg::SynthGraph g;
    std::unique_ptr<cg::QuerySampler> sampler;
    std::vector<uint32_t> edges;  // initial edges as u32 pairs
    std::vector<int32_t> ints;    // the stream as ints, two per pair
};

using EdgeList = std::vector<std::pair<uint32_t, uint32_t>>;

struct cgh_planner {
    cg::Query q;
    std::unique_ptr<cg::Plan> plan;
    uint32_t ebits = 5;
};

extern "C" {

int cgh_workload_open(const cgh_synth_spec* spec, cgh_workload** out) {
    CGH_TRY({
        auto* w = new cgh_workload;
        try {
            cg::ThreadPool pool(cg::DefaultThreads());
            w->g = cg::GenerateRmat(ToRmat(spec), &pool);
            w->edges.reserve(2 * w->g.initial.size());
            for (const auto& e : w->g.initial) {
                w->edges.push_back(e.first);
                w->edges.push_back(e.second);
            }
            w->ints.reserve(2 * w->g.stream.size());
            for (const auto& e : w->g.stream) {
                w->ints.push_back((int32_t)e.first);
                w->ints.push_back((int32_t)e.second);
            }
        } catch (...) {
            delete w;
            throw;
        }
        *out = w;
    })
}

void cgh_workload_close(cgh_workload* w) { delete w; }

int cgh_workload_graph(cgh_workload* w, uint32_t* n, const int32_t** labels, const uint32_t** edges, uint64_t* m) {
    CGH_TRY({
        *n = w->g.n;
        *labels = w->g.labels.data();
        *edges = w->edges.data();
        *m = w->edges.size() / 2;
    })
}

int cgh_workload_stream(cgh_workload* w, const int32_t** ints, uint64_t* n_ints) {
    CGH_TRY({
        *ints = w->ints.data();
        *n_ints = w->ints.size();
    })
}

int cgh_workload_query(cgh_workload* w, uint32_t size, uint32_t kind, uint64_t seed, int32_t* labels,
                       uint32_t* edges_uv, uint32_t* qv, uint32_t* qe) {
    CGH_TRY({
        if (!w->sampler) w->sampler = std::make_unique<cg::QuerySampler>(w->g);
        cg::SampledQuery sq = w->sampler->Sample(size, (cg::QueryKind)kind, seed);
        cg::Query q = cg::QueryFromEdges(sq.labels, sq.edges);
        *qv = q.vcount;
        *qe = q.ecount;
        for (uint32_t u = 0; u < q.vcount; u++) labels[u] = q.raw_labels[u];
        for (uint32_t e = 0; e < q.ecount; e++) {
            edges_uv[2 * e] = q.qe_list[e].first;
            edges_uv[2 * e + 1] = q.qe_list[e].second;
        }
    })
}

int cgh_planner_open(const int32_t* labels, uint32_t qv, const uint32_t* edges_uv, uint32_t qe, uint32_t ebits,
                     cgh_planner** out) {
    CGH_TRY({
        if (ebits != 5 && ebits != 7) throw cg::Error("ebits must be 5 or 7");
        if (2 * qe > (1u << ebits)) throw cg::Error("query has too many edges for the key format");
        std::vector<int> lab(labels, labels + qv);
        EdgeList ed;
        for (uint32_t e = 0; e < qe; e++) ed.push_back(std::make_pair(edges_uv[2 * e], edges_uv[2 * e + 1]));
        auto* p = new cgh_planner;
        try {
            p->q = cg::QueryFromEdges(lab, ed);
            p->plan = std::make_unique<cg::Plan>(p->q);
            p->plan->GenerateIndexingOrders();
            p->ebits = ebits;
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    })
}

void cgh_planner_close(cgh_planner* p) { delete p; }

int cgh_planner_tables(cgh_planner* p, cgbdsm_plan* out) { CGH_TRY(cg::FillQueryTables(p->q, p->plan.get(), p->ebits, out)) }

int cgh_planner_make_plan(cgh_planner* p, const uint32_t* card, const float* avg, cgbdsm_plan* out) {
    CGH_TRY({
        p->plan->GenerateMatchingOrders(card, avg);
        p->plan->Fill(out, p->ebits);
    })
}

}  // extern "C"
