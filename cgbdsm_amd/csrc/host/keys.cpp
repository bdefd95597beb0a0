#include "keys.h"

namespace cg {

KeyEncoder::KeyEncoder(const CaLiG& calig, const Query& q, uint32_t ebits) : c_(calig), q_(q), ebits_(ebits) {
    if (ebits != 5 && ebits != 7) throw Error("ebits must be 5 or 7");
    if (2 * q.ecount > (1u << ebits)) throw Error("query has too many edges for the key format");
    new_of_.assign(calig.NumVertices(), UINT32_MAX);
}

uint64_t KeyEncoder::Encode(uint32_t e, uint32_t u, uint32_t v) const {
    const uint32_t ubits = 32 - ebits_;
    if (u >= (1u << ubits)) throw Error("compressed vertex id exceeds the key's owner field");
    return ((uint64_t)e << (64 - ebits_)) | ((uint64_t)u << 32) | v;
}

void KeyEncoder::ExtendMapping() {
    const uint32_t n = c_.NumVertices();
    for (uint32_t v = 0; v < n; v++) {
        if (c_.Li(v) == 0 || new_of_[v] != UINT32_MAX) continue;
        new_of_[v] = (uint32_t)old_of_.size();
        old_of_.push_back(v);
    }
}

std::vector<uint64_t> KeyEncoder::InitialKeys(uint32_t* card, float* avg, ThreadPool& pool) {
    ExtendMapping();
    const int T = pool.size();
    const uint32_t ndir = 2 * q_.ecount;
    std::vector<std::vector<uint64_t>> local(T);
    std::vector<std::vector<uint64_t>> lcard(T, std::vector<uint64_t>(ndir, 0)), lnodes(T, std::vector<uint64_t>(ndir, 0));
    pool.For(old_of_.size(), 1024, [&](size_t b, size_t e, int tid) {
        for (size_t nu = b; nu < e; nu++) {
            uint32_t v = old_of_[nu];
            uint32_t li = c_.Li(v);
            if (li == 0) continue;
            int lb = c_.Label(v);
            for (int ui : c_.QueryVerticesWithLabel(lb)) {
                if (!(li & (1u << ui))) continue;
                for (int uj : c_.QueryNbrs(ui)) {
                    uint32_t dir = q_.E(ui, uj);
                    const auto* s = c_.Cand(v, ui, uj);
                    uint64_t cnt = 0;
                    for (int vj : *s) {
                        uint32_t nv = new_of_[vj];
                        if (nv == UINT32_MAX) continue;
                        local[tid].push_back(Encode(dir, (uint32_t)nu, nv));
                        cnt++;
                    }
                    lcard[tid][dir] += cnt;
                    lnodes[tid][dir] += cnt != 0;
                }
            }
        }
    });
    std::vector<uint64_t> keys;
    Gather(local, keys);
    for (uint32_t d = 0; d < ndir; d++) {
        uint64_t c = 0, nd = 0;
        for (int t = 0; t < T; t++) {
            c += lcard[t][d];
            nd += lnodes[t][d];
        }
        if (card) card[d] = (uint32_t)c;
        if (avg) avg[d] = nd > 0 ? (float)(uint32_t)c / (float)(uint32_t)nd : 0.0f;
    }
    return keys;
}

KeyBatch KeyEncoder::Prepare(const BatchDeltas& d, ThreadPool& pool) {
    ExtendMapping();
    KeyBatch kb;
    kb.dv = DV();
    const int T = pool.size();
    auto encode_all = [&](const std::vector<DeltaOp>& src, std::vector<uint64_t>& dst, bool need_lit) {
        std::vector<std::vector<uint64_t>> local(T);
        pool.For(src.size(), 4096, [&](size_t b, size_t e, int tid) {
            for (size_t i = b; i < e; i++) {
                const DeltaOp& x = src[i];
                if (need_lit && !(c_.Li(x.vi) & (1u << x.ui))) continue;
                uint32_t nu = NewId(x.vi), nv = NewId(x.vj);
                if (nu == UINT32_MAX || nv == UINT32_MAX) continue;
                local[tid].push_back(Encode(q_.E(x.ui, x.uj), nu, nv));
            }
        });
        Gather(local, dst);
    };
    encode_all(d.ins, kb.ins, true);
    encode_all(d.del, kb.del, false);
    encode_all(d.upd, kb.upd, false);
    std::vector<uint8_t> dir_nonempty(2 * q_.ecount, 0);
    const uint32_t shift = 64 - ebits_;
    for (uint64_t k : kb.upd) dir_nonempty[k >> shift] = 1;
    kb.edge_ok.assign(q_.ecount, 0);
    for (uint32_t e = 0; e < q_.ecount; e++)
        kb.edge_ok[e] = dir_nonempty[q_.qe_eidx[e].first] && dir_nonempty[q_.qe_eidx[e].second];
    return kb;
}

}  // namespace cg
