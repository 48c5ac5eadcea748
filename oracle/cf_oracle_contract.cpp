// Subgraph::contract restated on the oracle's vertex tables (test
// infrastructure only; see the header of cf_oracle.hpp).
//
// A literal restatement of the reference's per-subgraph unitig extraction:
//   Subgraph::contract            src/Subgraph.cpp:184-270
//   extract_maximal_unitig        include/Subgraph.hpp:414-467
//   walk_unitig                   include/Subgraph.hpp:470-543
//   Unitig_Scratch init/extend    include/Unitig_Scratch.hpp:106-146, 160-166
//   Maximal_Unitig_Scratch        include/Maximal_Unitig_Scratch.hpp:136-138, 178-330
//   cycle rotation                include/FASTA_Record.hpp:142-146
// Walks start from the vertices in the order of a hash of their keys, as the
// reference iterates its hash map (the extracted set does not depend on it).
//
// Pinning: the lm-tigs of all subgraphs, glued at their shared discontinuity
// vertices, and the trivial unitigs must give the maximal unitigs of the
// brute-force global graph (tests/unitig_ref.py), and the discontinuity-edge
// endpoints the walk reports must match those the normalised label implies.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "cf_oracle.hpp"

namespace cforacle {
namespace {

constexpr unsigned FRONT = 0, BACK = 1, PHI = 2;

struct KmHashC {
    size_t operator()(const Km& x) const {
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (uint64_t v : x.w) h = (h ^ v) * 0xBF58476D1CE4E5B9ull, h ^= h >> 31;
        return size_t(h);
    }
};

// Directed_Vertex (include/Directed_Vertex.hpp:189-193, 257-261, 302-320).
struct DV {
    Km kmer, bar;
    bool can;  // in_canonical_form
    const Km& canonical() const { return can ? kmer : bar; }
    unsigned entrance_side() const { return can ? FRONT : BACK; }
};

DV make_dv(const KmOps& ops, const Km& x) {
    DV v{x, ops.rc(x), false};
    v.can = ops.less(v.kmer, v.bar);
    return v;
}

void roll_forward(const KmOps& ops, DV& v, uint8_t b) {
    ops.roll_fwd(v.kmer, b);
    ops.roll_rc(v.bar, b);
    v.can = ops.less(v.kmer, v.bar);
}

struct Vst {
    uint32_t e_f;
    uint8_t status;  // 1 visited, 2 front-disc, 4 back-disc
};

// Unitig_Scratch: label in walk orientation, first and last vertex, minimum vertex.
struct Scratch {
    std::string label;
    DV anchor, endpoint, min_v;
    size_t idx = 0, min_idx = 0;
    bool cycle = false;
    size_t size() const { return idx + 1; }
};

std::string revcomp(const std::string& s) {
    std::string r(s.rbegin(), s.rend());
    for (char& c : r) c = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
    return r;
}

enum Term { NONE, BRANCHED, DEAD, EXITTED, CROSSED };

class Contractor {
public:
    Contractor(unsigned k, unsigned f_th, const VertexState* v, size_t n)
        : k_(k), f_th_(f_th), ops_(k) {
        M_.reserve(n * 2);
        order_.reserve(n);
        for (size_t i = 0; i < n; ++i) {
            M_[v[i].key] = Vst{v[i].st.e_f, uint8_t(v[i].st.status & 6)};
            order_.push_back(v[i].key);
        }
        // hash order, like iterating the reference's hash map: walks then start
        // anywhere inside a unitig, not only at its smallest vertex
        const KmHashC hs;
        std::sort(order_.begin(), order_.end(), [&](const Km& a, const Km& b) {
            const size_t ha = hs(a), hb = hs(b);
            return ha != hb ? ha < hb : ops_.less(a, b);
        });
    }

    void run(ContractResult& out) {
        out = ContractResult{};
        for (const Km& v : order_) {
            Vst& st = M_.at(v);
            if (st.status & 1) continue;
            const bool iso = edge_at(st.e_f, FRONT, f_th_) == BASE_E && edge_at(st.e_f, BACK, f_th_) == BASE_E;
            if (iso) {  // Subgraph.cpp:206-213
                if (st.status & 6) ++out.phantom;
                ++out.isolated;
                continue;
            }
            extract(v, out);
        }
    }

private:
    bool branching(uint32_t e_f, unsigned s) const { return edge_at(e_f, s, f_th_) == BASE_N; }

    // walk_unitig (include/Subgraph.hpp:470-543).
    Term walk(const Km& v_hat, unsigned s_hat, Scratch& u, DV& exit_v) {
        const unsigned s_icc = 1 - s_hat;
        DV v = make_dv(ops_, s_hat == BACK ? v_hat : ops_.rc(v_hat));
        unsigned s = s_hat;
        Vst* it = &M_.at(v.canonical());
        Vst state = *it;
        u.label = ops_.str(v.kmer);
        u.anchor = u.endpoint = u.min_v = v;
        u.idx = u.min_idx = 0;
        u.cycle = false;
        while (true) {
            it->status |= 1;  // mark_visited
            uint8_t b = edge_at(state.e_f, s, f_th_);
            if (b == BASE_N) return BRANCHED;
            if (b == BASE_E) {
                if (!(state.status & (s == FRONT ? 2 : 4))) return DEAD;
                exit_v = v;
                return EXITTED;
            }
            if (s == FRONT) b = complement(b);
            roll_forward(ops_, v, b);
            it = &M_.at(v.canonical());
            state = *it;
            s = v.entrance_side();
            if (branching(state.e_f, s)) return CROSSED;
            if (state.status & 1) {
                if (v.canonical() == v_hat && s == s_icc) u.cycle = true;
                return CROSSED;
            }
            // Unitig_Scratch::extend
            u.endpoint = v;
            ++u.idx;
            if (ops_.less(v.canonical(), u.min_v.canonical())) u.min_v = v, u.min_idx = u.idx;
            u.label.push_back("ACGT"[b]);
            s = 1 - s;
        }
    }

    // extract_maximal_unitig (include/Subgraph.hpp:414-467) with the output
    // orientation of Maximal_Unitig_Scratch::finalize / finalize_weak.
    void extract(const Km& v_hat, ContractResult& out) {
        Scratch ub, uf;
        DV vl, vr;
        Term tl = NONE;
        const Term tr = walk(v_hat, BACK, ub, vr);
        if (!ub.cycle) tl = walk(v_hat, FRONT, uf, vl);
        Unitig r;
        r.n_vertices = ub.cycle ? ub.size() : ub.size() + uf.size() - 1;
        out.vertex_count += r.n_vertices;
        out.label_sz += r.n_vertices + k_ - 1;
        if (tl == EXITTED || tr == EXITTED) {
            // finalize_weak: the front walk reversed, label in v_hat's orientation
            r.label = revcomp(uf.label) + ub.label.substr(k_);
            r.u_side = tl == EXITTED ? vl.entrance_side() : PHI;
            r.v_side = tr == EXITTED ? vr.entrance_side() : PHI;
            if (tl == EXITTED) r.u = vl.canonical();
            if (tr == EXITTED) r.v = vr.canonical();
            r.flags = (tl == EXITTED ? uint32_t(UNITIG_LEFT_EXIT) : 0u) | (tr == EXITTED ? uint32_t(UNITIG_RIGHT_EXIT) : 0u);
            ++out.disc_edges;
        } else if (!ub.cycle) {
            // finalize: canonical iff front endpoint's kmer_bar < back endpoint's kmer_bar
            const bool canon = ops_.less(uf.endpoint.bar, ub.endpoint.bar);
            const std::string u = revcomp(uf.label) + ub.label.substr(k_);
            r.label = canon ? u : revcomp(u);
            ++out.trivial;
        } else {
            std::string lab = ub.label;
            size_t pivot = ub.min_idx;
            if (!ub.min_v.can) lab = revcomp(lab), pivot = ub.size() - 1 - pivot;
            r.label = lab.substr(pivot) + lab.substr(k_ - 1, pivot);
            r.flags = UNITIG_CYCLE;
            ++out.trivial;
            ++out.icc;
        }
        out.unitigs.push_back(std::move(r));
    }

    unsigned k_, f_th_;
    KmOps ops_;
    std::unordered_map<Km, Vst, KmHashC> M_;
    std::vector<Km> order_;
};

}  // namespace

void contract_subgraph(unsigned k, unsigned f_th, const VertexState* v, size_t n, ContractResult& out) {
    Contractor c(k, f_th, v, n);
    c.run(out);
}

// Orient an lm-tig like a trivial linear unitig: the first k characters are at
// most those of the reverse complement; the ends swap with the orientation.
void normalise_unitig(unsigned k, Unitig& u) {
    if (u.flags & UNITIG_CYCLE) return;
    const std::string r = revcomp(u.label);
    if (u.label.compare(0, k, r, 0, k) <= 0) return;
    u.label = r;
    std::swap(u.u, u.v);
    std::swap(u.u_side, u.v_side);
    const uint32_t l = u.flags & UNITIG_LEFT_EXIT, rr = u.flags & UNITIG_RIGHT_EXIT;
    u.flags = (u.flags & ~(UNITIG_LEFT_EXIT | UNITIG_RIGHT_EXIT)) | (l ? uint32_t(UNITIG_RIGHT_EXIT) : 0u) |
              (rr ? uint32_t(UNITIG_LEFT_EXIT) : 0u);
}

}  // namespace cforacle
