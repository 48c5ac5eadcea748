// Global maximal unitigs (the compacted de Bruijn graph) from the per-subgraph
// output of cfgpu_contract.
//
// Every trivial maximal unitig (no end at a discontinuity) is already a global
// maximal unitig in the reference's output form.  Each lm-tig ends, on an
// exiting end, at a discontinuity vertex whose other side belongs to another
// subgraph; that subgraph's lm-tig through the vertex (if the vertex is not
// isolated there) ends at the same vertex on the opposite side.  Gluing lm-tigs
// at these shared (vertex, side) pairs is the contraction of the discontinuity
// graph and its expansion back into unitig labels
// (src/Discontinuity_Graph_Contractor.cpp, src/Contracted_Graph_Expander.cpp);
// a chain that ends at a vertex with no lm-tig on the far side ends there (the
// phantom edges, src/Discontinuity_Graph_Contractor.cpp:131-190).  Chains are
// written as Unitig_Collator writes maximal unitigs
// (include/Unitig_Collator.hpp:306-345 linear, :366-400 cycles).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cfgpu.h"
#include "../../include/cfhost.h"

struct cfhost_cdbg {
    std::string seq;
    std::vector<uint64_t> off;  // unitig i = seq[off[i], off[i+1])
};

namespace {

inline char comp(char c) { return c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A'; }

std::string revcomp(const char* s, size_t n) {
    std::string r(n, 'A');
    for (size_t i = 0; i < n; ++i) r[n - 1 - i] = comp(s[i]);
    return r;
}

inline uint64_t code(char c) { return c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : 3; }

// A k-mer (k <= 127) as 2-bit codes, first base most significant: word 0 holds
// bases 0..31, so comparing words in order compares the strings.
struct Key {
    uint64_t w[4];
    bool operator<(const Key& o) const {
        for (int i = 0; i < 4; ++i)
            if (w[i] != o.w[i]) return w[i] < o.w[i];
        return false;
    }
    bool operator==(const Key& o) const { return std::memcmp(w, o.w, sizeof w) == 0; }
};

Key encode(const char* s, uint32_t k) {
    Key x{{0, 0, 0, 0}};
    for (uint32_t i = 0; i < k; ++i) x.w[i >> 5] |= code(s[i]) << (62 - 2 * (i & 31));
    return x;
}

Key encode_rc(const char* s, uint32_t k) {
    Key x{{0, 0, 0, 0}};
    for (uint32_t i = 0; i < k; ++i) x.w[i >> 5] |= (3 - code(s[k - 1 - i])) << (62 - 2 * (i & 31));
    return x;
}

// The end of an lm-tig at a discontinuity vertex: the vertex's canonical
// k-mer and the side facing the lm-tig (0 front, 1 back).
struct End {
    Key key;
    uint32_t side;
    uint32_t lm;   // lm-tig index
    uint32_t at;   // 0 its first k-mer, 1 its last
};

// Maximal_Unitig_Scratch orientation of a linear unitig: its first k characters
// at most those of the reverse complement (include/Unitig_Collator.hpp:306-345).
std::string canonical_linear(std::string s, uint32_t k) {
    const std::string r = revcomp(s.data(), s.size());
    return std::memcmp(s.data(), r.data(), k) <= 0 ? s : r;
}

// A cycle of n vertices given as its n + k - 1 character label (the last k - 1
// repeat the first): rotated to start at its smallest vertex in canonical form
// (include/Unitig_Collator.hpp:366-400; FASTA_Record::append_rotated_cycle).
std::string canonical_cycle(const std::string& lab, uint32_t k) {
    const size_t n = lab.size() - (k - 1);
    Key best{};
    size_t best_p = 0;
    bool best_rc = false, have = false;
    for (size_t p = 0; p < n; ++p) {
        const Key f = encode(lab.data() + p, k), r = encode_rc(lab.data() + p, k);
        const bool use_rc = r < f;
        const Key& c = use_rc ? r : f;
        if (!have || c < best) best = c, best_p = p, best_rc = use_rc, have = true;
    }
    std::string src = lab;
    size_t p = best_p;
    if (best_rc) src = revcomp(lab.data(), lab.size()), p = n - 1 - best_p;
    return src.substr(p) + src.substr(k - 1, p);
}

}  // namespace

extern "C" int cfhost_cdbg_build(uint32_t k, const cfgpu_unitig* recs, uint64_t n_recs, const char* labels,
                                 uint64_t n_bytes, int sorted, cfhost_cdbg** out) {
    if (!out || k < 3 || k > 127 || !(k & 1) || (n_recs && (!recs || !labels))) return -1;
    std::vector<std::string> res;
    std::vector<uint64_t> lm;  // indices of lm-tigs in recs
    for (uint64_t i = 0; i < n_recs; ++i) {
        const cfgpu_unitig& u = recs[i];
        if (u.label_off + u.label_len > n_bytes || u.label_len < k) return -1;
        if (u.flags & (CFGPU_UNITIG_LEFT_EXIT | CFGPU_UNITIG_RIGHT_EXIT))
            lm.push_back(i);
        else
            res.emplace_back(labels + u.label_off, u.label_len);
    }
    // exiting ends, sorted by vertex: the two sides of one vertex pair up
    std::vector<End> ends;
    for (uint32_t j = 0; j < lm.size(); ++j) {
        const cfgpu_unitig& u = recs[lm[j]];
        const char* s = labels + u.label_off;
        for (uint32_t at = 0; at < 2; ++at) {
            if (!(u.flags & (at ? CFGPU_UNITIG_RIGHT_EXIT : CFGPU_UNITIG_LEFT_EXIT))) continue;
            const char* x = at ? s + u.label_len - k : s;
            const Key f = encode(x, k), r = encode_rc(x, k);
            const bool can = !(r < f);
            // first k-mer: the lm-tig continues out of its back side when it is
            // canonical; last k-mer: out of its front side
            const uint32_t side = at ? (can ? 0u : 1u) : (can ? 1u : 0u);
            ends.push_back(End{can ? f : r, side, j, at});
        }
    }
    std::sort(ends.begin(), ends.end(), [](const End& a, const End& b) {
        return a.key == b.key ? a.side < b.side : a.key < b.key;
    });
    // partner[(j, at)] = the (lm-tig, end) across the vertex, or -1
    std::vector<int64_t> partner(2 * lm.size(), -1);
    for (size_t e = 0; e < ends.size();) {
        size_t f = e + 1;
        while (f < ends.size() && ends[f].key == ends[e].key) ++f;
        if (f - e > 2 || (f - e == 2 && ends[e].side == ends[e + 1].side)) return -2;  // inconsistent input
        if (f - e == 2) {
            partner[2 * ends[e].lm + ends[e].at] = int64_t(2 * ends[e + 1].lm + ends[e + 1].at);
            partner[2 * ends[e + 1].lm + ends[e + 1].at] = int64_t(2 * ends[e].lm + ends[e].at);
        }
        e = f;
    }
    std::vector<uint8_t> used(lm.size(), 0);
    auto label = [&](uint32_t j) { return std::string(labels + recs[lm[j]].label_off, recs[lm[j]].label_len); };
    auto oriented = [&](uint32_t j, uint32_t at) {  // lm-tig j read from end `at`
        std::string s = label(j);
        return at ? revcomp(s.data(), s.size()) : s;
    };
    // follow the chain entered at (j, at); returns true if it closed on itself
    auto follow = [&](uint32_t j, uint32_t at, std::string& s) {
        s = oriented(j, at);
        used[j] = 1;
        uint32_t cur = j, far = 1 - at;
        for (;;) {
            const int64_t p = partner[2 * cur + far];
            if (p < 0) return false;
            const uint32_t nj = uint32_t(p >> 1), nat = uint32_t(p & 1);
            if (nj == j) return true;
            if (used[nj]) return false;  // cannot happen for consistent input
            used[nj] = 1;
            const std::string t = oriented(nj, nat);
            s.append(t, k, std::string::npos);
            cur = nj;
            far = 1 - nat;
        }
    };
    std::string s;
    for (uint32_t j = 0; j < lm.size(); ++j) {
        if (used[j]) continue;
        for (uint32_t at = 0; at < 2; ++at)
            if (partner[2 * j + at] < 0) {  // a chain end
                follow(j, at, s);
                res.push_back(canonical_linear(s, k));
                break;
            }
    }
    for (uint32_t j = 0; j < lm.size(); ++j) {
        if (used[j]) continue;
        if (!follow(j, 0, s)) return -2;
        s.pop_back();  // the first k-mer closed the cycle: keep its k - 1 overlap
        res.push_back(canonical_cycle(s, k));
    }
    if (sorted) std::sort(res.begin(), res.end());
    auto* c = new cfhost_cdbg;
    uint64_t total = 0;
    for (const auto& r : res) total += r.size();
    c->seq.reserve(total);
    c->off.reserve(res.size() + 1);
    c->off.push_back(0);
    for (const auto& r : res) {
        c->seq += r;
        c->off.push_back(c->seq.size());
    }
    *out = c;
    return 0;
}

extern "C" int cfhost_cdbg_get(const cfhost_cdbg* c, const char** seq, const uint64_t** off, uint64_t* n) {
    if (!c) return -1;
    if (seq) *seq = c->seq.data();
    if (off) *off = c->off.data();
    if (n) *n = c->off.size() - 1;
    return 0;
}

extern "C" int cfhost_cdbg_write_fasta(const cfhost_cdbg* c, const char* path) {
    if (!c || !path) return -1;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    for (size_t i = 0; i + 1 < c->off.size(); ++i) {
        std::fprintf(f, ">%zu\n", i);
        std::fwrite(c->seq.data() + c->off[i], 1, c->off[i + 1] - c->off[i], f);
        std::fputc('\n', f);
    }
    return std::fclose(f) == 0 ? 0 : -1;
}

extern "C" void cfhost_cdbg_free(cfhost_cdbg* c) { delete c; }
