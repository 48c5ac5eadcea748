// Maximal unitigs from the device-built subgraph tables (host side).
//
// Stitching.  A vertex whose two (k-1)-mers fall in the same subgraph has one
// table entry holding both sides.  Otherwise it has one entry in each of the two
// subgraphs, and each entry marks as discontinuous the side whose (k-1)-mer
// belongs to the other subgraph (the record boundary at a g change sets r_disc /
// l_disc, src/Graph_Partitioner.cpp:552-588; src/Subgraph.cpp:88-92).  Each
// side's counters are therefore taken from the entry that does not mark it.
//
// Walk.  Subgraph::walk_unitig / extract_maximal_unitig (include/Subgraph.hpp:414-543)
// applied to the whole graph: a walk leaves v through side s if edge_at(s) is a
// single base, enters u, stops before u if u's entrance side branches, and stops
// at a visited u (a cycle when u is the start, re-entered through the other side).
// Without subgraph boundaries no walk "exits", so every unitig is complete.
//
// Output orientation follows Unitig_Collator::Maximal_Unitig::canonicalize
// (include/Unitig_Collator.hpp:306-345) for linear unitigs and
// canonicalize_cycle (:366-400) for cycles.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cfgpu.h"

namespace {

struct Key {
    uint64_t w[4];
};

struct Vtx {
    uint64_t w[4];
    uint32_t e_front = 0, e_back = 0;  // counters of one side, low 16 bits
    std::atomic<uint8_t> visited{0};
};

class Graph {
public:
    Graph(unsigned k, unsigned f_th) : k_(k), kw_((k + 31) / 32), f_th_(f_th) {
        const unsigned kt = k - 32 * (kw_ - 1);
        top_mask_ = kt == 32 ? ~0ull : ((1ull << (2 * kt)) - 1);
    }

    unsigned k() const { return k_; }

    // add one table entry (keys in Kmer<k> word order)
    void add(const uint64_t* key, uint32_t e_f, uint8_t status) {
        Entry e;
        std::memset(e.w, 0, sizeof e.w);
        std::memcpy(e.w, key, kw_ * 8);
        e.e_f = e_f;
        e.status = status;
        entries_.push_back(e);
    }

    // Merge duplicate keys and build the open-addressing index.
    void finalize() {
        std::sort(entries_.begin(), entries_.end(), [&](const Entry& a, const Entry& b) { return less(a.w, b.w); });
        std::vector<Vtx> v;
        verts_ = std::vector<Vtx>(count_distinct());
        size_t j = 0;
        for (size_t i = 0; i < entries_.size();) {
            size_t e = i + 1;
            while (e < entries_.size() && equal(entries_[e].w, entries_[i].w)) ++e;
            Vtx& x = verts_[j++];
            std::memcpy(x.w, entries_[i].w, sizeof x.w);
            for (size_t t = i; t < e; ++t) {
                if (!(entries_[t].status & 2)) x.e_front |= entries_[t].e_f & 0xFFFFu;
                if (!(entries_[t].status & 4)) x.e_back |= entries_[t].e_f >> 16;
            }
            i = e;
        }
        entries_.clear();
        entries_.shrink_to_fit();
        uint64_t cap = 16;
        while (cap < 2 * verts_.size()) cap <<= 1;
        index_.assign(cap, UINT32_MAX);
        for (uint32_t i = 0; i < verts_.size(); ++i) {
            uint64_t s = hash(verts_[i].w) & (cap - 1);
            while (index_[s] != UINT32_MAX) s = (s + 1) & (cap - 1);
            index_[s] = i;
        }
    }

    size_t size() const { return verts_.size(); }

    // Maximal unitigs as canonical strings.  Walks run one after another, as
    // the reference's per-subgraph contract() does: concurrent walkers that
    // start inside the same unitig would each stop at the other's vertices.
    std::vector<std::string> unitigs() {
        std::vector<std::string> all;
        for (size_t i = 0; i < verts_.size(); ++i) {
            uint8_t z = 0;
            if (!verts_[i].visited.compare_exchange_strong(z, 1)) continue;
            all.push_back(extract(i));
        }
        return all;
    }

private:
    struct Entry {
        uint64_t w[4];
        uint32_t e_f;
        uint8_t status;
    };

    bool less(const uint64_t* a, const uint64_t* b) const {
        for (int i = int(kw_) - 1; i >= 0; --i)
            if (a[i] != b[i]) return a[i] < b[i];
        return false;
    }
    bool equal(const uint64_t* a, const uint64_t* b) const {
        for (unsigned i = 0; i < kw_; ++i)
            if (a[i] != b[i]) return false;
        return true;
    }
    size_t count_distinct() const {
        size_t n = 0;
        for (size_t i = 0; i < entries_.size(); ++i)
            if (i == 0 || !equal(entries_[i].w, entries_[i - 1].w)) ++n;
        return n;
    }
    uint64_t hash(const uint64_t* x) const {
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (unsigned w = 0; w < kw_; ++w) h = (h ^ x[w]) * 0xBF58476D1CE4E5B9ull, h ^= h >> 29;
        return h;
    }
    int64_t find(const uint64_t* x) const {
        const uint64_t cap = index_.size();
        for (uint64_t s = hash(x) & (cap - 1);; s = (s + 1) & (cap - 1)) {
            const uint32_t i = index_[s];
            if (i == UINT32_MAX) return -1;
            if (equal(verts_[i].w, x)) return i;
        }
    }

    // --- k-mer helpers (Kmer<k> layout, base 0 most significant)
    uint8_t base_at(const uint64_t* x, unsigned i) const {
        const unsigned bit = 2 * (k_ - 1 - i);
        return uint8_t((x[bit / 64] >> (bit % 64)) & 3);
    }
    void roll(uint64_t* x, uint8_t b) const {  // drop first base, append b
        for (unsigned i = kw_ - 1; i > 0; --i) x[i] = (x[i] << 2) | (x[i - 1] >> 62);
        x[0] = (x[0] << 2) | b;
        x[kw_ - 1] &= top_mask_;
    }
    void rc(const uint64_t* x, uint64_t* r) const {
        uint64_t t[4] = {0, 0, 0, 0};
        for (unsigned i = 0; i < k_; ++i) {
            const unsigned bit = 2 * i;
            t[bit / 64] |= uint64_t(3 - base_at(x, i)) << (bit % 64);
        }
        std::memcpy(r, t, sizeof t);
    }
    // edge_at (include/State_Config.hpp:232-252) on one side's 16 counter bits
    uint8_t edge_at(uint32_t side) const {
        unsigned c = 0, which = 0;
        for (unsigned b = 0; b < 4; ++b)
            if (((side >> (4 * b)) & 0xF) >= f_th_) ++c, which = b;
        return c == 0 ? 5 : (c == 1 ? uint8_t(which) : 4);
    }
    unsigned edge_count(uint32_t side) const {
        unsigned c = 0;
        for (unsigned b = 0; b < 4; ++b) c += ((side >> (4 * b)) & 0xF) >= f_th_;
        return c;
    }

    // Directed walk from canonical vertex i through side s (0 front, 1 back):
    // appends the bases of the walk to `ext`; returns true if it closed a cycle.
    bool walk(size_t i, unsigned s0, std::string& ext) {
        uint64_t v[4], vb[4];
        std::memcpy(v, verts_[i].w, sizeof v);
        if (s0 == 0) {  // walking out of the front side = walking the reverse complement forward
            rc(verts_[i].w, vb);
            std::memcpy(v, vb, sizeof v);
        }
        rc(v, vb);
        unsigned s = s0;
        size_t cur = i;
        while (true) {
            const Vtx& x = verts_[cur];
            uint8_t b = edge_at(s == 0 ? x.e_front : x.e_back);
            if (b >= 4) return false;  // branching or dead end
            if (s == 0) b = uint8_t(3 - b);
            roll(v, b);
            // rc of the rolled observed k-mer
            for (unsigned t = 0; t + 1 < kw_; ++t) vb[t] = (vb[t] >> 2) | (vb[t + 1] << 62);
            vb[kw_ - 1] = (vb[kw_ - 1] >> 2) | (uint64_t(3 - b) << (2 * ((k_ - 1) & 31)));
            const bool can = less(v, vb);
            const int64_t u = find(can ? v : vb);
            if (u < 0) return false;  // cannot happen for a consistent table set
            const unsigned ent = can ? 0 : 1;  // entrance side of the observed k-mer
            const Vtx& y = verts_[size_t(u)];
            if (edge_count(ent == 0 ? y.e_front : y.e_back) > 1) return false;  // crossed
            uint8_t z = 0;
            if (!verts_[size_t(u)].visited.compare_exchange_strong(z, 1))
                return size_t(u) == i && ent == (s0 == 1 ? 0u : 1u);  // cycle back to the start
            ext.push_back("ACGT"[b]);
            cur = size_t(u);
            s = 1 - ent;
        }
    }

    std::string label(size_t i) const {
        std::string s(k_, 'A');
        for (unsigned t = 0; t < k_; ++t) s[t] = "ACGT"[base_at(verts_[i].w, t)];
        return s;
    }

    static char comp(char c) { return c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A'; }
    static std::string revcomp(const std::string& s) {
        std::string r(s.rbegin(), s.rend());
        for (char& c : r) c = comp(c);
        return r;
    }

    std::string extract(size_t i) {
        std::string back, front;
        const bool cyc = walk(i, 1, back);
        if (cyc) return canonical_cycle(label(i) + back);
        walk(i, 0, front);
        // front walk went leftwards along the canonical label: prepend its rc
        std::string u = revcomp(front) + label(i) + back;
        // Unitig_Collator::Maximal_Unitig::canonicalize: compare the first k
        // characters with the reverse complement's.
        const std::string r = revcomp(u);
        for (unsigned t = 0; t < k_; ++t) {
            if (u[t] < r[t]) return u;
            if (u[t] > r[t]) return r;
        }
        return u;
    }

    // canonicalize_cycle: rotate so the minimum k-mer over both strands leads.
    std::string canonical_cycle(const std::string& lab) const {
        const size_t sz = lab.size();         // vertices + k - 1
        const size_t n = sz - (k_ - 1);       // distinct rotation starts
        const std::string cyc = lab.substr(0, n);
        std::string best;
        for (int strand = 0; strand < 2; ++strand) {
            const std::string c = strand ? revcomp(cyc) : cyc;
            for (size_t st = 0; st < n; ++st) {
                std::string r = c.substr(st) + c.substr(0, st);
                r += r.substr(0, k_ - 1);
                if (best.empty() || r.compare(0, k_, best, 0, k_) < 0 ||
                    (r.compare(0, k_, best, 0, k_) == 0 && r < best))
                    best = r;
            }
        }
        return best;
    }

    unsigned k_, kw_, f_th_;
    uint64_t top_mask_;
    std::vector<Entry> entries_;
    std::vector<Vtx> verts_;
    std::vector<uint32_t> index_;
};

}  // namespace
