// cf_oracle — CPU restatement of the reference's partition + subgraph construct.
// Test infrastructure only (parity checker and CPU baseline); see cf_oracle.hpp.
#include "cf_oracle.hpp"

#include <algorithm>
#include <cstring>
#include <limits>
#include <unordered_map>

namespace cforacle {

const uint64_t kWySalt[4] = {4167021922371662411llu, 7320285940802167691llu,
                             14307255741305819987llu, 10859488101230029397llu};

// ---------------------------------------------------------------------------
// wyhash v4.2 (include/wyhash/wyhash.h:45-142): 64x64->128 "mum", xor-fold "mix".
namespace {
inline void mum(uint64_t& A, uint64_t& B) {
    const unsigned __int128 r = static_cast<unsigned __int128>(A) * B;
    A = static_cast<uint64_t>(r);
    B = static_cast<uint64_t>(r >> 64);
}
inline uint64_t mix(uint64_t A, uint64_t B) { mum(A, B); return A ^ B; }
inline uint64_t rd8(const uint8_t* p) { uint64_t v; std::memcpy(&v, p, 8); return v; }
inline uint64_t rd4(const uint8_t* p) { uint32_t v; std::memcpy(&v, p, 4); return v; }
inline uint64_t rd3(const uint8_t* p, size_t n) {
    return (uint64_t(p[0]) << 16) | (uint64_t(p[n >> 1]) << 8) | p[n - 1];
}
}  // namespace

uint64_t wyhash(const void* key, size_t len, uint64_t seed, const uint64_t* s) {
    const uint8_t* p = static_cast<const uint8_t*>(key);
    seed ^= mix(seed ^ s[0], s[1]);
    uint64_t a, b;
    if (len <= 16) {
        if (len >= 4) {
            const size_t q = (len >> 3) << 2;
            a = (rd4(p) << 32) | rd4(p + q);
            b = (rd4(p + len - 4) << 32) | rd4(p + len - 4 - q);
        } else if (len > 0) {
            a = rd3(p, len);
            b = 0;
        } else {
            a = b = 0;
        }
    } else {
        size_t i = len;
        if (i >= 48) {
            uint64_t see1 = seed, see2 = seed;
            do {
                seed = mix(rd8(p) ^ s[1], rd8(p + 8) ^ seed);
                see1 = mix(rd8(p + 16) ^ s[2], rd8(p + 24) ^ see1);
                see2 = mix(rd8(p + 32) ^ s[3], rd8(p + 40) ^ see2);
                p += 48;
                i -= 48;
            } while (i >= 48);
            seed ^= see1 ^ see2;
        }
        while (i > 16) {
            seed = mix(rd8(p) ^ s[1], rd8(p + 8) ^ seed);
            i -= 16;
            p += 16;
        }
        a = rd8(p + i - 16);
        b = rd8(p + i - 8);
    }
    a ^= s[1];
    b ^= seed;
    mum(a, b);
    return mix(a ^ s[0] ^ len, b ^ s[1]);
}

uint64_t lmer_hash(uint64_t lmer) { return wyhash(&lmer, sizeof(lmer), 0, kWySalt); }

uint64_t canon_lmer_hash(uint64_t lmer, uint64_t lmer_bar) {
    const uint64_t a = lmer_hash(lmer), b = lmer_hash(lmer_bar);
    return a < b ? a : b;
}

// ---------------------------------------------------------------------------
// DNA_Utility (include/DNA_Utility.hpp:27-41, 110-191).
bool is_dna(char c) {
    const char b = static_cast<char>(c & 0x5F);
    return b == 'A' || b == 'C' || b == 'G' || b == 'T';
}

uint8_t map_base(char c) {
    switch (c) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': return 3;
        default: return BASE_N;
    }
}

uint8_t map_base_unchecked(char c) { return uint8_t(((c >> 2) ^ (c >> 1)) & 3); }

// ---------------------------------------------------------------------------
// Parsing (RabbitFX Formater.cpp).
namespace {
template <typename F>
void for_each_line(const char* buf, size_t n, F&& f) {
    size_t pos = 0;
    while (pos < n) {
        const void* nl = std::memchr(buf + pos, '\n', n - pos);
        const size_t end = nl ? size_t(static_cast<const char*>(nl) - buf) : n;
        f(buf + pos, end - pos);
        pos = end + 1;
    }
}
}  // namespace

std::vector<std::string> parse_fasta(const char* buf, size_t n) {
    std::vector<std::string> recs;
    bool have = false;
    for_each_line(buf, n, [&](const char* s, size_t len) {
        if (len > 0 && s[0] == '>') {
            recs.emplace_back();
            have = true;
        } else if (have) {
            recs.back().append(s, len);
        }
    });
    return recs;
}

std::vector<std::string> parse_fastq(const char* buf, size_t n) {
    std::vector<std::string> recs;
    size_t line = 0;
    for_each_line(buf, n, [&](const char* s, size_t len) {
        if (line % 4 == 1) recs.emplace_back(s, len);
        ++line;
    });
    return recs;
}

std::vector<std::pair<uint64_t, uint64_t>> fragments(const char* seq, size_t seq_len, unsigned k) {
    // src/Graph_Partitioner.cpp:476-501 and the extension loop at :523.
    std::vector<std::pair<uint64_t, uint64_t>> out;
    size_t last_end = 0;
    while (true) {
        size_t beg;
        for (beg = last_end; beg + (k + 1) <= seq_len; ++beg)
            if (is_dna(seq[beg])) break;
        if (beg + (k + 1) > seq_len) break;
        size_t flen;
        for (flen = 1; flen < k + 1; ++flen)
            if (!is_dna(seq[beg + flen])) break;
        if (flen < k + 1) {
            last_end = beg + flen;
            continue;
        }
        while (beg + flen < seq_len && is_dna(seq[beg + flen])) ++flen;
        out.emplace_back(beg, flen);
        last_end = beg + flen;
    }
    return out;
}

// ---------------------------------------------------------------------------
// Min_Iterator<k-1> (include/Minimizer_Iterator.hpp:236-378).
namespace {
class MinIter {
public:
    MinIter(unsigned km1, unsigned l)
        : km1_(km1), l_(l), clear_msn_(~(uint64_t(3) << (2 * (l - 1)))), H_(km1), Mpre_(km1) {}

    void reset(const char* seq) {
        last_ = last_bar_ = 0;
        for (unsigned i = 0; i < l_; ++i) {
            const uint8_t b = map_base(seq[i]);
            last_ |= uint64_t(b) << (2 * (l_ - 1 - i));
            last_bar_ |= uint64_t(complement(b)) << (2 * i);
        }
        pivot_ = km1_ - l_;
        H_[pivot_] = canon_lmer_hash(last_, last_bar_);
        for (unsigned i = l_; i < km1_; ++i) {
            roll(map_base(seq[i]));
            H_[--pivot_] = canon_lmer_hash(last_, last_bar_);
        }
        reset_windows();
    }

    uint64_t hash() const { return std::min(Mpre_[pivot_], Msuf_); }

    void advance(char ch) {
        roll(map_base(ch));
        H_[pivot_] = canon_lmer_hash(last_, last_bar_);
        Msuf_ = std::min(Msuf_, H_[pivot_]);
        if (pivot_ > 0)
            --pivot_;
        else
            reset_windows();
    }

private:
    void roll(uint8_t b) {
        last_ = ((last_ & clear_msn_) << 2) | b;
        last_bar_ = (last_bar_ >> 2) | (uint64_t(complement(b)) << (2 * (l_ - 1)));
    }
    void reset_windows() {
        const size_t ring = km1_ - l_;
        Mpre_[0] = H_[0];
        for (size_t i = 1; i <= ring; ++i) Mpre_[i] = std::min(Mpre_[i - 1], H_[i]);
        Msuf_ = std::numeric_limits<uint64_t>::max();
        pivot_ = ring;
    }

    unsigned km1_, l_;
    uint64_t clear_msn_;
    uint64_t last_ = 0, last_bar_ = 0;
    std::vector<uint64_t> H_, Mpre_;
    uint64_t Msuf_ = 0;
    size_t pivot_ = 0;
};

uint64_t lmer_at(const char* s, unsigned l) {
    uint64_t x = 0;
    for (unsigned i = 0; i < l; ++i) x = (x << 2) | map_base(s[i]);
    return x;
}
uint64_t lmer_rc(uint64_t x, unsigned l) {
    uint64_t r = 0;
    for (unsigned i = 0; i < l; ++i) {
        r = (r << 2) | (3 - (x & 3));
        x >>= 2;
    }
    return r;
}
}  // namespace

void minimizer_hashes_literal(const char* frag, size_t frag_len, unsigned km1, unsigned l,
                              std::vector<uint64_t>& H) {
    H.clear();
    if (frag_len < km1) return;
    MinIter it(km1, l);
    it.reset(frag);
    H.push_back(it.hash());
    for (size_t i = km1; i < frag_len; ++i) {
        it.advance(frag[i]);
        H.push_back(it.hash());
    }
}

void minimizer_hashes_brute(const char* frag, size_t frag_len, unsigned km1, unsigned l,
                            std::vector<uint64_t>& H) {
    H.clear();
    if (frag_len < km1) return;
    std::vector<uint64_t> lh(frag_len - l + 1);
    for (size_t q = 0; q + l <= frag_len; ++q) {
        const uint64_t f = lmer_at(frag + q, l);
        lh[q] = canon_lmer_hash(f, lmer_rc(f, l));
    }
    const size_t w = km1 - l + 1;
    for (size_t p = 0; p + km1 <= frag_len; ++p)
        H.push_back(*std::min_element(lh.begin() + p, lh.begin() + p + w));
}

// ---------------------------------------------------------------------------
// Weak super k-mer split (src/Graph_Partitioner.cpp:503-597).
void partition_fragment(const char* frag, size_t frag_len, unsigned k, unsigned l, uint32_t G,
                        std::vector<SuperKmer>& out) {
    const size_t th = 2 * (k - 1) - l;  // sup_km1_mer_len_th (:30)
    MinIter it(k - 1, l);
    it.reset(frag);
    uint32_t cur_g = uint32_t(it.hash() & (G - 1));
    uint32_t prev_g = G;
    size_t cur_off = 0, idx = 0, fl = k - 1;
    while (fl < frag_len) {
        const size_t len = idx + (k - 1);
        it.advance(frag[fl]);
        ++idx;
        ++fl;
        const uint32_t next_g = uint32_t(it.hash() & (G - 1));
        if (next_g != cur_g || len == th) {
            const bool lj = cur_off > 0;
            const bool l_disc = lj && prev_g != cur_g;
            const bool r_disc = next_g != cur_g;
            out.push_back({cur_g, cur_off - lj, uint32_t(lj + len + 1), l_disc, r_disc});
            cur_off += idx;
            prev_g = cur_g;
            cur_g = next_g;
            idx = 0;
        }
    }
    const size_t len = fl - cur_off;
    const bool lj = cur_off > 0;
    out.push_back({cur_g, cur_off - lj, uint32_t(lj + len), lj && prev_g != cur_g, false});
}

void partition_fragment_scan(const char* frag, size_t frag_len, unsigned k, unsigned l, uint32_t G,
                             std::vector<SuperKmer>& out) {
    std::vector<uint64_t> H;
    minimizer_hashes_brute(frag, frag_len, k - 1, l, H);
    const size_t n = H.size();
    const size_t w = k - l;
    std::vector<uint32_t> g(n);
    for (size_t p = 0; p < n; ++p) g[p] = uint32_t(H[p] & (G - 1));
    // run_start[p] = max{q <= p : q == 0 || g[q] != g[q-1]}
    std::vector<size_t> starts;
    size_t run_start = 0;
    for (size_t p = 0; p < n; ++p) {
        if (p == 0 || g[p] != g[p - 1]) run_start = p;
        if ((p - run_start) % w == 0) starts.push_back(p);
    }
    for (size_t i = 0; i < starts.size(); ++i) {
        const size_t p = starts[i];
        const size_t e = (i + 1 < starts.size()) ? starts[i + 1] : n;
        const bool lj = p > 0, rj = e < n;
        const size_t len = (e - p) + (k - 2);
        out.push_back({g[p], p - lj, uint32_t(lj + len + rj), lj && g[p - 1] != g[p],
                       rj && g[e] != g[p]});
    }
}

void encode_label(const char* seq, uint32_t len, unsigned W, uint64_t* words) {
    for (unsigned j = 0; j < W; ++j) {
        uint64_t x = 0;
        for (unsigned t = 0; t < 32; ++t) {
            const size_t i = size_t(j) * 32 + t;
            const uint64_t b = i < len ? map_base_unchecked(seq[i]) : 0;
            x |= b << (62 - 2 * t);
        }
        words[W - 1 - j] = x;
    }
}

// ---------------------------------------------------------------------------
// Kmer<k> operations.
KmOps::KmOps(unsigned k_) : k(k_), nw((k_ + 31) / 32) {
    const unsigned top_bases = k - 32 * (nw - 1);
    top_mask = top_bases == 32 ? ~uint64_t(0) : ((uint64_t(1) << (2 * top_bases)) - 1);
}

bool KmOps::less(const Km& a, const Km& b) const {
    for (int i = int(nw) - 1; i >= 0; --i)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i];
    return false;
}

void KmOps::roll_fwd(Km& x, uint8_t b) const {
    for (unsigned i = nw - 1; i > 0; --i) x.w[i] = (x.w[i] << 2) | (x.w[i - 1] >> 62);
    x.w[0] = (x.w[0] << 2) | b;
    x.w[nw - 1] &= top_mask;
}

void KmOps::roll_rc(Km& x, uint8_t b) const {
    for (unsigned i = 0; i + 1 < nw; ++i) x.w[i] = (x.w[i] >> 2) | (x.w[i + 1] << 62);
    x.w[nw - 1] >>= 2;
    x.w[nw - 1] |= uint64_t(complement(b)) << (2 * ((k - 1) & 31));
}

uint8_t KmOps::base_at(const Km& x, unsigned i) const {
    const unsigned bit = 2 * (k - 1 - i);
    return uint8_t((x.w[bit / 64] >> (bit % 64)) & 3);
}

Km KmOps::rc(const Km& x) const {
    Km r;
    for (unsigned i = 0; i < k; ++i) {
        const unsigned bit = 2 * i;  // rc base i (from the end) lands at position k-1-i
        r.w[bit / 64] |= uint64_t(3 - base_at(x, i)) << (bit % 64);
    }
    return r;
}

Km KmOps::from_codes(const uint8_t* c) const {
    Km x;
    for (unsigned i = 0; i < k; ++i) roll_fwd(x, c[i]);
    return x;
}

std::string KmOps::str(const Km& x) const {
    std::string s(k, 'N');
    for (unsigned i = 0; i < k; ++i) s[i] = "ACGT"[base_at(x, i)];
    return s;
}

uint8_t edge_at(uint32_t e_f, unsigned side, unsigned f_th) {
    const unsigned off = side == 0 ? 0 : 16;
    unsigned cnt = 0, which = 0;
    for (unsigned b = 0; b < 4; ++b)
        if (((e_f >> (off + 4 * b)) & 0xF) >= f_th) ++cnt, which = b;
    return cnt == 0 ? BASE_E : (cnt == 1 ? uint8_t(which) : BASE_N);
}

// ---------------------------------------------------------------------------
// Subgraph construct.
namespace {
struct KmHash {
    size_t operator()(const Km& x) const {
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (uint64_t v : x.w) h = (h ^ v) * 0xBF58476D1CE4E5B9ull, h ^= h >> 31;
        return size_t(h);
    }
};
using Map = std::unordered_map<Km, State, KmHash>;

constexpr unsigned FRONT = 0, BACK = 1, UNSPEC = 2;

void mark_disc(State& s, unsigned side) {
    if (side != UNSPEC) s.status |= (side == FRONT ? 2 : 4);
}

void emit(const Map& M, uint32_t g, std::vector<VertexState>& out) {
    for (const auto& kv : M) out.push_back({g, kv.first, kv.second});
}
}  // namespace

void construct_literal(unsigned k, unsigned W, uint32_t g, size_t n, const uint16_t* att,
                       const uint64_t* labels, std::vector<VertexState>& out) {
    const KmOps ops(k);
    Map M;
    for (size_t r = 0; r < n; ++r) {
        const uint64_t* label = labels + r * W;
        const unsigned len = att[r] & 0xFF;
        const bool l_disc = (att[r] >> 8) & 1, r_disc = (att[r] >> 9) & 1;
        // Directed_Vertex::from_super_kmer (include/Directed_Vertex.hpp:220-225).
        Km v, vbar;
        for (unsigned i = 0; i < k; ++i) {
            const uint8_t b = label_base(label, W, i);
            ops.roll_fwd(v, b);
        }
        vbar = ops.rc(v);
        Km pred_v;
        for (size_t idx = 0;; ++idx) {
            const bool is_can = ops.less(v, vbar);  // in_canonical_form (Directed_Vertex.hpp:189-193)
            const Km& can = is_can ? v : vbar;
            const uint8_t pred = idx == 0 ? BASE_E : label_base(label, W, idx - 1);
            const uint8_t succ = idx + k == len ? BASE_E : label_base(label, W, idx + k);
            uint8_t front = is_can ? pred : complement(succ);
            uint8_t back = is_can ? succ : complement(pred);
            if (idx > 0 && can == pred_v) (is_can ? front : back) = BASE_E;
            State& st = M[can];
            if (front != BASE_E) add_edge(st.e_f, FRONT, front);
            if (back != BASE_E) add_edge(st.e_f, BACK, back);
            const unsigned entrance = is_can ? FRONT : BACK, exit_ = is_can ? BACK : FRONT;
            mark_disc(st, (idx == 0 && l_disc) ? entrance : UNSPEC);
            mark_disc(st, (idx + k == len && r_disc) ? exit_ : UNSPEC);
            if (idx + k == len) break;
            pred_v = can;
            ops.roll_fwd(v, succ);
            ops.roll_rc(vbar, succ);
        }
    }
    emit(M, g, out);
}

// ---------------------------------------------------------------------------
// The same construct with an open-addressing map in the class of the
// reference's ankerl::unordered_dense (include/Subgraph.hpp:53): linear
// probing over a power-of-two slot array, grown by doubling past 3/4 load,
// kept by a worker across subgraphs and cleared, not freed, for the next one
// (Subgraph's constructor clears the worker's map, src/Subgraph.cpp:20-40).
// Only the speed differs from construct_literal: the states are the same.
struct FlatMap::Impl {
    std::vector<Km> key;
    std::vector<State> val;
    std::vector<uint8_t> used;
    size_t n = 0, mask = 0;
};
FlatMap::FlatMap() : p(new Impl) {}
FlatMap::~FlatMap() { delete p; }
namespace {
inline size_t km_hash_cpu(const Km& x) {
    uint64_t h = x.w[0] * 0x9E3779B97F4A7C15ull;
    for (int i = 1; i < 4; ++i) h = (h ^ (h >> 29) ^ x.w[i]) * 0xBF58476D1CE4E5B9ull;
    return size_t(h ^ (h >> 32));
}
}  // namespace
void FlatMap::clear(size_t expect) {
    size_t cap = 64;
    while (cap * 3 < expect * 4) cap <<= 1;
    if (cap > p->key.size()) {
        p->key.assign(cap, Km{});
        p->val.assign(cap, State{});
        p->used.assign(cap, 0);
    } else {
        cap = p->key.size();
        std::fill(p->used.begin(), p->used.end(), 0);
    }
    p->mask = cap - 1;
    p->n = 0;
}
State& FlatMap::at(const Km& x) {
    if ((p->n + 1) * 4 > (p->mask + 1) * 3) {  // grow by doubling, reinserting
        std::vector<Km> k2;
        std::vector<State> v2;
        for (size_t i = 0; i <= p->mask; ++i)
            if (p->used[i]) k2.push_back(p->key[i]), v2.push_back(p->val[i]);
        const size_t cap = (p->mask + 1) * 2;
        p->key.assign(cap, Km{});
        p->val.assign(cap, State{});
        p->used.assign(cap, 0);
        p->mask = cap - 1;
        p->n = 0;
        for (size_t i = 0; i < k2.size(); ++i) at(k2[i]) = v2[i];
    }
    size_t s = km_hash_cpu(x) & p->mask;
    while (p->used[s]) {
        if (p->key[s] == x) return p->val[s];
        s = (s + 1) & p->mask;
    }
    p->used[s] = 1;
    p->key[s] = x;
    p->val[s] = State{};
    ++p->n;
    return p->val[s];
}
void FlatMap::emit(uint32_t g, std::vector<VertexState>& out) const {
    for (size_t i = 0; i <= p->mask; ++i)
        if (p->used[i]) out.push_back({g, p->key[i], p->val[i]});
}

void construct_flat(unsigned k, unsigned W, uint32_t g, size_t n, const uint16_t* att, const uint64_t* labels,
                    FlatMap& M, std::vector<VertexState>& out) {
    const KmOps ops(k);
    size_t inst = 0;
    for (size_t r = 0; r < n; ++r) inst += (att[r] & 0xFF) - k + 1;
    M.clear(inst / 2);
    for (size_t r = 0; r < n; ++r) {
        const uint64_t* label = labels + r * W;
        const unsigned len = att[r] & 0xFF;
        const bool l_disc = (att[r] >> 8) & 1, r_disc = (att[r] >> 9) & 1;
        Km v, vbar;
        for (unsigned i = 0; i < k; ++i) ops.roll_fwd(v, label_base(label, W, i));
        vbar = ops.rc(v);
        Km pred_v;
        for (size_t idx = 0;; ++idx) {
            const bool is_can = ops.less(v, vbar);
            const Km& can = is_can ? v : vbar;
            const uint8_t pred = idx == 0 ? BASE_E : label_base(label, W, idx - 1);
            const uint8_t succ = idx + k == len ? BASE_E : label_base(label, W, idx + k);
            uint8_t front = is_can ? pred : complement(succ);
            uint8_t back = is_can ? succ : complement(pred);
            if (idx > 0 && can == pred_v) (is_can ? front : back) = BASE_E;
            State& st = M.at(can);
            if (front != BASE_E) add_edge(st.e_f, FRONT, front);
            if (back != BASE_E) add_edge(st.e_f, BACK, back);
            const unsigned entrance = is_can ? FRONT : BACK, exit_ = is_can ? BACK : FRONT;
            mark_disc(st, (idx == 0 && l_disc) ? entrance : UNSPEC);
            mark_disc(st, (idx + k == len && r_disc) ? exit_ : UNSPEC);
            if (idx + k == len) break;
            pred_v = can;
            ops.roll_fwd(v, succ);
            ops.roll_rc(vbar, succ);
        }
    }
    M.emit(g, out);
}

void construct_closed_form(unsigned k, unsigned W, uint32_t g, size_t n, const uint16_t* att,
                           const uint64_t* labels, std::vector<VertexState>& out) {
    const KmOps ops(k), eops(k + 1);
    std::unordered_map<Km, uint64_t, KmHash> ecount;  // canonical (k+1)-mer -> count, both strands
    uint64_t homo[4] = {0, 0, 0, 0};                  // oriented b^{k+1} counts
    Map M;
    std::vector<uint8_t> s;
    for (size_t r = 0; r < n; ++r) {
        const uint64_t* label = labels + r * W;
        const unsigned len = att[r] & 0xFF;
        const bool l_disc = (att[r] >> 8) & 1, r_disc = (att[r] >> 9) & 1;
        s.resize(len);
        for (unsigned i = 0; i < len; ++i) s[i] = label_base(label, W, i);
        for (unsigned i = 0; i + k <= len; ++i) {
            const Km v = ops.from_codes(&s[i]);
            const Km vb = ops.rc(v);
            const bool is_can = ops.less(v, vb);
            State& st = M[is_can ? v : vb];
            if (i == 0 && l_disc) mark_disc(st, is_can ? FRONT : BACK);
            if (i + k == len && r_disc) mark_disc(st, is_can ? BACK : FRONT);
        }
        for (unsigned i = 0; i + k + 1 <= len; ++i) {
            bool hp = true;
            for (unsigned j = 1; j <= k && hp; ++j) hp = s[i + j] == s[i];
            if (hp) {
                ++homo[s[i]];
                continue;
            }
            const Km e = eops.from_codes(&s[i]);
            const Km eb = eops.rc(e);
            ++ecount[eops.less(eb, e) ? eb : e];
        }
    }
    auto bump = [&](const Km& x, unsigned side, uint8_t b, uint64_t c) {
        auto it = M.find(x);
        if (it == M.end()) return;  // unreachable: endpoints are inserted above
        const uint32_t off = (side == FRONT ? 0 : 16) + 4 * b;
        const uint32_t f = uint32_t(std::min<uint64_t>(15, c));
        it->second.e_f = (it->second.e_f & ~(0xFu << off)) | (f << off);
    };
    std::vector<uint8_t> ec(k + 1);
    for (const auto& kv : ecount) {
        for (unsigned i = 0; i <= k; ++i) ec[i] = eops.base_at(kv.first, i);
        const Km u = ops.from_codes(&ec[0]), v = ops.from_codes(&ec[1]);
        const Km ub = ops.rc(u), vb = ops.rc(v);
        const bool ucan = ops.less(u, ub), vcan = ops.less(v, vb);
        const Km& uc = ucan ? u : ub;
        const Km& vc = vcan ? v : vb;
        // u's succ-derived slot.
        if (ucan) bump(uc, BACK, ec[k], kv.second);
        else bump(uc, FRONT, complement(ec[k]), kv.second);
        if (uc == vc) continue;  // palindromic (k+1)-mer: one slot only
        if (vcan) bump(vc, FRONT, ec[0], kv.second);
        else bump(vc, BACK, complement(ec[0]), kv.second);
    }
    for (uint8_t b = 0; b < 4; ++b) {
        if (!homo[b]) continue;
        std::vector<uint8_t> hb(k, b);
        const Km h = ops.from_codes(hb.data());
        const Km hbar = ops.rc(h);
        if (ops.less(h, hbar)) bump(h, BACK, b, homo[b]);
        else bump(hbar, FRONT, complement(b), homo[b]);
    }
    emit(M, g, out);
}

void sort_states(unsigned k, std::vector<VertexState>& v) {
    const KmOps ops(k);
    std::sort(v.begin(), v.end(), [&](const VertexState& a, const VertexState& b) {
        if (a.g != b.g) return a.g < b.g;
        return ops.less(a.key, b.key);
    });
}

}  // namespace cforacle
