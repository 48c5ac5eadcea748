// Subgraph::contract (src/Subgraph.cpp:184-270) on the device-built subgraph
// tables.
//
// The reference walks each unvisited vertex's unitig in both directions
// (extract_maximal_unitig / walk_unitig, include/Subgraph.hpp:414-543).  The
// walk leaves a vertex through side s when edge_at(s) is one base, and stops
// before the neighbour when the neighbour's entrance side branches, when the
// neighbour is already visited, or at an empty side (an exit when the side is
// discontinuous).  Edges are counted at both endpoints in the same subgraph, so
// "v continues through s to u" holds exactly when u continues back to v; the
// unitigs are the paths and cycles of this symmetric relation.  A subgraph is
// processed by one workgroup:
//
//   load    the vertices in no edge at the cutoff are counted (isolated, and
//           phantom when discontinuous, src/Subgraph.cpp:206-214); the others
//           go to an open-addressing table with their DFA byte;
//   link    each side that continues to a neighbour is marked (one lookup per
//           side; the edge back through the same side, a palindromic
//           (k+1)-mer, ends the unitig there);
//   paths   LDS tables (k_contract_rank): list ranking by pointer jumping over
//           the directed states (vertex, exit side) gives every vertex both
//           ends of its path and its distance to each, in log2(length) rounds;
//           the path is written in the orientation whose first k-mer is below
//           its reverse complement's (include/Maximal_Unitig_Scratch.hpp:
//           136-138, 198-217), each vertex storing its own base.
//           Dense tables (k_contract_big, subgraphs too large for LDS): the
//           same ranking over vertex arrays in global scratch, the key ->
//           vertex index held in LDS.
//           Global tables (k_contract_walk, larger still): a thread walks
//           from every path end to the other, and the end whose orientation
//           is canonical writes the path in a second walk;
//   cycles  what no path holds lies on cycles; each is walked once and written
//           from its smallest vertex in canonical form
//           (Maximal_Unitig_Scratch::finalize, :213-219; FASTA_Record.hpp:142-146).
//
// Labels are ASCII in that orientation: a trivial unitig exactly as the
// reference prints it, an lm-tig with the ends swapped when its walk ran the
// other way.  The flags give the ends that exit through a discontinuity
// (the discontinuity-edge endpoints of Discontinuity_Graph::add_edge,
// include/Discontinuity_Graph.hpp:144-152).
#include "cf_kernels.hpp"

namespace cfgpu {
namespace {

constexpr uint64_t EMPTY = ~0ull;
// state word: bits 0-7 the DFA byte (edge_at front | back << 3 | front-disc << 6 | back-disc << 7)
constexpr uint32_t M_LINK0 = 1u << 8;    // the front side continues the unitig (bit 9: the back side)
constexpr uint32_t M_VIS = 1u << 10;     // on a path walked from one of its ends
constexpr uint32_t M_WRITER = 1u << 11;  // writes its unitig (label offset in aux / the ranking word)
constexpr uint32_t M_CYCLE = 1u << 12;   // ... a cycle, from this vertex out of its back side
constexpr uint32_t M_START1 = 1u << 13;  // ... a path whose end side at this vertex is the back
constexpr uint32_t M_ONCYC = 1u << 14;   // lies on a cycle
constexpr int CT_CAND = 512;

template <int KW>
struct CtTab {
    uint64_t* key;  // cap * KW words; word KW-1 == EMPTY marks a free slot
    uint32_t* meta;
    uint32_t* aux;
    uint32_t cap;
};

__device__ __forceinline__ uint32_t side_edge(uint32_t m, uint32_t s) { return (m >> (3 * s)) & 7u; }
__device__ __forceinline__ bool side_exits(uint32_t m, uint32_t s) {
    return side_edge(m, s) == BASE_E && ((m >> (6 + s)) & 1u);
}

__device__ __forceinline__ uint32_t slot_of(uint64_t h, uint32_t cap) {
    return uint32_t((uint64_t(uint32_t(h >> 32)) * cap) >> 32);
}

// k-mer rolled one base forward: drop the first base, append b.
template <int KW>
__device__ __forceinline__ Kmer<KW> km_append(const Kmer<KW>& x, uint32_t b, uint32_t k) {
    Kmer<KW> r;
#pragma unroll
    for (int i = KW - 1; i > 0; --i) r.w[i] = (x.w[i] << 2) | (x.w[i - 1] >> 62);
    r.w[0] = (x.w[0] << 2) | b;
    const uint32_t kt = k - 32 * (KW - 1);
    r.w[KW - 1] &= (1ull << (2 * kt)) - 1;
    return r;
}

template <int KW>
__device__ __forceinline__ Kmer<KW> tab_key(const CtTab<KW>& t, uint32_t i) {
    Kmer<KW> x;
#pragma unroll
    for (int w = 0; w < KW; ++w) x.w[w] = t.key[size_t(i) * KW + w];
    return x;
}

// Unique keys: claim the top word, then fill the rest (read only after a
// barrier).  Returns the slot, or -1 when the table is full.
template <int KW>
__device__ __forceinline__ int tab_insert(const CtTab<KW>& t, const Kmer<KW>& x, uint32_t meta) {
    uint32_t s = slot_of(km_hash(x), t.cap);
    for (uint32_t n = 0; n < t.cap; ++n) {
        uint64_t* kp = &t.key[size_t(s) * KW];
        if (atomicCAS(reinterpret_cast<unsigned long long*>(kp + KW - 1), EMPTY, x.w[KW - 1]) == EMPTY) {
#pragma unroll
            for (int w = 0; w < KW - 1; ++w) kp[w] = x.w[w];
            t.meta[s] = meta;
            return int(s);
        }
        s = s + 1 == t.cap ? 0 : s + 1;
    }
    return -1;
}

template <int KW>
__device__ __forceinline__ int tab_find(const CtTab<KW>& t, const Kmer<KW>& x) {
    uint32_t s = slot_of(km_hash(x), t.cap);
    for (uint32_t n = 0; n < t.cap; ++n) {
        const uint64_t* kp = &t.key[size_t(s) * KW];
        const uint64_t top = kp[KW - 1];
        if (top == EMPTY) return -1;
        if (top == x.w[KW - 1]) {
            bool eq = true;
#pragma unroll
            for (int w = 0; w < KW - 1; ++w) eq &= kp[w] == x.w[w];
            if (eq) return int(s);
        }
        s = s + 1 == t.cap ? 0 : s + 1;
    }
    return -1;
}

// Dense vertex arrays in global scratch with an LDS index: subgraphs whose
// non-isolated vertices exceed an LDS table.  Vertex d (0 <= d < cap) has its
// key words at key[d * KW ..] and its state word at meta[d]; the LDS hash
// `ht` maps a key to d (entry d | the edge codes of d's two sides << 16 | 10
// fingerprint bits of the key's hash << 22, all-ones = free), so a lookup
// reads the one vertex whose fingerprint matches, and the link pass reads a
// neighbour's entrance side from the entry.
constexpr uint32_t CB_HCAP = 36864;  // LDS index slots (u32 each)
constexpr uint32_t CB_NMAX = 32767;  // vertices an index takes (load <= 0.89): directed states fit 16 bits
static_assert(CB_NMAX == uint32_t(ct_big_max_vertices()), "host and device agree on the dense table size");
static_assert(2 * CB_NMAX < (1u << 16), "directed states fit 16 bits");
constexpr uint32_t HT_FREE = 0xFFFFFFFFu;
constexpr uint32_t HT_FP = 0xFFC00000u;  // fingerprint bits of a dense-tier index entry

template <int KW>
struct BigTab {
    uint64_t* key;
    uint32_t* meta;
    uint32_t* aux;
    uint32_t cap;         // vertices
    uint32_t* ht;         // LDS index, CB_HCAP entries (until the links are made)
    const uint32_t* succ; // per directed state its successor state | 1 << 16, or itself (global scratch)
};

template <int KW>
__device__ __forceinline__ Kmer<KW> tab_key(const BigTab<KW>& t, uint32_t i) {
    Kmer<KW> x;
#pragma unroll
    for (int w = 0; w < KW; ++w) x.w[w] = t.key[size_t(i) * KW + w];
    return x;
}

// The index entry of key x (HT_FREE when x is not indexed), keys compared.
template <int KW>
__device__ __forceinline__ uint32_t big_entry(const BigTab<KW>& t, const Kmer<KW>& x) {
    const uint64_t h = km_hash(x);
    const uint32_t fp = uint32_t(h) & HT_FP;
    uint32_t s = slot_of(h, CB_HCAP);
    for (uint32_t n = 0; n < CB_HCAP; ++n) {
        const uint32_t e = t.ht[s];
        if (e == HT_FREE) return HT_FREE;
        if ((e & HT_FP) == fp) {
            const uint32_t d = e & 0xFFFFu;
            bool eq = true;
#pragma unroll
            for (int w = 0; w < KW; ++w) eq &= t.key[size_t(d) * KW + w] == x.w[w];
            if (eq) return e;
        }
        s = s + 1 == CB_HCAP ? 0 : s + 1;
    }
    return HT_FREE;
}

template <int KW>
__device__ __forceinline__ int tab_find(const BigTab<KW>& t, const Kmer<KW>& x) {
    const uint32_t e = big_entry(t, x);
    return e == HT_FREE ? -1 : int(e & 0xFFFFu);
}

// Lookup of a key known to be in the index (a neighbour through an edge at
// the cutoff: the edge is counted at both endpoints, so the neighbour is not
// isolated and was indexed): the probe chain is read to its end, and when one
// entry alone carries the key's fingerprint it is the key's, with no read of
// the vertex's key words.  Otherwise the keys are compared (tab_find), and
// always in an index loaded past 0.7, whose chains run long.
constexpr uint32_t CB_PRESENT_MAX = CB_HCAP * 7 / 10;
template <int KW>
__device__ __forceinline__ uint32_t big_entry_present(const BigTab<KW>& t, const Kmer<KW>& x) {
    if (t.cap > CB_PRESENT_MAX) return big_entry(t, x);
    const uint64_t h = km_hash(x);
    const uint32_t fp = uint32_t(h) & HT_FP;
    uint32_t s = slot_of(h, CB_HCAP), n_fp = 0, one = HT_FREE;
    for (uint32_t n = 0; n < CB_HCAP; ++n) {
        const uint32_t e = t.ht[s];
        if (e == HT_FREE) break;
        if ((e & HT_FP) == fp) {
            one = e;
            if (++n_fp > 1) break;
        }
        s = s + 1 == CB_HCAP ? 0 : s + 1;
    }
    return n_fp == 1 ? one : big_entry(t, x);
}

template <int KW>
__device__ __forceinline__ int tab_find_present(const BigTab<KW>& t, const Kmer<KW>& x) {
    const uint32_t e = big_entry_present(t, x);
    return e == HT_FREE ? -1 : int(e & 0xFFFFu);
}

template <int KW>
__device__ __forceinline__ int walk_next(const BigTab<KW>& t, uint32_t cur, uint32_t m, uint32_t k, Kmer<KW>& z,
                                         Kmer<KW>& zb, uint32_t& side, uint32_t& b) {
    const uint32_t e = (m >> (3 * side)) & 7u;
    b = side == 1 ? e : 3u - e;
    z = km_append(z, b, k);
    zb = km_prepend(zb, uint8_t(3u - b), k);
    const uint32_t x = 2 * cur + side, st = t.succ[x] & 0xFFFFu;
    if (st == x) return -1;
    side = st & 1u;
    return int(st >> 1);
}

// Index vertex d (state m) under key x (keys are unique): the first free
// slot from the key's home slot.
template <int KW>
__device__ __forceinline__ void big_index(uint32_t* ht, const Kmer<KW>& x, uint32_t d, uint32_t m) {
    const uint64_t h = km_hash(x);
    const uint32_t e = d | ((m & 0x3Fu) << 16) | (uint32_t(h) & HT_FP);
    uint32_t s = slot_of(h, CB_HCAP);
    while (atomicCAS(&ht[s], HT_FREE, e) != HT_FREE) s = s + 1 == CB_HCAP ? 0 : s + 1;
}

// Dense vertex arrays beyond CB_NMAX vertices (the wide dense tier): the
// key -> vertex index in the workgroup's global scratch (entry d | 12
// fingerprint bits << CW_DBITS, all-ones = free), 64-bit ranking words.
constexpr uint32_t CW_DBITS = 20;
constexpr uint32_t CW_DMASK = (1u << CW_DBITS) - 1;
static_assert(uint64_t(ct_wide_max_vertices()) <= CW_DMASK, "vertices fit the index entry");

template <int KW>
struct WideTab {
    uint64_t* key;
    uint32_t* meta;
    uint32_t* aux;
    uint32_t cap;          // vertices
    uint32_t* ht;          // global index, hcap entries
    uint32_t hcap;
    const uint64_t* succ;  // per directed state its successor state | 1 << 32, or itself
};

template <int KW>
__device__ __forceinline__ Kmer<KW> tab_key(const WideTab<KW>& t, uint32_t i) {
    Kmer<KW> x;
#pragma unroll
    for (int w = 0; w < KW; ++w) x.w[w] = t.key[size_t(i) * KW + w];
    return x;
}

template <int KW>
__device__ __forceinline__ int tab_find(const WideTab<KW>& t, const Kmer<KW>& x) {
    const uint64_t h = km_hash(x);
    const uint32_t fp = uint32_t(h) & ~CW_DMASK;
    uint32_t s = slot_of(h, t.hcap);
    for (uint32_t n = 0; n < t.hcap; ++n) {
        const uint32_t e = t.ht[s];
        if (e == HT_FREE) return -1;
        if ((e & ~CW_DMASK) == fp) {
            const uint32_t d = e & CW_DMASK;
            bool eq = true;
#pragma unroll
            for (int w = 0; w < KW; ++w) eq &= t.key[size_t(d) * KW + w] == x.w[w];
            if (eq) return int(d);
        }
        s = s + 1 == t.hcap ? 0 : s + 1;
    }
    return -1;
}

// The wide tier's index is in global memory: reading a probe chain to its end
// costs more there than the key words of the first fingerprint match.
template <int KW>
__device__ __forceinline__ int tab_find_present(const WideTab<KW>& t, const Kmer<KW>& x) {
    return tab_find(t, x);
}

template <int KW>
__device__ __forceinline__ int walk_next(const WideTab<KW>& t, uint32_t cur, uint32_t m, uint32_t k, Kmer<KW>& z,
                                         Kmer<KW>& zb, uint32_t& side, uint32_t& b) {
    const uint32_t e = (m >> (3 * side)) & 7u;
    b = side == 1 ? e : 3u - e;
    z = km_append(z, b, k);
    zb = km_prepend(zb, uint8_t(3u - b), k);
    const uint32_t x = 2 * cur + side, st = uint32_t(t.succ[x]);
    if (st == x) return -1;
    side = st & 1u;
    return int(st >> 1);
}

// Index vertex d under key x (keys are unique) in an index of hcap entries.
template <int KW>
__device__ __forceinline__ void wide_index(uint32_t* ht, uint32_t hcap, const Kmer<KW>& x, uint32_t d) {
    const uint64_t h = km_hash(x);
    const uint32_t e = d | (uint32_t(h) & ~CW_DMASK);
    uint32_t s = slot_of(h, hcap);
    while (atomicCAS(&ht[s], HT_FREE, e) != HT_FREE) s = s + 1 == hcap ? 0 : s + 1;
}

__device__ __forceinline__ uint32_t ld_meta(const uint32_t* m) {
    return __hip_atomic_load(m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One step of a walk: from the directed vertex z (rc zb) out of side `side`
// of slot `cur`, whose state `m` continues there.  Returns the next slot (-1
// if the table lacks it) and updates z, zb, side (the exit side there) and b.
template <int KW, class Tab>
__device__ __forceinline__ int walk_step(const Tab& t, uint32_t m, uint32_t k, Kmer<KW>& z, Kmer<KW>& zb,
                                         uint32_t& side, uint32_t& b) {
    const uint32_t e = side_edge(m, side);
    b = side == 1 ? e : 3u - e;
    z = km_append(z, b, k);
    zb = km_prepend(zb, uint8_t(3u - b), k);
    const bool can = km_less(z, zb);
    side = can ? 1u : 0u;  // entrance side front (canonical) -> leave through the back
    return tab_find(t, can ? z : zb);
}

// Tables whose keys are compared at every lookup.
template <int KW>
__device__ __forceinline__ int tab_find_present(const CtTab<KW>& t, const Kmer<KW>& x) {
    return tab_find(t, x);
}

// walk_step for the link pass: the neighbour through an edge at the cutoff is
// in the table (tab_find_present).
template <int KW, class Tab>
__device__ __forceinline__ int link_step(const Tab& t, uint32_t m, uint32_t k, Kmer<KW>& z, Kmer<KW>& zb,
                                         uint32_t& side, uint32_t& b) {
    const uint32_t e = side_edge(m, side);
    b = side == 1 ? e : 3u - e;
    z = km_append(z, b, k);
    zb = km_prepend(zb, uint8_t(3u - b), k);
    const bool can = km_less(z, zb);
    side = can ? 1u : 0u;
    return tab_find_present(t, can ? z : zb);
}

// One step of a walk that knows its current slot: the lookup of walk_step,
// or, for dense tables once linked, the successor word of the state.
template <int KW, class Tab>
__device__ __forceinline__ int walk_next(const Tab& t, uint32_t cur, uint32_t m, uint32_t k, Kmer<KW>& z, Kmer<KW>& zb,
                                         uint32_t& side, uint32_t& b) {
    return walk_step<KW>(t, m, k, z, zb, side, b);
}

// ASCII of base code b (A C G T).
__device__ __forceinline__ uint32_t base_char(uint32_t b) { return (0x54474341u >> (8 * b)) & 0xFFu; }

// The 2-bit codes of bases i .. i+3 of z (k bases, Kmer layout: the k-mer
// is the 2k-bit number w[KW-1] .. w[0], its first base highest), base i in
// bits 7-6.  Words are chosen by selects (no indexed register array).
template <int KW>
__device__ __forceinline__ uint32_t kmer_codes4(const Kmer<KW>& z, uint32_t k, uint32_t i) {
    const uint32_t p = 2 * (k - 4 - i);  // low bit of base i + 3
    const uint32_t wi = p >> 6, sh = p & 63u;
    uint64_t lo = z.w[0], hi = KW > 1 ? z.w[KW > 1 ? 1 : 0] : 0ull;
#pragma unroll
    for (int w = 1; w < KW; ++w)
        if (uint32_t(w) == wi) lo = z.w[w], hi = w + 1 < KW ? z.w[w + 1 < KW ? w + 1 : w] : 0ull;
    const uint64_t v = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
    return uint32_t(v) & 0xFFu;
}

// Four bases (codes in bits 7-6, 5-4, 3-2, 1-0) as four ASCII bytes, the
// first base's in the lowest byte.
__device__ __forceinline__ uint32_t chars4(uint32_t c) {
    return base_char(c >> 6) | (base_char((c >> 4) & 3u) << 8) | (base_char((c >> 2) & 3u) << 16) |
           (base_char(c & 3u) << 24);
}

// z's k bases as ASCII at out: single bytes up to the first 4-byte boundary,
// then four bases per aligned 4-byte store, then the remaining bytes.
template <int KW>
__device__ __forceinline__ void put_kmer(char* out, const Kmer<KW>& z, uint32_t k) {
    const uint32_t head = (4u - uint32_t(reinterpret_cast<uintptr_t>(out) & 3u)) & 3u;  // k >= 15 > head
    for (uint32_t i = 0; i < head; ++i) out[i] = char(base_char(kmer_codes4(z, k, i) >> 6));
    uint32_t i = head;
    for (; i + 4 <= k; i += 4) *reinterpret_cast<uint32_t*>(out + i) = chars4(kmer_codes4(z, k, i));
    for (; i < k; ++i) out[i] = char(base_char((kmer_codes4(z, k, i - 3) & 3u)));
}

// Reads bucket b's construct table: vertices in no edge at the cutoff are
// counted in s_cnt, the others inserted with their DFA byte (and their slot
// appended to `list` when given).  Sets *s_ovf when more than `limit` do not
// fit.  Several slots per thread are loaded before any is used, so each thread
// keeps several HBM reads in flight.  A table whose pass runs were recorded by
// the LDS construct (active vertices first in each run) is read only over the
// runs' active parts, its isolated and phantom counts taken from the
// construct's.  Every thread of the workgroup calls it (it synchronises).
template <int KW, class Insert>
__device__ __forceinline__ void load_bucket_with(const ContractArgs& a, uint32_t b, uint32_t limit, uint32_t* s_n,
                                                 uint32_t* s_ovf, uint32_t* s_cnt, Insert&& insert) {
    constexpr int U = KW <= 2 ? 8 : 4;
    __shared__ uint32_t s_pre[SEG_MAX + 1];  // active entries before each run
    __shared__ uint32_t s_beg[SEG_MAX];      // each run's first table entry
    const uint64_t* tbl = reinterpret_cast<const uint64_t*>(a.tbl_ptr[b]);
    const uint32_t ns = a.nseg ? a.nseg[b] : 0u;
    uint64_t n_idx = a.tbl_cap[b];
    if (ns) {
        if (threadIdx.x < 64) {  // one wave: the runs' active counts scanned
            const uint32_t l = threadIdx.x;
            const uint64_t sg = l < ns ? a.seg[uint64_t(b) * SEG_MAX + l] : 0ull;
            uint32_t v = uint32_t(sg >> 32);
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t u = uint32_t(__shfl_up(int(v), d));
                if (int(l) >= d) v += u;
            }
            s_beg[l] = uint32_t(sg);
            s_pre[l + 1] = v;
            if (l == 0) s_pre[0] = 0;
        }
        __syncthreads();
        n_idx = s_pre[ns];
    }
    // table entry of the v-th entry read (the runs' active parts back to back)
    const auto entry_of = [&](uint64_t v) -> uint64_t {
        if (!ns) return v;
        uint32_t lo = 0, hi = ns;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_pre[mid] <= uint32_t(v)) lo = mid;
            else hi = mid;
        }
        return uint64_t(s_beg[lo]) + (v - s_pre[lo]);
    };
    uint32_t iso = 0, ph = 0;
    for (uint64_t s0 = threadIdx.x; s0 < n_idx; s0 += uint64_t(U) * CT_WG) {
        uint64_t e[U][KW + 1];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t v = s0 + uint64_t(u) * CT_WG;
            if (v < n_idx) {
                const uint64_t s = entry_of(v);
#pragma unroll
                for (int w = 0; w <= KW; ++w) e[u][w] = tbl[s * (KW + 1) + w];
            } else {
                e[u][KW - 1] = EMPTY;
            }
        }
        bool stop = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (e[u][KW - 1] == EMPTY || stop) continue;
            const uint32_t ef = uint32_t(e[u][KW]), st = uint32_t(e[u][KW] >> 32);
            const uint32_t f = edge_at(ef, 0, a.cutoff), bk = edge_at(ef, 16, a.cutoff);
            if (f == BASE_E && bk == BASE_E) {
                ++iso;
                ph += (st & 6u) != 0;
                continue;
            }
            Kmer<KW> x;
#pragma unroll
            for (int w = 0; w < KW; ++w) x.w[w] = e[u][w];
            const uint32_t m = f | (bk << 3) | ((st & 2u) << 5) | ((st & 4u) << 5);
            const uint32_t d = atomicAdd(s_n, 1u);
            if (d >= limit || !insert(d, x, m)) {
                *s_ovf = 1;
                stop = true;
            }
        }
        if (stop || __hip_atomic_load(s_ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
    }
    if (ns && threadIdx.x == 0) iso += a.iso[2 * b], ph += a.iso[2 * b + 1];
    if (iso) atomicAdd(&s_cnt[CT_ISOLATED], iso);
    if (ph) atomicAdd(&s_cnt[CT_PHANTOM], ph);
}

// The LDS / global open-addressing tables: vertex d goes to a free slot of t,
// whose index is appended to `list` when given.
template <int KW>
__device__ __forceinline__ void load_bucket(const ContractArgs& a, uint32_t b, const CtTab<KW>& t, uint32_t limit,
                                            uint32_t* s_n, uint32_t* s_ovf, uint32_t* s_cnt, uint16_t* list) {
    load_bucket_with<KW>(a, b, limit, s_n, s_ovf, s_cnt, [&](uint32_t d, const Kmer<KW>& x, uint32_t m) {
        const int slot = tab_insert(t, x, m);
        if (slot >= 0 && list) list[d] = uint16_t(slot);
        return slot >= 0;
    });
}

// Side s of slot i (key x, rc xr, state m): the directed state it continues to
// as (slot << 1 | exit side there), or -1 when the unitig ends there.  Sets
// *err on an inconsistent table.
template <int KW, class Tab>
__device__ __forceinline__ int link_of(const Tab& t, uint32_t i, uint32_t s, const Kmer<KW>& x,
                                       const Kmer<KW>& xr, uint32_t m, uint32_t k, uint32_t* err) {
    if (side_edge(m, s) >= 4) return -1;
    Kmer<KW> z = s ? x : xr, zb = s ? xr : x;
    uint32_t side = s, bb;
    const int j = link_step<KW>(t, m, k, z, zb, side, bb);
    if (j < 0) {
        *err = 1;
        return -1;
    }
    const uint32_t ent = 1u - side;
    const uint32_t ej = side_edge(ld_meta(&t.meta[j]), ent);
    if (ej == BASE_E) *err = 1;                  // the edge is counted at both ends
    if (ej >= 4) return -1;                      // branching entrance: crossed
    if (uint32_t(j) == i && ent == s) return -1; // back through the same side
    return int((uint32_t(j) << 1) | side);
}

// link_of for the dense tier: the neighbour's entrance side read from its
// index entry instead of its state word in global scratch.
template <int KW>
__device__ __forceinline__ int link_of_big(const BigTab<KW>& t, uint32_t i, uint32_t s, const Kmer<KW>& x,
                                           const Kmer<KW>& xr, uint32_t m, uint32_t k, uint32_t* err) {
    if (side_edge(m, s) >= 4) return -1;
    Kmer<KW> z = s ? x : xr, zb = s ? xr : x;
    const uint32_t e = side_edge(m, s), b = s == 1 ? e : 3u - e;
    z = km_append(z, b, k);
    zb = km_prepend(zb, uint8_t(3u - b), k);
    const bool can = km_less(z, zb);
    const uint32_t side = can ? 1u : 0u;  // entrance side front (canonical) -> leave through the back
    const uint32_t ent_w = big_entry_present(t, can ? z : zb);
    if (ent_w == HT_FREE) {
        *err = 1;
        return -1;
    }
    const uint32_t j = ent_w & 0xFFFFu, ent = 1u - side;
    const uint32_t ej = side_edge(ent_w >> 16, ent);
    if (ej == BASE_E) *err = 1;            // the edge is counted at both ends
    if (ej >= 4) return -1;                // branching entrance: crossed
    if (j == i && ent == s) return -1;     // back through the same side
    return int((j << 1) | side);
}

// Thread 0: every cycle among the candidate slots (or all slots marked M_ONCYC
// when the list overflowed), walked once from a vertex out of its back side;
// the smallest vertex in canonical form writes it.  Its label offset goes to
// *off_of(slot).
template <int KW, class Tab, class OffOf>
__device__ void cycles_serial(const Tab& t, uint32_t k, const uint32_t* cand, uint32_t ncand, uint32_t* s_bytes,
                              uint32_t* s_nrec, uint32_t* s_cnt, uint32_t* s_err, OffOf off_of) {
    const bool all = ncand > uint32_t(CT_CAND);
    const uint32_t nc = all ? t.cap : ncand;
    for (uint32_t q = 0; q < nc; ++q) {
        const uint32_t c = all ? q : cand[q];
        if (t.key[size_t(c) * KW + KW - 1] == EMPTY || !(t.meta[c] & M_ONCYC) || (t.meta[c] & (M_VIS | M_WRITER)))
            continue;
        Kmer<KW> z = tab_key<KW>(t, c), zb = km_rc(z, k), mn = z;
        uint32_t side = 1, cur = c, n = 1, ms = c, bb;
        t.meta[c] |= M_VIS;
        for (uint32_t step = 0; step < t.cap; ++step) {
            const uint32_t m = t.meta[cur];
            if (!(m & (M_LINK0 << side))) {  // every side on a cycle continues
                *s_err = 1;
                break;
            }
            const int j = walk_next<KW>(t, cur, m, k, z, zb, side, bb);
            if (j < 0) {
                *s_err = 1;
                break;
            }
            if (uint32_t(j) == c) break;
            t.meta[j] |= M_VIS;
            Kmer<KW> key;  // canonical form
#pragma unroll
            for (int w = 0; w < KW; ++w) key.w[w] = side ? z.w[w] : zb.w[w];
            if (km_less(key, mn)) mn = key, ms = uint32_t(j);
            cur = uint32_t(j);
            ++n;
        }
        const uint32_t bytes = n + k - 1;
        *off_of(ms) = *s_bytes;
        *s_bytes += bytes;
        t.meta[ms] |= M_WRITER | M_CYCLE;
        *s_nrec += 1;
        s_cnt[CT_TRIVIAL] += 1;
        s_cnt[CT_ICC] += 1;
        s_cnt[CT_VERTICES] += n;
        s_cnt[CT_LABEL_BYTES] += bytes;
    }
}

// Writes the cycle whose smallest vertex is slot i (canonical, leaving through
// its back side) at out[0 .. n + k - 1), never past `end`.
template <int KW, class Tab>
__device__ __forceinline__ uint32_t write_cycle(const Tab& t, uint32_t i, uint32_t k, char* out, uint64_t off,
                                                uint64_t end, uint32_t* err) {
    Kmer<KW> z = tab_key<KW>(t, i), zb = km_rc(z, k);
    if (off + k > end) {
        atomicOr(err, 1u);
        return 0;
    }
    put_kmer(out, z, k);
    uint32_t side = 1, cur = i, n = 1, bb;
    for (uint32_t step = 0; step < t.cap; ++step) {
        const uint32_t m = t.meta[cur];
        if (!(m & (M_LINK0 << side))) break;
        const int j = walk_next<KW>(t, cur, m, k, z, zb, side, bb);
        if (j < 0 || uint32_t(j) == i) break;
        if (off + k + n > end) {
            atomicOr(err, 1u);
            break;
        }
        out[k - 1 + n] = char(base_char(bb));
        cur = uint32_t(j);
        ++n;
    }
    return n;
}

// list-ranking word: directed state (slot << 1 | exit side) in the low 14 bits,
// steps to the state it points at in the high 18 bits
constexpr uint32_t JP_BITS = 14, JP_MASK = (1u << JP_BITS) - 1, JP_DMAX = (1u << (32 - JP_BITS)) - 1;
__device__ __forceinline__ uint32_t jp_pack(uint32_t p, uint32_t d) { return p | (min(d, JP_DMAX) << JP_BITS); }
constexpr uint32_t RANK_ROUNDS = 16;  // > log2(2 * LDS capacity): every path converges
static_assert(2 * ct_lds_capacity(1) <= (1 << JP_BITS), "states fit the ranking word");

template <int KW>
__global__ __launch_bounds__(CT_WG) void k_contract_rank(ContractArgs a) {
    constexpr uint32_t CL = uint32_t(ct_lds_capacity(KW));
    constexpr uint32_t NV = CL * 4 / 5;                // vertices an LDS table takes
    constexpr uint32_t NQ = (NV + CT_WG - 1) / CT_WG;  // vertices per thread
    __shared__ uint64_t s_key[CL * KW];
    __shared__ uint16_t s_list[NV];  // slots in insertion order
    __shared__ uint32_t s_meta[CL];
    __shared__ uint32_t s_jp[2 * CL];
    __shared__ uint32_t s_cand[CT_CAND];
    __shared__ uint32_t s_cnt[8];
    __shared__ uint32_t s_chg[RANK_ROUNDS];
    __shared__ uint32_t s_item, s_n, s_ovf, s_ncand, s_bytes, s_nrec, s_rctr, s_err, s_skip;
    __shared__ unsigned long long s_brec, s_bbyte;
    const uint32_t tid = threadIdx.x;
    const uint32_t k = a.k;
    const uint32_t kt = k - 32 * (KW - 1);
    CtTab<KW> t;
    t.key = s_key;
    t.meta = s_meta;
    t.aux = nullptr;
    t.cap = CL;

    while (true) {
        __syncthreads();
        if (tid == 0) {
            s_item = atomicAdd(&a.queue_head[0], 1u);
            s_n = s_ovf = s_ncand = s_bytes = s_nrec = s_rctr = s_err = s_skip = 0;
        }
        if (tid < 8) s_cnt[tid] = 0;
        if (tid < RANK_ROUNDS) s_chg[tid] = 0;
        __syncthreads();
        if (s_item >= a.n_list) break;
        const uint32_t b = a.list[s_item];
        for (uint32_t i = tid; i < CL; i += CT_WG) {
            s_key[size_t(i) * KW + KW - 1] = EMPTY;
            s_meta[i] = 0;
        }
        __syncthreads();
        load_bucket(a, b, t, NV, &s_n, &s_ovf, s_cnt, s_list);
        __syncthreads();
        if (s_ovf) {
            if (tid == 0) a.ovf_list[atomicAdd(a.ovf_count, 1u)] = b;
            continue;
        }

        // this thread's vertices: list entries tid + q * CT_WG, bit q of occ
        const uint32_t nv = s_n;
        uint32_t occ = 0, slot[NQ];
#pragma unroll
        for (uint32_t q = 0; q < NQ; ++q) {
            const uint32_t d = tid + q * CT_WG;
            slot[q] = d < nv ? s_list[d] : 0u;
            if (d < nv) occ |= 1u << q;
        }

        // link: each directed state points at its successor (1 step) or at itself (an end)
#pragma unroll
        for (uint32_t q = 0; q < NQ; ++q) {
            if (!((occ >> q) & 1u)) continue;
            const uint32_t i = slot[q];
            const Kmer<KW> x = tab_key(t, i);
            const Kmer<KW> xr = km_rc(x, k);
            const uint32_t m = ld_meta(&s_meta[i]);
            uint32_t add = 0;
            for (uint32_t sd = 0; sd < 2; ++sd) {
                const int nx = link_of(t, i, sd, x, xr, m, k, &s_err);
                s_jp[2 * i + sd] = nx < 0 ? jp_pack(2 * i + sd, 0) : jp_pack(uint32_t(nx), 1);
                if (nx >= 0) add |= M_LINK0 << sd;
            }
            if (add) atomicOr(&s_meta[i], add);
        }
        __syncthreads();

        // rank: pointer jumping until no path state moves (cycle states never
        // settle; they stop counting once their distance passes the vertex count)
        for (uint32_t r = 0; r < RANK_ROUNDS; ++r) {
            uint32_t nw[NQ][2];
            bool any = false;
#pragma unroll
            for (uint32_t q = 0; q < NQ; ++q) {
                const uint32_t i = slot[q];
#pragma unroll
                for (uint32_t sd = 0; sd < 2; ++sd) {
                    nw[q][sd] = 0;
                    if (!((occ >> q) & 1u)) continue;
                    const uint32_t x = 2 * i + sd, w = s_jp[x];
                    const uint32_t p = w & JP_MASK;
                    nw[q][sd] = w;
                    if (p == x) continue;
                    const uint32_t w2 = s_jp[p];
                    const uint32_t p2 = w2 & JP_MASK, nd = (w >> JP_BITS) + (w2 >> JP_BITS);
                    nw[q][sd] = jp_pack(p2, nd);
                    any |= p2 != p && nd < s_n;
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t q = 0; q < NQ; ++q)
                if ((occ >> q) & 1u) {
                    const uint32_t i = slot[q];
                    s_jp[2 * i] = nw[q][0];
                    s_jp[2 * i + 1] = nw[q][1];
                }
            if (any) s_chg[r] = 1;
            __syncthreads();
            if (!s_chg[r]) break;
        }

        // per vertex: the ends of its path, the output orientation and its position
        uint32_t info[NQ], wn[NQ], wfl[NQ];  // info: A | pos << 13 | fwd << 26 | path << 27
#pragma unroll
        for (uint32_t q = 0; q < NQ; ++q) {
            info[q] = wn[q] = wfl[q] = 0;
            if (!((occ >> q) & 1u)) continue;
            const uint32_t i = slot[q];
            const uint32_t wf = s_jp[2 * i], wb = s_jp[2 * i + 1];
            const uint32_t Ff = wf & JP_MASK, Fb = wb & JP_MASK;
            if (s_jp[Fb] != jp_pack(Fb, 0)) {  // not a path end: the vertex lies on a cycle
                atomicOr(&s_meta[i], M_ONCYC);
                const uint32_t c = atomicAdd(&s_ncand, 1u);
                if (c < uint32_t(CT_CAND)) s_cand[c] = i;
                continue;
            }
            const uint32_t df = wf >> JP_BITS, db = wb >> JP_BITS;
            const uint32_t L = Ff >> 1, sL = Ff & 1u, R = Fb >> 1, sR = Fb & 1u;
            const Kmer<KW> cL = tab_key(t, L), cR = tab_key(t, R);
            // first k-mer of the path read with this vertex canonical, and the
            // reverse complement of its last k-mer
            const Kmer<KW> first = sL == 0 ? cL : km_rc(cL, k);
            const Kmer<KW> rlast = sR == 1 ? km_rc(cR, k) : cR;
            const bool fwd = km_less(first, rlast);
            const uint32_t pos = fwd ? df : db;
            const uint32_t A = fwd ? L : R;
            info[q] = A | (pos << 13) | (uint32_t(fwd) << 26) | (1u << 27);
            if (pos == 0) {
                const uint32_t sA = fwd ? sL : sR, B = fwd ? R : L, sB = fwd ? sR : sL;
                wn[q] = df + db + 1;
                wfl[q] = (side_exits(s_meta[A], sA) ? 2u : 0u) | (side_exits(s_meta[B], sB) ? 4u : 0u);
            }
        }
        __syncthreads();  // every ranking word read

        // the first vertex of each path reserves its label (offset in its front word)
        {
            uint32_t nrec = 0, verts = 0, lbytes = 0, triv = 0, disc = 0;
#pragma unroll
            for (uint32_t q = 0; q < NQ; ++q) {
                if (!info[q] || ((info[q] >> 13) & 0x1FFFu)) continue;
                const uint32_t i = slot[q];
                const uint32_t bytes = wn[q] + k - 1;
                s_jp[2 * i] = atomicAdd(&s_bytes, bytes);
                ++nrec;
                verts += wn[q];
                lbytes += bytes;
                disc += wfl[q] != 0;
                triv += wfl[q] == 0;
            }
            if (nrec) {
                atomicAdd(&s_nrec, nrec);
                atomicAdd(&s_cnt[CT_VERTICES], verts);
                atomicAdd(&s_cnt[CT_LABEL_BYTES], lbytes);
                atomicAdd(&s_cnt[CT_TRIVIAL], triv);
                atomicAdd(&s_cnt[CT_DISC_EDGES], disc);
            }
        }
        __syncthreads();
        if (tid == 0 && s_ncand)
            cycles_serial<KW>(t, k, s_cand, s_ncand, &s_bytes, &s_nrec, s_cnt, &s_err,
                          [&](uint32_t slot) { return &s_jp[2 * slot]; });
        __syncthreads();

        // reserve the bucket's output, publish its counters
        if (tid == 0) {
            s_brec = atomicAdd(&a.cursors[0], (unsigned long long)s_nrec);
            s_bbyte = atomicAdd(&a.cursors[1], (unsigned long long)s_bytes);
            s_skip = s_brec + s_nrec > a.rec_cap || s_bbyte + s_bytes > a.lab_cap;
            if (s_skip) atomicOr(a.retry, 1u);
            if (s_err) atomicOr(a.error, 1u);
            s_cnt[CT_FLAGS] = 1u | (s_err ? 4u : 0u);
        }
        __syncthreads();
        if (tid < 8) a.counts[size_t(b) * 8 + tid] = s_cnt[tid];
        if (s_skip) continue;

        // write: every path vertex stores its base, the first one its k-mer and the record
        const uint64_t lab_end = s_bbyte + s_bytes;
#pragma unroll
        for (uint32_t q = 0; q < NQ; ++q) {
            if (!((occ >> q) & 1u)) continue;
            const uint32_t i = slot[q];
            if (!info[q]) {
                if ((s_meta[i] & (M_WRITER | M_CYCLE)) != (M_WRITER | M_CYCLE)) continue;
                const uint64_t off = s_bbyte + s_jp[2 * i];
                const uint32_t n = write_cycle<KW>(t, i, k, a.labels + off, off, lab_end, a.error);
                const uint32_t ri = atomicAdd(&s_rctr, 1u);
                if (ri < s_nrec) a.recs[s_brec + ri] = UnitigRec{off, n + k - 1, uint16_t(b), 1};
                continue;
            }
            const uint32_t A = info[q] & 0x1FFFu, pos = (info[q] >> 13) & 0x1FFFu;
            const bool fwd = (info[q] >> 26) & 1u;
            const uint64_t off = s_bbyte + s_jp[2 * A];
            const Kmer<KW> x = tab_key(t, i);
            if (pos == 0) {
                if (off + wn[q] + k - 1 > lab_end) {
                    atomicOr(a.error, 1u);
                    continue;
                }
                put_kmer(a.labels + off, fwd ? x : km_rc(x, k), k);
                const uint32_t ri = atomicAdd(&s_rctr, 1u);
                if (ri < s_nrec) a.recs[s_brec + ri] = UnitigRec{off, wn[q] + k - 1, uint16_t(b), uint16_t(wfl[q])};
            } else {
                const uint32_t c = fwd ? uint32_t(x.w[0] & 3u) : 3u - uint32_t((x.w[KW - 1] >> (2 * (kt - 1))) & 3u);
                if (off + pos + k - 1 >= lab_end) {
                    atomicOr(a.error, 1u);
                    continue;
                }
                a.labels[off + pos + k - 1] = char(base_char(c));
            }
        }
    }
}

// Subgraphs whose non-isolated vertices exceed an LDS table but not
// CB_NMAX: the same list ranking as k_contract_rank over dense vertex arrays
// in this workgroup's global scratch region, with the key -> vertex index in
// LDS.  Ranking words pack a directed state (16 bits) and a distance (16
// bits); two buffers alternate between rounds.
constexpr uint32_t CB_ROUNDS = 17;  // > log2(2 * CB_NMAX): every path converges
constexpr uint32_t CB_M16 = 0xFFFFu;
__device__ __forceinline__ uint32_t cb_pack(uint32_t p, uint32_t d) { return p | (min(d, CB_M16) << 16); }

// Ranking words kept in LDS (over the index, no longer needed once linked)
// when the directed states fit: 2n <= CB_JP_LDS; otherwise the successor ids
// (u16) and a bit per state (from word CB_WALK_BITS on) for the path walks.
constexpr uint32_t CB_JP_LDS = 39936;
constexpr uint32_t CB_WALK_BITS = 32768;
static_assert(CB_WALK_BITS * 2 >= 2 * CB_NMAX && CB_WALK_BITS + (2 * CB_NMAX + 31) / 32 <= CB_JP_LDS, "walk buffers");
constexpr uint32_t CB_JQ = CB_JP_LDS / CT_WG;  // states per thread
static_assert(CB_JP_LDS >= CB_HCAP && CB_JP_LDS % CT_WG == 0, "LDS ranking buffer");
constexpr uint32_t CB_RU = 8;                        // states per thread in flight, global ranking
constexpr uint32_t CB_WU = 4;                        // vertices per thread in flight, label writes
constexpr uint32_t CB_STAGE = CB_JP_LDS * 4;         // label bytes staged in LDS (the same buffer)

template <int KW>
__global__ __launch_bounds__(CT_WG) void k_contract_big(ContractArgs a) {
    __shared__ __align__(16) uint32_t s_buf[CB_JP_LDS];  // the key index, the ranking words, then labels
    uint32_t* s_ht = s_buf;
    __shared__ uint32_t s_cand[CT_CAND];
    __shared__ uint32_t s_cnt[8];
    __shared__ uint32_t s_chg[CB_ROUNDS];
    __shared__ uint32_t s_item, s_n, s_ovf, s_ncand, s_bytes, s_nrec, s_rctr, s_err, s_skip, s_wq;
    __shared__ unsigned long long s_brec, s_bbyte;
    const uint32_t tid = threadIdx.x;
    const uint32_t k = a.k;
    const uint32_t kt = k - 32 * (KW - 1);
    // scratch: key words, state, two ranking buffers, path info (first vertex
    // | forward << 16 | on a path << 17), position, the first vertex's
    // vertex count | end flags << 29, label offsets
    constexpr uint64_t N = CB_NMAX;
    uint8_t* base = a.scratch + uint64_t(blockIdx.x) * a.scratch_per_wg;
    BigTab<KW> t;
    t.key = reinterpret_cast<uint64_t*>(base);
    t.meta = reinterpret_cast<uint32_t*>(base + N * KW * 8);
    uint32_t* jp0 = t.meta + N;
    uint32_t* jp1 = jp0 + 2 * N;
    uint32_t* ia = jp1 + 2 * N;
    uint32_t* posv = ia + N;
    uint32_t* wnv = posv + N;
    uint32_t* offv = wnv + N;
    t.aux = offv;
    t.ht = s_ht;
    t.succ = jp0;

    while (true) {
        __syncthreads();
        if (tid == 0) {
            s_item = atomicAdd(&a.queue_head[1], 1u);
            s_n = s_ovf = s_ncand = s_bytes = s_nrec = s_rctr = s_err = s_skip = 0;
        }
        if (tid < 8) s_cnt[tid] = 0;
        if (tid < CB_ROUNDS) s_chg[tid] = 0;
        __syncthreads();
        if (s_item >= __hip_atomic_load(a.ovf_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        const uint32_t b = a.ovf_list[s_item];
        for (uint32_t i = tid; i < CB_HCAP; i += CT_WG) s_ht[i] = HT_FREE;
        __syncthreads();
        load_bucket_with<KW>(a, b, CB_NMAX, &s_n, &s_ovf, s_cnt, [&](uint32_t d, const Kmer<KW>& x, uint32_t m) {
#pragma unroll
            for (int w = 0; w < KW; ++w) t.key[size_t(d) * KW + w] = x.w[w];
            t.meta[d] = m;
            big_index(s_ht, x, d, m);
            return true;
        });
        __syncthreads();
        if (s_ovf) {
            if (tid == 0) a.ovf2_list[atomicAdd(a.ovf2_count, 1u)] = b;
            continue;
        }
        const uint32_t n = s_n;
        t.cap = n;

        // link: each directed state points at its successor (1 step) or at itself (an end)
        for (uint32_t d = tid; d < n; d += CT_WG) {
            const Kmer<KW> x = tab_key<KW>(t, d);
            const Kmer<KW> xr = km_rc(x, k);
            const uint32_t m = ld_meta(&t.meta[d]);
            uint32_t add = 0;
            for (uint32_t sd = 0; sd < 2; ++sd) {
                const int nx = link_of_big<KW>(t, d, sd, x, xr, m, k, &s_err);
                jp0[2 * d + sd] = nx < 0 ? cb_pack(2 * d + sd, 0) : cb_pack(uint32_t(nx), 1);
                if (nx >= 0) add |= M_LINK0 << sd;
            }
            // only this thread writes vertex d's state here; the others read
            // its DFA bits, which the store keeps
            if (add) __hip_atomic_store(&t.meta[d], m | add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        __syncthreads();

        // rank: pointer jumping until no path state moves (cycle states stop
        // counting once past the vertex count).  The successor words in jp0
        // stay as they are (cycle walks follow them).
        const uint32_t* jp;
        if (2 * n <= CB_JP_LDS) {
            // in LDS, in place: new words in registers between two barriers
            for (uint32_t x = tid; x < 2 * n; x += CT_WG) s_buf[x] = jp0[x];
            __syncthreads();
            for (uint32_t r = 0; r < CB_ROUNDS; ++r) {
                uint32_t nw[CB_JQ];
                bool any = false;
#pragma unroll
                for (uint32_t q = 0; q < CB_JQ; ++q) {
                    const uint32_t x = tid + q * CT_WG;
                    nw[q] = 0;
                    if (x >= 2 * n) continue;
                    const uint32_t w = s_buf[x], p = w & CB_M16;
                    nw[q] = w;
                    if (p == x) continue;
                    const uint32_t w2 = s_buf[p];
                    const uint32_t p2 = w2 & CB_M16, nd = (w >> 16) + (w2 >> 16);
                    nw[q] = cb_pack(p2, nd);
                    any |= p2 != p && nd < n;
                }
                __syncthreads();
#pragma unroll
                for (uint32_t q = 0; q < CB_JQ; ++q) {
                    const uint32_t x = tid + q * CT_WG;
                    if (x < 2 * n) s_buf[x] = nw[q];
                }
                if (any) s_chg[r] = 1;
                __syncthreads();
                if (!s_chg[r]) break;
            }
            jp = s_buf;
        } else {
            // 2n states beyond the LDS ranking buffer (2n <= 65534): every
            // path is walked from its start (a state no other state leads to)
            // over successor ids held in LDS (16 bits each), twice: to its end
            // and length, then writing each state's end and distance.  States
            // on cycles have no start and keep their link words (successor,
            // distance 1), which the passes below tell from a path's.
            // Pointer jumping over these states in global scratch read a
            // random word per state and round.
            uint16_t* s_succ = reinterpret_cast<uint16_t*>(s_buf);
            uint32_t* s_led = s_buf + CB_WALK_BITS;  // bit x: some state leads to x
            for (uint32_t x = tid; x < 2 * n; x += CT_WG) s_succ[x] = uint16_t(jp0[x] & CB_M16);
            for (uint32_t i = tid; i < (2 * n + 31) / 32; i += CT_WG) s_led[i] = 0;
            if (tid == 0) s_wq = 0;
            __syncthreads();
            for (uint32_t x = tid; x < 2 * n; x += CT_WG) {
                const uint32_t y = s_succ[x];
                if (y != x) atomicOr(&s_led[y >> 5], 1u << (y & 31u));
                jp1[x] = jp0[x];
            }
            __syncthreads();
            while (true) {  // a wave takes 64 states at a time and walks from those that start a path
                uint32_t x0 = 0;
                if ((tid & 63u) == 0) x0 = atomicAdd(&s_wq, 64u);
                x0 = uint32_t(__builtin_amdgcn_readfirstlane(int(x0)));
                if (x0 >= 2 * n) break;
                const uint32_t x = x0 + (tid & 63u);
                if (x >= 2 * n || ((s_led[x >> 5] >> (x & 31u)) & 1u)) continue;
                uint32_t len = 1, y = x;
                for (uint32_t z = s_succ[y]; z != y && len <= 2 * n; z = s_succ[y]) y = z, ++len;
                if (len > 2 * n) {  // a path longer than the states: an inconsistent table
                    s_err = 1;
                    continue;
                }
                const uint32_t F = y;
                y = x;
                for (uint32_t i = 0;; ++i) {
                    jp1[y] = cb_pack(F, len - 1 - i);
                    const uint32_t z = s_succ[y];
                    if (z == y) break;
                    y = z;
                }
            }
            __syncthreads();
            const uint32_t* src = jp1;
            jp = src;
        }

        // Path ends first: the end state F = (vertex L, side s) of a path
        // with far end F' = (R, s') reads the path from F when E(F) <
        // E(F'), for E(F) = the key of L, or its reverse complement when s =
        // 1 (the first k-mer of the reading from F; E(F') is the reverse
        // complement of its last).  Bit s of offv[L] keeps that choice (offv
        // is free until the labels are reserved), so that the other vertices
        // of the path need not read the keys of both its ends.
        for (uint32_t d = tid; d < n; d += CT_WG) {
            const uint32_t w0 = jp[2 * d], w1 = jp[2 * d + 1];
            const bool e0 = w0 == cb_pack(2 * d, 0), e1 = w1 == cb_pack(2 * d + 1, 0);
            uint32_t bits = 0;
            if (e0 || e1) {
                const Kmer<KW> x = tab_key<KW>(t, d), xr = km_rc(x, k);
#pragma unroll
                for (uint32_t sd = 0; sd < 2; ++sd) {
                    if (!(sd ? e1 : e0)) continue;
                    const uint32_t Fo = (sd ? w0 : w1) & CB_M16;  // reached through the other side
                    const Kmer<KW> co = tab_key<KW>(t, Fo >> 1);
                    const Kmer<KW> eo = (Fo & 1u) ? km_rc(co, k) : co;
                    bits |= uint32_t(km_less(sd ? xr : x, eo)) << sd;
                }
            }
            offv[d] = bits;
        }
        __syncthreads();

        // per vertex: the ends of its path, the output orientation and its position
        for (uint32_t d = tid; d < n; d += CT_WG) {
            const uint32_t wf = jp[2 * d], wb = jp[2 * d + 1];
            const uint32_t Ff = wf & CB_M16, Fb = wb & CB_M16;
            if (jp[Fb] != cb_pack(Fb, 0)) {  // not a path end: the vertex lies on a cycle
                atomicOr(&t.meta[d], M_ONCYC);
                const uint32_t c = atomicAdd(&s_ncand, 1u);
                if (c < uint32_t(CT_CAND)) s_cand[c] = d;
                ia[d] = 0;
                continue;
            }
            const uint32_t df = wf >> 16, db = wb >> 16;
            const uint32_t L = Ff >> 1, sL = Ff & 1u, R = Fb >> 1, sR = Fb & 1u;
            const bool fwd = (offv[L] >> sL) & 1u;  // E(Ff) < E(Fb): Fb is Ff's far end
            const uint32_t pos = fwd ? df : db;
            const uint32_t A = fwd ? L : R;
            ia[d] = A | (uint32_t(fwd) << 16) | (1u << 17);
            posv[d] = pos;
            if (pos == 0) {
                const uint32_t sA = fwd ? sL : sR, B = fwd ? R : L, sB = fwd ? sR : sL;
                const uint32_t fl = (side_exits(ld_meta(&t.meta[A]), sA) ? 2u : 0u) |
                                    (side_exits(ld_meta(&t.meta[B]), sB) ? 4u : 0u);
                wnv[d] = (df + db + 1) | (fl << 29);
            }
        }
        __syncthreads();

        // the first vertex of each path reserves its label
        {
            uint32_t nrec = 0, verts = 0, lbytes = 0, triv = 0, disc = 0;
            for (uint32_t d = tid; d < n; d += CT_WG) {
                if (!ia[d] || posv[d]) continue;
                const uint32_t wn = wnv[d] & 0x1FFFFFFFu, fl = wnv[d] >> 29;
                const uint32_t bytes = wn + k - 1;
                offv[d] = atomicAdd(&s_bytes, bytes);
                ++nrec;
                verts += wn;
                lbytes += bytes;
                disc += fl != 0;
                triv += fl == 0;
            }
            if (nrec) {
                atomicAdd(&s_nrec, nrec);
                atomicAdd(&s_cnt[CT_VERTICES], verts);
                atomicAdd(&s_cnt[CT_LABEL_BYTES], lbytes);
                atomicAdd(&s_cnt[CT_TRIVIAL], triv);
                atomicAdd(&s_cnt[CT_DISC_EDGES], disc);
            }
        }
        __syncthreads();
        if (tid == 0 && s_ncand)
            cycles_serial<KW>(t, k, s_cand, s_ncand, &s_bytes, &s_nrec, s_cnt, &s_err,
                              [&](uint32_t slot) { return &offv[slot]; });
        __syncthreads();

        // reserve the bucket's output, publish its counters
        if (tid == 0) {
            s_brec = atomicAdd(&a.cursors[0], (unsigned long long)s_nrec);
            s_bbyte = atomicAdd(&a.cursors[1], (unsigned long long)s_bytes);
            s_skip = s_brec + s_nrec > a.rec_cap || s_bbyte + s_bytes > a.lab_cap;
            if (s_skip) atomicOr(a.retry, 1u);
            if (s_err) atomicOr(a.error, 1u);
            s_cnt[CT_FLAGS] = 8u | (s_err ? 4u : 0u);
        }
        __syncthreads();
        if (tid < 8) a.counts[size_t(b) * 8 + tid] = s_cnt[tid];
        if (s_skip) continue;

        // write: every path vertex stores its base, the first one its k-mer
        // and the record.  The ranking words are no longer needed: when the
        // subgraph's labels fit, they are assembled in that LDS buffer (at the
        // offset that keeps the output's 16-byte alignment) and copied out
        // with whole 16-byte stores.  Each thread loads CB_WU vertices'
        // words before it writes any byte.
        const uint64_t lab_end = s_bbyte + s_bytes;
        const uint32_t sh = uint32_t(s_bbyte & 15u);
        const bool staged = sh + s_bytes <= CB_STAGE;
        char* s_lab = reinterpret_cast<char*>(s_buf);
        char* out0 = staged ? s_lab + sh : a.labels + s_bbyte;  // label byte s_bbyte + j is out0[j]
        for (uint32_t d0 = tid; d0 < n; d0 += CT_WG * CB_WU) {
            uint32_t info[CB_WU], pos[CB_WU], offa[CB_WU];
            Kmer<KW> x[CB_WU];
#pragma unroll
            for (uint32_t u = 0; u < CB_WU; ++u) {
                const uint32_t d = d0 + u * CT_WG;
                info[u] = d < n ? ia[d] : 0u;
                pos[u] = d < n ? posv[d] : 0u;
            }
#pragma unroll
            for (uint32_t u = 0; u < CB_WU; ++u) {
                const uint32_t d = d0 + u * CT_WG;
                offa[u] = info[u] ? offv[info[u] & 0xFFFFu] : 0u;
                if (info[u]) x[u] = tab_key<KW>(t, d);
            }
#pragma unroll
            for (uint32_t u = 0; u < CB_WU; ++u) {
                const uint32_t d = d0 + u * CT_WG;
                if (d >= n) continue;
                if (!info[u]) {
                    if ((ld_meta(&t.meta[d]) & (M_WRITER | M_CYCLE)) != (M_WRITER | M_CYCLE)) continue;
                    const uint64_t off = s_bbyte + offv[d];
                    const uint32_t nc = write_cycle<KW>(t, d, k, out0 + offv[d], off, lab_end, a.error);
                    const uint32_t ri = atomicAdd(&s_rctr, 1u);
                    if (ri < s_nrec) a.recs[s_brec + ri] = UnitigRec{off, nc + k - 1, uint16_t(b), 1};
                    continue;
                }
                const bool fwd = (info[u] >> 16) & 1u;
                const uint64_t off = s_bbyte + offa[u];
                if (pos[u] == 0) {
                    const uint32_t wn = wnv[d] & 0x1FFFFFFFu, fl = wnv[d] >> 29;
                    if (off + wn + k - 1 > lab_end) {
                        atomicOr(a.error, 1u);
                        continue;
                    }
                    put_kmer(out0 + offa[u], fwd ? x[u] : km_rc(x[u], k), k);
                    const uint32_t ri = atomicAdd(&s_rctr, 1u);
                    if (ri < s_nrec) a.recs[s_brec + ri] = UnitigRec{off, wn + k - 1, uint16_t(b), uint16_t(fl)};
                } else {
                    const uint32_t c =
                        fwd ? uint32_t(x[u].w[0] & 3u) : 3u - uint32_t((x[u].w[KW - 1] >> (2 * (kt - 1))) & 3u);
                    if (off + pos[u] + k - 1 >= lab_end) {
                        atomicOr(a.error, 1u);
                        continue;
                    }
                    out0[offa[u] + pos[u] + k - 1] = char(base_char(c));
                }
            }
        }
        if (staged) {
            // LDS bytes [sh, sh + s_bytes) go to labels [s_bbyte, lab_end);
            // 16-byte chunk c covers LDS [16c, 16c + 16) and the same offsets
            // from s_bbyte - sh (16-byte aligned)
            __syncthreads();
            char* g0 = a.labels + (s_bbyte - sh);
            const uint32_t end = sh + s_bytes;
            for (uint32_t c = tid; c < (end + 15) / 16; c += CT_WG) {
                const uint32_t lo = 16 * c;
                if (lo >= sh && lo + 16 <= end) {
                    *reinterpret_cast<uint4*>(g0 + lo) = *reinterpret_cast<const uint4*>(s_lab + lo);
                } else {
                    for (uint32_t j = lo; j < lo + 16; ++j)
                        if (j >= sh && j < end) g0[j] = s_lab[j];
                }
            }
        }
    }
}

// Subgraphs beyond CB_NMAX non-isolated vertices (up to ct_wide_max_vertices()
// and the vertices this workgroup's scratch holds): the dense tier's steps with
// 32-bit directed states - the key -> vertex index in global scratch and
// 64-bit ranking words (state | distance << 32) in three global buffers.
constexpr uint32_t CW_ROUNDS = 22;  // > log2(2 * ct_wide_max_vertices()): every path converges
static_assert((1ull << CW_ROUNDS) > 2ull * ct_wide_max_vertices(), "ranking rounds");
__device__ __forceinline__ uint64_t cw_pack(uint32_t p, uint64_t d) { return uint64_t(p) | (d << 32); }

template <int KW>
__global__ __launch_bounds__(CT_WG) void k_contract_wide(ContractArgs a) {
    __shared__ __align__(16) uint32_t s_buf[CB_JP_LDS];  // labels, when they fit
    __shared__ uint32_t s_cand[CT_CAND];
    __shared__ uint32_t s_cnt[8];
    __shared__ uint32_t s_chg[CW_ROUNDS];
    __shared__ uint32_t s_item, s_n, s_ovf, s_ncand, s_bytes, s_nrec, s_rctr, s_err, s_skip;
    __shared__ unsigned long long s_brec, s_bbyte;
    const uint32_t tid = threadIdx.x;
    const uint32_t k = a.k;
    const uint32_t kt = k - 32 * (KW - 1);
    // scratch (wide_cap vertices): key words, state, three ranking buffers of
    // 64-bit words (state | distance << 32), path info (first vertex |
    // forward << 30 | on a path << 31), position, the first vertex's vertex
    // count | end flags << 29, label offsets, the key index
    const uint64_t N = a.wide_cap;
    uint8_t* base = a.scratch + uint64_t(blockIdx.x) * a.scratch_per_wg;
    WideTab<KW> t;
    t.key = reinterpret_cast<uint64_t*>(base);
    t.meta = reinterpret_cast<uint32_t*>(base + N * KW * 8);
    uint64_t* jp0 = reinterpret_cast<uint64_t*>(base + ((N * (8 * KW + 4) + 7) & ~7ull));
    uint64_t* jp1 = jp0 + 2 * N;
    uint64_t* jp2 = jp1 + 2 * N;
    uint32_t* ia = reinterpret_cast<uint32_t*>(jp2 + 2 * N);
    uint32_t* posv = ia + N;
    uint32_t* wnv = posv + N;
    uint32_t* offv = wnv + N;
    t.aux = offv;
    t.ht = offv + N;
    t.succ = jp0;

    while (true) {
        __syncthreads();
        if (tid == 0) {
            s_item = atomicAdd(a.wide_head, 1u);
            s_n = s_ovf = s_ncand = s_bytes = s_nrec = s_rctr = s_err = s_skip = 0;
        }
        if (tid < 8) s_cnt[tid] = 0;
        if (tid < CW_ROUNDS) s_chg[tid] = 0;
        __syncthreads();
        if (s_item >= __hip_atomic_load(a.ovf2_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        const uint32_t b = a.ovf2_list[s_item];
        // the index is sized by the subgraph's vertices (an upper bound on
        // the ones it takes); a subgraph with more than this workgroup's
        // scratch holds goes on to the global pass
        const uint64_t nv = a.n_vert[b];
        if (nv > N) {
            if (tid == 0) a.ovf3_list[atomicAdd(a.ovf3_count, 1u)] = b;
            continue;
        }
        t.hcap = uint32_t(ct_wide_index_slots(nv));
        for (uint32_t i = tid; i < t.hcap; i += CT_WG) t.ht[i] = HT_FREE;
        __syncthreads();
        load_bucket_with<KW>(a, b, uint32_t(N), &s_n, &s_ovf, s_cnt, [&](uint32_t d, const Kmer<KW>& x, uint32_t m) {
#pragma unroll
            for (int w = 0; w < KW; ++w) t.key[size_t(d) * KW + w] = x.w[w];
            t.meta[d] = m;
            wide_index(t.ht, t.hcap, x, d);
            return true;
        });
        __syncthreads();
        if (s_ovf) {
            if (tid == 0) a.ovf3_list[atomicAdd(a.ovf3_count, 1u)] = b;
            continue;
        }
        const uint32_t n = s_n;
        t.cap = n;

        // link: each directed state points at its successor (1 step) or at itself (an end)
        for (uint32_t d = tid; d < n; d += CT_WG) {
            const Kmer<KW> x = tab_key<KW>(t, d);
            const Kmer<KW> xr = km_rc(x, k);
            const uint32_t m = ld_meta(&t.meta[d]);
            uint32_t add = 0;
            for (uint32_t sd = 0; sd < 2; ++sd) {
                const int nx = link_of<KW>(t, d, sd, x, xr, m, k, &s_err);
                jp0[2 * d + sd] = nx < 0 ? cw_pack(2 * d + sd, 0) : cw_pack(uint32_t(nx), 1);
                if (nx >= 0) add |= M_LINK0 << sd;
            }
            // only this thread writes vertex d's state here; the others read
            // its DFA bits, which the store keeps
            if (add) __hip_atomic_store(&t.meta[d], m | add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        __syncthreads();

        // rank: pointer jumping until no path state moves (cycle states stop
        // counting once past the vertex count).  The successor words in jp0
        // stay as they are (cycle walks follow them).
        // in global scratch, three buffers (jp0 keeps the successor words,
        // which cycle walks follow); CB_RU states per thread at a time, their
        // loads issued before any store
        const uint64_t* jp;
        {
            const uint64_t* src = jp0;
            uint64_t* dst = jp1;
            for (uint32_t r = 0; r < CW_ROUNDS; ++r) {
                bool any = false;
                for (uint32_t x0 = tid; x0 < 2 * n; x0 += CT_WG * CB_RU) {
                    uint64_t w[CB_RU], w2[CB_RU];
#pragma unroll
                    for (uint32_t u = 0; u < CB_RU; ++u) {
                        const uint32_t x = x0 + u * CT_WG;
                        w[u] = x < 2 * n ? src[x] : x;
                    }
#pragma unroll
                    for (uint32_t u = 0; u < CB_RU; ++u) {
                        const uint32_t x = x0 + u * CT_WG, p = uint32_t(w[u]);
                        w2[u] = x < 2 * n && p != x ? src[p] : 0ull;
                    }
#pragma unroll
                    for (uint32_t u = 0; u < CB_RU; ++u) {
                        const uint32_t x = x0 + u * CT_WG;
                        if (x >= 2 * n) continue;
                        const uint32_t p = uint32_t(w[u]);
                        uint64_t nw = w[u];
                        if (p != x) {
                            const uint32_t p2 = uint32_t(w2[u]);
                            const uint64_t nd = (w[u] >> 32) + (w2[u] >> 32);
                            nw = cw_pack(p2, nd);
                            any |= p2 != p && nd < n;
                        }
                        dst[x] = nw;
                    }
                }
                if (any) s_chg[r] = 1;
                __syncthreads();
                src = dst;
                dst = dst == jp1 ? jp2 : jp1;
                if (!s_chg[r]) break;
            }
            jp = src;
        }

        // path ends first: the reading of each end state, as in k_contract_big
        for (uint32_t d = tid; d < n; d += CT_WG) {
            const uint64_t w0 = jp[2 * d], w1 = jp[2 * d + 1];
            const bool e0 = w0 == cw_pack(2 * d, 0), e1 = w1 == cw_pack(2 * d + 1, 0);
            uint32_t bits = 0;
            if (e0 || e1) {
                const Kmer<KW> x = tab_key<KW>(t, d), xr = km_rc(x, k);
#pragma unroll
                for (uint32_t sd = 0; sd < 2; ++sd) {
                    if (!(sd ? e1 : e0)) continue;
                    const uint32_t Fo = uint32_t(sd ? w0 : w1);  // reached through the other side
                    const Kmer<KW> co = tab_key<KW>(t, Fo >> 1);
                    const Kmer<KW> eo = (Fo & 1u) ? km_rc(co, k) : co;
                    bits |= uint32_t(km_less(sd ? xr : x, eo)) << sd;
                }
            }
            offv[d] = bits;
        }
        __syncthreads();

        // per vertex: the ends of its path, the output orientation and its position
        for (uint32_t d = tid; d < n; d += CT_WG) {
            const uint64_t wf = jp[2 * d], wb = jp[2 * d + 1];
            const uint32_t Ff = uint32_t(wf), Fb = uint32_t(wb);
            if (jp[Fb] != cw_pack(Fb, 0)) {  // not a path end: the vertex lies on a cycle
                atomicOr(&t.meta[d], M_ONCYC);
                const uint32_t c = atomicAdd(&s_ncand, 1u);
                if (c < uint32_t(CT_CAND)) s_cand[c] = d;
                ia[d] = 0;
                continue;
            }
            const uint32_t df = uint32_t(wf >> 32), db = uint32_t(wb >> 32);
            const uint32_t L = Ff >> 1, sL = Ff & 1u, R = Fb >> 1, sR = Fb & 1u;
            const bool fwd = (offv[L] >> sL) & 1u;  // E(Ff) < E(Fb): Fb is Ff's far end
            const uint32_t pos = fwd ? df : db;
            const uint32_t A = fwd ? L : R;
            ia[d] = A | (uint32_t(fwd) << 30) | (1u << 31);
            posv[d] = pos;
            if (pos == 0) {
                const uint32_t sA = fwd ? sL : sR, B = fwd ? R : L, sB = fwd ? sR : sL;
                const uint32_t fl = (side_exits(ld_meta(&t.meta[A]), sA) ? 2u : 0u) |
                                    (side_exits(ld_meta(&t.meta[B]), sB) ? 4u : 0u);
                wnv[d] = (df + db + 1) | (fl << 29);
            }
        }
        __syncthreads();

        // the first vertex of each path reserves its label
        {
            uint32_t nrec = 0, verts = 0, lbytes = 0, triv = 0, disc = 0;
            for (uint32_t d = tid; d < n; d += CT_WG) {
                if (!ia[d] || posv[d]) continue;
                const uint32_t wn = wnv[d] & 0x1FFFFFFFu, fl = wnv[d] >> 29;
                const uint32_t bytes = wn + k - 1;
                offv[d] = atomicAdd(&s_bytes, bytes);
                ++nrec;
                verts += wn;
                lbytes += bytes;
                disc += fl != 0;
                triv += fl == 0;
            }
            if (nrec) {
                atomicAdd(&s_nrec, nrec);
                atomicAdd(&s_cnt[CT_VERTICES], verts);
                atomicAdd(&s_cnt[CT_LABEL_BYTES], lbytes);
                atomicAdd(&s_cnt[CT_TRIVIAL], triv);
                atomicAdd(&s_cnt[CT_DISC_EDGES], disc);
            }
        }
        __syncthreads();
        if (tid == 0 && s_ncand)
            cycles_serial<KW>(t, k, s_cand, s_ncand, &s_bytes, &s_nrec, s_cnt, &s_err,
                              [&](uint32_t slot) { return &offv[slot]; });
        __syncthreads();

        // reserve the bucket's output, publish its counters
        if (tid == 0) {
            s_brec = atomicAdd(&a.cursors[0], (unsigned long long)s_nrec);
            s_bbyte = atomicAdd(&a.cursors[1], (unsigned long long)s_bytes);
            s_skip = s_brec + s_nrec > a.rec_cap || s_bbyte + s_bytes > a.lab_cap;
            if (s_skip) atomicOr(a.retry, 1u);
            if (s_err) atomicOr(a.error, 1u);
            s_cnt[CT_FLAGS] = 16u | (s_err ? 4u : 0u);
        }
        __syncthreads();
        if (tid < 8) a.counts[size_t(b) * 8 + tid] = s_cnt[tid];
        if (s_skip) continue;

        // write: every path vertex stores its base, the first one its k-mer
        // and the record.  The ranking words are no longer needed: when the
        // subgraph's labels fit, they are assembled in that LDS buffer (at the
        // offset that keeps the output's 16-byte alignment) and copied out
        // with whole 16-byte stores.  Each thread loads CB_WU vertices'
        // words before it writes any byte.
        const uint64_t lab_end = s_bbyte + s_bytes;
        const uint32_t sh = uint32_t(s_bbyte & 15u);
        const bool staged = sh + s_bytes <= CB_STAGE;
        char* s_lab = reinterpret_cast<char*>(s_buf);
        char* out0 = staged ? s_lab + sh : a.labels + s_bbyte;  // label byte s_bbyte + j is out0[j]
        for (uint32_t d0 = tid; d0 < n; d0 += CT_WG * CB_WU) {
            uint32_t info[CB_WU], pos[CB_WU], offa[CB_WU];
            Kmer<KW> x[CB_WU];
#pragma unroll
            for (uint32_t u = 0; u < CB_WU; ++u) {
                const uint32_t d = d0 + u * CT_WG;
                info[u] = d < n ? ia[d] : 0u;
                pos[u] = d < n ? posv[d] : 0u;
            }
#pragma unroll
            for (uint32_t u = 0; u < CB_WU; ++u) {
                const uint32_t d = d0 + u * CT_WG;
                offa[u] = info[u] ? offv[info[u] & 0x3FFFFFFFu] : 0u;
                if (info[u]) x[u] = tab_key<KW>(t, d);
            }
#pragma unroll
            for (uint32_t u = 0; u < CB_WU; ++u) {
                const uint32_t d = d0 + u * CT_WG;
                if (d >= n) continue;
                if (!info[u]) {
                    if ((ld_meta(&t.meta[d]) & (M_WRITER | M_CYCLE)) != (M_WRITER | M_CYCLE)) continue;
                    const uint64_t off = s_bbyte + offv[d];
                    const uint32_t nc = write_cycle<KW>(t, d, k, out0 + offv[d], off, lab_end, a.error);
                    const uint32_t ri = atomicAdd(&s_rctr, 1u);
                    if (ri < s_nrec) a.recs[s_brec + ri] = UnitigRec{off, nc + k - 1, uint16_t(b), 1};
                    continue;
                }
                const bool fwd = (info[u] >> 30) & 1u;
                const uint64_t off = s_bbyte + offa[u];
                if (pos[u] == 0) {
                    const uint32_t wn = wnv[d] & 0x1FFFFFFFu, fl = wnv[d] >> 29;
                    if (off + wn + k - 1 > lab_end) {
                        atomicOr(a.error, 1u);
                        continue;
                    }
                    put_kmer(out0 + offa[u], fwd ? x[u] : km_rc(x[u], k), k);
                    const uint32_t ri = atomicAdd(&s_rctr, 1u);
                    if (ri < s_nrec) a.recs[s_brec + ri] = UnitigRec{off, wn + k - 1, uint16_t(b), uint16_t(fl)};
                } else {
                    const uint32_t c =
                        fwd ? uint32_t(x[u].w[0] & 3u) : 3u - uint32_t((x[u].w[KW - 1] >> (2 * (kt - 1))) & 3u);
                    if (off + pos[u] + k - 1 >= lab_end) {
                        atomicOr(a.error, 1u);
                        continue;
                    }
                    out0[offa[u] + pos[u] + k - 1] = char(base_char(c));
                }
            }
        }
        if (staged) {
            // LDS bytes [sh, sh + s_bytes) go to labels [s_bbyte, lab_end);
            // 16-byte chunk c covers LDS [16c, 16c + 16) and the same offsets
            // from s_bbyte - sh (16-byte aligned)
            __syncthreads();
            char* g0 = a.labels + (s_bbyte - sh);
            const uint32_t end = sh + s_bytes;
            for (uint32_t c = tid; c < (end + 15) / 16; c += CT_WG) {
                const uint32_t lo = 16 * c;
                if (lo >= sh && lo + 16 <= end) {
                    *reinterpret_cast<uint4*>(g0 + lo) = *reinterpret_cast<const uint4*>(s_lab + lo);
                } else {
                    for (uint32_t j = lo; j < lo + 16; ++j)
                        if (j >= sh && j < end) g0[j] = s_lab[j];
                }
            }
        }
    }
}

// Subgraphs whose vertices do not fit an LDS table: a table in this
// workgroup's global scratch region, sized by the subgraph's vertex count.
template <int KW>
__global__ __launch_bounds__(CT_WG) void k_contract_walk(ContractArgs a) {
    __shared__ uint32_t s_cand[CT_CAND];
    __shared__ uint32_t s_cnt[8];
    __shared__ uint32_t s_item, s_n, s_ovf, s_ncand, s_bytes, s_nrec, s_rctr, s_err, s_skip;
    __shared__ unsigned long long s_brec, s_bbyte;
    const uint32_t tid = threadIdx.x;
    const uint32_t k = a.k;

    while (true) {
        __syncthreads();
        if (tid == 0) {
            s_item = atomicAdd(a.walk_head, 1u);
            s_n = s_ovf = s_ncand = s_bytes = s_nrec = s_rctr = s_err = s_skip = 0;
        }
        if (tid < 8) s_cnt[tid] = 0;
        __syncthreads();
        if (s_item >= __hip_atomic_load(a.ovf2_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        const uint32_t b = a.ovf2_list[s_item];

        CtTab<KW> t;
        const uint32_t cap = uint32_t(ct_global_capacity(a.n_vert[b]));
        uint8_t* base = a.scratch + uint64_t(blockIdx.x) * a.scratch_per_wg;
        t.key = reinterpret_cast<uint64_t*>(base);
        t.meta = reinterpret_cast<uint32_t*>(base + uint64_t(cap) * KW * 8);
        t.aux = t.meta + cap;
        t.cap = cap;
        for (uint32_t i = tid; i < t.cap; i += CT_WG) {
            t.key[size_t(i) * KW + KW - 1] = EMPTY;
            t.meta[i] = 0;
        }
        __syncthreads();
        load_bucket(a, b, t, cap, &s_n, &s_ovf, s_cnt, nullptr);  // the capacity exceeds the vertex count
        __syncthreads();
        if (s_ovf) {
            if (tid == 0) atomicOr(a.error, 1u);
            continue;
        }

        // link
        for (uint32_t i = tid; i < t.cap; i += CT_WG) {
            if (t.key[size_t(i) * KW + KW - 1] == EMPTY) continue;
            const Kmer<KW> x = tab_key(t, i);
            const Kmer<KW> xr = km_rc(x, k);
            const uint32_t m = ld_meta(&t.meta[i]);
            uint32_t add = 0;
            for (uint32_t sd = 0; sd < 2; ++sd)
                if (link_of(t, i, sd, x, xr, m, k, &s_err) >= 0) add |= M_LINK0 << sd;
            if (add) atomicOr(&t.meta[i], add);
        }
        __syncthreads();

        // paths: walk from every end; the canonical end writes
        {
            uint32_t nrec = 0, verts = 0, lbytes = 0, triv = 0, disc = 0;
            for (uint32_t i = tid; i < t.cap; i += CT_WG) {
                if (t.key[size_t(i) * KW + KW - 1] == EMPTY) continue;
                const uint32_t m0 = ld_meta(&t.meta[i]);
                for (uint32_t sd = 0; sd < 2; ++sd) {
                    if (m0 & (M_LINK0 << sd)) continue;
                    const Kmer<KW> x = tab_key(t, i);
                    const Kmer<KW> xr = km_rc(x, k);
                    Kmer<KW> z = sd ? xr : x, zb = sd ? x : xr;  // leave through side 1 - sd
                    const Kmer<KW> first = z;
                    uint32_t side = 1u - sd, cur = i, n = 1, bb;
                    atomicOr(&t.meta[i], M_VIS);
                    for (uint32_t step = 0; step < t.cap; ++step) {
                        const uint32_t m = ld_meta(&t.meta[cur]);
                        if (!(m & (M_LINK0 << side))) break;
                        const int j = walk_step(t, m, k, z, zb, side, bb);
                        if (j < 0) {
                            s_err = 1;
                            break;
                        }
                        atomicOr(&t.meta[j], M_VIS);
                        cur = uint32_t(j);
                        ++n;
                    }
                    if (!km_less(first, zb)) continue;  // the other end's orientation is canonical
                    const bool ex = side_exits(m0, sd) || side_exits(ld_meta(&t.meta[cur]), side);
                    const uint32_t bytes = n + k - 1;
                    t.aux[i] = atomicAdd(&s_bytes, bytes);
                    atomicOr(&t.meta[i], M_WRITER | (sd ? M_START1 : 0u));
                    ++nrec;
                    verts += n;
                    lbytes += bytes;
                    disc += ex;
                    triv += !ex;
                }
            }
            if (nrec) {
                atomicAdd(&s_nrec, nrec);
                atomicAdd(&s_cnt[CT_VERTICES], verts);
                atomicAdd(&s_cnt[CT_LABEL_BYTES], lbytes);
                atomicAdd(&s_cnt[CT_TRIVIAL], triv);
                atomicAdd(&s_cnt[CT_DISC_EDGES], disc);
            }
        }
        __syncthreads();

        // cycles: vertices no path reached
        for (uint32_t i = tid; i < t.cap; i += CT_WG) {
            if (t.key[size_t(i) * KW + KW - 1] == EMPTY || (ld_meta(&t.meta[i]) & M_VIS)) continue;
            atomicOr(&t.meta[i], M_ONCYC);
            const uint32_t c = atomicAdd(&s_ncand, 1u);
            if (c < uint32_t(CT_CAND)) s_cand[c] = i;
        }
        __syncthreads();
        if (tid == 0 && s_ncand)
            cycles_serial<KW>(t, k, s_cand, s_ncand, &s_bytes, &s_nrec, s_cnt, &s_err,
                          [&](uint32_t slot) { return &t.aux[slot]; });
        __syncthreads();

        // reserve the bucket's output, publish its counters
        if (tid == 0) {
            s_brec = atomicAdd(&a.cursors[0], (unsigned long long)s_nrec);
            s_bbyte = atomicAdd(&a.cursors[1], (unsigned long long)s_bytes);
            s_skip = s_brec + s_nrec > a.rec_cap || s_bbyte + s_bytes > a.lab_cap;
            if (s_skip) atomicOr(a.retry, 1u);
            if (s_err) atomicOr(a.error, 1u);
            s_cnt[CT_FLAGS] = 2u | (s_err ? 4u : 0u);
        }
        __syncthreads();
        if (tid < 8) a.counts[size_t(b) * 8 + tid] = s_cnt[tid];
        if (s_skip) continue;

        // write: each writer walks its unitig again (the same walk as when it
        // was measured; the bounds below only guard an inconsistent table)
        const uint64_t lab_end = s_bbyte + s_bytes;
        for (uint32_t i = tid; i < t.cap; i += CT_WG) {
            if (t.key[size_t(i) * KW + KW - 1] == EMPTY) continue;
            const uint32_t m0 = t.meta[i];
            if (!(m0 & M_WRITER)) continue;
            const uint64_t off = s_bbyte + t.aux[i];
            char* out = a.labels + off;
            if (m0 & M_CYCLE) {
                const uint32_t n = write_cycle<KW>(t, i, k, out, off, lab_end, a.error);
                const uint32_t ri = atomicAdd(&s_rctr, 1u);
                if (ri < s_nrec) a.recs[s_brec + ri] = UnitigRec{off, n + k - 1, uint16_t(b), 1};
                continue;
            }
            const uint32_t sd = (m0 & M_START1) ? 1u : 0u;
            const Kmer<KW> x = tab_key(t, i);
            const Kmer<KW> xr = km_rc(x, k);
            Kmer<KW> z = sd ? xr : x, zb = sd ? x : xr;
            if (off + k > lab_end) {
                atomicOr(a.error, 1u);
                continue;
            }
            put_kmer(out, z, k);
            uint32_t side = 1u - sd, cur = i, n = 1, bb;
            for (uint32_t step = 0; step < t.cap; ++step) {
                const uint32_t m = t.meta[cur];
                if (!(m & (M_LINK0 << side))) break;
                const int j = walk_step(t, m, k, z, zb, side, bb);
                if (j < 0) break;
                if (off + k + n > lab_end) {
                    atomicOr(a.error, 1u);
                    break;
                }
                out[k - 1 + n] = char(base_char(bb));
                cur = uint32_t(j);
                ++n;
            }
            const uint16_t fl = (side_exits(m0, sd) ? 2 : 0) | (side_exits(t.meta[cur], side) ? 4 : 0);
            const uint32_t ri = atomicAdd(&s_rctr, 1u);
            if (ri < s_nrec) a.recs[s_brec + ri] = UnitigRec{off, n + k - 1, uint16_t(b), fl};
        }
    }
}

template <int KW>
void launch_kw(const ContractArgs& a, uint32_t grid_lds, uint32_t grid_global, int part, hipStream_t s) {
    if (part == 0) {
        if (grid_lds) hipLaunchKernelGGL(k_contract_rank<KW>, dim3(grid_lds), dim3(CT_WG), 0, s, a);
        if (grid_global) hipLaunchKernelGGL(k_contract_big<KW>, dim3(grid_global), dim3(CT_WG), 0, s, a);
        return;
    }
    if (!grid_global) return;
    if (a.wide_cap) {
        // the wide pass takes the dense pass's overflow; the global pass its own
        hipLaunchKernelGGL(k_contract_wide<KW>, dim3(grid_global), dim3(CT_WG), 0, s, a);
        ContractArgs w = a;
        w.ovf2_list = a.ovf3_list;
        w.ovf2_count = a.ovf3_count;
        hipLaunchKernelGGL(k_contract_walk<KW>, dim3(grid_global), dim3(CT_WG), 0, s, w);
    } else {
        hipLaunchKernelGGL(k_contract_walk<KW>, dim3(grid_global), dim3(CT_WG), 0, s, a);
    }
}

}  // namespace

void launch_contract(const ContractArgs& a, uint32_t KW, uint32_t grid_lds, uint32_t grid_global, int part,
                     hipStream_t s) {
    switch (KW) {
        case 1: launch_kw<1>(a, grid_lds, grid_global, part, s); break;
        case 2: launch_kw<2>(a, grid_lds, grid_global, part, s); break;
        case 3: launch_kw<3>(a, grid_lds, grid_global, part, s); break;
        default: launch_kw<4>(a, grid_lds, grid_global, part, s); break;
    }
}

}  // namespace cfgpu
