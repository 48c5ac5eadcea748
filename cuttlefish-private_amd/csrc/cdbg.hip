// The compacted de Bruijn graph on the device: the per-subgraph unitigs of
// cfgpu_contract glued into every maximal unitig of the whole graph.
//
// An lm-tig end that exits lies on a discontinuity vertex; the lm-tig of the
// subgraph owning the vertex's other side (when the vertex is not isolated
// there) ends at the same vertex on the opposite side.  Gluing lm-tigs at these
// shared (vertex, side) pairs contracts the discontinuity graph and expands it
// back into labels (src/Discontinuity_Graph_Contractor.cpp,
// src/Contracted_Graph_Expander.cpp); a chain ends where a vertex has no lm-tig
// on the far side (the phantom edges, src/Discontinuity_Graph_Contractor.cpp:
// 131-190).  Every unitig record is a piece of exactly one chain (trivial
// unitigs are chains of one piece):
//
//   ends    each exiting end is keyed by its vertex's canonical k-mer and
//           moved into one of up to 8192 hash bins (a block histogram, scans,
//           a scatter: no global atomics); one workgroup per bin pairs the
//           ends of each vertex in an LDS table (the end on the other side
//           of the vertex is an end's partner);
//   rank    directed states (piece, end it is entered at) point at the next
//           piece of the chain; pointer jumping gives every state the last state
//           of its chain and the label characters of the pieces up to it;
//   place   a piece knows both chain ends, hence the orientation the chain is
//           written in (first k characters at most the reverse complement's,
//           include/Unitig_Collator.hpp:306-345), its offset in the chain and
//           the chain's length; the first piece reserves the output;
//   copy    each piece copies its label, reverse-complemented when read
//           backwards, to its place (neighbours rewrite the same k shared
//           characters).
// Chains that close on themselves (cycles through several subgraphs) are
// listed and finished on the host.
#include "cf_kernels.hpp"

namespace cfgpu {
namespace {

constexpr uint64_t EMPTY = ~0ull;
constexpr uint64_t BUSY = 1ull << 63;  // never set in a key word: top words hold < 63 bits
constexpr uint32_t NONE = 0xFFFFFFFFu;
// list-ranking word: the state pointed at (low 30 bits) | label characters up
// to it (high 34 bits, saturating)
constexpr uint32_t JT_BITS = 30;
constexpr uint64_t JT_MASK = (1ull << JT_BITS) - 1, JC_MAX = (1ull << (64 - JT_BITS)) - 1;
__device__ __forceinline__ uint64_t jword(uint64_t t, uint64_t c) { return t | (min(c, JC_MAX) << JT_BITS); }

// k characters at s as a k-mer (Kmer layout: word 0 holds the last 32
// bases).  The characters are read as aligned 4-byte words (the label buffer
// has at least 4 bytes of slack after its last label) and decoded four at a
// time: ((c >> 1) ^ (c >> 2)) & 3 maps A C G T to 0 1 2 3 in every byte.
template <int KW>
__device__ __forceinline__ Kmer<KW> enc_kmer(const char* s, uint32_t k) {
    const uint32_t* p = reinterpret_cast<const uint32_t*>(reinterpret_cast<uintptr_t>(s) & ~uintptr_t(3));
    const uint32_t sh = uint32_t(reinterpret_cast<uintptr_t>(s) & 3u);
    const uint32_t nw = (sh + k + 3) / 4;
    // every word is requested before the first is decoded (a loop over the
    // runtime count would wait for each load in turn)
    constexpr uint32_t NWM = 8 * KW + 1;  // (3 + 32 KW + 3) / 4 >= nw
    uint32_t words[NWM];
#pragma unroll
    for (uint32_t j = 0; j < NWM; ++j) words[j] = j < nw ? p[j] : 0u;
    uint64_t acc[KW + 1];
#pragma unroll
    for (int w = 0; w <= KW; ++w) acc[w] = 0;
#pragma unroll
    for (uint32_t j = 0; j < NWM; ++j) {
        const uint32_t v = words[j];
        const uint32_t c = ((v >> 1) ^ (v >> 2)) & 0x03030303u;  // byte q: code of character 4j + q
        const uint32_t b8 = ((c & 3u) << 6) | (((c >> 8) & 3u) << 4) | (((c >> 16) & 3u) << 2) | (c >> 24);
        if (j < nw) {
#pragma unroll
            for (int w = KW; w > 0; --w) acc[w] = (acc[w] << 8) | (acc[w - 1] >> 56);
            acc[0] = (acc[0] << 8) | b8;
        }
    }
    // drop the characters read past the k-mer (fewer than 4), then the ones
    // before it (the mask)
    const uint32_t t = 2 * (4 * nw - sh - k);
    Kmer<KW> f;
#pragma unroll
    for (int w = 0; w < KW; ++w) f.w[w] = t ? (acc[w] >> t) | (acc[w + 1] << (64 - t)) : acc[w];
    const uint32_t kt = k - 32 * (KW - 1);
    f.w[KW - 1] &= (1ull << (2 * kt)) - 1;
    return f;
}

// The vertex and side of piece r's end `at` (0 first k-mer, 1 last), from
// kend: the first k characters (at 0) or the reverse complement of the last
// k (at 1).  The side is the one facing the label: first k-mer -> back side
// if canonical; last k-mer -> front side if canonical.
template <int KW>
__device__ __forceinline__ void end_vertex(const Kmer<KW>& e, uint32_t at, uint32_t k, Kmer<KW>& key,
                                           uint32_t& side) {
    const Kmer<KW> er = km_rc(e, k);  // at 1: the last k characters themselves
    if (at == 0) {
        const bool can = km_less(e, er);
        key = can ? e : er;
        side = can ? 1u : 0u;
    } else {
        const bool can = km_less(er, e);
        key = can ? er : e;
        side = can ? 0u : 1u;
    }
}

__device__ __forceinline__ uint32_t bin_of(uint64_t h, uint32_t nb_bits) {
    return nb_bits ? uint32_t(h >> (64 - nb_bits)) : 0u;
}

// Pass 1, block b owns pieces [b * per_block, (b + 1) * per_block): both ends'
// k-mers into kend, partner[] cleared, and the block's count of exiting ends
// per hash bin into hist[bin * n_blocks + b].
template <int KW>
__global__ __launch_bounds__(256) void k_ends_encode(CdbgArgs a) {
    __shared__ uint32_t s_hist[CD_MAX_BINS];
    Kmer<KW>* kend = reinterpret_cast<Kmer<KW>*>(a.kend);
    const uint32_t nb = 1u << a.nb_bits;
    for (uint32_t i = threadIdx.x; i < nb; i += 256) s_hist[i] = 0;
    __syncthreads();
    const uint64_t r0 = uint64_t(blockIdx.x) * a.per_block, r1 = min(a.n, r0 + a.per_block);
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += 256) {
        const UnitigRec u = a.recs[r];
        a.piece_len[r] = u.label_len - a.k;  // for k_rank_init: 4 bytes per piece instead of a record
        const char* lab = a.labels + u.label_off;
        const Kmer<KW> A = enc_kmer<KW>(lab, a.k);
        const Kmer<KW> B = km_rc(enc_kmer<KW>(lab + u.label_len - a.k, a.k), a.k);
        kend[2 * r] = A;
        kend[2 * r + 1] = B;
        a.partner[2 * r] = a.partner[2 * r + 1] = NONE;
        for (uint32_t at = 0; at < 2; ++at) {
            if (!(u.flags & (at ? 4u : 2u))) continue;
            Kmer<KW> key;
            uint32_t side;
            end_vertex(at ? B : A, at, a.k, key, side);
            atomicAdd(&s_hist[bin_of(km_hash(key), a.nb_bits)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += 256) a.bin_hist[uint64_t(i) * a.n_blocks + blockIdx.x] = s_hist[i];
}

// Per bin (one workgroup each): exclusive scan of hist[bin][*] over the
// blocks, in place, and the bin's total.
__global__ __launch_bounds__(1024) void k_bin_row_scan(uint32_t* hist, uint32_t n_blocks, uint32_t* tot) {
    __shared__ uint32_t s[1024];
    uint32_t* row = hist + uint64_t(blockIdx.x) * n_blocks;
    const uint32_t tid = threadIdx.x;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < n_blocks; c0 += 1024) {
        const uint32_t i = c0 + tid;
        const uint32_t v = i < n_blocks ? row[i] : 0u;
        s[tid] = v;
        __syncthreads();
        for (uint32_t d = 1; d < 1024; d <<= 1) {
            const uint32_t x = tid >= d ? s[tid - d] : 0u;
            __syncthreads();
            s[tid] += x;
            __syncthreads();
        }
        if (i < n_blocks) row[i] = carry + s[tid] - v;
        carry += s[1023];
        __syncthreads();
    }
    if (tid == 0) tot[blockIdx.x] = carry;
}

// Exclusive scan of the bin totals -> off[0..nb] (one workgroup).
__global__ __launch_bounds__(1024) void k_bin_off_scan(const uint32_t* tot, uint32_t nb, uint32_t* off) {
    __shared__ uint32_t s[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (nb + 1023) / 1024;
    uint32_t sum = 0;
    for (uint32_t i = tid * per; i < min(nb, (tid + 1) * per); ++i) sum += tot[i];
    s[tid] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint32_t x = tid >= d ? s[tid - d] : 0u;
        __syncthreads();
        s[tid] += x;
        __syncthreads();
    }
    uint32_t run = s[tid] - sum;
    for (uint32_t i = tid * per; i < min(nb, (tid + 1) * per); ++i) {
        off[i] = run;
        run += tot[i];
    }
    if (tid == 1023) off[nb] = s[1023];
}

// Pass 2: every exiting end (vertex key, end id << 1 | side) into its bin's
// run, block after block (the same block ranges as pass 1).
template <int KW>
__global__ __launch_bounds__(256) void k_ends_scatter(CdbgArgs a) {
    __shared__ uint32_t s_cur[CD_MAX_BINS];
    const Kmer<KW>* kend = reinterpret_cast<const Kmer<KW>*>(a.kend);
    Kmer<KW>* bkey = reinterpret_cast<Kmer<KW>*>(a.bin_key);
    const uint32_t nb = 1u << a.nb_bits;
    for (uint32_t i = threadIdx.x; i < nb; i += 256)
        s_cur[i] = a.bin_off[i] + a.bin_hist[uint64_t(i) * a.n_blocks + blockIdx.x];
    __syncthreads();
    const uint64_t r0 = uint64_t(blockIdx.x) * a.per_block, r1 = min(a.n, r0 + a.per_block);
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += 256) {
        const uint32_t fl = a.recs[r].flags;
        for (uint32_t at = 0; at < 2; ++at) {
            if (!(fl & (at ? 4u : 2u))) continue;
            Kmer<KW> key;
            uint32_t side;
            end_vertex(kend[2 * r + at], at, a.k, key, side);
            const uint32_t pos = atomicAdd(&s_cur[bin_of(km_hash(key), a.nb_bits)], 1u);
            bkey[pos] = key;
            a.bin_id[pos] = uint32_t((2 * r + at) << 1) | side;
        }
    }
}

// Pass 3, one workgroup per bin: the bin's ends are entered in an LDS table
// keyed by vertex (P hash passes when the bin is large), one end id per side
// of each vertex; then every end reads the id on the other side of its
// vertex, its partner.  A vertex side claimed twice is an inconsistent input.
template <int KW>
struct PairTab {
    // 144 KiB of LDS for KW 1, 2 and 4 (keys + two ids per slot), 128 KiB for
    // KW 3: the largest tables one workgroup's LDS holds, so that most bins
    // take one pass
    static constexpr uint32_t CAP = KW == 1 ? 9216u : KW == 2 ? 6144u : KW == 3 ? 4096u : 3584u;
    static_assert(CAP * (8 * KW + 8) <= 147456, "pairing table in LDS");
};

template <int KW>
__device__ __forceinline__ int pair_slot(uint64_t* s_key, const Kmer<KW>& key, uint32_t h, bool insert) {
    constexpr uint32_t CAP = PairTab<KW>::CAP;
    const uint64_t top = key.w[KW - 1];
    uint32_t s = uint32_t((uint64_t(h) * CAP) >> 32);
    for (uint32_t n = 0; n < CAP;) {
        uint64_t* kp = &s_key[s * KW];
        uint64_t cur = __hip_atomic_load(kp + KW - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == EMPTY) {
            if (!insert) return -1;
            const uint64_t claim = KW == 1 ? top : (top | BUSY);
            cur = atomicCAS(reinterpret_cast<unsigned long long*>(kp + KW - 1), EMPTY, claim);
            if (cur == EMPTY) {
                if constexpr (KW > 1) {
#pragma unroll
                    for (int w = 0; w < KW - 1; ++w)
                        __hip_atomic_store(kp + w, key.w[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __hip_atomic_store(kp + KW - 1, top, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                return int(s);
            }
        }
        if (KW > 1 && cur == (top | BUSY)) continue;  // being written: the claimer publishes in its own iteration
        if (cur == top) {
            bool eq = true;
#pragma unroll
            for (int w = 0; w < KW - 1; ++w)
                eq &= __hip_atomic_load(kp + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == key.w[w];
            if (eq) return int(s);
        }
        s = s + 1 == CAP ? 0 : s + 1;
        ++n;
    }
    return -1;
}

template <int KW>
__global__ __launch_bounds__(1024) void k_ends_pair(CdbgArgs a) {
    constexpr uint32_t CAP = PairTab<KW>::CAP;
    constexpr uint32_t LOAD = CAP * 3 / 5;  // ends per pass (each vertex holds one or two)
    __shared__ uint64_t s_key[CAP * KW];
    __shared__ uint32_t s_id[2 * CAP];
    const Kmer<KW>* bkey = reinterpret_cast<const Kmer<KW>*>(a.bin_key);
    const uint32_t tid = threadIdx.x;
    const uint32_t nb = 1u << a.nb_bits;
    for (uint32_t bin = blockIdx.x; bin < nb; bin += gridDim.x) {
        const uint32_t e0 = a.bin_off[bin], e1 = a.bin_off[bin + 1];
        const uint32_t P = (e1 - e0 + LOAD - 1) / LOAD;
        for (uint32_t p = 0; p < P; ++p) {
            __syncthreads();  // the previous pass is done with the table
            for (uint32_t i = tid; i < CAP; i += 1024) {
                s_key[i * KW + KW - 1] = EMPTY;
                s_id[2 * i] = s_id[2 * i + 1] = NONE;
            }
            __syncthreads();
            for (uint32_t e = e0 + tid; e < e1; e += 1024) {
                const Kmer<KW> key = bkey[e];
                const uint32_t h = uint32_t(km_hash(key));
                if (P > 1 && uint32_t((uint64_t(fmix32(h)) * P) >> 32) != p) continue;
                const int sl = pair_slot<KW>(s_key, key, h, true);
                const uint32_t id = a.bin_id[e];
                if (sl < 0 || atomicCAS(&s_id[2 * sl + (id & 1u)], NONE, id >> 1) != NONE)
                    atomicOr(a.error, sl < 0 ? 1u : 2u);  // full table / two lm-tigs on one side of a vertex
            }
            __syncthreads();
            for (uint32_t e = e0 + tid; e < e1; e += 1024) {
                const Kmer<KW> key = bkey[e];
                const uint32_t h = uint32_t(km_hash(key));
                if (P > 1 && uint32_t((uint64_t(fmix32(h)) * P) >> 32) != p) continue;
                const int sl = pair_slot<KW>(s_key, key, h, false);
                const uint32_t id = a.bin_id[e];
                if (sl >= 0) a.partner[id >> 1] = s_id[2 * sl + 1 - (id & 1u)];
            }
        }
    }
}

// State 2r + e: piece r entered at end e.  Its successor is the piece across
// the vertex at r's other end, entered at its matching end; a state with no
// partner across the end it leaves through is a chain end.

__global__ __launch_bounds__(256) void k_rank_init(CdbgArgs a) {
    for (uint64_t st = uint64_t(blockIdx.x) * 256 + threadIdx.x; st < 2 * a.n; st += uint64_t(gridDim.x) * 256) {
        const uint32_t next = a.partner[st ^ 1];  // the end across the vertex it leaves through
        a.jp[st] = next == NONE ? jword(st, 0) : jword(next, a.piece_len[next >> 1]);
    }
}

// One pointer-jumping pass, in place: a state's word is replaced by the
// combination with the word of the state it points at.  Words are read and
// written whole, so every word always names a state further along the chain
// with the characters up to it, whatever other states have already jumped.
// A state that is a chain end, or points at one, never changes again: its bit
// in `settled` is set (one atomic per 32 states: a wave's lanes cover 64
// consecutive states) and later passes skip it without reading the word it
// points at.
__global__ __launch_bounds__(256) void k_rank_round(CdbgArgs a, uint32_t* changed) {
    bool any = false;
    const int lane = int(threadIdx.x & 63u);
    for (uint64_t st = uint64_t(blockIdx.x) * 256 + threadIdx.x; st < 2 * a.n; st += uint64_t(gridDim.x) * 256) {
        bool done = (a.settled[st >> 5] >> (st & 31)) & 1u;
        bool now = false;
        if (!done) {
            const uint64_t w = __hip_atomic_load(&a.jp[st], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint64_t t = w & JT_MASK;
            if (t == st) {  // chain end
                now = true;
            } else {
                const uint64_t w2 = __hip_atomic_load(&a.jp[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t t2 = w2 & JT_MASK;
                if (t2 == t) {  // points at a chain end already
                    now = true;
                } else {
                    const uint64_t c = (w >> JT_BITS) + (w2 >> JT_BITS);
                    __hip_atomic_store(&a.jp[st], jword(t2, c), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    any |= c < a.total_chars;  // chains settle; cycles stop counting past every character
                }
            }
        }
        const unsigned long long m = __ballot(now);
        if ((lane == 0 || lane == 32) && uint32_t(m >> lane)) atomicOr(&a.settled[st >> 5], uint32_t(m >> lane));
    }
    if (any) atomicOr(changed, 1u);
}

// Per state, as bits of 64-state words: whether it is a chain end (no
// partner across the end it leaves through) and, for an end E, whether the
// chain read from E comes first: K(E) < K(F) for F = E's far end (where the
// reading of E's piece the other way stops) and K(E) = the first k characters
// of that reading (kend of E's piece on the side it leaves through backwards).
// k_place then reads two bits per piece instead of the partners of both its
// ends and the k-mers at both ends of its chain.
template <int KW>
__global__ __launch_bounds__(256) void k_chain_bits(CdbgArgs a) {
    const Kmer<KW>* kend = reinterpret_cast<const Kmer<KW>*>(a.kend);
    const uint32_t lane = threadIdx.x & 63u;
    const auto K = [&](uint64_t e) { return kend[2 * (e >> 1) + 1 - (e & 1u)]; };
    for (uint64_t s0 = uint64_t(blockIdx.x) * 256 + (threadIdx.x & ~63u); s0 < 2 * a.n;
         s0 += uint64_t(gridDim.x) * 256) {  // wave-uniform, 64-state words
        const uint64_t st = s0 + lane;
        bool end = false, first = false;
        if (st < 2 * a.n) {
            end = a.partner[st ^ 1] == NONE;
            if (end) first = km_less(K(st), K(a.jp[st ^ 1] & JT_MASK));
        }
        const unsigned long long me = __ballot(end), mf = __ballot(first);
        if (lane == 0) {
            a.end_bits[s0 >> 6] = me;
            a.first_bits[s0 >> 6] = mf;
        }
    }
}

// Per piece: chain ends, orientation, offset and length of its chain.  The
// first piece of each chain (its head) reserves the chain's output bytes and
// record: a workgroup takes PL_Q pieces per thread, scans its heads' lengths
// and makes one reservation per PL_Q * 256 pieces (reservations per wave made
// the two counters' device-scope atomics the kernel's bound).
constexpr uint32_t PL_Q = 8;
template <int KW>
__global__ __launch_bounds__(256) void k_place(CdbgArgs a) {
    __shared__ unsigned long long s_wb[4], s_wc[4], s_base, s_rbase;
    const uint64_t* jp = a.jp;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
    const auto is_end = [&](uint64_t e) { return (a.end_bits[e >> 6] >> (e & 63u)) & 1u; };
    for (uint64_t r0 = uint64_t(blockIdx.x) * 256 * PL_Q; r0 < a.n; r0 += uint64_t(gridDim.x) * 256 * PL_Q) {
        uint64_t len[PL_Q];
        uint32_t heads = 0;  // bit i: piece r0 + 256 i + tid heads its chain
        uint64_t my_bytes = 0;
#pragma unroll
        for (uint32_t i = 0; i < PL_Q; ++i) {
            const uint64_t r = r0 + 256 * i + tid;
            len[i] = 0;
            if (r >= a.n) continue;
            const UnitigRec u = a.recs[r];
            const uint64_t w0 = jp[2 * r], w1 = jp[2 * r + 1];
            const uint64_t E0 = w0 & JT_MASK, E1 = w1 & JT_MASK;
            // E0 and E1 are where the piece's two readings stop: a chain's
            // ends, or, on a cycle, wherever the jumping stopped (a cycle
            // state can land on itself)
            if (!is_end(E0) || !is_end(E1)) {  // a cycle
                a.place_head[r] = NONE;
                atomicAdd(a.n_cycle_pieces, 1u);
                continue;
            }
            const uint64_t c0 = w0 >> JT_BITS, c1 = w1 >> JT_BITS;
            // read the chain with this piece forward when the reading from E1
            // (the end of direction 1) comes first: its first k characters
            // below the reverse complement of the last k at E0 (k_chain_bits);
            // a cycle unitig of one subgraph is as written
            const bool fwd = (u.flags & 1u) || ((a.first_bits[E1 >> 6] >> (E1 & 63u)) & 1u);
            a.place_head[r] = uint32_t(fwd ? (E1 >> 1) : (E0 >> 1));
            // the piece ends its chain when nothing follows it in that reading
            const uint64_t last = (fwd ? c0 : c1) == 0 ? 1u : 0u;
            a.place_pos[r] = (fwd ? c1 : c0) | (last << 62) | (uint64_t(fwd ? 0u : 1u) << 63);
            if (fwd ? E1 == 2 * r + 1 : E0 == 2 * r) {  // nothing before it in that reading
                len[i] = c0 + c1 + u.label_len;
                heads |= 1u << i;
                my_bytes += len[i];
            }
        }
        // exclusive prefix of the heads' bytes and counts over the workgroup
        uint64_t ib = my_bytes, ic = uint64_t(__popc(heads));
        const uint64_t my_cnt = ic;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t vb = __shfl_up(ib, d), vc = __shfl_up(ic, d);
            if (int(lane) >= d) ib += vb, ic += vc;
        }
        if (lane == 63) s_wb[wv] = ib, s_wc[wv] = ic;
        __syncthreads();
        if (tid == 0) {
            const unsigned long long tb = s_wb[0] + s_wb[1] + s_wb[2] + s_wb[3];
            const unsigned long long tc = s_wc[0] + s_wc[1] + s_wc[2] + s_wc[3];
            s_base = tb ? atomicAdd(&a.cursors[1], tb) : 0ull;
            s_rbase = tc ? atomicAdd(&a.cursors[0], tc) : 0ull;
        }
        __syncthreads();
        uint64_t off = s_base + ib - my_bytes, ri = s_rbase + ic - my_cnt;
        for (uint32_t wb = 0; wb < wv; ++wb) off += s_wb[wb], ri += s_wc[wb];
        __syncthreads();  // s_wb, s_wc and the bases read before the next round writes them
#pragma unroll
        for (uint32_t i = 0; i < PL_Q; ++i) {
            if (!((heads >> i) & 1u)) continue;
            const uint64_t r = r0 + 256 * i + tid;
            a.head_off[r] = off;
            if (off + len[i] <= a.out_cap && ri < a.out_n_cap) {
                a.out_rec[2 * ri] = off;
                a.out_rec[2 * ri + 1] = len[i];
            } else {
                atomicOr(a.retry, 1u);
            }
            off += len[i];
            ++ri;
        }
    }
}

// Four characters at s (any alignment) as one little-endian word: two aligned
// loads and a byte funnel shift (the label buffer has slack after its last
// label, and s - (s & 3) is inside it).
__device__ __forceinline__ uint32_t load4(const char* s) {
    const uintptr_t ad = reinterpret_cast<uintptr_t>(s);
    const uint32_t* q = reinterpret_cast<const uint32_t*>(ad & ~uintptr_t(3));
    return __builtin_amdgcn_alignbyte(q[1], q[0], uint32_t(ad & 3u));  // ({q1, q0} >> 8 * (ad & 3))[31:0]
}

// Sixteen characters at s (any alignment): five aligned loads and four byte
// funnel shifts (the label buffer's slack covers the fifth word).
__device__ __forceinline__ uint4 load16(const char* s) {
    const uintptr_t ad = reinterpret_cast<uintptr_t>(s);
    const uint32_t* q = reinterpret_cast<const uint32_t*>(ad & ~uintptr_t(3));
    const uint32_t sh = uint32_t(ad & 3u);
    const uint32_t q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4];
    return make_uint4(__builtin_amdgcn_alignbyte(q1, q0, sh), __builtin_amdgcn_alignbyte(q2, q1, sh),
                      __builtin_amdgcn_alignbyte(q3, q2, sh), __builtin_amdgcn_alignbyte(q4, q3, sh));
}

// Complement of four ASCII bases A C G T: C and G (bit 1 set) differ by
// 0x04, A and T by 0x15.
__device__ __forceinline__ uint32_t comp4(uint32_t v) {
    return v ^ (0x15151515u ^ (((v >> 1) & 0x01010101u) * 0x11u));
}

// Eight lanes per piece.  Every aligned W-byte word of the output that lies
// inside a piece's range is written by that piece (W = 16 for k >= 16, else
// 4: neighbours overlap by k >= W - 1 characters, so each aligned word inside
// a chain lies inside one of its pieces, and two pieces writing one word
// write the same characters); the bytes of the partial words at the chain's
// two ends are written one by one by its first and its last piece.
__global__ __launch_bounds__(256) void k_copy(CdbgArgs a) {
    const uint32_t j = threadIdx.x & 7u;
    for (uint64_t r = (uint64_t(blockIdx.x) * 256 + threadIdx.x) >> 3; r < a.n; r += (uint64_t(gridDim.x) * 256) >> 3) {
        const uint32_t hd = a.place_head[r];
        if (hd == NONE) continue;
        const UnitigRec u = a.recs[r];
        const uint64_t pp = a.place_pos[r];
        const uint64_t pos = pp & ((1ull << 62) - 1);
        const bool rev = pp >> 63, last = (pp >> 62) & 1u;
        const uint64_t hoff = a.head_off[hd];  // all-ones unless the head placed its chain
        if (hoff >= a.out_cap || pos >= a.out_cap - hoff || a.out_cap - hoff - pos < u.label_len) {
            if (j == 0) atomicOr(a.error, 4u);
            continue;
        }
        const char* src = a.labels + u.label_off;
        const uint32_t L = u.label_len;
        const uint64_t d0 = hoff + pos, d1 = d0 + L;  // the piece's output bytes
        // whole aligned words of W bytes: 16 when neighbours overlap by k >= 16
        // characters (every aligned 16-byte word inside a chain then lies
        // inside one of its pieces), else 4
        const uint64_t W = a.k >= 16 ? 16 : 4;
        const uint64_t w0 = (d0 + W - 1) & ~(W - 1), w1 = d1 & ~(W - 1);
        if (W == 16) {
            for (uint64_t x = w0 + 16 * j; x < w1; x += 128) {
                const uint32_t i = uint32_t(x - d0);  // output characters i .. i+15 of the piece
                uint4 v;
                if (rev) {
                    const uint4 f = load16(src + (L - 16 - i));
                    v = make_uint4(comp4(__builtin_bswap32(f.w)), comp4(__builtin_bswap32(f.z)),
                                   comp4(__builtin_bswap32(f.y)), comp4(__builtin_bswap32(f.x)));
                } else {
                    v = load16(src + i);
                }
                *reinterpret_cast<uint4*>(a.out + x) = v;
            }
        } else {
            for (uint64_t x = w0 + 4 * j; x < w1; x += 32) {
                const uint32_t i = uint32_t(x - d0);  // output characters i .. i+3 of the piece
                const uint32_t v = rev ? comp4(__builtin_bswap32(load4(src + (L - 4 - i)))) : load4(src + i);
                *reinterpret_cast<uint32_t*>(a.out + x) = v;
            }
        }
        const auto put = [&](uint64_t x) {
            const uint32_t i = uint32_t(x - d0);
            const char c = rev ? src[L - 1 - i] : src[i];
            a.out[x] = rev ? (c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A') : c;
        };
        if (pos == 0)
            for (uint64_t x = d0 + j; x < min(w0, d1); x += 8) put(x);
        if (last)
            for (uint64_t x = max(w1, d0) + j; x < d1; x += 8) put(x);
    }
}

__global__ __launch_bounds__(256) void k_cycle_list(CdbgArgs a, uint32_t* list, uint32_t* n) {
    for (uint64_t r = uint64_t(blockIdx.x) * 256 + threadIdx.x; r < a.n; r += uint64_t(gridDim.x) * 256)
        if (a.place_head[r] == NONE) list[atomicAdd(n, 1u)] = uint32_t(r);
}

template <int KW>
void launch_kw(const CdbgArgs& a, uint32_t grid, int phase, uint32_t src, uint32_t* changed, hipStream_t s) {
    switch (phase) {
        case 0: {
            const uint32_t nb = 1u << a.nb_bits;
            hipLaunchKernelGGL(k_ends_encode<KW>, dim3(a.n_blocks), dim3(256), 0, s, a);
            hipLaunchKernelGGL(k_bin_row_scan, dim3(nb), dim3(1024), 0, s, a.bin_hist, a.n_blocks, a.bin_tot);
            hipLaunchKernelGGL(k_bin_off_scan, dim3(1), dim3(1024), 0, s, a.bin_tot, nb, a.bin_off);
            hipLaunchKernelGGL(k_ends_scatter<KW>, dim3(a.n_blocks), dim3(256), 0, s, a);
            hipLaunchKernelGGL(k_ends_pair<KW>, dim3(grid < nb ? grid : nb), dim3(1024), 0, s, a);
            break;
        }
        case 1: hipLaunchKernelGGL(k_rank_init, dim3(grid), dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_rank_round, dim3(grid), dim3(256), 0, s, a, changed); break;
        case 3:
            hipLaunchKernelGGL(k_chain_bits<KW>, dim3(grid), dim3(256), 0, s, a);
            hipLaunchKernelGGL(k_place<KW>, dim3(grid), dim3(256), 0, s, a);
            break;
        default: hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, a); break;
    }
}

}  // namespace

void launch_cdbg_phase(const CdbgArgs& a, uint32_t KW, uint32_t grid, int phase, uint32_t src, uint32_t* changed,
                       hipStream_t s) {
    switch (KW) {
        case 1: launch_kw<1>(a, grid, phase, src, changed, s); break;
        case 2: launch_kw<2>(a, grid, phase, src, changed, s); break;
        case 3: launch_kw<3>(a, grid, phase, src, changed, s); break;
        default: launch_kw<4>(a, grid, phase, src, changed, s); break;
    }
}

void launch_cdbg_cycle_list(const CdbgArgs& a, uint32_t grid, uint32_t* list, uint32_t* n, hipStream_t s) {
    hipLaunchKernelGGL(k_cycle_list, dim3(grid), dim3(256), 0, s, a, list, n);
}

}  // namespace cfgpu
