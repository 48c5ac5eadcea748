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
//   ends    each exiting end is entered in a hash table keyed by its vertex's
//           canonical k-mer, in the slot half of its side; the other half holds
//           the partner end;
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

__device__ __forceinline__ uint32_t ccode(char c) { return c == 'A' ? 0u : c == 'C' ? 1u : c == 'G' ? 2u : 3u; }

// k characters at s as a k-mer (Kmer layout: word 0 holds the last 32 bases),
// f, and its reverse complement, r, in one pass.
// The characters are read as aligned 4-byte words (the label buffer has at
// least 4 bytes of slack after its last label).
template <int KW>
__device__ __forceinline__ void enc2(const char* s, uint32_t k, Kmer<KW>& f, Kmer<KW>& r) {
#pragma unroll
    for (int w = 0; w < KW; ++w) f.w[w] = r.w[w] = 0;
    const uint32_t kt = k - 32 * (KW - 1);
    const uint32_t* p = reinterpret_cast<const uint32_t*>(reinterpret_cast<uintptr_t>(s) & ~uintptr_t(3));
    const uint32_t sh = uint32_t(reinterpret_cast<uintptr_t>(s) & 3u);
    const uint32_t nw = (sh + k + 3) / 4;
#pragma unroll 4
    for (uint32_t j = 0; j < nw; ++j) {
        const uint32_t v = p[j];
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t i = 4 * j + q;  // character i - sh of the k-mer
            if (i < sh || i >= sh + k) continue;
            const uint32_t b = ccode(char((v >> (8 * q)) & 0xFFu));
#pragma unroll
            for (int w = KW - 1; w > 0; --w) f.w[w] = (f.w[w] << 2) | (f.w[w - 1] >> 62);
            f.w[0] = (f.w[0] << 2) | b;
            r = km_prepend(r, uint8_t(3u - b), k);
        }
    }
    f.w[KW - 1] &= (1ull << (2 * kt)) - 1;
}

__device__ __forceinline__ uint32_t slot_of(uint64_t h, uint64_t cap) {
    return uint32_t((uint64_t(uint32_t(h >> 32)) * cap) >> 32);
}

template <int KW>
__global__ __launch_bounds__(256) void k_ends_insert(CdbgArgs a) {
    Kmer<KW>* kend = reinterpret_cast<Kmer<KW>*>(a.kend);
    for (uint64_t r = uint64_t(blockIdx.x) * 256 + threadIdx.x; r < a.n; r += uint64_t(gridDim.x) * 256) {
        const UnitigRec u = a.recs[r];
        const char* lab = a.labels + u.label_off;
        Kmer<KW> A0, A1, B0, B1;  // first k characters and their rc; last k characters and their rc
        enc2<KW>(lab, a.k, A0, A1);
        enc2<KW>(lab + u.label_len - a.k, a.k, B0, B1);
        kend[2 * r] = A0;
        kend[2 * r + 1] = B1;
        for (uint32_t at = 0; at < 2; ++at) {
            a.end_slot[2 * r + at] = NONE;
            if (!(u.flags & (at ? 4u : 2u))) continue;
            // the end's vertex (canonical k-mer) and the side facing the label:
            // first k-mer -> back side if canonical; last k-mer -> front side if canonical
            const bool can = at ? km_less(B0, B1) : km_less(A0, A1);
            const uint32_t side = at ? (can ? 0u : 1u) : (can ? 1u : 0u);
            Kmer<KW> x;
#pragma unroll
            for (int w = 0; w < KW; ++w) x.w[w] = at ? (can ? B0.w[w] : B1.w[w]) : (can ? A0.w[w] : A1.w[w]);
            uint32_t s = slot_of(km_hash(x), a.hcap);
            // claim a slot: the top key word with BUSY set while the lower words
            // are written (a lane that finds BUSY retries the slot; it never
            // waits inside this iteration)
            const uint64_t top = KW == 1 ? x.w[0] : (x.w[KW - 1] | BUSY);
            for (uint64_t n = 0; n < 2 * a.hcap;) {
                uint64_t* kp = &a.hkey[uint64_t(s) * KW];
                uint64_t cur = __hip_atomic_load(kp + KW - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cur == EMPTY) {
                    cur = atomicCAS(reinterpret_cast<unsigned long long*>(kp + KW - 1), EMPTY, top);
                    if (cur == EMPTY) {
#pragma unroll
                        for (int w = 0; w < KW - 1; ++w)
                            __hip_atomic_store(kp + w, x.w[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (KW > 1)
                            __hip_atomic_store(kp + KW - 1, x.w[KW - 1], __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                        cur = x.w[KW - 1];
                    } else if (KW > 1) {
                        continue;  // lost the claim: look again
                    }
                }
                if (KW > 1 && cur == (x.w[KW - 1] | BUSY)) continue;  // being written: look again
                bool eq = cur == x.w[KW - 1];
                if (KW > 1 && eq) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // published: its lower words
#pragma unroll
                for (int w = 0; w < KW - 1; ++w)
                    eq &= eq && __hip_atomic_load(kp + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == x.w[w];
                if (eq) {
                    if (atomicCAS(&a.hval[2 * uint64_t(s) + side], NONE, uint32_t(2 * r + at)) != NONE)
                        atomicOr(a.error, 2u);  // two lm-tigs on one side of a vertex
                    a.end_slot[2 * r + at] = (s << 1) | side;
                    break;
                }
                s = s + 1 == a.hcap ? 0 : s + 1;
                if (++n == a.hcap) atomicOr(a.error, 1u);
            }
        }
    }
}

// State 2r + e: piece r entered at end e.  Its successor is the piece across
// the vertex at r's other end, entered at its matching end.
// State st has no partner across the end it leaves through: a chain end.
__device__ __forceinline__ bool rank_end(const CdbgArgs& a, uint64_t st) {
    const uint32_t es = a.end_slot[st ^ 1];
    return es == NONE || a.hval[2 * uint64_t(es >> 1) + (1u - (es & 1u))] == NONE;
}

__global__ __launch_bounds__(256) void k_rank_init(CdbgArgs a) {
    for (uint64_t st = uint64_t(blockIdx.x) * 256 + threadIdx.x; st < 2 * a.n; st += uint64_t(gridDim.x) * 256) {
        const uint32_t es = a.end_slot[st ^ 1];  // the end it leaves through
        const uint32_t next = es == NONE ? NONE : a.hval[2 * uint64_t(es >> 1) + (1u - (es & 1u))];
        a.jp[st] = next == NONE ? jword(st, 0) : jword(next, a.recs[next >> 1].label_len - a.k);
    }
}

// One pointer-jumping pass, in place: a state's word is replaced by the
// combination with the word of the state it points at.  Words are read and
// written whole, so every word always names a state further along the chain
// with the characters up to it, whatever other states have already jumped.
__global__ __launch_bounds__(256) void k_rank_round(CdbgArgs a, uint32_t* changed) {
    bool any = false;
    for (uint64_t st = uint64_t(blockIdx.x) * 256 + threadIdx.x; st < 2 * a.n; st += uint64_t(gridDim.x) * 256) {
        const uint64_t w = __hip_atomic_load(&a.jp[st], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t t = w & JT_MASK;
        if (t == st) continue;  // chain end
        const uint64_t w2 = __hip_atomic_load(&a.jp[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t t2 = w2 & JT_MASK;
        if (t2 == t) continue;  // points at a chain end already
        const uint64_t c = (w >> JT_BITS) + (w2 >> JT_BITS);
        __hip_atomic_store(&a.jp[st], jword(t2, c), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        any |= c < a.total_chars;  // chains settle; cycles stop counting past every character
    }
    if (any) atomicOr(changed, 1u);
}

// Per piece: chain ends, orientation, offset and length of its chain; the
// first piece of each chain reserves the output (wave-aggregated).
template <int KW>
__global__ __launch_bounds__(256) void k_place(CdbgArgs a) {
    const uint64_t* jp = a.jp;
    const int lane = int(threadIdx.x & 63u);
    for (uint64_t r0 = uint64_t(blockIdx.x) * 256; r0 < a.n; r0 += uint64_t(gridDim.x) * 256) {
        const uint64_t r = r0 + threadIdx.x;
        bool head = false;
        uint64_t len = 0;
        if (r < a.n) {
            const UnitigRec u = a.recs[r];
            const uint64_t w0 = jp[2 * r], w1 = jp[2 * r + 1];
            const uint64_t E0 = w0 & JT_MASK, E1 = w1 & JT_MASK;
            // E0 and E1 are where the piece's two readings stop: a chain's ends,
            // or, on a cycle, wherever the jumping stopped (a cycle state can land
            // on itself).  Ends are the states with no partner (as in k_rank_init).
            if (!rank_end(a, E0) || !rank_end(a, E1)) {  // a cycle
                a.place_head[r] = NONE;
                atomicAdd(a.n_cycle_pieces, 1u);
            } else {
                const uint64_t c0 = w0 >> JT_BITS, c1 = w1 >> JT_BITS;
                bool fwd = true;  // read the chain with this piece forward
                if (!(u.flags & 1u)) {
                    // first piece in that reading: the end state of direction 1,
                    // read from its unpaired end; last piece: the end of direction 0
                    // kend[2j] = first k characters of piece j, kend[2j+1] = rc of its last k
                    const Kmer<KW>* kend = reinterpret_cast<const Kmer<KW>*>(a.kend);
                    const uint32_t jf = E1 >> 1, af = 1u - (E1 & 1u), jl = E0 >> 1, al = E0 & 1u;
                    const Kmer<KW> first = kend[2 * uint64_t(jf) + af];
                    const Kmer<KW> rlast = kend[2 * uint64_t(jl) + (1u - al)];  // rc of the last k characters
                    fwd = km_less(first, rlast);
                }
                const uint32_t hd = uint32_t(fwd ? (E1 >> 1) : (E0 >> 1));
                a.place_head[r] = hd;
                a.place_pos[r] = (fwd ? c1 : c0) | (uint64_t(fwd ? 0u : 1u) << 63);
                head = fwd ? E1 == 2 * r + 1 : E0 == 2 * r;  // nothing before it in that reading
                len = c0 + c1 + u.label_len;
            }
        }
        // heads reserve their chain's bytes: one atomic per wave
        uint64_t incl = head ? len : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t v = __shfl_up(incl, d);
            if (lane >= d) incl += v;
        }
        const uint64_t tot = __shfl(incl, 63);
        const unsigned long long hb = __ballot(head);
        uint64_t base = 0, rbase = 0;
        if (lane == 0 && tot) {
            base = atomicAdd(&a.cursors[1], (unsigned long long)tot);
            rbase = atomicAdd(&a.cursors[0], (unsigned long long)__popcll(hb));
        }
        base = __shfl(base, 0);
        rbase = __shfl(rbase, 0);
        if (head) {
            const uint64_t off = base + incl - len;
            const uint64_t ri = rbase + uint64_t(__popcll(hb & ((1ull << lane) - 1)));
            a.head_off[r] = off;
            if (off + len <= a.out_cap && ri < a.out_n_cap) {
                a.out_rec[2 * ri] = off;
                a.out_rec[2 * ri + 1] = len;
            } else {
                atomicOr(a.retry, 1u);
            }
        }
    }
}

// Eight lanes per piece, each copying every eighth character.
__global__ __launch_bounds__(256) void k_copy(CdbgArgs a) {
    const uint32_t j = threadIdx.x & 7u;
    for (uint64_t r = (uint64_t(blockIdx.x) * 256 + threadIdx.x) >> 3; r < a.n; r += (uint64_t(gridDim.x) * 256) >> 3) {
        const uint32_t hd = a.place_head[r];
        if (hd == NONE) continue;
        const UnitigRec u = a.recs[r];
        const uint64_t pos = a.place_pos[r] & ~(1ull << 63);
        const bool rev = a.place_pos[r] >> 63;
        const uint64_t hoff = a.head_off[hd];  // all-ones unless the head placed its chain
        if (hoff >= a.out_cap || pos >= a.out_cap - hoff || a.out_cap - hoff - pos < u.label_len) {
            if (j == 0) atomicOr(a.error, 4u);
            continue;
        }
        const char* src = a.labels + u.label_off;
        char* dst = a.out + hoff + pos;
        if (!rev) {
            for (uint32_t i = j; i < u.label_len; i += 8) dst[i] = src[i];
        } else {
            for (uint32_t i = j; i < u.label_len; i += 8) {
                const char c = src[u.label_len - 1 - i];
                dst[i] = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_cycle_list(CdbgArgs a, uint32_t* list, uint32_t* n) {
    for (uint64_t r = uint64_t(blockIdx.x) * 256 + threadIdx.x; r < a.n; r += uint64_t(gridDim.x) * 256)
        if (a.place_head[r] == NONE) list[atomicAdd(n, 1u)] = uint32_t(r);
}

template <int KW>
void launch_kw(const CdbgArgs& a, uint32_t grid, int phase, uint32_t src, uint32_t* changed, hipStream_t s) {
    switch (phase) {
        case 0: hipLaunchKernelGGL(k_ends_insert<KW>, dim3(grid), dim3(256), 0, s, a); break;
        case 1: hipLaunchKernelGGL(k_rank_init, dim3(grid), dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_rank_round, dim3(grid), dim3(256), 0, s, a, changed); break;
        case 3: hipLaunchKernelGGL(k_place<KW>, dim3(grid), dim3(256), 0, s, a); break;
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
