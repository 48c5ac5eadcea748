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
//   paths   every side that does not continue is a path end; a thread walks
//           from it to the other end, and the end whose orientation is
//           canonical (first k-mer below the reverse complement's,
//           include/Maximal_Unitig_Scratch.hpp:136-138, 198-217) writes it;
//   cycles  what no path reached lies on cycles; each is walked once and
//           written from its smallest vertex in canonical form
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
constexpr uint32_t M_WRITER = 1u << 11;  // writes its unitig (label offset in aux)
constexpr uint32_t M_CYCLE = 1u << 12;   // ... a cycle, from this vertex out of its back side
constexpr uint32_t M_START1 = 1u << 13;  // ... a path whose end side at this vertex is the back
constexpr int CT_CAND = 1024;

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

// Unique keys: claim the top word, then fill the rest (read only after a barrier).
template <int KW>
__device__ __forceinline__ bool tab_insert(const CtTab<KW>& t, const Kmer<KW>& x, uint32_t meta) {
    uint32_t s = slot_of(km_hash(x), t.cap);
    for (uint32_t n = 0; n < t.cap; ++n) {
        uint64_t* kp = &t.key[size_t(s) * KW];
        if (atomicCAS(reinterpret_cast<unsigned long long*>(kp + KW - 1), EMPTY, x.w[KW - 1]) == EMPTY) {
#pragma unroll
            for (int w = 0; w < KW - 1; ++w) kp[w] = x.w[w];
            t.meta[s] = meta;
            return true;
        }
        s = s + 1 == t.cap ? 0 : s + 1;
    }
    return false;
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

__device__ __forceinline__ uint32_t ld_meta(const uint32_t* m) {
    return __hip_atomic_load(m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One step of a walk: from the directed vertex z (rc zb) out of side `side`
// of slot `cur`, whose state `m` continues there.  Returns the next slot (-1
// if the table lacks it) and updates z, zb, side (the exit side there) and b.
template <int KW>
__device__ __forceinline__ int walk_step(const CtTab<KW>& t, uint32_t m, uint32_t k, Kmer<KW>& z, Kmer<KW>& zb,
                                         uint32_t& side, uint32_t& b) {
    const uint32_t e = side_edge(m, side);
    b = side == 1 ? e : 3u - e;
    z = km_append(z, b, k);
    zb = km_prepend(zb, uint8_t(3u - b), k);
    const bool can = km_less(z, zb);
    side = can ? 1u : 0u;  // entrance side front (canonical) -> leave through the back
    return tab_find(t, can ? z : zb);
}

template <int KW>
__device__ __forceinline__ void put_kmer(char* out, const Kmer<KW>& z, uint32_t k) {
    uint32_t p = 0;
#pragma unroll
    for (int w = KW - 1; w >= 0; --w) {  // word KW-1 holds the first k - 32(KW-1) bases
        const uint32_t nb = w == KW - 1 ? k - 32 * (KW - 1) : 32u;
        const uint64_t v = z.w[w];
        for (uint32_t q = 0; q < nb; ++q) out[p++] = "ACGT"[(v >> (2 * (nb - 1 - q))) & 3u];
    }
}

template <int KW, bool GLOBAL>
__global__ __launch_bounds__(CT_WG) void k_contract(ContractArgs a) {
    constexpr uint32_t CL = GLOBAL ? 1u : uint32_t(ct_lds_capacity(KW));
    __shared__ uint64_t s_key[CL * KW];
    __shared__ uint32_t s_meta[CL];
    __shared__ uint32_t s_aux[CL];
    __shared__ uint32_t s_cand[CT_CAND];
    __shared__ uint32_t s_cnt[8];
    __shared__ uint32_t s_item, s_n, s_ovf, s_ncand, s_bytes, s_nrec, s_rctr, s_err, s_skip;
    __shared__ unsigned long long s_brec, s_bbyte;
    const uint32_t tid = threadIdx.x;
    const uint32_t k = a.k;

    while (true) {
        __syncthreads();
        if (tid == 0) {
            s_item = atomicAdd(&a.queue_head[GLOBAL ? 1 : 0], 1u);
            s_n = s_ovf = s_ncand = s_bytes = s_nrec = s_rctr = s_err = s_skip = 0;
        }
        if (tid < 8) s_cnt[tid] = 0;
        __syncthreads();
        const uint32_t n_items =
            GLOBAL ? __hip_atomic_load(a.ovf_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : a.n_list;
        if (s_item >= n_items) break;
        const uint32_t b = GLOBAL ? a.ovf_list[s_item] : a.list[s_item];

        CtTab<KW> t;
        uint32_t limit;
        if constexpr (GLOBAL) {
            const uint32_t cap = uint32_t(ct_global_capacity(a.n_vert[b]));
            uint8_t* base = a.scratch + uint64_t(blockIdx.x) * a.scratch_per_wg;
            t.key = reinterpret_cast<uint64_t*>(base);
            t.meta = reinterpret_cast<uint32_t*>(base + uint64_t(cap) * KW * 8);
            t.aux = t.meta + cap;
            t.cap = cap;
            limit = cap;  // holds every vertex of the bucket
        } else {
            t.key = s_key;
            t.meta = s_meta;
            t.aux = s_aux;
            t.cap = CL;
            limit = CL * 4 / 5;
        }
        for (uint32_t i = tid; i < t.cap; i += CT_WG) {
            t.key[size_t(i) * KW + KW - 1] = EMPTY;
            t.meta[i] = 0;
        }
        __syncthreads();

        // load: isolated vertices counted, the others into the table
        {
            const uint8_t* tbl = reinterpret_cast<const uint8_t*>(a.tbl_ptr[b]);
            const uint64_t tcap = a.tbl_cap[b];
            uint32_t iso = 0, ph = 0;
            for (uint64_t s = tid; s < tcap; s += CT_WG) {
                const uint64_t* p = reinterpret_cast<const uint64_t*>(tbl + s * uint64_t(8 * KW + 8));
                Kmer<KW> x;
#pragma unroll
                for (int w = 0; w < KW; ++w) x.w[w] = p[w];
                if (x.w[KW - 1] == EMPTY) continue;
                const uint64_t es = p[KW];
                const uint32_t ef = uint32_t(es), st = uint32_t(es >> 32);
                const uint32_t f = edge_at(ef, 0, a.cutoff), bk = edge_at(ef, 16, a.cutoff);
                if (f == BASE_E && bk == BASE_E) {
                    ++iso;
                    ph += (st & 6u) != 0;
                    continue;
                }
                if (__hip_atomic_load(&s_ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
                const uint32_t m = f | (bk << 3) | ((st & 2u) << 5) | ((st & 4u) << 5);
                if (atomicAdd(&s_n, 1u) >= limit || !tab_insert(t, x, m)) {
                    s_ovf = 1;
                    break;
                }
            }
            if (iso) atomicAdd(&s_cnt[CT_ISOLATED], iso);
            if (ph) atomicAdd(&s_cnt[CT_PHANTOM], ph);
        }
        __syncthreads();
        if (s_ovf) {
            if (tid == 0) {
                if constexpr (GLOBAL) {
                    atomicOr(a.error, 1u);
                } else {
                    a.ovf_list[atomicAdd(a.ovf_count, 1u)] = b;
                }
            }
            continue;
        }

        // link: sides that continue the unitig
        for (uint32_t i = tid; i < t.cap; i += CT_WG) {
            if (t.key[size_t(i) * KW + KW - 1] == EMPTY) continue;
            const Kmer<KW> x = tab_key(t, i);
            const Kmer<KW> xr = km_rc(x, k);
            const uint32_t m = ld_meta(&t.meta[i]);
            uint32_t add = 0;
            for (uint32_t s = 0; s < 2; ++s) {
                const uint32_t e = side_edge(m, s);
                if (e >= 4) continue;
                Kmer<KW> z = s ? x : xr, zb = s ? xr : x;
                uint32_t side = s, bb;
                const int j = walk_step(t, m, k, z, zb, side, bb);
                if (j < 0) {
                    s_err = 1;
                    continue;
                }
                const uint32_t ent = 1u - side;
                const uint32_t ej = side_edge(ld_meta(&t.meta[j]), ent);
                if (ej == BASE_E) s_err = 1;                 // the edge is counted at both ends
                if (ej >= 4) continue;                       // branching entrance: crossed
                if (uint32_t(j) == i && ent == s) continue;  // back through the same side
                add |= M_LINK0 << s;
            }
            if (add) atomicOr(&t.meta[i], add);
        }
        __syncthreads();

        // paths: walk from every end; the canonical end writes
        {
            uint32_t nrec = 0, verts = 0, lbytes = 0, triv = 0, disc = 0;
            for (uint32_t i = tid; i < t.cap; i += CT_WG) {
                if (t.key[size_t(i) * KW + KW - 1] == EMPTY) continue;
                const uint32_t m0 = ld_meta(&t.meta[i]);
                for (uint32_t s = 0; s < 2; ++s) {
                    if (m0 & (M_LINK0 << s)) continue;
                    const Kmer<KW> x = tab_key(t, i);
                    const Kmer<KW> xr = km_rc(x, k);
                    Kmer<KW> z = s ? xr : x, zb = s ? x : xr;  // leave through side 1 - s
                    const Kmer<KW> first = z;
                    uint32_t side = 1u - s, cur = i, n = 1, bb;
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
                    const bool ex = side_exits(m0, s) || side_exits(ld_meta(&t.meta[cur]), side);
                    const uint32_t bytes = n + k - 1;
                    t.aux[i] = atomicAdd(&s_bytes, bytes);
                    atomicOr(&t.meta[i], M_WRITER | (s ? M_START1 : 0u));
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
            const uint32_t c = atomicAdd(&s_ncand, 1u);
            if (c < CT_CAND) s_cand[c] = i;
        }
        __syncthreads();
        if (tid == 0 && s_ncand) {
            const bool all = s_ncand > uint32_t(CT_CAND);
            const uint32_t nc = all ? t.cap : s_ncand;
            for (uint32_t q = 0; q < nc; ++q) {
                const uint32_t c = all ? q : s_cand[q];
                if (t.key[size_t(c) * KW + KW - 1] == EMPTY || (t.meta[c] & M_VIS)) continue;
                Kmer<KW> z = tab_key(t, c), zb = km_rc(z, k), mn = z;
                uint32_t side = 1, cur = c, n = 1, ms = c, bb;
                t.meta[c] |= M_VIS;
                for (uint32_t step = 0; step < t.cap; ++step) {
                    const uint32_t m = t.meta[cur];
                    if (!(m & (M_LINK0 << side))) {  // every side on a cycle continues
                        s_err = 1;
                        break;
                    }
                    const int j = walk_step(t, m, k, z, zb, side, bb);
                    if (j < 0) {
                        s_err = 1;
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
                t.aux[ms] = s_bytes;
                s_bytes += bytes;
                t.meta[ms] |= M_WRITER | M_CYCLE;
                s_nrec += 1;
                s_cnt[CT_TRIVIAL] += 1;
                s_cnt[CT_ICC] += 1;
                s_cnt[CT_VERTICES] += n;
                s_cnt[CT_LABEL_BYTES] += bytes;
            }
        }
        __syncthreads();

        // reserve the bucket's output, publish its counters
        if (tid == 0) {
            s_brec = atomicAdd(&a.cursors[0], (unsigned long long)s_nrec);
            s_bbyte = atomicAdd(&a.cursors[1], (unsigned long long)s_bytes);
            s_skip = s_brec + s_nrec > a.rec_cap || s_bbyte + s_bytes > a.lab_cap;
            if (s_skip) atomicOr(a.retry, 1u);
            if (s_err) atomicOr(a.error, 1u);
            s_cnt[CT_FLAGS] = (GLOBAL ? 2u : 1u) | (s_err ? 4u : 0u);
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
            const bool cyc = m0 & M_CYCLE;
            const uint32_t s = cyc ? 0u : ((m0 & M_START1) ? 1u : 0u);
            const Kmer<KW> x = tab_key(t, i);
            const Kmer<KW> xr = km_rc(x, k);
            Kmer<KW> z = s ? xr : x, zb = s ? x : xr;
            const uint64_t off = s_bbyte + t.aux[i];
            if (off + k > lab_end) {
                atomicOr(a.error, 1u);
                continue;
            }
            char* out = a.labels + off;
            put_kmer(out, z, k);
            uint32_t side = 1u - s, cur = i, n = 1, bb;
            for (uint32_t step = 0; step < t.cap; ++step) {
                const uint32_t m = t.meta[cur];
                if (!(m & (M_LINK0 << side))) break;
                const int j = walk_step(t, m, k, z, zb, side, bb);
                if (j < 0 || (cyc && uint32_t(j) == i)) break;
                if (off + k + n > lab_end) {
                    atomicOr(a.error, 1u);
                    break;
                }
                out[k - 1 + n] = "ACGT"[bb];
                cur = uint32_t(j);
                ++n;
            }
            uint16_t fl = cyc ? 1 : 0;
            if (!cyc) fl |= (side_exits(m0, s) ? 2 : 0) | (side_exits(t.meta[cur], side) ? 4 : 0);
            UnitigRec r;
            r.label_off = off;
            r.label_len = n + k - 1;
            r.g = uint16_t(b);
            r.flags = fl;
            const uint32_t ri = atomicAdd(&s_rctr, 1u);
            if (ri < s_nrec) a.recs[s_brec + ri] = r;
        }
    }
}

template <int KW>
void launch_kw(const ContractArgs& a, uint32_t grid_lds, uint32_t grid_global, hipStream_t s) {
    if (grid_lds) hipLaunchKernelGGL((k_contract<KW, false>), dim3(grid_lds), dim3(CT_WG), 0, s, a);
    if (grid_global) hipLaunchKernelGGL((k_contract<KW, true>), dim3(grid_global), dim3(CT_WG), 0, s, a);
}

}  // namespace

void launch_contract(const ContractArgs& a, uint32_t KW, uint32_t grid_lds, uint32_t grid_global, hipStream_t s) {
    switch (KW) {
        case 1: launch_kw<1>(a, grid_lds, grid_global, s); break;
        case 2: launch_kw<2>(a, grid_lds, grid_global, s); break;
        case 3: launch_kw<3>(a, grid_lds, grid_global, s); break;
        default: launch_kw<4>(a, grid_lds, grid_global, s); break;
    }
}

}  // namespace cfgpu
