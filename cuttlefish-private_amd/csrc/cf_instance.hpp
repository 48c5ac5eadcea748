// Per-k-mer-instance decode shared by the construct kernels: the body of the
// loop in Subgraph::construct (src/Subgraph.cpp:57-105) for one k-mer of a
// weak super k-mer staged in LDS.
#pragma once

#include "cf_device.hpp"

namespace cfgpu {

// Exclusive sum-scan of one u32 per thread across a WG-thread workgroup;
// returns the exclusive prefix and writes the total to *total.
template <int WG>
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t* s_wave, uint32_t* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(x, d);
        if (lane >= d) x += u;
    }
    if (lane == 63) s_wave[wid] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int j = 0; j < WG / 64; ++j) {
        const uint32_t t = s_wave[j];
        pre += j < wid ? t : 0;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

template <int KW>
struct Inst {
    Kmer<KW> key;     // canonical k-mer
    uint8_t front;    // base added to the front side's counters (BASE_E: none)
    uint8_t back;     // base added to the back side's counters
    uint32_t disc;    // Vertex_Status discontinuity bits: 2 front, 4 back
};

// Instance j of the records staged in LDS: s_kofs[r] is the index of record r's
// first k-mer (exclusive scan of len-k+1), ns records.
template <int KW>
__device__ __forceinline__ Inst<KW> decode_instance(const uint32_t* s_att, const uint64_t* s_lab,
                                                    const uint32_t* s_kofs, uint32_t ns, uint32_t W,
                                                    uint32_t k, uint32_t j) {
    int lo = 0, hi = int(ns) - 1;  // largest r with s_kofs[r] <= j
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_kofs[mid] <= j) lo = mid;
        else hi = mid - 1;
    }
    const uint32_t at = s_att[lo];
    const uint32_t i = j - s_kofs[lo];
    const uint32_t len = at & 0xFFu;
    const uint64_t* L = s_lab + lo * W;

    const Kmer<KW> v = km_from_label<KW>(L, W, i, k);
    const Kmer<KW> vb = km_rc<KW>(v, k);
    const bool is_can = km_less<KW>(v, vb);  // Directed_Vertex::in_canonical_form
    const uint8_t pred = i > 0 ? label_base(L, W, i - 1) : BASE_E;
    const uint8_t succ = i + k < len ? label_base(L, W, i + k) : BASE_E;
    Inst<KW> r;
    r.front = is_can ? pred : comp_base(succ);
    r.back = is_can ? succ : comp_base(pred);
    if (i > 0) {
        // Self-loop rule (src/Subgraph.cpp:82-83): the previous k-mer has the
        // same canonical form iff it equals v or rc(v).
        const Kmer<KW> u = km_prepend<KW>(v, pred, k);
        if (km_eq<KW>(u, v) || km_eq<KW>(u, vb)) (is_can ? r.front : r.back) = BASE_E;
    }
    r.disc = 0;  // src/Subgraph.cpp:88-92 with entrance/exit sides
    if (i == 0 && ((at >> 8) & 1u)) r.disc |= is_can ? 2u : 4u;
    if (i + k == len && ((at >> 9) & 1u)) r.disc |= is_can ? 4u : 2u;
    r.key = is_can ? v : vb;
    return r;
}

// Stage records [rbase, rbase+ns) of the bucket arrays into LDS and scan their
// k-mer counts; returns the number of k-mer instances.  Ends with a barrier.
template <int WG>
__device__ __forceinline__ uint32_t stage_records(const uint32_t* att, const uint64_t* labels,
                                                  uint64_t rbase, uint32_t ns, uint32_t W, uint32_t k,
                                                  uint32_t* s_att, uint64_t* s_lab, uint32_t* s_kofs,
                                                  uint32_t* s_wave) {
    const int tid = threadIdx.x;
    uint32_t kc = 0;
    if (uint32_t(tid) < ns) {
        const uint32_t at = att[rbase + tid];
        s_att[tid] = at;
        kc = (at & 0xFFu) - (k - 1);
    }
    for (uint32_t i = tid; i < ns * W; i += WG) s_lab[i] = labels[rbase * W + i];
    uint32_t total;
    const uint32_t ex = block_excl_sum<WG>(kc, s_wave, &total);
    s_kofs[tid] = ex;
    if (tid == WG - 1) s_kofs[WG] = ex + kc;
    __syncthreads();
    return total;
}

}  // namespace cfgpu
