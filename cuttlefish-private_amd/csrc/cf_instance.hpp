// Per-k-mer-instance decode shared by the construct kernels: the body of the
// loop in Subgraph::construct (src/Subgraph.cpp:57-105), walked record by
// record.
#pragma once

#include "cf_device.hpp"

namespace cfgpu {

template <int KW>
struct Inst {
    Kmer<KW> key;     // canonical k-mer
    uint8_t front;    // base added to the front side's counters (BASE_E: none)
    uint8_t back;     // base added to the back side's counters
    uint32_t disc;    // Vertex_Status discontinuity bits: 2 front, 4 back
};

// One weak super k-mer held in registers and walked k-mer by k-mer: the loop of
// Subgraph::construct (src/Subgraph.cpp:57-105) with the k-mer and its reverse
// complement rolled one base per step (Kmer::roll_to_next_kmer,
// include/Kmer.hpp:510-521; Directed_Vertex::roll_forward, Directed_Vertex.hpp:257-269)
// instead of re-extracted.  WMAX bounds the label words (W <= 2 * KW).
template <int KW, int WMAX>
struct RecordWalker {
    uint64_t S[WMAX];      // bases from position i + k on, MSB first (shift register)
    Kmer<KW> v, vb, prev;  // current k-mer, its reverse complement, previous k-mer
    uint32_t len, k, kt, i, n, at;

    // Fetch record r of a bucket array with W label words per record (word
    // W-1 holds bases 0..31) into registers: att and the label words with
    // bases 0..31 first.
    __device__ __forceinline__ static void fetch(const uint32_t* att, const uint64_t* labels, uint64_t r,
                                                 uint32_t W, uint32_t& at_out, uint64_t (&L)[WMAX]) {
        at_out = att[r];
#pragma unroll
        for (int q = 0; q < WMAX; ++q) L[q] = uint32_t(q) < W ? labels[r * W + (W - 1 - q)] : 0;
    }

    // Start walking a fetched record.  Every index below is static, so the
    // walker stays in registers.
    __device__ __forceinline__ void init(uint32_t at_in, const uint64_t (&L)[WMAX], uint32_t k_) {
        at = at_in;
        len = at & 0xFFu;
        k = k_;
        kt = k - 32 * (KW - 1);
#pragma unroll
        for (int q = 0; q < WMAX; ++q) S[q] = L[q];
        // First k-mer: the top word holds bases 0..kt-1, word j < KW-1 holds
        // the 32 bases from kt + 32(KW-2-j).
        v.w[KW - 1] = S[0] >> (64 - 2 * kt);
#pragma unroll
        for (int j = KW - 2; j >= 0; --j) {
            const int m = KW - 2 - j;
            v.w[j] = (S[m] << (2 * kt)) | (S[m + 1] >> (64 - 2 * kt));
        }
        vb = km_rc<KW>(v, k);
        // Advance the shift register by k bases: whole words, then bits.
        for (uint32_t c = 0; c < (k >> 5); ++c) {
#pragma unroll
            for (int q = 0; q + 1 < WMAX; ++q) S[q] = S[q + 1];
            S[WMAX - 1] = 0;
        }
        const uint32_t bs = 2 * (k & 31);
        if (bs) {
#pragma unroll
            for (int q = 0; q < WMAX; ++q) S[q] = (S[q] << bs) | (q + 1 < WMAX ? S[q + 1] >> (64 - bs) : 0);
        }
        i = 0;
        n = len - k + 1;
    }

    __device__ __forceinline__ void load(const uint32_t* att, const uint64_t* labels, uint64_t r,
                                         uint32_t W, uint32_t k_) {
        uint32_t a;
        uint64_t L[WMAX];
        fetch(att, labels, r, W, a, L);
        init(a, L, k_);
    }

    // Decode the current instance and advance.
    __device__ __forceinline__ Inst<KW> next() {
        const uint8_t succ = i + k < len ? uint8_t(S[0] >> 62) : BASE_E;
        const uint8_t pred = i > 0 ? uint8_t((prev.w[KW - 1] >> (2 * (kt - 1))) & 3u) : BASE_E;
        const bool is_can = km_less<KW>(v, vb);  // Directed_Vertex::in_canonical_form
        Inst<KW> r;
        r.front = is_can ? pred : comp_base(succ);
        r.back = is_can ? succ : comp_base(pred);
        // Self-loop rule (src/Subgraph.cpp:82-83): the previous k-mer has the
        // same canonical form iff it equals v or rc(v).
        if (i > 0 && (km_eq<KW>(prev, v) || km_eq<KW>(prev, vb))) {
            if (is_can) r.front = BASE_E;
            else r.back = BASE_E;
        }
        r.disc = 0;  // src/Subgraph.cpp:88-92 with entrance/exit sides
        if (i == 0 && ((at >> 8) & 1u)) r.disc |= is_can ? 2u : 4u;
        if (i + k == len && ((at >> 9) & 1u)) r.disc |= is_can ? 4u : 2u;
#pragma unroll
        for (int t = 0; t < KW; ++t) r.key.w[t] = is_can ? v.w[t] : vb.w[t];
        prev = v;
        if (succ < 4) {
            // v <- v[1..k) + succ; vb <- comp(succ) + vb[0..k-1)
#pragma unroll
            for (int t = KW - 1; t > 0; --t) v.w[t] = (v.w[t] << 2) | (v.w[t - 1] >> 62);
            v.w[0] = (v.w[0] << 2) | succ;
            v.w[KW - 1] &= (1ull << (2 * kt)) - 1;
#pragma unroll
            for (int t = 0; t < KW - 1; ++t) vb.w[t] = (vb.w[t] >> 2) | (vb.w[t + 1] << 62);
            vb.w[KW - 1] = (vb.w[KW - 1] >> 2) | (uint64_t(3 - succ) << (2 * (kt - 1)));
#pragma unroll
            for (int q = 0; q < WMAX; ++q) S[q] = (S[q] << 2) | (q + 1 < WMAX ? S[q + 1] >> 62 : 0);
        }
        ++i;
        return r;
    }
};

}  // namespace cfgpu
