// Device-side primitives shared by the partition and construct kernels (gfx950).
//
// Encodings (all integer work):
//  * base codes A=0 C=1 G=2 T=3 (include/DNA.hpp:17-25); E (empty) = 5, N = 4.
//  * input stream: 2-bit bases packed MSB-first, 32 per 64-bit word: base j at
//    bits 63-2(j%32) .. 62-2(j%32) of word j/32.
//  * super k-mer label: W words MSB-aligned, word W-1 holds bases 0..31
//    (include/Super_Kmer_Chunk.hpp:394-405, include/Subgraph.hpp:403-411).
//  * k-mer keys: Kmer<k> layout (include/Kmer.hpp:45-54), KW = ceil(k/32) words,
//    w[0] = least-significant 64 bits, base 0 of the k-mer most significant.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cfgpu {

constexpr uint8_t BASE_N = 4;
constexpr uint8_t BASE_E = 5;

// Cuttlefish-salted wyhash constants for 8-byte keys with seed 0
// (include/Minimizer_Utility.hpp:26, 97-107); seed0 = wymix(s0, s1).
struct WyConst {
    uint64_t s0, s1, seed0;
};

// Full 64x64 -> 128-bit product from four 32x32+64 multiply-adds
// (v_mad_u64_u32); no partial sum overflows 64 bits.
__device__ __forceinline__ void mul128(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
    const uint64_t a0 = uint32_t(a), a1 = a >> 32, b0 = uint32_t(b), b1 = b >> 32;
    const uint64_t p00 = a0 * b0;
    const uint64_t t = a0 * b1 + (p00 >> 32);
    const uint64_t u = a1 * b0 + uint32_t(t);
    lo = (u << 32) | uint32_t(p00);
    hi = a1 * b1 + (t >> 32) + (u >> 32);
}

__device__ __forceinline__ uint64_t lmer_hash(uint64_t x, const WyConst& c) {
    // wyhash(&x, 8, 0, salt) (include/wyhash/wyhash.h:118-142) specialised to
    // len == 8: a = rot32(x), b = x, then two 64x64->128 multiply-folds.
    uint64_t a = ((x << 32) | (x >> 32)) ^ c.s1;
    uint64_t b = x ^ c.seed0;
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    a = lo ^ c.s0 ^ 8ull;
    b = hi ^ c.s1;
    mul128(a, b, lo, hi);
    return lo ^ hi;
}

// Reverse the order of the 32 2-bit groups of a word.
__device__ __forceinline__ uint64_t rev2(uint64_t x) {
    const uint64_t r = __builtin_bitreverse64(x);
    return ((r >> 1) & 0x5555555555555555ull) | ((r & 0x5555555555555555ull) << 1);
}

// Reverse complement of an MSB-first l-mer of 2l bits held in the low bits.
__device__ __forceinline__ uint64_t lmer_rc(uint64_t x, unsigned l) {
    const uint64_t mask = l == 32 ? ~0ull : ((1ull << (2 * l)) - 1);
    return (rev2(x) >> (64 - 2 * l)) ^ mask;
}

// 64 bases starting at base `pos` of a forward-ordered MSB-first stream whose
// word 0 holds bases [base0, base0+32).  `words` must hold word idx+1.
__device__ __forceinline__ uint64_t stream_get64(const uint64_t* words, uint64_t pos) {
    const uint64_t wi = pos >> 5;
    const unsigned sh = unsigned(pos & 31) << 1;
    const uint64_t hi = words[wi];
    return sh ? (hi << sh) | (words[wi + 1] >> (64 - sh)) : hi;
}

// 64 bases starting at base `pos` of an MSB-aligned label of W words.  Words
// past the label's end read as zero.
__device__ __forceinline__ uint64_t label_get64(const uint64_t* L, unsigned W, unsigned pos) {
    const unsigned wi = pos >> 5;
    const unsigned sh = (pos & 31) << 1;
    const uint64_t hi = wi < W ? L[W - 1 - wi] : 0;
    if (!sh) return hi;
    const uint64_t lo = wi + 1 < W ? L[W - 2 - wi] : 0;
    return (hi << sh) | (lo >> (64 - sh));
}

__device__ __forceinline__ uint8_t label_base(const uint64_t* L, unsigned W, unsigned i) {
    return uint8_t((L[W - 1 - (i >> 5)] >> (62 - ((i & 31) << 1))) & 3);
}

__device__ __forceinline__ uint8_t comp_base(uint8_t b) { return b < 4 ? uint8_t(3 - b) : b; }

// ---------------------------------------------------------------------------
// Multi-word k-mers (KW words, runtime k with 32(KW-1) < k <= 32 KW).
template <int KW>
struct Kmer {
    uint64_t w[KW];
};

template <int KW>
__device__ __forceinline__ bool km_less(const Kmer<KW>& a, const Kmer<KW>& b) {
#pragma unroll
    for (int i = KW - 1; i > 0; --i)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i];
    return a.w[0] < b.w[0];
}

template <int KW>
__device__ __forceinline__ bool km_eq(const Kmer<KW>& a, const Kmer<KW>& b) {
    bool e = true;
#pragma unroll
    for (int i = 0; i < KW; ++i) e &= a.w[i] == b.w[i];
    return e;
}

// k-mer at label offset i (bases i..i+k-1).
template <int KW>
__device__ __forceinline__ Kmer<KW> km_from_label(const uint64_t* L, unsigned W, unsigned i, unsigned k) {
    Kmer<KW> x;
    const unsigned kt = k - 32 * (KW - 1);  // bases in the top word (1..31 for odd k)
    x.w[KW - 1] = label_get64(L, W, i) >> (64 - 2 * kt);
#pragma unroll
    for (int j = KW - 2; j >= 0; --j) x.w[j] = label_get64(L, W, i + kt + 32 * (KW - 2 - j));
    return x;
}

// Reverse complement (include/Kmer.hpp:479-506).
template <int KW>
__device__ __forceinline__ Kmer<KW> km_rc(const Kmer<KW>& x, unsigned k) {
    Kmer<KW> r;
    // Reverse all 32*KW groups, then shift right by s = 64*KW - 2k bits.
    uint64_t t[KW];
#pragma unroll
    for (int i = 0; i < KW; ++i) t[i] = rev2(x.w[KW - 1 - i]);
    const unsigned s = 64 * KW - 2 * k;  // 2..62, never 0 (k odd)
#pragma unroll
    for (int i = 0; i < KW; ++i) {
        uint64_t v = t[i] >> s;
        if (i + 1 < KW) v |= t[i + 1] << (64 - s);
        r.w[i] = ~v;
    }
    const unsigned kt = k - 32 * (KW - 1);
    r.w[KW - 1] &= (1ull << (2 * kt)) - 1;
    return r;
}

// k-mer rolled one base backward: prepend base b, drop the last base.
template <int KW>
__device__ __forceinline__ Kmer<KW> km_prepend(const Kmer<KW>& x, uint8_t b, unsigned k) {
    Kmer<KW> r;
#pragma unroll
    for (int i = 0; i < KW; ++i) {
        uint64_t v = x.w[i] >> 2;
        if (i + 1 < KW) v |= x.w[i + 1] << 62;
        r.w[i] = v;
    }
    const unsigned kt = k - 32 * (KW - 1);
    r.w[KW - 1] |= uint64_t(b) << (2 * (kt - 1));
    return r;
}

// Table hash (any hash is admissible: the reference's unitig extraction is
// independent of table iteration order, include/Kmer_Hasher.hpp:12-21).
__device__ __forceinline__ uint64_t fmix64(uint64_t h) {
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return h;
}

// murmur3's 32-bit finaliser.
// Wave-wide scans by DPP lane moves (no LDS traffic).  A 64-lane wave is 4
// rows of 16 lanes: row_shr / row_shl move values within a row, row_bcast
// 15 / 31 carry a row's last lane into the rows above, wave_shr / wave_shl
// move every lane by one.  A lane whose source is outside the row (or the
// wave) receives `old`.
namespace dpp {
constexpr int ROW_SHL(int n) { return 0x100 + n; }
constexpr int ROW_SHR(int n) { return 0x110 + n; }
constexpr int WAVE_SHL1 = 0x130, WAVE_SHR1 = 0x138, ROW_BCAST15 = 0x142, ROW_BCAST31 = 0x143;

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t move(uint32_t old, uint32_t src) {
    return uint32_t(__builtin_amdgcn_update_dpp(int(old), int(src), CTRL, ROW_MASK, 0xF, false));
}

// Inclusive scan over lanes 0..lane with an associative op whose identity is id.
template <class Op>
__device__ __forceinline__ uint32_t incl_scan(uint32_t x, uint32_t id, Op op) {
    x = op(x, move<ROW_SHR(1)>(id, x));
    x = op(x, move<ROW_SHR(2)>(id, x));
    x = op(x, move<ROW_SHR(4)>(id, x));
    x = op(x, move<ROW_SHR(8)>(id, x));
    x = op(x, move<ROW_BCAST15, 0xA>(id, x));  // rows 1, 3 take rows 0, 2
    x = op(x, move<ROW_BCAST31, 0xC>(id, x));  // rows 2, 3 take rows 0..1
    return x;
}

// Inclusive scan over lanes lane..63 (a suffix scan) with identity id.
template <class Op>
__device__ __forceinline__ uint32_t suffix_scan(uint32_t x, uint32_t id, Op op) {
    x = op(x, move<ROW_SHL(1)>(id, x));
    x = op(x, move<ROW_SHL(2)>(id, x));
    x = op(x, move<ROW_SHL(4)>(id, x));
    x = op(x, move<ROW_SHL(8)>(id, x));
    // a row's first lane now holds the whole row; rows above lane's row
    const uint32_t r1 = __builtin_amdgcn_readlane(x, 16), r2 = __builtin_amdgcn_readlane(x, 32),
                   r3 = __builtin_amdgcn_readlane(x, 48);
    const uint32_t row = __lane_id() >> 4;
    const uint32_t from3 = row < 3 ? r3 : id, from2 = row < 2 ? op(r2, from3) : from3,
                   from1 = row < 1 ? op(r1, from2) : from2;
    return op(x, from1);
}

// The value of lane - 1 (lane 0: old) and of lane + 1 (lane 63: old).
__device__ __forceinline__ uint32_t from_prev_lane(uint32_t old, uint32_t x) { return move<WAVE_SHR1>(old, x); }
__device__ __forceinline__ uint32_t from_next_lane(uint32_t old, uint32_t x) { return move<WAVE_SHL1>(old, x); }
}  // namespace dpp

__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// 32-bit hash of N u64 words for choosing a pass and a probe group of an
// LDS table: each word folded to 32 bits (low half xor an odd multiple of the
// high half), chained through fmix32.  Pass = top of h * P, group = the
// fraction below it times the group count, so the two are independent.
template <int N>
__device__ __forceinline__ uint32_t pass_hash32(const uint64_t (&w)[N]) {
    uint32_t h = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const uint32_t f = uint32_t(w[i]) ^ (uint32_t(w[i] >> 32) * 0x9e3779b1u);
        h = i ? fmix32(h) ^ f : f;
    }
    return fmix32(h);
}

template <int KW>
__device__ __forceinline__ uint64_t km_hash(const Kmer<KW>& x) {
    uint64_t h = x.w[0];
#pragma unroll
    for (int i = 1; i < KW; ++i) h = fmix64(h) ^ x.w[i];
    return fmix64(h);
}

// Saturating 4-bit counters (include/State_Config.hpp:207-217): add m to the
// front field `fb` (if < 4) and the back field `bb` (if < 4), each field
// saturating at 15 (m single adds give the same word).
__device__ __forceinline__ uint32_t ef_add(uint32_t e, uint8_t fb, uint8_t bb, uint32_t m = 1) {
    if (fb < 4) {
        const unsigned off = 4u * fb;
        const uint32_t f = (e >> off) & 0xFu;
        const uint32_t nf = f + m < 15u ? f + m : 15u;
        e += (nf - f) << off;
    }
    if (bb < 4) {
        const unsigned off = 16u + 4u * bb;
        const uint32_t f = (e >> off) & 0xFu;
        const uint32_t nf = f + m < 15u ? f + m : 15u;
        e += (nf - f) << off;
    }
    return e;
}

// Edge_Frequency::edge_at (include/State_Config.hpp:232-252) of the side whose
// counters start at bit `off` (0 front, 16 back): E, the one base, or N.
__device__ __forceinline__ uint32_t edge_at(uint32_t e_f, unsigned off, uint32_t f_th) {
    uint32_t c = 0, which = 0;
#pragma unroll
    for (uint32_t bse = 0; bse < 4; ++bse)
        if (((e_f >> (off + 4 * bse)) & 0xFu) >= f_th) ++c, which = bse;
    return c == 0 ? BASE_E : (c == 1 ? which : BASE_N);
}

// True when some counter of e_f is at least f_th (1 <= f_th <= 15): the
// vertex is in an edge at the cutoff on one of its sides (edge_at of that side
// is not E).  The even and the odd 4-bit counters as bytes v, each v + 128 -
// f_th reaching bit 7 exactly when v >= f_th (no carry leaves a byte).
__device__ __forceinline__ bool ef_active(uint32_t e_f, uint32_t f_th) {
    const uint32_t add = (128u - f_th) * 0x01010101u;
    return (((e_f & 0x0F0F0F0Fu) + add) | (((e_f >> 4) & 0x0F0F0F0Fu) + add)) & 0x80808080u;
}

}  // namespace cfgpu
