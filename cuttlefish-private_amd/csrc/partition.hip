// Partition kernels: (k-1)-mer minimizer hashing, weak super k-mer split and
// bucket scatter (src/Graph_Partitioner.cpp:424-614 restated data-parallel).
//
// Work is tiled over the concatenated fragment stream: workgroup t owns the
// (k-1)-mer start positions [t*P_TILE, (t+1)*P_TILE).  A position p is a
// (k-1)-mer iff [p, p+k-2] lies inside one fragment.  Pass 1 writes, per
// position, g = H & (G-1) (and 4 more hash bits for the bucket features) where
// H is the minimum canonical l-mer hash over the (k-1)-mer's w = k-l l-mers.
// Pieces are then cut at every g change and every w (k-1)-mers from a run's
// start, which is the length cap `2(k-1)-l` of the reference's sequential loop
// (src/Graph_Partitioner.cpp:30, 541).
//
// Records reach their buckets in two levels, with no per-record global
// atomics: big tiles (P_SUB sub-tiles) count their records per coarse bin
// (g >> FINE_BITS); a scan gives each (coarse bin, big tile) a contiguous run,
// which the emit pass fills; then one workgroup per coarse bin moves its
// records into the 2^FINE_BITS final buckets.  Each output run is written by
// a single workgroup, so HBM lines are written whole.
#include <algorithm>

#include "cf_kernels.hpp"

namespace cfgpu {
namespace {

typedef __attribute__((ext_vector_type(2))) uint64_t u64x2;


constexpr int WIN_MAX = 128;                      // w = k - l <= 126
constexpr int NWORDS = (P_TILE + 2 * WIN_MAX) / 32 + 4;
constexpr int S_MAX = P_TILE + WIN_MAX;           // l-mer hashes per tile
constexpr int SPAN_MAX = 64;                      // largest power of two <= w (w <= 126)
constexpr int PER_T = (S_MAX + P_WG - 1) / P_WG;

// tile_frag[t] = the fragment holding position max(t * P_TILE - 1, 0): the
// largest f with frag_off[f] <= that position.  Fragment f covers the tiles
// whose position falls in [frag_off[f], frag_off[f+1]) (frag_off[0] = 0,
// strictly increasing), so every tile is written exactly once.  One lane per
// fragment; the tiles of a fragment spanning more than 8 tiles (a genome
// record) are written by its whole wave.
__global__ __launch_bounds__(256) void k_tile_frag(const uint64_t* __restrict__ frag_off, uint64_t n_frag,
                                                   uint64_t tiles, uint64_t* __restrict__ tile_frag) {
    const uint32_t lane = threadIdx.x & 63u;
    for (uint64_t fw = uint64_t(blockIdx.x) * 256 + (threadIdx.x & ~63u); fw < n_frag;
         fw += uint64_t(gridDim.x) * 256) {  // wave-uniform
        const uint64_t f = fw + lane;
        uint64_t t0 = 0, t1 = 0;
        if (f < n_frag) {
            const uint64_t b = frag_off[f], e = frag_off[f + 1];
            // tiles t >= 1 with t * P_TILE - 1 in [b, e); tile 0 (position 0) belongs to f = 0
            t0 = f == 0 ? 0 : (b + P_TILE) / P_TILE;  // ceil((b + 1) / P_TILE)
            t1 = max(t0, min((e + P_TILE) / P_TILE, tiles));  // ceil((e + 1) / P_TILE)
        }
        const bool wide = t1 - t0 > 8;
        if (!wide)
            for (uint64_t t = t0; t < t1; ++t) tile_frag[t] = f;
        for (unsigned long long m = __ballot(wide); m; m &= m - 1) {
            const int src = __ffsll((long long)m) - 1;
            const uint64_t a0 = __shfl(t0, src), a1 = __shfl(t1, src);
            for (uint64_t t = a0 + lane; t < a1; t += 64) tile_frag[t] = fw + uint64_t(src);
        }
    }
}

__global__ __launch_bounds__(P_WG) void k_minimizer_g(PartitionArgs a, uint16_t* __restrict__ gbin,
                                                      uint32_t* __restrict__ mbn, uint64_t* __restrict__ tile_lc) {
    __shared__ uint64_t s_words[NWORDS];
    __shared__ uint64_t s_m[S_MAX + SPAN_MAX];  // hashes, then window minima; ~0 past the tile's l-mers
    __shared__ uint32_t s_inv[(P_TILE + 1 + 31) / 32];  // bit i: position q0 + i starts no (k-1)-mer
    __shared__ unsigned long long s_lc;
    uint32_t* s_g = reinterpret_cast<uint32_t*>(s_m);  // reuses s_m once the minima are read

    const int tid = threadIdx.x;
    const int64_t t0 = int64_t(blockIdx.x) * P_TILE;
    const int64_t q0 = t0 - 1;  // position held at LDS index 0
    const int64_t N = int64_t(a.n_bases);
    const unsigned k = a.k, l = a.l, w = a.w;
    const int S = P_TILE + int(w);  // l-mers q0 .. q0 + P_TILE + w - 1

    // Fragments from the one containing max(q0, 0) (found for every tile
    // beforehand, k_tile_frag) to the one containing the next tile's first
    // position, plus two: fragments are longer than k, so no later boundary
    // lies within k - 2 of this tile.
    if (tid == 0) s_lc = 0;
    for (int i = tid; i < (P_TILE + 1 + 31) / 32; i += P_WG) s_inv[i] = 0;
    const uint64_t f0 = a.tile_frag[blockIdx.x];
    const uint64_t f1 = blockIdx.x + 1 < gridDim.x ? a.tile_frag[blockIdx.x + 1] + 2 : a.n_frag;
    const int64_t wlo = (q0 < 0 ? 0 : q0) >> 5;
    for (int i = tid; i < NWORDS; i += P_WG) {
        const uint64_t wi = uint64_t(wlo) + uint64_t(i);
        s_words[i] = wi < a.n_words ? a.packed[wi] : 0;
    }
    // Position p starts a (k-1)-mer iff no fragment boundary b has
    // p < b <= p + k - 2: each boundary marks [b - k + 2, b - 1].
    __syncthreads();  // s_inv cleared
    for (uint64_t f = f0 + uint64_t(tid); f <= min(f1, a.n_frag); f += P_WG) {
        const int64_t b = int64_t(a.frag_off[f]);
        const int64_t lo = max(int64_t(0), b - int64_t(k) + 2 - q0), hi = min(int64_t(P_TILE), b - 1 - q0);
        for (int64_t wi = lo >> 5; wi <= (hi >> 5) && lo <= hi; ++wi) {  // bits lo..hi, a word at a time
            const int64_t a0 = max(lo, wi << 5), a1 = min(hi, (wi << 5) + 31);
            const uint32_t n = uint32_t(a1 - a0 + 1);
            const uint32_t mask = (n == 32 ? ~0u : ((1u << n) - 1)) << (a0 & 31);
            atomicOr(&s_inv[wi], mask);
        }
    }
    __syncthreads();

    // Canonical l-mer hashes (include/Minimizer_Iterator.hpp:291-296).  The
    // l-mer at q0 + i exists for i in [i_lo, i_hi]; it starts at stream bit
    // pair i + soff of s_words.  Tile-relative 32-bit indices, no branches.
    const int i_lo = q0 < 0 ? 1 : 0;
    const int i_hi = int(min(N - int64_t(l) - q0, int64_t(S)));
    const int soff = int(q0 - (wlo << 5));  // -1 only in tile 0, where i = 0 has no l-mer
    const uint64_t lmask = l == 32 ? ~0ull : ((1ull << (2 * l)) - 1);
    // Eight consecutive l-mers per thread: their 32 bases as one MSB-first
    // word `seg` and its reverse complement `rseg` (base q of rseg = the
    // complement of base 31 - q of seg); l-mer j is bits [64 - 2(j+l), 64 - 2j)
    // of seg and its reverse complement bits [2j, 2j + 2l) of rseg (l <= 25).
    for (int g0 = tid * 8; g0 < S; g0 += P_WG * 8) {
        const int p0 = g0 + soff;
        if (l <= 25 && p0 >= 0) {
            const uint32_t wi = uint32_t(p0) >> 5, sh = (uint32_t(p0) & 31u) << 1;
            const uint64_t seg = (s_words[wi] << sh) | ((s_words[wi + 1] >> 1) >> (63u - sh));
            const uint64_t rseg = ~rev2(seg);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int i = g0 + j;
                const bool ok = i < S && i >= i_lo && i <= i_hi;
                const uint64_t x = (seg >> (64 - 2 * l - 2 * j)) & lmask;
                const uint64_t xr = (rseg >> (2 * j)) & lmask;
                const uint64_t h1 = lmer_hash(x, a.wy), h2 = lmer_hash(xr, a.wy);
                if (i < S) s_m[i] = ok ? (h1 < h2 ? h1 : h2) : ~0ull;
            }
        } else {
            for (int i = g0; i < g0 + 8 && i < S; ++i) {
                const bool ok = i >= i_lo && i <= i_hi;
                const uint32_t pos = ok ? uint32_t(i + soff) : 0u;
                const uint32_t wi = pos >> 5, sh = (pos & 31u) << 1;
                const uint64_t x = ((s_words[wi] << sh) | ((s_words[wi + 1] >> 1) >> (63u - sh))) >> (64 - 2 * l);
                const uint64_t h1 = lmer_hash(x, a.wy), h2 = lmer_hash(lmer_rc(x, l), a.wy);
                s_m[i] = ok ? (h1 < h2 ? h1 : h2) : ~0ull;
            }
        }
    }
    if (tid < SPAN_MAX) s_m[S + tid] = ~0ull;  // reads past the last l-mer see the maximum
    __syncthreads();

    // Window minimum over w consecutive l-mers by doubling: after the loop
    // s_m[i] = min(h[i .. i+span-1]) with span the largest power of two <= w.
    unsigned span = 1;
    for (; 2 * span <= w; span *= 2) {
        uint64_t tmp[PER_T];
#pragma unroll
        for (int c = 0; c < PER_T; ++c) {
            const int i = tid + c * P_WG;
            if (i < S) {
                const uint64_t v = s_m[i], u = s_m[i + span];
                tmp[c] = u < v ? u : v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < PER_T; ++c) {
            const int i = tid + c * P_WG;
            if (i < S) s_m[i] = tmp[c];
        }
        __syncthreads();
    }

    constexpr int GPT = (P_TILE + 1 + P_WG - 1) / P_WG;
    const int i_end = int(min(N - q0, int64_t(P_TILE + 1)));  // positions q0 + i < N: i < i_end
    uint32_t gvs[GPT];
#pragma unroll
    for (int c = 0; c < GPT; ++c) {
        const int i = tid + c * P_WG;
        const int ic = i <= P_TILE ? i : P_TILE;
        const bool ok = i >= i_lo && i < i_end && !((s_inv[ic >> 5] >> (ic & 31)) & 1u);
        const uint64_t m0 = s_m[ic], m1 = s_m[ic + w - span];
        const uint64_t H = m0 < m1 ? m0 : m1;
        const uint32_t gv = ok ? uint32_t(H & a.g_mask) | (uint32_t((H >> a.g_bits) & 15u) << 16) : G_INVALID;
        gvs[c] = gv;
        if (i >= 1 && i < i_end) gbin[q0 + i] = uint16_t(gv);  // G16_INVALID for G_INVALID
    }
    __syncthreads();  // every minimum read: s_g may overwrite s_m
#pragma unroll
    for (int c = 0; c < GPT; ++c) {
        const int i = tid + c * P_WG;
        if (i <= P_TILE) s_g[i] = gvs[c];
    }
    __syncthreads();

    {  // the 4 hash bits above g of positions t0 + 8 tid .. t0 + 8 tid + 7, a nibble each
        static_assert(P_TILE == 8 * P_WG, "one nibble word per thread");
        uint32_t word = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) word |= ((s_g[1 + 8 * tid + j] >> 16) & 15u) << (4 * j);
        if (t0 + 8 * tid < N) mbn[(t0 >> 3) + tid] = word;
    }
    unsigned long long my = 0;
    for (int i = tid + 1; i <= P_TILE; i += P_WG) {
        const uint32_t cur = s_g[i];
        if (cur == G_INVALID) continue;
        const uint32_t prv = s_g[i - 1];
        if (prv == G_INVALID || ((prv ^ cur) & 0xFFFFu)) my = (unsigned long long)(q0 + i + 1);
    }
    if (my) atomicMax(&s_lc, my);
    __syncthreads();
    if (tid == 0) tile_lc[blockIdx.x] = s_lc;
}

// Packed words staged per tile: records start at t0 - 1 or later and end
// before t0 + TILE + 256 (len <= 255), and stream_get64 reads one word ahead.
#define TILE_WORDS(TILE) ((TILE) / 32 + 16)

// The g values of a tile's positions q0 .. q0 + TILE + w + 1 (q0 = t0 - 1),
// loaded into registers one tile ahead and stored to LDS when the tile starts.
template <int WG, int TILE>
struct GAhead {
    static constexpr int N = (TILE + WIN_MAX + 2 + WG - 1) / WG;
    uint16_t v[N];
    __device__ __forceinline__ void load(const PartitionArgs& a, const uint16_t* __restrict__ gbin, uint64_t t0) {
        const int64_t q0 = int64_t(t0) - 1, n = int64_t(a.n_bases);
        const int span = TILE + int(a.w) + 2;
#pragma unroll
        for (int c = 0; c < N; ++c) {
            const int i = int(threadIdx.x) + c * WG;
            const int64_t p = q0 + i;
            v[c] = (i < span && p >= 0 && p < n) ? gbin[p] : G16_INVALID;
        }
    }
    __device__ __forceinline__ void store(uint16_t* s_g) const {
#pragma unroll
        for (int c = 0; c < N; ++c) {
            const int i = int(threadIdx.x) + c * WG;
            if (i < TILE + WIN_MAX + 2) s_g[i] = v[c];
        }
    }
};

// Piece descriptor (one u64 per weak super k-mer): idx (bits 0-15, its first
// (k-1)-mer at tile position t0 - 1 + idx), g (16-31), the 4 hash bits above
// g (32-35), len (36-43), l_disc (44), r_disc (45), l_joined (46); the
// record starts at base t0 - 1 + idx - l_joined.
__host__ __device__ constexpr uint64_t piece_desc(uint32_t idx, uint32_t g, uint32_t mb, uint32_t len,
                                                  uint32_t l_disc, uint32_t r_disc, uint32_t lj) {
    return uint64_t(idx) | (uint64_t(g) << 16) | (uint64_t(mb) << 32) | (uint64_t(len) << 36) |
           (uint64_t(l_disc) << 44) | (uint64_t(r_disc) << 45) | (uint64_t(lj) << 46);
}

// Weak super k-mers with their first (k-1)-mer in [t0, t0 + TILE) (t0 a
// multiple of P_TILE), whose g values are in s_g (G16_INVALID: no (k-1)-mer)
// and the 4 hash bits above g in s_mbw (nibble o = position t0 + o): calls
// emit(o, descriptor) once per piece, o = 0 .. n_out - 1 in position order.
// Every thread of the WG-thread workgroup must call it (it synchronises).
// `carry` is the last g change (+1) before t0 (0: none); the same value for
// the next tile is returned.
template <int WG, int TILE, class Emit>
__device__ __forceinline__ unsigned long long tile_pieces(const PartitionArgs& a, uint64_t t0,
                                                          unsigned long long carry, const uint16_t* s_g,
                                                          const uint32_t* s_mbw, uint32_t& n_out, Emit&& emit) {
    constexpr int PPT = TILE / WG;  // positions per thread
    static_assert(PPT <= 32 && TILE % P_TILE == 0, "tile shape");
    const int tid = threadIdx.x;
    const int64_t q0 = int64_t(t0) - 1;
    const unsigned k = a.k, w = a.w;

    // My PPT contiguous positions' g values (and the one before) in
    // registers: gv[j + 1] is position base_idx + j, read with a 128-bit load.
    static_assert(PPT == 8, "register layout below");
    typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
    const int base_idx = 1 + tid * PPT;
    uint32_t gv[PPT + 1];
    {
        const u32x4 r0 = *reinterpret_cast<const u32x4*>(s_g + tid * PPT);
        gv[0] = r0.x & 0xFFFFu; gv[1] = r0.x >> 16; gv[2] = r0.y & 0xFFFFu; gv[3] = r0.y >> 16;
        gv[4] = r0.z & 0xFFFFu; gv[5] = r0.z >> 16; gv[6] = r0.w & 0xFFFFu; gv[7] = r0.w >> 16;
        gv[8] = s_g[tid * PPT + PPT];
    }
    // Change flags for my positions (an invalid previous position differs
    // from every g).
    uint32_t chg = 0, last_chg = 0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const uint32_t cur = gv[j + 1], prv = gv[j];
        if (cur == G16_INVALID) continue;
        if (prv != cur) {
            chg |= 1u << j;
            last_chg = uint32_t(base_idx + j);
        }
    }
    // Run boundaries (a change flag or an invalid position) and, per thread,
    // the first boundary after its positions: a piece starting at idx ends at
    // min(next boundary, idx + w).
    uint32_t bnd = chg;
#pragma unroll
    for (int j = 0; j < PPT; ++j) bnd |= uint32_t(gv[j + 1] == G16_INVALID) << j;
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    constexpr int NW = WG / 64;
    __shared__ uint32_t s_wmax[NW], s_wmin[NW], s_wsum[NW], s_tail;
    const int lane = tid & 63, wid = tid >> 6;

    // One barrier for three things: per wave the running maximum of the last
    // change (run starts) and the suffix minimum of the first boundary (piece
    // ends), and the first boundary past the tile found by wave 0.  The wave
    // scans are DPP lane moves (every lane of the workgroup is active here).
    const auto umax = [](uint32_t x, uint32_t y) { return x > y ? x : y; };
    const auto umin = [](uint32_t x, uint32_t y) { return x < y ? x : y; };
    const uint32_t vmax = dpp::incl_scan(last_chg, 0u, umax);
    const uint32_t vmin = dpp::suffix_scan(bnd ? uint32_t(base_idx + __ffs(int(bnd)) - 1) : NONE, NONE, umin);
    if (lane == 63) s_wmax[wid] = vmax;
    if (lane == 0) s_wmin[wid] = vmin;
    if (tid < 64) {  // first boundary in (TILE, TILE + w + 1]
        uint32_t tail = NONE;
        for (int i0 = TILE + 1; i0 <= TILE + int(w) + 1 && tail == NONE; i0 += 64) {
            const int i = i0 + tid;
            bool b = false;
            if (i <= TILE + int(w) + 1) {
                const uint32_t cur = s_g[i], prv = s_g[i - 1];
                b = cur == G16_INVALID || prv != cur;
            }
            const unsigned long long m = __ballot(b);
            if (m) tail = uint32_t(i0 + __ffsll((long long)m) - 1);
        }
        if (tid == 0) s_tail = tail;
    }
    __syncthreads();
    uint32_t before = 0, tile_last = 0, after = s_tail;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
        const uint32_t m = s_wmax[j], n = s_wmin[j];
        before = j < wid && m > before ? m : before;
        tile_last = m > tile_last ? m : tile_last;
        after = j > wid && n < after ? n : after;
    }
    const uint32_t incl = vmax > before ? vmax : before;
    // the last change before my positions (0: none), the first boundary after them
    const uint32_t excl = dpp::from_prev_lane(before, incl);
    const uint32_t next_bnd = umin(dpp::from_next_lane(NONE, vmin), after);
    const unsigned long long carry_out = tile_last ? (unsigned long long)(q0 + int64_t(tile_last) + 1) : carry;
    const int64_t rs0 = excl ? q0 + int64_t(excl) : (carry ? int64_t(carry) - 1 : -1);

    // A piece starts at every change and every w (k-1)-mers after one.  The
    // phase (p - run start) mod w is carried from position to position; a
    // valid position after an invalid one always starts a run (it has a change
    // flag), so only the thread's first position can need a division.
    uint32_t ph0 = 0;
    if (rs0 >= 0 && !(chg & 1u) && gv[1] != G16_INVALID)
        ph0 = uint32_t(uint64_t(q0 + base_idx - rs0) % w);
    // Piece starts among my positions (bit j: position base_idx + j), then
    // one descriptor per set bit: the loop runs as often as the wave's
    // busiest lane has pieces, not once per position.
    uint32_t starts = 0;
    {
        uint32_t ph = ph0;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            if (j) ph = ph + 1 == w ? 0u : ph + 1;
            if ((chg >> j) & 1u) ph = 0;
            starts |= uint32_t(ph == 0 && gv[j + 1] != G16_INVALID) << j;
        }
    }
    // exclusive sum of the piece counts (one barrier; the arrays are rewritten
    // only after the caller's next barrier)
    const uint32_t mine = uint32_t(__popc(starts));
    const uint32_t xs = dpp::incl_scan(mine, 0u, [](uint32_t x, uint32_t y) { return x + y; });
    if (lane == 63) s_wsum[wid] = xs;
    __syncthreads();
    uint32_t n_tile = 0, o_desc = xs - mine;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
        const uint32_t t = s_wsum[j];
        n_tile += t;
        o_desc += j < wid ? t : 0u;
    }
    for (uint32_t m = starts; m;) {
        const uint32_t j = uint32_t(__ffs(int(m))) - 1u;
        m &= m - 1u;
        const uint32_t idx = uint32_t(base_idx) + j;
        // Piece [idx, e): (k-1)-mers with this g, at most w of them.
        const uint32_t g = s_g[idx], pv = s_g[idx - 1];
        const uint32_t o = idx - 1u;
        const uint32_t mb = (s_mbw[o >> 3] >> (4 * (o & 7u))) & 15u;
        const uint32_t later = bnd >> (j + 1);  // my boundaries after position j
        const uint32_t nb = later ? idx + uint32_t(__ffs(int(later))) : next_bnd;
        const uint32_t e = min(nb, idx + w);
        const uint32_t nv = s_g[e];
        const uint32_t rj = nv != G16_INVALID, lj = pv != G16_INVALID;
        const uint32_t l_disc = lj && pv != g;
        const uint32_t r_disc = rj && nv != g;
        const uint32_t len = lj + (e - idx) + (k - 2) + rj;
        emit(o_desc++, piece_desc(idx, g, mb, len, l_disc, r_disc, lj));
    }
    n_out = n_tile;
    return carry_out;
}

// Last g change (+1) before minimizer tile t, from tile_lc: a wave looks back
// 64 tiles at a time; only runs longer than a tile need more than one step.
// Called by wave 0; the value is returned to every lane of that wave.
__device__ __forceinline__ unsigned long long carry_before(const uint64_t* __restrict__ tile_lc, uint64_t t) {
    int64_t j = int64_t(t) - 1 - int64_t(threadIdx.x & 63);
    while (true) {
        const unsigned long long v = j >= 0 ? tile_lc[j] : 0ull;
        const unsigned long long m = __ballot(v != 0 || j < 0);
        if (m) return __shfl(v, __ffsll((long long)m) - 1);
        j -= 64;
    }
}

// Span of workgroup blockIdx.x: minimizer tiles (P_TILE positions)
// [b * per, (b + 1) * per).
struct Span {
    uint64_t t_begin, t_end;
};
__device__ __forceinline__ Span span_of(uint64_t n_bases, uint64_t per) {
    const uint64_t n_t = (n_bases + P_TILE - 1) / P_TILE;
    const uint64_t b = min(n_t, uint64_t(blockIdx.x) * per);
    return {b, min(n_t, b + per)};
}

// Pass 2: the pieces of every tile: descriptors to desc[t * P_TILE + o]
// (o < tile_n[t] <= P_TILE), and records per coarse bin (g >> FINE_BITS) of
// each count span into hist[c * gridDim.x + span].  No global atomics.
__global__ __launch_bounds__(H_WG, 5) void k_pieces_hist(PartitionArgs a, const uint16_t* __restrict__ gbin,
                                                      const uint32_t* __restrict__ mbn,
                                                      const uint64_t* __restrict__ tile_lc, uint64_t per,
                                                      uint32_t* __restrict__ hist, uint64_t* __restrict__ desc,
                                                      uint32_t* __restrict__ tile_n) {
    static_assert(P_TILE / 8 == H_WG, "one nibble word per thread");
    __shared__ __attribute__((aligned(16))) uint16_t s_g[P_TILE + WIN_MAX + 2];
    __shared__ uint32_t s_mbw[P_TILE / 8];
    __shared__ unsigned long long s_carry;
    __shared__ uint32_t s_hist[N_COARSE];
    __shared__ uint64_t s_desc[P_TILE];  // the tile's descriptors, written out coalesced
    for (int i = threadIdx.x; i < N_COARSE; i += H_WG) s_hist[i] = 0;
    const Span sp = span_of(a.n_bases, per);
    const uint64_t n_mbw = (a.n_bases + 7) / 8;
    // g values and hash-bit words of the next tile are loaded while the
    // current one is processed
    GAhead<H_WG, P_TILE> ga;
    uint32_t mbw_next = 0;
    auto load_mbw = [&](uint64_t t) {
        const uint64_t mi = t * (P_TILE / 8) + threadIdx.x;
        mbw_next = mi < n_mbw ? mbn[mi] : 0u;
    };
    if (sp.t_begin < sp.t_end) {
        ga.load(a, gbin, sp.t_begin * P_TILE);
        load_mbw(sp.t_begin);
        if (threadIdx.x < 64) {
            const unsigned long long c = carry_before(tile_lc, sp.t_begin);
            if (threadIdx.x == 0) s_carry = c;
        }
    }
    __syncthreads();
    unsigned long long carry = s_carry;
    for (uint64_t t = sp.t_begin; t < sp.t_end; ++t) {
        __syncthreads();  // the previous tile is done with s_g, s_mbw and s_desc
        ga.store(s_g);
        s_mbw[threadIdx.x] = mbw_next;
        if (t + 1 < sp.t_end) {
            ga.load(a, gbin, (t + 1) * P_TILE);
            load_mbw(t + 1);
        }
        __syncthreads();
        uint32_t n_t;
        carry = tile_pieces<H_WG, P_TILE>(a, t * P_TILE, carry, s_g, s_mbw, n_t,
                                          [&](uint32_t o, uint64_t d) { s_desc[o] = d; });
        __syncthreads();
        uint64_t* out = desc + t * P_TILE;
        for (uint32_t i = threadIdx.x; i < n_t; i += H_WG) {
            const uint64_t d = s_desc[i];
            out[i] = d;
            atomicAdd(&s_hist[(uint32_t(d >> 16) & 0xFFFFu) >> FINE_BITS], 1u);
        }
        if (threadIdx.x == 0) tile_n[t] = n_t;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N_COARSE; i += H_WG) hist[uint64_t(i) * gridDim.x + blockIdx.x] = s_hist[i];
}

// Per coarse bin c (one workgroup each): exclusive scan of hist[c][*] over
// the big tiles, in place, and the bin's total into ctot[c].
__global__ __launch_bounds__(1024) void k_hist_scan(uint32_t* __restrict__ hist, uint64_t n_big,
                                                    uint64_t* __restrict__ ctot) {
    __shared__ uint64_t s[1024];
    const int tid = threadIdx.x;
    uint32_t* row = hist + uint64_t(blockIdx.x) * n_big;
    const uint64_t per = (n_big + 1023) / 1024;
    const uint64_t b0 = min(n_big, uint64_t(tid) * per), b1 = min(n_big, b0 + per);
    uint64_t sum = 0;
    for (uint64_t i = b0; i < b1; ++i) sum += row[i];
    s[tid] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const uint64_t v = tid >= d ? s[tid - d] : 0;
        __syncthreads();
        s[tid] += v;
        __syncthreads();
    }
    uint64_t run = s[tid] - sum;
    for (uint64_t i = b0; i < b1; ++i) {
        const uint32_t c = row[i];
        row[i] = uint32_t(run);  // a coarse bin holds < 2^32 records
        run += c;
    }
    if (tid == 1023) ctot[blockIdx.x] = s[1023];
}

// Exclusive scan of the N_COARSE bin totals -> coff[0..N_COARSE].
__global__ __launch_bounds__(N_COARSE) void k_coarse_scan(const uint64_t* __restrict__ ctot,
                                                          uint64_t* __restrict__ coff) {
    __shared__ uint64_t s[N_COARSE];
    const int tid = threadIdx.x;
    s[tid] = ctot[tid];
    __syncthreads();
    for (int d = 1; d < N_COARSE; d <<= 1) {
        const uint64_t v = tid >= d ? s[tid - d] : 0;
        __syncthreads();
        s[tid] += v;
        __syncthreads();
    }
    coff[tid + 1] = s[tid];
    if (tid == 0) coff[0] = 0;
}

// Pass 3: the pieces the count pass described; emit span s (H_SPLIT count
// spans) appends its records of coarse bin c to the run starting at coff[c] +
// hist[c][s * H_SPLIT].  A span is a contiguous part of the input, so each
// run is written front to back by one workgroup.  The span is taken EB_TILES
// tiles at a time: their descriptors (at most EB_CAP per round) are staged in
// LDS and counting-sorted by coarse bin, so consecutive lanes write
// consecutive slots of a bin and HBM lines leave whole instead of being
// assembled from single records.  Level-1 records: the header att | mbin <<
// L1_MB_SHIFT (u32) in hdr, the label in lab as whole 16-byte pairs
// (l1_label_words(W) words).
constexpr int EB_TILES = 32;     // tiles per batch (their packed words staged)
constexpr int EB_CAP = 10240;    // descriptors per sorting round
constexpr int EB_WORDS = EB_TILES * (P_TILE / 32) + 16;
__global__ __launch_bounds__(E_WG) void k_pieces_emit(PartitionArgs a, const uint64_t* __restrict__ desc,
                                                      const uint32_t* __restrict__ tile_n, uint64_t per,
                                                      uint32_t n_hist, const uint32_t* __restrict__ hist,
                                                      const uint64_t* __restrict__ coff, uint32_t W,
                                                      uint32_t* __restrict__ hdr, uint64_t* __restrict__ lab) {
    static_assert(N_COARSE == 256 && E_WG >= N_COARSE && EB_TILES <= 64, "shapes below");
    __shared__ uint64_t s_words[EB_WORDS];
    __shared__ uint64_t s_d[EB_CAP];
    __shared__ uint16_t s_perm[EB_CAP];
    __shared__ unsigned long long s_cur[N_COARSE];
    __shared__ uint32_t s_bcnt[N_COARSE], s_bbase[N_COARSE], s_bpos[N_COARSE];
    __shared__ uint32_t s_toff[EB_TILES + 1];
    __shared__ uint32_t s_wave[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    if (tid < N_COARSE) s_cur[tid] = coff[tid] + hist[uint64_t(tid) * n_hist + uint64_t(blockIdx.x) * H_SPLIT];
    const Span sp = span_of(a.n_bases, per);
    for (uint64_t ta = sp.t_begin; ta < sp.t_end; ta += EB_TILES) {
        const uint32_t nt = uint32_t(min(uint64_t(EB_TILES), sp.t_end - ta));
        const uint64_t t0 = ta * P_TILE;
        const uint64_t wlo = t0 ? t0 / 32 - 1 : 0;  // s_words[0] holds bases wlo*32 ..
        const uint64_t base0 = wlo * 32;
        __syncthreads();  // the previous batch is done with s_words and s_toff
        // the batch's tile piece counts, loaded with its packed words
        const uint32_t tn = tid < 64 && lane < nt ? tile_n[ta + lane] : 0u;
        {  // every packed-word load of the batch in flight at once
            constexpr int WU = (EB_WORDS + E_WG - 1) / E_WG;
            const uint32_t nw = nt * (P_TILE / 32) + 16;
            uint64_t wv[WU];
#pragma unroll
            for (int u = 0; u < WU; ++u) {
                const uint32_t i = tid + uint32_t(u) * E_WG;
                wv[u] = i < nw && wlo + i < a.n_words ? a.packed[wlo + i] : 0ull;
            }
#pragma unroll
            for (int u = 0; u < WU; ++u) {
                const uint32_t i = tid + uint32_t(u) * E_WG;
                if (i < nw) s_words[i] = wv[u];
            }
        }
        if (tid < 64) {  // descriptor offsets of the batch's tiles (the whole first wave)
            const uint32_t x = dpp::incl_scan(tn, 0u, [](uint32_t p, uint32_t q) { return p + q; });
            if (lane < nt) s_toff[lane + 1] = x;
            if (lane == 0) s_toff[0] = 0;
        }
        __syncthreads();
        const uint32_t n_b = s_toff[nt];
        for (uint32_t x0 = 0; x0 < n_b; x0 += EB_CAP) {
            const uint32_t m = min(uint32_t(EB_CAP), n_b - x0);
            if (tid < N_COARSE) s_bcnt[tid] = 0;
            __syncthreads();
            // stage, tagged with the tile (bits 48+): SU descriptors per thread
            // located (independent searches) and loaded together, indices
            // clamped to the last one so that no load is conditional
            constexpr int SU = 4;
            for (uint32_t j0 = tid; j0 < m; j0 += SU * E_WG) {
                uint64_t dv[SU];
                uint32_t tv[SU];
#pragma unroll
                for (int u = 0; u < SU; ++u) {
                    const uint32_t jj = x0 + min(j0 + uint32_t(u) * E_WG, m - 1);
                    uint32_t lo = 0, hi = nt;  // tile = last tt with s_toff[tt] <= jj
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (s_toff[mid] <= jj) lo = mid;
                        else hi = mid;
                    }
                    tv[u] = lo;
                    dv[u] = desc[(ta + lo) * P_TILE + (jj - s_toff[lo])];
                }
#pragma unroll
                for (int u = 0; u < SU; ++u) {
                    const uint32_t j = j0 + uint32_t(u) * E_WG;
                    if (j < m) {
                        s_d[j] = dv[u] | (uint64_t(tv[u]) << 48);
                        atomicAdd(&s_bcnt[(uint32_t(dv[u] >> 16) & 0xFFFFu) >> FINE_BITS], 1u);
                    }
                }
            }
            __syncthreads();
            uint32_t v = 0, x = 0;  // exclusive scan of the bin counts by the whole first 4 waves
            if (tid < N_COARSE) {
                v = s_bcnt[tid];
                x = dpp::incl_scan(v, 0u, [](uint32_t p, uint32_t q) { return p + q; });
                if (lane == 63) s_wave[tid >> 6] = x;
            }
            __syncthreads();
            if (tid < N_COARSE) {
                uint32_t pre = 0;
                for (uint32_t j = 0; j < (tid >> 6); ++j) pre += s_wave[j];
                s_bbase[tid] = s_bpos[tid] = pre + x - v;
            }
            __syncthreads();
            for (uint32_t j = tid; j < m; j += E_WG)
                s_perm[atomicAdd(&s_bpos[(uint32_t(s_d[j] >> 16) & 0xFFFFu) >> FINE_BITS], 1u)] = uint16_t(j);
            __syncthreads();
            for (uint32_t q = tid; q < m; q += E_WG) {  // consecutive lanes, consecutive slots of a bin
                const uint64_t d = s_d[s_perm[q]];
                const uint32_t idx = uint32_t(d & 0xFFFFu), g = uint32_t(d >> 16) & 0xFFFFu;
                const uint32_t mb = uint32_t(d >> 32) & 15u, len = uint32_t(d >> 36) & 0xFFu;
                const uint32_t l_disc = uint32_t(d >> 44) & 1u, r_disc = uint32_t(d >> 45) & 1u;
                const uint32_t lj = uint32_t(d >> 46) & 1u, tile = uint32_t(d >> 48);
                const uint32_t c = g >> FINE_BITS;
                const uint64_t slot = s_cur[c] + (q - s_bbase[c]);
                // first base: tile position idx - 1 (idx >= 1; lj = 0 at base 0)
                const uint64_t start = t0 + uint64_t(tile) * P_TILE - 1 + idx - lj;
                hdr[slot] = len | (l_disc << 8) | (r_disc << 9) | (mb << L1_MB_SHIFT) | (g << 16);
                // label words two at a time (16-byte stores)
                u64x2* out = reinterpret_cast<u64x2*>(lab + slot * l1_label_words(W));
                uint64_t prev = 0;
                for (uint32_t qq = 0; qq < W; ++qq) {
                    const uint32_t wj = W - 1 - qq;  // label word qq holds bases 32wj ..
                    uint64_t xw = 0;
                    if (32 * wj < len) {
                        xw = stream_get64(s_words, start - base0 + 32ull * wj);
                        const uint32_t nb = len - 32 * wj;
                        if (nb < 32) xw &= ~0ull << (64 - 2 * nb);
                    }
                    if (!(qq & 1)) prev = xw;
                    else out[qq / 2] = u64x2{prev, xw};
                }
                if (W & 1) out[W / 2] = u64x2{prev, 0ull};
            }
            __syncthreads();
            if (tid < N_COARSE) s_cur[tid] += s_bcnt[tid];
        }
    }
}

// Level 2: one workgroup per coarse bin moves its records into the final
// per-g order (the bin's 2^FINE_BITS buckets) and writes their offsets.
// Records are staged a chunk at a time in LDS and counting-sorted by bucket,
// so each bucket's records leave as one contiguous run per chunk (coalesced
// stores instead of 64 interleaved streams).  Order inside a bucket is free.
constexpr uint32_t FINE_WG = 1024;
constexpr uint32_t FINE_LDS = 144 * 1024;  // staging bytes
constexpr uint32_t CH_MAX = 4096;
__host__ __device__ constexpr uint32_t fine_chunk(uint32_t W) {
    // largest power of two <= FINE_LDS / (record bytes + 2 permutation bytes), at most 4096
    return (FINE_LDS / (13 + 8 * W)) >= 4096 ? 4096 : (FINE_LDS / (13 + 8 * W)) >= 2048 ? 2048 : 1024;
}

__global__ __launch_bounds__(FINE_WG) void k_bucket_fine(const uint32_t* __restrict__ hdr,
                                                         const uint64_t* __restrict__ lab,
                                                         const uint64_t* __restrict__ coff, uint32_t W,
                                                         uint64_t* __restrict__ boff, uint32_t* __restrict__ att,
                                                         uint64_t* __restrict__ labels, uint8_t* __restrict__ mbin) {
    constexpr uint32_t NF = 1u << FINE_BITS;
    static_assert(NF == 64, "one wave scans the bucket counts");
    __shared__ uint32_t s_cnt[NF], s_ccnt[NF], s_cbase[NF];
    __shared__ unsigned long long s_cur[NF];
    __shared__ __attribute__((aligned(16))) uint8_t s_stage[FINE_LDS];
    const uint32_t c = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    const uint64_t r0 = coff[c], r1 = coff[c + 1];
    const uint64_t LW = l1_label_words(W);
    const uint32_t CH = fine_chunk(W);
    uint64_t* s_lab = reinterpret_cast<uint64_t*>(s_stage);                // CH * W
    uint32_t* s_att = reinterpret_cast<uint32_t*>(s_lab + size_t(CH) * W);  // CH
    uint16_t* s_perm = reinterpret_cast<uint16_t*>(s_att + CH);             // CH: sorted slot -> staged record
    uint8_t* s_mb = reinterpret_cast<uint8_t*>(s_perm + CH);                // CH

    for (uint32_t i = tid; i < NF; i += FINE_WG) s_cnt[i] = 0;
    __syncthreads();
    // bucket counts of the bin: 8 independent header loads per thread in flight
    for (uint64_t rb = r0 + tid; rb < r1; rb += 8 * FINE_WG) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = hdr[min(rb + uint64_t(j) * FINE_WG, r1 - 1)];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (rb + uint64_t(j) * FINE_WG < r1) atomicAdd(&s_cnt[(v[j] >> 16) & (NF - 1)], 1u);
    }
    __syncthreads();
    // exclusive scan of the bucket counts by the whole first wave (lane f: bucket f)
    const auto wave_excl = [](uint32_t v) {
        return dpp::incl_scan(v, 0u, [](uint32_t x, uint32_t y) { return x + y; }) - v;
    };
    if (tid < NF) {
        const uint64_t run = r0 + wave_excl(s_cnt[tid]);
        boff[c * NF + tid] = run;
        s_cur[tid] = run;
        if (tid == 0 && c + 1 == gridDim.x) boff[(c + 1) * NF] = r1;
    }
    // A chunk's headers and first two label words are loaded into registers
    // while the previous chunk is written out (one workgroup per CU has no
    // other loads in flight then); later label words are read when staged.
    static_assert(CH_MAX == 4 * FINE_WG, "4 records per thread");
    uint32_t h_next[4];
    u64x2 l_next[4];
    const auto prefetch = [&](uint64_t q0) {
        const uint32_t n = uint32_t(min(uint64_t(CH), r1 - q0));
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t i = tid + j * FINE_WG;
            if (i < n) {
                h_next[j] = hdr[q0 + i];
                l_next[j] = *reinterpret_cast<const u64x2*>(lab + (q0 + i) * LW);
            }
        }
    };
    if (r0 < r1) prefetch(r0);
    for (uint64_t q0 = r0; q0 < r1; q0 += CH) {
        const uint32_t n = uint32_t(min(uint64_t(CH), r1 - q0));
        if (tid < NF) s_ccnt[tid] = 0;
        __syncthreads();
        // stage the chunk and rank each record in its bucket
        uint32_t rank[4], fb[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t i = tid + j * FINE_WG;
            rank[j] = fb[j] = 0;
            if (i >= n) continue;
            const uint32_t h0 = h_next[j];
            const uint32_t at = h0 & ~(15u << L1_MB_SHIFT);
            s_att[i] = at;
            s_mb[i] = uint8_t((h0 >> L1_MB_SHIFT) & 15u);
            s_lab[size_t(i) * W] = l_next[j].x;
            if (W > 1) s_lab[size_t(i) * W + 1] = l_next[j].y;
            const u64x2* in = reinterpret_cast<const u64x2*>(lab + (q0 + i) * LW);
            for (uint32_t q = 2; q < W; q += 2) {
                const u64x2 pq = in[q / 2];
                s_lab[size_t(i) * W + q] = pq.x;
                if (q + 1 < W) s_lab[size_t(i) * W + q + 1] = pq.y;
            }
            fb[j] = (at >> 16) & (NF - 1);
            rank[j] = atomicAdd(&s_ccnt[fb[j]], 1u);
        }
        __syncthreads();
        if (tid < NF) s_cbase[tid] = wave_excl(s_ccnt[tid]);
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t i = tid + j * FINE_WG;
            if (i < n) s_perm[s_cbase[fb[j]] + rank[j]] = uint16_t(i);
        }
        __syncthreads();
        if (q0 + CH < r1) prefetch(q0 + CH);
        // write the sorted chunk: consecutive threads, consecutive slots of a bucket
        for (uint32_t s2 = tid; s2 < n; s2 += FINE_WG) {
            const uint32_t i = s_perm[s2];
            const uint32_t at = s_att[i];
            const uint32_t f = (at >> 16) & (NF - 1);
            const uint64_t slot = s_cur[f] + (s2 - s_cbase[f]);
            att[slot] = at;
            mbin[slot] = s_mb[i];
            for (uint32_t q = 0; q < W; ++q) labels[slot * W + q] = s_lab[size_t(i) * W + q];
        }
        __syncthreads();
        if (tid < NF) s_cur[tid] += s_ccnt[tid];
    }
}

__global__ __launch_bounds__(256) void k_bucket_features(const uint64_t* __restrict__ off,
                                                         const uint32_t* __restrict__ att,
                                                         const uint8_t* __restrict__ mbin,
                                                         uint32_t k, uint64_t* __restrict__ stats,
                                                         float* __restrict__ feat) {
    __shared__ uint32_t s_hist[16];
    __shared__ unsigned long long s_km, s_bases;
    const uint32_t b = blockIdx.x;
    const int tid = threadIdx.x;
    if (tid < 16) s_hist[tid] = 0;
    if (tid == 0) s_km = 0, s_bases = 0;
    __syncthreads();
    const uint64_t r0 = off[b], r1 = off[b + 1];
    unsigned long long km = 0, bases = 0;
    for (uint64_t r = r0 + tid; r < r1; r += 256) {
        const uint32_t len = att[r] & 0xFFu;
        km += len - (k - 1);
        bases += len;
        atomicAdd(&s_hist[mbin[r] & 15u], 1u);
    }
    atomicAdd(&s_km, km);
    atomicAdd(&s_bases, bases);
    __syncthreads();
    if (tid == 0) {
        const uint64_t n = r1 - r0;
        stats[4 * b + 0] = n;
        stats[4 * b + 1] = s_km;
        stats[4 * b + 2] = s_bases;
        stats[4 * b + 3] = 0;
        float* f = feat + size_t(b) * N_FEAT;
        for (int i = 0; i < 16; ++i) f[i] = log1pf(float(s_hist[i]));
        f[16] = log1pf(float(n));
        f[17] = log1pf(float(s_km));
        f[18] = n ? float(s_bases) / float(n) / float(2 * k) : 0.f;
    }
}

// Per-bucket sums of one partition chunk added into sums (SUM_WORDS u32 words
// per bucket: the 16-bin histogram of the 4 hash bits above g, then u64 record,
// k-mer instance and base counts), for buckets that leave device memory
// chunk by chunk; k_features_from_sums then gives what k_bucket_features gives
// for the merged buckets.
__global__ __launch_bounds__(256) void k_bucket_sums(const uint64_t* __restrict__ off,
                                                     const uint32_t* __restrict__ att,
                                                     const uint8_t* __restrict__ mbin, uint32_t k,
                                                     uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_hist[16];
    __shared__ unsigned long long s_km, s_bases;
    const uint32_t b = blockIdx.x;
    const int tid = threadIdx.x;
    if (tid < 16) s_hist[tid] = 0;
    if (tid == 0) s_km = 0, s_bases = 0;
    __syncthreads();
    const uint64_t r0 = off[b], r1 = off[b + 1];
    unsigned long long km = 0, bases = 0;
    for (uint64_t r = r0 + tid; r < r1; r += 256) {
        const uint32_t len = att[r] & 0xFFu;
        km += len - (k - 1);
        bases += len;
        atomicAdd(&s_hist[mbin[r] & 15u], 1u);
    }
    atomicAdd(&s_km, km);
    atomicAdd(&s_bases, bases);
    __syncthreads();
    uint32_t* sb = sums + size_t(b) * SUM_WORDS;
    if (tid < 16 && s_hist[tid]) atomicAdd(&sb[tid], s_hist[tid]);
    if (tid == 0 && r1 > r0) {
        unsigned long long* s64 = reinterpret_cast<unsigned long long*>(sb + 16);
        atomicAdd(&s64[0], (unsigned long long)(r1 - r0));
        atomicAdd(&s64[1], s_km);
        atomicAdd(&s64[2], s_bases);
    }
}

__global__ __launch_bounds__(256) void k_features_from_sums(const uint32_t* __restrict__ sums, uint32_t G,
                                                            uint32_t k, uint64_t* __restrict__ stats,
                                                            float* __restrict__ feat) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b >= G) return;
    const uint32_t* sb = sums + size_t(b) * SUM_WORDS;
    const uint64_t* s64 = reinterpret_cast<const uint64_t*>(sb + 16);
    const uint64_t n = s64[0], km = s64[1], bases = s64[2];
    stats[4 * b + 0] = n;
    stats[4 * b + 1] = km;
    stats[4 * b + 2] = bases;
    stats[4 * b + 3] = 0;
    float* f = feat + size_t(b) * N_FEAT;  // the expressions of k_bucket_features
    for (int i = 0; i < 16; ++i) f[i] = log1pf(float(sb[i]));
    f[16] = log1pf(float(n));
    f[17] = log1pf(float(km));
    f[18] = n ? float(bases) / float(n) / float(2 * k) : 0.f;
}

}  // namespace

void launch_bucket_sums(const uint64_t* off, const uint32_t* att, const uint8_t* mbin, uint32_t G, uint32_t k,
                        uint32_t* sums, hipStream_t s) {
    hipLaunchKernelGGL(k_bucket_sums, dim3(G), dim3(256), 0, s, off, att, mbin, k, sums);
}

void launch_features_from_sums(const uint32_t* sums, uint32_t G, uint32_t k, uint64_t* stats, float* feat,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_features_from_sums, dim3((G + 255) / 256), dim3(256), 0, s, sums, G, k, stats, feat);
}

void launch_minimizer_g(const PartitionArgs& a, uint16_t* gbin, uint32_t* mbn, uint64_t* tile_lc, hipStream_t s) {
    const uint64_t tiles = (a.n_bases + P_TILE - 1) / P_TILE;
    if (!tiles) return;
    const uint32_t fg = uint32_t(std::min<uint64_t>((a.n_frag + 255) / 256, 65536));
    hipLaunchKernelGGL(k_tile_frag, dim3(fg), dim3(256), 0, s, a.frag_off, a.n_frag, tiles, a.tile_frag);
    hipLaunchKernelGGL(k_minimizer_g, dim3(uint32_t(tiles)), dim3(P_WG), 0, s, a, gbin, mbn, tile_lc);
}

PieceSpans pieces_spans(uint64_t n_bases, uint32_t n_cu) {
    PieceSpans ps{0, 0, 0, 0};
    const uint64_t n_t = (n_bases + P_TILE - 1) / P_TILE;
    if (!n_t) return ps;
    // emit spans: E_SPANS_PER_CU per CU, each a multiple of the emit tile
    // (E_TILE / P_TILE minimizer tiles) and of H_SPLIT count spans
    constexpr uint64_t Q = uint64_t(E_TILE / P_TILE) * H_SPLIT;
    const uint64_t n_e = std::max<uint64_t>(1, std::min<uint64_t>(uint64_t(n_cu) * E_SPANS_PER_CU, (n_t + Q - 1) / Q));
    ps.per_emit = (n_t + n_e - 1) / n_e;
    ps.per_emit = (ps.per_emit + Q - 1) / Q * Q;
    ps.per_count = ps.per_emit / H_SPLIT;
    ps.n_emit = uint32_t((n_t + ps.per_emit - 1) / ps.per_emit);
    ps.n_count = uint32_t((n_t + ps.per_count - 1) / ps.per_count);
    return ps;
}

void launch_pieces_hist(const PartitionArgs& a, const uint16_t* gbin, const uint32_t* mbn, const uint64_t* tile_lc,
                        const PieceSpans& ps, uint32_t* hist, uint64_t* ctot, uint64_t* coff, uint64_t* desc,
                        uint32_t* tile_n, hipStream_t s) {
    if (!ps.n_count) return;
    hipLaunchKernelGGL(k_pieces_hist, dim3(ps.n_count), dim3(H_WG), 0, s, a, gbin, mbn, tile_lc, ps.per_count, hist,
                       desc, tile_n);
    hipLaunchKernelGGL(k_hist_scan, dim3(N_COARSE), dim3(1024), 0, s, hist, uint64_t(ps.n_count), ctot);
    hipLaunchKernelGGL(k_coarse_scan, dim3(1), dim3(N_COARSE), 0, s, ctot, coff);
}

void launch_pieces_emit(const PartitionArgs& a, const uint64_t* desc, const uint32_t* tile_n, const PieceSpans& ps,
                        const uint32_t* hist, const uint64_t* coff, uint32_t W, uint32_t* hdr1, uint64_t* lab1,
                        hipStream_t s) {
    if (!ps.n_emit) return;
    hipLaunchKernelGGL(k_pieces_emit, dim3(ps.n_emit), dim3(E_WG), 0, s, a, desc, tile_n, ps.per_emit,
                       ps.n_count, hist, coff, W, hdr1, lab1);
}

void launch_bucket_fine(const uint32_t* hdr1, const uint64_t* lab1, const uint64_t* coff, uint32_t W, uint64_t* boff,
                        uint32_t* att,
                        uint64_t* labels, uint8_t* mbin, hipStream_t s) {
    hipLaunchKernelGGL(k_bucket_fine, dim3(N_COARSE), dim3(FINE_WG), 0, s, hdr1, lab1, coff, W, boff, att, labels, mbin);
}

void launch_bucket_features(const uint64_t* off, const uint32_t* att, const uint8_t* mbin,
                            uint32_t G, uint32_t k, uint64_t* stats, float* feat, hipStream_t s) {
    hipLaunchKernelGGL(k_bucket_features, dim3(G), dim3(256), 0, s, off, att, mbin, k, stats, feat);
}

namespace {
// One workgroup per bucket: each chunk's run of the bucket's records (headers,
// minimizer bins, W label words each) copied to its place, coalesced.
__global__ __launch_bounds__(256) void k_merge_chunks(const ChunkSrc* __restrict__ src, uint32_t n_chunks,
                                                      uint32_t G, uint32_t W, const uint64_t* __restrict__ dst_off,
                                                      uint32_t* __restrict__ att, uint64_t* __restrict__ labels,
                                                      uint8_t* __restrict__ mbin) {
    const uint32_t b = blockIdx.x;
    for (uint32_t c = 0; c < n_chunks; ++c) {
        const ChunkSrc sc = src[c];
        const uint64_t s0 = sc.boff[b], n = sc.boff[b + 1] - s0, d0 = dst_off[uint64_t(c) * G + b];
        for (uint64_t r = threadIdx.x; r < n; r += 256) {
            att[d0 + r] = sc.att[s0 + r];
            mbin[d0 + r] = sc.mbin[s0 + r];
        }
        for (uint64_t i = threadIdx.x; i < n * W; i += 256) labels[d0 * W + i] = sc.labels[s0 * W + i];
    }
}
}  // namespace

void launch_merge_chunks(const ChunkSrc* src, uint32_t n_chunks, uint32_t G, uint32_t W, const uint64_t* dst_off,
                         uint32_t* att, uint64_t* labels, uint8_t* mbin, hipStream_t s) {
    if (!n_chunks) return;
    hipLaunchKernelGGL(k_merge_chunks, dim3(G), dim3(256), 0, s, src, n_chunks, G, W, dst_off, att, labels, mbin);
}

}  // namespace cfgpu
