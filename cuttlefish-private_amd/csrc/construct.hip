// Subgraph construct kernels: per weak super k-mer, every k-mer instance is
// inserted into its subgraph's open-addressing vertex table in HBM and its
// State_Config is updated (src/Subgraph.cpp:44-108, include/Subgraph.hpp:546-560,
// include/State_Config.hpp:207-268).
//
// Table entry (stride 8*KW + 8 bytes): KW key words (Kmer<k> layout), a u32
// Edge_Frequency word (8 saturating 4-bit counters) and a u32 status word
// (Vertex_Status bits: 2 = front-discontinuous, 4 = back-discontinuous).
//
// Every 4-bit counter slot (vertex, side, base) is fed by exactly one canonical
// (k+1)-mer, so the counters are the per-subgraph (k+1)-mer counts; counting
// saturates at 15 exactly like Edge_Frequency::add_edge, and saturating adds
// commute, so the atomic update order cannot change any result.
//
// Key claim protocol.  The top key word of a canonical k-mer never uses bit 63
// (k odd => at most 31 bases in the top word), so ~0 marks an empty slot and
// `top | 1<<63` marks a slot whose lower key words are still being written.
// KW == 1: one 64-bit CAS.  KW > 1: CAS the top word EMPTY -> top|BUSY, store
// the lower words, release fence, store top.  A prober that meets BUSY with a
// matching top word re-reads the same slot on its next loop iteration; the
// claimer finishes its stores within its own iteration, so lanes of one
// wavefront never wait on each other inside an iteration.
#include "cf_instance.hpp"
#include "cf_kernels.hpp"

#include <algorithm>

namespace cfgpu {
namespace {

constexpr uint64_t EMPTY = ~0ull;
constexpr uint64_t BUSY = 1ull << 63;
constexpr uint32_t MAX_SPINS = 1u << 22;
constexpr uint32_t MAX_PROBES = 1024;  // longer probe chains count as overflow

// Table memory is addressed through global-address-space pointers so probes
// and atomics compile to global_* (not flat_*) instructions.
typedef __attribute__((address_space(1))) uint8_t g_u8;
typedef __attribute__((address_space(1))) uint64_t g_u64;
typedef __attribute__((address_space(1))) uint32_t g_u32;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
typedef __attribute__((address_space(1))) u32x4 g_u32x4;

template <int KW>
__device__ __forceinline__ g_u64* slot_ptr(g_u8* tbl, uint64_t s) {
    return reinterpret_cast<g_u64*>(tbl + s * uint64_t(8 * KW + 8));
}

__device__ __forceinline__ uint32_t ld_relaxed32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_coherent(g_u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas64(g_u64* p, uint64_t& expected, uint64_t desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Home slot of a hash in a table of `cap` slots (multiply-shift range
// reduction, so capacities need not be powers of two).
__device__ __forceinline__ uint64_t home_slot(uint64_t h, uint64_t cap) { return __umul64hi(h, cap); }

struct Probe {
    int64_t slot;      // -1: no room within MAX_PROBES
    uint32_t ef_seen;  // Edge_Frequency word as loaded (may lag the true value)
    bool inserted;
};

// Finds `key`'s slot, inserting the key if absent.  Plain loads may return a
// stale line: a slot's key only ever changes from EMPTY to a key, and counters
// only grow, so a stale read can only show EMPTY (the CAS then returns the
// truth) or a smaller counter word (the counter CAS then fails and retries).
// A long chain polls the bucket's overflow flag: once the table is known to
// be too small its contents are discarded, so the search gives up.
constexpr uint32_t POLL_PROBES = 64;
template <int KW>
__device__ Probe find_or_insert(g_u8* tbl, uint64_t cap, const Kmer<KW>& key, const uint32_t* ovf) {
    uint64_t s = home_slot(km_hash(key), cap);
    const uint64_t lim = cap < MAX_PROBES ? cap : MAX_PROBES;
    if constexpr (KW == 1) {
        for (uint64_t n = 0; n < lim; ++n) {
            if (n % POLL_PROBES == POLL_PROBES - 1 && ld_relaxed32(ovf)) return {-1, 0u, false};
            g_u64* kp = slot_ptr<1>(tbl, s);
            const u32x4 e = *reinterpret_cast<g_u32x4*>(kp);  // key, e_f, status
            const uint64_t cur = uint64_t(e.x) | (uint64_t(e.y) << 32);
            if (cur == key.w[0]) return {int64_t(s), e.z, false};
            if (cur == EMPTY) {
                uint64_t exp = EMPTY;
                if (cas64(kp, exp, key.w[0])) return {int64_t(s), 0u, true};
                if (exp == key.w[0]) return {int64_t(s), 0u, false};
            }
            s = s + 1 == cap ? 0 : s + 1;
        }
        return {-1, 0u, false};
    } else {
        const uint64_t top = key.w[KW - 1];
        uint32_t spins = 0;
        uint64_t n = 0;
        bool fresh = false;  // re-read a claimed slot coherently
        while (n < lim) {
            g_u64* kp = slot_ptr<KW>(tbl, s);
            uint64_t cur = fresh ? ld_coherent(kp + KW - 1) : kp[KW - 1];
            fresh = false;
            if (cur == EMPTY) {
                uint64_t exp = EMPTY;
                if (cas64(kp + KW - 1, exp, top | BUSY)) {
#pragma unroll
                    for (int i = 0; i < KW - 1; ++i)
                        __hip_atomic_store(kp + i, key.w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __threadfence();
                    __hip_atomic_store(kp + KW - 1, top, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    return {int64_t(s), 0u, true};
                }
                cur = exp;
            }
            if ((cur & ~BUSY) == top) {
                if (cur & BUSY) {
                    if (++spins > MAX_SPINS) return {-1, 0u, false};
                    fresh = true;
                    continue;  // the claimer publishes within its own iteration
                }
                bool eq = true;
#pragma unroll
                for (int i = 0; i < KW - 1; ++i) eq &= ld_coherent(kp + i) == key.w[i];
                if (eq) return {int64_t(s), uint32_t(kp[KW]), false};
            }
            ++n;
            if (n % POLL_PROBES == 0 && ld_relaxed32(ovf)) return {-1, 0u, false};
            s = s + 1 == cap ? 0 : s + 1;
        }
        return {-1, 0u, false};
    }
}

template <int KW, int WG>
__device__ void construct_item(const ConstructArgs& a, uint32_t item) {
    const int tid = threadIdx.x;
    const uint32_t b = a.item_bucket[item];
    // Buckets already known to overflow are skipped (they are rebuilt with a
    // larger table); each thread polls the flag between its records.
    if (ld_relaxed32(&a.overflow[b])) return;
    const uint32_t first = a.item_first[item], cnt = a.item_count[item];
    const uint32_t W = a.W, k = a.k;
    const uint64_t rbase = a.bucket_off[b] + first;
    g_u8* tbl = reinterpret_cast<g_u8*>(a.tbl_ptr[b]);
    const uint64_t cap = a.tbl_cap[b];
    const uint32_t limit = a.tbl_limit[b];
    const int lane = tid & 63;

    // One thread walks one record (RecordWalker rolls the k-mer per base).
    for (uint32_t r = tid; r < cnt; r += WG) {
        if (r >= uint32_t(WG) && ld_relaxed32(&a.overflow[b])) break;
        RecordWalker<KW, 2 * KW> rw;
        rw.load(a.att, a.labels, rbase + r, W, k);
        const uint32_t m = a.mult ? a.mult[rbase + r] : 1u;
        for (uint32_t t = 0; t < rw.n; ++t) {
            // a table flagged too small is discarded: stop adding to it
            if (t % 16 == 15 && ld_relaxed32(&a.overflow[b])) break;
            const Inst<KW> in = rw.next();
            const Probe pr = find_or_insert<KW>(tbl, cap, in.key, &a.overflow[b]);
            if (pr.slot < 0) {
                atomicOr(&a.overflow[b], 1u);
            } else {
                g_u32* ef = reinterpret_cast<g_u32*>(slot_ptr<KW>(tbl, uint64_t(pr.slot)) + KW);
                // Saturating update; counters never decrease, so a (possibly
                // stale) word that is already saturated needs no atomic.
                uint32_t old = pr.ef_seen;
                while (true) {
                    const uint32_t nw = ef_add(old, in.front, in.back, m);
                    if (nw == old) break;
                    if (__hip_atomic_compare_exchange_strong(ef, &old, nw, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT))
                        break;
                }
                if (in.disc) __hip_atomic_fetch_or(ef + 1, in.disc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            // distinct-key accounting: one atomic per wavefront
            const unsigned long long m = __ballot(pr.inserted);
            if (m) {
                const int leader = __ffsll((long long)m) - 1;
                if (lane == leader) {
                    const uint32_t add = uint32_t(__popcll(m));
                    const uint32_t before = atomicAdd(&a.n_vert[b], add);
                    if (before + add > limit) atomicOr(&a.overflow[b], 1u);
                }
            }
        }
    }
}

template <int KW, int WG>
__global__ __launch_bounds__(WG) void k_construct(ConstructArgs a, uint32_t item0, int persistent) {
    __shared__ uint32_t s_item;
    if (!persistent) {
        construct_item<KW, WG>(a, item0 + blockIdx.x);
        return;
    }
    while (true) {
        if (threadIdx.x == 0) s_item = atomicAdd(a.queue_head, 1u);
        __syncthreads();
        const uint32_t it = s_item;
        __syncthreads();
        if (it >= a.n_items) break;
        construct_item<KW, WG>(a, it);
    }
}

template <int KW>
__global__ __launch_bounds__(256) void k_clear(uint8_t* tables, uint64_t n) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        g_u64* p = slot_ptr<KW>((g_u8*)(tables), i);
#pragma unroll
        for (int j = 0; j < KW; ++j) p[j] = EMPTY;
        p[KW] = 0;  // e_f = 0, status = 0
    }
}

__global__ __launch_bounds__(256) void k_zero_buckets(const uint32_t* list, uint32_t n,
                                                      uint32_t* n_vert, uint32_t* overflow, uint32_t* iso) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const uint32_t b = list[i];
        n_vert[b] = 0;
        overflow[b] = 0;
        if (iso) iso[2 * b] = iso[2 * b + 1] = 0;
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

template <int KW>
__global__ __launch_bounds__(256) void k_export(const uint64_t* tbl_ptr,
                                                const uint64_t* tbl_cap, const uint32_t* buckets,
                                                const uint64_t* out_off, uint32_t* out_cursor,
                                                uint64_t* keys_out, uint32_t* ef_out,
                                                uint8_t* st_out, uint8_t* dfa_out,
                                                uint32_t cutoff) {
    const uint32_t bi = blockIdx.y;
    const uint32_t b = buckets[bi];
    const uint64_t cap = tbl_cap[b];
    const uint8_t* tbl = reinterpret_cast<const uint8_t*>(tbl_ptr[b]);
    const int lane = threadIdx.x & 63;
    for (uint64_t s = uint64_t(blockIdx.x) * 256 + threadIdx.x; s - threadIdx.x < cap;
         s += uint64_t(gridDim.x) * 256) {
        const bool in = s < cap;
        const uint64_t* p = in ? reinterpret_cast<const uint64_t*>(tbl + s * uint64_t(8 * KW + 8)) : nullptr;
        const bool occ = in && p[KW - 1] != EMPTY;
        const unsigned long long m = __ballot(occ);
        if (!m) continue;
        uint32_t base = 0;
        const int leader = __ffsll((long long)m) - 1;
        if (lane == leader) base = atomicAdd(&out_cursor[bi], uint32_t(__popcll(m)));
        base = __shfl(base, leader);
        // out_off has n_buckets + 1 entries: entries past a bucket's expected
        // count are counted but not written (the host reports the mismatch).
        const uint32_t rank = uint32_t(__popcll(m & ((1ull << lane) - 1)));
        const uint64_t o = out_off[bi] + base + rank;
        if (occ && o < out_off[bi + 1]) {
#pragma unroll
            for (int j = 0; j < KW; ++j) keys_out[o * KW + j] = p[j];
            const uint32_t e_f = uint32_t(p[KW]), st = uint32_t(p[KW] >> 32);
            ef_out[o] = e_f;
            st_out[o] = uint8_t(st);
            dfa_out[o] = uint8_t(edge_at(e_f, 0, cutoff) | (edge_at(e_f, 16, cutoff) << 3) |
                                 ((st & 2u) << 5) | ((st & 4u) << 5));
        }
    }
}

}  // namespace

void launch_clear_tables(uint8_t* tables, uint64_t n_slots, uint32_t KW, hipStream_t s) {
    if (!n_slots) return;
    const uint32_t grid = uint32_t(std::min<uint64_t>((n_slots + 255) / 256, 65536));
    switch (KW) {
        case 1: hipLaunchKernelGGL(k_clear<1>, dim3(grid), dim3(256), 0, s, tables, n_slots); break;
        case 2: hipLaunchKernelGGL(k_clear<2>, dim3(grid), dim3(256), 0, s, tables, n_slots); break;
        case 3: hipLaunchKernelGGL(k_clear<3>, dim3(grid), dim3(256), 0, s, tables, n_slots); break;
        default: hipLaunchKernelGGL(k_clear<4>, dim3(grid), dim3(256), 0, s, tables, n_slots); break;
    }
}

template <int KW>
static void launch_kw(const ConstructArgs& a, uint32_t item0, uint32_t grid, uint32_t wg,
                      int persistent, hipStream_t s) {
    if (wg == 512)
        hipLaunchKernelGGL((k_construct<KW, 512>), dim3(grid), dim3(512), 0, s, a, item0, persistent);
    else
        hipLaunchKernelGGL((k_construct<KW, 256>), dim3(grid), dim3(256), 0, s, a, item0, persistent);
}

static void launch_any(const ConstructArgs& a, uint32_t KW, uint32_t item0, uint32_t grid,
                       uint32_t wg, int persistent, hipStream_t s) {
    if (!grid) return;
    switch (KW) {
        case 1: launch_kw<1>(a, item0, grid, wg, persistent, s); break;
        case 2: launch_kw<2>(a, item0, grid, wg, persistent, s); break;
        case 3: launch_kw<3>(a, item0, grid, wg, persistent, s); break;
        default: launch_kw<4>(a, item0, grid, wg, persistent, s); break;
    }
}

void launch_construct(const ConstructArgs& a, uint32_t KW, uint32_t item0, uint32_t n, uint32_t wg,
                      hipStream_t s) {
    launch_any(a, KW, item0, n, wg, 0, s);
}

void launch_construct_persistent(const ConstructArgs& a, uint32_t KW, uint32_t grid, uint32_t wg,
                                 hipStream_t s) {
    launch_any(a, KW, 0, grid, wg, 1, s);
}

void launch_zero_buckets(const uint32_t* list, uint32_t n, uint32_t* n_vert, uint32_t* overflow, uint32_t* iso,
                         hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_zero_buckets, dim3((n + 255) / 256), dim3(256), 0, s, list, n, n_vert, overflow, iso);
}

void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
    if (!n) return;
    const uint32_t grid = uint32_t(std::min<uint64_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, s, p, v, n);
}

void launch_export(const uint64_t* tbl_ptr, const uint64_t* tbl_cap,
                   uint32_t KW, const uint32_t* buckets, uint32_t n_buckets, const uint64_t* out_off,
                   uint32_t* out_cursor, uint64_t* keys_out, uint32_t* ef_out, uint8_t* st_out,
                   uint8_t* dfa_out, uint32_t cutoff, hipStream_t s) {
    if (!n_buckets) return;
    const dim3 grid(64, n_buckets);
    switch (KW) {
        case 1: hipLaunchKernelGGL(k_export<1>, grid, dim3(256), 0, s, tbl_ptr, tbl_cap, buckets, out_off, out_cursor, keys_out, ef_out, st_out, dfa_out, cutoff); break;
        case 2: hipLaunchKernelGGL(k_export<2>, grid, dim3(256), 0, s, tbl_ptr, tbl_cap, buckets, out_off, out_cursor, keys_out, ef_out, st_out, dfa_out, cutoff); break;
        case 3: hipLaunchKernelGGL(k_export<3>, grid, dim3(256), 0, s, tbl_ptr, tbl_cap, buckets, out_off, out_cursor, keys_out, ef_out, st_out, dfa_out, cutoff); break;
        default: hipLaunchKernelGGL(k_export<4>, grid, dim3(256), 0, s, tbl_ptr, tbl_cap, buckets, out_off, out_cursor, keys_out, ef_out, st_out, dfa_out, cutoff); break;
    }
}

}  // namespace cfgpu
