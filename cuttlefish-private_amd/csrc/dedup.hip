// Exact weak super k-mer deduplication per bucket.
//
// Two records of one bucket with the same length, discontinuity bits and
// label yield the same k-mer instances with the same neighbours and marks in
// Subgraph::construct (src/Subgraph.cpp:57-105).  Counters saturate
// (include/State_Config.hpp:207-217), so adding a record m times equals adding
// min(15, count + m) once, and status bits are ORed.  Reads sequenced at
// coverage c repeat most records ~c/2 times per strand, so construct walks the
// distinct records with their multiplicities.
//
// One work item = one hash pass of one bucket: records whose key hash selects
// the pass are counted in an LDS table (key = the label words with bases past
// `len` cleared, then len | l_disc | r_disc | hash << 10 as the top word; or,
// when the label's last word holds at most 15 bases, that word shifted right
// by one with len | disc bits | 22 hash bits in its free low bits as the top
// word, one word fewer), and the distinct
// records are appended to the bucket's region of the output arrays.  A pass
// whose table fills past 7/8 flags the bucket; the host then copies its
// records unchanged (multiplicity 1), so results never depend on the sizing.
#include "cf_kernels.hpp"

namespace cfgpu {
namespace {

constexpr uint64_t EMPTY = ~0ull;
constexpr uint64_t BUSY = 1ull << 63;

// Table layout: the top words of all C slots first, then the lower KR - 1
// words slot by slot (dd_low).  Slot j (0..3, the fill order) of group g is
// slot 4g + ((j + r) & 3) for r = dd_rot(g), so that the j-th top words of the
// unrelated groups a wave probes in one instruction fall on LDS banks spread
// over the whole bank set (at a fixed offset modulo the group size they would
// share a few banks).
__device__ __forceinline__ uint32_t dd_rot(uint32_t g) { return ((g >> 2) + (g >> 3)) & 3u; }
template <int KR, int C>
__device__ __forceinline__ uint64_t* dd_low(uint64_t* s_key, uint32_t s, int w) {
    return &s_key[C + s * uint32_t(KR - 1) + uint32_t(w)];
}

// Find or insert a KR-word key (word KR-1 is the top word, < 2^63) in an LDS
// table of C slots by 4-slot group probing; the claim protocol of the
// construct tables (top word claimed with BUSY, lower words written, top word
// published).  Returns the slot or -1 when every group is full.
template <int KR, int C>
__device__ __forceinline__ int dd_find_or_insert(uint64_t* s_key, const uint64_t (&key)[KR], uint32_t g,
                                                 bool& inserted) {
    constexpr uint32_t NG = uint32_t(C) / 4;
    const uint64_t top = key[KR - 1];
    inserted = false;
    for (uint32_t n = 0; n < NG;) {
        const uint32_t rot = dd_rot(g);
        uint64_t cur[4];  // in the group's fill order
#pragma unroll
        for (int j = 0; j < 4; ++j)
            cur[j] = __hip_atomic_load(&s_key[g * 4 + ((uint32_t(j) + rot) & 3u)], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        const int empty = cur[0] == EMPTY ? 0 : cur[1] == EMPTY ? 1 : cur[2] == EMPTY ? 2 : cur[3] == EMPTY ? 3 : 4;
        uint32_t cand = 0;  // published slots whose top word matches (scanned without branches)
        bool busy = false;  // a matching top word is still being published
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool c = j < empty && (cur[j] & ~BUSY) == top;
            cand |= uint32_t(c && !(cur[j] & BUSY)) << j;
            busy |= c && (cur[j] & BUSY);
        }
        while (cand) {
            const uint32_t slot = g * 4 + ((uint32_t(__builtin_ctz(cand)) + rot) & 3u);
            cand &= cand - 1;
            bool eq = true;
#pragma unroll
            for (int w = 0; w < KR - 1; ++w)
                eq &= __hip_atomic_load(dd_low<KR, C>(s_key, slot, w), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) ==
                      key[w];
            if (eq) return int(slot);
        }
        if (busy) continue;
        if (empty == 4) {
            g = g + 1 == NG ? 0 : g + 1;
            ++n;
            continue;
        }
        const uint32_t slot = g * 4 + ((uint32_t(empty) + rot) & 3u);
        uint64_t* kp = &s_key[slot];
        if (atomicCAS(reinterpret_cast<unsigned long long*>(kp), EMPTY, top | BUSY) == EMPTY) {
#pragma unroll
            for (int w = 0; w < KR - 1; ++w)
                __hip_atomic_store(dd_low<KR, C>(s_key, slot, w), key[w], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // LDS table: the workgroup is the only observer
            __hip_atomic_store(kp, top, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            inserted = true;
            return int(slot);
        }
    }
    return -1;
}

// Key of a record (att, W label words): bases past len cleared; PK: label
// words 1..W-1 then word 0 (the record's last bases, low bits free) packed as
// the top word, else the W label words then a top word of their own.  The top
// word also carries h (a function of the rest of the key), so that it alone
// tells almost every pair of distinct keys apart; h is returned (its top
// selects the pass, the fraction below it the group).
template <int W, bool PK, int KR>
__device__ __forceinline__ uint32_t dd_key(uint32_t at, const uint64_t (&lab)[W], uint64_t (&key)[KR]) {
    const uint32_t len = at & 0xFFu;
#pragma unroll
    for (int q = 0; q < W; ++q) {  // word W-1-i holds bases 32i .. 32i+31
        const int i = W - 1 - q;
        const int nb = int(len) - 32 * i;
        const uint64_t x = lab[q];
        key[PK ? (q + W - 1) % W : q] = nb >= 32 ? x : (nb <= 0 ? 0ull : x & (~0ull << (64 - 2 * nb)));
    }
    uint32_t h;
    if constexpr (PK) {
        key[KR - 1] = (key[KR - 1] >> 1) | (at & 0x3FFu);  // bases in bits 33..62
        h = pass_hash32<KR>(key);
        key[KR - 1] |= uint64_t(h & ~0x3FFu);
    } else {
        key[KR - 1] = at & 0x3FFu;
        h = pass_hash32<KR>(key);
        key[KR - 1] |= uint64_t(h) << 10;
    }
    return h;
}

template <int W>
__device__ __forceinline__ void dd_load(const DedupArgs& a, uint64_t r, uint32_t& at, uint64_t (&lab)[W]) {
    at = a.att[r];
#pragma unroll
    for (int q = 0; q < W; ++q) lab[q] = a.labels[r * W + q];
}

// One work item is one hash pass of one bucket, or (item_pass ==
// DD_ALL_PASSES) every pass of a bucket of at most DD_REGION_RECORDS records:
// the workgroup then hashes each record once, lists the records of each pass
// in its region of `region` (a pass byte and a u32 index per record, the
// indices grouped by pass at exact offsets), and probes pass after pass over
// its list with every lane busy.
template <int W, bool PK>
__global__ __launch_bounds__(DD_WG) void k_dedup(DedupArgs a) {
    constexpr int KR = PK ? W : W + 1;  // PK: len, disc bits and hash packed into the last label word
    constexpr int C = dedup_capacity(KR);
    static_assert(C % 64 == 0, "capacity");
    __shared__ uint64_t s_key[C * KR];  // top words, then lower words (dd_low)
    __shared__ uint32_t s_cnt[C];
    __shared__ uint32_t s_item, s_n, s_ovf, s_base;
    __shared__ uint32_t s_len[256], s_wsum[4];  // per-len counts, then output positions
    __shared__ uint32_t s_pc[DD_ALL_PASS_MAX], s_po[DD_ALL_PASS_MAX + 1];  // records per pass, their offsets
    __shared__ unsigned long long s_km;
    const int tid = threadIdx.x;
    constexpr uint32_t LIMIT = (uint32_t(C) * 7u) / 8u;
    uint8_t* rq = a.region + uint64_t(blockIdx.x) * DD_REGION_RECORDS * 5;  // pass of each record
    uint32_t* ridx = reinterpret_cast<uint32_t*>(rq + DD_REGION_RECORDS);   // records grouped by pass

    while (true) {
        __syncthreads();
        if (tid == 0) {
            s_item = atomicAdd(a.queue_head, 1u);
            s_ovf = 0;
        }
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= a.n_items) break;
        const uint32_t b = a.item_bucket[item], p_it = a.item_pass[item], P = a.item_passes[item];
        const bool all = p_it == DD_ALL_PASSES;
        if (tid == 0) s_ovf = __hip_atomic_load(&a.overflow[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t r0 = a.bucket_off[b], r1 = a.bucket_off[b + 1];
        const uint32_t n = uint32_t(r1 - r0);
        if (all) {
            // each record's pass (one hash per record), then its index at its
            // pass's exact place
            for (int i = tid; i < DD_ALL_PASS_MAX; i += DD_WG) s_pc[i] = 0;
            __syncthreads();
            for (uint32_t r = tid; r < n; r += DD_WG) {
                uint32_t at;
                uint64_t lab[W], key[KR];
                dd_load<W>(a, r0 + r, at, lab);
                const uint32_t q = uint32_t((uint64_t(dd_key<W, PK, KR>(at, lab, key)) * P) >> 32);
                rq[r] = uint8_t(q);
                atomicAdd(&s_pc[q], 1u);
            }
            __syncthreads();
            if (tid == 0) {
                uint32_t o = 0;
                for (uint32_t q = 0; q < P; ++q) {
                    s_po[q] = o;
                    o += s_pc[q];
                    s_pc[q] = s_po[q];  // now the cursor of pass q
                }
                s_po[P] = o;
            }
            __syncthreads();
            for (uint32_t r = tid; r < n; r += DD_WG) ridx[atomicAdd(&s_pc[rq[r]], 1u)] = r;
        }
        for (uint32_t p = all ? 0u : p_it; p <= (all ? P - 1 : p_it); ++p) {
            if (tid == 0) s_n = 0;
            for (int i = tid; i < C; i += DD_WG) s_key[i] = EMPTY;  // top words (lower ones are written at a claim)
            for (int i = tid; i < C; i += DD_WG) s_cnt[i] = 0;
            __syncthreads();
            if (s_ovf) break;
            // one pass: the records of the pass's list (all), or every record
            // of the bucket, keeping those of pass p; the next record is
            // loaded while the current one is probed
            const uint32_t e0 = all ? s_po[p] : 0u, e1 = all ? s_po[p + 1] : n;
            uint32_t my_ins = 0, seen = 0;
            uint32_t at_n = 0, ix_n = 0;
            uint64_t lab_n[W];
            if (e0 + tid < e1) {
                const uint32_t r = all ? ridx[e0 + tid] : e0 + tid;
                dd_load<W>(a, r0 + r, at_n, lab_n);
                if (all && e0 + tid + DD_WG < e1) ix_n = ridx[e0 + tid + DD_WG];
            }
            for (uint32_t e = e0 + tid; e < e1; e += DD_WG) {
                const uint32_t at = at_n;
                uint64_t lab[W];
#pragma unroll
                for (int q = 0; q < W; ++q) lab[q] = lab_n[q];
                if (e + DD_WG < e1) {
                    const uint32_t r = all ? ix_n : e + DD_WG;
                    dd_load<W>(a, r0 + r, at_n, lab_n);
                    if (all && e + 2 * DD_WG < e1) ix_n = ridx[e + 2 * DD_WG];
                }
                if ((++seen & 7u) == 0) {  // account inserts every 8 records
                    if (my_ins && atomicAdd(&s_n, my_ins) + my_ins > LIMIT) s_ovf = 1;
                    my_ins = 0;
                    if (__hip_atomic_load(&s_ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
                }
                uint64_t key[KR];
                const uint64_t hp = uint64_t(dd_key<W, PK, KR>(at, lab, key)) * P;  // pass = top, group from the fraction
                if (!all && uint32_t(hp >> 32) != p) continue;
                bool inserted;
                const int slot = dd_find_or_insert<KR, C>(
                    s_key, key, uint32_t((uint64_t(uint32_t(hp)) * (uint32_t(C) / 4)) >> 32), inserted);
                if (slot < 0) {
                    s_ovf = 1;
                    break;
                }
                atomicAdd(&s_cnt[slot], 1u);
                my_ins += inserted;
            }
            if (my_ins && atomicAdd(&s_n, my_ins) + my_ins > LIMIT) s_ovf = 1;
            __syncthreads();
            if (s_ovf) break;
            // Append the pass's distinct records to the bucket's output region,
            // longest first (a counting sort on len), so that a construct lane
            // assignment that alternates direction over blocks of records gives
            // every lane about the same number of k-mers to walk.
            if (tid == 0) {
                s_base = atomicAdd(&a.out_count[b], s_n);
                s_km = 0;
            }
            if (tid < 256) s_len[tid] = 0;
            __syncthreads();
            for (int i = tid; i < C; i += DD_WG)
                if (s_cnt[i]) atomicAdd(&s_len[s_key[i] & 0xFFu], 1u);
            __syncthreads();
            // exclusive prefix over len = 255, 254, ..., 0 by the first 4 waves
            const int lane = tid & 63, wv = tid >> 6;
            uint32_t v = 0, x = 0;
            if (tid < 256) {
                v = x = s_len[255 - tid];
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const uint32_t u = __shfl_up(x, d);
                    if (lane >= d) x += u;
                }
                if (lane == 63) s_wsum[wv] = x;
            }
            __syncthreads();
            if (tid < 256) {
                uint32_t pre = 0;
                for (int j = 0; j < wv; ++j) pre += s_wsum[j];
                s_len[255 - tid] = pre + x - v;  // first output position of this len
            }
            __syncthreads();
            unsigned long long my_km = 0;
            const uint64_t base = r0 + s_base;
            const uint32_t g16 = uint32_t(b) << 16;
            for (int i = tid; i < C; i += DD_WG) {
                const uint32_t c = s_cnt[i];
                if (!c) continue;
                const uint64_t top = s_key[i];
                const uint64_t* kp = dd_low<KR, C>(s_key, uint32_t(i), 0);  // the KR - 1 lower words
                const uint64_t o = base + atomicAdd(&s_len[top & 0xFFu], 1u);
                my_km += (top & 0xFFu) - (a.k - 1);
                a.out_att[o] = uint32_t(top & 0x3FFu) | g16;
                a.out_mult[o] = c;
#pragma unroll
                for (int q = 0; q < W; ++q)
                    a.out_labels[o * W + q] = !PK ? (q < KR - 1 ? kp[q] : top) : q == 0 ? (top >> 32) << 33 : kp[q - 1];
            }
            if (my_km) atomicAdd(&s_km, my_km);
            __syncthreads();
            if (tid == 0 && s_km) atomicAdd(&a.out_kmers[b], s_km);
        }
        if (s_ovf && tid == 0) atomicOr(&a.overflow[b], 1u);
    }
}

// Copy the records of flagged buckets unchanged (multiplicity 1).
__global__ __launch_bounds__(256) void k_dedup_copy(DedupArgs a, const uint32_t* __restrict__ list, uint32_t W) {
    const uint32_t b = list[blockIdx.x];
    const uint64_t r0 = a.bucket_off[b], r1 = a.bucket_off[b + 1];
    __shared__ unsigned long long s_km;
    if (threadIdx.x == 0) s_km = 0;
    __syncthreads();
    unsigned long long km = 0;
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += 256) {
        const uint32_t at = a.att[r];
        a.out_att[r] = at;
        a.out_mult[r] = 1u;
        km += (at & 0xFFu) - (a.k - 1);
        for (uint32_t q = 0; q < W; ++q) a.out_labels[r * W + q] = a.labels[r * W + q];
    }
    atomicAdd(&s_km, km);
    __syncthreads();
    if (threadIdx.x == 0) {
        a.out_count[b] = uint32_t(r1 - r0);
        a.out_kmers[b] = s_km;
    }
}

// bucket_end[b] = bucket_off[b] + count[b] for every b.
__global__ __launch_bounds__(256) void k_dedup_end(const uint64_t* __restrict__ off, const uint32_t* __restrict__ cnt,
                                                   uint32_t G, uint64_t* __restrict__ end) {
    const uint32_t b = blockIdx.x * 256 + threadIdx.x;
    if (b < G) end[b] = off[b] + cnt[b];
}

}  // namespace

void launch_dedup(const DedupArgs& a, uint32_t W, bool packed, uint32_t grid, hipStream_t s) {
    if (!grid) return;
#define CF_DD(w)                                                                                        \
    case w:                                                                                             \
        if (packed) hipLaunchKernelGGL((k_dedup<w, true>), dim3(grid), dim3(DD_WG), 0, s, a);           \
        else hipLaunchKernelGGL((k_dedup<w, false>), dim3(grid), dim3(DD_WG), 0, s, a);                 \
        break;
    switch (W) {
        CF_DD(1) CF_DD(2) CF_DD(3) CF_DD(4) CF_DD(5) CF_DD(6) CF_DD(7)
        default: CF_DD(8)
    }
#undef CF_DD
}

void launch_dedup_copy(const DedupArgs& a, const uint32_t* list, uint32_t n, uint32_t W, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_dedup_copy, dim3(n), dim3(256), 0, s, a, list, W);
}

void launch_dedup_end(const uint64_t* off, const uint32_t* cnt, uint32_t G, uint64_t* end, hipStream_t s) {
    hipLaunchKernelGGL(k_dedup_end, dim3((G + 255) / 256), dim3(256), 0, s, off, cnt, G, end);
}

}  // namespace cfgpu
