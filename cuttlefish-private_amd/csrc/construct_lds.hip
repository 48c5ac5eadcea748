// LDS-resident subgraph construct.
//
// Device-scope atomics on HBM-resident tables are serviced past the per-XCD L2,
// so a table built with global CAS costs a memory transaction per k-mer
// instance.  This kernel instead builds one bucket's table inside a workgroup's
// LDS and writes the finished table to HBM with plain, coalesced stores.
//
// A bucket whose distinct k-mers do not fit one LDS table is split by hash into
// P passes: pass p keeps the k-mers whose 32-bit pass hash h has
// (h * P) >> 32 == p.  Each finished pass appends its occupied slots to the
// bucket's HBM table, which ends up holding exactly the bucket's vertices
// (its allocation is P * C entries; the rest is never written or read).  A bucket of 2..LDS_ALL_PASS_MAX passes is one work
// item: its records are walked once, every instance going to its pass's
// sub-list of a per-workgroup spill region, and each pass probes its own
// sub-list with every lane busy (each instance is written and read once,
// whatever P is).  Other buckets are one item per pass (P = 1, or buckets too
// large for the spill regions), each walking the records.
//
// Each lane walks a stream of whole records (RecordWalker: the k-mer and its
// reverse complement are rolled one base per instance), so no staging
// barriers or per-instance searches are needed and the LDS holds only the
// table: u64 key words, the u32 Edge_Frequency word and one Vertex_Status byte
// per slot.  The key words are split: the top words of all slots first (a
// group's four top words are what every probe step reads), then the lower
// words, slot by slot.
//
// The per-instance semantics are those of Subgraph::construct
// (src/Subgraph.cpp:44-108); counters saturate exactly as
// Edge_Frequency::add_edge (include/State_Config.hpp:207-217).  A record with
// multiplicity m (deduplicated records) adds m to each counter it touches.  A pass whose
// table fills past 7/8 flags the bucket as overflowed; the host rebuilds it on
// the global path, so capacity choices never change results.
#include "cf_instance.hpp"
#include "cf_kernels.hpp"

namespace cfgpu {
namespace {

constexpr uint64_t EMPTY = ~0ull;
constexpr uint64_t BUSY = 1ull << 63;
constexpr int LWG = LDS_WG;

// Pass and group choice of these tables: pass_hash32 of the key words.  Only
// the spread of keys over passes and groups depends on it: keys are always
// compared whole, and nothing looks a key up in a finished table by it.
template <int KW>
__device__ __forceinline__ uint32_t pass_hash(const Kmer<KW>& x) {
    return pass_hash32<KW>(x.w);
}

constexpr int UPSERT_FULL = -1;  // every group is full without the key

// Group probe order of a key: its home group g0, then g0 + s, g0 + 2s, ...
// modulo the NG groups, for an odd stride s < 64 coprime to NG (so the order
// visits every group) taken from 5 hash bits; the few odd values that share a
// factor with NG are replaced by 1.  Keys that share a home group mostly go
// on with different strides, so the full groups do not gather into the runs
// that every later search through them must cross (the clusters of a probe by
// consecutive groups).
constexpr uint32_t gcd_u32(uint32_t a, uint32_t b) { return b ? gcd_u32(b, a % b) : a; }
constexpr uint64_t stride_bad_mask(uint32_t ng) {  // bit i: i shares a factor with ng
    uint64_t m = 0;
    for (uint32_t i = 0; i < 64; ++i)
        if (gcd_u32(i, ng) != 1) m |= 1ull << i;
    return m;
}
template <uint32_t NG>
__device__ __forceinline__ uint32_t group_stride(uint32_t h) {
    static_assert(NG > 64, "strides below the group count");
    constexpr uint64_t BAD = stride_bad_mask(NG);
    const uint32_t s = 2 * (h & 31u) + 1;
    return (BAD >> s) & 1u ? 1u : s;
}

// Add one k-mer instance to an LDS table of C slots (KW u64 words each): find
// or insert `key`, add m to the counters of its front / back bases (saturating
// at 15, Edge_Frequency::add_edge, include/State_Config.hpp:207-217) and OR
// its discontinuity bits into the status byte.  Returns the slot, or
// UPSERT_FULL.
//
// Groups of 4 slots are probed in the key's group order (group_stride): the
// 4 top words are read together, and a group fills from its first slot on
// (claims take the first empty slot; entries are never removed), so a group
// with an empty slot ends an unsuccessful search.  The group is scanned
// without branches (select chains, a candidate bit mask); a top word is below
// 2^62 (k odd: at most 31 bases in the top word), so neither EMPTY nor a
// claimed word of another key equals the key's plain or claimed top word.  A slot is claimed by writing its top word
// with BUSY set; its lower key words and its first counters are written, and
// a release fence orders them before the plain top word publishes the slot.
// A lane that finds the key reads its counters together with its lower key
// words and updates them by CAS; a lane that meets a claimed slot with its own
// top word re-reads the group.
//
// One group read is one step (upsert_step, STEP_DONE / STEP_AGAIN /
// STEP_FULL): lds_upsert repeats it until the instance is added; the pass
// phase's lanes each take one step per loop trip and a new instance when
// theirs is added.  Every group before g in the key's probe order is full
// without the key, and stays so, so a search that goes on from g adds the
// instance where a search from its home group would.
constexpr int STEP_DONE = 0, STEP_AGAIN = 1, STEP_FULL = 2;
// Slot j (0..3, the group's fill order) of group g is slot 4g + ((j + r) & 3)
// for r = group_rot(g): the lanes of a wave read the j-th top words of
// unrelated groups in one instruction, and with the rotation those words fall
// on LDS banks spread over the whole bank set (unrotated, every group's j-th
// top word would sit at the same offset modulo 64 bytes: a few banks, many
// conflicts).
__device__ __forceinline__ uint32_t group_rot(uint32_t g) { return ((g >> 2) + (g >> 3)) & 3u; }

// Key words of slot s in a table of C slots: the top word at s_key[s], word w
// < KW - 1 at s_key[C + s (KW - 1) + w].
template <int KW, int C>
__device__ __forceinline__ uint64_t* low_word(uint64_t* s_key, uint32_t s, int w) {
    return &s_key[C + s * uint32_t(KW - 1) + uint32_t(w)];
}

template <int KW, int C>
__device__ __forceinline__ int upsert_step(uint64_t* s_key, uint32_t* s_ef, uint32_t* s_st, const Kmer<KW>& key,
                                           uint32_t& g, uint32_t stride, uint32_t& n_full, uint32_t front,
                                           uint32_t back, uint32_t disc, uint32_t m, uint32_t& my_ins) {
    constexpr uint32_t NG = uint32_t(C) / 4;
    const uint64_t top = key.w[KW - 1];
    {
        const uint32_t rot = group_rot(g);
        uint64_t cur[4];  // in the group's fill order
#pragma unroll
        for (int j = 0; j < 4; ++j)
            cur[j] = __hip_atomic_load(&s_key[g * 4 + ((uint32_t(j) + rot) & 3u)], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        const int empty = cur[0] == EMPTY ? 0 : cur[1] == EMPTY ? 1 : cur[2] == EMPTY ? 2 : cur[3] == EMPTY ? 3 : 4;
        uint32_t cand = 0;  // published slots whose top word matches
        bool busy = false;  // a matching top word is still being published
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cand |= uint32_t(cur[j] == top) << j;
            busy |= cur[j] == (top | BUSY);
        }
        int hit = -1;
        uint32_t old = 0;
        while (cand) {
            const uint32_t slot = g * 4 + ((uint32_t(__builtin_ctz(cand)) + rot) & 3u);
            cand &= cand - 1;
            old = __hip_atomic_load(&s_ef[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            bool eq = true;
#pragma unroll
            for (int w = 0; w < KW - 1; ++w)
                eq &= __hip_atomic_load(low_word<KW, C>(s_key, slot, w), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_WORKGROUP) == key.w[w];
            if (eq) {
                hit = int(slot);
                break;
            }
        }
        if (hit >= 0) {
            // a stale counter word only costs a CAS retry
            while (true) {
                const uint32_t nw = ef_add(old, uint8_t(front), uint8_t(back), m);
                if (nw == old) break;
                const uint32_t prev = atomicCAS(&s_ef[hit], old, nw);
                if (prev == old) break;
                old = prev;
            }
            if (disc) atomicOr(&s_st[hit >> 2], disc << (8 * (hit & 3)));
            return STEP_DONE;
        }
        if (busy) return STEP_AGAIN;  // re-read the group once it is published
        if (empty == 4) {             // full group without the key: next group
            g += stride;
            g = g >= NG ? g - NG : g;
            return ++n_full == NG ? STEP_FULL : STEP_AGAIN;
        }
        const uint32_t slot = g * 4 + ((uint32_t(empty) + rot) & 3u);
        uint64_t* kp = &s_key[slot];
        if (atomicCAS(reinterpret_cast<unsigned long long*>(kp), EMPTY, top | BUSY) == EMPTY) {
#pragma unroll
            for (int w = 0; w < KW - 1; ++w)
                __hip_atomic_store(low_word<KW, C>(s_key, slot, w), key.w[w], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store(&s_ef[slot], ef_add(0u, uint8_t(front), uint8_t(back), m), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // LDS table: the workgroup is the only observer
            __hip_atomic_store(kp, top, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (disc) atomicOr(&s_st[slot >> 2], disc << (8 * (slot & 3)));
            ++my_ins;
            return STEP_DONE;
        }
        return STEP_AGAIN;  // lost the slot to another key: re-read the group
    }
}

template <int KW, int C>
__device__ __forceinline__ int lds_upsert(uint64_t* s_key, uint32_t* s_ef, uint32_t* s_st, const Kmer<KW>& key,
                                          uint32_t g, uint32_t stride, uint32_t front, uint32_t back, uint32_t disc,
                                          uint32_t m, uint32_t& my_ins) {
    uint32_t n_full = 0;
    while (true) {
        const int r = upsert_step<KW, C>(s_key, s_ef, s_st, key, g, stride, n_full, front, back, disc, m, my_ins);
        if (r != STEP_AGAIN) return r == STEP_DONE ? 0 : UPSERT_FULL;
    }
}

// Spill region of a workgroup for one all-pass item ((2 KW + 1) spill_cap <
// 2^32 words: the host bounds the regions' bytes): P sub-lists of c =
// spill_cap / P entries, sub-list q holding the instances of pass q.  Entry e
// is 2 KW + 1 consecutive u32 words at 4 (2 KW + 1) e bytes: the KW key words
// (low half first), then meta = front | back << 3 | disc << 6 | min(m, 15) <<
// 9 | group << 13 | stride << 25 (counters saturate at 15, so adding min(m,
// 15) equals adding m; the group of the pass table, below 2^12, and the group
// stride, below 2^6, are computed once by the walk).  A lane's entry is one
// contiguous span, so a wave's appends to a few sub-lists touch few lines.
template <int KW>
struct SpillEnt {
    uint32_t d[2 * KW + 1];
};
struct SpillRegion {
    uint32_t* base;
};

template <int KW>
__device__ __forceinline__ SpillRegion spill_region(const ConstructArgs& a) {
    return {reinterpret_cast<uint32_t*>(a.spill_key) + uint64_t(blockIdx.x) * (2 * KW + 1) * a.spill_cap};
}

template <int KW>
__device__ __forceinline__ void spill_store(const SpillRegion& R, uint32_t e, const Kmer<KW>& key, uint32_t meta) {
    SpillEnt<KW> v;
#pragma unroll
    for (int w = 0; w < KW; ++w) {
        v.d[2 * w] = uint32_t(key.w[w]);
        v.d[2 * w + 1] = uint32_t(key.w[w] >> 32);
    }
    v.d[2 * KW] = meta;
    *reinterpret_cast<SpillEnt<KW>*>(R.base + e * uint32_t(2 * KW + 1)) = v;
}

template <int KW>
__device__ __forceinline__ void spill_load(const SpillRegion& R, uint32_t e, Kmer<KW>& key, uint32_t& meta) {
    const SpillEnt<KW> v = *reinterpret_cast<const SpillEnt<KW>*>(R.base + e * uint32_t(2 * KW + 1));
#pragma unroll
    for (int w = 0; w < KW; ++w) key.w[w] = uint64_t(v.d[2 * w]) | (uint64_t(v.d[2 * w + 1]) << 32);
    meta = v.d[2 * KW];
}

// Append the calling lanes' instances to their passes' sub-lists.  With more
// than 4 passes each lane reserves its entry with its own LDS atomic (lanes
// of a wave spread over many counters; the LDS unit resolves the few that
// share one, and the VALU, the kernel's bottleneck, issues one instruction).
// With at most 4 passes, where up to 32 lanes would share a counter, the
// lanes are split by pass with nbits ballots (one per bit of q): each group's
// first lane reserves the group's entries with one LDS atomic, and the lanes
// take consecutive entries in lane order.  A full sub-list sets *ovf (the
// bucket is then rebuilt on the global path).
template <int KW>
__device__ __forceinline__ void spill_push_pass(uint32_t q, uint32_t nbits, const Kmer<KW>& key, uint32_t meta,
                                                const SpillRegion& R, uint32_t cap, uint32_t c, uint32_t* s_n,
                                                uint32_t* s_ovf) {
    uint32_t idx;
    if (nbits > 2) {
        idx = atomicAdd(&s_n[q], 1u);
    } else {
        unsigned long long same = __ballot(true);
        for (uint32_t bit = 0; bit < nbits; ++bit) {
            const bool set = (q >> bit) & 1u;
            const unsigned long long m = __ballot(set);
            same &= set ? m : ~m;
        }
        const int lane = int(__lane_id());
        const int leader = __ffsll((long long)same) - 1;
        const uint32_t rank =
            __builtin_amdgcn_mbcnt_hi(uint32_t(same >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(same), 0u));
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&s_n[q], uint32_t(__popcll(same)));
        base = __builtin_amdgcn_ds_bpermute(leader << 2, int(base));
        idx = base + rank;
    }
    if (idx >= c) {
        *s_ovf = 1;
        return;
    }
    spill_store<KW>(R, q * c + idx, key, meta);
}

template <int KW, int WMAX>
__global__ __launch_bounds__(LWG) void k_construct_lds(ConstructArgs a) {
    constexpr int C = lds_capacity(KW);
    static_assert(C % 64 == 0, "capacity");
    static_assert(C / 4 <= (1 << 12), "a group index fits the spill meta word");
    __shared__ uint64_t s_key[C * KW];  // top words, then lower words (low_word)
    __shared__ uint32_t s_ef[C];
    __shared__ uint32_t s_st[C / 4];  // one Vertex_Status byte per slot
    __shared__ uint32_t s_item, s_cnt, s_ovf, s_base, s_wp;
    __shared__ uint32_t s_next;  // pass phase: entries of the sub-list handed out
    __shared__ unsigned long long s_wai;  // write-out: active entries | isolated entries << 32
    __shared__ uint32_t s_sn[LDS_ALL_PASS_MAX];  // sub-list lengths of an all-pass item

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t W = a.W, k = a.k;
    constexpr uint32_t LIMIT = (uint32_t(C) * 7u) / 8u;
    constexpr uint32_t NG = uint32_t(C) / 4;
    using Walker = RecordWalker<KW, WMAX>;

    while (true) {
        __syncthreads();  // every thread is done with the previous item's flags
        if (tid == 0) {
            s_item = atomicAdd(a.queue_head, 1u);
            s_ovf = 0;
        }
        __syncthreads();
        const uint32_t item = s_item;
        if (item >= a.n_items) break;
        const uint32_t b = a.item_bucket[item];
        const uint32_t P = a.item_count[item];  // passes of this bucket
        // item_first: the pass of this item, or LDS_ALL_PASSES: every pass
        // here.  The records are then walked once, every instance going to
        // its pass's sub-list, and each pass probes its own sub-list with
        // every lane busy.
        const bool all = a.item_first[item] == LDS_ALL_PASSES;
        const uint32_t p_lo = all ? 0u : a.item_first[item], p_hi = all ? P - 1 : p_lo;
        const uint32_t c_sub = all ? a.spill_cap / P : 0u;  // entries per sub-list
        const SpillRegion R = spill_region<KW>(a);
        if (tid == 0) s_ovf = __hip_atomic_load(&a.overflow[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (all)
            for (uint32_t i = tid; i < P; i += LWG) s_sn[i] = 0;

        for (uint32_t p = p_lo; p <= p_hi; ++p) {
            if (!(all && p == p_lo)) {  // the all-pass walk needs no table
                for (int i = tid; i < C; i += LWG) s_key[i] = EMPTY;  // top words (lower ones are written at a claim)
                for (int i = tid; i < C; i += LWG) s_ef[i] = 0u;
                for (int i = tid; i < C / 4; i += LWG) s_st[i] = 0u;
            }
            if (tid == 0) s_cnt = s_next = 0;
            __syncthreads();
            if (s_ovf) break;  // bucket already flagged
            uint32_t my_ins = 0;  // keys this lane inserted since its last flush into s_cnt
            const auto flush = [&]() {
                if (my_ins) {
                    if (atomicAdd(&s_cnt, my_ins) + my_ins > LIMIT) s_ovf = 1;
                    my_ins = 0;
                }
            };

            if (p == p_lo) {
                // Each lane walks one record of every block of LWG records,
                // taking them by lane index in even blocks and in reverse in
                // odd blocks, and prefetches its next record while walking the
                // current one.  Records sorted by length (deduplicated buckets
                // are) then give every lane nearly the same number of k-mers,
                // and the lanes of a wave switch records together.
                const uint64_t r_end = a.bucket_end ? a.bucket_end[b] : a.bucket_off[b + 1];
                const uint64_t r0 = a.bucket_off[b];
                const auto rec_of = [&](uint32_t j) {  // strictly increasing in j
                    return r0 + uint64_t(j) * LWG + ((j & 1u) ? uint32_t(LWG - 1 - tid) : uint32_t(tid));
                };
                const uint32_t nbits = all ? 32u - uint32_t(__builtin_clz((P - 1) | 1u)) : 0u;  // bits of q < P
                uint32_t j = 0;
                uint64_t r = rec_of(0), rn = rec_of(1);
                Walker rw;
                uint32_t at_next = 0, m = 1, m_next = 1;
                uint64_t lab_next[WMAX];
                if (r < r_end) {
                    rw.load(a.att, a.labels, r, W, k);
                    if (a.mult) m = a.mult[r];
                    if (rn < r_end) {
                        Walker::fetch(a.att, a.labels, rn, W, at_next, lab_next);
                        if (a.mult) m_next = a.mult[rn];
                    }
                }
                while (r < r_end) {
                    const Inst<KW> in = rw.next();
                    // pass = top of h * P; the group comes from the fraction
                    // below it, the group stride from the low bits of h
                    const uint32_t h = pass_hash<KW>(in.key);
                    const uint64_t hp = uint64_t(h) * P;
                    const uint32_t q = uint32_t(hp >> 32);
                    const uint32_t g0 = uint32_t((uint64_t(uint32_t(hp)) * NG) >> 32);
                    const uint32_t sg = group_stride<NG>(h);
                    if (all) {
                        spill_push_pass<KW>(q, nbits, in.key,
                                            in.front | (in.back << 3) | (in.disc << 6) | (min(m, 15u) << 9) |
                                                (g0 << 13) | (sg << 25),
                                            R, a.spill_cap, c_sub, s_sn, &s_ovf);
                    } else if (q == p) {
                        if (lds_upsert<KW, C>(s_key, s_ef, s_st, in.key, g0, sg, in.front, in.back, in.disc, m,
                                              my_ins) == UPSERT_FULL)
                            s_ovf = 1;
                    }
                    if (rw.i == rw.n) {  // record done: account inserts, start the prefetched record
                        flush();
                        r = rn;
                        rn = rec_of(++j + 1);
                        if (r >= r_end) break;
                        if (__hip_atomic_load(&s_ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
                        rw.init(at_next, lab_next, k);
                        m = m_next;
                        if (rn < r_end) {
                            Walker::fetch(a.att, a.labels, rn, W, at_next, lab_next);
                            if (a.mult) m_next = a.mult[rn];
                        }
                    }
                }
                if (all) {
                    // The walk filled every sub-list: pass p_lo starts over
                    // with a clear table (the next iteration of a loop over
                    // the same p would need one more flag; the table is
                    // cleared here instead).
                    __syncthreads();
                    if (s_ovf) break;
                    for (int i = tid; i < C; i += LWG) s_key[i] = EMPTY;
                    for (int i = tid; i < C; i += LWG) s_ef[i] = 0u;
                    for (int i = tid; i < C / 4; i += LWG) s_st[i] = 0u;
                    __syncthreads();
                }
            }
            if (all) {
                // Pass p over its own sub-list (a sub-list that overflowed
                // has flagged the bucket: nothing past its capacity is read).
                // Each lane takes one group-read step of its current search
                // per loop trip and, once the instance is added, goes on with
                // the entry it prefetched, so a wave's trips follow the
                // average search length, not its longest search.  The lanes
                // that consumed their prefetched entry take the next ones
                // from the sub-list together (one LDS atomic per wave, the
                // entries consecutive).  Counter adds commute, so the order
                // of the adds never shows.
                const uint32_t n_in = min(s_sn[p], c_sub);
                const uint32_t e0 = p * c_sub;
                // entries for the lanes with `want` (all lanes of the wave call this)
                const auto take = [&](bool want) -> uint32_t {
                    const unsigned long long mw = __ballot(want);
                    uint32_t base = 0;
                    if (lane == 0 && mw) base = atomicAdd(&s_next, uint32_t(__popcll(mw)));
                    base = uint32_t(__builtin_amdgcn_readfirstlane(int(base)));
                    return base + __builtin_amdgcn_mbcnt_hi(uint32_t(mw >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mw), 0u));
                };
                Kmer<KW> key_n, key;
                uint32_t meta_n = 0, meta = 0, g = 0, n_full = 0;
                bool pre = false, have = false;  // a lane with neither has no entries left to take
                const auto fetch = [&](uint32_t e) {  // prefetch entry e, if it exists
                    pre = e < n_in;
                    if (pre) spill_load<KW>(R, e0 + e, key_n, meta_n);
                };
                fetch(take(true));
                for (uint32_t trip = 1;; ++trip) {
                    bool want = false;
                    if (!have && pre) {  // start the prefetched instance, prefetch the next
                        key = key_n;
                        meta = meta_n;
                        g = (meta >> 13) & 0xFFFu;
                        n_full = 0;
                        have = true;
                        pre = false;
                        want = true;
                    }
                    const uint32_t e = take(want);
                    if (want) fetch(e);
                    if (!__ballot(have)) break;  // every entry handed out, and this wave's are added
                    if (have) {
                        const int r = upsert_step<KW, C>(s_key, s_ef, s_st, key, g, meta >> 25, n_full, meta & 7u,
                                                         (meta >> 3) & 7u, (meta >> 6) & 7u, (meta >> 9) & 15u,
                                                         my_ins);
                        if (r != STEP_AGAIN) have = false;
                        if (r == STEP_FULL) s_ovf = 1;
                    }
                    if ((trip & 7u) == 0) {
                        flush();
                        if (__hip_atomic_load(&s_ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
                    }
                }
                flush();
            }
            __syncthreads();
            if (s_ovf) {
                if (tid == 0) atomicOr(&a.overflow[b], 1u);
                break;
            }
            // Append the pass's vertices (its occupied slots, s_cnt of them)
            // after those of the bucket's other passes: the bucket's table is
            // its n_vert[b] vertices, dense.  Within the pass's run the
            // vertices in an edge at the cutoff (active) come first and the
            // isolated ones fill it from its end; each wave reserves its
            // entries of either kind with one LDS atomic.
            if (tid == 0) {
                s_base = atomicAdd(&a.n_vert[b], s_cnt);
                s_wai = 0;
                s_wp = 0;
            }
            __syncthreads();
            const uint32_t n_run = s_cnt;
            typedef __attribute__((address_space(1))) uint64_t g_u64;
            g_u64* tbl = reinterpret_cast<g_u64*>(a.tbl_ptr[b]) + uint64_t(s_base) * (KW + 1);
            for (int i = tid; i < C; i += LWG) {  // C % 64 == 0: a wave's slots are all below C or none
                const bool occ = s_key[i] != EMPTY;
                const uint32_t ef = s_ef[i];
                const uint32_t st = (s_st[i >> 2] >> (8 * (i & 3))) & 0xFFu;
                const bool act = occ && ef_active(ef, a.cutoff);
                const unsigned long long ma = __ballot(act), mi = __ballot(occ) & ~ma,
                                         mp = __ballot(occ && !act && (st & 6u));
                // one LDS atomic reserves the wave's active (low half) and isolated (high half) entries
                unsigned long long rr = 0;
                if (lane == 0) {
                    const unsigned long long inc =
                        (unsigned long long)__popcll(ma) | ((unsigned long long)__popcll(mi) << 32);
                    if (inc) rr = atomicAdd(&s_wai, inc);
                    if (mp) atomicAdd(&s_wp, uint32_t(__popcll(mp)));
                }
                const uint32_t ra = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(rr))));
                const uint32_t ri = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(rr >> 32))));
                if (!occ) continue;
                const unsigned long long mk = act ? ma : mi;
                const uint32_t r = __builtin_amdgcn_mbcnt_hi(uint32_t(mk >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mk), 0u));
                const uint32_t pos = act ? ra + r : n_run - 1 - (ri + r);
                g_u64* e = tbl + uint64_t(pos) * (KW + 1);
#pragma unroll
                for (int w = 0; w < KW - 1; ++w) e[w] = *low_word<KW, C>(s_key, uint32_t(i), w);
                e[KW - 1] = s_key[i];
                e[KW] = uint64_t(ef) | (uint64_t(st) << 32);
            }
            __syncthreads();  // table read out and counters used before the next pass clears them
            if (tid == 0) {
                const uint32_t n_act = uint32_t(s_wai), n_iso = uint32_t(s_wai >> 32);
                if (a.seg && P <= SEG_MAX) a.seg[uint64_t(b) * SEG_MAX + p] = uint64_t(s_base) | (uint64_t(n_act) << 32);
                if (a.iso) {
                    if (n_iso) atomicAdd(&a.iso[2 * b], n_iso);
                    if (s_wp) atomicAdd(&a.iso[2 * b + 1], s_wp);
                }
            }
        }
        if (s_ovf && tid == 0) atomicOr(&a.overflow[b], 1u);
    }
}

}  // namespace

void launch_construct_lds(const ConstructArgs& a, uint32_t KW, uint32_t grid, hipStream_t s) {
    if (!grid) return;
    // label words: W = ceil((2k - l) / 32) <= 2 * KW for l >= 1
    switch (KW) {
        case 1: hipLaunchKernelGGL((k_construct_lds<1, 2>), dim3(grid), dim3(LWG), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_construct_lds<2, 4>), dim3(grid), dim3(LWG), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_construct_lds<3, 6>), dim3(grid), dim3(LWG), 0, s, a); break;
        default: hipLaunchKernelGGL((k_construct_lds<4, 8>), dim3(grid), dim3(LWG), 0, s, a); break;
    }
}

}  // namespace cfgpu
