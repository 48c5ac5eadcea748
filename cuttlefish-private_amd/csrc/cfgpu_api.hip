// libcfgpu: C ABI (include/cfgpu.h) over the gfx950 partition / construct kernels.
//
// Orchestration only: input checks, device buffers, table sizing, work-queue
// construction, hipGraph capture and timing.  All per-base / per-k-mer work runs
// in the HIP kernels (partition.hip, construct.hip, mlp.hip).
#include "../../include/cfgpu.h"
#include "../../include/cfhost.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <memory>
#include <unistd.h>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "cf_kernels.hpp"

using namespace cfgpu;

namespace {

constexpr uint32_t G_FIXED = 16384;
constexpr uint64_t SAMPLE_TARGET = MLP_TRAIN_MAX + MLP_TRAIN_MAX / 3;  // 3/4 train, 1/4 held out
// The MLP is trained only when the constant mean ratio misses the held-out
// buckets by more than this (mean |log error|).
constexpr double CONST_RATIO_GOOD = 0.05;
constexpr uint32_t MAX_ROUNDS = 8;
// MLP weight initialisation seed (epochs and learning rate are tuning knobs)
constexpr uint32_t MLP_SEED = 12345u;

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t ensure(size_t bytes) {
        bytes = std::max<size_t>(bytes, 256);
        if (bytes <= n) return hipSuccess;
        release();
        const hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        n = bytes;
        return hipSuccess;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

void swap_buf(DevBuf& a, DevBuf& b) {
    std::swap(a.p, b.p);
    std::swap(a.n, b.n);
}

uint64_t wymix_host(uint64_t a, uint64_t b) {
    const unsigned __int128 r = static_cast<unsigned __int128>(a) * b;
    return uint64_t(r) ^ uint64_t(r >> 64);
}

// Table capacities are multiples of 64 slots (any size works: the kernels
// reduce hashes with a multiply-shift).
uint64_t round_cap(double slots) {
    const uint64_t c = uint64_t(std::ceil(std::max(64.0, slots)));
    return (c + 63) & ~uint64_t(63);
}

struct Event {
    hipEvent_t e = nullptr;
    Event() { (void)hipEventCreate(&e); }
    ~Event() { if (e) (void)hipEventDestroy(e); }
};

}  // namespace

// Scratch of cfgpu_table_copy_noalloc: the per-bucket output offsets, the
// export cursors and the bucket list, for at most NOALLOC_MAX_BUCKETS buckets.
// One array per GPU, held until the process ends and shared by every handle on
// that GPU; host code uses the device address hipGetSymbolAddress returns
// (noalloc_scratch in the handle), never the array's name.
constexpr uint32_t NOALLOC_MAX_BUCKETS = 16384;
constexpr size_t NOALLOC_OFF_BYTES = (size_t(NOALLOC_MAX_BUCKETS) + 1) * 8;
constexpr size_t NOALLOC_CURSOR_AT = (NOALLOC_OFF_BYTES + 255) & ~size_t(255);
constexpr size_t NOALLOC_LIST_AT = NOALLOC_CURSOR_AT + size_t(NOALLOC_MAX_BUCKETS) * 4;
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_LIST_AT + size_t(NOALLOC_MAX_BUCKETS) * 4;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

struct cfgpu_handle {
    cfgpu_params prm{};
    uint32_t KW = 1, W = 1, w = 1, G = G_FIXED, g_bits = 14;
    WyConst wy{};
    int cus = 256;
    hipStream_t stream = nullptr;
    std::string err;
    // target fill of an LDS pass table 0.8 (a table overflows at 7/8)
    cfgpu_tuning tune{0.6f, 256, 2048, 8, 1.15f, 1, LDS_ALL_PASS_MAX, 0, 1, 120, 3e-3f, 0.8f, 1};

    // partition state
    DevBuf packed, frag, gbin, mbn, tile_lc, tile_frag, boff, att, labels, mbin, stats, feat;
    DevBuf pdesc, tile_n;  // piece descriptors (P_TILE per tile) and per-tile counts
    DevBuf hist, ctot, coff, rec1;  // level-1 bucketing
    // Chunks partitioned with CFGPU_PARTITION_APPEND, waiting for cfgpu_partition_finish:
    // each chunk's buckets (records, minimizer bins, offsets) and its counters.
    struct PartChunk {
        DevBuf att, labels, mbin, boff;
        std::vector<uint64_t> h_boff;
        uint64_t n_rec = 0, n_bases = 0, n_frag = 0;
        double t_minimizer = 0, t_count = 0, t_emit = 0, t_features = 0, t_partition = 0;
    };
    std::vector<std::unique_ptr<PartChunk>> chunks;
    DevBuf ch_src, ch_dst;  // gather arguments
    // External bucket store (CFGPU_PARTITION_SPILL): per spilled chunk its
    // records in bucket order (headers | labels | minimizer bins) in pinned
    // host memory or in a file, and its bucket offsets; the per-bucket sums
    // of every chunk (statistics and features) on the device.
    struct SpillSeg {
        uint64_t n_rec = 0, n_bases = 0, n_frag = 0;
        std::vector<uint64_t> h_boff;
        uint8_t* host = nullptr;
        std::string path;
        double t_partition = 0, t_minimizer = 0, t_count = 0, t_emit = 0, t_features = 0;
    };
    std::vector<SpillSeg> spill;
    std::string spill_dir;
    uint64_t spill_serial = 0;     // file names
    DevBuf spill_sums;
    DevBuf rst_att, rst_lab, rst_mb, rst_boff;  // restore: staged ranges of the segments
    bool spilled = false;          // the buckets live in the store
    std::vector<uint8_t> resident; // spilled: buckets in device memory
    uint8_t* spill_stage = nullptr;  // pinned staging of file segments and restores
    size_t spill_stage_cap = 0;
    double t_spill_ms = 0, t_restore_ms = 0;
    uint64_t n_bases = 0, n_frag = 0, n_records = 0;
    uint64_t n_dev_records = 0;    // records in the device bucket arrays
    bool partitioned = false;
    std::vector<uint64_t> h_boff, h_stats;

    // construct state
    DevBuf d_tbl_ptr, d_cap, d_limit, d_nvert, d_overflow, d_items_b, d_items_f, d_items_c,
        d_qhead, d_round_list;
    DevBuf d_seg, d_iso, d_nseg;  // LDS-built tables: pass runs, isolated / phantom counts (contract reads)
    std::vector<std::unique_ptr<DevBuf>> pools;  // one per sizing round, reused across calls
    uint32_t round_idx = 0;
    DevBuf spill_key;  // LDS construct: per-workgroup spill regions of all-pass items
    uint64_t spill_need = 0;       // region entries the largest all-pass item of the round needs
    uint32_t n_all_pass = 0;       // all-pass items of the round
    uint8_t* stage = nullptr;      // pinned staging of small copies (h2d, d2h); reused after sync_stream
    size_t stage_cap = 0, stage_off = 0;
    struct PendingCopy {
        void* dst;
        const void* src;
        size_t n;
    };
    std::vector<PendingCopy> stage_pending;  // d2h copies finished on the host by sync_stream
    bool stage_direct = false;                // an h2d copy since the last sync read pageable memory
    std::vector<uint64_t> h_tbl_ptr;
    std::vector<uint64_t> h_cap;
    std::vector<uint8_t> h_dense;             // 1: built in LDS, the table is its n_vert vertices
    std::vector<uint32_t> h_passes;           // LDS-built tables: hash passes (runs of the table)
    std::vector<uint32_t> h_limit, h_nvert, h_flags;
    std::vector<float> h_pred;
    bool constructed = false;

    // record dedup: distinct records and multiplicities in each bucket's region
    DevBuf dd_att, dd_lab, dd_mult, dd_cnt, dd_km, dd_end, dd_ovf, dd_ib, dd_ip, dd_iP, dd_list, dd_region;
    bool dd_active = false;
    std::vector<uint32_t> h_ddcnt;
    std::vector<uint64_t> h_ddkm;

    // MLP: device state, and the training set and settings of the last run
    DevBuf mlp_params, mlp_adam, mlp_idx, mlp_target, mlp_pred, mlp_loss;
    std::vector<uint32_t> h_mlp_idx;
    std::vector<float> h_mlp_target;
    bool mlp_trained = false;
    float mlp_loss_last = 0.f, mlp_lr_last = 0.f;
    uint32_t mlp_epochs_last = 0;

    // export scratch
    DevBuf ex_keys, ex_ef, ex_st, ex_dfa, ex_off, ex_cursor, ex_list;

    // contract: per-subgraph counters, unitig records and labels
    DevBuf ct_list, ct_ovf, ct_ovf2, ct_ovf3, ct_counts, ct_ctl, ct_scratch, ct_recs, ct_labels;
    uint64_t n_unitigs = 0, n_ubytes = 0;
    uint64_t ct_need[2] = {0, 0};  // records and label bytes the last contract needed
    bool contracted = false;

    // compacted graph: glue table, list-ranking states, output; cycles through
    // several subgraphs finished on the host (appended after the device part)
    DevBuf cd_settled;
    DevBuf cd_partner, cd_bhist, cd_btot, cd_boff, cd_bkey, cd_bid, cd_kend, cd_jp, cd_head, cd_pos, cd_hoff, cd_out,
        cd_rec, cd_ctl, cd_cyc;
    uint64_t cd_n = 0, cd_bytes = 0;
    std::string cd_host_seq;
    std::vector<uint64_t> cd_host_rec;
    bool cdbg_done = false;

    // graph cache
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;

    cfgpu_run_stats rs{};

    // cfgpu_table_copy_noalloc: the device address of this GPU's `scratch`
    // array and host arrays sized at open (that call allocates nothing)
    uint8_t* noalloc_scratch = nullptr;
    std::vector<uint64_t> na_off;
    std::vector<uint32_t> na_list, na_cnt;

    ~cfgpu_handle() {
        for (SpillSeg& sg : spill) {
            if (sg.host) (void)hipHostFree(sg.host);
            if (!sg.path.empty()) (void)unlink(sg.path.c_str());
        }
        if (spill_stage) (void)hipHostFree(spill_stage);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Host-to-device copy of a small host array through a pinned staging buffer:
// a copy from pageable memory blocks the host for tens of microseconds while
// the GPU idles between launches.  Staged bytes are reused only after
// sync_stream (every copy has completed); an upload that does not fit is
// copied directly.  Not for stream capture (a captured copy would read the
// staging later).
constexpr size_t STAGE_BYTES = 8u << 20;
hipError_t h2d(cfgpu_handle* h, void* dst, const void* src, size_t n) {
    if (!h->stage && hipHostMalloc(reinterpret_cast<void**>(&h->stage), STAGE_BYTES, hipHostMallocDefault) == hipSuccess)
        h->stage_cap = STAGE_BYTES;
    if (h->stage && n <= h->stage_cap - h->stage_off) {
        uint8_t* p = h->stage + h->stage_off;
        std::memcpy(p, src, n);
        h->stage_off += (n + 255) & ~size_t(255);
        if (h->stage_off > h->stage_cap) h->stage_off = h->stage_cap;
        return hipMemcpyAsync(dst, p, n, hipMemcpyHostToDevice, h->stream);
    }
    h->stage_direct = true;
    return hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, h->stream);
}

// Device-to-host copy into a host array through the staging buffer; the
// array holds the data once sync_stream returns (every caller synchronises
// with it before reading).
hipError_t d2h(cfgpu_handle* h, void* dst, const void* src, size_t n) {
    if (!h->stage && hipHostMalloc(reinterpret_cast<void**>(&h->stage), STAGE_BYTES, hipHostMallocDefault) == hipSuccess)
        h->stage_cap = STAGE_BYTES;
    if (h->stage && n <= h->stage_cap - h->stage_off) {
        uint8_t* p = h->stage + h->stage_off;
        h->stage_off += (n + 255) & ~size_t(255);
        if (h->stage_off > h->stage_cap) h->stage_off = h->stage_cap;
        h->stage_pending.push_back({dst, p, n});
        return hipMemcpyAsync(p, src, n, hipMemcpyDeviceToHost, h->stream);
    }
    return hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, h->stream);
}

hipError_t sync_stream(cfgpu_handle* h) {
    const hipError_t e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) {
        for (const auto& c : h->stage_pending) std::memcpy(c.dst, c.src, c.n);
        h->stage_pending.clear();
        h->stage_off = 0;
        h->stage_direct = false;
    }
    return e;
}

}  // namespace

namespace {

// Memory pressure: an allocation that runs out of device memory releases
// the buffers no later step of the current call reads and is tried once
// more.  Partition scratch is dead once the records are bucketed (the next
// cfgpu_partition allocates it again), pools past the current construct's
// rounds hold no live table, and each later step makes the scratch of the
// earlier ones dead.
// Phase::Partition is the gather of appended chunks, after which every table,
// and every later step's scratch, is dead too.  Phase::Input is a partition
// of new input: every table, every later step's scratch and the unitigs of
// the previous input are dead, and only the partition's own scratch and the
// appended chunks' buckets are live.
enum class Phase { Input, Partition, Construct, Contract, Cdbg };

void trim(cfgpu_handle* h, Phase ph) {
    if (ph == Phase::Input) {
        for (auto& pool : h->pools) pool->release();
        for (DevBuf* b : {&h->ex_keys, &h->ex_ef, &h->ex_st, &h->ex_dfa, &h->spill_key, &h->dd_att,
                          &h->dd_lab, &h->dd_mult, &h->dd_region, &h->ct_scratch, &h->ct_recs, &h->ct_labels,
                          &h->cd_settled, &h->cd_partner, &h->cd_bhist, &h->cd_bkey, &h->cd_bid, &h->cd_kend,
                          &h->cd_jp, &h->cd_head, &h->cd_pos, &h->cd_hoff, &h->cd_out, &h->cd_rec})
            b->release();
        return;
    }
    for (DevBuf* b : {&h->gbin, &h->mbn, &h->tile_lc, &h->pdesc, &h->tile_n, &h->tile_frag, &h->hist, &h->rec1,
                      &h->ex_keys, &h->ex_ef, &h->ex_st, &h->ex_dfa})
        b->release();
    for (size_t i = ph == Phase::Partition ? 0 : h->round_idx; i < h->pools.size(); ++i) h->pools[i]->release();
    if (ph == Phase::Partition) h->ct_scratch.release();
    if (ph != Phase::Cdbg)
        for (DevBuf* b : {&h->cd_settled, &h->cd_partner, &h->cd_bhist, &h->cd_bkey, &h->cd_bid, &h->cd_kend, &h->cd_jp, &h->cd_head,
                          &h->cd_pos, &h->cd_hoff, &h->cd_out, &h->cd_rec})
            b->release();
    if (ph != Phase::Construct)
        for (DevBuf* b : {&h->spill_key, &h->dd_att, &h->dd_lab, &h->dd_mult, &h->dd_region}) b->release();
    if (ph == Phase::Cdbg) h->ct_scratch.release();
}

hipError_t ensure_mem(cfgpu_handle* h, DevBuf& b, size_t bytes, Phase ph) {
    hipError_t e = b.ensure(bytes);
    if (e != hipErrorOutOfMemory) return e;
    (void)hipGetLastError();  // the failed allocation is handled here
    e = sync_stream(h);       // nothing queued may still use what is released
    if (e != hipSuccess) return e;
    trim(h, ph);
    e = b.ensure(bytes);
    if (e == hipErrorOutOfMemory) (void)hipGetLastError();
    return e;
}

}  // namespace

namespace {

int fail(cfgpu_handle* h, int code, const char* fmt, ...) {
    if (h) {
        h->stage_pending.clear();  // staged copies whose host arrays the failing call leaves
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        h->err = buf;
    }
    return code;
}

#define HIPCK(h, x)                                                                         \
    do {                                                                                    \
        const hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess)                                                               \
            return fail((h), e_ == hipErrorOutOfMemory ? CFGPU_E_NOMEM : CFGPU_E_HIP,       \
                        "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

#define CK(x)                    \
    do {                         \
        const int r_ = (x);      \
        if (r_ != CFGPU_OK) return r_; \
    } while (0)

double elapsed_ms(const Event& a, const Event& b) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a.e, b.e);
    return ms;
}

ConstructArgs construct_args(cfgpu_handle* h) {
    ConstructArgs a{};
    a.att = h->att.as<uint32_t>();
    a.labels = h->labels.as<uint64_t>();
    a.W = h->W;
    a.k = h->prm.k;
    a.bucket_off = h->boff.as<uint64_t>();
    a.tbl_ptr = h->d_tbl_ptr.as<uint64_t>();
    a.tbl_cap = h->d_cap.as<uint64_t>();
    a.tbl_limit = h->d_limit.as<uint32_t>();
    a.n_vert = h->d_nvert.as<uint32_t>();
    a.overflow = h->d_overflow.as<uint32_t>();
    a.item_bucket = h->d_items_b.as<uint32_t>();
    a.item_first = h->d_items_f.as<uint32_t>();
    a.item_count = h->d_items_c.as<uint32_t>();
    a.queue_head = h->d_qhead.as<uint32_t>();
    a.cutoff = h->prm.cutoff;
    a.seg = h->d_seg.as<uint64_t>();
    a.iso = h->d_iso.as<uint32_t>();
    if (h->dd_active) {
        a.att = h->dd_att.as<uint32_t>();
        a.labels = h->dd_lab.as<uint64_t>();
        a.mult = h->dd_mult.as<uint32_t>();
        a.bucket_end = h->dd_end.as<uint64_t>();
    }
    return a;
}

uint64_t bucket_records(const cfgpu_handle* h, uint32_t b) { return h->h_stats[4 * b + 0]; }
// Records the construct kernels walk: the distinct ones after dedup.
uint64_t walked_records(const cfgpu_handle* h, uint32_t b) {
    return h->dd_active ? h->h_ddcnt[b] : h->h_stats[4 * b + 0];
}
uint64_t bucket_kmers(const cfgpu_handle* h, uint32_t b) { return h->h_stats[4 * b + 1]; }
// k-mer instances the construct kernels walk (of the distinct records after dedup).
uint64_t walked_kmers(const cfgpu_handle* h, uint32_t b) {
    return h->dd_active ? h->h_ddkm[b] : bucket_kmers(h, b);
}

// An LDS bucket of 2 <= P <= LDS_ALL_PASS_MAX passes is one all-pass item
// (records walked once, every instance spilled to its pass's sub-list) when
// its sub-lists fit a per-workgroup region of at most SPILL_BYTES_MAX over all
// workgroups; otherwise one item per pass, each walking the records.
constexpr uint64_t SPILL_BYTES_MAX = 16ull << 30;
uint64_t spill_cap_max(const cfgpu_handle* h) {
    // entries are addressed with 32-bit word offsets within a region
    return std::min<uint64_t>(SPILL_BYTES_MAX / (uint64_t(std::max(h->cus, 1)) * (8 * h->KW + 4)),
                              0xFFFFFFFFull / (2 * h->KW + 1));
}
// Region entries an all-pass item of `inst` walked instances over P passes
// needs: P sub-lists with 50 % headroom over the mean plus 8 standard
// deviations of a binomial count (instances of one k-mer share a pass, so
// counts spread more than independent draws; a sub-list that still fills
// flags the bucket, which is then rebuilt exactly).
uint64_t spill_need_for(uint64_t inst, uint32_t P) {
    const double mean = double(inst) / double(P);
    const uint64_t c = uint64_t(std::ceil(1.5 * mean + 8.0 * std::sqrt(mean) + 64.0));
    return uint64_t(P) * c;
}

// Capacity that holds every k-mer instance of the bucket at the target load:
// such a table never overflows by load (a probe chain past MAX_PROBES is
// still detected and rebuilt at twice the size).
uint64_t upper_cap(const cfgpu_handle* h, uint32_t b) {
    return round_cap(double(std::max<uint64_t>(bucket_kmers(h, b), 8)) / h->tune.load_factor);
}

uint32_t limit_for(uint64_t cap) { return uint32_t(std::min<uint64_t>((cap * 7) / 8, 0xFFFFFFFFull)); }

// One sizing round: allocate a pool for `buckets` with capacities `caps`,
// upload pointers / limits, zero their counters, clear their tables and build
// the work-item list in the given order.  Returns the number of work items.
int prepare_round(cfgpu_handle* h, const std::vector<uint32_t>& buckets,
                  const std::vector<uint64_t>& caps, uint32_t& n_items, uint64_t& n_slots_out,
                  DevBuf*& pool_out, uint32_t chunk, uint32_t lds_cap = 0) {
    const uint32_t eb = entry_bytes(h->KW);
    uint64_t n_slots = 0;
    for (uint64_t c : caps) n_slots += c;
    if (h->pools.size() <= h->round_idx) h->pools.emplace_back(new DevBuf());
    DevBuf* pool = h->pools[h->round_idx++].get();
    // a pool much larger than this round needs (left by another mode) is
    // given back first, so that the rounds of one call fit together
    if (pool->n > 2 * n_slots * eb + (size_t(1) << 30)) pool->release();
    HIPCK(h, ensure_mem(h, *pool, n_slots * eb, Phase::Construct));
    uint64_t off = 0;
    std::vector<uint32_t> ib, ifirst, icnt;
    ib.reserve(buckets.size());
    ifirst.reserve(buckets.size());
    icnt.reserve(buckets.size());
    h->spill_need = 0;
    h->n_all_pass = 0;
    const uint64_t spill_max = spill_cap_max(h);
    for (size_t i = 0; i < buckets.size(); ++i) {
        const uint32_t b = buckets[i];
        h->h_tbl_ptr[b] = reinterpret_cast<uint64_t>(pool->as<uint8_t>() + off * eb);
        h->h_cap[b] = caps[i];
        h->h_dense[b] = lds_cap ? 1 : 0;
        h->h_passes[b] = lds_cap ? uint32_t(caps[i] / lds_cap) : 0u;
        h->h_limit[b] = limit_for(caps[i]);
        off += caps[i];
        if (lds_cap) {
            const uint32_t P = uint32_t(caps[i] / lds_cap);
            const uint64_t need = spill_need_for(walked_kmers(h, b), std::max(P, 1u));
            if (P >= 2 && P <= LDS_ALL_PASS_MAX && need <= spill_max) {  // one item: every pass, records walked once
                ib.push_back(b), ifirst.push_back(LDS_ALL_PASSES), icnt.push_back(P);
                h->spill_need = std::max(h->spill_need, need);
                ++h->n_all_pass;
                continue;
            }
            // one item per hash pass: (bucket, pass, passes)
            for (uint32_t q = 0; q < P; ++q) ib.push_back(b), ifirst.push_back(q), icnt.push_back(P);
            continue;
        }
        const uint64_t nr = walked_records(h, b);
        for (uint64_t f = 0; f < nr; f += chunk) {
            ib.push_back(b);
            ifirst.push_back(uint32_t(f));
            icnt.push_back(uint32_t(std::min<uint64_t>(chunk, nr - f)));
        }
    }
    n_items = uint32_t(ib.size());
    HIPCK(h, h->d_items_b.ensure(ib.size() * 4));
    HIPCK(h, h->d_items_f.ensure(ib.size() * 4));
    HIPCK(h, h->d_items_c.ensure(ib.size() * 4));
    HIPCK(h, h->d_round_list.ensure(buckets.size() * 4));
    const size_t G = h->G;
    HIPCK(h, h2d(h, h->d_tbl_ptr.p, h->h_tbl_ptr.data(), G * 8));
    HIPCK(h, h2d(h, h->d_cap.p, h->h_cap.data(), G * 8));
    HIPCK(h, h2d(h, h->d_limit.p, h->h_limit.data(), G * 4));
    if (n_items) {
        HIPCK(h, h2d(h, h->d_items_b.p, ib.data(), ib.size() * 4));
        HIPCK(h, h2d(h, h->d_items_f.p, ifirst.data(), ib.size() * 4));
        HIPCK(h, h2d(h, h->d_items_c.p, icnt.data(), ib.size() * 4));
    }
    if (!buckets.empty())
        HIPCK(h, h2d(h, h->d_round_list.p, buckets.data(), buckets.size() * 4));
    // the host vectors above are reused: wait only if one was copied directly
    // (staged copies took their bytes already)
    if (h->stage_direct) HIPCK(h, sync_stream(h));
    n_slots_out = n_slots;
    pool_out = pool;
    return CFGPU_OK;
}

// Zero the per-bucket counters of a round and clear its tables (kernels, so
// the sequence can be captured in a graph).
void enqueue_round_reset(cfgpu_handle* h, uint32_t n_buckets, DevBuf* pool, uint64_t n_slots) {
    launch_zero_buckets(h->d_round_list.as<uint32_t>(), n_buckets, h->d_nvert.as<uint32_t>(),
                        h->d_overflow.as<uint32_t>(), h->d_iso.as<uint32_t>(), h->stream);
    launch_clear_tables(pool->as<uint8_t>(), n_slots, h->KW, h->stream);
}

int finish_round(cfgpu_handle* h, const std::vector<uint32_t>& buckets,
                 std::vector<uint32_t>& overflowed) {
    const size_t G = h->G;
    std::vector<uint32_t> ovf(G);
    HIPCK(h, d2h(h, ovf.data(), h->d_overflow.p, G * 4));
    HIPCK(h, d2h(h, h->h_nvert.data(), h->d_nvert.p, G * 4));
    HIPCK(h, sync_stream(h));
    overflowed.clear();
    for (uint32_t b : buckets) {
        if (ovf[b]) {
            overflowed.push_back(b);
            h->h_flags[b] |= CFGPU_BUCKET_OVERFLOWED;
            h->h_flags[b] &= ~uint32_t(CFGPU_BUCKET_BUILT);
        } else {
            h->h_flags[b] |= CFGPU_BUCKET_BUILT;
        }
    }
    return CFGPU_OK;
}

// Plain path: one launch per bucket, in bucket order.
int run_round_plain(cfgpu_handle* h, const std::vector<uint32_t>& buckets,
                    const std::vector<uint64_t>& caps, std::vector<uint32_t>& overflowed) {
    // one workgroup per LDS tile of records, so each bucket's launch spreads
    // over as many CUs as the bucket has tiles
    const uint32_t chunk = h->tune.workgroup;
    uint32_t n_items = 0;
    uint64_t n_slots = 0;
    DevBuf* pool = nullptr;
    CK(prepare_round(h, buckets, caps, n_items, n_slots, pool, chunk));
    enqueue_round_reset(h, uint32_t(buckets.size()), pool, n_slots);
    const ConstructArgs a = construct_args(h);
    uint32_t item = 0;
    for (uint32_t b : buckets) {
        const uint32_t n = uint32_t((bucket_records(h, b) + chunk - 1) / chunk);
        ConstructArgs ab = a;
        ab.n_items = n_items;
        launch_construct(ab, h->KW, item, n, h->tune.workgroup, h->stream);
        item += n;
    }
    HIPCK(h, hipGetLastError());
    return finish_round(h, buckets, overflowed);
}

// Batched path: reset + persistent work-queue launch, captured in a hipGraph.
// Batched path: one persistent work-queue launch for the round, optionally
// captured in a hipGraph.  lds = true builds each bucket in LDS (caps are
// passes * lds_capacity, one work item per pass); otherwise record chunks
// insert into HBM tables with global atomics.
int run_round_batched(cfgpu_handle* h, const std::vector<uint32_t>& buckets,
                      const std::vector<uint64_t>& caps, std::vector<uint32_t>& overflowed,
                      bool use_graph, bool lds = false) {
    const uint32_t chunk = std::max<uint32_t>(h->tune.workgroup, h->tune.chunk_records);
    const uint32_t C = uint32_t(lds_capacity(int(h->KW)));
    uint32_t n_items = 0;
    uint64_t n_slots = 0;
    DevBuf* pool = nullptr;
    CK(prepare_round(h, buckets, caps, n_items, n_slots, pool, chunk, lds ? C : 0));
    ConstructArgs a = construct_args(h);
    a.n_items = n_items;
    const uint32_t per_cu = lds ? 1u : h->tune.wgs_per_cu;  // the LDS kernel fills a CU's LDS
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(n_items, uint32_t(h->cus) * per_cu));
    if (lds && h->spill_need) {  // one region per workgroup, sized for the largest all-pass item
        const uint64_t cap = (h->spill_need + 63) & ~63ull;
        HIPCK(h, ensure_mem(h, h->spill_key, uint64_t(grid) * cap * (h->KW * 8 + 4), Phase::Construct));
        a.spill_key = h->spill_key.as<uint64_t>();
        a.spill_cap = uint32_t(cap);
    }
    auto reset = [&]() {
        launch_fill_u32(h->d_qhead.as<uint32_t>(), 0u, 1, h->stream);
        launch_zero_buckets(h->d_round_list.as<uint32_t>(), uint32_t(buckets.size()), h->d_nvert.as<uint32_t>(),
                            h->d_overflow.as<uint32_t>(), h->d_iso.as<uint32_t>(), h->stream);
        if (!lds) launch_clear_tables(pool->as<uint8_t>(), n_slots, h->KW, h->stream);  // LDS passes write every slot
    };
    auto kernel = [&]() {
        if (lds) launch_construct_lds(a, h->KW, grid, h->stream);
        else launch_construct_persistent(a, h->KW, grid, h->tune.workgroup, h->stream);
    };
    if (!use_graph || !h->tune.use_graph) {
        // direct launches; the construct kernel alone is timed with events
        Event k0, k1;
        reset();
        HIPCK(h, hipEventRecord(k0.e, h->stream));
        kernel();
        HIPCK(h, hipEventRecord(k1.e, h->stream));
        HIPCK(h, hipGetLastError());
        const int r = finish_round(h, buckets, overflowed);
        if (use_graph) h->rs.t_construct_kernel_ms += elapsed_ms(k0, k1);
        return r;
    }
    // The graph captures kernel arguments (device pointers and sizes); array
    // contents are read when it runs.  Re-capture when any captured value moves.
    const void* ptrs[] = {pool->p, a.att, a.labels, a.mult, a.bucket_end, h->boff.p, h->d_tbl_ptr.p, h->d_cap.p,
                          h->d_limit.p, h->d_nvert.p, h->d_overflow.p, h->d_items_b.p,
                          h->d_items_f.p, h->d_items_c.p, h->d_qhead.p, h->d_round_list.p,
                          a.spill_key, a.seg, a.iso};
    std::vector<uint64_t> key = {n_slots, n_items, grid, h->KW, h->W, h->prm.k, h->tune.workgroup,
                                 buckets.size(), uint64_t(lds), a.spill_cap};
    for (const void* q : ptrs) key.push_back(reinterpret_cast<uint64_t>(q));
    if (!h->graph_exec || key != h->graph_key) {
        if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec), h->graph_exec = nullptr;
        if (h->graph) (void)hipGraphDestroy(h->graph), h->graph = nullptr;
        HIPCK(h, sync_stream(h));  // capture starts on an idle stream
        HIPCK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        reset();
        kernel();
        HIPCK(h, hipStreamEndCapture(h->stream, &h->graph));
        HIPCK(h, hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
        h->graph_key = key;
    }
    HIPCK(h, hipGraphLaunch(h->graph_exec, h->stream));
    return finish_round(h, buckets, overflowed);
}

// Split a set of buckets (with distinct-count estimates) between the LDS
// kernel (estimated passes <= max_passes) and the global-atomic kernel, run
// both, and collect the overflowed buckets.
int run_main(cfgpu_handle* h, const std::vector<uint32_t>& order, const std::vector<double>& est,
             double margin, std::vector<uint32_t>& overflowed) {
    const double C = double(lds_capacity(int(h->KW)));
    std::vector<uint32_t> lds_b, glob_b;
    std::vector<uint64_t> lds_c, glob_c;
    lds_b.reserve(order.size());
    lds_c.reserve(order.size());
    for (uint32_t b : order) {
        const double need = std::max(1.0, est[b] * margin);
        const uint32_t P = uint32_t(std::ceil(need / (C * double(h->tune.lds_load))));
        if (P <= h->tune.max_passes) {
            lds_b.push_back(b);
            lds_c.push_back(uint64_t(P) * uint64_t(C));
        } else {
            glob_b.push_back(b);
            glob_c.push_back(std::min(upper_cap(h, b), round_cap(need / h->tune.load_factor)));
        }
    }
    std::vector<uint32_t> o1, o2;
    if (!glob_b.empty()) CK(run_round_batched(h, glob_b, glob_c, o1, !lds_b.empty() ? false : true));
    if (!lds_b.empty()) CK(run_round_batched(h, lds_b, lds_c, o2, true, true));
    overflowed = o1;
    overflowed.insert(overflowed.end(), o2.begin(), o2.end());
    h->rs.n_lds_buckets = uint32_t(lds_b.size());
    h->rs.n_lds_all_pass = lds_b.empty() ? 0u : h->n_all_pass;  // the LDS round was the last prepared
    return CFGPU_OK;
}

// Rebuild overflowed buckets with never-overflowing capacities until done.
int fallback(cfgpu_handle* h, std::vector<uint32_t> todo, bool plain_growth) {
    for (uint32_t round = 0; !todo.empty(); ++round) {
        if (round >= MAX_ROUNDS) return fail(h, CFGPU_E_OVERFLOW, "table overflow persists after %u rounds", round);
        ++h->rs.n_fallback_rounds;
        std::vector<uint64_t> l2(todo.size());
        for (size_t i = 0; i < todo.size(); ++i) {
            const uint32_t b = todo[i];
            // plain: grow 4x (bounded by the never-overflowing size); otherwise
            // that size, doubled if the last attempt already had it
            const uint64_t ub = upper_cap(h, b);
            l2[i] = plain_growth ? std::min(ub, round_cap(4.0 * double(h->h_cap[b])))
                                 : (h->h_cap[b] >= ub ? round_cap(2.0 * double(h->h_cap[b])) : ub);
        }
        std::vector<uint32_t> ovf;
        if (plain_growth) CK(run_round_plain(h, todo, l2, ovf));
        else CK(run_round_batched(h, todo, l2, ovf, false));
        todo.swap(ovf);
    }
    return CFGPU_OK;
}

std::vector<uint32_t> nonempty_buckets(const cfgpu_handle* h) {
    std::vector<uint32_t> v;
    for (uint32_t b = 0; b < h->G; ++b)
        if (bucket_records(h, b)) v.push_back(b);
    return v;
}

// Largest cost first, ties in the given order: a stable LSD radix sort (3
// passes of 11 bits) on the inverted bits of the cost as a float (costs are
// >= 0, so the bit patterns order like the values).  A comparison sort of the
// 16384 buckets costs about a millisecond of host time between launches.
void order_largest_first(std::vector<uint32_t>& v, const std::vector<double>& cost) {
    const size_t n = v.size();
    std::vector<uint32_t> key(n), key2(n), v2(n);
    for (size_t i = 0; i < n; ++i) {
        const float f = float(std::max(0.0, cost[v[i]]));
        uint32_t bits;
        std::memcpy(&bits, &f, 4);
        key[i] = ~bits;
    }
    std::vector<uint32_t> cnt(2049);
    for (int shift = 0; shift < 33; shift += 11) {
        std::fill(cnt.begin(), cnt.end(), 0u);
        for (size_t i = 0; i < n; ++i) ++cnt[((key[i] >> shift) & 2047u) + 1];
        for (size_t d = 1; d < cnt.size(); ++d) cnt[d] += cnt[d - 1];
        for (size_t i = 0; i < n; ++i) {
            const uint32_t pos = cnt[(key[i] >> shift) & 2047u]++;
            key2[pos] = key[i];
            v2[pos] = v[i];
        }
        key.swap(key2);
        v.swap(v2);
    }
}

// Record dedup before the batched constructs.  A probe set of evenly spread
// buckets is deduplicated with passes sized for all-distinct records; when
// its records repeat less than DEDUP_MIN_GAIN times (auto mode) the
// construct walks the original records.  Otherwise the other buckets get
// passes sized from the measured distinct fraction, and buckets whose pass
// tables overflowed keep their records unchanged (copied, multiplicity 1).
// The deduplicated records also come out longest first within a bucket,
// which evens the lanes of the construct's walk: at 25x reads with k = 63
// (1.22 repeats per record) the pass costs 15 ms and saves 33 ms of walk.
constexpr double DEDUP_LOAD = 0.85;
constexpr double DEDUP_MIN_GAIN = 1.1;
constexpr size_t DEDUP_PROBES = 64;

// all_pass: buckets of 2 .. DD_ALL_PASS_MAX passes are one work item each (the
// probe round, a few buckets, keeps one item per pass: more workgroups busy).
int dedup_round(cfgpu_handle* h, const std::vector<uint32_t>& buckets, double distinct_frac, bool all_pass) {
    const uint32_t max_len = std::min<uint32_t>(255, 2 * h->prm.k - h->prm.l);  // record length bound
    const bool packed = dedup_packed(h->W, max_len);
    const double C = double(dedup_capacity(dedup_key_words(h->W, max_len))) * DEDUP_LOAD;
    std::vector<uint32_t> ib, ip, iP;
    bool any_all = false;
    for (uint32_t b : buckets) {
        const uint64_t nr = bucket_records(h, b);
        const double need = std::max(1.0, double(nr) * distinct_frac);
        const uint32_t P = uint32_t(std::min(4096.0, std::ceil(need / C)));
        if (all_pass && P >= 2 && P <= DD_ALL_PASS_MAX && nr <= DD_REGION_RECORDS) {  // every pass in one workgroup
            ib.push_back(b), ip.push_back(DD_ALL_PASSES), iP.push_back(P);
            any_all = true;
            continue;
        }
        for (uint32_t q = 0; q < P; ++q) ib.push_back(b), ip.push_back(q), iP.push_back(P);
    }
    if (ib.empty()) return CFGPU_OK;
    HIPCK(h, h->dd_ib.ensure(ib.size() * 4));
    HIPCK(h, h->dd_ip.ensure(ib.size() * 4));
    HIPCK(h, h->dd_iP.ensure(ib.size() * 4));
    HIPCK(h, h2d(h, h->dd_ib.p, ib.data(), ib.size() * 4));
    HIPCK(h, h2d(h, h->dd_ip.p, ip.data(), ip.size() * 4));
    HIPCK(h, h2d(h, h->dd_iP.p, iP.data(), iP.size() * 4));
    launch_fill_u32(h->d_qhead.as<uint32_t>(), 0u, 1, h->stream);
    DedupArgs a{};
    a.att = h->att.as<uint32_t>();
    a.labels = h->labels.as<uint64_t>();
    a.bucket_off = h->boff.as<uint64_t>();
    a.overflow = h->dd_ovf.as<uint32_t>();
    a.out_count = h->dd_cnt.as<uint32_t>();
    a.out_kmers = h->dd_km.as<unsigned long long>();
    a.k = h->prm.k;
    a.out_att = h->dd_att.as<uint32_t>();
    a.out_labels = h->dd_lab.as<uint64_t>();
    a.out_mult = h->dd_mult.as<uint32_t>();
    a.item_bucket = h->dd_ib.as<uint32_t>();
    a.item_pass = h->dd_ip.as<uint32_t>();
    a.item_passes = h->dd_iP.as<uint32_t>();
    a.n_items = uint32_t(ib.size());
    a.queue_head = h->d_qhead.as<uint32_t>();
    const uint32_t grid = std::min<uint32_t>(a.n_items, uint32_t(h->cus));
    if (any_all) {
        HIPCK(h, ensure_mem(h, h->dd_region, size_t(grid) * DD_REGION_RECORDS * 5, Phase::Construct));
        a.region = h->dd_region.as<uint8_t>();
    }
    launch_dedup(a, h->W, packed, grid, h->stream);
    HIPCK(h, hipGetLastError());
    return CFGPU_OK;
}

int dedup_buckets(cfgpu_handle* h, const std::vector<uint32_t>& all) {
    h->dd_active = false;
    cfgpu_run_stats& rs = h->rs;
    rs.n_records_walked = h->n_records;
    rs.t_dedup_ms = 0;
    if (!h->tune.dedup || all.empty()) return CFGPU_OK;
    const size_t G = h->G;
    const uint64_t n = std::max<uint64_t>(h->n_dev_records, 1);
    HIPCK(h, ensure_mem(h, h->dd_att, n * 4, Phase::Construct));
    HIPCK(h, ensure_mem(h, h->dd_lab, n * h->W * 8, Phase::Construct));
    HIPCK(h, ensure_mem(h, h->dd_mult, n * 4, Phase::Construct));
    HIPCK(h, h->dd_cnt.ensure(G * 4));
    HIPCK(h, h->dd_km.ensure(G * 8));
    HIPCK(h, h->dd_end.ensure(G * 8));
    HIPCK(h, h->dd_ovf.ensure(G * 4));
    HIPCK(h, h->dd_list.ensure(G * 4));
    Event e0, e1;
    HIPCK(h, hipEventRecord(e0.e, h->stream));
    launch_fill_u32(h->dd_cnt.as<uint32_t>(), 0u, G, h->stream);
    launch_fill_u32(h->dd_km.as<uint32_t>(), 0u, 2 * G, h->stream);
    launch_fill_u32(h->dd_ovf.as<uint32_t>(), 0u, G, h->stream);

    std::vector<uint32_t> probe, rest;
    if (all.size() >= 4 * DEDUP_PROBES) {
        const double step = double(all.size()) / double(DEDUP_PROBES);
        std::vector<char> in(all.size(), 0);
        for (size_t i = 0; i < DEDUP_PROBES; ++i) in[size_t(i * step)] = 1;
        for (size_t i = 0; i < all.size(); ++i) (in[i] ? probe : rest).push_back(all[i]);
    } else {
        probe = all;
    }
    CK(dedup_round(h, probe, 1.0, false));
    std::vector<uint32_t> cnt(G), ovf(G);
    HIPCK(h, d2h(h, cnt.data(), h->dd_cnt.p, G * 4));
    HIPCK(h, d2h(h, ovf.data(), h->dd_ovf.p, G * 4));
    HIPCK(h, sync_stream(h));
    double d = 0, t = 0;
    for (uint32_t b : probe)
        if (!ovf[b]) d += cnt[b], t += double(bucket_records(h, b));
    const double frac = t > 0 ? d / t : 1.0;
    if (h->tune.dedup == 1 && frac * DEDUP_MIN_GAIN > 1.0) {
        HIPCK(h, hipEventRecord(e1.e, h->stream));
        HIPCK(h, sync_stream(h));
        rs.t_dedup_ms = elapsed_ms(e0, e1);
        return CFGPU_OK;  // records barely repeat: walk them as they are
    }
    // margin < 1 under-sizes on purpose here too (exercises the copy fallback)
    CK(dedup_round(h, rest, std::min(1.0, frac * 1.1) * std::min(1.0, double(h->tune.margin)), true));
    HIPCK(h, d2h(h, ovf.data(), h->dd_ovf.p, G * 4));
    HIPCK(h, sync_stream(h));
    std::vector<uint32_t> list;
    for (uint32_t b : all)
        if (ovf[b]) list.push_back(b);
    if (!list.empty()) {
        HIPCK(h, h2d(h, h->dd_list.p, list.data(), list.size() * 4));
        DedupArgs a{};
        a.att = h->att.as<uint32_t>();
        a.labels = h->labels.as<uint64_t>();
        a.bucket_off = h->boff.as<uint64_t>();
        a.out_count = h->dd_cnt.as<uint32_t>();
        a.out_kmers = h->dd_km.as<unsigned long long>();
        a.k = h->prm.k;
        a.out_att = h->dd_att.as<uint32_t>();
        a.out_labels = h->dd_lab.as<uint64_t>();
        a.out_mult = h->dd_mult.as<uint32_t>();
        launch_dedup_copy(a, h->dd_list.as<uint32_t>(), uint32_t(list.size()), h->W, h->stream);
    }
    launch_dedup_end(h->boff.as<uint64_t>(), h->dd_cnt.as<uint32_t>(), uint32_t(G), h->dd_end.as<uint64_t>(),
                     h->stream);
    h->h_ddcnt.resize(G);
    h->h_ddkm.resize(G);
    HIPCK(h, d2h(h, h->h_ddcnt.data(), h->dd_cnt.p, G * 4));
    HIPCK(h, d2h(h, h->h_ddkm.data(), h->dd_km.p, G * 8));
    HIPCK(h, hipEventRecord(e1.e, h->stream));
    HIPCK(h, sync_stream(h));
    HIPCK(h, hipGetLastError());
    uint64_t walked = 0;
    for (uint32_t b = 0; b < G; ++b) walked += h->h_ddcnt[b];
    rs.n_records_walked = walked;
    rs.n_dedup_fallback = uint32_t(list.size());
    rs.t_dedup_ms = elapsed_ms(e0, e1);
    h->dd_active = true;
    return CFGPU_OK;
}

// Construct the listed buckets (all non-empty ones when subset is null);
// every other bucket is left without a table.
int construct_impl(cfgpu_handle* h, uint32_t mode, const uint32_t* subset = nullptr, uint32_t n_subset = 0) {
    const size_t G = h->G;
    h->round_idx = 0;
    h->h_tbl_ptr.assign(G, 0);
    h->h_cap.assign(G, 0);
    h->h_dense.assign(G, 0);
    h->h_passes.assign(G, 0);
    h->h_limit.assign(G, 0);
    h->h_nvert.assign(G, 0);
    h->h_flags.assign(G, 0);
    h->h_pred.assign(G, 0.f);
    HIPCK(h, h->d_tbl_ptr.ensure(G * 8));
    HIPCK(h, h->d_cap.ensure(G * 8));
    HIPCK(h, h->d_limit.ensure(G * 4));
    HIPCK(h, h->d_nvert.ensure(G * 4));
    HIPCK(h, h->d_overflow.ensure(G * 4));
    HIPCK(h, h->d_qhead.ensure(4));
    HIPCK(h, h->d_seg.ensure(G * SEG_MAX * 8));
    HIPCK(h, h->d_iso.ensure(G * 8));
    launch_fill_u32(h->d_nvert.as<uint32_t>(), 0u, G, h->stream);
    launch_fill_u32(h->d_overflow.as<uint32_t>(), 0u, G, h->stream);

    cfgpu_run_stats& rs = h->rs;
    rs.n_overflowed = rs.n_fallback_rounds = rs.n_samples = 0;
    rs.t_sample_ms = rs.t_train_ms = rs.t_infer_ms = rs.t_construct_main_ms = rs.t_fallback_ms = 0;
    rs.mlp_loss = rs.mispredict_mean_abs_log = rs.mispredict_under_frac = rs.fallback_rate = 0;
    rs.holdout_err_mlp = rs.holdout_err_const = 0;
    rs.predictor = 0;
    rs.t_construct_kernel_ms = 0;
    rs.n_lds_buckets = 0;

    std::vector<uint32_t> all;
    if (subset) {
        std::vector<char> in(G, 0);
        for (uint32_t i = 0; i < n_subset; ++i) in[subset[i]] = 1;
        for (uint32_t b = 0; b < G; ++b)
            if (in[b] && bucket_records(h, b)) all.push_back(b);
    } else {
        all = nonempty_buckets(h);
    }
    if (h->spilled)
        for (uint32_t b : all)
            if (!h->resident[b])
                return fail(h, CFGPU_E_STATE, "construct: bucket %u is in the bucket store (restore it first)", b);
    rs.n_buckets_nonempty = uint32_t(all.size());
    h->dd_active = false;
    rs.n_records_walked = h->n_records;
    rs.n_dedup_fallback = 0;
    rs.t_dedup_ms = 0;
    if (mode != CFGPU_MODE_PLAIN) CK(dedup_buckets(h, all));
    Event e0, e1, e2, e3, e4, e5;
    HIPCK(h, hipEventRecord(e0.e, h->stream));

    std::vector<uint32_t> ovf;
    std::vector<uint32_t> main_set = all;
    std::vector<uint64_t> main_l2(all.size());
    std::vector<double> cost(G, 0.0);
    for (uint32_t b : all) cost[b] = double(bucket_kmers(h, b));

    if (mode == CFGPU_MODE_PLAIN) {
        // grow-and-retry: start at a quarter of the k-mer instances
        for (size_t i = 0; i < all.size(); ++i) {
            const uint32_t b = all[i];
            const double guess = double(std::max<uint64_t>(bucket_kmers(h, b) / 4, 8)) / h->tune.load_factor;
            main_l2[i] = std::min(upper_cap(h, b), round_cap(guess));
        }
        HIPCK(h, hipEventRecord(e2.e, h->stream));
        CK(run_round_plain(h, main_set, main_l2, ovf));
        HIPCK(h, hipEventRecord(e3.e, h->stream));
        rs.n_overflowed = uint32_t(ovf.size());
        CK(fallback(h, ovf, true));
    } else if (mode == CFGPU_MODE_BATCHED) {
        order_largest_first(main_set, cost);
        HIPCK(h, hipEventRecord(e2.e, h->stream));
        CK(run_main(h, main_set, cost, 1.0, ovf));
        HIPCK(h, hipEventRecord(e3.e, h->stream));
        rs.n_overflowed = uint32_t(ovf.size());
        CK(fallback(h, ovf, false));
    } else if (mode == CFGPU_MODE_BATCHED_LEARNED) {
        // 1. exact tables for an evenly spread sample of buckets: their distinct
        //    counts are the training labels (and their tables are final).
        std::vector<uint32_t> sample, rest;
        const size_t n_s = std::min<size_t>(SAMPLE_TARGET, all.size() / 4);
        if (n_s >= 8) {
            const double step = double(all.size()) / double(n_s);
            std::vector<char> in(all.size(), 0);
            for (size_t i = 0; i < n_s; ++i) in[size_t(i * step)] = 1;
            for (size_t i = 0; i < all.size(); ++i) (in[i] ? sample : rest).push_back(all[i]);
        } else {
            rest = all;
        }
        rs.n_samples = uint32_t(sample.size());
        if (!sample.empty()) {
            // LDS passes sized by the k-mer instances the construct walks (an
            // upper bound on the distinct vertices), the global kernel for
            // buckets needing more than max_passes; direct launches (the main
            // round keeps the captured graph).
            order_largest_first(sample, cost);
            const double C = double(lds_capacity(int(h->KW)));
            std::vector<uint32_t> lds_s, glob_s, so, so2;
            std::vector<uint64_t> lds_c, glob_c;
            for (uint32_t b : sample) {
                const uint64_t inst = std::max<uint64_t>(1, walked_kmers(h, b));
                const uint32_t P = uint32_t(std::ceil(double(inst) / (C * double(h->tune.lds_load))));
                if (P <= h->tune.max_passes) {
                    lds_s.push_back(b);
                    lds_c.push_back(uint64_t(P) * uint64_t(C));
                } else {
                    glob_s.push_back(b);
                    glob_c.push_back(upper_cap(h, b));
                }
            }
            if (!glob_s.empty()) CK(run_round_batched(h, glob_s, glob_c, so, false));
            if (!lds_s.empty()) CK(run_round_batched(h, lds_s, lds_c, so2, false, true));
            so.insert(so.end(), so2.begin(), so2.end());
            CK(fallback(h, so, false));
            for (uint32_t b : sample) h->h_flags[b] |= CFGPU_BUCKET_SAMPLE;
        }
        HIPCK(h, hipEventRecord(e1.e, h->stream));
        // 2. train on (features, log(distinct / instances)), predict all buckets.
        bool have_model = false;
        if (sample.size() >= 8) {
            // Every fourth sample bucket is held out to choose between the MLP
            // and a constant distinct/instance ratio (the training mean): with
            // few informative features the network can memorise its training
            // buckets and predict worse than the mean.
            auto log_ratio = [&](uint32_t b) {
                return std::log(double(std::max<uint32_t>(1, h->h_nvert[b])) /
                                double(std::max<uint64_t>(1, bucket_kmers(h, b))));
            };
            std::vector<uint32_t> idx, hold;
            for (size_t i = 0; i < sample.size(); ++i)
                (i % 4 == 3 || idx.size() == MLP_TRAIN_MAX ? hold : idx).push_back(sample[i]);
            std::vector<float> tgt(idx.size());
            double mean_t = 0.0;
            for (size_t i = 0; i < idx.size(); ++i) {
                tgt[i] = float(log_ratio(idx[i]));
                mean_t += tgt[i];
            }
            mean_t /= double(idx.size());
            double err_const = 0.0;
            for (uint32_t b : hold) err_const += std::fabs(std::min(0.0, mean_t) - log_ratio(b));
            rs.holdout_err_const = hold.empty() ? 0.0 : err_const / double(hold.size());
            std::vector<float> pred(G, 0.f);
            bool use_mlp = false;
            const uint32_t force = h->tune.predictor;
            if (force == 1 || (force == 0 && (hold.empty() || rs.holdout_err_const > CONST_RATIO_GOOD))) {
                HIPCK(h, h->mlp_params.ensure(size_t(mlp_param_floats()) * 4));
                HIPCK(h, h->mlp_adam.ensure(size_t(2 * MLP_PARAMS) * 4));
                HIPCK(h, h->mlp_idx.ensure(idx.size() * 4));
                HIPCK(h, h->mlp_target.ensure(idx.size() * 4));
                HIPCK(h, h->mlp_pred.ensure(G * 4));
                HIPCK(h, h->mlp_loss.ensure(4));
                HIPCK(h, h2d(h, h->mlp_idx.p, idx.data(), idx.size() * 4));
                HIPCK(h, h2d(h, h->mlp_target.p, tgt.data(), tgt.size() * 4));
                launch_mlp_train(h->mlp_params.as<float>(), h->feat.as<float>(), h->mlp_idx.as<uint32_t>(),
                                 h->mlp_target.as<float>(), uint32_t(idx.size()), h->tune.mlp_epochs, h->tune.mlp_lr, MLP_SEED,
                                 h->mlp_adam.as<float>(), h->mlp_loss.as<float>(), h->stream);
                h->h_mlp_idx = idx;
                h->h_mlp_target = tgt;
                h->mlp_epochs_last = h->tune.mlp_epochs;
                h->mlp_lr_last = h->tune.mlp_lr;
                HIPCK(h, hipEventRecord(e4.e, h->stream));
                launch_mlp_infer(h->mlp_params.as<float>(), h->feat.as<float>(), uint32_t(G),
                                 h->mlp_pred.as<float>(), h->stream);
                float loss = 0.f;
                HIPCK(h, hipMemcpyAsync(pred.data(), h->mlp_pred.p, G * 4, hipMemcpyDeviceToHost, h->stream));
                HIPCK(h, hipMemcpyAsync(&loss, h->mlp_loss.p, 4, hipMemcpyDeviceToHost, h->stream));
                HIPCK(h, hipEventRecord(e5.e, h->stream));
                HIPCK(h, sync_stream(h));
                rs.mlp_loss = loss;
                h->mlp_trained = true;
                h->mlp_loss_last = loss;
                rs.t_train_ms = elapsed_ms(e1, e4);
                rs.t_infer_ms = elapsed_ms(e4, e5);
                auto mlp_log = [&](uint32_t b) {
                    return std::isfinite(pred[b]) ? std::min(0.0, double(pred[b])) : 0.0;
                };
                double err_mlp = 0.0;
                for (uint32_t b : hold) err_mlp += std::fabs(mlp_log(b) - log_ratio(b));
                rs.holdout_err_mlp = hold.empty() ? 0.0 : err_mlp / double(hold.size());
                use_mlp = force == 1 || hold.empty() || err_mlp <= err_const;
            }
            rs.predictor = use_mlp ? 1u : 2u;
            const double ratio_const = std::exp(std::min(0.0, mean_t));
            for (uint32_t b : rest) {
                const double km = double(bucket_kmers(h, b));
                const double ratio = use_mlp ? std::exp(std::isfinite(pred[b]) ? std::min(0.0, double(pred[b])) : 0.0)
                                             : ratio_const;
                h->h_pred[b] = float(std::max(1.0, km * ratio));
            }
            have_model = true;
        }
        // 3. size and order the remaining buckets by the prediction.
        main_set = rest;
        std::vector<double> pcost(G, 0.0);
        for (uint32_t b : rest) pcost[b] = have_model ? double(h->h_pred[b]) : cost[b];
        order_largest_first(main_set, pcost);
        HIPCK(h, hipEventRecord(e2.e, h->stream));
        CK(run_main(h, main_set, pcost, have_model ? double(h->tune.margin) : 1.0, ovf));
        HIPCK(h, hipEventRecord(e3.e, h->stream));
        rs.n_overflowed = uint32_t(ovf.size());
        rs.fallback_rate = main_set.empty() ? 0.0 : double(ovf.size()) / double(main_set.size());
        CK(fallback(h, ovf, false));
    } else {
        return fail(h, CFGPU_E_ARG, "unknown construct mode %u", mode);
    }
    Event ef;
    HIPCK(h, hipEventRecord(ef.e, h->stream));
    HIPCK(h, d2h(h, h->h_nvert.data(), h->d_nvert.p, G * 4));
    HIPCK(h, sync_stream(h));

    if (mode == CFGPU_MODE_BATCHED_LEARNED) {
        rs.t_sample_ms = elapsed_ms(e0, e1);
        double sum = 0;
        uint64_t n = 0, under = 0;
        for (uint32_t b : main_set) {
            if (h->h_pred[b] <= 0.f || !h->h_nvert[b]) continue;
            sum += std::fabs(std::log(double(h->h_pred[b]) / double(h->h_nvert[b])));
            under += double(h->h_pred[b]) < double(h->h_nvert[b]);
            ++n;
        }
        rs.mispredict_mean_abs_log = n ? sum / double(n) : 0.0;
        rs.mispredict_under_frac = n ? double(under) / double(n) : 0.0;
    }
    rs.t_construct_main_ms = elapsed_ms(e2, e3);
    rs.t_fallback_ms = elapsed_ms(e3, ef);
    rs.t_construct_ms = elapsed_ms(e0, ef);
    rs.n_vertices = 0;
    rs.table_slots = 0;
    for (uint32_t b : all) {
        rs.n_vertices += h->h_nvert[b];
        rs.table_slots += h->h_cap[b];
        if (!(h->h_flags[b] & CFGPU_BUCKET_BUILT))
            return fail(h, CFGPU_E_OVERFLOW, "bucket %u was not built", b);
    }
    rs.table_bytes = rs.table_slots * entry_bytes(h->KW);
    // An LDS-built table holds exactly its bucket's vertices: its readers
    // (contract, table copies and views) see that many entries.
    bool dense = false;
    for (uint32_t b : all)
        if (h->h_dense[b]) h->h_cap[b] = h->h_nvert[b], dense = true;
    if (dense) {
        HIPCK(h, h2d(h, h->d_cap.p, h->h_cap.data(), G * 8));
        if (h->stage_direct) HIPCK(h, sync_stream(h));
    }
    h->constructed = true;
    return CFGPU_OK;
}

// ---------------------------------------------------------------------------
// External bucket store.

void spill_clear(cfgpu_handle* h) {
    for (auto& sg : h->spill) {
        if (sg.host) (void)hipHostFree(sg.host);
        if (!sg.path.empty()) (void)unlink(sg.path.c_str());
    }
    h->spill.clear();
    h->resident.clear();
    h->t_spill_ms = h->t_restore_ms = 0;
}

hipError_t spill_stage_ensure(cfgpu_handle* h, size_t bytes) {
    if (bytes <= h->spill_stage_cap) return hipSuccess;
    if (h->spill_stage) (void)hipHostFree(h->spill_stage);
    h->spill_stage = nullptr;
    h->spill_stage_cap = 0;
    const hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&h->spill_stage), bytes, hipHostMallocDefault);
    if (e == hipSuccess) h->spill_stage_cap = bytes;
    return e;
}

// Segment byte layout: headers (4 n) | labels (8 W n) | minimizer bins (n).
uint64_t seg_bytes(const cfgpu_handle* h, uint64_t n) { return n * (4 + 8 * uint64_t(h->W) + 1); }

bool write_all(int fd, const uint8_t* p, uint64_t n) {
    while (n) {
        const ssize_t w = write(fd, p, size_t(std::min<uint64_t>(n, 1ull << 30)));
        if (w <= 0) return false;
        p += w;
        n -= uint64_t(w);
    }
    return true;
}

bool pread_all(int fd, uint8_t* p, uint64_t n, uint64_t off) {
    while (n) {
        const ssize_t r = pread(fd, p, size_t(std::min<uint64_t>(n, 1ull << 30)), off_t(off));
        if (r <= 0) return false;
        p += r;
        n -= uint64_t(r);
        off += uint64_t(r);
    }
    return true;
}

// The chunk just partitioned (its buckets in h->att / labels / mbin / boff,
// offsets in h->h_boff) becomes a segment of the store; its per-bucket sums
// are added to the store's.
int spill_chunk(cfgpu_handle* h) {
    const uint32_t G = h->G;
    const uint64_t n = h->n_records, W = h->W;
    const size_t sum_bytes = size_t(G) * SUM_WORDS * 4;
    Event e0, e1;
    HIPCK(h, hipEventRecord(e0.e, h->stream));
    if (h->spill.empty()) {
        HIPCK(h, h->spill_sums.ensure(sum_bytes));
        launch_fill_u32(h->spill_sums.as<uint32_t>(), 0u, uint64_t(G) * SUM_WORDS, h->stream);
    }
    launch_bucket_sums(h->boff.as<uint64_t>(), h->att.as<uint32_t>(), h->mbin.as<uint8_t>(), G, h->prm.k,
                       h->spill_sums.as<uint32_t>(), h->stream);
    HIPCK(h, hipGetLastError());
    cfgpu_handle::SpillSeg sg;
    sg.n_rec = n;
    sg.n_bases = h->n_bases;
    sg.n_frag = h->n_frag;
    sg.h_boff = h->h_boff;
    const cfgpu_run_stats& rs = h->rs;
    sg.t_partition = rs.t_partition_ms;
    sg.t_minimizer = rs.t_minimizer_ms;
    sg.t_count = rs.t_count_ms;
    sg.t_emit = rs.t_emit_ms;
    sg.t_features = rs.t_features_ms;
    const uint64_t bytes = seg_bytes(h, n);
    uint8_t* dst = nullptr;
    if (h->spill_dir.empty()) {
        HIPCK(h, hipHostMalloc(reinterpret_cast<void**>(&sg.host), std::max<uint64_t>(bytes, 1), hipHostMallocDefault));
        dst = sg.host;
    } else {
        HIPCK(h, spill_stage_ensure(h, std::max<uint64_t>(bytes, 1)));
        dst = h->spill_stage;
    }
    if (n) {
        HIPCK(h, hipMemcpyAsync(dst, h->att.p, n * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCK(h, hipMemcpyAsync(dst + n * 4, h->labels.p, n * W * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCK(h, hipMemcpyAsync(dst + n * (4 + 8 * W), h->mbin.p, n, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCK(h, hipEventRecord(e1.e, h->stream));
    HIPCK(h, sync_stream(h));
    if (!h->spill_dir.empty()) {
        char name[64];
        std::snprintf(name, sizeof name, "/cfgpu_bucket_segment_%p_%llu.bin", static_cast<void*>(h),
                      (unsigned long long)h->spill_serial++);
        sg.path = h->spill_dir + name;
        const int fd = open(sg.path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0600);
        if (fd < 0) return fail(h, CFGPU_E_ARG, "spill: cannot create %s", sg.path.c_str());
        const bool ok = write_all(fd, dst, bytes);
        const bool closed = close(fd) == 0;
        if (!ok || !closed) {
            (void)unlink(sg.path.c_str());
            return fail(h, CFGPU_E_ARG, "spill: cannot write %s", sg.path.c_str());
        }
    }
    h->t_spill_ms += elapsed_ms(e0, e1);
    h->spill.push_back(std::move(sg));
    h->rs.n_partition_chunks = uint32_t(h->spill.size());
    return CFGPU_OK;
}

// Statistics and features of the store's buckets; none is in device memory.
int spill_finish(cfgpu_handle* h) {
    const uint32_t G = h->G;
    HIPCK(h, h->stats.ensure(size_t(G) * 4 * 8));
    HIPCK(h, h->feat.ensure(size_t(G) * N_FEAT * 4));
    HIPCK(h, h->boff.ensure((G + 1) * 8));
    launch_features_from_sums(h->spill_sums.as<uint32_t>(), G, h->prm.k, h->stats.as<uint64_t>(),
                              h->feat.as<float>(), h->stream);
    launch_fill_u32(h->boff.as<uint32_t>(), 0u, uint64_t(G + 1) * 2, h->stream);  // nothing resident
    HIPCK(h, hipGetLastError());
    h->h_stats.resize(size_t(G) * 4);
    HIPCK(h, d2h(h, h->h_stats.data(), h->stats.p, size_t(G) * 32));
    HIPCK(h, sync_stream(h));
    // the device bucket arrays held the last chunk only
    h->att.release();
    h->labels.release();
    h->mbin.release();
    h->h_boff.assign(G + 1, 0);
    h->resident.assign(G, 0);
    h->n_dev_records = 0;
    cfgpu_run_stats sum{};
    uint64_t n = 0, nb = 0, nf = 0;
    for (const auto& sg : h->spill) {
        n += sg.n_rec;
        nb += sg.n_bases;
        nf += sg.n_frag;
        sum.t_partition_ms += sg.t_partition;
        sum.t_minimizer_ms += sg.t_minimizer;
        sum.t_count_ms += sg.t_count;
        sum.t_emit_ms += sg.t_emit;
        sum.t_features_ms += sg.t_features;
    }
    h->n_bases = nb;
    h->n_frag = nf;
    h->n_records = n;
    cfgpu_run_stats& rs = h->rs;
    rs = sum;
    rs.n_bases = nb;
    rs.n_fragments = nf;
    rs.n_records = n;
    for (uint32_t b = 0; b < G; ++b) {
        rs.n_kmers += h->h_stats[4 * b + 1];
        rs.n_edges += h->h_stats[4 * b + 1] - h->h_stats[4 * b + 0];
    }
    rs.n_partition_chunks = uint32_t(h->spill.size());
    h->spilled = true;
    h->partitioned = true;
    return CFGPU_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

int cfgpu_open(const cfgpu_params* p, cfgpu_handle** out) {
    if (!p || !out) return CFGPU_E_ARG;
    *out = nullptr;
    const uint32_t k = p->k, l = p->l;
    if (k < 3 || k > 127 || !(k & 1) || l < 1 || l > 32 || l >= k || p->cutoff < 1 ||
        p->cutoff > 15 || p->subgraph_count != G_FIXED || 2 * k - l > 255)
        return CFGPU_E_ARG;
    std::unique_ptr<cfgpu_handle> h(new cfgpu_handle());
    h->prm = *p;
    h->KW = (k + 31) / 32;
    h->W = ((2 * (k - 1) - l + 2) + 31) / 32;
    h->w = k - l;
    h->G = G_FIXED;
    h->g_bits = 14;
    static const uint64_t salt[4] = {4167021922371662411ull, 7320285940802167691ull,
                                     14307255741305819987ull, 10859488101230029397ull};
    h->wy = WyConst{salt[0], salt[1], wymix_host(salt[0], salt[1])};
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= p->device || p->device < 0)
        return CFGPU_E_HIP;
    if (hipSetDevice(p->device) != hipSuccess) return CFGPU_E_HIP;
    if (hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess)
        h->cus = 256;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return CFGPU_E_HIP;
    // this device's copy of the no-alloc scratch array (the device is current)
    if (hipGetSymbolAddress(reinterpret_cast<void**>(&h->noalloc_scratch), HIP_SYMBOL(scratch)) != hipSuccess ||
        !h->noalloc_scratch)
        return CFGPU_E_HIP;
    h->na_off.resize(NOALLOC_MAX_BUCKETS + 1);
    h->na_list.resize(NOALLOC_MAX_BUCKETS);
    h->na_cnt.resize(NOALLOC_MAX_BUCKETS);
    *out = h.release();
    return CFGPU_OK;
}

int cfgpu_close(cfgpu_handle* h) {
    if (!h) return CFGPU_E_ARG;
    (void)hipSetDevice(h->prm.device);
    (void)hipStreamSynchronize(h->stream);
    if (h->stage) (void)hipHostFree(h->stage);
    delete h;
    return CFGPU_OK;
}

int cfgpu_last_error(const cfgpu_handle* h, char* buf, size_t n) {
    if (!h || !buf || !n) return CFGPU_E_ARG;
    std::snprintf(buf, n, "%s", h->err.c_str());
    return CFGPU_OK;
}

int cfgpu_params_get(const cfgpu_handle* h, cfgpu_params* out) {
    if (!h || !out) return CFGPU_E_ARG;
    *out = h->prm;
    return CFGPU_OK;
}

int cfgpu_partition(cfgpu_handle* h, const uint64_t* packed, uint64_t n_bases,
                    const uint64_t* frag_off, uint64_t n_frag, uint32_t flags) {
    if (!h) return CFGPU_E_ARG;
    HIPCK(h, hipSetDevice(h->prm.device));
    const uint32_t k = h->prm.k;
    if ((n_frag && !packed) || (!frag_off && n_frag) ||
        (flags & ~uint32_t(CFGPU_INPUT_DEVICE | CFGPU_PARTITION_APPEND | CFGPU_PARTITION_SPILL)))
        return fail(h, CFGPU_E_ARG, "partition: bad pointers or flags");
    const bool dev = flags & CFGPU_INPUT_DEVICE;
    const bool to_store = flags & CFGPU_PARTITION_SPILL;
    const bool append = (flags & CFGPU_PARTITION_APPEND) || to_store;
    if (to_store && !h->chunks.empty())
        return fail(h, CFGPU_E_STATE, "partition: chunks kept in device memory are pending (finish them first)");
    if (append && !to_store && !h->spill.empty() && !h->spilled)
        return fail(h, CFGPU_E_STATE, "partition: spilled chunks are pending (finish them first)");
    if (!dev) {
        if (n_frag == 0 && n_bases != 0) return fail(h, CFGPU_E_ARG, "partition: bases without fragments");
        if (n_frag) {
            if (frag_off[0] != 0 || frag_off[n_frag] != n_bases)
                return fail(h, CFGPU_E_ARG, "partition: frag_off must run from 0 to n_bases");
            for (uint64_t f = 0; f < n_frag; ++f)
                if (frag_off[f + 1] < frag_off[f] + k + 1)
                    return fail(h, CFGPU_E_ARG, "partition: fragment %llu shorter than k+1",
                                (unsigned long long)f);
        }
    }
    h->partitioned = h->constructed = h->contracted = h->cdbg_done = false;
    h->mlp_trained = false;
    if (!append || h->spilled) {  // new input replaces a finished store
        h->chunks.clear();
        spill_clear(h);
    }
    h->spilled = false;
    const uint32_t G = h->G;
    const uint64_t n_words = (n_bases + 31) / 32 + 2;
    const uint64_t tiles = (n_bases + P_TILE - 1) / P_TILE;
    const uint64_t* d_packed = packed;
    const uint64_t* d_frag = frag_off;
    if (!dev) {
        HIPCK(h, ensure_mem(h, h->packed, n_words * 8, Phase::Input));
        HIPCK(h, ensure_mem(h, h->frag, (n_frag + 1) * 8, Phase::Input));
        if (n_words > 2)
            HIPCK(h, hipMemcpyAsync(h->packed.p, packed, (n_words - 2) * 8, hipMemcpyHostToDevice, h->stream));
        HIPCK(h, hipMemsetAsync(h->packed.as<uint64_t>() + (n_words - 2), 0, 16, h->stream));
        HIPCK(h, hipMemcpyAsync(h->frag.p, frag_off, (n_frag + 1) * 8, hipMemcpyHostToDevice, h->stream));
        d_packed = h->packed.as<uint64_t>();
        d_frag = h->frag.as<uint64_t>();
    }
    const PieceSpans ps = pieces_spans(n_bases, uint32_t(h->cus));
    HIPCK(h, ensure_mem(h, h->gbin, n_bases * 2, Phase::Input));
    HIPCK(h, ensure_mem(h, h->mbn, ((n_bases + 7) / 8) * 4, Phase::Input));
    HIPCK(h, ensure_mem(h, h->tile_lc, tiles * 8, Phase::Input));
    HIPCK(h, ensure_mem(h, h->pdesc, tiles * P_TILE * 8, Phase::Input));
    HIPCK(h, ensure_mem(h, h->tile_n, tiles * 4, Phase::Input));
    HIPCK(h, ensure_mem(h, h->tile_frag, tiles * 8, Phase::Input));
    HIPCK(h, ensure_mem(h, h->hist, size_t(N_COARSE) * ps.n_count * 4, Phase::Input));
    HIPCK(h, h->ctot.ensure(size_t(N_COARSE) * 8));
    HIPCK(h, h->coff.ensure(size_t(N_COARSE + 1) * 8));
    HIPCK(h, h->boff.ensure((G + 1) * 8));
    HIPCK(h, h->stats.ensure(size_t(G) * 4 * 8));
    HIPCK(h, h->feat.ensure(size_t(G) * N_FEAT * 4));

    PartitionArgs a{};
    a.packed = d_packed;
    a.n_words = n_words;
    a.n_bases = n_bases;
    a.frag_off = d_frag;
    a.tile_frag = h->tile_frag.as<uint64_t>();
    a.n_frag = n_frag;
    a.k = k;
    a.l = h->prm.l;
    a.w = h->w;
    a.g_mask = G - 1;
    a.g_bits = h->g_bits;
    a.wy = h->wy;

    Event e0, e1, e2, e3, e4, e5;
    HIPCK(h, hipMemsetAsync(h->coff.p, 0, size_t(N_COARSE + 1) * 8, h->stream));
    HIPCK(h, hipEventRecord(e0.e, h->stream));
    if (n_frag) launch_minimizer_g(a, h->gbin.as<uint16_t>(), h->mbn.as<uint32_t>(), h->tile_lc.as<uint64_t>(), h->stream);
    HIPCK(h, hipEventRecord(e1.e, h->stream));
    if (n_frag)
        launch_pieces_hist(a, h->gbin.as<uint16_t>(), h->mbn.as<uint32_t>(), h->tile_lc.as<uint64_t>(), ps,
                           h->hist.as<uint32_t>(), h->ctot.as<uint64_t>(), h->coff.as<uint64_t>(),
                           h->pdesc.as<uint64_t>(), h->tile_n.as<uint32_t>(), h->stream);
    HIPCK(h, hipEventRecord(e2.e, h->stream));
    uint64_t n_rec = 0;
    HIPCK(h, d2h(h, &n_rec, h->coff.as<uint64_t>() + N_COARSE, 8));
    HIPCK(h, sync_stream(h));
    const uint64_t lab_off = l1_label_offset(n_rec);
    HIPCK(h, ensure_mem(h, h->rec1, (lab_off + n_rec * l1_label_words(h->W)) * 8, Phase::Input));
    HIPCK(h, ensure_mem(h, h->att, n_rec * 4, Phase::Input));
    HIPCK(h, ensure_mem(h, h->labels, n_rec * h->W * 8, Phase::Input));
    HIPCK(h, ensure_mem(h, h->mbin, n_rec, Phase::Input));
    HIPCK(h, hipEventRecord(e3.e, h->stream));
    if (n_frag)
        launch_pieces_emit(a, h->pdesc.as<uint64_t>(), h->tile_n.as<uint32_t>(), ps, h->hist.as<uint32_t>(),
                           h->coff.as<uint64_t>(), h->W, h->rec1.as<uint32_t>(), h->rec1.as<uint64_t>() + lab_off,
                           h->stream);
    launch_bucket_fine(h->rec1.as<uint32_t>(), h->rec1.as<uint64_t>() + lab_off, h->coff.as<uint64_t>(), h->W,
                       h->boff.as<uint64_t>(), h->att.as<uint32_t>(), h->labels.as<uint64_t>(), h->mbin.as<uint8_t>(),
                       h->stream);
    HIPCK(h, hipEventRecord(e4.e, h->stream));
    launch_bucket_features(h->boff.as<uint64_t>(), h->att.as<uint32_t>(), h->mbin.as<uint8_t>(), G, k,
                           h->stats.as<uint64_t>(), h->feat.as<float>(), h->stream);
    HIPCK(h, hipEventRecord(e5.e, h->stream));
    HIPCK(h, hipGetLastError());
    h->h_boff.resize(G + 1);
    h->h_stats.resize(size_t(G) * 4);
    HIPCK(h, d2h(h, h->h_boff.data(), h->boff.p, (G + 1) * 8));
    HIPCK(h, d2h(h, h->h_stats.data(), h->stats.p, size_t(G) * 32));
    HIPCK(h, sync_stream(h));

    h->n_bases = n_bases;
    h->n_frag = n_frag;
    h->n_records = n_rec;
    cfgpu_run_stats& rs = h->rs;
    rs = cfgpu_run_stats{};
    rs.n_bases = n_bases;
    rs.n_fragments = n_frag;
    rs.n_records = n_rec;
    for (uint32_t b = 0; b < G; ++b) {
        rs.n_kmers += h->h_stats[4 * b + 1];
        rs.n_edges += h->h_stats[4 * b + 1] - h->h_stats[4 * b + 0];
    }
    rs.t_minimizer_ms = elapsed_ms(e0, e1);
    rs.t_count_ms = elapsed_ms(e1, e2);
    rs.t_emit_ms = elapsed_ms(e3, e4);
    rs.t_features_ms = elapsed_ms(e4, e5);
    rs.t_partition_ms = elapsed_ms(e0, e5);
    rs.n_partition_chunks = 1;
    if (to_store) return spill_chunk(h);
    if (append) {
        // the chunk's buckets move aside; the next chunk allocates its own
        auto c = std::make_unique<cfgpu_handle::PartChunk>();
        swap_buf(c->att, h->att);
        swap_buf(c->labels, h->labels);
        swap_buf(c->mbin, h->mbin);
        swap_buf(c->boff, h->boff);
        c->h_boff = h->h_boff;
        c->n_rec = n_rec;
        c->n_bases = n_bases;
        c->n_frag = n_frag;
        c->t_minimizer = rs.t_minimizer_ms;
        c->t_count = rs.t_count_ms;
        c->t_emit = rs.t_emit_ms;
        c->t_features = rs.t_features_ms;
        c->t_partition = rs.t_partition_ms;
        h->chunks.push_back(std::move(c));
        rs.n_partition_chunks = uint32_t(h->chunks.size());
        return CFGPU_OK;
    }
    h->partitioned = true;
    h->n_dev_records = n_rec;
    return CFGPU_OK;
}

int cfgpu_partition_finish(cfgpu_handle* h) {
    if (!h) return CFGPU_E_ARG;
    if (!h->spill.empty() && !h->spilled) {
        HIPCK(h, hipSetDevice(h->prm.device));
        return spill_finish(h);
    }
    if (h->chunks.empty()) return fail(h, CFGPU_E_STATE, "partition_finish without appended chunks");
    HIPCK(h, hipSetDevice(h->prm.device));
    const uint32_t G = h->G, nc = uint32_t(h->chunks.size());
    // bucket b: chunk 0's records, then chunk 1's, ...
    std::vector<uint64_t> boff(G + 1, 0), dst(size_t(nc) * G);
    for (uint32_t b = 0; b < G; ++b) {
        uint64_t at = boff[b];
        for (uint32_t c = 0; c < nc; ++c) {
            const auto& hb = h->chunks[c]->h_boff;
            dst[size_t(c) * G + b] = at;
            at += hb[b + 1] - hb[b];
        }
        boff[b + 1] = at;
    }
    const uint64_t n = boff[G];
    std::vector<ChunkSrc> src(nc);
    uint64_t n_bases = 0, n_frag = 0;
    cfgpu_run_stats sum{};
    for (uint32_t c = 0; c < nc; ++c) {
        const auto& ch = *h->chunks[c];
        src[c] = {ch.att.as<uint32_t>(), ch.labels.as<uint64_t>(), ch.mbin.as<uint8_t>(), ch.boff.as<uint64_t>()};
        n_bases += ch.n_bases;
        n_frag += ch.n_frag;
        sum.t_minimizer_ms += ch.t_minimizer;
        sum.t_count_ms += ch.t_count;
        sum.t_emit_ms += ch.t_emit;
        sum.t_features_ms += ch.t_features;
        sum.t_partition_ms += ch.t_partition;
    }
    HIPCK(h, ensure_mem(h, h->att, n * 4, Phase::Partition));
    HIPCK(h, ensure_mem(h, h->labels, n * h->W * 8, Phase::Partition));
    HIPCK(h, ensure_mem(h, h->mbin, n, Phase::Partition));
    HIPCK(h, h->boff.ensure((G + 1) * 8));
    HIPCK(h, h->stats.ensure(size_t(G) * 4 * 8));
    HIPCK(h, h->feat.ensure(size_t(G) * N_FEAT * 4));
    HIPCK(h, h->ch_src.ensure(nc * sizeof(ChunkSrc)));
    HIPCK(h, h->ch_dst.ensure(dst.size() * 8));
    HIPCK(h, h2d(h, h->ch_src.p, src.data(), nc * sizeof(ChunkSrc)));
    HIPCK(h, h2d(h, h->ch_dst.p, dst.data(), dst.size() * 8));
    HIPCK(h, h2d(h, h->boff.p, boff.data(), (G + 1) * 8));
    Event e0, e1;
    HIPCK(h, hipEventRecord(e0.e, h->stream));
    launch_merge_chunks(h->ch_src.as<ChunkSrc>(), nc, G, h->W, h->ch_dst.as<uint64_t>(), h->att.as<uint32_t>(),
                        h->labels.as<uint64_t>(), h->mbin.as<uint8_t>(), h->stream);
    HIPCK(h, hipEventRecord(e1.e, h->stream));
    launch_bucket_features(h->boff.as<uint64_t>(), h->att.as<uint32_t>(), h->mbin.as<uint8_t>(), G, h->prm.k,
                           h->stats.as<uint64_t>(), h->feat.as<float>(), h->stream);
    HIPCK(h, hipGetLastError());
    h->h_boff = boff;
    h->h_stats.resize(size_t(G) * 4);
    HIPCK(h, d2h(h, h->h_stats.data(), h->stats.p, size_t(G) * 32));
    HIPCK(h, sync_stream(h));
    h->chunks.clear();

    h->n_bases = n_bases;
    h->n_frag = n_frag;
    h->n_records = n;
    cfgpu_run_stats& rs = h->rs;
    rs = sum;
    rs.n_bases = n_bases;
    rs.n_fragments = n_frag;
    rs.n_records = n;
    for (uint32_t b = 0; b < G; ++b) {
        rs.n_kmers += h->h_stats[4 * b + 1];
        rs.n_edges += h->h_stats[4 * b + 1] - h->h_stats[4 * b + 0];
    }
    rs.n_partition_chunks = nc;
    rs.t_partition_merge_ms = elapsed_ms(e0, e1);
    h->partitioned = true;
    h->n_dev_records = n;
    return CFGPU_OK;
}

int cfgpu_spill_dir(cfgpu_handle* h, const char* dir) {
    if (!h) return CFGPU_E_ARG;
    if (!h->spill.empty()) return fail(h, CFGPU_E_STATE, "spill_dir: the bucket store is not empty");
    h->spill_dir = dir ? dir : "";
    while (h->spill_dir.size() > 1 && h->spill_dir.back() == '/') h->spill_dir.pop_back();
    if (!h->spill_dir.empty() && access(h->spill_dir.c_str(), W_OK) != 0)
        return fail(h, CFGPU_E_ARG, "spill_dir: %s is not a writable directory", h->spill_dir.c_str());
    return CFGPU_OK;
}

int cfgpu_buckets_restore(cfgpu_handle* h, const uint32_t* buckets, uint32_t n) {
    if (!h || (n && !buckets)) return CFGPU_E_ARG;
    if (!h->spilled) return fail(h, CFGPU_E_STATE, "restore: the handle's buckets are not in a bucket store");
    const uint32_t G = h->G;
    std::vector<uint8_t> want(G, 0);
    for (uint32_t i = 0; i < n; ++i) {
        if (buckets[i] >= G) return fail(h, CFGPU_E_ARG, "restore: bucket %u out of range", buckets[i]);
        want[buckets[i]] = 1;
    }
    HIPCK(h, hipSetDevice(h->prm.device));
    h->constructed = h->contracted = h->cdbg_done = false;
    const uint64_t W = h->W;
    const uint32_t nc = uint32_t(h->spill.size());
    // Device layout: the listed buckets in bucket order, each one's records
    // segment after segment (as cfgpu_partition_finish gathers kept chunks).
    // Each segment's records of the listed buckets (one range per run of
    // consecutive listed buckets) are copied into a device staging area,
    // segment after segment, and the chunk gather kernel puts them in place.
    std::vector<uint64_t> boff(G + 1, 0), dst(size_t(nc) * G), sboff(size_t(nc) * (G + 1));
    for (uint32_t b = 0; b < G; ++b) {
        uint64_t at = boff[b];
        for (uint32_t c = 0; c < nc; ++c) {
            dst[size_t(c) * G + b] = at;
            if (want[b]) at += h->spill[c].h_boff[b + 1] - h->spill[c].h_boff[b];
        }
        boff[b + 1] = at;
    }
    const uint64_t tot = boff[G];
    struct Run {
        uint32_t c;
        uint64_t r_lo, n, to;  // segment records [r_lo, r_lo + n) to staging record `to`
    };
    std::vector<Run> runs;
    uint64_t st_at = 0;
    for (uint32_t c = 0; c < nc; ++c) {
        const auto& hb = h->spill[c].h_boff;
        uint64_t* sb = &sboff[size_t(c) * (G + 1)];
        sb[0] = st_at;
        for (uint32_t b1 = 0; b1 < G;) {
            if (!want[b1]) {
                sb[b1 + 1] = sb[b1];
                ++b1;
                continue;
            }
            uint32_t b2 = b1;
            while (b2 < G && want[b2]) {
                sb[b2 + 1] = sb[b2] + (hb[b2 + 1] - hb[b2]);
                ++b2;
            }
            if (hb[b2] > hb[b1]) runs.push_back({c, hb[b1], hb[b2] - hb[b1], sb[b1]});
            b1 = b2;
        }
        st_at = sb[G];
    }
    Event e0, e1;
    HIPCK(h, hipEventRecord(e0.e, h->stream));
    const uint64_t m1 = std::max<uint64_t>(tot, 1);
    HIPCK(h, ensure_mem(h, h->att, m1 * 4, Phase::Input));
    HIPCK(h, ensure_mem(h, h->labels, m1 * W * 8, Phase::Input));
    HIPCK(h, ensure_mem(h, h->mbin, m1, Phase::Input));
    HIPCK(h, ensure_mem(h, h->rst_att, m1 * 4, Phase::Input));
    HIPCK(h, ensure_mem(h, h->rst_lab, m1 * W * 8, Phase::Input));
    HIPCK(h, ensure_mem(h, h->rst_mb, m1, Phase::Input));
    const bool files = nc && !h->spill.front().path.empty();
    uint8_t* const d_att = h->rst_att.as<uint8_t>();
    uint8_t* const d_lab = h->rst_lab.as<uint8_t>();
    uint8_t* const d_mb = h->rst_mb.as<uint8_t>();
    if (!files) {
        // pinned segments: straight to the device
        for (const Run& r : runs) {
            const auto& sg = h->spill[r.c];
            const uint64_t m = sg.n_rec;
            HIPCK(h, hipMemcpyAsync(d_att + r.to * 4, sg.host + r.r_lo * 4, r.n * 4, hipMemcpyHostToDevice,
                                    h->stream));
            HIPCK(h, hipMemcpyAsync(d_lab + r.to * 8 * W, sg.host + m * 4 + r.r_lo * 8 * W, r.n * 8 * W,
                                    hipMemcpyHostToDevice, h->stream));
            HIPCK(h, hipMemcpyAsync(d_mb + r.to, sg.host + m * (4 + 8 * W) + r.r_lo, r.n, hipMemcpyHostToDevice,
                                    h->stream));
        }
    } else {
        // file segments: read into pinned memory (host threads over the runs), then to the device
        HIPCK(h, spill_stage_ensure(h, std::max<uint64_t>(seg_bytes(h, tot), 1)));
        uint8_t* const s_att = h->spill_stage;
        uint8_t* const s_lab = s_att + tot * 4;
        uint8_t* const s_mb = s_lab + tot * W * 8;
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::string> errs(nt);
        const auto work = [&](unsigned t) {
            int fd = -1;
            uint32_t open_c = ~0u;
            for (size_t i = t; i < runs.size(); i += nt) {
                const Run& r = runs[i];
                const auto& sg = h->spill[r.c];
                if (r.c != open_c) {
                    if (fd >= 0) close(fd);
                    fd = open(sg.path.c_str(), O_RDONLY);
                    open_c = r.c;
                    if (fd < 0) {
                        errs[t] = "cannot open " + sg.path;
                        return;
                    }
                }
                const uint64_t m = sg.n_rec;
                if (!pread_all(fd, s_att + r.to * 4, r.n * 4, r.r_lo * 4) ||
                    !pread_all(fd, s_lab + r.to * 8 * W, r.n * 8 * W, m * 4 + r.r_lo * 8 * W) ||
                    !pread_all(fd, s_mb + r.to, r.n, m * (4 + 8 * W) + r.r_lo)) {
                    errs[t] = "cannot read " + sg.path;
                    break;
                }
            }
            if (fd >= 0) close(fd);
        };
        std::vector<std::thread> th;
        for (unsigned t = 1; t < nt; ++t) th.emplace_back(work, t);
        work(0);
        for (auto& x : th) x.join();
        for (const auto& e : errs)
            if (!e.empty()) return fail(h, CFGPU_E_ARG, "restore: %s", e.c_str());
        if (tot) {
            HIPCK(h, hipMemcpyAsync(d_att, s_att, tot * 4, hipMemcpyHostToDevice, h->stream));
            HIPCK(h, hipMemcpyAsync(d_lab, s_lab, tot * W * 8, hipMemcpyHostToDevice, h->stream));
            HIPCK(h, hipMemcpyAsync(d_mb, s_mb, tot, hipMemcpyHostToDevice, h->stream));
        }
    }
    // gather: segment c's staged records of bucket b, at sboff[c][b], go to dst[c][b]
    std::vector<ChunkSrc> src(nc);
    HIPCK(h, h->ch_src.ensure(std::max<size_t>(nc, 1) * sizeof(ChunkSrc)));
    HIPCK(h, h->ch_dst.ensure(std::max<size_t>(dst.size(), 1) * 8));
    HIPCK(h, h->rst_boff.ensure(std::max<size_t>(sboff.size(), 1) * 8));
    for (uint32_t c = 0; c < nc; ++c)
        src[c] = {h->rst_att.as<uint32_t>(), h->rst_lab.as<uint64_t>(), h->rst_mb.as<uint8_t>(),
                  h->rst_boff.as<uint64_t>() + size_t(c) * (G + 1)};
    HIPCK(h, hipMemcpyAsync(h->rst_boff.p, sboff.data(), sboff.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCK(h, hipMemcpyAsync(h->ch_dst.p, dst.data(), dst.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPCK(h, hipMemcpyAsync(h->ch_src.p, src.data(), nc * sizeof(ChunkSrc), hipMemcpyHostToDevice, h->stream));
    HIPCK(h, hipMemcpyAsync(h->boff.p, boff.data(), (G + 1) * 8, hipMemcpyHostToDevice, h->stream));
    launch_merge_chunks(h->ch_src.as<ChunkSrc>(), nc, G, h->W, h->ch_dst.as<uint64_t>(), h->att.as<uint32_t>(),
                        h->labels.as<uint64_t>(), h->mbin.as<uint8_t>(), h->stream);
    HIPCK(h, hipGetLastError());
    HIPCK(h, hipEventRecord(e1.e, h->stream));
    HIPCK(h, sync_stream(h));  // the host arrays above are read by the copies until here
    h->t_restore_ms = elapsed_ms(e0, e1);
    h->rst_att.release();  // staging is not kept between restores
    h->rst_lab.release();
    h->rst_mb.release();
    h->h_boff = boff;
    h->resident = want;
    h->n_dev_records = tot;
    return CFGPU_OK;
}

int cfgpu_spill_stats(cfgpu_handle* h, cfgpu_spill_stat* out) {
    if (!h || !out) return CFGPU_E_ARG;
    cfgpu_spill_stat s{};
    s.n_segments = h->spill.size();
    for (const auto& sg : h->spill) {
        s.n_records += sg.n_rec;
        s.bytes += seg_bytes(h, sg.n_rec);
    }
    s.in_files = !h->spill.empty() && !h->spill.front().path.empty();
    for (uint8_t r : h->resident) s.n_resident += r;
    s.n_resident_records = h->spilled ? h->n_dev_records : 0;
    s.t_spill_ms = h->t_spill_ms;
    s.t_restore_ms = h->t_restore_ms;
    *out = s;
    return CFGPU_OK;
}

int cfgpu_bucket_stats(cfgpu_handle* h, cfgpu_bucket_stat* out, uint32_t n) {
    if (!h || !out || n < h->G) return CFGPU_E_ARG;
    if (!h->partitioned) return fail(h, CFGPU_E_STATE, "bucket_stats before partition");
    for (uint32_t b = 0; b < h->G; ++b) {
        cfgpu_bucket_stat& s = out[b];
        s.n_records = h->h_stats[4 * b + 0];
        s.n_kmers = h->h_stats[4 * b + 1];
        s.n_bases = h->h_stats[4 * b + 2];
        const bool c = h->constructed && s.n_records;
        s.n_vertices = c ? h->h_nvert[b] : 0;
        s.capacity = c ? h->h_cap[b] : 0;
        s.predicted = c ? h->h_pred[b] : 0.f;
        s.flags = c ? h->h_flags[b] : 0;
        const bool dd = c && h->dd_active;
        s.n_records_walked = dd ? h->h_ddcnt[b] : (c ? s.n_records : 0);
        s.n_kmers_walked = dd ? h->h_ddkm[b] : (c ? s.n_kmers : 0);
    }
    return CFGPU_OK;
}

int cfgpu_records_copy(cfgpu_handle* h, uint32_t* att, uint64_t* labels, uint64_t* bucket_off,
                       uint64_t cap_records, uint64_t* n_records) {
    if (!h) return CFGPU_E_ARG;
    if (!h->partitioned) return fail(h, CFGPU_E_STATE, "records_copy before partition");
    HIPCK(h, hipSetDevice(h->prm.device));
    const uint64_t nd = h->n_dev_records;  // spilled buckets: the restored ones
    if (n_records) *n_records = nd;
    if (cap_records < nd) return fail(h, CFGPU_E_ARG, "records_copy: capacity too small");
    if (att && nd) HIPCK(h, hipMemcpyAsync(att, h->att.p, nd * 4, hipMemcpyDeviceToHost, h->stream));
    if (labels && nd)
        HIPCK(h, hipMemcpyAsync(labels, h->labels.p, nd * h->W * 8, hipMemcpyDeviceToHost, h->stream));
    if (bucket_off) std::memcpy(bucket_off, h->h_boff.data(), (h->G + 1) * 8);
    HIPCK(h, sync_stream(h));
    return CFGPU_OK;
}

int cfgpu_load_buckets(cfgpu_handle* h, const uint32_t* att, const uint64_t* labels,
                       const uint64_t* bucket_off) {
    if (!h) return CFGPU_E_ARG;
    if (!att || !labels || !bucket_off) return fail(h, CFGPU_E_ARG, "load_buckets: null pointer");
    HIPCK(h, hipSetDevice(h->prm.device));
    const uint32_t G = h->G, k = h->prm.k, l = h->prm.l;
    if (bucket_off[0] != 0) return fail(h, CFGPU_E_ARG, "load_buckets: bucket_off[0] != 0");
    const uint32_t max_len = std::min<uint32_t>(255, 2 * k - l);
    for (uint32_t g = 0; g < G; ++g) {
        if (bucket_off[g + 1] < bucket_off[g]) return fail(h, CFGPU_E_ARG, "load_buckets: offsets decrease");
        for (uint64_t r = bucket_off[g]; r < bucket_off[g + 1]; ++r) {
            const uint32_t a = att[r], len = a & 0xFFu;
            if ((a >> 16) != g || len < k || len > max_len || ((a >> 10) & 0x3Fu))
                return fail(h, CFGPU_E_ARG, "load_buckets: record %llu of bucket %u is malformed",
                            (unsigned long long)r, g);
        }
    }
    h->partitioned = h->constructed = h->contracted = h->cdbg_done = false;
    h->chunks.clear();
    spill_clear(h);
    h->spilled = false;
    const uint64_t n = bucket_off[G];
    HIPCK(h, h->att.ensure(n * 4));
    HIPCK(h, h->labels.ensure(n * h->W * 8));
    HIPCK(h, h->mbin.ensure(n));
    HIPCK(h, h->boff.ensure((G + 1) * 8));
    HIPCK(h, h->stats.ensure(size_t(G) * 4 * 8));
    HIPCK(h, h->feat.ensure(size_t(G) * N_FEAT * 4));
    if (n) {
        HIPCK(h, hipMemcpyAsync(h->att.p, att, n * 4, hipMemcpyHostToDevice, h->stream));
        HIPCK(h, hipMemcpyAsync(h->labels.p, labels, n * h->W * 8, hipMemcpyHostToDevice, h->stream));
        HIPCK(h, hipMemsetAsync(h->mbin.p, 0, n, h->stream));  // no minimizer bins for loaded records
    }
    HIPCK(h, hipMemcpyAsync(h->boff.p, bucket_off, (G + 1) * 8, hipMemcpyHostToDevice, h->stream));
    launch_bucket_features(h->boff.as<uint64_t>(), h->att.as<uint32_t>(), h->mbin.as<uint8_t>(), G, k,
                           h->stats.as<uint64_t>(), h->feat.as<float>(), h->stream);
    HIPCK(h, hipGetLastError());
    h->h_boff.assign(bucket_off, bucket_off + G + 1);
    h->h_stats.resize(size_t(G) * 4);
    HIPCK(h, d2h(h, h->h_stats.data(), h->stats.p, size_t(G) * 32));
    HIPCK(h, sync_stream(h));
    h->n_bases = 0;
    h->n_frag = 0;
    h->n_records = n;
    h->n_dev_records = n;
    cfgpu_run_stats& rs = h->rs;
    rs = cfgpu_run_stats{};
    rs.n_records = n;
    for (uint32_t b = 0; b < G; ++b) {
        rs.n_kmers += h->h_stats[4 * b + 1];
        rs.n_edges += h->h_stats[4 * b + 1] - h->h_stats[4 * b + 0];
    }
    h->partitioned = true;
    return CFGPU_OK;
}

int cfgpu_construct(cfgpu_handle* h, uint32_t mode) {
    if (!h) return CFGPU_E_ARG;
    if (!h->partitioned) return fail(h, CFGPU_E_STATE, "construct before partition");
    HIPCK(h, hipSetDevice(h->prm.device));
    h->constructed = h->contracted = h->cdbg_done = false;
    return construct_impl(h, mode);
}

int cfgpu_construct_buckets(cfgpu_handle* h, uint32_t mode, const uint32_t* buckets, uint32_t n) {
    if (!h || (n && !buckets)) return CFGPU_E_ARG;
    if (!h->partitioned) return fail(h, CFGPU_E_STATE, "construct before partition");
    for (uint32_t i = 0; i < n; ++i)
        if (buckets[i] >= h->G) return fail(h, CFGPU_E_ARG, "construct: bucket %u out of range", buckets[i]);
    HIPCK(h, hipSetDevice(h->prm.device));
    h->constructed = h->contracted = h->cdbg_done = false;
    return construct_impl(h, mode, buckets, n);
}

static_assert(sizeof(cfgpu_unitig) == sizeof(UnitigRec) && sizeof(cfgpu_unitig) == 16, "cfgpu_unitig layout");
static_assert(sizeof(cfgpu_contract_stat) == 32, "cfgpu_contract_stat layout");

int cfgpu_contract(cfgpu_handle* h) {
    if (!h) return CFGPU_E_ARG;
    if (!h->constructed) return fail(h, CFGPU_E_STATE, "contract before construct");
    HIPCK(h, hipSetDevice(h->prm.device));
    h->contracted = h->cdbg_done = false;
    // Subgraphs whose vertex count passes 16 LDS tables go straight to the
    // dense tables (their non-isolated vertices would not fit an LDS table
    // unless nearly every vertex were isolated; the dense pass takes any
    // subgraph, so the choice changes only where the work runs).
    std::vector<uint32_t> list, big;
    uint64_t sum_nv = 0;
    uint32_t max_nv = 0;
    const uint32_t direct = 16u * uint32_t(ct_lds_capacity(int(h->KW))) * 4u / 5u;
    for (uint32_t g = 0; g < h->G; ++g)
        if (h->h_tbl_ptr[g] && h->h_nvert[g]) {
            (h->h_nvert[g] > direct ? big : list).push_back(g);
            sum_nv += h->h_nvert[g];
            max_nv = std::max(max_nv, h->h_nvert[g]);
        }
    const uint32_t KW = h->KW, n_list = uint32_t(list.size()), n_big = uint32_t(big.size());
    const uint32_t grid = std::min<uint32_t>(n_list + n_big, uint32_t(h->cus));  // one 1024-thread workgroup per CU
    // scratch per workgroup: the dense tables or a global table for the
    // largest subgraph, whichever is larger (256-byte aligned regions)
    const uint64_t per_wg =
        (std::max(ct_global_capacity(max_nv) * uint64_t(8 * KW + 8), ct_big_scratch_bytes(KW)) + 255) & ~255ull;
    HIPCK(h, h->ct_list.ensure(n_list * 4ull));
    HIPCK(h, h->ct_ovf.ensure((n_list + n_big) * 4ull));
    HIPCK(h, h->ct_ovf2.ensure((n_list + n_big) * 4ull));
    HIPCK(h, h->ct_ovf3.ensure((n_list + n_big) * 4ull));
    HIPCK(h, h->ct_counts.ensure(size_t(h->G) * sizeof(cfgpu_contract_stat)));
    HIPCK(h, h->ct_ctl.ensure(64));
    HIPCK(h, ensure_mem(h, h->ct_scratch, per_wg * std::max<uint32_t>(grid, 1), Phase::Contract));
    // tables built in LDS passes (at most SEG_MAX) are read over their runs' active vertices
    std::vector<uint32_t> nseg(h->G, 0u);
    for (uint32_t g = 0; g < h->G; ++g)
        if (h->h_tbl_ptr[g] && h->h_dense[g] && h->h_passes[g] <= SEG_MAX) nseg[g] = h->h_passes[g];
    HIPCK(h, h->d_nseg.ensure(size_t(h->G) * 4));
    HIPCK(h, h2d(h, h->d_nseg.p, nseg.data(), size_t(h->G) * 4));
    if (n_list) HIPCK(h, h2d(h, h->ct_list.p, list.data(), n_list * 4ull));
    if (n_big) HIPCK(h, h2d(h, h->ct_ovf.p, big.data(), n_big * 4ull));  // the dense pass's list starts with them
    // Output sized from the vertex count (most vertices of read sets are
    // isolated) or the last call's exact need, whichever is larger; re-run
    // once at the exact size when short.
    uint64_t rec_cap = std::max<uint64_t>({4096, sum_nv / 16, h->ct_need[0]}),
             lab_cap = std::max<uint64_t>({1 << 20, sum_nv * 2, h->ct_need[1]});
    Event e0, e1;
    HIPCK(h, hipEventRecord(e0.e, h->stream));
    uint32_t ctl[16];
    uint32_t run = 0;
    uint32_t wide_cap = 0;
    for (;;) {
        ++run;
        HIPCK(h, ensure_mem(h, h->ct_recs, rec_cap * sizeof(cfgpu_unitig), Phase::Contract));
        HIPCK(h, ensure_mem(h, h->ct_labels, lab_cap + 64, Phase::Contract));  // slack: labels are read as aligned words
        launch_fill_u32(h->ct_counts.as<uint32_t>(), 0u, uint64_t(h->G) * 8, h->stream);
        launch_fill_u32(h->ct_ctl.as<uint32_t>(), 0u, 16, h->stream);
        launch_fill_u32(h->ct_ctl.as<uint32_t>() + 2, n_big, 1, h->stream);  // ovf_count
        ContractArgs a{};
        a.tbl_ptr = h->d_tbl_ptr.as<uint64_t>();
        a.tbl_cap = h->d_cap.as<uint64_t>();
        a.n_vert = h->d_nvert.as<uint32_t>();
        a.list = h->ct_list.as<uint32_t>();
        a.n_list = n_list;
        a.k = h->prm.k;
        a.cutoff = h->prm.cutoff;
        uint32_t* ctl_d = h->ct_ctl.as<uint32_t>();
        a.queue_head = ctl_d;       // [0], [1]
        a.ovf_count = ctl_d + 2;
        a.walk_head = ctl_d + 5;
        a.ovf2_count = ctl_d + 6;
        a.ovf2_list = h->ct_ovf2.as<uint32_t>();
        a.retry = ctl_d + 3;
        a.error = ctl_d + 4;
        a.cursors = reinterpret_cast<unsigned long long*>(ctl_d + 8);  // [8..11]
        a.ovf_list = h->ct_ovf.as<uint32_t>();
        a.scratch = h->ct_scratch.as<uint8_t>();
        a.scratch_per_wg = per_wg;
        a.counts = h->ct_counts.as<uint32_t>();
        a.recs = h->ct_recs.as<UnitigRec>();
        a.rec_cap = rec_cap;
        a.labels = h->ct_labels.as<char>();
        a.lab_cap = lab_cap;
        a.seg = h->d_seg.as<uint64_t>();
        a.nseg = h->d_nseg.as<uint32_t>();
        a.iso = h->d_iso.as<uint32_t>();
        a.wide_head = ctl_d + 7;
        a.ovf3_list = h->ct_ovf3.as<uint32_t>();
        a.ovf3_count = ctl_d + 12;
        launch_contract(a, KW, grid, grid, 0, h->stream);
        HIPCK(h, hipGetLastError());
        // The dense pass's overflow: the wide dense pass takes the subgraphs
        // its scratch holds (sized here for the largest of them), the global
        // pass the rest.
        uint32_t n2 = 0;
        wide_cap = 0;
        if (grid) {
            HIPCK(h, hipMemcpyAsync(&n2, ctl_d + 6, 4, hipMemcpyDeviceToHost, h->stream));
            HIPCK(h, sync_stream(h));
        }
        if (n2 && h->tune.contract_wide) {
            std::vector<uint32_t> l2(n2);
            HIPCK(h, hipMemcpyAsync(l2.data(), h->ct_ovf2.p, n2 * 4ull, hipMemcpyDeviceToHost, h->stream));
            HIPCK(h, sync_stream(h));
            uint32_t m2 = 0;
            for (uint32_t g : l2) m2 = std::max(m2, h->h_nvert[g]);
            wide_cap = std::min<uint32_t>(m2, ct_wide_max_vertices());
            const uint64_t per_wg2 = std::max(per_wg, ct_wide_scratch_bytes(KW, wide_cap));
            HIPCK(h, ensure_mem(h, h->ct_scratch, per_wg2 * std::max<uint32_t>(grid, 1), Phase::Contract));
            a.scratch = h->ct_scratch.as<uint8_t>();
            a.scratch_per_wg = per_wg2;
            a.wide_cap = wide_cap;
        }
        if (n2) launch_contract(a, KW, grid, grid, 1, h->stream);
        HIPCK(h, hipGetLastError());
        HIPCK(h, hipMemcpyAsync(ctl, h->ct_ctl.p, sizeof ctl, hipMemcpyDeviceToHost, h->stream));
        HIPCK(h, sync_stream(h));
        if (ctl[4]) return fail(h, CFGPU_E_STATE, "contract: inconsistent subgraph table");
        uint64_t cur[2];
        std::memcpy(cur, ctl + 8, sizeof cur);
        h->ct_need[0] = cur[0];
        h->ct_need[1] = cur[1];
        if (!ctl[3]) {
            h->n_unitigs = cur[0];
            h->n_ubytes = cur[1];
            h->rs.n_contract_global = wide_cap ? ctl[12] : ctl[6];
            h->rs.n_contract_dense = ctl[2] - ctl[6];
            h->rs.n_contract_wide = wide_cap ? ctl[6] - ctl[12] : 0;
            break;
        }
        if (run == 2) return fail(h, CFGPU_E_OVERFLOW, "contract: output still short after resizing");
        rec_cap = cur[0];
        lab_cap = cur[1];
    }
    HIPCK(h, hipEventRecord(e1.e, h->stream));
    HIPCK(h, hipEventSynchronize(e1.e));
    h->rs.t_contract_ms = elapsed_ms(e0, e1);
    h->rs.n_unitigs = h->n_unitigs;
    h->rs.n_unitig_bytes = h->n_ubytes;
    h->rs.n_contract_runs = run;
    h->contracted = true;
    return CFGPU_OK;
}

int cfgpu_cdbg(cfgpu_handle* h) {
    if (!h) return CFGPU_E_ARG;
    if (!h->contracted) return fail(h, CFGPU_E_STATE, "cdbg before contract");
    HIPCK(h, hipSetDevice(h->prm.device));
    h->cdbg_done = false;
    h->cd_host_seq.clear();
    h->cd_host_rec.clear();
    const uint64_t n = h->n_unitigs;
    const uint32_t KW = h->KW;
    Event e0, e1;
    HIPCK(h, hipEventRecord(e0.e, h->stream));
    uint32_t ctl[16] = {};
    uint32_t rounds = 0;
    if (n) {
        if (2 * n >= (1ull << 30)) return fail(h, CFGPU_E_ARG, "cdbg: %llu unitigs", (unsigned long long)n);
        // pairing bins: about 4096 ends each, at most CD_MAX_BINS
        uint32_t nb_bits = 0;
        while (nb_bits < 13 && (uint64_t(4096) << nb_bits) < 2 * n) ++nb_bits;
        static_assert((1u << 13) == CD_MAX_BINS, "bin bits");
        const uint32_t nb = 1u << nb_bits, per_block = 8192;
        const uint32_t n_blocks = uint32_t((n + per_block - 1) / per_block);
        HIPCK(h, ensure_mem(h, h->cd_partner, 2 * n * 4, Phase::Cdbg));
        HIPCK(h, ensure_mem(h, h->cd_bhist, uint64_t(nb) * n_blocks * 4, Phase::Cdbg));
        HIPCK(h, h->cd_btot.ensure(uint64_t(nb) * 4));
        HIPCK(h, h->cd_boff.ensure((uint64_t(nb) + 1) * 4));
        HIPCK(h, ensure_mem(h, h->cd_bkey, 2 * n * KW * 8, Phase::Cdbg));
        HIPCK(h, ensure_mem(h, h->cd_bid, 2 * n * 4, Phase::Cdbg));
        HIPCK(h, ensure_mem(h, h->cd_kend, 2 * n * KW * 8, Phase::Cdbg));
        HIPCK(h, ensure_mem(h, h->cd_jp, 2 * n * 8, Phase::Cdbg));
        HIPCK(h, h->cd_head.ensure(n * 4));
        HIPCK(h, h->cd_pos.ensure(n * 8));
        HIPCK(h, h->cd_hoff.ensure(n * 8));
        HIPCK(h, ensure_mem(h, h->cd_out, h->n_ubytes, Phase::Cdbg));  // glued labels share k characters: never longer
        HIPCK(h, h->cd_rec.ensure(n * 16));
        HIPCK(h, h->cd_ctl.ensure(64));
        launch_fill_u32(reinterpret_cast<uint32_t*>(h->cd_hoff.p), 0xFFFFFFFFu, 2 * n, h->stream);
        launch_fill_u32(h->cd_ctl.as<uint32_t>(), 0u, 16, h->stream);
        uint32_t* c = h->cd_ctl.as<uint32_t>();
        CdbgArgs a{};
        a.recs = h->ct_recs.as<const UnitigRec>();
        a.n = n;
        a.labels = h->ct_labels.as<const char>();
        a.k = h->prm.k;
        a.partner = h->cd_partner.as<uint32_t>();
        a.bin_hist = h->cd_bhist.as<uint32_t>();
        a.bin_tot = h->cd_btot.as<uint32_t>();
        a.bin_off = h->cd_boff.as<uint32_t>();
        a.bin_key = h->cd_bkey.as<uint64_t>();
        a.bin_id = h->cd_bid.as<uint32_t>();
        a.nb_bits = nb_bits;
        a.n_blocks = n_blocks;
        a.per_block = per_block;
        a.kend = h->cd_kend.as<uint64_t>();
        a.jp = h->cd_jp.as<uint64_t>();
        a.total_chars = h->n_ubytes;
        a.place_head = h->cd_head.as<uint32_t>();
        a.place_pos = h->cd_pos.as<uint64_t>();
        a.piece_len = h->cd_pos.as<uint32_t>();  // n words of the n u64 place_pos words
        a.head_off = h->cd_hoff.as<uint64_t>();
        a.out = h->cd_out.as<char>();
        a.out_cap = h->n_ubytes;
        a.out_rec = h->cd_rec.as<uint64_t>();
        a.out_n_cap = n;
        a.cursors = reinterpret_cast<unsigned long long*>(c + 8);  // [8..11]
        a.retry = c + 2;
        a.error = c + 3;
        a.n_cycle_pieces = c + 4;
        const uint32_t grid = uint32_t(std::min<uint64_t>((2 * n + 255) / 256, uint64_t(h->cus) * 8));
        HIPCK(h, ensure_mem(h, h->cd_settled, ((2 * n + 63) / 64) * 8 * 3, Phase::Cdbg));  // whole words of 64 states
        a.settled = h->cd_settled.as<uint32_t>();
        a.end_bits = h->cd_settled.as<uint64_t>() + (2 * n + 63) / 64;
        a.first_bits = a.end_bits + (2 * n + 63) / 64;
        launch_fill_u32(a.settled, 0u, ((2 * n + 63) / 64) * 2, h->stream);
        launch_cdbg_phase(a, KW, grid, 0, 0, nullptr, h->stream);
        launch_cdbg_phase(a, KW, grid, 1, 0, nullptr, h->stream);
        for (rounds = 0; rounds < 64;) {
            HIPCK(h, hipMemsetAsync(c + 5, 0, 4, h->stream));
            launch_cdbg_phase(a, KW, grid, 2, 0, c + 5, h->stream);
            uint32_t changed = 0;
            HIPCK(h, hipMemcpyAsync(&changed, c + 5, 4, hipMemcpyDeviceToHost, h->stream));
            HIPCK(h, sync_stream(h));
            ++rounds;
            if (!changed) break;
        }
        launch_cdbg_phase(a, KW, grid, 3, 0, nullptr, h->stream);
        launch_cdbg_phase(a, KW, grid, 4, 0, nullptr, h->stream);
        HIPCK(h, hipGetLastError());
        HIPCK(h, hipMemcpyAsync(ctl, c, sizeof ctl, hipMemcpyDeviceToHost, h->stream));
        HIPCK(h, sync_stream(h));
        if (ctl[3]) return fail(h, CFGPU_E_STATE, "cdbg: inconsistent unitig ends (%u)", ctl[3]);
        if (ctl[2]) return fail(h, CFGPU_E_STATE, "cdbg: output longer than the unitig labels");
        if (ctl[4]) {  // cycles through several subgraphs: their pieces to the host
            HIPCK(h, h->cd_cyc.ensure(uint64_t(ctl[4]) * 4));
            HIPCK(h, hipMemsetAsync(c + 6, 0, 4, h->stream));
            launch_cdbg_cycle_list(a, grid, h->cd_cyc.as<uint32_t>(), c + 6, h->stream);
            std::vector<uint32_t> list(ctl[4]);
            HIPCK(h, hipMemcpyAsync(list.data(), h->cd_cyc.p, list.size() * 4, hipMemcpyDeviceToHost, h->stream));
            HIPCK(h, sync_stream(h));
            std::vector<cfgpu_unitig> recs(list.size());
            std::string labs;
            for (size_t i = 0; i < list.size(); ++i) {
                HIPCK(h, hipMemcpy(&recs[i], h->ct_recs.as<cfgpu_unitig>() + list[i], sizeof(cfgpu_unitig),
                                   hipMemcpyDeviceToHost));
                std::string lab(recs[i].label_len, 'A');
                HIPCK(h, hipMemcpy(&lab[0], h->ct_labels.as<char>() + recs[i].label_off, lab.size(),
                                   hipMemcpyDeviceToHost));
                recs[i].label_off = labs.size();
                labs += lab;
            }
            cfhost_cdbg* cc = nullptr;
            if (cfhost_cdbg_build(h->prm.k, recs.data(), recs.size(), labs.data(), labs.size(), 0, &cc) != 0)
                return fail(h, CFGPU_E_STATE, "cdbg: cycles through several subgraphs do not close");
            const char* seq;
            const uint64_t* off;
            uint64_t m;
            cfhost_cdbg_get(cc, &seq, &off, &m);
            h->cd_host_seq.assign(seq, off[m]);
            for (uint64_t i = 0; i < m; ++i) {
                h->cd_host_rec.push_back(off[i]);
                h->cd_host_rec.push_back(off[i + 1] - off[i]);
            }
            cfhost_cdbg_free(cc);
        }
    }
    uint64_t cur[2] = {0, 0};
    std::memcpy(cur, ctl + 8, sizeof cur);
    h->cd_n = cur[0];
    h->cd_bytes = cur[1];
    HIPCK(h, hipEventRecord(e1.e, h->stream));
    HIPCK(h, hipEventSynchronize(e1.e));
    h->rs.t_cdbg_ms = elapsed_ms(e0, e1);
    h->rs.n_cdbg_unitigs = h->cd_n + h->cd_host_rec.size() / 2;
    h->rs.n_cdbg_bytes = h->cd_bytes + h->cd_host_seq.size();
    h->rs.n_cdbg_rounds = rounds;
    h->rs.n_cdbg_cycle_pieces = ctl[4];
    h->cdbg_done = true;
    return CFGPU_OK;
}

int cfgpu_cdbg_copy(cfgpu_handle* h, char* seq, uint64_t seq_cap, uint64_t* rec, uint64_t rec_cap, uint64_t* n,
                    uint64_t* n_bytes) {
    if (!h) return CFGPU_E_ARG;
    if (!h->cdbg_done) return fail(h, CFGPU_E_STATE, "cdbg copy before cdbg");
    const uint64_t tn = h->cd_n + h->cd_host_rec.size() / 2, tb = h->cd_bytes + h->cd_host_seq.size();
    if (n) *n = tn;
    if (n_bytes) *n_bytes = tb;
    if ((seq && seq_cap < tb) || (rec && rec_cap < tn))
        return fail(h, CFGPU_E_ARG, "cdbg copy: capacity %llu / %llu < %llu / %llu", (unsigned long long)seq_cap,
                    (unsigned long long)rec_cap, (unsigned long long)tb, (unsigned long long)tn);
    HIPCK(h, hipSetDevice(h->prm.device));
    if (seq && h->cd_bytes)
        HIPCK(h, hipMemcpyAsync(seq, h->cd_out.p, h->cd_bytes, hipMemcpyDeviceToHost, h->stream));
    if (rec && h->cd_n) HIPCK(h, hipMemcpyAsync(rec, h->cd_rec.p, h->cd_n * 16, hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, sync_stream(h));
    if (seq && !h->cd_host_seq.empty()) std::memcpy(seq + h->cd_bytes, h->cd_host_seq.data(), h->cd_host_seq.size());
    if (rec)
        for (size_t i = 0; i < h->cd_host_rec.size() / 2; ++i) {
            rec[2 * (h->cd_n + i)] = h->cd_bytes + h->cd_host_rec[2 * i];
            rec[2 * (h->cd_n + i) + 1] = h->cd_host_rec[2 * i + 1];
        }
    return CFGPU_OK;
}

int cfgpu_contract_stats(cfgpu_handle* h, cfgpu_contract_stat* out, uint32_t n) {
    if (!h || (n && !out) || n > h->G) return CFGPU_E_ARG;
    if (!h->contracted) return fail(h, CFGPU_E_STATE, "contract stats before contract");
    HIPCK(h, hipSetDevice(h->prm.device));
    if (n) {
        HIPCK(h, hipMemcpyAsync(out, h->ct_counts.p, size_t(n) * sizeof(cfgpu_contract_stat), hipMemcpyDeviceToHost,
                                h->stream));
        HIPCK(h, sync_stream(h));
    }
    return CFGPU_OK;
}

int cfgpu_unitigs_view(cfgpu_handle* h, const cfgpu_unitig** recs, const char** labels, uint64_t* n_unitigs,
                       uint64_t* n_bytes) {
    if (!h) return CFGPU_E_ARG;
    if (!h->contracted) return fail(h, CFGPU_E_STATE, "unitigs before contract");
    if (recs) *recs = h->ct_recs.as<const cfgpu_unitig>();
    if (labels) *labels = h->ct_labels.as<const char>();
    if (n_unitigs) *n_unitigs = h->n_unitigs;
    if (n_bytes) *n_bytes = h->n_ubytes;
    return CFGPU_OK;
}

int cfgpu_unitigs_copy(cfgpu_handle* h, cfgpu_unitig* recs, uint64_t rec_cap, char* labels, uint64_t lab_cap,
                       uint64_t* n_unitigs, uint64_t* n_bytes) {
    if (!h) return CFGPU_E_ARG;
    if (!h->contracted) return fail(h, CFGPU_E_STATE, "unitigs before contract");
    if (n_unitigs) *n_unitigs = h->n_unitigs;
    if (n_bytes) *n_bytes = h->n_ubytes;
    if ((recs && rec_cap < h->n_unitigs) || (labels && lab_cap < h->n_ubytes))
        return fail(h, CFGPU_E_ARG, "unitigs copy: capacity %llu / %llu < %llu / %llu", (unsigned long long)rec_cap,
                    (unsigned long long)lab_cap, (unsigned long long)h->n_unitigs, (unsigned long long)h->n_ubytes);
    HIPCK(h, hipSetDevice(h->prm.device));
    if (recs && h->n_unitigs)
        HIPCK(h, hipMemcpyAsync(recs, h->ct_recs.p, h->n_unitigs * sizeof(cfgpu_unitig), hipMemcpyDeviceToHost,
                                h->stream));
    if (labels && h->n_ubytes)
        HIPCK(h, hipMemcpyAsync(labels, h->ct_labels.p, h->n_ubytes, hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, sync_stream(h));
    return CFGPU_OK;
}

int cfgpu_table_view_get(cfgpu_handle* h, uint32_t g, cfgpu_table_view* out) {
    if (!h || !out || g >= h->G) return CFGPU_E_ARG;
    if (!h->constructed) return fail(h, CFGPU_E_STATE, "table view before construct");
    out->entries = reinterpret_cast<const void*>(h->h_tbl_ptr[g]);
    out->capacity = h->h_tbl_ptr[g] ? h->h_cap[g] : 0;
    out->entry_bytes = entry_bytes(h->KW);
    out->key_words = h->KW;
    out->n_vertices = h->h_nvert[g];
    return CFGPU_OK;
}

int cfgpu_table_copy(cfgpu_handle* h, const uint32_t* buckets, uint32_t n_buckets, uint64_t* keys,
                     uint32_t* e_f, uint8_t* status, uint8_t* dfa, uint32_t* bucket_n, uint64_t cap,
                     uint64_t* n_out) {
    if (!h || (n_buckets && !buckets)) return CFGPU_E_ARG;
    if (!h->constructed) return fail(h, CFGPU_E_STATE, "table copy before construct");
    HIPCK(h, hipSetDevice(h->prm.device));
    std::vector<uint64_t> off(n_buckets + 1, 0);
    std::vector<uint32_t> list;
    for (uint32_t i = 0; i < n_buckets; ++i) {
        if (buckets[i] >= h->G) return fail(h, CFGPU_E_ARG, "table copy: bucket %u out of range", buckets[i]);
        off[i + 1] = off[i] + (h->h_tbl_ptr[buckets[i]] ? h->h_nvert[buckets[i]] : 0);
    }
    const uint64_t total = off[n_buckets];
    if (n_out) *n_out = total;
    if (total > cap) return fail(h, CFGPU_E_ARG, "table copy: capacity %llu < %llu", (unsigned long long)cap,
                                 (unsigned long long)total);
    // only buckets with a table take part in the export grid
    std::vector<uint32_t> ex_b;
    std::vector<uint64_t> ex_o;
    for (uint32_t i = 0; i < n_buckets; ++i)
        if (h->h_tbl_ptr[buckets[i]]) ex_b.push_back(buckets[i]), ex_o.push_back(off[i]);
    std::vector<uint64_t> ex_end(ex_o.size());
    for (size_t i = 0, j = 0; i < n_buckets; ++i)
        if (h->h_tbl_ptr[buckets[i]]) ex_end[j++] = off[i + 1];
    ex_o.push_back(0);  // room for the end offset of the last bucket
    const uint32_t KW = h->KW;
    HIPCK(h, h->ex_keys.ensure(total * KW * 8));
    HIPCK(h, h->ex_ef.ensure(total * 4));
    HIPCK(h, h->ex_st.ensure(total));
    HIPCK(h, h->ex_dfa.ensure(total));
    HIPCK(h, h->ex_off.ensure((ex_b.size() + 1) * 8));
    HIPCK(h, h->ex_cursor.ensure(ex_b.size() * 4));
    HIPCK(h, h->ex_list.ensure(ex_b.size() * 4));
    if (!ex_b.empty()) {
        // out_off[i] = start of bucket i, out_off[i+1] = its end: buckets are
        // exported back to back, so ends are the next starts except the last
        ex_o.back() = ex_end.back();
        for (size_t i = 0; i + 1 < ex_b.size(); ++i)
            if (ex_o[i + 1] != ex_end[i]) return fail(h, CFGPU_E_STATE, "table copy: offsets");
        HIPCK(h, hipMemcpyAsync(h->ex_off.p, ex_o.data(), ex_o.size() * 8, hipMemcpyHostToDevice, h->stream));
        HIPCK(h, hipMemcpyAsync(h->ex_list.p, ex_b.data(), ex_b.size() * 4, hipMemcpyHostToDevice, h->stream));
        launch_fill_u32(h->ex_cursor.as<uint32_t>(), 0u, ex_b.size(), h->stream);
        launch_export(h->d_tbl_ptr.as<uint64_t>(), h->d_cap.as<uint64_t>(), KW, h->ex_list.as<uint32_t>(),
                      uint32_t(ex_b.size()), h->ex_off.as<uint64_t>(), h->ex_cursor.as<uint32_t>(),
                      h->ex_keys.as<uint64_t>(), h->ex_ef.as<uint32_t>(), h->ex_st.as<uint8_t>(),
                      h->ex_dfa.as<uint8_t>(), h->prm.cutoff, h->stream);
        HIPCK(h, hipGetLastError());
    }
    std::vector<uint32_t> cnt(ex_b.size());
    if (!ex_b.empty())
        HIPCK(h, hipMemcpyAsync(cnt.data(), h->ex_cursor.p, cnt.size() * 4, hipMemcpyDeviceToHost, h->stream));
    if (total) {
        if (keys) HIPCK(h, hipMemcpyAsync(keys, h->ex_keys.p, total * KW * 8, hipMemcpyDeviceToHost, h->stream));
        if (e_f) HIPCK(h, hipMemcpyAsync(e_f, h->ex_ef.p, total * 4, hipMemcpyDeviceToHost, h->stream));
        if (status) HIPCK(h, hipMemcpyAsync(status, h->ex_st.p, total, hipMemcpyDeviceToHost, h->stream));
        if (dfa) HIPCK(h, hipMemcpyAsync(dfa, h->ex_dfa.p, total, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCK(h, sync_stream(h));
    size_t j = 0;
    for (uint32_t i = 0; i < n_buckets; ++i) {
        uint32_t c = 0;
        if (h->h_tbl_ptr[buckets[i]]) c = cnt[j++];
        if (c != off[i + 1] - off[i])
            return fail(h, CFGPU_E_STATE, "table copy: bucket %u exported %u of %llu entries", buckets[i], c,
                        (unsigned long long)(off[i + 1] - off[i]));
        if (bucket_n) bucket_n[i] = c;
    }
    return CFGPU_OK;
}

int cfgpu_table_copy_noalloc(cfgpu_handle* h, const uint32_t* buckets, uint32_t n_buckets, uint64_t* keys,
                             uint32_t* e_f, uint8_t* status, uint8_t* dfa, uint32_t* bucket_n, uint64_t cap,
                             uint64_t* n_out) {
    if (!h || (n_buckets && !buckets)) return CFGPU_E_ARG;
    if (!h->constructed) return fail(h, CFGPU_E_STATE, "table copy before construct");
    if (n_buckets > NOALLOC_MAX_BUCKETS)
        return fail(h, CFGPU_E_ARG, "table copy (no alloc): %u buckets, at most %u", n_buckets, NOALLOC_MAX_BUCKETS);
    if (!keys || !e_f || !status || !dfa) return fail(h, CFGPU_E_ARG, "table copy (no alloc): null output");
    HIPCK(h, hipSetDevice(h->prm.device));
    // offsets of the listed buckets that have a table, back to back
    uint64_t total = 0;
    uint32_t m = 0;
    for (uint32_t i = 0; i < n_buckets; ++i) {
        const uint32_t b = buckets[i];
        if (b >= h->G) return fail(h, CFGPU_E_ARG, "table copy: bucket %u out of range", b);
        if (!h->h_tbl_ptr[b]) continue;
        h->na_list[m] = b;
        h->na_off[m] = total;
        total += h->h_nvert[b];
        ++m;
    }
    h->na_off[m] = total;
    if (n_out) *n_out = total;
    if (total > cap) return fail(h, CFGPU_E_ARG, "table copy: capacity %llu < %llu", (unsigned long long)cap,
                                 (unsigned long long)total);
    uint64_t* d_off = reinterpret_cast<uint64_t*>(h->noalloc_scratch);
    uint32_t* d_cursor = reinterpret_cast<uint32_t*>(h->noalloc_scratch + NOALLOC_CURSOR_AT);
    uint32_t* d_list = reinterpret_cast<uint32_t*>(h->noalloc_scratch + NOALLOC_LIST_AT);
    if (m) {
        HIPCK(h, hipMemcpyAsync(d_off, h->na_off.data(), (m + 1) * 8ull, hipMemcpyHostToDevice, h->stream));
        HIPCK(h, hipMemcpyAsync(d_list, h->na_list.data(), m * 4ull, hipMemcpyHostToDevice, h->stream));
        launch_fill_u32(d_cursor, 0u, m, h->stream);  // a fill kernel: the cursors start at zero
        launch_export(h->d_tbl_ptr.as<uint64_t>(), h->d_cap.as<uint64_t>(), h->KW, d_list, m, d_off, d_cursor, keys,
                      e_f, status, dfa, h->prm.cutoff, h->stream);
        HIPCK(h, hipGetLastError());
        HIPCK(h, hipMemcpyAsync(h->na_cnt.data(), d_cursor, m * 4ull, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCK(h, hipStreamSynchronize(h->stream));
    for (uint32_t i = 0, j = 0; i < n_buckets; ++i) {
        const uint32_t b = buckets[i];
        uint32_t c = 0;
        if (h->h_tbl_ptr[b]) {
            c = h->na_cnt[j];
            if (c != h->na_off[j + 1] - h->na_off[j])
                return fail(h, CFGPU_E_STATE, "table copy: bucket %u exported %u of %llu entries", b, c,
                            (unsigned long long)(h->na_off[j + 1] - h->na_off[j]));
            ++j;
        }
        if (bucket_n) bucket_n[i] = c;
    }
    return CFGPU_OK;
}

int cfgpu_release(cfgpu_handle* h, uint32_t g) {
    if (!h || g >= h->G) return CFGPU_E_ARG;
    if (!h->constructed) return fail(h, CFGPU_E_STATE, "release before construct");
    // The table memory belongs to a pool shared with other buckets and is
    // reclaimed at the next construct; the view of g stops being valid.
    h->h_tbl_ptr[g] = 0;
    h->h_flags[g] &= ~uint32_t(CFGPU_BUCKET_BUILT);
    return CFGPU_OK;
}

int cfgpu_sizing_model_get(cfgpu_handle* h, cfgpu_sizing_model* info, float* params, float* features,
                           uint32_t* train_buckets, float* train_targets, float* predictions) {
    if (!h) return CFGPU_E_ARG;
    if (!h->mlp_trained) return fail(h, CFGPU_E_STATE, "sizing model: no MLP trained since the last partition");
    HIPCK(h, hipSetDevice(h->prm.device));
    const uint32_t n_train = uint32_t(h->h_mlp_idx.size());
    if (info) {
        info->n_features = N_FEAT;
        info->hidden = MLP_H;
        info->n_params = uint32_t(mlp_param_floats());
        info->n_train = n_train;
        info->epochs = h->mlp_epochs_last;
        info->seed = MLP_SEED;
        info->lr = h->mlp_lr_last;
        info->loss = h->mlp_loss_last;
    }
    if (params) HIPCK(h, hipMemcpyAsync(params, h->mlp_params.p, size_t(mlp_param_floats()) * 4, hipMemcpyDeviceToHost,
                                        h->stream));
    if (features) HIPCK(h, hipMemcpyAsync(features, h->feat.p, size_t(h->G) * N_FEAT * 4, hipMemcpyDeviceToHost,
                                          h->stream));
    if (predictions) HIPCK(h, hipMemcpyAsync(predictions, h->mlp_pred.p, size_t(h->G) * 4, hipMemcpyDeviceToHost,
                                             h->stream));
    HIPCK(h, sync_stream(h));
    if (train_buckets) std::memcpy(train_buckets, h->h_mlp_idx.data(), n_train * 4ull);
    if (train_targets) std::memcpy(train_targets, h->h_mlp_target.data(), n_train * 4ull);
    return CFGPU_OK;
}

int cfgpu_run_stats_get(cfgpu_handle* h, cfgpu_run_stats* out) {
    if (!h || !out) return CFGPU_E_ARG;
    *out = h->rs;
    return CFGPU_OK;
}

int cfgpu_tuning_get(cfgpu_handle* h, cfgpu_tuning* out) {
    if (!h || !out) return CFGPU_E_ARG;
    *out = h->tune;
    return CFGPU_OK;
}

int cfgpu_tuning_set(cfgpu_handle* h, const cfgpu_tuning* in) {
    if (!h || !in) return CFGPU_E_ARG;
    if (!(in->load_factor >= 0.2f && in->load_factor <= 0.85f) ||
        (in->workgroup != 256 && in->workgroup != 512) || in->chunk_records < 64 ||
        in->chunk_records > (1u << 24) || in->wgs_per_cu < 1 || in->wgs_per_cu > 32 ||
        !(in->margin >= 0.01f && in->margin <= 4.0f) || in->use_graph > 1 || in->max_passes > 4096 ||
        in->predictor > 2 || in->dedup > 2 || in->mlp_epochs < 1 || in->mlp_epochs > 10000 ||
        !(in->mlp_lr > 0.f && in->mlp_lr <= 1.f) || !(in->lds_load >= 0.5f && in->lds_load <= 0.85f) ||
        in->contract_wide > 1)
        return fail(h, CFGPU_E_ARG, "tuning out of range");
    h->tune = *in;
    return CFGPU_OK;
}

int cfgpu_autotune(cfgpu_handle* h, uint32_t mode, cfgpu_tune_trial* trials_out, uint32_t n_trials,
                   uint32_t* n_written) {
    if (!h) return CFGPU_E_ARG;
    if (!h->partitioned) return fail(h, CFGPU_E_STATE, "autotune before partition");
    const cfgpu_tuning base = h->tune;
    int r = construct_impl(h, mode);  // its vertex counts are the ones every trial must give
    if (r != CFGPU_OK) return r;
    const std::vector<uint32_t> nv0 = h->h_nvert;
    // the learned mode's sample buckets take the global kernel whatever the
    // tuning; its shape is tuned only when main-round buckets take it too
    const bool global_used = h->rs.n_lds_buckets + h->rs.n_samples < h->rs.n_buckets_nonempty;
    cfgpu_tuning best = base;
    double best_ms = 1e300;
    uint32_t w = 0;
    auto trial = [&](const std::vector<cfgpu_tuning>& grid) {
        for (const cfgpu_tuning& t : grid) {
            h->tune = t;
            // an untimed run first: a tuning that needs larger pools than the
            // last one pays for their allocation once, not in its score
            int rr = construct_impl(h, mode);
            if (rr == CFGPU_OK) rr = construct_impl(h, mode);
            if (rr != CFGPU_OK) {
                h->tune = base;
                return rr;
            }
            const double ms = h->rs.t_construct_ms + h->rs.t_dedup_ms;
            const uint32_t same = h->h_nvert == nv0 ? 1u : 0u;
            if (trials_out && w < n_trials) trials_out[w++] = {t, ms, same, 0};
            if (!same) {
                h->tune = base;
                return fail(h, CFGPU_E_STATE, "autotune: a tuning changed the vertex counts");
            }
            if (ms < best_ms) best_ms = ms, best = t;
        }
        return int(CFGPU_OK);
    };
    // stage 1: record dedup as the auto rule decides it, and always on
    std::vector<cfgpu_tuning> grid = {base};
    if (base.dedup == 1) {
        cfgpu_tuning t = base;
        t.dedup = 2;
        grid.push_back(t);
    }
    CK(trial(grid));
    // stage 2: LDS pass-table load x sizing margin (+ the global kernel's
    // shape when some buckets take it) around the stage-1 winner
    const cfgpu_tuning s1 = best;
    grid.clear();
    for (float ll : {0.7f, 0.8f, 0.85f})
        for (float mg : {1.05f, 1.15f, 1.3f}) {
            cfgpu_tuning t = s1;
            t.lds_load = ll, t.margin = mg;
            grid.push_back(t);
        }
    if (global_used)
        for (float lf : {0.5f, 0.65f})
            for (uint32_t wg : {256u, 512u})
                for (uint32_t ch : {1024u, 4096u})
                    for (uint32_t wpc : {4u, 8u}) {
                        cfgpu_tuning t = s1;
                        t.load_factor = lf, t.workgroup = wg, t.chunk_records = ch, t.wgs_per_cu = wpc;
                        grid.push_back(t);
                    }
    CK(trial(grid));
    if (n_written) *n_written = w;
    h->tune = best;
    return construct_impl(h, mode);
}

}  // extern "C"
