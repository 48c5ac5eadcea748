// Chunked ingest overlapped with the device partition: the text is cut into
// chunks of whole records (each chunk's end found while it is parsed: the
// first record start past chunk_bytes); while the device copies chunk i from pinned host
// memory and partitions it (cfgpu_partition with CFGPU_PARTITION_APPEND), the
// host threads parse and pack chunk i+1 into the other pinned buffer pair.
// cfgpu_partition_finish then gathers the chunks' buckets.  The reference reads
// its input in chunks on parallel workers the same way and finalizes the
// buckets once every chunk is in (src/Graph_Partitioner.cpp:121-184, 424-614;
// include/Subgraphs_Manager.hpp:113).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <hip/hip_runtime_api.h>
#include <thread>
#include <vector>

#include "../../include/cfhost.h"
#include "ingest_core.hpp"

using namespace cfhost_detail;

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Page-locked host words, grown on demand (DMA copies run from them at full
// link rate, and asynchronously to the host).
struct Pinned {
    uint64_t* p = nullptr;
    size_t cap = 0;
    Pinned() = default;
    Pinned(const Pinned&) = delete;
    Pinned& operator=(const Pinned&) = delete;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
    bool ensure(size_t words) {
        if (words <= cap) return true;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (hipHostMalloc(reinterpret_cast<void**>(&p), words * 8, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return false;
        }
        cap = words;
        return true;
    }
};

// One chunk parsed and packed: the records of [b, end).
struct Packed {
    Pinned packed, frag;
    uint64_t n_bases = 0, n_frag = 0, n_rec = 0, end = 0;
    double ms = 0;
    int err = 0;
};

// The chunk from b (a record start) to the first record start at or after
// b + chunk_bytes, found while it is parsed.
void pack_chunk(const char* buf, uint64_t n, int is_fastq, uint64_t b, uint64_t chunk_bytes, uint32_t k, uint32_t T,
                Packed& c) {
    const double t0 = now_ms();
    const std::vector<Seg> segs = parse_chunk(buf, n, is_fastq, b, b + chunk_bytes, T, &c.end);
    Packer pk;
    pk.plan(segs, k);
    c.err = 0;
    if (!c.packed.ensure(pk.n_words()) || !c.frag.ensure(pk.n_frag() + 1)) {
        c.err = -3;
        return;
    }
    pk.write(c.packed.p, c.frag.p);
    c.n_bases = pk.n_bases();
    c.n_frag = pk.n_frag();
    c.n_rec = 0;
    for (const Seg& s : segs) c.n_rec += s.rec.size();
    c.ms = now_ms() - t0;
}

}  // namespace

extern "C" {

int cfhost_partition_buffer_ex(cfgpu_handle* h, const char* buf, uint64_t n, int is_fastq, uint32_t threads,
                               uint64_t chunk_bytes, uint32_t flags, cfhost_ingest_stats* st) {
    if (!h || (!buf && n) || (flags & ~uint32_t(CFGPU_PARTITION_SPILL))) return -1;
    cfgpu_params prm{};
    if (cfgpu_params_get(h, &prm) != CFGPU_OK) return -1;
    const uint32_t T = pick_threads(threads);
    if (!chunk_bytes) chunk_bytes = uint64_t(1) << 30;
    const double t0 = now_ms();
    Packed c[2];
    pack_chunk(buf, n, is_fastq, 0, chunk_bytes, prm.k, T, c[0]);
    double t_host = c[0].ms, t_dev = 0;
    uint64_t n_rec = 0, n_bases = 0, n_frag = 0, m = 0;
    for (uint64_t i = 0;; ++i) {
        Packed& cur = c[i & 1];
        Packed& nxt = c[(i + 1) & 1];
        if (cur.err) return cur.err;
        const bool more = cur.end < n;
        std::thread th;
        if (more) th = std::thread([&]() { pack_chunk(buf, n, is_fastq, cur.end, chunk_bytes, prm.k, T, nxt); });
        const double d0 = now_ms();
        const int rc =
            cfgpu_partition(h, cur.packed.p, cur.n_bases, cur.frag.p, cur.n_frag, CFGPU_PARTITION_APPEND | flags);
        t_dev += now_ms() - d0;
        if (th.joinable()) th.join();
        if (rc != CFGPU_OK) return rc;
        n_rec += cur.n_rec;
        n_bases += cur.n_bases;
        n_frag += cur.n_frag;
        ++m;
        if (!more) break;
        t_host += nxt.ms;
    }
    const double f0 = now_ms();
    const int rc = cfgpu_partition_finish(h);
    const double f1 = now_ms();
    if (rc != CFGPU_OK) return rc;
    if (st) {
        st->n_bytes = n;
        st->n_records = n_rec;
        st->n_bases = n_bases;
        st->n_fragments = n_frag;
        st->n_chunks = uint32_t(m);
        st->threads = T;
        st->t_total_ms = f1 - t0;
        st->t_host_ms = t_host;
        st->t_device_ms = t_dev;
        st->t_finish_ms = f1 - f0;
    }
    return 0;
}

int cfhost_partition_buffer(cfgpu_handle* h, const char* buf, uint64_t n, int is_fastq, uint32_t threads,
                            uint64_t chunk_bytes, cfhost_ingest_stats* st) {
    return cfhost_partition_buffer_ex(h, buf, n, is_fastq, threads, chunk_bytes, 0, st);
}

int cfhost_partition_file_ex(cfgpu_handle* h, const char* path, int is_fastq, uint32_t threads, uint64_t chunk_bytes,
                             uint32_t flags, cfhost_ingest_stats* st) {
    if (!h || !path) return -1;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return -2;
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        close(fd);
        return -2;
    }
    const uint64_t n = uint64_t(sb.st_size);
    const char* buf = nullptr;
    void* map = nullptr;
    if (n) {
        // pages are read as the parser reaches them, during the device's work on earlier chunks
        map = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (map == MAP_FAILED) {
            close(fd);
            return -2;
        }
        (void)madvise(map, n, MADV_SEQUENTIAL);
        buf = static_cast<const char*>(map);
    }
    const int rc = cfhost_partition_buffer_ex(h, buf, n, is_fastq, threads, chunk_bytes, flags, st);
    if (map) munmap(map, n);
    close(fd);
    return rc;
}

int cfhost_partition_file(cfgpu_handle* h, const char* path, int is_fastq, uint32_t threads, uint64_t chunk_bytes,
                          cfhost_ingest_stats* st) {
    return cfhost_partition_file_ex(h, path, is_fastq, threads, chunk_bytes, 0, st);
}

}  // extern "C"
