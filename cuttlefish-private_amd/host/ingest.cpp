// Host ingest: FASTA/FASTQ records -> fragments -> 2-bit MSB-first packed stream
// (the input format of cfgpu_partition).  See include/cfhost.h for the rules
// and the reference lines they follow.
#include "../../include/cfhost.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

struct cfhost_input {
    std::vector<uint64_t> packed;   // n_words = ceil(n_bases/32) + 2
    std::vector<uint64_t> frag_off; // n_frag + 1
    uint64_t n_bases = 0, n_records = 0;
};

namespace {

// DNA_Utility::is_DNA_base (include/DNA_Utility.hpp:167-172) and the 2-bit code
// of include/DNA.hpp:17-25; 0xFF marks a non-DNA byte.
struct CodeLut {
    uint8_t c[256];
    CodeLut() {
        for (int i = 0; i < 256; ++i) {
            switch (i & 0x5F) {
                case 'A': c[i] = 0; break;
                case 'C': c[i] = 1; break;
                case 'G': c[i] = 2; break;
                case 'T': c[i] = 3; break;
                default: c[i] = 0xFF;
            }
        }
    }
};
const CodeLut LUT;

struct Frag {
    uint64_t beg, len;  // in the record byte buffer
};

// Maximal DNA runs of length >= k+1 in [b, e) (src/Graph_Partitioner.cpp:476-501).
void find_fragments(const char* s, uint64_t b, uint64_t e, uint32_t k, std::vector<Frag>& out) {
    uint64_t i = b;
    while (i < e) {
        while (i < e && LUT.c[uint8_t(s[i])] == 0xFF) ++i;
        const uint64_t st = i;
        while (i < e && LUT.c[uint8_t(s[i])] != 0xFF) ++i;
        if (i - st >= uint64_t(k) + 1) out.push_back({st, i - st});
    }
}

uint32_t pick_threads(uint32_t t) {
    if (t) return t;
    const unsigned hw = std::thread::hardware_concurrency();
    return hw ? std::min(hw, 16u) : 1u;
}

template <class F>
void parallel(uint32_t T, F&& f) {
    if (T <= 1) {
        f(0u);
        return;
    }
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < T; ++t) th.emplace_back(f, t);
    for (auto& x : th) x.join();
}

int build(const char* seq, const uint64_t* rec_off, uint64_t n_rec, uint32_t k, uint32_t threads,
          cfhost_input** out) {
    if (!out || (n_rec && (!seq || !rec_off)) || k < 3 || k > 127) return -1;
    const uint32_t T = std::max<uint32_t>(1, std::min<uint64_t>(pick_threads(threads), std::max<uint64_t>(1, n_rec)));
    std::unique_ptr<cfhost_input> in(new cfhost_input());
    in->n_records = n_rec;
    // record ranges per thread, balanced by bytes
    const uint64_t total = n_rec ? rec_off[n_rec] - rec_off[0] : 0;
    std::vector<uint64_t> rb(T + 1, n_rec);
    rb[0] = 0;
    for (uint32_t t = 1; t < T; ++t) {
        const uint64_t target = rec_off[0] + total * t / T;
        rb[t] = uint64_t(std::lower_bound(rec_off, rec_off + n_rec, target) - rec_off);
        rb[t] = std::max(rb[t], rb[t - 1]);
    }
    std::vector<std::vector<Frag>> fr(T);
    parallel(T, [&](uint32_t t) {
        for (uint64_t r = rb[t]; r < rb[t + 1]; ++r) find_fragments(seq, rec_off[r], rec_off[r + 1], k, fr[t]);
    });
    std::vector<uint64_t> tb(T + 1, 0), tf(T + 1, 0);
    for (uint32_t t = 0; t < T; ++t) {
        uint64_t s = 0;
        for (const Frag& f : fr[t]) s += f.len;
        tb[t + 1] = tb[t] + s;
        tf[t + 1] = tf[t] + fr[t].size();
    }
    in->n_bases = tb[T];
    in->packed.assign((in->n_bases + 31) / 32 + 2, 0);
    in->frag_off.resize(tf[T] + 1);
    in->frag_off[tf[T]] = in->n_bases;
    uint64_t* P = in->packed.data();
    parallel(T, [&](uint32_t t) {
        uint64_t pos = tb[t], fi = tf[t];
        const uint64_t first_w = tb[t] / 32, last_w = tb[t + 1] ? (tb[t + 1] - 1) / 32 : 0;
        uint64_t cur = 0, cur_w = first_w;
        auto flush = [&](uint64_t w, uint64_t v) {
            // words shared with a neighbouring thread's range are merged with OR
            if (w == first_w || w == last_w) __atomic_fetch_or(&P[w], v, __ATOMIC_RELAXED);
            else P[w] = v;
        };
        for (const Frag& f : fr[t]) {
            in->frag_off[fi++] = pos;
            const char* s = seq + f.beg;
            for (uint64_t i = 0; i < f.len; ++i, ++pos) {
                const uint64_t w = pos >> 5;
                if (w != cur_w) {
                    flush(cur_w, cur);
                    cur = 0;
                    cur_w = w;
                }
                cur |= uint64_t(LUT.c[uint8_t(s[i])]) << (62 - 2 * (pos & 31));
            }
        }
        if (tb[t + 1] > tb[t]) flush(cur_w, cur);
    });
    *out = in.release();
    return 0;
}

// FASTA / FASTQ -> concatenated record bytes + offsets.
void parse(const char* buf, uint64_t n, int is_fastq, std::vector<char>& seq,
           std::vector<uint64_t>& off) {
    seq.clear();
    off.assign(1, 0);
    seq.reserve(n);
    uint64_t pos = 0, line = 0;
    bool have = false;
    while (pos < n) {
        const void* nl = std::memchr(buf + pos, '\n', n - pos);
        const uint64_t end = nl ? uint64_t(static_cast<const char*>(nl) - buf) : n;
        const char* s = buf + pos;
        const uint64_t len = end - pos;
        if (is_fastq) {
            if (line % 4 == 1) {
                seq.insert(seq.end(), s, s + len);
                off.push_back(seq.size());
            }
        } else if (len > 0 && s[0] == '>') {
            if (have) off.push_back(seq.size());
            have = true;
        } else if (have) {
            seq.insert(seq.end(), s, s + len);
        }
        ++line;
        pos = end + 1;
    }
    if (!is_fastq && have) off.push_back(seq.size());
}

}  // namespace

extern "C" {

int cfhost_input_from_records(const char* seq, const uint64_t* rec_off, uint64_t n_rec, uint32_t k,
                              uint32_t threads, cfhost_input** out) {
    return build(seq, rec_off, n_rec, k, threads, out);
}

int cfhost_input_from_buffer(const char* buf, uint64_t n, int is_fastq, uint32_t k, uint32_t threads,
                             cfhost_input** out) {
    if (!buf && n) return -1;
    std::vector<char> seq;
    std::vector<uint64_t> off;
    parse(buf, n, is_fastq, seq, off);
    return build(seq.data(), off.data(), off.size() - 1, k, threads, out);
}

int cfhost_input_from_file(const char* path, int is_fastq, uint32_t k, uint32_t threads,
                           cfhost_input** out) {
    if (!path) return -1;
    FILE* f = std::fopen(path, "rb");
    if (!f) return -2;
    std::vector<char> buf;
    char tmp[1 << 16];
    size_t r;
    while ((r = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + r);
    const bool err = std::ferror(f);
    std::fclose(f);
    if (err) return -2;
    return cfhost_input_from_buffer(buf.data(), buf.size(), is_fastq, k, threads, out);
}

int cfhost_input_get(const cfhost_input* in, const uint64_t** packed, uint64_t* n_words,
                     uint64_t* n_bases, const uint64_t** frag_off, uint64_t* n_frag,
                     uint64_t* n_records) {
    if (!in) return -1;
    if (packed) *packed = in->packed.data();
    if (n_words) *n_words = in->packed.size();
    if (n_bases) *n_bases = in->n_bases;
    if (frag_off) *frag_off = in->frag_off.data();
    if (n_frag) *n_frag = in->frag_off.size() - 1;
    if (n_records) *n_records = in->n_records;
    return 0;
}

void cfhost_input_free(cfhost_input* in) { delete in; }

}  // extern "C"
