// Host ingest: FASTA/FASTQ records -> fragments -> 2-bit MSB-first packed stream
// (the input format of cfgpu_partition).  See include/cfhost.h for the rules
// and the reference lines they follow.
#include "ingest_core.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/cfhost.h"

struct cfhost_input {
    std::vector<uint64_t> packed;   // n_words = ceil(n_bases/32) + 2
    std::vector<uint64_t> frag_off; // n_frag + 1
    uint64_t n_bases = 0, n_records = 0;
};

namespace cfhost_detail {
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

// End of the line starting at p ('\n' position, or n).
uint64_t line_end(const char* buf, uint64_t n, uint64_t p) {
    const void* q = std::memchr(buf + p, '\n', n - p);
    return q ? uint64_t(static_cast<const char*>(q) - buf) : n;
}

// First line start at or after p (n when none).
uint64_t line_start_from(const char* buf, uint64_t n, uint64_t p) {
    if (p == 0) return 0;
    if (p >= n) return n;
    if (buf[p - 1] == '\n') return p;
    const uint64_t e = line_end(buf, n, p);
    return e < n ? e + 1 : n;
}

// First '>' line starting at or after p (n when none).
uint64_t header_from(const char* buf, uint64_t n, uint64_t p) {
    p = line_start_from(buf, n, p);
    if (p < n && buf[p] == '>') return p;
    while (p < n) {  // p is a line start holding no '>', or just past a '>' that starts no line
        const void* q = std::memchr(buf + p, '>', n - p);
        if (!q) return n;
        const uint64_t x = uint64_t(static_cast<const char*>(q) - buf);  // x > 0
        if (buf[x - 1] == '\n') return x;
        p = x + 1;
    }
    return n;
}

// FASTQ records whose lines start in [b, e); `li` is the number of the first
// line starting at or after b (lines counted from the text's start).
void fastq_range(const char* buf, uint64_t n, uint64_t b, uint64_t e, uint64_t nl_before, Seg& sg) {
    uint64_t p = line_start_from(buf, n, b);
    uint64_t li = nl_before + (p > b ? 1 : 0);  // one '\n' in [b, p) when p moved
    sg.seq = buf;
    while (p < e && p < n) {
        const uint64_t end = line_end(buf, n, p);
        if (li % 4 == 1) sg.rec.emplace_back(p, end);
        ++li;
        p = end + 1;
    }
}

// FASTA records whose '>' line starts in [b, e), each running to the next
// '>' line (wherever it starts) or the end of the text.
void fasta_range(const char* buf, uint64_t n, uint64_t b, uint64_t e, Seg& sg) {
    uint64_t p = header_from(buf, n, b);
    bool have = false;
    uint64_t start = 0;
    while (p < n) {
        const uint64_t end = line_end(buf, n, p);
        if (end > p && buf[p] == '>') {
            if (p >= e) break;  // a later segment's record
            if (have) sg.rec.emplace_back(start, sg.own.size());
            have = true;
            start = sg.own.size();
        } else if (have) {
            sg.own.insert(sg.own.end(), buf + p, buf + end);
        }
        p = end + 1;
    }
    if (have) sg.rec.emplace_back(start, sg.own.size());
    sg.seq = sg.own.data();
}

// Maximal DNA runs of length >= k+1 in s[b, e) (src/Graph_Partitioner.cpp:476-501).
// Runs are crossed 8 bytes per branch: the OR of 8 codes has bit 7 set when
// any of the 8 bytes is not a base.
template <class V>
void find_fragments(const char* s, uint64_t b, uint64_t e, uint32_t k, V& out) {
    const uint8_t* c = LUT.c;
    uint64_t i = b;
    while (i < e) {
        while (i < e && c[uint8_t(s[i])] == 0xFF) ++i;
        const uint64_t st = i;
        while (i + 8 <= e &&
               !((c[uint8_t(s[i])] | c[uint8_t(s[i + 1])] | c[uint8_t(s[i + 2])] | c[uint8_t(s[i + 3])] |
                  c[uint8_t(s[i + 4])] | c[uint8_t(s[i + 5])] | c[uint8_t(s[i + 6])] | c[uint8_t(s[i + 7])]) &
                 0x80))
            i += 8;
        while (i < e && c[uint8_t(s[i])] != 0xFF) ++i;
        if (i - st >= uint64_t(k) + 1) out.push_back({st, i - st});
    }
}

// 2-bit codes of 8 bases (A/C/G/T in either case) at s, the first in the top
// two bits: ((c >> 1) ^ (c >> 2)) & 3 maps A C G T (and a c g t) to 0 1 2 3.
inline uint32_t codes8(const char* s) {
    uint64_t v;
    std::memcpy(&v, s, 8);
    v = __builtin_bswap64(v);  // the first base in the top byte
    uint64_t c = ((v >> 1) ^ (v >> 2)) & 0x0303030303030303ull;
    c = (c | (c >> 6)) & 0x000F000F000F000Full;
    c = (c | (c >> 12)) & 0x000000FF000000FFull;
    return uint32_t((c | (c >> 24)) & 0xFFFFu);
}

}  // namespace

uint32_t pick_threads(uint32_t t) {
    if (t) return t;
    const unsigned hw = std::thread::hardware_concurrency();
    return hw ? std::min(hw, 16u) : 1u;
}

std::vector<Seg> parse_chunk(const char* buf, uint64_t n, int is_fastq, uint64_t b, uint64_t x, uint32_t T,
                             uint64_t* end) {
    x = std::min(std::max(x, b), n);
    T = std::max<uint32_t>(1, std::min<uint64_t>(T, std::max<uint64_t>(1, (x - b) >> 16)));  // >= 64 KiB each
    std::vector<uint64_t> r(T + 1);
    for (uint32_t t = 0; t <= T; ++t) r[t] = b + (x - b) * t / T;
    std::vector<Seg> segs(T);
    uint64_t e = n;
    if (is_fastq) {
        // line numbers from b (a record start: its line number is a multiple of 4)
        std::vector<uint64_t> nl(T + 1, 0);
        parallel(T, [&](uint32_t t) { nl[t + 1] = uint64_t(std::count(buf + r[t], buf + r[t + 1], '\n')); });
        for (uint32_t t = 0; t < T; ++t) nl[t + 1] += nl[t];
        // the chunk ends at the first record start at or after x
        e = line_start_from(buf, n, x);
        for (uint64_t li = nl[T] + (e > x ? 1 : 0); e < n && li % 4 != 0; ++li) e = line_end(buf, n, e) + 1;
        e = std::min(e, n);
        r[T] = e;  // the last segment takes the lines from x to the boundary
        parallel(T, [&](uint32_t t) { fastq_range(buf, n, r[t], r[t + 1], nl[t], segs[t]); });
    } else {
        e = header_from(buf, n, x);  // records whose '>' line starts before x run up to it
        parallel(T, [&](uint32_t t) { fasta_range(buf, n, r[t], r[t + 1], segs[t]); });
    }
    if (end) *end = e;
    return segs;
}

std::vector<Seg> parse_text(const char* buf, uint64_t n, int is_fastq, uint32_t T) {
    return parse_chunk(buf, n, is_fastq, 0, n, T, nullptr);
}

std::vector<Seg> segs_of_records(const char* seq, const uint64_t* rec_off, uint64_t n_rec, uint32_t T) {
    T = std::max<uint32_t>(1, std::min<uint64_t>(T, std::max<uint64_t>(1, n_rec)));
    const uint64_t total = n_rec ? rec_off[n_rec] - rec_off[0] : 0;
    std::vector<uint64_t> rb(T + 1, n_rec);
    rb[0] = 0;
    for (uint32_t t = 1; t < T; ++t) {
        const uint64_t target = rec_off[0] + total * t / T;
        rb[t] = uint64_t(std::lower_bound(rec_off, rec_off + n_rec, target) - rec_off);
        rb[t] = std::max(rb[t], rb[t - 1]);
    }
    std::vector<Seg> segs(T);
    for (uint32_t t = 0; t < T; ++t) {
        segs[t].seq = seq;
        for (uint64_t i = rb[t]; i < rb[t + 1]; ++i) segs[t].rec.emplace_back(rec_off[i], rec_off[i + 1]);
    }
    return segs;
}

void Packer::plan(const std::vector<Seg>& segs, uint32_t k) {
    segs_ = &segs;
    const uint32_t T = uint32_t(segs.size());
    fr_.assign(T, {});
    parallel(T, [&](uint32_t t) {
        for (const auto& r : segs[t].rec) find_fragments(segs[t].seq, r.first, r.second, k, fr_[t]);
    });
    tb_.assign(T + 1, 0);
    tf_.assign(T + 1, 0);
    for (uint32_t t = 0; t < T; ++t) {
        uint64_t s = 0;
        for (const Frag& f : fr_[t]) s += f.len;
        tb_[t + 1] = tb_[t] + s;
        tf_[t + 1] = tf_[t] + fr_[t].size();
    }
}

void Packer::write(uint64_t* P, uint64_t* frag_off) const {
    const std::vector<Seg>& segs = *segs_;
    const uint32_t T = uint32_t(segs.size());
    const uint64_t nb = n_bases(), nw = n_words();
    // words shared by two segments are ORed into: zero them (and the two pad words) first
    P[nw - 1] = P[nw - 2] = 0;
    for (uint32_t t = 0; t < T; ++t)
        if (tb_[t + 1] > tb_[t]) P[tb_[t] / 32] = P[(tb_[t + 1] - 1) / 32] = 0;
    frag_off[tf_[T]] = nb;
    parallel(T, [&](uint32_t t) {
        if (tb_[t + 1] == tb_[t]) return;
        uint64_t pos = tb_[t], fi = tf_[t];
        const uint64_t first_w = tb_[t] / 32, last_w = (tb_[t + 1] - 1) / 32;
        uint64_t cur = 0, cur_w = first_w;
        auto flush = [&](uint64_t w, uint64_t v) {
            if (w == first_w || w == last_w) __atomic_fetch_or(&P[w], v, __ATOMIC_RELAXED);
            else P[w] = v;
        };
        for (const Frag& f : fr_[t]) {
            frag_off[fi++] = pos;
            const char* s = segs[t].seq + f.beg;
            uint64_t i = 0;
            // one base at a time up to a multiple of 8 in the word, then 8 at a time
            for (; i < f.len && ((pos & 7) || i + 8 > f.len); ++i, ++pos) {
                const uint64_t w = pos >> 5;
                if (w != cur_w) {
                    flush(cur_w, cur);
                    cur = 0;
                    cur_w = w;
                }
                cur |= uint64_t(LUT.c[uint8_t(s[i])]) << (62 - 2 * (pos & 31));
            }
            for (; i + 8 <= f.len; i += 8, pos += 8) {
                const uint64_t w = pos >> 5;
                if (w != cur_w) {
                    flush(cur_w, cur);
                    cur = 0;
                    cur_w = w;
                }
                cur |= uint64_t(codes8(s + i)) << (48 - 2 * (pos & 31));
            }
            for (; i < f.len; ++i, ++pos) {
                const uint64_t w = pos >> 5;
                if (w != cur_w) {
                    flush(cur_w, cur);
                    cur = 0;
                    cur_w = w;
                }
                cur |= uint64_t(LUT.c[uint8_t(s[i])]) << (62 - 2 * (pos & 31));
            }
        }
        flush(cur_w, cur);
    });
}

std::vector<uint64_t> split_records(const char* buf, uint64_t n, int is_fastq, uint64_t chunk_bytes, uint32_t T) {
    std::vector<uint64_t> off{0};
    if (n == 0) {
        off.push_back(0);
        return off;
    }
    chunk_bytes = std::max<uint64_t>(chunk_bytes, 1);
    const uint64_t m = (n + chunk_bytes - 1) / chunk_bytes;
    std::vector<uint64_t> cand(m + 1, n);
    cand[0] = 0;
    if (is_fastq) {
        // line numbers at the targets: '\n' counts per 64 MiB block on every
        // thread, then per target the rest of its block
        constexpr uint64_t BLK = uint64_t(64) << 20;
        const uint64_t nblk = (n + BLK - 1) / BLK;
        std::vector<uint64_t> nl(nblk + 1, 0);
        const uint32_t Tb = std::max<uint32_t>(1, std::min<uint64_t>(T, nblk));
        parallel(Tb, [&](uint32_t t) {
            for (uint64_t j = t; j < nblk; j += Tb)
                nl[j + 1] = uint64_t(std::count(buf + j * BLK, buf + std::min(n, (j + 1) * BLK), '\n'));
        });
        for (uint64_t j = 0; j < nblk; ++j) nl[j + 1] += nl[j];
        uint64_t at = 0, nl_at = 0;  // '\n' in [0, at), carried from target to target
        for (uint64_t j = 1; j < m; ++j) {
            const uint64_t x = j * chunk_bytes, xb = x / BLK;
            if (cand[j - 1] >= x) {  // no record starts in [x_{j-1}, x): the same boundary
                cand[j] = cand[j - 1];
                continue;
            }
            if (xb * BLK > at) at = xb * BLK, nl_at = nl[xb];
            const uint64_t nl_x = nl_at + uint64_t(std::count(buf + at, buf + x, '\n'));
            at = x;
            nl_at = nl_x;
            uint64_t p = line_start_from(buf, n, x), li = nl_x + (p > x ? 1 : 0);
            while (p < n && li % 4 != 0) {
                p = line_end(buf, n, p) + 1;
                ++li;
            }
            cand[j] = std::min(p, n);
        }
    } else {
        for (uint64_t j = 1; j < m; ++j)  // a boundary past the next target is that target's too
            cand[j] = cand[j - 1] >= j * chunk_bytes ? cand[j - 1] : header_from(buf, n, j * chunk_bytes);
    }
    for (uint64_t j = 1; j <= m; ++j)
        if (cand[j] > off.back()) off.push_back(cand[j]);
    if (off.back() != n) off.push_back(n);
    return off;
}

}  // namespace cfhost_detail

using namespace cfhost_detail;

namespace {

int build(std::vector<Seg> segs, uint64_t n_rec, uint32_t k, cfhost_input** out) {
    std::unique_ptr<cfhost_input> in(new cfhost_input());
    Packer pk;
    pk.plan(segs, k);
    in->n_records = n_rec;
    in->n_bases = pk.n_bases();
    in->packed.assign(pk.n_words(), 0);
    in->frag_off.resize(pk.n_frag() + 1);
    pk.write(in->packed.data(), in->frag_off.data());
    *out = in.release();
    return 0;
}

}  // namespace

extern "C" {

int cfhost_input_from_records(const char* seq, const uint64_t* rec_off, uint64_t n_rec, uint32_t k,
                              uint32_t threads, cfhost_input** out) {
    if (!out || (n_rec && (!seq || !rec_off)) || k < 3 || k > 127) return -1;
    return build(segs_of_records(seq, rec_off, n_rec, pick_threads(threads)), n_rec, k, out);
}

int cfhost_input_from_buffer(const char* buf, uint64_t n, int is_fastq, uint32_t k, uint32_t threads,
                             cfhost_input** out) {
    if (!out || (!buf && n) || k < 3 || k > 127) return -1;
    std::vector<Seg> segs = parse_text(buf, n, is_fastq, pick_threads(threads));
    uint64_t n_rec = 0;
    for (const Seg& s : segs) n_rec += s.rec.size();
    return build(std::move(segs), n_rec, k, out);
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

int cfhost_split_records(const char* buf, uint64_t n, int is_fastq, uint64_t chunk_bytes, uint32_t threads,
                         uint64_t* offsets, uint64_t cap, uint64_t* n_chunks) {
    if ((!buf && n) || !n_chunks) return -1;
    const std::vector<uint64_t> off = split_records(buf, n, is_fastq, chunk_bytes, pick_threads(threads));
    *n_chunks = off.size() - 1;
    if (!offsets) return 0;
    if (cap < off.size()) return -1;
    std::copy(off.begin(), off.end(), offsets);
    return 0;
}

}  // extern "C"
