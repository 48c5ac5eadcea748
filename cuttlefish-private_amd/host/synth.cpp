// Seeded synthetic inputs for tests and benchmarks (include/cfhost.h).
#include "../../include/cfhost.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// splitmix64: stateless stream keyed by (seed, index).
inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(mix64(seed ^ 0xD1B54A32D192ED03ull)) {}
    uint64_t next() { return mix64(s++); }
    double unif() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

const char ACGT[4] = {'A', 'C', 'G', 'T'};

inline char comp(char c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        default: return 'A';
    }
}

inline char substitute(char c, Rng& r) {
    const char* p = std::strchr(ACGT, c);
    const int cur = p ? int(p - ACGT) : 0;
    return ACGT[(cur + 1 + int(r.below(3))) & 3];
}

void fill_iid(char* out, uint64_t n, uint64_t seed) {
    // 32 bases per draw, chunked by index so the result is thread-count free
    for (uint64_t i = 0; i < n; i += 32) {
        uint64_t x = mix64(seed * 0x9E3779B97F4A7C15ull + i);
        const uint64_t e = std::min<uint64_t>(n, i + 32);
        for (uint64_t j = i; j < e; ++j, x >>= 2) out[j] = ACGT[x & 3];
    }
}

}  // namespace

extern "C" {

int cfhost_synth_genome(uint64_t len, double repeat_frac, uint32_t fam_min, uint32_t fam_max,
                        double divergence, uint64_t seed, char* out) {
    if (!out || repeat_frac < 0 || repeat_frac >= 1 || (repeat_frac > 0 && (fam_min < 1 || fam_max < fam_min)))
        return -1;
    fill_iid(out, len, seed);
    if (repeat_frac <= 0 || len < fam_max) return 0;
    Rng r(seed ^ 0x5EEDull);
    const uint64_t target = uint64_t(repeat_frac * double(len));
    uint64_t covered = 0;
    std::vector<char> fam;
    while (covered < target) {
        // one family: log-uniform length, 2..64 copies
        const double lf = std::log(double(fam_min)) + r.unif() * (std::log(double(fam_max)) - std::log(double(fam_min)));
        const uint32_t flen = std::max<uint32_t>(fam_min, std::min<uint32_t>(fam_max, uint32_t(std::exp(lf))));
        fam.resize(flen);
        for (uint32_t i = 0; i < flen; ++i) fam[i] = ACGT[r.below(4)];
        const uint32_t copies = 2 + uint32_t(r.below(63));
        for (uint32_t c = 0; c < copies && covered < target; ++c) {
            const uint64_t pos = r.below(len - flen + 1);
            const bool rcc = r.below(2);
            for (uint32_t i = 0; i < flen; ++i) {
                char b = rcc ? comp(fam[flen - 1 - i]) : fam[i];
                if (r.unif() < divergence) b = substitute(b, r);
                out[pos + i] = b;
            }
            covered += flen;
        }
    }
    return 0;
}

int cfhost_synth_insert_repeats(const char* base, uint64_t base_len, uint32_t n_copies, uint32_t rep_len,
                                double sub_rate, uint64_t seed, char* out) {
    if ((!base && base_len) || !out) return -1;
    Rng r(seed);
    std::vector<char> rep(rep_len);
    for (uint32_t i = 0; i < rep_len; ++i) rep[i] = ACGT[r.below(4)];
    std::vector<uint64_t> pos(n_copies);
    for (auto& p : pos) p = r.below(base_len + 1);
    std::sort(pos.begin(), pos.end());
    uint64_t src = 0, dst = 0;
    for (uint32_t c = 0; c < n_copies; ++c) {
        std::memcpy(out + dst, base + src, pos[c] - src);
        dst += pos[c] - src;
        src = pos[c];
        for (uint32_t i = 0; i < rep_len; ++i) {
            char b = rep[i];
            if (r.unif() < sub_rate) b = substitute(b, r);
            out[dst++] = b;
        }
    }
    std::memcpy(out + dst, base + src, base_len - src);
    return 0;
}

int cfhost_synth_reads(const char* genome, uint64_t glen, uint64_t n_reads, uint32_t read_len,
                       double sub_rate, uint64_t seed, uint32_t threads, char* out) {
    if (!genome || !out || glen < read_len || read_len == 0) return -1;
    uint32_t T = threads ? threads : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    T = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(T, n_reads)));
    auto work = [&](uint32_t t) {
        const uint64_t b = n_reads * t / T, e = n_reads * (t + 1) / T;
        for (uint64_t i = b; i < e; ++i) {
            Rng r(seed * 0x9E3779B97F4A7C15ull + i);
            const uint64_t st = r.below(glen - read_len + 1);
            const bool rcs = r.next() & 1;
            char* o = out + i * read_len;
            for (uint32_t j = 0; j < read_len; ++j) {
                char c = rcs ? comp(genome[st + read_len - 1 - j]) : genome[st + j];
                if (r.unif() < sub_rate) c = substitute(c, r);
                o[j] = c;
            }
        }
    };
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < T; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
    return 0;
}

int cfhost_synth_mutate(const char* base, uint64_t len, double snp_rate, double indel_rate,
                        uint32_t indel_max, uint64_t seed, char* out, uint64_t cap, uint64_t* out_len) {
    if (!base || !out || !out_len || indel_max < 1) return -1;
    Rng r(seed);
    uint64_t o = 0;
    for (uint64_t i = 0; i < len;) {
        const double u = r.unif();
        if (u < indel_rate) {
            const uint32_t n = 1 + uint32_t(r.below(indel_max));
            if (r.below(2)) {  // insertion before base i
                for (uint32_t j = 0; j < n; ++j) {
                    if (o >= cap) return -2;
                    out[o++] = ACGT[r.below(4)];
                }
            } else {  // deletion of n bases
                i += n;
                continue;
            }
        }
        if (o >= cap) return -2;
        char c = base[i++];
        if (r.unif() < snp_rate) c = substitute(c, r);
        out[o++] = c;
    }
    *out_len = o;
    return 0;
}

int cfhost_synth_bucket_workload(uint32_t n_buckets, uint64_t n_records, double zipf_s, uint32_t k,
                                 uint32_t l, uint64_t seed, uint32_t* att, uint64_t* labels,
                                 uint64_t* bucket_off) {
    const uint32_t G = 16384;
    if (!att || !labels || !bucket_off || n_buckets == 0 || n_buckets > G || k < 3 || l >= k ||
        2 * k - l > 255)
        return -1;
    const uint32_t W = ((2 * (k - 1) - l + 2) + 31) / 32;
    // record counts per bucket: largest remainder of n_records * w_b / sum(w)
    std::vector<double> w(n_buckets);
    double tot = 0;
    for (uint32_t b = 0; b < n_buckets; ++b) tot += (w[b] = std::pow(double(b + 1), -zipf_s));
    std::vector<uint64_t> cnt(n_buckets);
    uint64_t assigned = 0;
    for (uint32_t b = 0; b < n_buckets; ++b) assigned += (cnt[b] = uint64_t(double(n_records) * w[b] / tot));
    for (uint32_t b = 0; assigned < n_records; b = (b + 1) % n_buckets) ++cnt[b], ++assigned;
    bucket_off[0] = 0;
    for (uint32_t g = 0; g < G; ++g) bucket_off[g + 1] = bucket_off[g] + (g < n_buckets ? cnt[g] : 0);
    const uint32_t max_len = 2 * k - l;
    std::vector<uint8_t> src;
    for (uint32_t b = 0; b < n_buckets; ++b) {
        Rng r(seed * 0x9E3779B97F4A7C15ull + b);
        const uint64_t slen = std::max<uint64_t>(2 * max_len, 2 * cnt[b]);
        src.resize(slen);
        for (auto& x : src) x = uint8_t(r.below(4));
        for (uint64_t i = bucket_off[b]; i < bucket_off[b + 1]; ++i) {
            const uint32_t len = k + uint32_t(r.below(max_len - k + 1));
            const uint64_t pos = r.below(slen - len + 1);
            const bool ld = r.below(8) == 0, rd = r.below(8) == 0;
            att[i] = len | (uint32_t(ld) << 8) | (uint32_t(rd) << 9) | (b << 16);
            uint64_t* L = labels + i * W;
            for (uint32_t j = 0; j < W; ++j) L[j] = 0;
            for (uint32_t t = 0; t < len; ++t)
                L[W - 1 - t / 32] |= uint64_t(src[pos + t]) << (62 - 2 * (t % 32));
        }
    }
    return 0;
}

}  // extern "C"
