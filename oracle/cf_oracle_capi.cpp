// C ABI over cf_oracle for the Python tests (ctypes) and the CPU baseline leg of
// bench.py.  Test infrastructure only: the product library never links this.
#include "cf_oracle.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

using namespace cforacle;

// ---------------------------------------------------------------------------
// Order-independent checksums for comparisons at sizes where rows cannot be
// compared one by one: per bucket g, the number of items and the wrapping sum
// of a 64-bit mix of each item.  Equal multisets give equal sums; an item
// that differs changes its bucket's sum except with probability about 2^-64.
namespace {
uint64_t mix64(uint64_t z) {  // the splitmix64 finaliser
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint64_t record_mix(uint16_t att, const uint64_t* lab, unsigned W) {
    uint64_t h = mix64(att);
    for (unsigned w = 0; w < W; ++w) h = mix64(h ^ lab[w] ^ (uint64_t(w + 1) << 56));
    return h;
}
uint64_t state_mix(const uint64_t* key, unsigned kw, uint32_t e_f, uint8_t status) {
    uint64_t h = mix64(uint64_t(e_f) | (uint64_t(status) << 32));
    for (unsigned w = 0; w < kw; ++w) h = mix64(h ^ key[w] ^ (uint64_t(w + 1) << 56));
    return h;
}
template <class F>
void per_thread_sums(uint64_t n, int threads, uint32_t G, uint64_t* sum_out, uint64_t* cnt_out, F&& item) {
    if (threads < 1) threads = 1;
    std::vector<std::vector<uint64_t>> S(threads, std::vector<uint64_t>(G)), N(threads, std::vector<uint64_t>(G));
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t)
        th.emplace_back([&, t]() {
            for (uint64_t i = n * t / threads; i < n * (t + 1) / threads; ++i) item(i, S[t], N[t]);
        });
    for (auto& x : th) x.join();
    for (uint32_t g = 0; g < G; ++g) {
        sum_out[g] = cnt_out[g] = 0;
        for (int t = 0; t < threads; ++t) sum_out[g] += S[t][g], cnt_out[g] += N[t][g];
    }
}
}  // namespace

extern "C" {

uint64_t cfo_wyhash(const void* key, size_t len, uint64_t seed) { return wyhash(key, len, seed, kWySalt); }
uint64_t cfo_lmer_hash(uint64_t lmer) { return lmer_hash(lmer); }

int64_t cfo_minimizers(const char* frag, size_t len, int km1, int l, int brute, uint64_t* out,
                       size_t cap) {
    std::vector<uint64_t> H;
    if (brute) minimizer_hashes_brute(frag, len, km1, l, H);
    else minimizer_hashes_literal(frag, len, km1, l, H);
    if (H.size() > cap) return -int64_t(H.size());
    std::memcpy(out, H.data(), H.size() * 8);
    return int64_t(H.size());
}

// Records -> sequence bytes concatenated, rec_off[n+1].
int64_t cfo_parse(const char* buf, size_t n, int fastq, char* seq_out, uint64_t* rec_off,
                  size_t cap_bytes, size_t cap_recs) {
    const auto recs = fastq ? parse_fastq(buf, n) : parse_fasta(buf, n);
    size_t tot = 0;
    for (const auto& r : recs) tot += r.size();
    if (recs.size() > cap_recs || tot > cap_bytes) return -1;
    size_t p = 0;
    rec_off[0] = 0;
    for (size_t i = 0; i < recs.size(); ++i) {
        std::memcpy(seq_out + p, recs[i].data(), recs[i].size());
        p += recs[i].size();
        rec_off[i + 1] = p;
    }
    return int64_t(recs.size());
}

int64_t cfo_fragments(const char* seq, const uint64_t* rec_off, size_t n_rec, int k,
                      uint64_t* beg_out, uint64_t* len_out, size_t cap) {
    size_t c = 0;
    for (size_t r = 0; r < n_rec; ++r) {
        const auto fr = fragments(seq + rec_off[r], rec_off[r + 1] - rec_off[r], k);
        for (const auto& f : fr) {
            if (c < cap) beg_out[c] = rec_off[r] + f.first, len_out[c] = f.second;
            ++c;
        }
    }
    return c > cap ? -int64_t(c) : int64_t(c);
}

// Partition all records.  Outputs per weak super k-mer: g, packed attribute,
// W label words, global position of its first base. Emission order.
int64_t cfo_partition(const char* seq, const uint64_t* rec_off, size_t n_rec, int k, int l,
                      uint32_t G, int scan_form, uint32_t* g_out, uint16_t* att_out,
                      uint64_t* label_out, uint64_t* pos_out, size_t cap) {
    const unsigned W = label_words(k, l);
    size_t c = 0;
    std::vector<SuperKmer> sk;
    for (size_t r = 0; r < n_rec; ++r) {
        const char* rs = seq + rec_off[r];
        for (const auto& f : fragments(rs, rec_off[r + 1] - rec_off[r], k)) {
            sk.clear();
            if (scan_form) partition_fragment_scan(rs + f.first, f.second, k, l, G, sk);
            else partition_fragment(rs + f.first, f.second, k, l, G, sk);
            for (const auto& s : sk) {
                if (c < cap) {
                    g_out[c] = s.g;
                    att_out[c] = att_pack(s.len, s.l_disc, s.r_disc);
                    encode_label(rs + f.first + s.off, s.len, W, label_out + c * W);
                    pos_out[c] = rec_off[r] + f.first + s.off;
                }
                ++c;
            }
        }
    }
    return c > cap ? -int64_t(c) : int64_t(c);
}

// Construct all subgraphs from bucketed records (bucket_off[G+1], records of
// bucket g at [bucket_off[g], bucket_off[g+1])).  mode 0 = literal, 1 = closed
// form, 2 = literal over the reusable open-addressing map (the CPU baseline's).  Output sorted by (g, key); keys as 4 words (Kmer layout, w[0] low).
int64_t cfo_construct(int k, int l, uint32_t G, int mode, const uint64_t* bucket_off,
                      const uint16_t* att, const uint64_t* labels, int threads, uint32_t* g_out,
                      uint64_t* key_out, uint32_t* ef_out, uint8_t* st_out, size_t cap) {
    const unsigned W = label_words(k, l);
    if (threads < 1) threads = 1;
    std::vector<std::vector<VertexState>> per_t(threads);
    std::atomic<uint32_t> next{0};
    auto work = [&](int t) {
        FlatMap M;
        while (true) {
            const uint32_t g = next.fetch_add(1);
            if (g >= G) break;
            const uint64_t b = bucket_off[g], e = bucket_off[g + 1];
            if (b == e) continue;
            if (mode == 0) construct_literal(k, W, g, e - b, att + b, labels + b * W, per_t[t]);
            else if (mode == 1) construct_closed_form(k, W, g, e - b, att + b, labels + b * W, per_t[t]);
            else construct_flat(k, W, g, e - b, att + b, labels + b * W, M, per_t[t]);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t) th.emplace_back(work, t);
    for (auto& x : th) x.join();
    std::vector<VertexState> all;
    for (auto& v : per_t) all.insert(all.end(), v.begin(), v.end());
    sort_states(k, all);
    if (all.size() > cap) return -int64_t(all.size());
    for (size_t i = 0; i < all.size(); ++i) {
        g_out[i] = all[i].g;
        std::memcpy(key_out + 4 * i, all[i].key.w.data(), 32);
        ef_out[i] = all[i].st.e_f;
        st_out[i] = all[i].st.status;
    }
    return int64_t(all.size());
}


// Partition every record (the oracle's sequential split, records spread over
// `threads` workers) and checksum the weak super k-mers of each bucket:
// mix of (attribute pack, label words, bits past the length zero).
int cfo_partition_hash(const char* seq, const uint64_t* rec_off, size_t n_rec, int k, int l, int threads,
                       uint64_t* sum_out, uint64_t* cnt_out) {
    const unsigned W = label_words(k, l);
    const uint32_t G = 16384;
    per_thread_sums(n_rec, threads, G, sum_out, cnt_out,
                    [&](uint64_t r, std::vector<uint64_t>& S, std::vector<uint64_t>& N) {
                        thread_local std::vector<SuperKmer> sk;
                        uint64_t lab[8];
                        const char* rs = seq + rec_off[r];
                        for (const auto& f : fragments(rs, rec_off[r + 1] - rec_off[r], k)) {
                            sk.clear();
                            partition_fragment(rs + f.first, f.second, k, l, G, sk);
                            for (const auto& x : sk) {
                                encode_label(rs + f.first + x.off, x.len, W, lab);
                                S[x.g] += record_mix(att_pack(x.len, x.l_disc, x.r_disc), lab, W);
                                ++N[x.g];
                            }
                        }
                    });
    return 0;
}

// The same checksums of records given as the device returns them: att with
// the bucket in its high half, W label words each.
int cfo_records_hash(const uint32_t* att, const uint64_t* labels, uint64_t n, int W, int threads,
                     uint64_t* sum_out, uint64_t* cnt_out) {
    per_thread_sums(n, threads, 16384, sum_out, cnt_out,
                    [&](uint64_t i, std::vector<uint64_t>& S, std::vector<uint64_t>& N) {
                        const uint32_t g = att[i] >> 16;
                        S[g] += record_mix(uint16_t(att[i] & 0xFFFF), labels + i * uint64_t(W), unsigned(W));
                        ++N[g];
                    });
    return 0;
}

// Checksums of vertex states per bucket: mix of (key words, e_f, status).
int cfo_states_hash(const uint32_t* g, const uint64_t* keys, int kw, const uint32_t* e_f, const uint8_t* status,
                    uint64_t n, int threads, uint64_t* sum_out, uint64_t* cnt_out) {
    per_thread_sums(n, threads, 16384, sum_out, cnt_out,
                    [&](uint64_t i, std::vector<uint64_t>& S, std::vector<uint64_t>& N) {
                        S[g[i]] += state_mix(keys + i * uint64_t(kw), unsigned(kw), e_f[i], status[i]);
                        ++N[g[i]];
                    });
    return 0;
}

// CPU baseline in the reference's shape, per phase:
//  1. partition on part_threads workers over contiguous record ranges (the
//     reference partitions one input file per worker, src/Graph_Partitioner.cpp:
//     121-184: part_threads = 1 for a single read file; = threads is the
//     parallel variant);
//  2. bucketize by g (Atlas / Super_Kmer_Bucket);
//  3. construct every subgraph with each worker's reusable open-addressing
//     map, buckets largest first (src/Subgraphs_Manager.cpp:222-255);
//  4. contract every subgraph's table (Subgraph::contract restated).
// out[0..8] = partition ms, bucketize ms, construct ms, contract ms, k-mer
// instances, weak super k-mers, vertex states, unitigs, construct + contract
// wall ms when run together per subgraph.
int cfo_baseline(const char* seq, const uint64_t* rec_off, size_t n_rec, int k, int l, int f_th, int part_threads,
                 int threads, double* out) {
    using clk = std::chrono::steady_clock;
    const unsigned W = label_words(k, l);
    const uint32_t G = 16384;
    if (threads < 1) threads = 1;
    if (part_threads < 1) part_threads = 1;
    struct Part { std::vector<uint32_t> g; std::vector<uint16_t> att; std::vector<uint64_t> lab; };
    std::vector<Part> parts(part_threads);
    const auto t0 = clk::now();
    {
        std::vector<std::thread> th;
        for (int t = 0; t < part_threads; ++t)
            th.emplace_back([&, t]() {
                Part& P = parts[t];
                std::vector<SuperKmer> sk;
                const size_t r0 = n_rec * t / part_threads, r1 = n_rec * (t + 1) / part_threads;
                for (size_t r = r0; r < r1; ++r) {
                    const char* rs = seq + rec_off[r];
                    for (const auto& f : fragments(rs, rec_off[r + 1] - rec_off[r], k)) {
                        sk.clear();
                        partition_fragment(rs + f.first, f.second, k, l, G, sk);
                        for (const auto& x : sk) {
                            P.g.push_back(x.g);
                            P.att.push_back(att_pack(x.len, x.l_disc, x.r_disc));
                            const size_t o = P.lab.size();
                            P.lab.resize(o + W);
                            encode_label(rs + f.first + x.off, x.len, W, &P.lab[o]);
                        }
                    }
                }
            });
        for (auto& x : th) x.join();
    }
    const auto t1 = clk::now();
    std::vector<uint64_t> off(G + 1, 0);
    for (const auto& P : parts) for (uint32_t g : P.g) ++off[g + 1];
    for (uint32_t g = 0; g < G; ++g) off[g + 1] += off[g];
    const uint64_t n_sk = off[G];
    std::vector<uint16_t> att(n_sk);
    std::vector<uint64_t> lab(n_sk * W);
    {
        std::vector<uint64_t> cur(off.begin(), off.end() - 1);
        for (const auto& P : parts)
            for (size_t i = 0; i < P.g.size(); ++i) {
                const uint64_t d = cur[P.g[i]]++;
                att[d] = P.att[i];
                std::memcpy(&lab[d * W], &P.lab[i * W], W * 8);
            }
    }
    parts.clear();
    const auto t2 = clk::now();
    std::vector<uint32_t> order(G);
    for (uint32_t g = 0; g < G; ++g) order[g] = g;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return off[a + 1] - off[a] > off[b + 1] - off[b];
    });
    std::atomic<uint32_t> next{0};
    std::atomic<uint64_t> states{0}, unitigs{0}, ns_con{0}, ns_ctr{0};
    {
        std::vector<std::thread> th;
        for (int t = 0; t < threads; ++t)
            th.emplace_back([&]() {
                FlatMap M;
                std::vector<VertexState> vs;
                ContractResult res;
                uint64_t s = 0, u = 0, c0 = 0, c1 = 0;
                while (true) {
                    const uint32_t i = next.fetch_add(1);
                    if (i >= G) break;
                    const uint32_t g = order[i];
                    if (off[g + 1] == off[g]) continue;
                    const auto a = clk::now();
                    vs.clear();
                    construct_flat(k, W, g, off[g + 1] - off[g], &att[off[g]], &lab[off[g] * W], M, vs);
                    const auto b = clk::now();
                    res = ContractResult{};
                    contract_subgraph(k, f_th, vs.data(), vs.size(), res);
                    const auto c = clk::now();
                    s += vs.size();
                    u += res.unitigs.size();
                    c0 += uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(b - a).count());
                    c1 += uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(c - b).count());
                }
                states += s;
                unitigs += u;
                ns_con += c0;
                ns_ctr += c1;
            });
        for (auto& x : th) x.join();
    }
    const auto t3 = clk::now();
    uint64_t n_km = 0;
    for (uint64_t i = 0; i < n_sk; ++i) n_km += (att[i] & 0xFF) - (k - 1);
    auto ms = [](clk::duration d) { return std::chrono::duration<double, std::milli>(d).count(); };
    const double wall = ms(t3 - t2), share = double(ns_con) / double(std::max<uint64_t>(1, ns_con + ns_ctr));
    out[0] = ms(t1 - t0);
    out[1] = ms(t2 - t1);
    out[2] = wall * share;          // the workers' construct share of the joint phase
    out[3] = wall * (1.0 - share);  // ... and their contract share
    out[4] = double(n_km);
    out[5] = double(n_sk);
    out[6] = double(states.load());
    out[7] = double(unitigs.load());
    out[8] = wall;
    return 0;
}

uint32_t cfo_edge_at(uint32_t e_f, int side, int f_th) { return edge_at(e_f, side, f_th); }

// CPU baseline: the reference's pipeline shape on `threads` workers.
//  1. partition: records split into `threads` contiguous ranges (the reference
//     parallelises across input files, src/Graph_Partitioner.cpp:121-184; a
//     single read file is split here so every core partitions);
//  2. bucketize the weak super k-mers by g (Atlas / Super_Kmer_Bucket);
//  3. construct every subgraph with a per-worker hash map, buckets handed out
//     largest first (src/Subgraphs_Manager.cpp:222-255).
// out[0..5] = partition ms, bucketize ms, construct ms, k-mer instances,
//             weak super k-mers, distinct (g, vertex) states.
int cfo_pipeline(const char* seq, const uint64_t* rec_off, size_t n_rec, int k, int l, int threads,
                 double* out) {
    using clk = std::chrono::steady_clock;
    const unsigned W = label_words(k, l);
    const uint32_t G = 16384;
    if (threads < 1) threads = 1;
    struct Part { std::vector<uint32_t> g; std::vector<uint16_t> att; std::vector<uint64_t> lab; };
    std::vector<Part> parts(threads);
    const auto t0 = clk::now();
    {
        std::vector<std::thread> th;
        for (int t = 0; t < threads; ++t)
            th.emplace_back([&, t]() {
                Part& P = parts[t];
                std::vector<SuperKmer> sk;
                const size_t r0 = n_rec * t / threads, r1 = n_rec * (t + 1) / threads;
                for (size_t r = r0; r < r1; ++r) {
                    const char* rs = seq + rec_off[r];
                    for (const auto& f : fragments(rs, rec_off[r + 1] - rec_off[r], k)) {
                        sk.clear();
                        partition_fragment(rs + f.first, f.second, k, l, G, sk);
                        for (const auto& x : sk) {
                            P.g.push_back(x.g);
                            P.att.push_back(att_pack(x.len, x.l_disc, x.r_disc));
                            const size_t o = P.lab.size();
                            P.lab.resize(o + W);
                            encode_label(rs + f.first + x.off, x.len, W, &P.lab[o]);
                        }
                    }
                }
            });
        for (auto& x : th) x.join();
    }
    const auto t1 = clk::now();
    std::vector<uint64_t> off(G + 1, 0);
    for (const auto& P : parts) for (uint32_t g : P.g) ++off[g + 1];
    for (uint32_t g = 0; g < G; ++g) off[g + 1] += off[g];
    const uint64_t n_sk = off[G];
    std::vector<uint16_t> att(n_sk);
    std::vector<uint64_t> lab(n_sk * W);
    {
        std::vector<uint64_t> cur(off.begin(), off.end() - 1);
        for (const auto& P : parts)
            for (size_t i = 0; i < P.g.size(); ++i) {
                const uint64_t d = cur[P.g[i]]++;
                att[d] = P.att[i];
                std::memcpy(&lab[d * W], &P.lab[i * W], W * 8);
            }
    }
    parts.clear();
    const auto t2 = clk::now();
    std::vector<uint32_t> order(G);
    for (uint32_t g = 0; g < G; ++g) order[g] = g;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return off[a + 1] - off[a] > off[b + 1] - off[b];
    });
    std::atomic<uint32_t> next{0};
    std::atomic<uint64_t> states{0};
    {
        std::vector<std::thread> th;
        for (int t = 0; t < threads; ++t)
            th.emplace_back([&]() {
                std::vector<VertexState> vs;
                uint64_t mine = 0;
                while (true) {
                    const uint32_t i = next.fetch_add(1);
                    if (i >= G) break;
                    const uint32_t g = order[i];
                    if (off[g + 1] == off[g]) continue;
                    vs.clear();
                    construct_literal(k, W, g, off[g + 1] - off[g], &att[off[g]], &lab[off[g] * W], vs);
                    mine += vs.size();
                }
                states += mine;
            });
        for (auto& x : th) x.join();
    }
    const auto t3 = clk::now();
    uint64_t n_km = 0;
    for (uint64_t i = 0; i < n_sk; ++i) n_km += (att[i] & 0xFF) - (k - 1);
    auto ms = [](clk::duration d) { return std::chrono::duration<double, std::milli>(d).count(); };
    out[0] = ms(t1 - t0);
    out[1] = ms(t2 - t1);
    out[2] = ms(t3 - t2);
    out[3] = double(n_km);
    out[4] = double(n_sk);
    out[5] = double(states.load());
    return 0;
}

// Subgraph::contract over vertex states sorted by g (as cfo_construct returns
// them).  Per bucket g, counts[7g..7g+6] = isolated, phantom, trivial unitigs,
// cycles, lm-tigs (discontinuity edges), vertices in unitigs, label bytes.  Per
// unitig: rec_g, rec_flags (1 cycle, 2 left exit, 4 right exit), rec_off /
// rec_len into `labels`, ends[8i..8i+7] = the left and right discontinuity
// vertices (4 words each, zero for phi), sides[2i..2i+1] (0 front, 1 back,
// 2 phi).  `normalise` orients lm-tigs canonically.  Unitigs of a bucket are
// listed in the order of their walks.  Returns the unitig count, or -1 when a
// capacity is short (*lab_need then holds the label bytes needed).
int64_t cfo_contract(int k, int f_th, uint32_t G, size_t n, const uint32_t* g, const uint64_t* keys,
                     const uint32_t* e_f, const uint8_t* status, int normalise, int threads, uint64_t* counts,
                     uint32_t* rec_g, uint32_t* rec_flags, uint64_t* rec_off, uint32_t* rec_len,
                     uint64_t* ends, uint8_t* sides, size_t rec_cap, char* labels, size_t lab_cap,
                     uint64_t* lab_need) {
    std::vector<VertexState> v(n);
    for (size_t i = 0; i < n; ++i) {
        v[i].g = g[i];
        std::memcpy(v[i].key.w.data(), keys + 4 * i, 32);
        v[i].st.e_f = e_f[i];
        v[i].st.status = status[i];
    }
    std::vector<size_t> beg(G + 1, n);
    for (size_t i = n; i-- > 0;) beg[g[i]] = i;
    for (uint32_t b = G; b-- > 0;) beg[b] = std::min(beg[b], beg[b + 1]);
    std::vector<ContractResult> res(G);
    if (threads < 1) threads = 1;
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        while (true) {
            const uint32_t b = next.fetch_add(1);
            if (b >= G) break;
            if (beg[b] == beg[b + 1]) continue;
            contract_subgraph(k, f_th, v.data() + beg[b], beg[b + 1] - beg[b], res[b]);
            if (normalise)
                for (auto& u : res[b].unitigs) normalise_unitig(k, u);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t) th.emplace_back(work);
    for (auto& x : th) x.join();
    size_t nr = 0, nl = 0;
    for (uint32_t b = 0; b < G; ++b) {
        const auto& r = res[b];
        const uint64_t c[7] = {r.isolated, r.phantom, r.trivial, r.icc, r.disc_edges, r.vertex_count, r.label_sz};
        std::memcpy(counts + 7 * size_t(b), c, sizeof c);
        nr += r.unitigs.size();
        for (const auto& u : r.unitigs) nl += u.label.size();
    }
    *lab_need = nl;
    if (nr > rec_cap || nl > lab_cap) return -1;
    size_t i = 0, o = 0;
    for (uint32_t b = 0; b < G; ++b)
        for (const auto& u : res[b].unitigs) {
            rec_g[i] = b;
            rec_flags[i] = u.flags;
            rec_off[i] = o;
            rec_len[i] = uint32_t(u.label.size());
            std::memcpy(labels + o, u.label.data(), u.label.size());
            o += u.label.size();
            std::memcpy(ends + 8 * i, (u.u_side == 2 ? Km{} : u.u).w.data(), 32);
            std::memcpy(ends + 8 * i + 4, (u.v_side == 2 ? Km{} : u.v).w.data(), 32);
            sides[2 * i] = uint8_t(u.u_side);
            sides[2 * i + 1] = uint8_t(u.v_side);
            ++i;
        }
    return int64_t(nr);
}

}  // extern "C"
