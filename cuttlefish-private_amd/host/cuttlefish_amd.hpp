// C++ host layer over the C ABI, shaped like the reference classes it stands in
// for, so reference-side code (and tests written like the reference's) drive
// the GPU path through familiar names:
//
//   Subgraphs_Manager   include/Subgraphs_Manager.hpp:38-135 (graph_count, graph_ID,
//                       finalize, estimate_size_max, process) plus the partition
//                       entry of Graph_Partitioner (src/Graph_Partitioner.cpp:45)
//   Subgraph            include/Subgraph.hpp (construct, kmer_count, size, and the
//                       HT_Router accessors contract() uses: find / iteration /
//                       get_key / get_val)
//   State_Config        include/State_Config.hpp:24-268 (edge_at, is_branching_side,
//                       is_empty_side, is_discontinuous, mark_visited, is_visited)
//
// Errors: the ABI never throws; this layer turns a failed call into
// std::runtime_error carrying cfgpu_last_error's message.  The reference exits
// on I/O failure (include/Super_Kmer_Chunk.hpp:298-303); callers that want that
// behaviour catch and std::exit, as tools/cf_build does.
#pragma once

#include <cstdint>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cfgpu.h"
#include "../../include/cfhost.h"

namespace cuttlefish_amd {

enum class side_t : uint8_t { front = 0, back = 1, unspecified = 2 };  // include/globals.hpp:69-75
enum base_t : uint8_t { A = 0, C = 1, G = 2, T = 3, N = 4, E = 5 };     // include/DNA.hpp:17-25

// State_Config<false>: Edge_Frequency word + Vertex_Status byte.
class State_Config {
public:
    State_Config() = default;
    State_Config(uint32_t e_f, uint8_t status, uint8_t f_th) : e_f_(e_f), status_(status), f_th_(f_th) {}

    base_t edge_at(side_t s) const {
        const unsigned off = s == side_t::front ? 0 : 16;
        unsigned c = 0, which = 0;
        for (unsigned b = 0; b < 4; ++b)
            if (((e_f_ >> (off + 4 * b)) & 0xF) >= f_th_) ++c, which = b;
        return c == 0 ? E : (c == 1 ? base_t(which) : N);
    }
    uint32_t edge_count(side_t s) const {
        const unsigned off = s == side_t::front ? 0 : 16;
        uint32_t c = 0;
        for (unsigned b = 0; b < 4; ++b) c += ((e_f_ >> (off + 4 * b)) & 0xF) >= f_th_;
        return c;
    }
    bool is_branching_side(side_t s) const { return edge_count(s) > 1; }
    bool is_empty_side(side_t s) const { return edge_count(s) == 0; }
    bool is_isolated() const { return is_empty_side(side_t::back) && is_empty_side(side_t::front); }
    bool is_discontinuous(side_t s) const { return status_ & (s == side_t::front ? 2 : 4); }
    bool is_discontinuity() const { return status_ & 6; }
    bool is_visited() const { return status_ & 1; }
    void mark_visited() { status_ |= 1; }
    uint32_t frequency_word() const { return e_f_; }
    uint8_t status_byte() const { return status_; }

private:
    uint32_t e_f_ = 0;
    uint8_t status_ = 0;
    uint8_t f_th_ = 1;
};

// One subgraph's table, copied from HBM after construct.  Keys are Kmer<k>
// words (word 0 least significant); find() is an open-addressing index built
// over the copy.
class Subgraph {
public:
    Subgraph(uint32_t g, unsigned k, uint8_t f_th, std::vector<uint64_t>&& keys,
             std::vector<uint32_t>&& e_f, std::vector<uint8_t>&& status, uint64_t kmer_count)
        : g_(g), k_(k), kw_((k + 31) / 32), keys_(std::move(keys)), kmer_count_(kmer_count) {
        states_.reserve(e_f.size());
        for (size_t i = 0; i < e_f.size(); ++i) states_.emplace_back(e_f[i], status[i], f_th);
        uint64_t cap = 16;
        while (cap < 2 * states_.size()) cap <<= 1;
        index_.assign(cap, UINT32_MAX);
        for (uint32_t i = 0; i < states_.size(); ++i) {
            uint64_t s = hash(&keys_[size_t(i) * kw_]) & (cap - 1);
            while (index_[s] != UINT32_MAX) s = (s + 1) & (cap - 1);
            index_[s] = i;
        }
    }

    // The table was built on the GPU by cfgpu_construct; construct() is kept so
    // call sites read like src/Subgraphs_Manager.cpp:154-155.
    void construct() {}
    uint32_t id() const { return g_; }
    uint64_t kmer_count() const { return kmer_count_; }  // k-mer instances (Subgraph::kmer_count)
    std::size_t size() const { return states_.size(); }  // distinct vertices (Subgraph::size)

    // HT_Router-style access (include/Subgraph.hpp:336-358).
    const uint64_t* get_key(std::size_t i) const { return &keys_[i * kw_]; }
    State_Config& get_val(std::size_t i) { return states_[i]; }
    const State_Config& get_val(std::size_t i) const { return states_[i]; }
    // Index of the vertex with canonical key `kmer` (kw words), or size() if absent.
    std::size_t find(const uint64_t* kmer) const {
        const uint64_t cap = index_.size();
        for (uint64_t s = hash(kmer) & (cap - 1);; s = (s + 1) & (cap - 1)) {
            const uint32_t i = index_[s];
            if (i == UINT32_MAX) return size();
            bool eq = true;
            for (unsigned w = 0; w < kw_; ++w) eq &= keys_[size_t(i) * kw_ + w] == kmer[w];
            if (eq) return i;
        }
    }

private:
    uint64_t hash(const uint64_t* x) const {
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (unsigned w = 0; w < kw_; ++w) h = (h ^ x[w]) * 0xBF58476D1CE4E5B9ull, h ^= h >> 29;
        return h;
    }
    uint32_t g_;
    unsigned k_, kw_;
    std::vector<uint64_t> keys_;
    std::vector<State_Config> states_;
    std::vector<uint32_t> index_;
    uint64_t kmer_count_;
};

class Subgraphs_Manager {
public:
    static constexpr std::size_t graph_count() { return 16384; }  // include/Atlas.hpp:33-35
    uint64_t graph_ID(uint64_t h) const { return h & (graph_count() - 1); }  // Subgraphs_Manager.hpp:130

    Subgraphs_Manager(uint16_t k, uint16_t l, uint8_t cutoff, int device = 0)
        : k_(k), l_(l), cutoff_(cutoff) {
        cfgpu_params p{k, l, cutoff, uint32_t(graph_count()), device, 0};
        const int r = cfgpu_open(&p, &h_);
        if (r != CFGPU_OK) throw std::runtime_error("cfgpu_open failed (" + std::to_string(r) + ")");
    }
    ~Subgraphs_Manager() {
        if (h_) cfgpu_close(h_);
    }
    Subgraphs_Manager(const Subgraphs_Manager&) = delete;
    Subgraphs_Manager& operator=(const Subgraphs_Manager&) = delete;

    // Graph_Partitioner::partition over the input files (one FASTA/FASTQ group).
    void partition(const std::vector<std::string>& paths, bool is_fastq, unsigned threads = 0) {
        constructed_ = false;
        std::vector<uint64_t> packed, frag{0};
        uint64_t bases = 0;
        for (const std::string& p : paths) {
            cfhost_input* in = nullptr;
            if (cfhost_input_from_file(p.c_str(), is_fastq, k_, threads, &in) != 0)
                throw std::runtime_error("cannot read input " + p);
            const uint64_t *pk, *fo;
            uint64_t nw, nb, nf, nr;
            cfhost_input_get(in, &pk, &nw, &nb, &fo, &nf, &nr);
            append_stream(packed, bases, pk, nb);
            for (uint64_t f = 1; f <= nf; ++f) frag.push_back(fo[f] + (bases - nb));
            cfhost_input_free(in);
        }
        packed.resize((bases + 31) / 32 + 2, 0);
        check(cfgpu_partition(h_, packed.data(), bases, frag.data(), frag.size() - 1, 0));
        stats_.assign(graph_count(), cfgpu_bucket_stat{});
        check(cfgpu_bucket_stats(h_, stats_.data(), uint32_t(graph_count())));
    }

    // Subgraphs_Manager::finalize: buckets already sit in HBM.
    void finalize() {}

    // Largest possible subgraph size (its k-mer instance count): the bound a
    // table of that subgraph can never exceed.
    uint64_t estimate_size_max() const {
        uint64_t m = 0;
        for (const auto& s : stats_) m = s.n_kmers > m ? s.n_kmers : m;
        return m;
    }

    // Every Subgraph::construct on the GPU (src/Subgraphs_Manager.cpp:147-160).
    void construct(uint32_t mode = CFGPU_MODE_BATCHED_LEARNED) {
        check(cfgpu_construct(h_, mode));
        check(cfgpu_bucket_stats(h_, stats_.data(), uint32_t(graph_count())));
        constructed_ = true;
    }

    // Every Subgraph::contract, then the discontinuity-graph contraction, expansion and
    // collation (src/dBG_Contractor.cpp:75-113), on the GPU; the maximal unitigs are
    // written to `fasta_path` as one FASTA record each.  Returns their count.
    uint64_t contract_and_collate(const std::string& fasta_path) {
        if (!constructed_) construct();
        check(cfgpu_contract(h_));
        check(cfgpu_cdbg(h_));
        uint64_t n = 0, nb = 0;
        check(cfgpu_cdbg_copy(h_, nullptr, 0, nullptr, 0, &n, &nb));
        std::string seq(nb, '\0');
        std::vector<uint64_t> rec(2 * n);
        check(cfgpu_cdbg_copy(h_, &seq[0], nb, rec.data(), n, &n, &nb));
        std::FILE* f = std::fopen(fasta_path.c_str(), "wb");
        if (!f) throw std::runtime_error("cannot write " + fasta_path);
        for (uint64_t i = 0; i < n; ++i) {
            std::fprintf(f, ">%llu\n", static_cast<unsigned long long>(i));
            std::fwrite(seq.data() + rec[2 * i], 1, rec[2 * i + 1], f);
            std::fputc('\n', f);
        }
        if (std::fclose(f) != 0) throw std::runtime_error("cannot write " + fasta_path);
        return n;
    }

    // Subgraphs_Manager::process: construct every subgraph on the GPU (unless already
    // constructed), then hand each non-empty one to `contract` (the reference calls
    // Subgraph::contract) and release it.
    void process(const std::function<void(Subgraph&)>& contract,
                 uint32_t mode = CFGPU_MODE_BATCHED_LEARNED) {
        if (!constructed_) construct(mode);
        constructed_ = false;
        const unsigned kw = (k_ + 31) / 32;
        for (uint32_t g = 0; g < graph_count(); ++g) {
            if (!stats_[g].n_records) continue;
            const uint64_t n = stats_[g].n_vertices;
            std::vector<uint64_t> keys(n * kw);
            std::vector<uint32_t> e_f(n);
            std::vector<uint8_t> st(n);
            uint64_t got = 0;
            check(cfgpu_table_copy(h_, &g, 1, keys.data(), e_f.data(), st.data(), nullptr, nullptr, n, &got));
            Subgraph sub(g, k_, cutoff_, std::move(keys), std::move(e_f), std::move(st), stats_[g].n_kmers);
            sub.construct();
            check(cfgpu_release(h_, g));  // b.remove()
            if (contract) contract(sub);
        }
    }

    const std::vector<cfgpu_bucket_stat>& bucket_stats() const { return stats_; }
    cfgpu_run_stats run_stats() const {
        cfgpu_run_stats r{};
        cfgpu_run_stats_get(h_, &r);
        return r;
    }
    cfgpu_handle* handle() const { return h_; }

private:
    void check(int r) const {
        if (r == CFGPU_OK) return;
        char buf[512];
        cfgpu_last_error(h_, buf, sizeof buf);
        throw std::runtime_error(std::string("cfgpu error ") + std::to_string(r) + ": " + buf);
    }
    static void append_stream(std::vector<uint64_t>& dst, uint64_t& bases, const uint64_t* src, uint64_t nb) {
        // append nb MSB-first 2-bit bases of src at base offset `bases`
        const uint64_t total = bases + nb;
        dst.resize((total + 31) / 32 + 2, 0);
        const unsigned sh = unsigned(bases & 31) * 2;
        uint64_t w = bases / 32;
        for (uint64_t i = 0; i < (nb + 31) / 32; ++i) {
            const uint64_t x = src[i];
            if (!sh) {
                dst[w + i] = x;
            } else {
                dst[w + i] |= x >> sh;
                dst[w + i + 1] |= x << (64 - sh);
            }
        }
        // clear bits past the end (src words beyond nb are zero-padded by cfhost)
        bases = total;
    }

    uint16_t k_, l_;
    uint8_t cutoff_;
    cfgpu_handle* h_ = nullptr;
    bool constructed_ = false;
    std::vector<cfgpu_bucket_stat> stats_;
};

}  // namespace cuttlefish_amd
