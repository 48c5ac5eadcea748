// cf_oracle — CPU restatement of Cuttlefish 3's minimizer-partitioned subgraph
// construction, used ONLY as test infrastructure: the parity checker for the
// HIP path and the CPU baseline timed beside it. Nothing in the product
// (libcfgpu) links, loads or calls this code.
//
// Each function cites the reference file:line whose behaviour it restates
// (paths relative to the reference repository root).
//
// Parity status of each part (how it is pinned, see tests/test_oracle_*.py):
//   wyhash / lmer_hash        — golden vectors produced by the reference's own
//                               include/wyhash/wyhash.h compiled in oracle/_ref.
//   MinIter (sliding min)     — brute force over all l-mers of each (k-1)-mer.
//   parse / fragments         — closed form + hand cases.
//   partition_fragment        — invariants (each (k+1)-mer in exactly one record
//                               whose g = g(middle (k-1)-mer), disc <=> g change,
//                               piece length <= w, labels decode to input) and a
//                               second, data-parallel formulation.
//   construct_literal         — closed form (canonical (k+1)-mer counting with the
//                               self-loop / homopolymer exceptions) and a
//                               brute-force global graph over all (k+1)-mers.
//   contract_subgraph         — lm-tigs glued across subgraphs plus trivial unitigs
//                               equal the brute-force global maximal unitigs;
//                               discontinuity-edge endpoints match the labels.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace cforacle {

// Cuttlefish's wyhash salt (include/Minimizer_Utility.hpp:26, include/Kmer.hpp:43).
extern const uint64_t kWySalt[4];

// wyhash "final version 4.2" (include/wyhash/wyhash.h:118-142), WYHASH_CONDOM 1,
// 128-bit multiply, little-endian reads.
uint64_t wyhash(const void* key, size_t len, uint64_t seed, const uint64_t* secret);

// Minimizer_Utility::hash, release build (include/Minimizer_Utility.hpp:97-107).
uint64_t lmer_hash(uint64_t lmer);

// Min_Iterator::lmer_hash (include/Minimizer_Iterator.hpp:291-296).
uint64_t canon_lmer_hash(uint64_t lmer, uint64_t lmer_bar);

// DNA_Utility (include/DNA_Utility.hpp:110-191).
bool is_dna(char c);
uint8_t map_base(char c);            // A/a=0 C/c=1 G/g=2 T/t=3, else N=4
uint8_t map_base_unchecked(char c);  // ((c>>2)^(c>>1))&3
inline uint8_t complement(uint8_t b) { return b < 4 ? uint8_t(3 - b) : (b == 4 ? 4 : 5); }

constexpr uint8_t BASE_N = 4;
constexpr uint8_t BASE_E = 5;

// Records: FASTA = header line starting '>', sequence lines concatenated (split
// on '\n' only); FASTQ = 4-line records, sequence = line 2.
// (RabbitFX include/RabbitFX/io/Formater.cpp:63-79, 87-130, 238-263.)
std::vector<std::string> parse_fasta(const char* buf, size_t n);
std::vector<std::string> parse_fastq(const char* buf, size_t n);

// A fragment is a maximal run of ACGT/acgt of length >= k+1 inside a record
// (src/Graph_Partitioner.cpp:476-501).  Returns (begin, length) pairs.
std::vector<std::pair<uint64_t, uint64_t>> fragments(const char* seq, size_t len, unsigned k);

// Two-stack sliding-window minimum of canonical l-mer hashes over the (k-1)-mers
// of a fragment (include/Minimizer_Iterator.hpp:299-378). Output H[p] for each
// (k-1)-mer start p in [0, frag_len - (k-1)].
void minimizer_hashes_literal(const char* frag, size_t frag_len, unsigned km1, unsigned l,
                              std::vector<uint64_t>& H);
// Brute force: H[p] = min over the w l-mers of the (k-1)-mer at p.
void minimizer_hashes_brute(const char* frag, size_t frag_len, unsigned km1, unsigned l,
                            std::vector<uint64_t>& H);

// One weak super k-mer (src/Graph_Partitioner.cpp:541-595).
struct SuperKmer {
    uint32_t g;        // subgraph id = H & (G-1)  (include/Subgraphs_Manager.hpp:130)
    uint64_t off;      // offset of its first base inside the fragment
    uint32_t len;      // length in bases (weak: includes the joined neighbour bases)
    bool l_disc, r_disc;
};

// Sequential split of one fragment (src/Graph_Partitioner.cpp:503-597).
void partition_fragment(const char* frag, size_t frag_len, unsigned k, unsigned l, uint32_t G,
                        std::vector<SuperKmer>& out);
// Data-parallel formulation: run starts from a max-scan over g changes, pieces
// cut every w (k-1)-mers from the run start. Must equal partition_fragment.
void partition_fragment_scan(const char* frag, size_t frag_len, unsigned k, unsigned l, uint32_t G,
                             std::vector<SuperKmer>& out);

// Super_Kmer_Attributes<false> 16-bit pack (include/Super_Kmer_Attributes.hpp:25-55,108-113).
inline uint16_t att_pack(uint32_t len, bool l_disc, bool r_disc) {
    return uint16_t((len & 0xFF) | (uint32_t(l_disc) << 8) | (uint32_t(r_disc) << 9));
}
// Label word count W = ceil((2(k-1)-l+2)/32) (src/Super_Kmer_Chunk.cpp:11-15).
inline unsigned label_words(unsigned k, unsigned l) { return ((2 * (k - 1) - l + 2) + 31) / 32; }
// MSB-aligned label: word W-1 holds bases 0..31, base i at bits 62-2(i%32) of
// word W-1-i/32 (include/Super_Kmer_Chunk.hpp:394-405). Words/bits past `len`
// are written as zero here (the reference leaves them unspecified).
void encode_label(const char* seq, uint32_t len, unsigned W, uint64_t* words);
// Subgraph::get_base (include/Subgraph.hpp:403-411).
inline uint8_t label_base(const uint64_t* label, unsigned W, size_t idx) {
    return uint8_t((label[(W - 1) - (idx >> 5)] >> (62 - ((idx & 31) << 1))) & 3);
}

// k-mer value, Kmer<k> layout (include/Kmer.hpp:45-54): w[0] holds the last 32
// bases (least significant), base 0 of the k-mer is the most significant.
struct Km {
    std::array<uint64_t, 4> w{};
    bool operator==(const Km& o) const { return w == o.w; }
    bool operator!=(const Km& o) const { return w != o.w; }
};
struct KmOps {
    unsigned k, nw;
    uint64_t top_mask;   // valid bits of w[nw-1]
    explicit KmOps(unsigned k);
    bool less(const Km& a, const Km& b) const;           // Kmer::operator< (Kmer.hpp:509-521)
    void roll_fwd(Km& x, uint8_t b) const;                // roll_to_next_kmer (Kmer.hpp:609-621)
    void roll_rc(Km& x, uint8_t b) const;                 // reverse-complement side of the roll
    Km rc(const Km& x) const;                             // as_reverse_complement (Kmer.hpp:479-506)
    uint8_t base_at(const Km& x, unsigned i) const;       // base i (0 = first)
    std::string str(const Km& x) const;
    Km from_codes(const uint8_t* c) const;
};

// State_Config<false> (include/State_Config.hpp:24-268).
struct State {
    uint32_t e_f = 0;    // 8 saturating 4-bit counters, (side==front?0:16)+4*base
    uint8_t status = 0;  // 1 visited, 2 front-disc, 4 back-disc
};
// Edge_Frequency::add_edge (State_Config.hpp:207-217).
inline void add_edge(uint32_t& e_f, unsigned side, uint8_t base) {
    const uint32_t off = (side == 0 ? 0 : 16) + 4 * base;
    const uint32_t f = (e_f >> off) & 0xF;
    if (f < 0xF) e_f = (e_f & ~(0xFu << off)) | ((f + 1) << off);
}
// Edge_Frequency::edge_at (State_Config.hpp:232-252): 0..3 base, 4 N, 5 E.
uint8_t edge_at(uint32_t e_f, unsigned side, unsigned f_th);

struct VertexState {
    uint32_t g;
    Km key;
    State st;
};

// Per-subgraph construct, literal restatement of Subgraph::construct
// (src/Subgraph.cpp:44-108) + HT_Router::update (include/Subgraph.hpp:546-560).
// `att` are packed attributes, `labels` W words per record, all of subgraph g.
void construct_literal(unsigned k, unsigned W, uint32_t g, size_t n, const uint16_t* att,
                       const uint64_t* labels, std::vector<VertexState>& out);
// The same construct over a worker's reusable open-addressing map (the CPU
// baseline's shape: the reference's ankerl::unordered_dense class).
class FlatMap {
  public:
    FlatMap();
    ~FlatMap();
    FlatMap(const FlatMap&) = delete;
    FlatMap& operator=(const FlatMap&) = delete;
    void clear(size_t expect);
    State& at(const Km& x);
    void emit(uint32_t g, std::vector<VertexState>& out) const;
  private:
    struct Impl;
    Impl* p;
};
void construct_flat(unsigned k, unsigned W, uint32_t g, size_t n, const uint16_t* att, const uint64_t* labels,
                    FlatMap& M, std::vector<VertexState>& out);
// Closed form: canonical (k+1)-mer counts (both orientations) with the
// palindrome and homopolymer exceptions, then e_f = min(15,count) per slot.
void construct_closed_form(unsigned k, unsigned W, uint32_t g, size_t n, const uint16_t* att,
                           const uint64_t* labels, std::vector<VertexState>& out);

// Sort helper: by (g, key).
void sort_states(unsigned k, std::vector<VertexState>& v);

// One unitig extracted by Subgraph::contract (src/Subgraph.cpp:184-270).
enum : uint32_t { UNITIG_CYCLE = 1, UNITIG_LEFT_EXIT = 2, UNITIG_RIGHT_EXIT = 4 };
struct Unitig {
    std::string label;       // trivial: the reference's output orientation; lm-tig: v_hat's
    uint32_t flags = 0;      // UNITIG_*; an exit flag marks an lm-tig end at a discontinuity
    uint64_t n_vertices = 0;
    Km u, v;                 // lm-tig: discontinuity-edge endpoints (left, right) ...
    unsigned u_side = 2, v_side = 2;  // ... and their sides (0 front, 1 back, 2 phi)
};
struct ContractResult {
    std::vector<Unitig> unitigs;
    uint64_t isolated = 0, phantom = 0, trivial = 0, icc = 0, disc_edges = 0, vertex_count = 0, label_sz = 0;
};
// Subgraph::contract over one subgraph's vertex states (cf_oracle_contract.cpp).
void contract_subgraph(unsigned k, unsigned f_th, const VertexState* v, size_t n, ContractResult& out);
// Orients an lm-tig so that its first k characters are at most those of its
// reverse complement (ends swapped accordingly); cycles are left as they are.
void normalise_unitig(unsigned k, Unitig& u);

}  // namespace cforacle
