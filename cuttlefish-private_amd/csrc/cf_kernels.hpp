// Kernel launch interface between the C-ABI layer (cfgpu.cpp) and the HIP
// kernels.  Every launcher validates nothing: the callers check shapes first.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cf_device.hpp"

namespace cfgpu {

constexpr int P_TILE = 2048;            // (k-1)-mer positions per partition workgroup
constexpr int P_WG = 256;               // partition workgroup size
constexpr uint32_t G_INVALID = 0xFFFFFFFFu;
constexpr uint32_t G16_INVALID = 0xFFFFu;  // per-position g array: no (k-1)-mer starts here
constexpr int N_FEAT = 19;              // 16-bin minimizer histogram + 3 size features

struct PartitionArgs {
    const uint64_t* packed;    // MSB-first 2-bit stream, padded with >= 2 zero words
    uint64_t n_words;          // readable words of `packed` (incl. padding)
    uint64_t n_bases;
    const uint64_t* frag_off;  // n_frag + 1 base offsets, frag_off[n_frag] == n_bases
    uint64_t n_frag;
    uint64_t* tile_frag;       // per minimizer tile (ceil(n_bases / P_TILE)): its first fragment
    uint32_t k, l, w;          // w = k - l: (k-1)-mers per piece at most
    uint32_t g_mask, g_bits;   // G - 1, log2 G
    WyConst wy;
};

// Pass 1: per position p the minimizer hash's g (gbin[p], G16_INVALID where
// no (k-1)-mer starts) and the 4 hash bits above it (nibble p % 8 of
// mbn[p / 8]), and per tile the last g-change position (+1, 0 = none).
void launch_minimizer_g(const PartitionArgs& a, uint16_t* gbin, uint32_t* mbn, uint64_t* tile_lc, hipStream_t s);

// Records are bucketed in two levels.  Level 1 splits by coarse bin
// g >> FINE_BITS: the input is cut into contiguous spans (count spans, and
// emit spans of H_SPLIT count spans, one workgroup each); level 2 splits each
// coarse bin into its 2^FINE_BITS buckets.
constexpr int FINE_BITS = 6;
constexpr int N_COARSE = 256;
constexpr int E_WG = 1024;         // emit workgroup (its LDS staging fills most of a CU)
constexpr int E_TILE = P_TILE;     // emit spans are whole minimizer tiles
constexpr int E_SPANS_PER_CU = 1;  // emit spans (workgroups) per CU
constexpr int H_WG = 256;          // count workgroup, one minimizer tile per step
static_assert((N_COARSE << FINE_BITS) == 16384, "coarse x fine bins cover the 16384 subgraphs");
constexpr int H_SPLIT = 8;  // count spans per emit span
// Level-1 records: a header array (one u32 per record: att with the 4 hash
// bits above g in its free bits 10-13) and a label array of l1_label_words(W)
// words per record (W padded to even, so labels are 16-byte aligned pairs).
__host__ __device__ constexpr uint32_t l1_label_words(uint32_t W) { return (W + 1) & ~1u; }
// Offset in u64 words of the label array in a level-1 buffer of n records
// (after 4n header bytes, 16-byte aligned).
__host__ __device__ constexpr uint64_t l1_label_offset(uint64_t n) { return (n + 3) / 4 * 2; }
constexpr uint32_t L1_MB_SHIFT = 10;
struct PieceSpans {
    uint64_t per_count, per_emit;  // minimizer tiles (P_TILE positions) per count / emit span
    uint32_t n_count, n_emit;      // spans (workgroups)
};
PieceSpans pieces_spans(uint64_t n_bases, uint32_t n_cu);
// Pass 2: every tile's piece descriptors into desc[t * P_TILE ..] (tile_n[t]
// of them; desc holds P_TILE per tile), record counts per (coarse bin, count
// span) in hist[N_COARSE][n_count], scanned in place; coff[N_COARSE+1] =
// coarse bin offsets (coff[N_COARSE] = records); ctot[N_COARSE] scratch.
void launch_pieces_hist(const PartitionArgs& a, const uint16_t* gbin, const uint32_t* mbn, const uint64_t* tile_lc,
                        const PieceSpans& ps, uint32_t* hist, uint64_t* ctot, uint64_t* coff, uint64_t* desc,
                        uint32_t* tile_n, hipStream_t s);
// Pass 3: the described pieces as records into the level-1 arrays (hdr1,
// lab1), grouped by coarse bin.
void launch_pieces_emit(const PartitionArgs& a, const uint64_t* desc, const uint32_t* tile_n, const PieceSpans& ps,
                        const uint32_t* hist, const uint64_t* coff, uint32_t W, uint32_t* hdr1, uint64_t* lab1,
                        hipStream_t s);
// Level 2: level-1 records into per-bucket order; boff[G+1] bucket offsets.
void launch_bucket_fine(const uint32_t* hdr1, const uint64_t* lab1, const uint64_t* coff, uint32_t W, uint64_t* boff,
                        uint32_t* att, uint64_t* labels, uint8_t* mbin, hipStream_t s);

// Buckets of separately partitioned input chunks gathered into one bucket
// array: bucket b receives chunk 0's records of b, then chunk 1's, and so on;
// chunk c's records of bucket b start at dst_off[c * G + b] of the output.
struct ChunkSrc {
    const uint32_t* att;
    const uint64_t* labels;
    const uint8_t* mbin;
    const uint64_t* boff;  // G + 1 offsets
};
void launch_merge_chunks(const ChunkSrc* src, uint32_t n_chunks, uint32_t G, uint32_t W, const uint64_t* dst_off,
                         uint32_t* att, uint64_t* labels, uint8_t* mbin, hipStream_t s);

// Per-bucket statistics: [n_rec, n_kmer_instances, n_bases, 0] as u64 and the
// MLP feature vector (N_FEAT floats).
void launch_bucket_features(const uint64_t* off, const uint32_t* att, const uint8_t* mbin,
                            uint32_t G, uint32_t k, uint64_t* stats, float* feat, hipStream_t s);
// The same statistics and features for buckets that leave device memory chunk by
// chunk: each chunk's per-bucket sums added into `sums` (SUM_WORDS u32 per bucket,
// zeroed before the first chunk), then computed from the sums.
constexpr uint32_t SUM_WORDS = 22;  // 16 histogram words, records, k-mer instances, bases (u64)
void launch_bucket_sums(const uint64_t* off, const uint32_t* att, const uint8_t* mbin, uint32_t G, uint32_t k,
                        uint32_t* sums, hipStream_t s);
void launch_features_from_sums(const uint32_t* sums, uint32_t G, uint32_t k, uint64_t* stats, float* feat,
                               hipStream_t s);

// ---------------------------------------------------------------------------
struct ConstructArgs {
    const uint32_t* att;        // bucketed records: len | l_disc<<8 | r_disc<<9 | g<<16
    const uint64_t* labels;     // W words per record
    uint32_t W, k;
    const uint64_t* bucket_off; // G+1 record offsets
    const uint64_t* tbl_ptr;    // per bucket table base (device address)
    const uint64_t* tbl_cap;    // per bucket capacity (slots)
    const uint32_t* tbl_limit;  // per bucket max distinct keys before overflow
    uint32_t* n_vert;           // per bucket distinct vertex count
    uint32_t* overflow;         // per bucket overflow flag
    const uint32_t* item_bucket;  // work items: bucket id
    const uint32_t* item_first;   //   first record (relative to the bucket)
    const uint32_t* item_count;   //   record count
    uint32_t n_items;
    uint32_t* queue_head;       // persistent mode: next item
    const uint32_t* mult;       // per record multiplicity (null: every record counts once)
    const uint64_t* bucket_end; // per bucket end of its records (null: bucket_off[b + 1])
    uint64_t* spill_key;        // LDS kernel, all-pass items: per workgroup a region of spill_cap
    uint32_t spill_cap;         //   entries (KW key words + one meta word each), P sub-lists of
                                //   spill_cap / P entries for a P-pass item
    uint32_t cutoff;            // LDS kernel: edge count threshold of the active / isolated split
    uint64_t* seg;              // LDS kernel: per bucket SEG_MAX pass runs (start | active << 32)
    uint32_t* iso;              // LDS kernel: per bucket isolated and phantom vertex counts
};

// The LDS kernel writes each pass's vertices as one run of the bucket's table,
// the vertices in an edge at the cutoff (active) first; for buckets of at most
// SEG_MAX passes it records each run's start and active count, and it counts
// the bucket's isolated (and phantom: isolated and discontinuous) vertices, so
// that contract reads only the active vertices.
constexpr uint32_t SEG_MAX = 64;

// Entry stride in bytes for KW key words: KW u64 key words, u32 e_f, u32 status.
inline uint32_t entry_bytes(uint32_t KW) { return 8 * KW + 8; }

void launch_clear_tables(uint8_t* tables, uint64_t n_slots, uint32_t KW, hipStream_t s);
// Non-persistent: one workgroup per work item in [item0, item0 + n).
void launch_construct(const ConstructArgs& a, uint32_t KW, uint32_t item0, uint32_t n,
                      uint32_t wg, hipStream_t s);
// Persistent: `grid` workgroups pull items from queue_head until exhausted.
void launch_construct_persistent(const ConstructArgs& a, uint32_t KW, uint32_t grid, uint32_t wg,
                                 hipStream_t s);
// LDS-resident construct: one work item = one hash pass of one bucket
// (item_first = pass index, item_count = number of passes P; the bucket's
// table holds P * lds_capacity(KW) slots), or every pass of one bucket
// (item_first = LDS_ALL_PASSES, 2 <= P <= LDS_ALL_PASS_MAX): the records are
// walked once, every instance going to its pass's sub-list of the
// workgroup's spill region, then each pass probes its own sub-list.
// Persistent over queue_head; grid at most the number of workgroups the
// spill regions are sized for.
// The LDS holds only the pass table: 8*KW key bytes + 4 e_f bytes + 1 status
// byte per slot, sized to fill the 160 KiB of a gfx950 CU.
constexpr int LDS_WG = 1024;
constexpr uint32_t LDS_ALL_PASSES = 0xFFFFFFFFu;
constexpr uint32_t LDS_ALL_PASS_MAX = 64;  // sub-list counters kept in LDS beside the table
constexpr int lds_capacity(int KW) { return KW == 1 ? 12544 : KW == 2 ? 7744 : KW == 3 ? 5632 : 4416; }
void launch_construct_lds(const ConstructArgs& a, uint32_t KW, uint32_t grid, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
// n_vert[b] = overflow[b] = 0 (and iso[2b] = iso[2b+1] = 0 when given) for every b in list[0..n).
void launch_zero_buckets(const uint32_t* list, uint32_t n, uint32_t* n_vert, uint32_t* overflow, uint32_t* iso,
                         hipStream_t s);

// Compact the occupied slots of bucket tables into key/e_f/status/dfa arrays.
// out_cursor[i] receives the number of occupied entries of buckets[i]; they
// are written from out_off[i] up to (excluding) out_off[i+1].
void launch_export(const uint64_t* tbl_ptr, const uint64_t* tbl_cap,
                   uint32_t KW, const uint32_t* buckets, uint32_t n_buckets,
                   const uint64_t* out_off, uint32_t* out_cursor, uint64_t* keys_out,
                   uint32_t* ef_out, uint8_t* st_out, uint8_t* dfa_out, uint32_t cutoff,
                   hipStream_t s);

// ---------------------------------------------------------------------------
// Exact record dedup per bucket (dedup.hip): distinct (len, disc bits, label)
// records with their multiplicities, written to the bucket's own region of
// the out_* arrays; out_count[b] distinct records.  One work item = one hash
// pass of one bucket; the LDS table holds dedup_capacity(key words) records.
// A key is the W label words plus a word for len | disc bits | hash, or, when
// the label's last word holds at most 15 bases (records are at most max_len =
// min(255, 2k - l) bases), just the W label words: the last one shifted right
// by one (bit 63 stays free for the claim flag) with len | disc bits in bits
// 0-9 and 22 hash bits in bits 10-31.
constexpr int DD_WG = 1024;
constexpr int dedup_capacity(int KR) { return ((163840 - 2048) / (8 * KR + 4)) & ~63; }
constexpr bool dedup_packed(uint32_t W, uint32_t max_len) { return max_len - 32 * (W - 1) <= 15; }
constexpr int dedup_key_words(uint32_t W, uint32_t max_len) { return int(W) + (dedup_packed(W, max_len) ? 0 : 1); }
struct DedupArgs {
    const uint32_t* att;
    const uint64_t* labels;
    const uint64_t* bucket_off;
    uint32_t* overflow;      // per bucket: a pass overflowed (records are then copied)
    uint32_t* out_count;     // per bucket distinct records (zeroed before the launch)
    unsigned long long* out_kmers;  // per bucket k-mer instances of those records (zeroed)
    uint32_t k;
    uint32_t* out_att;
    uint64_t* out_labels;
    uint32_t* out_mult;
    const uint32_t* item_bucket;
    const uint32_t* item_pass;
    const uint32_t* item_passes;
    uint32_t n_items;
    uint32_t* queue_head;
    uint8_t* region;  // per workgroup DD_REGION_RECORDS * 5 bytes (all-pass items)
};
// An all-pass dedup item (item_pass == DD_ALL_PASSES): a bucket of 2 ..
// DD_ALL_PASS_MAX passes and at most DD_REGION_RECORDS records, its passes in
// one workgroup.
constexpr uint32_t DD_ALL_PASSES = 0xFFFFFFFFu;
constexpr uint32_t DD_ALL_PASS_MAX = 64;
constexpr uint32_t DD_REGION_RECORDS = 1u << 16;
void launch_dedup(const DedupArgs& a, uint32_t W, bool packed, uint32_t grid, hipStream_t s);
// Buckets in list[0..n): records copied unchanged with multiplicity 1.
void launch_dedup_copy(const DedupArgs& a, const uint32_t* list, uint32_t n, uint32_t W, hipStream_t s);
void launch_dedup_end(const uint64_t* off, const uint32_t* cnt, uint32_t G, uint64_t* end, hipStream_t s);

// ---------------------------------------------------------------------------
// Subgraph::contract on the device tables (contract.hip).  Every built
// subgraph's vertices in no edge at the cutoff are counted; the others go to
// a per-subgraph table (LDS when they fit, else a per-workgroup global
// scratch region), and every maximal unitig of the subgraph is written as an
// ASCII label with a record.
constexpr int CT_WG = 1024;
// LDS table slots: 8*KW key bytes + 4 state bytes + 2 x 4 list-ranking bytes each
// (a slot's two directed states are numbered below 2^14).
constexpr int ct_lds_capacity(int KW) { return KW == 1 ? 7360 : KW == 2 ? 5248 : KW == 3 ? 4096 : 3328; }
// Dense tables (non-isolated vertices beyond an LDS table): at most this many
// vertices, and per workgroup this many scratch bytes.
__host__ __device__ constexpr uint32_t ct_big_max_vertices() { return 32767; }
__host__ __device__ constexpr uint64_t ct_big_scratch_bytes(uint32_t KW) {
    return uint64_t(ct_big_max_vertices()) * (8 * KW + 4 + 16 + 16);
}
// Wide dense tables (beyond ct_big_max_vertices(), up to ct_wide_max_vertices()):
// vertex arrays, 64-bit ranking words and the key index all in the workgroup's
// global scratch; index slots for nv vertices (load <= 0.8) and scratch bytes for
// a workgroup holding up to nv vertices.
__host__ __device__ constexpr uint32_t ct_wide_max_vertices() { return (1u << 20) - 1; }
__host__ __device__ constexpr uint64_t ct_wide_index_slots(uint64_t nv) { return ((nv * 5 / 4 + 64) + 63) & ~63ull; }
__host__ __device__ constexpr uint64_t ct_wide_scratch_bytes(uint32_t KW, uint64_t nv) {
    // keys, state words, three ranking buffers of 2 nv u64, four u32 arrays, the index
    return (((nv * (8 * KW + 4) + 7) & ~7ull) + 3 * 16 * nv + 16 * nv + 4 * ct_wide_index_slots(nv) + 255) & ~255ull;
}
// Global scratch table slots for a subgraph of nv vertices (load <= 0.7).
__host__ __device__ constexpr uint64_t ct_global_capacity(uint64_t nv) { return ((nv * 10 / 7 + 64) + 63) & ~63ull; }
struct UnitigRec {             // layout of cfgpu_unitig
    uint64_t label_off;
    uint32_t label_len;
    uint16_t g;
    uint16_t flags;            // 1 cycle, 2 left end exits, 4 right end exits
};
enum : uint32_t { CT_ISOLATED, CT_PHANTOM, CT_TRIVIAL, CT_ICC, CT_DISC_EDGES, CT_VERTICES, CT_LABEL_BYTES, CT_FLAGS };
struct ContractArgs {
    const uint64_t* tbl_ptr;       // construct tables (entry_bytes(KW) per slot)
    const uint64_t* tbl_cap;
    const uint32_t* n_vert;        // per bucket distinct vertices (global scratch sizing)
    const uint32_t* list;          // buckets for the LDS pass
    uint32_t n_list;
    uint32_t k, cutoff;
    uint32_t* queue_head;          // [0] LDS pass, [1] dense pass
    uint32_t* ovf_list;            // buckets the LDS pass could not hold
    uint32_t* ovf_count;
    uint32_t* walk_head;           // global pass
    uint32_t* ovf2_list;           // buckets the dense pass could not hold
    uint32_t* ovf2_count;
    uint32_t* wide_head;           // wide dense pass (over ovf2_list)
    uint32_t* ovf3_list;           // buckets the wide dense pass could not hold (the global pass's)
    uint32_t* ovf3_count;
    uint64_t wide_cap;             // vertices a workgroup's scratch holds in the wide pass (0: no wide pass)
    uint8_t* scratch;              // dense and global passes: scratch_per_wg bytes per workgroup
    uint64_t scratch_per_wg;
    uint32_t* counts;              // per bucket 8 counters (CT_*), written by the pass that finishes it
    UnitigRec* recs;
    uint64_t rec_cap;
    char* labels;
    uint64_t lab_cap;
    unsigned long long* cursors;   // [0] records, [1] label bytes (totals, even past capacity)
    uint32_t* retry;               // set when the output capacity was short
    uint32_t* error;               // set on an inconsistent table
    const uint64_t* seg;           // per bucket SEG_MAX runs (start | active << 32) of the LDS construct
    const uint32_t* nseg;          // per bucket runs recorded (0: scan the whole table)
    const uint32_t* iso;           // per bucket isolated and phantom counts (buckets with runs)
};
// Four passes: LDS tables for every listed bucket that fits, dense tables for
// the rest that fit those, wide dense tables for the rest that fit those
// (wide_cap > 0), then global scratch tables (the counts of the later passes
// are read on the device).
// part 0: the LDS and dense passes; part 1: the wide dense pass (when
// wide_cap > 0) and the global pass, over the dense pass's overflow.
void launch_contract(const ContractArgs& a, uint32_t KW, uint32_t grid_lds, uint32_t grid_global, int part,
                     hipStream_t s);

// ---------------------------------------------------------------------------
// The compacted graph (cdbg.hip): contract's unitigs glued into the maximal
// unitigs of the whole graph.  Phases 0 ends binned and paired, 1 chain
// successors, 2 one pointer-jumping pass (sets *changed), 3 placement, 4 label
// copies.
constexpr uint32_t CD_MAX_BINS = 8192;  // pairing bins (LDS counters in the binning passes)
struct CdbgArgs {
    const UnitigRec* recs;
    uint64_t n;
    const char* labels;
    uint32_t k;
    uint32_t* partner;          // per piece end: the end across its vertex (all-ones: none, or not exiting)
    uint32_t* bin_hist;         // pairing bins: exiting ends per (bin, block), scanned in place
    uint32_t* bin_tot;          //   per bin total
    uint32_t* bin_off;          //   bin offsets (1 << nb_bits) + 1
    uint64_t* bin_key;          //   vertex keys of the exiting ends, bin after bin (KW words each)
    uint32_t* bin_id;           //   their end ids << 1 | side
    uint32_t nb_bits;           // log2 bins (<= log2 CD_MAX_BINS)
    uint32_t n_blocks;          // pass 1 / pass 2 blocks, per_block pieces each
    uint32_t per_block;
    uint64_t* kend;             // per piece: its first k characters and the rc of its last k (KW words each)
    uint64_t* jp;               // per state: the state it points at | label characters up to it << 30
    uint32_t* settled;          // a bit per state: settled (zeroed before the ranking passes)
    uint64_t* end_bits;         // a bit per state: a chain end (k_chain_bits)
    uint64_t* first_bits;       // a bit per state: an end whose reading comes first
    uint64_t total_chars;       // label bytes of all pieces
    uint32_t* place_head;       // per piece: first piece of its chain (all-ones: on a cycle)
    uint64_t* place_pos;        //   ... its offset in the chain | ends the chain << 62 | reversed << 63
    uint32_t* piece_len;        // per piece: label characters after its first k-mer (place_pos's memory until k_place)
    uint64_t* head_off;         // per chain head: output offset
    char* out;
    uint64_t out_cap;
    uint64_t* out_rec;          // (offset, length) per maximal unitig
    uint64_t out_n_cap;
    unsigned long long* cursors;  // [0] unitigs, [1] bytes
    uint32_t* retry;
    uint32_t* error;
    uint32_t* n_cycle_pieces;
};
void launch_cdbg_phase(const CdbgArgs& a, uint32_t KW, uint32_t grid, int phase, uint32_t src, uint32_t* changed,
                       hipStream_t s);
// list[0..*n) = pieces on cycles
void launch_cdbg_cycle_list(const CdbgArgs& a, uint32_t grid, uint32_t* list, uint32_t* n, hipStream_t s);

// ---------------------------------------------------------------------------
// FP32 MLP bucket-size predictor: N_FEAT -> H -> H -> 1 (ReLU), predicts
// log(distinct vertices / k-mer instances) of a bucket.
constexpr int MLP_H = 64;
constexpr int MLP_PARAMS = N_FEAT * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1;

constexpr int MLP_TRAIN_MAX = 128;  // training buckets (one workgroup holds all activations)
int mlp_param_floats();             // MLP_PARAMS + 2 * N_FEAT (standardisation)
// `adam` holds 2 * MLP_PARAMS floats of optimiser state.
void launch_mlp_train(float* params, const float* feat, const uint32_t* idx, const float* target,
                      uint32_t n, uint32_t epochs, float lr, uint32_t seed, float* adam,
                      float* loss_out, hipStream_t s);
void launch_mlp_infer(const float* params, const float* feat, uint32_t n, float* pred,
                      hipStream_t s);

}  // namespace cfgpu
