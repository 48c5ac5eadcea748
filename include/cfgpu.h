/*
 * cfgpu.h — C ABI of libcfgpu, the MI355X (gfx950) implementation of Cuttlefish 3's
 * minimizer-partitioned subgraph construction.
 *
 * It replaces, behind the reference's C++ class contract, this call sequence of
 * dBG_Contractor::construct (src/dBG_Contractor.cpp:43-66, 75-113):
 *
 *   Graph_Partitioner<k,...>(G, logistics, l).partition();   -> cfgpu_partition
 *   G.finalize();                                             -> (buckets stay in HBM)
 *   G.process(): per g  Subgraph(B,G,op_buf,space).construct() -> cfgpu_construct
 *                        b.remove()                            -> cfgpu_release
 *                        contract() reads the table            -> cfgpu_table_view_get /
 *                                                                 cfgpu_table_copy
 *                        contract() itself                     -> cfgpu_contract
 *   Discontinuity_Graph_Contractor::contract, Contracted_Graph_Expander::expand,
 *   Unitig_Collator::collate                                   -> cfgpu_cdbg
 *
 * Conventions: every function returns CFGPU_OK (0) or a negative CFGPU_E_* code and
 * never throws; the message of the last failure is available from cfgpu_last_error.
 * A handle is used from one host thread at a time.  The handle owns all device
 * memory; host buffers passed in are only read during the call.
 */
#ifndef CFGPU_H
#define CFGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CFGPU_OK = 0,
    CFGPU_E_ARG = -1,      /* invalid argument or input shape */
    CFGPU_E_HIP = -2,      /* HIP runtime error (includes "no GPU") */
    CFGPU_E_NOMEM = -3,    /* device allocation failed */
    CFGPU_E_STATE = -4,    /* call out of order (e.g. construct before partition) */
    CFGPU_E_OVERFLOW = -5  /* table overflow not recovered (internal retry cap hit) */
};

/* Build parameters.  k, l and cutoff carry the meaning of `cuttlefish build -k`,
 * `--min-len` and `-c` (src/commands.cpp:38-106, include/Build_Params.hpp):
 * odd k with l < k <= 127, 1 <= l <= 32, 1 <= cutoff <= 15.  subgraph_count must be
 * 16384, the reference's fixed subgraph count (include/Atlas.hpp:33-35). */
typedef struct cfgpu_params {
    uint32_t k;
    uint32_t l;
    uint32_t cutoff;
    uint32_t subgraph_count;
    int32_t device;
    uint32_t reserved;
} cfgpu_params;

typedef struct cfgpu_handle cfgpu_handle;

/* Opens a handle on `device`.  Replaces the construction of Subgraphs_Manager
 * (src/Subgraphs_Manager.cpp:26-52) and Edge_Frequency::set_edge_threshold
 * (src/State_Config.cpp:10-14, called at src/dBG_Contractor.cpp:24). */
int cfgpu_open(const cfgpu_params* params, cfgpu_handle** out);
/* The parameters the handle was opened with. */
int cfgpu_params_get(const cfgpu_handle* h, cfgpu_params* out);
int cfgpu_close(cfgpu_handle* h);
/* Copies the last error message (NUL-terminated, truncated to n bytes). */
int cfgpu_last_error(const cfgpu_handle* h, char* buf, size_t n);

/* Input: the fragments of the input (maximal ACGT runs of length >= k+1,
 * src/Graph_Partitioner.cpp:476-501) concatenated into one stream of 2-bit bases
 * (A=0 C=1 G=2 T=3) packed MSB-first, 32 per word: base j at bits 63-2(j%32)..62-2(j%32)
 * of word j/32.  frag_off[n_frag+1] gives each fragment's first base; frag_off[0] == 0
 * and frag_off[n_frag] == n_bases.  With CFGPU_INPUT_DEVICE the two pointers are device
 * memory and `packed` must hold ceil(n_bases/32)+2 readable words (the last two zero).
 *
 * Replaces Graph_Partitioner::partition()/process_chunk (src/Graph_Partitioner.cpp:45,
 * 424-614) and Subgraphs_Manager::add_super_kmer/finalize
 * (include/Subgraphs_Manager.hpp:93, 113; src/Subgraphs_Manager.cpp:84-95): the weak
 * super k-mers are left in 16384 device-resident buckets in the reference record
 * format (Super_Kmer_Attributes<false> + MSB-aligned label). */
enum { CFGPU_INPUT_DEVICE = 1, CFGPU_PARTITION_APPEND = 2, CFGPU_PARTITION_SPILL = 4 };
int cfgpu_partition(cfgpu_handle* h, const uint64_t* packed, uint64_t n_bases,
                    const uint64_t* frag_off, uint64_t n_frag, uint32_t flags);
/* Input in chunks: with CFGPU_PARTITION_APPEND, cfgpu_partition partitions one chunk of
 * whole fragments and keeps its records aside; cfgpu_partition_finish gathers every
 * appended chunk's records into the buckets, which are then what one cfgpu_partition of
 * all the chunks' fragments gives (record order inside a bucket aside).  This is the
 * reference's Subgraphs_Manager::finalize after the partitioner's workers have added
 * their chunks' super k-mers (include/Subgraphs_Manager.hpp:93, 113; src/Graph_Partitioner.cpp:
 * 121-184).  Until the finish, the handle holds no buckets; a partition without APPEND
 * or a cfgpu_load_buckets drops the appended chunks. */
int cfgpu_partition_finish(cfgpu_handle* h);

/* External-memory buckets, for inputs whose buckets do not fit device memory.  With
 * CFGPU_PARTITION_SPILL (appending, like CFGPU_PARTITION_APPEND), cfgpu_partition
 * partitions one chunk and moves its buckets out of device memory into the handle's
 * bucket store: one segment per chunk (its records in bucket order and the bucket
 * offsets), in pinned host memory or, after cfgpu_spill_dir, in one file per segment.
 * This replaces the reference's Super_Kmer_Bucket files, to which every partitioner
 * worker flushes its per-bucket chunks (src/Super_Kmer_Bucket.cpp:27-40,
 * include/Super_Kmer_Chunk.hpp:281-351; a segment holds every bucket's chunk of one
 * input chunk, uncompressed: the labels are already 2-bit packed).  cfgpu_partition_finish
 * then computes the bucket statistics and features from the segments without gathering
 * them; the handle's buckets live in the store and none is in device memory.  A
 * partition without SPILL or APPEND, a cfgpu_load_buckets or cfgpu_close empties the
 * store (and removes its files).  Segments of spilled and kept chunks are not mixed. */
int cfgpu_spill_dir(cfgpu_handle* h, const char* dir);  /* NULL or "": pinned host memory */
/* Loads the listed buckets' records from the store into device memory, replacing the
 * buckets there (every other bucket is then empty in device memory, and the tables,
 * unitigs and compacted graph of the previous buckets are dropped): the reader side of
 * the reference's bucket files (Super_Kmer_Bucket::Iterator::read_chunk,
 * src/Super_Kmer_Bucket.cpp:85-95).  cfgpu_construct_buckets then builds any of them;
 * a construct of a bucket of the store that is not loaded is CFGPU_E_STATE. */
int cfgpu_buckets_restore(cfgpu_handle* h, const uint32_t* buckets, uint32_t n);
typedef struct cfgpu_spill_stat {
    uint64_t n_segments;    /* chunks in the store */
    uint64_t n_records;     /* records in the store */
    uint64_t bytes;         /* bytes of the store (headers, labels, minimizer bins) */
    uint32_t in_files;      /* 1: the segments are files under the spill directory */
    uint32_t n_resident;    /* buckets in device memory (the last restore's) */
    uint64_t n_resident_records;
    double t_spill_ms;      /* device-to-store time summed over the segments */
    double t_restore_ms;    /* the last restore: store to device */
} cfgpu_spill_stat;
int cfgpu_spill_stats(cfgpu_handle* h, cfgpu_spill_stat* out);

typedef struct cfgpu_bucket_stat {
    uint64_t n_records;   /* weak super k-mers in the bucket */
    uint64_t n_kmers;     /* k-mer instances: sum(len - k + 1) */
    uint64_t n_bases;     /* sum of record lengths */
    uint64_t n_vertices;  /* distinct canonical k-mers (after construct) */
    uint64_t capacity;    /* table slots (after construct) */
    float predicted;      /* MLP-predicted distinct vertices (learned mode) */
    uint32_t flags;       /* CFGPU_BUCKET_* */
    uint64_t n_records_walked;  /* records the last construct walked (distinct after dedup) */
    uint64_t n_kmers_walked;    /* their k-mer instances */
} cfgpu_bucket_stat;
enum {
    CFGPU_BUCKET_BUILT = 1,       /* table holds the bucket's final states */
    CFGPU_BUCKET_OVERFLOWED = 2,  /* first-attempt table was too small and was rebuilt */
    CFGPU_BUCKET_SAMPLE = 4       /* used as an MLP training sample */
};
/* Per-subgraph statistics (the reference logs these at src/Subgraphs_Manager.cpp:264-305).
 * `out` holds n >= 16384 entries. */
int cfgpu_bucket_stats(cfgpu_handle* h, cfgpu_bucket_stat* out, uint32_t n);

/* Copies every bucket's records to the host: att[n_records] (16-bit attribute pack of
 * Super_Kmer_Attributes<false> in the low half, g in the high half, the reference's
 * in-memory layout, include/Super_Kmer_Attributes.hpp:25-55), labels[n_records*W]
 * (W = ceil((2(k-1)-l+2)/32) words per record, include/Super_Kmer_Chunk.hpp:394-405;
 * bits past the record length are zero) and bucket_off[16385].  Record order inside
 * a bucket is unspecified, as in the reference (parallel workers append). */
int cfgpu_records_copy(cfgpu_handle* h, uint32_t* att, uint64_t* labels, uint64_t* bucket_off,
                       uint64_t cap_records, uint64_t* n_records);

/* Loads bucketed weak super k-mer records (the same layout cfgpu_records_copy returns) in
 * place of a partition: the reader side of the reference's bucket files
 * (Super_Kmer_Bucket::Iterator, include/Super_Kmer_Bucket.hpp:144-175).  Every record of
 * bucket g must carry g in its attribute's high half and a length in [k, 2k-l]; bits of
 * a label past the record length are ignored.  Host pointers. */
int cfgpu_load_buckets(cfgpu_handle* h, const uint32_t* att, const uint64_t* labels,
                       const uint64_t* bucket_off);

/* Builds every non-empty subgraph's vertex table: Subgraph::construct
 * (src/Subgraph.cpp:44-108) for all buckets, scheduled as in Subgraphs_Manager::process
 * (src/Subgraphs_Manager.cpp:113-255).  Modes (results are identical in all modes):
 *  PLAIN            one launch per bucket in bucket order, tables grown and retried;
 *  BATCHED          one persistent work-queue launch (hipGraph) over all buckets,
 *                   largest first, tables sized to the exact upper bound;
 *  BATCHED_LEARNED  as BATCHED with tables sized and ordered by the MLP prediction,
 *                   undersized tables detected exactly and rebuilt. */
enum { CFGPU_MODE_PLAIN = 0, CFGPU_MODE_BATCHED = 1, CFGPU_MODE_BATCHED_LEARNED = 2 };
int cfgpu_construct(cfgpu_handle* h, uint32_t mode);
/* cfgpu_construct of the listed buckets only (duplicates and empty buckets ignored):
 * the reference constructs subgraph by subgraph (Subgraphs_Manager::process_graph,
 * src/Subgraphs_Manager.cpp:147-160), and a caller bounds device memory by building
 * the buckets in batches - construct a batch, contract it, copy its unitigs out, go
 * on (cfhost_cdbg_build glues the batches' unitigs on the host).  Every bucket not
 * listed is left without a table, as after cfgpu_release. */
int cfgpu_construct_buckets(cfgpu_handle* h, uint32_t mode, const uint32_t* buckets, uint32_t n);

/* Device view of subgraph g's table: the contract() side of HT_Router
 * (include/Subgraph.hpp:336-358).  Each entry is key_words u64 key words (Kmer<k>
 * layout, include/Kmer.hpp:45-54: word 0 least significant), a u32 Edge_Frequency word
 * and a u32 status word (State_Config<false>, include/State_Config.hpp:24-163).  Each
 * of the `capacity` entries is a vertex or empty (all-ones in key word key_words-1);
 * a table built in LDS holds exactly the vertices (capacity == n_vertices), one built
 * with global atomics is open addressing.  Entry order is unspecified.  Valid until
 * cfgpu_release(g), the next cfgpu_partition/cfgpu_construct, or cfgpu_close. */
typedef struct cfgpu_table_view {
    const void* entries;
    uint64_t capacity;
    uint32_t entry_bytes;
    uint32_t key_words;
    uint64_t n_vertices;
} cfgpu_table_view;
int cfgpu_table_view_get(cfgpu_handle* h, uint32_t g, cfgpu_table_view* out);

/* Copies the occupied entries of the listed buckets to host arrays, bucket after
 * bucket (entry order inside a bucket unspecified): keys[n*key_words], e_f[n],
 * status[n] and dfa[n] = edge_at(front) | edge_at(back)<<3 | front_disc<<6 |
 * back_disc<<7 under the handle's cutoff (Edge_Frequency::edge_at,
 * include/State_Config.hpp:232-252: 0..3 a base, 4 N, 5 E).  bucket_n[i] receives
 * the count for buckets[i].  Any output pointer may be NULL. */
int cfgpu_table_copy(cfgpu_handle* h, const uint32_t* buckets, uint32_t n_buckets,
                     uint64_t* keys, uint32_t* e_f, uint8_t* status, uint8_t* dfa,
                     uint32_t* bucket_n, uint64_t cap, uint64_t* n_out);

/* As cfgpu_table_copy, but every output (keys, e_f, status, dfa) is a DEVICE array of
 * the caller's and the call allocates no memory: no hipMalloc, hipFree or allocating
 * library routine (its bucket offsets, export cursors and bucket list live in a static
 * device array of 262,400 bytes per GPU, kept until the process ends).  No counterpart in the
 * reference: for callers that cannot allocate (code under stream capture, serving loops,
 * harnesses that forbid it).  n_buckets is at most 16384; bucket_n and n_out are host
 * pointers (may be NULL).  All four outputs are required.  Every handle on one GPU shares
 * that GPU's scratch array: calls that overlap on one GPU (other host threads or
 * handles) race. */
int cfgpu_table_copy_noalloc(cfgpu_handle* h, const uint32_t* buckets, uint32_t n_buckets, uint64_t* keys,
                             uint32_t* e_f, uint8_t* status, uint8_t* dfa, uint32_t* bucket_n, uint64_t cap,
                             uint64_t* n_out);

/* Releases subgraph g's table view (Super_Kmer_Bucket::remove,
 * src/Subgraphs_Manager.cpp:158, frees bucket files; here the view becomes invalid). */
int cfgpu_release(cfgpu_handle* h, uint32_t g);

/* Subgraph::contract (src/Subgraph.cpp:184-270; extract_maximal_unitig and
 * walk_unitig, include/Subgraph.hpp:414-543) of every subgraph the last
 * cfgpu_construct built and that was not released: the maximal unitigs of each
 * subgraph's graph at the handle's cutoff, as ASCII labels on the device.
 *   - a trivial maximal unitig (no end at a discontinuity) in the orientation
 *     and rotation the reference prints it (Maximal_Unitig_Scratch::finalize,
 *     include/Maximal_Unitig_Scratch.hpp:198-222; FASTA_Record.hpp:142-146);
 *   - an lm-tig (an end exits through a discontinuity vertex): the label with
 *     its first k characters at most its reverse complement's, and the flags
 *     of the ends that exit.  Its discontinuity edge (Discontinuity_Graph::
 *     add_edge, include/Discontinuity_Graph.hpp:144-152) joins the exiting ends:
 *     the left end is the canonical form of the first k-mer, on its back side
 *     if that k-mer is canonical, else its front; the right end is the last
 *     k-mer's canonical form, on its front side if canonical, else its back;
 *     an end that does not exit is phi.
 * Vertices in no edge at the cutoff are counted (isolated; phantom when one
 * side is discontinuous) and produce no unitig.  Unitig order is unspecified. */
typedef struct cfgpu_unitig {
    uint64_t label_off;   /* byte offset of the label */
    uint32_t label_len;   /* vertices + k - 1 */
    uint16_t g;           /* subgraph */
    uint16_t flags;       /* CFGPU_UNITIG_* */
} cfgpu_unitig;
enum { CFGPU_UNITIG_CYCLE = 1, CFGPU_UNITIG_LEFT_EXIT = 2, CFGPU_UNITIG_RIGHT_EXIT = 4 };
int cfgpu_contract(cfgpu_handle* h);
/* Per-subgraph counters of the last contract (Subgraph's isolated, trivial_mtig_c,
 * icc_count, disc_edge_c, the vertices in unitigs and the label bytes;
 * include/Subgraph.hpp:178-186).  flags: 1 LDS table, 8 dense table (vertex arrays in
 * global memory, index in LDS), 16 wide dense table (32-bit states, index in global
 * memory), 2 global table, 0 not contracted; 4 is added when the subgraph's table was
 * inconsistent. */
typedef struct cfgpu_contract_stat {
    uint32_t isolated, phantom, trivial, icc, disc_edges, vertices, label_bytes, flags;
} cfgpu_contract_stat;
int cfgpu_contract_stats(cfgpu_handle* h, cfgpu_contract_stat* out, uint32_t n);
/* Device pointers to the unitig records and labels (valid until the next
 * cfgpu_partition / cfgpu_construct / cfgpu_contract or cfgpu_close). */
int cfgpu_unitigs_view(cfgpu_handle* h, const cfgpu_unitig** recs, const char** labels,
                       uint64_t* n_unitigs, uint64_t* n_bytes);
/* Copies the records and labels to host arrays (label_off indexes `labels`). */
int cfgpu_unitigs_copy(cfgpu_handle* h, cfgpu_unitig* recs, uint64_t rec_cap, char* labels,
                       uint64_t lab_cap, uint64_t* n_unitigs, uint64_t* n_bytes);

/* The compacted de Bruijn graph on the device: the unitigs of the last cfgpu_contract
 * glued into every maximal unitig of the whole graph, as cfhost_cdbg_build does on the
 * host (include/cfhost.h): lm-tigs joined at the discontinuity vertices they share
 * across subgraphs (src/Discontinuity_Graph_Contractor.cpp,
 * src/Contracted_Graph_Expander.cpp), each unitig in the form Unitig_Collator writes it
 * (include/Unitig_Collator.hpp:306-400).  Cycles through several subgraphs, rare, are
 * finished on the host. */
int cfgpu_cdbg(cfgpu_handle* h);
/* Unitig i = seq[rec[2i], rec[2i] + rec[2i+1]), i < *n (order unspecified).  Either
 * output may be NULL to query *n and *n_bytes. */
int cfgpu_cdbg_copy(cfgpu_handle* h, char* seq, uint64_t seq_cap, uint64_t* rec, uint64_t rec_cap,
                    uint64_t* n, uint64_t* n_bytes);

/* Counters and timings of the last partition / construct / contract. */
typedef struct cfgpu_run_stats {
    uint64_t n_bases, n_fragments, n_records, n_kmers, n_edges, n_vertices;
    uint64_t table_slots, table_bytes;
    uint32_t n_buckets_nonempty, n_overflowed, n_fallback_rounds, n_samples;
    double t_minimizer_ms, t_count_ms, t_emit_ms, t_features_ms, t_partition_ms;
    double t_sample_ms, t_train_ms, t_infer_ms, t_construct_main_ms, t_fallback_ms;
    double t_construct_ms;
    double t_construct_kernel_ms;     /* persistent construct kernels alone, main round
                                         (recorded when use_graph == 0) */
    uint32_t n_lds_buckets;           /* main-round buckets built in LDS */
    uint32_t predictor;               /* sizing used: 0 none, 1 MLP, 2 constant ratio */
    double mlp_loss;                  /* final training MSE (log space) */
    double mispredict_mean_abs_log;   /* mean |log(pred/true)| over predicted buckets */
    double mispredict_under_frac;     /* fraction of predicted buckets with pred < true */
    double fallback_rate;             /* overflowed / predicted buckets */
    double holdout_err_mlp;           /* mean |log error| on held-out sample buckets: MLP */
    double holdout_err_const;         /*   ... and the constant mean-ratio predictor */
    uint64_t n_records_walked;        /* records the construct kernels walked (distinct after dedup) */
    uint32_t n_dedup_fallback;        /* buckets whose dedup tables overflowed (walked as they are) */
    uint32_t n_lds_all_pass;          /* LDS buckets of >= 2 passes built in one workgroup, records walked once */
    double t_dedup_ms;                /* record dedup, incl. its probe round */
    /* last cfgpu_contract */
    double t_contract_ms;             /* both passes, incl. an output re-run when it was short */
    uint64_t n_unitigs, n_unitig_bytes;
    uint32_t n_contract_global;       /* subgraphs too large for the LDS and dense tables */
    uint32_t n_contract_runs;         /* 1, or 2 after growing the output */
    /* last cfgpu_cdbg */
    double t_cdbg_ms;
    uint64_t n_cdbg_unitigs, n_cdbg_bytes;
    uint32_t n_cdbg_rounds;           /* pointer-jumping rounds */
    uint32_t n_cdbg_cycle_pieces;     /* lm-tigs on cycles through several subgraphs (host) */
    uint32_t n_contract_dense;        /* subgraphs contracted with dense tables (too large for LDS) */
    uint32_t n_partition_chunks;      /* input chunks the last partition gathered (1 without APPEND) */
    double t_partition_merge_ms;      /* cfgpu_partition_finish's gather of the chunks' buckets */
    uint32_t n_contract_wide;         /* subgraphs contracted with wide dense tables (beyond the dense ones) */
} cfgpu_run_stats;
int cfgpu_run_stats_get(cfgpu_handle* h, cfgpu_run_stats* out);

/* The FP32 MLP bucket-size predictor of the last learned construct that trained one
 * (tuning.predictor 1, or 0 when the constant ratio missed the held-out buckets): the
 * slot of the reference's HyperLogLog estimate (include/HyperLogLog.hpp:29-58;
 * Subgraphs_Manager::estimate_size, disabled at src/Subgraphs_Manager.cpp:98-115).
 * Network n_features -> hidden -> hidden -> 1 with ReLU, predicting
 * log(distinct vertices / k-mer instances) of a bucket from its features, trained
 * full-batch with Adam (lr, epochs, weight init from seed) on the listed buckets.
 * params: W1[hidden][n_features] b1 W2[hidden][hidden] b2 w3 b3, then the training
 * set's feature mean[n_features] and 1/std[n_features]; features: 16384 x n_features
 * (16-bin minimizer-hash histogram, records, k-mer instances, mean length);
 * predictions: 16384 (every bucket).  Any output pointer may be NULL. */
typedef struct cfgpu_sizing_model {
    uint32_t n_features, hidden, n_params, n_train, epochs, seed;
    float lr, loss;
} cfgpu_sizing_model;
int cfgpu_sizing_model_get(cfgpu_handle* h, cfgpu_sizing_model* info, float* params, float* features,
                           uint32_t* train_buckets, float* train_targets, float* predictions);

/* Launch-shape knobs (defaults chosen by cfgpu_autotune or built-in). */
typedef struct cfgpu_tuning {
    float load_factor;        /* target table load for sized tables (0.3 .. 0.8) */
    uint32_t workgroup;       /* construct workgroup size: 256 or 512 */
    uint32_t chunk_records;   /* records per work item (LDS tiles of `workgroup`) */
    uint32_t wgs_per_cu;      /* persistent grid = CUs * wgs_per_cu */
    float margin;             /* learned sizing: capacity = pred * margin / load_factor;
                                 below 1 under-sizes on purpose, tables and dedup passes
                                 (exercises the exact overflow fallbacks; results do not
                                 change) */
    uint32_t use_graph;       /* 1: batched launches are captured and replayed as a hipGraph */
    uint32_t max_passes;      /* batched paths: buckets needing at most this many LDS hash
                                 passes are built in LDS, larger ones with global atomics
                                 (0: always global) */
    uint32_t predictor;       /* learned sizing: 0 chooses on held-out sample buckets,
                                 1 always the MLP, 2 always the constant mean ratio */
    uint32_t dedup;           /* batched paths: walk each bucket's distinct records once with
                                 their multiplicity.  0 off, 1 when sampled buckets repeat
                                 records at least 1.1x, 2 always (results never change) */
    uint32_t mlp_epochs;      /* learned sizing: MLP training epochs (1 .. 10000) */
    float mlp_lr;             /* learned sizing: Adam learning rate (0 < lr <= 1) */
    float lds_load;           /* batched paths: target fill of an LDS pass table (0.5 .. 0.85);
                                 fewer, fuller passes or more, emptier ones */
    uint32_t contract_wide;   /* contract: 1 subgraphs beyond the dense tables take wide dense
                                 tables (32-bit states, index in global memory), 0 they take
                                 the global-table walk (results never change) */
} cfgpu_tuning;
int cfgpu_tuning_get(cfgpu_handle* h, cfgpu_tuning* out);
int cfgpu_tuning_set(cfgpu_handle* h, const cfgpu_tuning* in);
/* Times the construct in `mode` on the current partition for a grid of tunings and
 * keeps the fastest (whole construct: dedup, samples, main round, fallbacks; each
 * trial timed on its second run, after its pools are allocated).  Stage
 * one times the current tuning and, when its dedup is automatic, the same with dedup
 * always on; stage two, around the faster, is the LDS pass-table load x the learned
 * sizing margin, plus the global-atomic kernel's load factor x workgroup x chunk x
 * workgroups per CU when buckets other than the learned mode's samples take that
 * kernel.  Results never change with the tuning: each trial's per-bucket vertex
 * counts are compared with the first trial's (`identical`).  `trials_out` (may be
 * NULL) receives up to n_trials trials. */
typedef struct cfgpu_tune_trial {
    cfgpu_tuning tuning;
    double ms;
    uint32_t identical;   /* 1: every bucket's vertex count equals the first trial's */
    uint32_t reserved;
} cfgpu_tune_trial;
int cfgpu_autotune(cfgpu_handle* h, uint32_t mode, cfgpu_tune_trial* trials_out,
                   uint32_t n_trials, uint32_t* n_written);

#ifdef __cplusplus
}
#endif

#endif /* CFGPU_H */
