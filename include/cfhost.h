/*
 * cfhost.h — host-side ingest, synthetic inputs and the compacted graph for libcfgpu.
 *
 * Ingest restates the reference's record -> fragment rules and 2-bit encoding on the
 * host, in parallel, and produces exactly the stream cfgpu_partition consumes:
 *   records  FASTA: '>' header lines, sequence lines concatenated, '\n' line ends only
 *            (RabbitFX include/RabbitFX/io/Formater.cpp:63-130); FASTQ: 4-line records,
 *            sequence = line 2 (Formater.cpp:238-263);
 *   fragments maximal runs of A/C/G/T (either case) of length >= k+1 inside a record
 *            (src/Graph_Partitioner.cpp:476-501, DNA_Utility::is_DNA_base);
 *   bases    A=0 C=1 G=2 T=3 (include/DNA.hpp:17-25), packed MSB-first 32 per word.
 */
#ifndef CFHOST_H
#define CFHOST_H

#include <stddef.h>
#include <stdint.h>

#include "cfgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cfhost_input cfhost_input;

/* Parse a FASTA (is_fastq = 0) or FASTQ (is_fastq = 1) file / buffer. */
int cfhost_input_from_file(const char* path, int is_fastq, uint32_t k, uint32_t threads,
                           cfhost_input** out);
int cfhost_input_from_buffer(const char* buf, uint64_t n, int is_fastq, uint32_t k,
                             uint32_t threads, cfhost_input** out);
/* Records given directly: record r is seq[rec_off[r], rec_off[r+1]). */
int cfhost_input_from_records(const char* seq, const uint64_t* rec_off, uint64_t n_rec, uint32_t k,
                              uint32_t threads, cfhost_input** out);
/* packed holds n_words = ceil(n_bases/32) + 2 words (the last two zero). */
int cfhost_input_get(const cfhost_input* in, const uint64_t** packed, uint64_t* n_words,
                     uint64_t* n_bases, const uint64_t** frag_off, uint64_t* n_frag,
                     uint64_t* n_records);
void cfhost_input_free(cfhost_input* in);

/* Chunk boundaries at record starts about chunk_bytes apart: offsets[0] = 0 < ... <
 * offsets[*n_chunks] = n (offsets may be NULL to query *n_chunks; it then holds
 * *n_chunks + 1 entries, cap of them at most).  Each chunk parses on its own to the
 * records the whole text gives there: a FASTQ boundary starts a line whose number is a
 * multiple of 4, a FASTA boundary a '>' line (a record longer than chunk_bytes is one
 * chunk). */
int cfhost_split_records(const char* buf, uint64_t n, int is_fastq, uint64_t chunk_bytes, uint32_t threads,
                         uint64_t* offsets, uint64_t cap, uint64_t* n_chunks);

/* The reference's partition of an input file, end to end, with host and device
 * overlapped (src/Graph_Partitioner.cpp:121-184, 424-614; Subgraphs_Manager::finalize,
 * include/Subgraphs_Manager.hpp:113): the text is cut into chunks of whole records (the
 * boundaries cfhost_split_records gives, each found while its chunk is parsed;
 * chunk_bytes 0 means 1 GiB); while the device copies chunk i
 * from pinned host memory and partitions it (cfgpu_partition with
 * CFGPU_PARTITION_APPEND), `threads` host threads parse and pack chunk i+1; then
 * cfgpu_partition_finish.  The handle then holds the buckets one cfgpu_partition of the
 * whole input gives.  The file is memory-mapped and read as the parser reaches it.
 * Returns 0, a CFGPU_E_* code from the device calls, -1 on bad arguments, -2 when the
 * file cannot be read, -3 when pinned host memory runs out. */
typedef struct cfhost_ingest_stats {
    uint64_t n_bytes, n_records, n_bases, n_fragments;
    uint32_t n_chunks, threads;
    double t_total_ms;   /* text -> buckets finished */
    double t_host_ms;    /* parse + pack, summed over chunks (all but the first overlap the device) */
    double t_device_ms;  /* the cfgpu_partition calls (copy from pinned memory + kernels), summed */
    double t_finish_ms;  /* cfgpu_partition_finish */
} cfhost_ingest_stats;
int cfhost_partition_buffer(cfgpu_handle* h, const char* buf, uint64_t n, int is_fastq, uint32_t threads,
                            uint64_t chunk_bytes, cfhost_ingest_stats* stats);
int cfhost_partition_file(cfgpu_handle* h, const char* path, int is_fastq, uint32_t threads, uint64_t chunk_bytes,
                          cfhost_ingest_stats* stats);
/* The same with partition flags for the chunks: CFGPU_PARTITION_SPILL sends every
 * chunk's buckets to the handle's bucket store (cfgpu_spill_dir), so that inputs whose
 * buckets exceed device memory are partitioned as the reference partitions into its
 * bucket files; 0 is cfhost_partition_buffer / _file. */
int cfhost_partition_buffer_ex(cfgpu_handle* h, const char* buf, uint64_t n, int is_fastq, uint32_t threads,
                               uint64_t chunk_bytes, uint32_t flags, cfhost_ingest_stats* stats);
int cfhost_partition_file_ex(cfgpu_handle* h, const char* path, int is_fastq, uint32_t threads, uint64_t chunk_bytes,
                             uint32_t flags, cfhost_ingest_stats* stats);

/* Seeded synthetic sequences (ACGT only).  All are deterministic in `seed`
 * regardless of `threads`. */
/* iid genome of length len; then repeat families (lengths log-uniform in
 * [fam_min, fam_max]) copied with per-copy iid substitutions at `divergence`
 * over positions until repeat_frac * len bases are covered. */
int cfhost_synth_genome(uint64_t len, double repeat_frac, uint32_t fam_min, uint32_t fam_max,
                        double divergence, uint64_t seed, char* out);
/* base + n_copies inserted copies of one random rep_len repeat, each copy with
 * iid substitutions at sub_rate; out holds base_len + n_copies * rep_len. */
int cfhost_synth_insert_repeats(const char* base, uint64_t base_len, uint32_t n_copies,
                                uint32_t rep_len, double sub_rate, uint64_t seed, char* out);
/* n_reads reads of read_len from uniform starts on either strand (50/50) with iid
 * substitutions at sub_rate, written back to back into out[n_reads*read_len]. */
int cfhost_synth_reads(const char* genome, uint64_t glen, uint64_t n_reads, uint32_t read_len,
                       double sub_rate, uint64_t seed, uint32_t threads, char* out);
/* A derivative of base with SNPs at snp_rate and indels (1..indel_max bp, half
 * insertions, half deletions) at indel_rate per base. */
int cfhost_synth_mutate(const char* base, uint64_t len, double snp_rate, double indel_rate,
                        uint32_t indel_max, uint64_t seed, char* out, uint64_t cap,
                        uint64_t* out_len);

/* Bucketed records for scheduling stress: n_records weak super k-mers over buckets
 * 0..n_buckets-1 with counts proportional to (rank+1)^-zipf_s (rank = bucket id), each a
 * substring (length k..2k-l) of a per-bucket random sequence twice as long as the
 * bucket's record count, disc flags at random.  Layout as cfgpu_load_buckets takes:
 * att[n_records], labels[n_records*W] (W = ceil((2k-l)/32)), bucket_off[16385]. */
int cfhost_synth_bucket_workload(uint32_t n_buckets, uint64_t n_records, double zipf_s,
                                 uint32_t k, uint32_t l, uint64_t seed, uint32_t* att,
                                 uint64_t* labels, uint64_t* bucket_off);

/* The compacted de Bruijn graph: every maximal unitig of the whole input, from the
 * per-subgraph output of cfgpu_contract (records and labels as cfgpu_unitigs_copy
 * returns them).  Trivial unitigs are kept; lm-tigs are glued at the discontinuity
 * vertices they share across subgraphs - the contraction of the discontinuity graph and
 * its expansion into labels (src/Discontinuity_Graph_Contractor.cpp,
 * src/Contracted_Graph_Expander.cpp) - and written in the form Unitig_Collator gives
 * maximal unitigs (include/Unitig_Collator.hpp:306-400): a linear unitig with its first
 * k characters at most its reverse complement's, a cycle rotated to its smallest vertex
 * in canonical form.  `sorted` != 0 sorts the unitigs.  Returns 0, -1 on bad arguments,
 * -2 on records that do not glue consistently. */
typedef struct cfhost_cdbg cfhost_cdbg;
int cfhost_cdbg_build(uint32_t k, const cfgpu_unitig* recs, uint64_t n_recs, const char* labels,
                      uint64_t n_bytes, int sorted, cfhost_cdbg** out);
/* unitig i = seq[off[i], off[i+1]), i < n */
int cfhost_cdbg_get(const cfhost_cdbg* c, const char** seq, const uint64_t** off, uint64_t* n);
/* FASTA, one line per unitig, headers are the unitig index. */
int cfhost_cdbg_write_fasta(const cfhost_cdbg* c, const char* path);
void cfhost_cdbg_free(cfhost_cdbg* c);

#ifdef __cplusplus
}
#endif

#endif /* CFHOST_H */
