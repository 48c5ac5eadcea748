// Host ingest internals shared by the cfhost_input_* calls (host/ingest.cpp) and
// the chunked ingest pipeline (host/pipeline.cpp): FASTA/FASTQ text -> records
// -> fragments -> the 2-bit packed stream cfgpu_partition takes.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

namespace cfhost_detail {

// Records of one thread's share of the input: record r is seq[rec[r].first,
// rec[r].second).  seq points into the caller's text (FASTQ, records given
// directly) or into `own` (FASTA, whose sequence lines are concatenated).
struct Seg {
    std::vector<char> own;
    const char* seq = nullptr;
    std::vector<std::pair<uint64_t, uint64_t>> rec;
};

uint32_t pick_threads(uint32_t t);

// The records of FASTA / FASTQ text in up to T segments, in input order.  The
// result is the sequential parse's whatever T is: FASTQ records are the
// second line of every 4 (line numbers from the text's start, counted in
// parallel); FASTA records are a '>' line's following lines concatenated,
// lines before the first '>' line ignored; '\n' ends lines
// (RabbitFX include/RabbitFX/io/Formater.cpp:87-130, 238-263).
std::vector<Seg> parse_text(const char* buf, uint64_t n, int is_fastq, uint32_t T);

// One chunk of that text: the records from b (0 or a record start) up to the
// first record start at or after x (*end; n when none), with the same rules
// (FASTQ line numbers counted from b, whose line number is a multiple of 4).
std::vector<Seg> parse_chunk(const char* buf, uint64_t n, int is_fastq, uint64_t b, uint64_t x, uint32_t T,
                             uint64_t* end);

// Records given directly, split into up to T segments balanced by bytes.
std::vector<Seg> segs_of_records(const char* seq, const uint64_t* rec_off, uint64_t n_rec, uint32_t T);

// Fragments (maximal A/C/G/T runs of length >= k+1 inside a record,
// src/Graph_Partitioner.cpp:476-501) of the segments, then their 2-bit
// MSB-first packing, one thread per segment.
class Packer {
public:
    void plan(const std::vector<Seg>& segs, uint32_t k);
    uint64_t n_bases() const { return tb_.back(); }
    uint64_t n_frag() const { return tf_.back(); }
    uint64_t n_words() const { return (n_bases() + 31) / 32 + 2; }
    // packed[n_words()] (the last two words zero) and frag_off[n_frag() + 1].
    void write(uint64_t* packed, uint64_t* frag_off) const;

private:
    struct Frag {
        uint64_t beg, len;
    };
    const std::vector<Seg>* segs_ = nullptr;
    std::vector<std::vector<Frag>> fr_;
    std::vector<uint64_t> tb_{0}, tf_{0};
};

// Chunk boundaries at record starts, about chunk_bytes apart: off[0] = 0 <
// off[1] < ... < off[m] = n.  Each chunk parses, on its own, to the records
// the whole text gives there (FASTQ: a boundary starts a line whose number is
// a multiple of 4; FASTA: a '>' line).
std::vector<uint64_t> split_records(const char* buf, uint64_t n, int is_fastq, uint64_t chunk_bytes, uint32_t T);

}  // namespace cfhost_detail
