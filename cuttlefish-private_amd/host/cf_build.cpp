// cf_build: `cuttlefish build` front end for the GPU partition + construct.
//
// Flags keep the reference's meaning (src/commands.cpp:38-106):
//   -s/--seq FILE...   input files      -l/--list FILE   file of input paths
//   -k/--kmer-len K    odd k (default 27, include/Input_Defaults.hpp:20)
//   --min-len L        minimizer length (default 20)
//   --read | --ref     FASTQ reads / FASTA references
//   -c/--cutoff C      (k+1)-mer frequency cutoff (default reads 2, refs 1)
//   -t/--threads T     host threads for the ingest
//   -o/--output PREFIX output prefix: writes PREFIX.fa, the maximal unitigs
//                      (src/Build_Params.cpp:72-75, include/File_Extensions.hpp:16)
// Extra: --mode plain|batched|learned (default learned), --device N, --states (with -o:
// also PREFIX.states, every (g, canonical k-mer, Edge_Frequency, status) row).
//
// Output: the reference's per-subgraph statistics on stderr
// (src/Subgraphs_Manager.cpp:264-305).  Exit status EXIT_FAILURE on any error, as the
// reference.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "cuttlefish_amd.hpp"

namespace {

void usage() {
    std::cerr << "usage: cf_build (-s FILE... | -l LIST) -k K [--min-len L] (--read|--ref) [-c C] "
                 "[-t T] [-o PREFIX [--states]] [--mode plain|batched|learned] [--device N]\n";
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> seqs;
    unsigned k = 27, l = 20, threads = 0;
    int cutoff = -1, device = 0;
    bool read = false, ref = false, states = false;
    std::string out, mode = "learned";
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto need = [&](const char* what) -> std::string {
            if (i + 1 >= argc) {
                std::cerr << "missing value for " << what << "\n";
                std::exit(EXIT_FAILURE);
            }
            return argv[++i];
        };
        if (a == "-s" || a == "--seq") {
            seqs.push_back(need("--seq"));
            while (i + 1 < argc && argv[i + 1][0] != '-') seqs.push_back(argv[++i]);
        } else if (a == "-l" || a == "--list") {
            std::ifstream f(need("--list"));
            if (!f) { std::cerr << "cannot open list\n"; return EXIT_FAILURE; }
            for (std::string p; std::getline(f, p);)
                if (!p.empty()) seqs.push_back(p);
        } else if (a == "-k" || a == "--kmer-len") {
            k = unsigned(std::stoul(need("-k")));
        } else if (a == "--min-len") {
            l = unsigned(std::stoul(need("--min-len")));
        } else if (a == "-c" || a == "--cutoff") {
            cutoff = std::stoi(need("-c"));
        } else if (a == "-t" || a == "--threads") {
            threads = unsigned(std::stoul(need("-t")));
        } else if (a == "-o" || a == "--output") {
            out = need("-o");
        } else if (a == "--read") {
            read = true;
        } else if (a == "--ref") {
            ref = true;
        } else if (a == "--mode") {
            mode = need("--mode");
        } else if (a == "--states") {
            states = true;
        } else if (a == "--device") {
            device = std::stoi(need("--device"));
        } else if (a == "-h" || a == "--help") {
            usage();
            return EXIT_SUCCESS;
        } else {
            std::cerr << "unknown option " << a << "\n";
            usage();
            return EXIT_FAILURE;
        }
    }
    // Build_Params::is_valid (src/Build_Params.cpp:97-218): one of --read/--ref,
    // odd k, cutoff != 0.
    if (seqs.empty() || read == ref || !(k & 1) || cutoff == 0) {
        usage();
        return EXIT_FAILURE;
    }
    if (cutoff < 0) cutoff = read ? 2 : 1;
    const uint32_t m = mode == "plain" ? CFGPU_MODE_PLAIN
                     : mode == "batched" ? CFGPU_MODE_BATCHED
                     : CFGPU_MODE_BATCHED_LEARNED;
    try {
        cuttlefish_amd::Subgraphs_Manager G(uint16_t(k), uint16_t(l), uint8_t(cutoff), device);
        G.partition(seqs, read, threads);
        G.finalize();
        G.construct(m);
        if (!out.empty()) {
            const uint64_t n = G.contract_and_collate(out + ".fa");
            std::cerr << "Maximal unitig count: " << n << ".\n";
        }
        std::FILE* dump = nullptr;
        if (!out.empty() && states) {
            dump = std::fopen((out + ".states").c_str(), "wb");
            if (!dump) { std::cerr << "cannot write " << out << ".states\n"; return EXIT_FAILURE; }
        }
        uint64_t max_kc = 0, min_kc = UINT64_MAX, sum_sz = 0, max_sz = 0, min_sz = UINT64_MAX;
        const unsigned kw = (k + 31) / 32;
        G.process([&](cuttlefish_amd::Subgraph& s) {
            max_kc = std::max(max_kc, s.kmer_count());
            min_kc = std::min(min_kc, s.kmer_count());
            sum_sz += s.size();
            max_sz = std::max<uint64_t>(max_sz, s.size());
            min_sz = std::min<uint64_t>(min_sz, s.size());
            if (dump)
                for (std::size_t i = 0; i < s.size(); ++i) {
                    const uint32_t g = s.id();
                    const uint32_t e = s.get_val(i).frequency_word();
                    const uint8_t st = s.get_val(i).status_byte();
                    std::fwrite(&g, 4, 1, dump);
                    std::fwrite(s.get_key(i), 8, kw, dump);
                    std::fwrite(&e, 4, 1, dump);
                    std::fwrite(&st, 1, 1, dump);
                }
        }, m);
        if (dump) std::fclose(dump);
        const cfgpu_run_stats rs = G.run_stats();
        std::cerr << "Number of weak super k-mers: " << rs.n_records << ".\n"
                  << "Maximum k-mer count in bucket: " << max_kc << ".\n"
                  << "Minimum k-mer count in bucket: " << (min_kc == UINT64_MAX ? 0 : min_kc) << ".\n"
                  << "Sum graph size: " << sum_sz << ".\n"
                  << "Largest graph size: " << max_sz << ".\n"
                  << "Smallest graph size: " << (min_sz == UINT64_MAX ? 0 : min_sz) << ".\n"
                  << "Partition time (GPU): " << rs.t_partition_ms << " ms, construct time (GPU): "
                  << rs.t_construct_ms << " ms.\n";
    } catch (const std::exception& e) {
        std::cerr << "cf_build: " << e.what() << "\n";
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
