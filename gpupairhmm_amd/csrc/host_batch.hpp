// Host-side batch of sequence groups: the reference's BatchOfSequenceGroups
// (common.cuh:109-201), its line-record parser (file_parser_worker.cu:31-218),
// its score writer (output_writer_worker.cu:29-159) and the seeded synthetic
// workloads used by tests and the benchmark.
#pragma once

#include <cstddef>
#include <cstdint>
#include <istream>
#include <string>
#include <vector>

#include "gpupairhmm.h"

namespace phmm {

// One record ("sequence group"): R reads x H haps, as parsed.
struct SequenceGroup {
    int num_reads = 0;
    int num_haps = 0;
    std::vector<uint8_t> read_data, hap_data, base_quals, ins_quals, del_quals, gcp_quals;
    std::vector<int32_t> read_lengths, hap_lengths;
};

struct HostBatch {
    std::vector<uint8_t> read_data, hap_data, base_quals, ins_quals, del_quals, gcp_quals;
    std::vector<int32_t> read_lengths, hap_lengths, read_offsets, hap_offsets;
    std::vector<int32_t> num_reads, num_haps, reads_prefix, haps_prefix;
    int64_t num_alignments = 0;

    void clear();
    phmm_batch_view view() const;
    // Append groups and rebuild offsets/prefix sums
    // (convertGroupsToBatchInplace, file_parser_worker.cu:155-218).
    void assign(const std::vector<SequenceGroup>& groups);
};

// Padding of one sequence: 4 - len%4 bytes, 1..4 (file_parser_worker.cu:81-86).
inline int padding_of(int len) { return 4 - (len % 4); }
inline int padded_length(int len) { return len + padding_of(len); }

// Parse the next record whose header line has been read.  Throws
// std::runtime_error with a message on malformed input.
SequenceGroup parse_group(std::istream& in, const std::string& header, int64_t* line_no);

// Checks the structural invariants of a view (host memory).  Returns an
// empty string when valid, else a description.
std::string validate_host_view(const phmm_batch_view& v);

// Score text: per group "R H\n" then R lines of H "%.5f" (writer format).
void format_scores(const phmm_batch_view& v, const float* scores, std::string& out);

// Synthetic workloads 0..4 (see synth.cpp).
void synth_generate(int config, uint64_t seed, int32_t num_records, HostBatch& out);

}  // namespace phmm
