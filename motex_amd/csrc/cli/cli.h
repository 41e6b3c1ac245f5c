// Host side of the motex command-line tool: options, input readers and writers.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace motex_cli {

// The options of the reference's struct TSwitch (motexdefs.h:32-61) that this tool uses.
struct Switches {
    std::string alphabet, input_filename, output_filename;
    std::string background_in_filename, un_in_filename, un_out_filename;
    std::string boxes_in_filename, smile_out_filename, un_smile_out_filename;
    bool has_background = false, has_un_in = false, has_un_out = false;
    bool has_boxes = false, has_smile = false, has_un_smile = false;
    unsigned int q = 0, Q = 100, l = 0, d = 0, e = 0, n = 1, N = 10000, t = 4, L = 0;
    unsigned int total_length = 0;
    unsigned int nb_gaps = 0, nb_boxes = 1, nb_structs = 1;
    std::vector<unsigned int> bgaps_min, bgaps_max, blens, berrs;
    std::vector<unsigned int> S;  // nb_structs x nb_gaps, row-major
};

// Returns optind (the reference's main requires >= 15), 0 for -h or a malformed
// number; prints the usage and exits(1) when fewer than 7 mandatory options are seen.
int decode_switches(int argc, char **argv, Switches *sw);
void usage();

// FASTA input (motex.cc:225-300, 409-435): returns 0 or 1 (after printing the error).
int read_fasta(const Switches &sw, const char *alphabet, std::vector<std::string> &seqs);
// Boxes file (motex.cc:303-363).
int read_boxes(Switches &sw, unsigned int *total_mot_length);

// Writers (functions.cc:1010-1142, 1147-1260, 1265-1461, 1466-1608).
int write_motifs(const Switches &sw, const std::vector<std::string> &seqs,
                 const std::vector<std::vector<uint32_t>> &u,
                 const std::vector<std::vector<uint32_t>> &v, double exectime, int P);
int write_motifs_smile(const Switches &sw, const std::vector<std::string> &seqs,
                       const std::vector<std::vector<uint32_t>> &u,
                       const std::vector<std::vector<uint32_t>> &v, double exectime);

}  // namespace motex_cli
