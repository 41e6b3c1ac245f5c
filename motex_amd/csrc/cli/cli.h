// Host side of the motex command-line tool: options, input readers and writers.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "motex.h"

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

// Pieces shared by the writers (writer.cpp).
const char *alphabet_letters(const std::string &name);
void write_header(FILE *f, const Switches &sw, unsigned int num_seqs, double exectime, int P);
void write_smile_header(FILE *f, const Switches &sw, unsigned int num_seqs, const char *astr);
bool passes(const Switches &sw, unsigned int num_seqs, uint32_t u, uint32_t v);
std::string motif_text(const Switches &sw, const std::string &seq, unsigned int j, unsigned int s);
std::string smile_id(const Switches &sw, const std::string &mot, const char *astr);

// Binomial upper tail P[X >= x], X ~ Bin(N, p), as binomial.cc:25-74 computes and prints it
// (stats.cpp); available() is false in builds without MPFR.
class BinomialTail {
  public:
    explicit BinomialTail(unsigned long n_max);  // factorials 0..n_max
    ~BinomialTail();
    BinomialTail(const BinomialTail &) = delete;
    BinomialTail &operator=(const BinomialTail &) = delete;
    static bool available();
    void print_upper_tail(FILE *f, unsigned long x, unsigned long N, long double p);

  private:
    struct Impl;
    Impl *impl;
};

// One motif line of a .motex file read back (struct Tdata, motexdefs.h:63-69): the motif
// string and its u, n (= number of sequences), ratio and v.
struct MotifRecord {
    std::string motif;
    unsigned int u = 0, n = 0, v = 0;
    double r = 0;
};

// Background statistics (-b, -u, -U): functions.cc:1613-1905 (single), 1906-2241
// (structured), 2243-2466 and 2471-2745 (SMILE forms).
int write_motifs_back(const Switches &sw, const std::vector<std::string> &seqs,
                      const std::vector<std::vector<uint32_t>> &u,
                      const std::vector<std::vector<uint32_t>> &v, double exectime, int P);
int write_motifs_back_smile(const Switches &sw, const std::vector<std::string> &seqs,
                            const std::vector<std::vector<uint32_t>> &u,
                            const std::vector<std::vector<uint32_t>> &v, double exectime);
// Re-search of unmatched motifs (-I): reads the motif lines of an unmatched-motif file
// (motex.cc:883-925) and writes the foreground statistics (functions.cc:2750-2832).
int read_motif_records(const std::string &path, std::vector<MotifRecord> &out);
int write_motifs_fore(const Switches &sw, const std::vector<MotifRecord> &fg,
                      const std::vector<std::vector<uint32_t>> &u,
                      const std::vector<std::vector<uint32_t>> &v, double exectime, int P,
                      unsigned int num_seqs);

// The extraction step behind the command line: the product binary runs it on the GPU
// through libmotex (main.cpp).  Both calls accumulate into caller-zeroed rows and return
// MOTEX_OK or a negative code with last_error() set.
class Extractor {
  public:
    virtual ~Extractor() = default;
    virtual int all_pairs(const char *const *seqs, const uint32_t *lens, uint32_t num, const motex_params &p,
                          uint32_t *const *u, uint32_t *const *v) = 0;
    virtual int cross(const char *const *xs, const uint32_t *xl, uint32_t nx, const char *const *ys,
                      const uint32_t *yl, uint32_t ny, const motex_params &p, uint32_t *const *u,
                      uint32_t *const *v) = 0;
    virtual const char *last_error() = 0;
};

// The whole command line (app.cpp): returns the process exit code.
int run_cli(int argc, char **argv, Extractor &ex);

}  // namespace motex_cli
