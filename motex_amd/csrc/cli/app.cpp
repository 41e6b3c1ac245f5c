// The motex command line's driver.  Follows the control flow of the reference's main
// (motex.cc:36-1168): decode and validate options, read the FASTA input and the boxes file,
// run the extraction, write the .motex output (SMILE output with -S, the background
// statistics with -b/-u/-U, the re-search statistics with -I).  The extraction -- the
// all-pairs loop of motex.cc:494-583, or the motif-by-sequence loop of motex.cc:940-990 -- is
// one call of the Extractor (main.cpp: into libmotex on the GPU).
#include <sys/time.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "motex.h"
#include "cli.h"

namespace motex_cli {

static double gettime()
{
    struct timeval tv;
    gettimeofday(&tv, 0);
    return tv.tv_sec + tv.tv_usec * 0.000001;
}

// -I: re-search the motifs of an unmatched-motif file in the input sequences
// (motex.cc:883-1137): each motif string is an x sequence searched in every input sequence,
// and the foreground statistics are written (functions.cc:2750-2832).
static int research_unmatched(const Switches &sw, const std::vector<std::string> &seqs, const motex_params &p,
                              unsigned int total_mot_length, int P, Extractor &ex)
{
    std::vector<MotifRecord> fg;
    if (read_motif_records(sw.un_in_filename, fg)) return 1;
    const unsigned int num = (unsigned int)seqs.size(), nf = (unsigned int)fg.size();
    double start = gettime();
    std::vector<std::vector<uint32_t>> u(nf), v(nf);
    std::vector<const char *> xs(nf), ys(num);
    std::vector<uint32_t> xl(nf), yl(num);
    std::vector<uint32_t *> up(nf), vp(nf);
    for (unsigned int i = 0; i < nf; ++i) {
        u[i].assign(fg[i].motif.size() + 1, 0);
        v[i].assign(fg[i].motif.size() + 1, 0);
        xs[i] = fg[i].motif.data();
        xl[i] = (uint32_t)fg[i].motif.size();
        up[i] = u[i].data();
        vp[i] = v[i].data();
    }
    uint64_t short_seqs = 0;
    for (unsigned int j = 0; j < num; ++j) {
        ys[j] = seqs[j].data();
        yl[j] = (uint32_t)seqs[j].size();
        short_seqs += yl[j] < total_mot_length;
    }
    // motex.cc:948-952: one warning per (motif, short sequence)
    for (uint64_t t = 0; t < short_seqs * nf; ++t)
        fprintf(stderr, " Omitting short sequence in file %s!\n", sw.input_filename.c_str());
    if (ex.cross(xs.data(), xl.data(), nf, ys.data(), yl.data(), num, p, up.data(), vp.data()) != MOTEX_OK) {
        fprintf(stderr, " Error: motif extraction failed: %s\n", ex.last_error());
        return 1;
    }
    double end = gettime();
    return write_motifs_fore(sw, fg, u, v, end - start, P, num) ? 0 : 1;
}

int run_cli(int argc, char **argv, Extractor &ex)
{
    Switches sw;
    int i = decode_switches(argc, argv, &sw);
    if (i < 15) {
        usage();
        return 1;
    }
    // validation, motex.cc:105-141
    if (sw.q > 100) {
        fprintf(stderr, " Error: quorum argument q (%%) should be less or equal to 100!\n");
        return 1;
    }
    if (sw.Q > 100) {
        fprintf(stderr, " Error: max-quorum argument Q (%%) should be less or equal to 100!\n");
        return 1;
    }
    if (sw.e >= sw.l) {
        fprintf(stderr, " Error: error threshold e must be less than the fixed-length of motifs k!\n");
        return 1;
    }
    if (sw.d != 0 && sw.d != 1) {
        fprintf(stderr, " Error: distance argument d should be `0' for Hamming distance or `1' for edit distance!\n");
        return 1;
    }
    const char *alphabet;
    uint32_t alpha_id;
    if (sw.alphabet == "DNA") { alphabet = "ACGTN"; alpha_id = MOTEX_ALPHA_DNA; }
    else if (sw.alphabet == "PROT") { alphabet = "ARNDCQEGHILKMFPSTWYVOUBZJX*"; alpha_id = MOTEX_ALPHA_PROT; }
    else if (sw.alphabet == "USR") { alphabet = "0123456789"; alpha_id = MOTEX_ALPHA_USR; }
    else {
        fprintf(stderr, " Error: alphabet argument a should be `DNA' for nucleotide sequences or `PROT' for protein sequences or `USR' for sequences over a user-defined alphabet!\n");
        return 1;
    }
    // option combinations, motex.cc:156-201
    if (!sw.has_background) {
        if (sw.has_un_out) { fprintf(stderr, " Error: `-u' option must be used with `-b' option!\n"); return 1; }
        if (sw.has_un_smile) { fprintf(stderr, " Error: `-U' option must be used with `-b' option!\n"); return 1; }
    } else {
        if (!sw.has_un_out) { fprintf(stderr, " Error: `-b' option must be used with `-u' option!\n"); return 1; }
        if (sw.has_un_smile && !sw.has_smile) { fprintf(stderr, " Error: `-U' option must be used with `-S' option!\n"); return 1; }
        if (sw.has_un_in) { fprintf(stderr, " Error: `-I' option cannot be used with `-b' option!\n"); return 1; }
    }
    if (sw.has_boxes && sw.has_un_in) {
        fprintf(stderr, " Error: `-I' option cannot be used with `-s' option!\n");
        return 1;
    }
    if ((sw.has_background || sw.has_un_in) && !BinomialTail::available()) {
        fprintf(stderr, " Error: `-%c' needs the MPFR library, which this build of motex was compiled without!\n",
                sw.has_background ? 'b' : 'I');
        return 1;
    }

    std::vector<std::string> seqs;
    if (read_fasta(sw, alphabet, seqs)) return 1;
    unsigned int total_mot_length = sw.l;
    if (sw.has_boxes) {
        if (read_boxes(sw, &total_mot_length)) return 1;
    } else {
        sw.nb_gaps = 0;
        sw.nb_boxes = 1;
        sw.nb_structs = 1;
    }
    unsigned int num = (unsigned int)seqs.size();
    uint64_t total = 0;
    for (const std::string &s : seqs) total += s.size();
    sw.total_length = (unsigned int)total;

    motex_params p{};
    p.k = sw.l;
    p.e = sw.e;
    p.distance = sw.d;
    p.alphabet = alpha_id;
    p.nb_gaps = sw.nb_gaps;
    p.gap_min = sw.bgaps_min.data();
    p.gap_max = sw.bgaps_max.data();
    p.box_len = sw.blens.data();
    p.box_err = sw.berrs.data();
    p.long_mode = sw.L;
    const int P = 1;  // one GPU
    if (sw.has_un_in) return research_unmatched(sw, seqs, p, total_mot_length, P, ex);

    double start = gettime();
    std::vector<std::vector<uint32_t>> u(num), v(num);
    std::vector<const char *> ptrs(num);
    std::vector<uint32_t> lens(num);
    std::vector<uint32_t *> up(num), vp(num);
    for (unsigned int s = 0; s < num; ++s) {
        u[s].assign((size_t)seqs[s].size() * sw.nb_structs + 1, 0);
        v[s].assign((size_t)seqs[s].size() * sw.nb_structs + 1, 0);
        ptrs[s] = seqs[s].data();
        lens[s] = (uint32_t)seqs[s].size();
        up[s] = u[s].data();
        vp[s] = v[s].data();
    }
    // motex.cc:536-540 prints a warning for every skipped (x, y) pair
    uint64_t short_seqs = 0;
    for (unsigned int s = 0; s < num; ++s) short_seqs += lens[s] < total_mot_length;
    uint64_t skipped = short_seqs * num * 2 - short_seqs * short_seqs;
    for (uint64_t t = 0; t < skipped; ++t)
        fprintf(stderr, " Omitting short sequence in file %s!\n", sw.input_filename.c_str());

    if (ex.all_pairs(ptrs.data(), lens.data(), num, p, up.data(), vp.data()) != MOTEX_OK) {
        fprintf(stderr, " Error: motif extraction failed: %s\n", ex.last_error());
        return 1;
    }
    double end = gettime();

    // writer dispatch, motex.cc:837-867: with a background file the statistics writers run
    // instead, and the SMILE form only with -U
    if (sw.has_background) {
        if (!write_motifs_back(sw, seqs, u, v, end - start, P)) return 1;
        if (sw.has_un_smile && !write_motifs_back_smile(sw, seqs, u, v, end - start)) return 1;
        return 0;
    }
    if (!write_motifs(sw, seqs, u, v, end - start, P)) return 1;
    if (sw.has_smile && !write_motifs_smile(sw, seqs, u, v, end - start)) return 1;
    return 0;
}

}  // namespace motex_cli
