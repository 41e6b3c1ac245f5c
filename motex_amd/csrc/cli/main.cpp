// motex: the command-line tool.  Follows the control flow of the reference's main
// (motex.cc:36-1168): decode and validate options, read the FASTA input and the boxes
// file, run the extraction, write the .motex output (and SMILE output with -S).  The
// extraction -- the all-pairs loop of motex.cc:494-583 -- is one call into libmotex.
#include <sys/time.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "motex.h"
#include "cli.h"

using namespace motex_cli;

static double gettime()
{
    struct timeval tv;
    gettimeofday(&tv, 0);
    return tv.tv_sec + tv.tv_usec * 0.000001;
}

int main(int argc, char **argv)
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
    if (sw.has_background || sw.has_un_in) {
        fprintf(stderr, " Error: `-%c' (%s) is not supported by this build of motex!\n",
                sw.has_background ? 'b' : 'I',
                sw.has_background ? "background statistics" : "re-search of unmatched motifs");
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
    motex_ctx *ctx = nullptr;
    int rc = motex_ctx_create(0, 0, &ctx);
    if (rc == MOTEX_OK)
        rc = motex_extract_all_pairs(ctx, ptrs.data(), lens.data(), num, &p, up.data(), vp.data());
    if (rc != MOTEX_OK) {
        fprintf(stderr, " Error: motif extraction failed: %s\n", motex_last_error());
        motex_ctx_destroy(ctx);
        return 1;
    }
    motex_ctx_destroy(ctx);
    double end = gettime();

    const int P = 1;  // one GPU
    if (!write_motifs(sw, seqs, u, v, end - start, P)) return 1;
    if (sw.has_smile && !write_motifs_smile(sw, seqs, u, v, end - start)) return 1;
    return 0;
}
