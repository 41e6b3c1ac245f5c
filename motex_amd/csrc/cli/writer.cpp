// Output writers with the reference's formats:
//   .motex header  functions.cc:1010-1068
//   motifs         functions.cc:1073-1142 (single) and 1147-1260 (structured)
//   SMILE          functions.cc:1265-1365 (header), 1370-1461, 1466-1608
// The quorum filter is evaluated in double precision exactly as written there, and a
// motif string is printed only the first time it passes (the reference keeps a trie of
// printed motifs; a hash set gives the same membership answers).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <unordered_set>

#include "cli.h"

namespace motex_cli {

static const char *kDNA = "ACGTN";
static const char *kPROT = "ARNDCQEGHILKMFPSTWYVOUBZJX*";
static const char *kUSR = "0123456789";

const char *alphabet_letters(const std::string &name)
{
    if (name == "DNA") return kDNA;
    if (name == "PROT") return kPROT;
    return kUSR;
}

void write_header(FILE *f, const Switches &sw, unsigned int num_seqs, double exectime, int P)
{
    time_t t;
    time(&t);
    fprintf(f, "####################################\n");
    fprintf(f, "# Program: MoTeX\n");
    fprintf(f, "# Rundate: %s", ctime(&t));
    fprintf(f, "# Input file: %s\n", sw.input_filename.c_str());
    fprintf(f, "# Output file: %s\n", sw.output_filename.c_str());
    fprintf(f, "# For %d input sequences\n", (int)num_seqs);
    fprintf(f, "#     d = %d\n", (int)sw.d);
    fprintf(f, "#     q = %d\n", (int)sw.q);
    if (sw.Q != 100) fprintf(f, "#     Q = %d\n", (int)sw.Q);
    if (sw.n != 1) fprintf(f, "#     n = %d\n", (int)sw.n);
    if (sw.N != 10000) fprintf(f, "#     N = %d\n", (int)sw.N);
    if (sw.nb_boxes == 1) {
        fprintf(f, "#     l = %d\n", (int)sw.l);
        fprintf(f, "#     e = %d\n", (int)sw.e);
    } else {
        fprintf(f, "# Number of boxes: %d\n", (int)sw.nb_boxes);
        fprintf(f, "#     l0 = %d\n", (int)sw.l);
        fprintf(f, "#     e0 = %d\n", (int)sw.e);
        for (unsigned int i = 0; i < sw.nb_gaps; ++i) {
            fprintf(f, "#     g%d = [%d,%d]\n", (int)i + 1, (int)sw.bgaps_min[i], (int)sw.bgaps_max[i]);
            fprintf(f, "#     l%d = %d\n", (int)i + 1, (int)sw.blens[i]);
            fprintf(f, "#     e%d = %d\n", (int)i + 1, (int)sw.berrs[i]);
        }
        fprintf(f, "# Boxes input file: %s\n", sw.boxes_in_filename.c_str());
    }
    fprintf(f, "# Run on %d proc(s) in %lf secs\n", P, exectime);
    fprintf(f, "####################################\n\n");
}

bool passes(const Switches &sw, unsigned int num_seqs, uint32_t u, uint32_t v)
{
    double r = ((double)u / num_seqs) * 100;
    return r >= sw.q && v >= sw.n && r <= sw.Q && v <= sw.N;
}

// The printed motif of entry (j, s) of sequence seq: the k-mer ending at j, or for
// structured motifs every box joined by '_' (functions.cc:1191-1231).
std::string motif_text(const Switches &sw, const std::string &seq, unsigned int j, unsigned int s)
{
    unsigned int jj = j - sw.l + 1;
    std::string out = seq.substr(jj, sw.l);
    unsigned int offset = 0, ell = sw.l;
    for (unsigned int b = 1; b < sw.nb_boxes; ++b) {
        offset += sw.S[(size_t)s * sw.nb_gaps + b - 1] + ell;
        ell = sw.blens[b - 1];
        out.push_back('_');
        size_t p = (size_t)jj + offset;
        // the reference reads past the sequence here for entries whose structure does not
        // fit; such entries have u = 0 and can only be printed with -q 0 -n 0
        out += p < seq.size() ? seq.substr(p, ell) : std::string();
    }
    return out;
}

int write_motifs(const Switches &sw, const std::vector<std::string> &seqs,
                 const std::vector<std::vector<uint32_t>> &u, const std::vector<std::vector<uint32_t>> &v,
                 double exectime, int P)
{
    FILE *f = fopen(sw.output_filename.c_str(), "w");
    if (!f) {
        fprintf(stderr, " Error: cannot open file!\n");
        return 0;
    }
    write_header(f, sw, (unsigned int)seqs.size(), exectime, P);
    std::unordered_set<std::string> printed;
    unsigned int num = (unsigned int)seqs.size(), valid = 0;
    for (unsigned int i = 0; i < num; ++i) {
        unsigned int m = (unsigned int)seqs[i].size();
        for (unsigned int j = sw.l - 1; j < m; ++j) {
            for (unsigned int s = 0; s < sw.nb_structs; ++s) {
                size_t o = (size_t)s * m + j;
                if (!passes(sw, num, u[i][o], v[i][o])) continue;
                std::string mot = motif_text(sw, seqs[i], j, s);
                if (!printed.insert(mot).second) continue;
                fprintf(f, "%s %d %d %lf %d\n", mot.c_str(), (int)u[i][o], (int)num,
                        (double)u[i][o] / num, (int)v[i][o]);
                valid++;
            }
        }
    }
    if (sw.nb_boxes == 1)
        fprintf(f, "\n#A total of %d valid motifs were extracted.\n\n", (int)valid);
    else
        fprintf(f, "\n#A total of %d valid structured motifs were extracted.\n\n", (int)valid);
    if (fclose(f)) {
        fprintf(stderr, " Error: file close error!\n");
        return 0;
    }
    return 1;
}

void write_smile_header(FILE *f, const Switches &sw, unsigned int num_seqs, const char *astr)
{
    std::string line1;
    char buf[512];
    int need = (int)ceil((double)((double)sw.q * num_seqs) / (double)100);
    if (sw.nb_boxes == 1) {
        snprintf(buf, sizeof buf, "%%%%%% 1 %d/%d %d %d %d %d alphabet %s$", need, (int)num_seqs,
                 (int)sw.total_length, (int)sw.l, (int)sw.l, (int)sw.e, astr);
        line1 = buf;
    } else {
        unsigned int mn = sw.l, mx = sw.l, te = sw.e;
        for (unsigned int b = 0; b < sw.nb_gaps; ++b) { mn += sw.blens[b]; mx += sw.blens[b]; te += sw.berrs[b]; }
        snprintf(buf, sizeof buf, "%%%%%% %d %d/%d %d %d %d %d %d %d %d ", (int)sw.nb_boxes, need,
                 (int)num_seqs, (int)sw.total_length, (int)mn, (int)mx, (int)te, (int)sw.l, (int)sw.l,
                 (int)sw.e);
        line1 = buf;
        for (unsigned int b = 0; b < sw.nb_gaps; ++b) {
            snprintf(buf, sizeof buf, "%d %d %d %d %d ", (int)sw.bgaps_min[b], (int)sw.bgaps_max[b],
                     (int)sw.blens[b], (int)sw.blens[b], (int)sw.berrs[b]);
            line1 += buf;
        }
        line1 += "alphabet ";
        line1 += astr;
        line1 += "$";
    }
    fputs(line1.c_str(), f);
    if (sw.alphabet == "DNA") {
        for (size_t i = line1.size(); i < 79; ++i) fputc(' ', f);
        fputc('\n', f);
        for (int r = 0; r < 2; ++r) fprintf(f, "%79s\n", "");
    }
    if (sw.alphabet == "PROT") {
        for (size_t i = line1.size(); i < 159; ++i) fputc(' ', f);
        fputc('\n', f);
        fprintf(f, "%79s\n", "");
    }
    for (int r = 0; r < 79; ++r) fputc('=', f);
    fputc('\n', f);
}

// SMILE digit id of a motif: each letter's index in the alphabet string ('0' + index; the
// index one past the end for a letter outside it), and '-' for the '_' box separators of
// structured motifs (functions.cc:1424-1432, 1564-1577).
std::string smile_id(const Switches &sw, const std::string &mot, const char *astr)
{
    const size_t alen = strlen(astr);
    std::string id(mot.size(), '0');
    for (size_t c = 0; c < mot.size(); ++c) {
        size_t a = 0;
        while (a < alen && mot[c] != astr[a]) ++a;
        id[c] = (sw.nb_boxes > 1 && a >= alen) ? '-' : (char)('0' + a);
    }
    return id;
}

int write_motifs_smile(const Switches &sw, const std::vector<std::string> &seqs,
                       const std::vector<std::vector<uint32_t>> &u,
                       const std::vector<std::vector<uint32_t>> &v, double exectime)
{
    const char *astr = alphabet_letters(sw.alphabet);
    FILE *f = fopen(sw.smile_out_filename.c_str(), "w");
    if (!f) {
        fprintf(stderr, " Error: cannot open file!\n");
        return 0;
    }
    unsigned int num = (unsigned int)seqs.size(), valid = 0;
    write_smile_header(f, sw, num, astr);
    std::unordered_set<std::string> printed;
    for (unsigned int i = 0; i < num; ++i) {
        unsigned int m = (unsigned int)seqs[i].size();
        for (unsigned int j = sw.l - 1; j < m; ++j) {
            for (unsigned int s = 0; s < sw.nb_structs; ++s) {
                size_t o = (size_t)s * m + j;
                if (!passes(sw, num, u[i][o], v[i][o])) continue;
                std::string mot = motif_text(sw, seqs[i], j, s);
                if (!printed.insert(mot).second) continue;
                const std::string id = smile_id(sw, mot, astr);
                fprintf(f, "%s %s %d\t%d\n", mot.c_str(), id.c_str(), (int)u[i][o], (int)v[i][o]);
                valid++;
            }
        }
    }
    fprintf(f, "Nb models: %d\n", (int)valid);
    fprintf(f, "User time: %lf sec.\n", exectime);
    if (fclose(f)) {
        fprintf(stderr, " Error: file close error!\n");
        return 0;
    }
    return 1;
}

}  // namespace motex_cli
