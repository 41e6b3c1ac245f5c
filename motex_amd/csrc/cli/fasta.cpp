// Input readers with the reference's byte-level behaviour:
//   FASTA  -- motex.cc:225-300 (read, upper-case, alphabet check, empty-sequence handling)
//             and motex.cc:409-435 (split into sequences);
//   boxes  -- motex.cc:303-363 (gap count, then min gap, max gap, box length, box errors).
#include <cctype>
#include <cstdio>
#include <cstring>

#include "cli.h"

namespace motex_cli {

int read_fasta(const Switches &sw, const char *alphabet, std::vector<std::string> &seqs)
{
    const char *fname = sw.input_filename.c_str();
    FILE *in = fopen(fname, "r");
    if (!in) {
        fprintf(stderr, " Error: Cannot open file %s!\n", fname);
        return 1;
    }
    seqs.clear();
    std::string cur;
    // `char c` compared with EOF, as the reference does: a 0xFF byte ends the input too.
    char c = (char)fgetc(in);
    do {
        if (c != '>') {
            fprintf(stderr, " Error: input file %s is not in FASTA format!\n", fname);
            fclose(in);
            return 1;
        }
        while ((c = (char)fgetc(in)) != EOF && c != '\n') {
        }
        cur.clear();
        while ((c = (char)fgetc(in)) != EOF && c != '>') {
            if (cur.empty() && c == '\n') {
                fprintf(stderr, " Omitting empty sequence in file %s!\n", fname);
                c = (char)fgetc(in);
                break;
            }
            if (c == '\n') continue;
            c = (char)toupper(c);
            if (strchr(alphabet, c)) {
                cur.push_back(c);
            } else {
                fprintf(stderr, " Error: input file %s contains an unexpected character!\n", fname);
                fclose(in);
                return 1;
            }
        }
        if (!cur.empty()) seqs.push_back(cur);
    } while (c != EOF);
    if (fclose(in)) {
        fprintf(stderr, " Error: file close error!\n");
        return 1;
    }
    // A NUL byte passes the alphabet check above (strchr finds the terminator); the
    // reference then splits its concatenated text at it.  Splitting here keeps that.
    std::vector<std::string> split;
    for (const std::string &s : seqs) {
        size_t p = s.find('\0');
        split.push_back(p == std::string::npos ? s : s.substr(0, p));
    }
    seqs.swap(split);
    if (seqs.empty()) {
        fprintf(stderr, " Error: input file %s contains no sequences!\n", fname);
        return 1;
    }
    return 0;
}

int read_boxes(Switches &sw, unsigned int *total_mot_length)
{
    const char *fname = sw.boxes_in_filename.c_str();
    FILE *f = fopen(fname, "r");
    if (!f) {
        fprintf(stderr, " Error: Cannot open file %s!\n", fname);
        return 1;
    }
    int v;
    auto bad = [&]() {
        fprintf(stderr, " Error: file %s is not in the correct format!\n", fname);
        fclose(f);
        return 1;
    };
    if (fscanf(f, "%d", &v) != 1) return bad();
    sw.nb_gaps = (unsigned int)v;
    if (v < 0 || v > 64) return bad();
    sw.nb_boxes = sw.nb_gaps + 1;
    *total_mot_length = sw.l;
    sw.bgaps_min.assign(sw.nb_gaps, 0);
    sw.bgaps_max.assign(sw.nb_gaps, 0);
    sw.blens.assign(sw.nb_gaps, 0);
    sw.berrs.assign(sw.nb_gaps, 0);
    for (unsigned int i = 0; i < sw.nb_gaps; ++i) {
        if (fscanf(f, "%d", &v) != 1) return bad();
        sw.bgaps_min[i] = (unsigned int)v;
        *total_mot_length += sw.bgaps_min[i];
        if (fscanf(f, "%d", &v) != 1) return bad();
        sw.bgaps_max[i] = (unsigned int)v;
        if (fscanf(f, "%d", &v) != 1) return bad();
        sw.blens[i] = (unsigned int)v;
        *total_mot_length += sw.blens[i];
        if (fscanf(f, "%d", &v) != 1) return bad();
        sw.berrs[i] = (unsigned int)v;
        if (sw.bgaps_max[i] < sw.bgaps_min[i] || sw.blens[i] == 0) return bad();
    }
    if (fclose(f)) {
        fprintf(stderr, " Error: file close error!\n");
        return 1;
    }
    // gap combinations, last gap fastest (functions.cc:710-729)
    sw.nb_structs = 1;
    for (unsigned int g = 0; g < sw.nb_gaps; ++g) sw.nb_structs *= sw.bgaps_max[g] - sw.bgaps_min[g] + 1;
    sw.S.assign((size_t)sw.nb_structs * sw.nb_gaps, 0);
    for (unsigned int s = 0; s < sw.nb_structs; ++s) {
        unsigned int rem = s;
        for (int g = (int)sw.nb_gaps - 1; g >= 0; --g) {
            unsigned int span = sw.bgaps_max[g] - sw.bgaps_min[g] + 1;
            sw.S[(size_t)s * sw.nb_gaps + g] = sw.bgaps_min[g] + rem % span;
            rem /= span;
        }
    }
    return 0;
}

}  // namespace motex_cli
