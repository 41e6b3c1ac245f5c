// Background and foreground statistics writers:
//   -b/-u   write_motifs_back, functions.cc:1613-1905; structured form 1906-2241
//   -U      write_motifs_back_smile, functions.cc:2243-2466; structured form 2471-2745
//   -I      write_motifs_fore, functions.cc:2750-2832, over the motif lines read back as
//           motex.cc:883-925 reads them
//
// A background file is an earlier .motex output.  Its header block (up to the first empty
// line) gives the background's sequence count (6th line, 3rd field) and quorum percentage
// (8th line, 4th field); the motif lines that follow, up to the next empty line, give
// (motif, u, n, ratio, v).  Lines are read 511 bytes at a time and split at spaces, as the
// reference's fgets/strtok loop does.  A motif line's key is its first k characters (for
// structured motifs, the boxes and their '_' separators); a later line with the same key
// replaces an earlier one, as a trie store does.
//
// Each foreground motif passing the filters is looked up once (first occurrence, like the
// plain writers):
//   - absent from the background: "unmatched", p-value of its u under Bin(N, q_b / N_b) --
//     the background's quorum percentage over its sequence count, as written at
//     functions.cc:1753;
//   - present: "matched", p-value under Bin(N, ratio of the background line).
// The p-value is P[X >= u] (stats.cpp).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>

#include "cli.h"

namespace motex_cli {

static const int kLineSize = 512;  // LINE_SIZE, motexdefs.h:20

// Splits at spaces like strtok(line, " "): empty fields are skipped.
static std::vector<std::string> fields(const char *line, const char *delims = " ")
{
    std::vector<std::string> out;
    std::string cur;
    for (const char *c = line; *c; ++c) {
        if (strchr(delims, *c)) {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(*c);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

// A motif line: motif, u, n, ratio, v (missing fields read as 0).
static MotifRecord parse_record(const char *line)
{
    std::vector<std::string> f = fields(line);
    MotifRecord r;
    if (f.size() > 0) r.motif = f[0];
    if (f.size() > 1) r.u = (unsigned int)atoi(f[1].c_str());
    if (f.size() > 2) r.n = (unsigned int)atoi(f[2].c_str());
    if (f.size() > 3) r.r = atof(f[3].c_str());
    if (f.size() > 4) r.v = (unsigned int)atoi(f[4].c_str());
    return r;
}

struct Background {
    int size = 0, quorum = 0;                        // header lines 6 and 8
    std::vector<MotifRecord> recs;
    std::unordered_map<std::string, long> index;     // key -> record, or -1 once reported
};

static int read_background(const Switches &sw, size_t key_len, Background &bg, bool smile)
{
    FILE *in = fopen(sw.background_in_filename.c_str(), "r");
    if (!in) {
        if (smile) fprintf(stderr, " Error: cannot open file %s!\n", sw.background_in_filename.c_str());
        else fprintf(stderr, " Error: cannot open file!\n");
        return 0;
    }
    char line[kLineSize];
    unsigned int line_num = 0;
    while (fgets(line, kLineSize, in) && line[0] != '\n') {
        ++line_num;
        if (line_num == 6) {
            std::vector<std::string> f = fields(line);
            if (f.size() > 2) bg.size = atoi(f[2].c_str());
        }
        if (line_num == 8) {
            // fields 1-3 split at spaces, the 4th runs to the end of the line
            std::vector<std::string> f = fields(line);
            if (f.size() > 3) {
                const char *p = strstr(line, f[2].c_str()) + f[2].size();
                while (*p == ' ') ++p;
                std::string rest(p);
                if (!rest.empty() && rest.back() == '\n') rest.pop_back();
                bg.quorum = atoi(rest.c_str());
            }
        }
    }
    while (fgets(line, kLineSize, in) && line[0] != '\n') {
        MotifRecord r = parse_record(line);
        bg.index[r.motif.substr(0, key_len)] = (long)bg.recs.size();
        bg.recs.push_back(r);
    }
    if (fclose(in)) {
        fprintf(stderr, " Error: file close error!\n");
        return 0;
    }
    return 1;
}

// Characters of a printed motif: k, or the boxes plus one '_' per gap.
static size_t motif_key_len(const Switches &sw)
{
    size_t len = sw.l;
    for (unsigned int b = 0; b < sw.nb_gaps; ++b) len += sw.blens[b] + 1;
    return len;
}

int write_motifs_back(const Switches &sw, const std::vector<std::string> &seqs,
                      const std::vector<std::vector<uint32_t>> &u,
                      const std::vector<std::vector<uint32_t>> &v, double exectime, int P)
{
    const unsigned int num = (unsigned int)seqs.size();
    BinomialTail tail(num);
    FILE *out = fopen(sw.output_filename.c_str(), "w");
    if (!out) {
        fprintf(stderr, " Error: cannot open file!\n");
        return 0;
    }
    write_header(out, sw, num, exectime, P);
    FILE *un = fopen(sw.un_out_filename.c_str(), "w");
    if (!un) {
        fprintf(stderr, " Error: cannot open file!\n");
        fclose(out);
        return 0;
    }
    write_header(un, sw, num, exectime, P);
    Background bg;
    if (!read_background(sw, motif_key_len(sw), bg, false)) {
        fclose(out);
        fclose(un);
        return 0;
    }
    if (!(bg.size > 0 && bg.quorum > 0)) {
        // the reference asserts this (functions.cc:1708) and aborts
        fprintf(stderr, " Error: background file %s gives no sequence count and quorum in its header!\n",
                sw.background_in_filename.c_str());
        fclose(out);
        fclose(un);
        return 0;
    }
    const long double p_unmatched = (long double)bg.quorum / bg.size;
    unsigned int pvalid = 0, uvalid = 0;
    for (unsigned int i = 0; i < num; ++i) {
        const unsigned int m = (unsigned int)seqs[i].size();
        for (unsigned int j = sw.l - 1; j < m; ++j) {
            for (unsigned int s = 0; s < sw.nb_structs; ++s) {
                const size_t o = (size_t)s * m + j;
                if (!passes(sw, num, u[i][o], v[i][o])) continue;
                const std::string mot = motif_text(sw, seqs[i], j, s);
                auto it = bg.index.find(mot);
                if (it == bg.index.end()) {
                    bg.index.emplace(mot, -1);
                    fprintf(out, "%s %d %d %lf %d %d %d ", mot.c_str(), (int)u[i][o], (int)num,
                            (double)u[i][o] / num, (int)v[i][o], 0, 0);
                    tail.print_upper_tail(out, u[i][o], num, p_unmatched);
                    fprintf(out, "\n");
                    fprintf(un, "%s %d %d %lf %d\n", mot.c_str(), (int)u[i][o], (int)num,
                            (double)u[i][o] / num, (int)v[i][o]);
                    uvalid++;
                } else if (it->second != -1) {
                    const MotifRecord &b = bg.recs[(size_t)it->second];
                    fprintf(out, "%s %d %d %lf %d %lf %d ", mot.c_str(), (int)u[i][o], (int)num,
                            (double)u[i][o] / num, (int)v[i][o], b.r, (int)b.v);
                    tail.print_upper_tail(out, u[i][o], num, (long double)b.r);
                    fprintf(out, "\n");
                    it->second = -1;
                    pvalid++;
                }
            }
        }
    }
    fprintf(out, "\n#A total of %d background motifs were read.\n", (int)bg.recs.size());
    fprintf(out, "#A total of %d valid motifs were extracted:\n", (int)(pvalid + uvalid));
    fprintf(out, "#  %d of them were matched with background motifs\n", (int)pvalid);
    fprintf(out, "#  %d of them were not matched with any background motif.\n\n", (int)uvalid);
    fprintf(un, "\n#A total of %d background motifs were read.\n", (int)bg.recs.size());
    fprintf(un, "#A total of %d valid motifs were extracted:\n", (int)(pvalid + uvalid));
    fprintf(un, "#  %d of them were not matched with any background motif.\n\n", (int)uvalid);
    const bool bad = fclose(out) != 0;
    if (fclose(un) || bad) {
        fprintf(stderr, " Error: file close error!\n");
        return 0;
    }
    return 1;
}

int write_motifs_back_smile(const Switches &sw, const std::vector<std::string> &seqs,
                            const std::vector<std::vector<uint32_t>> &u,
                            const std::vector<std::vector<uint32_t>> &v, double exectime)
{
    const unsigned int num = (unsigned int)seqs.size();
    const char *astr = alphabet_letters(sw.alphabet);
    FILE *un = fopen(sw.un_smile_out_filename.c_str(), "w");
    if (!un) {
        fprintf(stderr, " Error: cannot open file %s!\n", sw.un_smile_out_filename.c_str());
        return 0;
    }
    write_smile_header(un, sw, num, astr);
    FILE *out = fopen(sw.smile_out_filename.c_str(), "w");
    if (!out) {
        fprintf(stderr, " Error: cannot open file %s!\n", sw.smile_out_filename.c_str());
        fclose(un);
        return 0;
    }
    write_smile_header(out, sw, num, astr);
    Background bg;
    if (!read_background(sw, motif_key_len(sw), bg, true)) {
        fclose(out);
        fclose(un);
        return 0;
    }
    unsigned int pvalid = 0, uvalid = 0;
    for (unsigned int i = 0; i < num; ++i) {
        const unsigned int m = (unsigned int)seqs[i].size();
        for (unsigned int j = sw.l - 1; j < m; ++j) {
            for (unsigned int s = 0; s < sw.nb_structs; ++s) {
                const size_t o = (size_t)s * m + j;
                if (!passes(sw, num, u[i][o], v[i][o])) continue;
                const std::string mot = motif_text(sw, seqs[i], j, s);
                auto it = bg.index.find(mot);
                if (it == bg.index.end()) {
                    bg.index.emplace(mot, -1);
                    const std::string id = smile_id(sw, mot, astr);
                    fprintf(un, "%s %s %d\t%d\n", mot.c_str(), id.c_str(), (int)u[i][o], (int)v[i][o]);
                    fprintf(out, "%s %s %d\t%d\n", mot.c_str(), id.c_str(), (int)u[i][o], (int)v[i][o]);
                    uvalid++;
                } else if (it->second != -1) {
                    const std::string id = smile_id(sw, mot, astr);
                    fprintf(out, "%s %s %d\t%d\n", mot.c_str(), id.c_str(), (int)u[i][o], (int)v[i][o]);
                    it->second = -1;
                    pvalid++;
                }
            }
        }
    }
    fprintf(out, "Nb models: %d\n", (int)(pvalid + uvalid));
    fprintf(out, "User time: %lf sec.\n", exectime);
    fprintf(un, "Nb models: %d\n", (int)uvalid);
    fprintf(un, "User time: %lf sec.\n", exectime);
    const bool bad = fclose(un) != 0;
    if (fclose(out) || bad) {
        fprintf(stderr, " Error: file close error!\n");
        return 0;
    }
    return 1;
}

int read_motif_records(const std::string &path, std::vector<MotifRecord> &out)
{
    FILE *in = fopen(path.c_str(), "r");
    if (!in) {
        fprintf(stderr, " Error: Cannot open file %s!\n", path.c_str());
        return 1;
    }
    char line[kLineSize];
    // the header block, up to the first empty line (motex.cc:902 means to skip it in this
    // file; as shipped it reads the already-closed FASTA input instead)
    while (fgets(line, kLineSize, in) && line[0] != '\n') {
    }
    out.clear();
    while (fgets(line, kLineSize, in) && line[0] != '\n') out.push_back(parse_record(line));
    if (fclose(in)) {
        fprintf(stderr, " Error: file close error!\n");
        return 1;
    }
    return 0;
}

int write_motifs_fore(const Switches &sw, const std::vector<MotifRecord> &fg,
                      const std::vector<std::vector<uint32_t>> &u,
                      const std::vector<std::vector<uint32_t>> &v, double exectime, int P,
                      unsigned int num_seqs)
{
    // the tail is taken over each record's own n (its run's sequence count): the factorial
    // table covers the largest
    unsigned long n_max = num_seqs;
    for (const MotifRecord &r : fg) n_max = std::max<unsigned long>(n_max, r.n);
    BinomialTail tail(n_max);
    FILE *out = fopen(sw.output_filename.c_str(), "w");
    if (!out) {
        fprintf(stderr, " Error: cannot open file!\n");
        return 0;
    }
    write_header(out, sw, num_seqs, exectime, P);
    unsigned int valid = 0;
    for (size_t i = 0; i < fg.size(); ++i) {
        const unsigned int m = (unsigned int)fg[i].motif.size();
        if (m == 0) continue;
        const unsigned int j = m - 1;  // the whole motif string
        if (!passes(sw, num_seqs, u[i][j], v[i][j])) continue;
        fprintf(out, "%s %d %d %lf %d %lf %d ", fg[i].motif.c_str(), (int)fg[i].u, (int)fg[i].n, fg[i].r,
                (int)fg[i].v, (double)u[i][j] / num_seqs, (int)v[i][j]);
        tail.print_upper_tail(out, fg[i].u, fg[i].n, (long double)u[i][j] / num_seqs);
        fprintf(out, "\n");
        valid++;
    }
    fprintf(out, "\n#A total of %d valid motifs were extracted.\n\n", (int)valid);
    if (fclose(out)) {
        fprintf(stderr, " Error: file close error!\n");
        return 0;
    }
    return 1;
}

}  // namespace motex_cli
