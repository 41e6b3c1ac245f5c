// Command-line options: the flag set, defaults and failure behaviour of the reference's
// decode_switches (functions.cc:2835-3043), plus -L, which the reference documents
// (README.md:50-54) but compiles out; here it is accepted as a scheduling hint.
#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "cli.h"

namespace motex_cli {

static struct option long_options[] = {
    {"alphabet", required_argument, NULL, 'a'},
    {"background-in-file", required_argument, NULL, 'b'},
    {"input-file", required_argument, NULL, 'i'},
    {"output-file", required_argument, NULL, 'o'},
    {"quorum", required_argument, NULL, 'q'},
    {"max-quorum", required_argument, NULL, 'Q'},
    {"motifs-length", required_argument, NULL, 'k'},
    {"distance", required_argument, NULL, 'd'},
    {"errors", required_argument, NULL, 'e'},
    {"num-of-occ", required_argument, NULL, 'n'},
    {"max-num-of-occ", required_argument, NULL, 'N'},
    {"num-of-threads", required_argument, NULL, 't'},
    {"long-sequences", required_argument, NULL, 'L'},
    {"boxes-in-file", required_argument, NULL, 's'},
    {"unmatched-in-file", required_argument, NULL, 'I'},
    {"unmatched-out-file", required_argument, NULL, 'u'},
    {"SMILE-out-file", required_argument, NULL, 'S'},
    {"unmatched-SMILE-out-file", required_argument, NULL, 'U'},
    {"help", no_argument, NULL, 'h'},
    {NULL, 0, NULL, 0},
};

// strtol into an unsigned field; false when no digits were read.
static bool parse_uint(const char *s, unsigned int *out)
{
    char *ep = nullptr;
    long val = strtol(s, &ep, 10);
    if (!ep || ep == s) return false;
    *out = (unsigned int)val;
    return true;
}

int decode_switches(int argc, char **argv, Switches *sw)
{
    int oi = 0, opt, args = 0;
    *sw = Switches();
    while ((opt = getopt_long(argc, argv, "a:b:i:o:q:Q:k:d:e:n:N:s:t:I:S:u:U:L:h", long_options,
                              &oi)) != -1) {
        switch (opt) {
        case 'a': sw->alphabet = optarg; args++; break;
        case 'i': sw->input_filename = optarg; args++; break;
        case 'o': sw->output_filename = optarg; args++; break;
        case 'b': sw->background_in_filename = optarg; sw->has_background = true; break;
        case 'I': sw->un_in_filename = optarg; sw->has_un_in = true; break;
        case 'u': sw->un_out_filename = optarg; sw->has_un_out = true; break;
        case 'S': sw->smile_out_filename = optarg; sw->has_smile = true; break;
        case 'U': sw->un_smile_out_filename = optarg; sw->has_un_smile = true; break;
        case 's': sw->boxes_in_filename = optarg; sw->has_boxes = true; break;
        case 'q': if (!parse_uint(optarg, &sw->q)) return 0; args++; break;
        case 'Q': if (!parse_uint(optarg, &sw->Q)) return 0; break;
        case 'k': if (!parse_uint(optarg, &sw->l)) return 0; args++; break;
        case 'n': if (!parse_uint(optarg, &sw->n)) return 0; break;
        case 'N': if (!parse_uint(optarg, &sw->N)) return 0; break;
        case 'd': if (!parse_uint(optarg, &sw->d)) return 0; args++; break;
        case 't': if (!parse_uint(optarg, &sw->t)) return 0; break;
        case 'L': if (!parse_uint(optarg, &sw->L)) return 0; break;
        case 'e': if (!parse_uint(optarg, &sw->e)) return 0; args++; break;
        case 'h': return 0;
        default: break;  // getopt has printed "invalid option"; like the reference, go on
        }
    }
    if (args < 7) {
        usage();
        exit(1);
    }
    return optind;
}

void usage()
{
    fprintf(stdout,
            "\n motex: all-pairs motif extraction on AMD Instinct MI355X (gfx950)\n\n"
            " Usage: motex <options>\n"
            " Standard (Mandatory):\n"
            "  -a, --alphabet            <str>     DNA (ACGTN), PROT (protein letters) or USR\n"
            "                                      (digits 0-9).\n"
            "  -i, --input-file          <str>     (Multi)FASTA input filename.\n"
            "  -o, --output-file         <str>     Output filename.\n"
            "  -d, --distance            <int>     0 for Hamming distance, 1 for edit distance.\n"
            "  -k, --motifs-length       <int>     Motif length.\n"
            "  -e, --errors              <int>     Maximum number of errors per motif occurrence.\n"
            "  -q, --quorum              <int>     Minimum percentage (%%) of sequences in which a\n"
            "                                      motif must occur.\n\n"
            " Optional:\n"
            "  -Q, --max-quorum          <int>     Maximum percentage (%%) of sequences in which a\n"
            "                                      motif may occur (default: 100).\n"
            "  -n, --num-of-occ          <int>     Minimum number of occurrences of a reported\n"
            "                                      motif (default: 1).\n"
            "  -N, --max-num-of-occ      <int>     Maximum number of occurrences of a reported\n"
            "                                      motif (default: 10000).\n"
            "  -s, --structured-motifs   <str>     Boxes file describing structured motifs.\n"
            "  -S, --SMILE-out-file      <str>     Also write SMILE-like output to this file.\n"
            "  -t, --threads             <int>     Accepted for compatibility; the extraction\n"
            "                                      runs on one GPU (default: 4).\n"
            "  -L, --long-sequences      <int>     1 for a few very long sequences; a scheduling\n"
            "                                      hint that never changes results (default: 0).\n"
            "  -b, --background-in-file  <str>     A .motex output of an earlier run, used as the\n"
            "                                      background: each motif gets a binomial p-value.\n"
            "  -u, --unmatched-out-file  <str>     Output file of the motifs not found in the -b file.\n"
            "  -U, --unmatched-SMILE-out-file <str> SMILE-like form of the -u file (needs -S).\n"
            "  -I, --unmatched-in-file   <str>     A -u file to search again in this input, with\n"
            "                                      foreground p-values.\n");
}

}  // namespace motex_cli
