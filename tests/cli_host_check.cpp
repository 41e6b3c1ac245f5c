// Host-side check program for the motex command line's writers and statistics (no GPU).
//
//   cli_host_check tail            reads "x N kind a b" lines on stdin and prints the p-value
//                                  the command line writes for them: kind q -> p = (long
//                                  double)a / b with integers a, b; kind d -> p = (long double)
//                                  strtod(a)
//   cli_host_check run UV ARGS...  runs the command line's driver (options, readers, writers)
//                                  with the extraction replaced by u/v read from the file UV:
//                                  the rows in order, all u rows then all v rows, uint32 each
//                                  (len x structures words per row; for -I, one row per motif)
//
// Built by tests/test_stats.py from the command line's own sources.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cli/cli.h"

using namespace motex_cli;

namespace {

class FileExtractor : public Extractor {
  public:
    explicit FileExtractor(const char *path) : path_(path) {}
    int all_pairs(const char *const *, const uint32_t *lens, uint32_t num, const motex_params &p,
                  uint32_t *const *u, uint32_t *const *v) override
    {
        uint64_t ns = 1;
        for (uint32_t g = 0; g < p.nb_gaps; ++g) ns *= p.gap_max[g] - p.gap_min[g] + 1;
        return load(lens, num, ns, u, v);
    }
    int cross(const char *const *, const uint32_t *xl, uint32_t nx, const char *const *, const uint32_t *,
              uint32_t, const motex_params &, uint32_t *const *u, uint32_t *const *v) override
    {
        return load(xl, nx, 1, u, v);
    }
    const char *last_error() override { return err_.c_str(); }

  private:
    int load(const uint32_t *lens, uint32_t num, uint64_t ns, uint32_t *const *u, uint32_t *const *v)
    {
        FILE *f = fopen(path_, "rb");
        if (!f) { err_ = "cannot open the u/v file"; return MOTEX_EINVAL; }
        for (int pass = 0; pass < 2; ++pass)
            for (uint32_t i = 0; i < num; ++i) {
                const size_t n = (size_t)lens[i] * ns;
                std::vector<uint32_t> row(n);
                if (n && fread(row.data(), 4, n, f) != n) { fclose(f); err_ = "short u/v file"; return MOTEX_EINVAL; }
                uint32_t *dst = pass ? v[i] : u[i];
                for (size_t t = 0; t < n; ++t) dst[t] += row[t];
            }
        fclose(f);
        return MOTEX_OK;
    }
    const char *path_;
    std::string err_;
};

}  // namespace

int main(int argc, char **argv)
{
    if (argc >= 2 && !strcmp(argv[1], "tail")) {
        if (!BinomialTail::available()) return 2;
        unsigned long x, N;
        char kind[8], a[128], b[128];
        std::vector<std::string> lines;
        while (scanf("%lu %lu %7s %127s %127s", &x, &N, kind, a, b) == 5) {
            long double p = kind[0] == 'q' ? (long double)strtoul(a, nullptr, 10) / strtoul(b, nullptr, 10)
                                           : (long double)strtod(a, nullptr);
            BinomialTail t(N);
            t.print_upper_tail(stdout, x, N, p);
            fputc('\n', stdout);
        }
        return 0;
    }
    if (argc >= 3 && !strcmp(argv[1], "run")) {
        FileExtractor ex(argv[2]);
        // argv[2] becomes the program name the driver sees
        return run_cli(argc - 2, argv + 2, ex);
    }
    fprintf(stderr, "usage: cli_host_check tail | run UV ARGS...\n");
    return 2;
}
