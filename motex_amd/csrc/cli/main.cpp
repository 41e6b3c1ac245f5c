// motex: the command-line tool (drop-in for motexCPU/motexOMP).  The driver is app.cpp;
// here the extraction runs on GPU 0 through the C ABI of libmotex.
#include "motex.h"
#include "cli.h"

namespace {

class GpuExtractor : public motex_cli::Extractor {
  public:
    ~GpuExtractor() override { motex_ctx_destroy(ctx_); }
    int all_pairs(const char *const *seqs, const uint32_t *lens, uint32_t num, const motex_params &p,
                  uint32_t *const *u, uint32_t *const *v) override
    {
        int rc = open();
        return rc ? rc : motex_extract_all_pairs(ctx_, seqs, lens, num, &p, u, v);
    }
    int cross(const char *const *xs, const uint32_t *xl, uint32_t nx, const char *const *ys, const uint32_t *yl,
              uint32_t ny, const motex_params &p, uint32_t *const *u, uint32_t *const *v) override
    {
        int rc = open();
        return rc ? rc : motex_extract_cross(ctx_, xs, xl, nx, ys, yl, ny, &p, u, v);
    }
    const char *last_error() override { return motex_last_error(); }

  private:
    int open() { return ctx_ ? MOTEX_OK : motex_ctx_create(0, 0, &ctx_); }
    motex_ctx *ctx_ = nullptr;
};

}  // namespace

int main(int argc, char **argv)
{
    GpuExtractor gpu;
    return motex_cli::run_cli(argc, argv, gpu);
}
