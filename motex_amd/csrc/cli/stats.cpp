// Binomial upper-tail p-values of the background/foreground statistics (-b/-u/-U, -I).
//
// binomial.cc:25-74 (the MPFR build, -D_USE_MPFR, which every shipped Makefile selects):
//   LUT[i]  = i!                                    (LUT[0] = 1, LUT[i] = LUT[i-1] * i)
//   cdf     = sum_{i < x} LUT[N] / LUT[i] / LUT[N-i] * p^i * (1-p)^(N-i)
//   p-value = 1 - cdf, printed as mpfr_out_str(f, 10, 10, ., round up) prints it
// Every value is a 200-bit MPFR number and every operation rounds upward, in that order of
// operations; MPFR rounds each operation correctly, so the same sequence gives the same bits
// and the same 10 printed digits.  The per-term evaluation order is therefore kept as written
// there.  p >= 1 gives cdf = 0.  The reference reads its table out of bounds for x > N + 1
// (C(N, i) = 0 there); here the sum stops at i = N.
#include <cstdio>

#include "cli.h"

#ifdef MOTEX_HAVE_MPFR
#include <mpfr.h>

namespace motex_cli {

static const mpfr_prec_t kPrec = 200;  // ACC, motexdefs.h:82

struct BinomialTail::Impl {
    std::vector<__mpfr_struct> lut;  // sized once; entry i is i!
};

BinomialTail::BinomialTail(unsigned long n_max) : impl(new Impl)
{
    impl->lut.resize(n_max + 1);
    mpfr_init2(&impl->lut[0], kPrec);
    mpfr_set_ui(&impl->lut[0], 1, MPFR_RNDU);
    for (unsigned long i = 1; i <= n_max; ++i) {
        mpfr_init2(&impl->lut[i], kPrec);
        mpfr_mul_ui(&impl->lut[i], &impl->lut[i - 1], i, MPFR_RNDU);
    }
}

BinomialTail::~BinomialTail()
{
    for (auto &v : impl->lut) mpfr_clear(&v);
    delete impl;
}

bool BinomialTail::available() { return true; }

// The text mpfr_out_str(f, 10, 10, v, MPFR_RNDU) prints in MPFR 3.1.2, the version the
// reference ships and builds against (mpfr-3.1.2 src/out_str.c): the ten significant digits
// rounded up, as "d.ddddddddd", then "e<exponent>" unless the exponent is 0.  (MPFR 4 prints
// "e0" too, so the digits come from mpfr_get_str and the layout is written here.)
static void print_out_str(FILE *f, mpfr_srcptr v)
{
    if (mpfr_nan_p(v)) { fputs("@NaN@", f); return; }
    if (mpfr_inf_p(v)) { fputs(mpfr_sgn(v) > 0 ? "@Inf@" : "-@Inf@", f); return; }
    if (mpfr_zero_p(v)) { fputs(mpfr_signbit(v) ? "-0" : "0", f); return; }
    mpfr_exp_t e = 0;
    char *s = mpfr_get_str(nullptr, &e, 10, 10, v, MPFR_RNDU);
    const char *d = s;
    if (*d == '-') fputc(*d++, f);
    fputc(*d++, f);
    fputc('.', f);
    fputs(d, f);
    if (--e) fprintf(f, "e%ld", (long)e);
    mpfr_free_str(s);
}

void BinomialTail::print_upper_tail(FILE *f, unsigned long x, unsigned long N, long double p)
{
    mpfr_t t, mp, om, pr, a, b;
    for (mpfr_ptr z : {t, mp, om, pr, a, b}) mpfr_init2(z, kPrec);
    mpfr_set_ld(t, 0.0L, MPFR_RNDU);
    if (p < 1.0L && N < impl->lut.size()) {
        mpfr_set_ld(mp, p, MPFR_RNDU);
        mpfr_ui_sub(om, 1, mp, MPFR_RNDU);
        const unsigned long last = x <= N ? x : N + 1;
        for (unsigned long i = 0; i < last; ++i) {
            mpfr_pow_ui(a, mp, i, MPFR_RNDU);
            mpfr_pow_ui(b, om, N - i, MPFR_RNDU);
            mpfr_set(pr, &impl->lut[N], MPFR_RNDU);
            mpfr_div(pr, pr, &impl->lut[i], MPFR_RNDU);
            mpfr_div(pr, pr, &impl->lut[N - i], MPFR_RNDU);
            mpfr_mul(pr, pr, a, MPFR_RNDU);
            mpfr_mul(pr, pr, b, MPFR_RNDU);
            mpfr_add(t, t, pr, MPFR_RNDU);
        }
    }
    mpfr_ui_sub(t, 1, t, MPFR_RNDU);
    print_out_str(f, t);
    for (mpfr_ptr z : {t, mp, om, pr, a, b}) mpfr_clear(z);
}

}  // namespace motex_cli

#else  // built without MPFR: the statistics modes are refused

namespace motex_cli {

struct BinomialTail::Impl {};
BinomialTail::BinomialTail(unsigned long) : impl(nullptr) {}
BinomialTail::~BinomialTail() {}
bool BinomialTail::available() { return false; }
void BinomialTail::print_upper_tail(FILE *, unsigned long, unsigned long, long double) {}

}  // namespace motex_cli

#endif
