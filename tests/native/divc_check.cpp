// Host check of fmr_divc and fmr_divc1 (formura_hip_rt.h): the division of a by
// a constant b through y = RN(1/b) and two fused corrections (one, where
// |b*y - 1| <= 2^-54) must equal IEEE a/b bit for bit.  b*y - 1 is exact as a
// fused multiply-add: it is below 2^-52 in magnitude and needs at most 53 bits.
// Cases: random dividends over the fast range, dividends whose quotient lies
// next to a rounding midpoint (the hard cases of correct rounding), stencil-like
// differences, and specials (signed zeros, subnormals, the range limits, Inf, NaN).
// The one-correction sequence (remainder at scale 2^120) is checked without a
// range test on tiny and subnormal dividends, near midpoints included.
#include "formura_hip_rt.h"
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
static uint64_t s = 0x9E3779B97F4A7C15ull;
static inline uint64_t nx() { uint64_t z = (s += 0x9E3779B97F4A7C15ull); z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull; return z ^ (z >> 31); }
static inline uint64_t bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double dbl(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }
// the remainder formed unscaled (it underflows for tiny dividends): the tiny
// cases below must catch it
static inline double divc1_unscaled(double a, double b, double y) {
  const double q = a * y;
  return fma(-fma(b, q, -a), y, q);
}
static long unscaled_bad = 0;
int main(int argc, char** argv) {
  long n = argc > 1 ? atol(argv[1]) : 10000000;
  if (argc > 2) s = strtoull(argv[2], nullptr, 0);
  // fixed divisors, then random ones over [2^-100, 2^100) with fewer dividends each
  const int NFIX = 12, NRND = 48;
  double bs[NFIX + NRND] = {1e-4, 0.01, 3.0, 7.0, 0.1, 1.0/3.0, 20.48, 1e-5, 1.7976931348623157e+30, 0x1.fffffffffffffp-1, 0x1.0000000000001p0, 9.999999999999999e-5};
  for (int k = 0; k < NRND; ++k) bs[NFIX + k] = ldexp(1.0 + (double)(nx() >> 12) * 0x1p-52, (int)(nx() % 200) - 100);
  long bad = 0, tot = 0, one_step = 0;
  for (int bi = 0; bi < NFIX + NRND; ++bi) {
    const double b = bs[bi];
    double y = 1.0 / b;
    const bool one = fabs(fma(b, y, -1.0)) <= 0x1p-54;
    one_step += one;
    const long nb = bi < NFIX ? n : n / 16;
    for (long i = 0; i < nb; ++i) {
      uint64_t r = nx();
      double a;
      int kind = i % 3;
      if (kind == 0) {  // random bits over the fast range
        int e = (int)(r % 1800) - 900;
        a = ldexp(1.0 + (double)(nx() >> 12) * 0x1p-52, e);
        if (r & (1ull << 63)) a = -a;
      } else if (kind == 1) {  // near midpoints: a = RN(b*m) with m a midpoint
        double q = ldexp(1.0 + (double)(nx() >> 12) * 0x1p-52, (int)(r % 200) - 100);
        double m = q + ldexp(1.0, ilogb(q) - 53);
        a = b * m;
        if (r & 1) a = nextafter(a, 0.0);
        if (r & 2) a = nextafter(a, 1e308);
      } else {  // values like stencil differences
        a = ((double)(nx() >> 11) * 0x1p-53 - 0.5) * ldexp(1.0, (int)(r % 60) - 50);
      }
      double q1 = fmr_divc(a, b, y), q2 = a / b;
      ++tot;
      if (bits(q1) != bits(q2)) { if (bad < 10) printf("MISMATCH b=%a a=%a got %a want %a\n", b, a, q1, q2); ++bad; }
      if (fmr_divc_in_range(a)) {  // the branch-free sequence the speculative forms use
        const double q4 = fmr_divc_fast(a, b, y);
        ++tot;
        if (bits(q4) != bits(q2)) { if (bad < 10) printf("MISMATCH (fast) b=%a a=%a got %a want %a\n", b, a, q4, q2); ++bad; }
        if (one && bits(fmr_divc1_fast(a, b, y)) != bits(q2)) { if (bad < 10) printf("MISMATCH (fast, one step) b=%a a=%a\n", b, a); ++bad; }
      }
      if (one) {
        const double q3 = fmr_divc1(a, b, y);
        ++tot;
        if (bits(q3) != bits(q2)) { if (bad < 10) printf("MISMATCH (one step) b=%a a=%a got %a want %a\n", b, a, q3, q2); ++bad; }
      }
    }
    // the redo of a wave with out-of-range dividends: tiny and subnormal ones
    // (scaled by 2^600 where |a| >= lo, else the division), huge ones
    {
      int e;
      const double m = frexp(b, &e);
      const double lo = ldexp(1.0, (m == 0.5 ? e - 1 : e) - 1021);
      const long nr = bi < NFIX ? n / 4 : n / 64;
      for (long i = 0; i < nr; ++i) {
        uint64_t r = nx();
        double a;
        if (i % 4 == 3) a = dbl(r >> 12);  // subnormals and tiny normals, random bits
        else a = ldexp(1.0 + (double)(nx() >> 12) * 0x1p-52, (int)(r % 200) - 1100);  // 2^-1100 .. 2^-901
        if (i % 5 == 4) a = ldexp(1.0 + (double)(nx() >> 12) * 0x1p-52, 900 + (int)(r % 120));  // huge
        if (r & (1ull << 62)) a = -a;
        const double want = a / b;
        ++tot;
        if (bits(fmr_divc_redo(a, b, y, lo)) != bits(want)) { if (bad < 10) printf("MISMATCH (redo) b=%a a=%a got %a want %a\n", b, a, fmr_divc_redo(a, b, y, lo), want); ++bad; }
        if (one && bits(fmr_divc1_redo(a, b, y, lo)) != bits(want)) { if (bad < 10) printf("MISMATCH (redo, one step) b=%a a=%a\n", b, a); ++bad; }
        // the one-correction sequence forms its remainder at scale 2^120: exact
        // without a range test for every |a| <= 2^900, tiny and subnormal too
        if (one && fmr_divc1_in_range(a)) {
          ++tot;
          if (bits(fmr_divc1_fast(a, b, y)) != bits(want)) { if (bad < 10) printf("MISMATCH (one step, tiny) b=%a a=%a got %a want %a\n", b, a, fmr_divc1_fast(a, b, y), want); ++bad; }
        }
      }
      // tiny dividends whose quotient lies next to a rounding midpoint, in the
      // normal range and on the subnormal grid
      if (one) {
        for (long i = 0; i < nr; ++i) {
          uint64_t r = nx();
          double m;
          if (i & 1) {  // quotient a subnormal midpoint: (k + 1/2) 2^-1074
            m = ldexp((double)((r >> 12) & ((1ull << 51) - 1)) + 0.5, -1074);
          } else {  // a normal quotient's midpoint, 2^-1022 .. 2^-900 / b
            double q = ldexp(1.0 + (double)(nx() >> 12) * 0x1p-52, -1022 + (int)(r % 100));
            m = q + ldexp(1.0, ilogb(q) - 53);
          }
          double a = b * m;
          if (r & 4) a = nextafter(a, 0.0);
          if (r & 8) a = nextafter(a, 1.0);
          if (r & 16) a = -a;
          if (!(fabs(a) <= 0x1p-890)) continue;
          const double want = a / b;
          ++tot;
          if (bits(fmr_divc1_fast(a, b, y)) != bits(want)) { if (bad < 10) printf("MISMATCH (one step, tiny midpoint) b=%a a=%a got %a want %a\n", b, a, fmr_divc1_fast(a, b, y), want); ++bad; }
          unscaled_bad += bits(divc1_unscaled(a, b, y)) != bits(want);
        }
      }
    }
    double sp[] = {0.0, -0.0, 1e-310, -1e-310, 1e308, -1e308, INFINITY, -INFINITY, NAN, 0x1p-900, 0x1p900, 0x1p-901, 0x1.fffffffffffffp900, 5e-324};
    // the range test on the specials: in range exactly the signed zeros and 2^-900, 2^900
    for (double a : sp) {
      const bool want = a == 0.0 || fabs(a) == 0x1p-900 || fabs(a) == 0x1p900;
      if (fmr_divc_in_range(a) != want) { printf("RANGE MISMATCH a=%a\n", a); ++bad; }
      // the one-correction range: every |a| <= 2^900 (zeros, subnormals, tiny)
      const bool want1 = fabs(a) <= 0x1p900;
      if (fmr_divc1_in_range(a) != want1) { printf("RANGE MISMATCH (one step) a=%a\n", a); ++bad; }
    }
    for (double a : sp) {
      double q1 = fmr_divc(a, b, y), q2 = a / b, q3 = one ? fmr_divc1(a, b, y) : q2;
      ++tot;
      if (bits(q1) != bits(q2) && !(q1 != q1 && q2 != q2)) { printf("SPECIAL MISMATCH b=%a a=%a got %a want %a\n", b, a, q1, q2); ++bad; }
      if (bits(q3) != bits(q2) && !(q3 != q3 && q2 != q2)) { printf("SPECIAL MISMATCH (one step) b=%a a=%a got %a want %a\n", b, a, q3, q2); ++bad; }
      if (one && fmr_divc1_in_range(a) && bits(fmr_divc1_fast(a, b, y)) != bits(q2)) { printf("SPECIAL MISMATCH (one step, fast) b=%a a=%a\n", b, a); ++bad; }
      int e;
      const double mb = frexp(b, &e);
      const double q5 = fmr_divc_redo(a, b, y, ldexp(1.0, (mb == 0.5 ? e - 1 : e) - 1021));
      if (bits(q5) != bits(q2) && !(q5 != q5 && q2 != q2)) { printf("SPECIAL MISMATCH (redo) b=%a a=%a got %a want %a\n", b, a, q5, q2); ++bad; }
    }
  }
  printf("divisors %d, one-step divisors %ld\n", NFIX + NRND, one_step);
  printf("checked %ld, mismatches %ld\n", tot, bad);
  printf("unscaled remainder, tiny midpoint cases wrong: %ld\n", unscaled_bad);
  return bad != 0 || unscaled_bad == 0;
}
