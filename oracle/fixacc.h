// Exact fixed-point accumulation for the oracle (test infrastructure only).
//
// Every sum the H.v path forms (matrix elements with many integral terms, the
// per-determinant merge of H.v contributions, dots and norms) is accumulated exactly
// in a signed 128-bit integer whose unit is 2^-84, and rounded to double once at the
// end.  Terms are cut to the 2^-84 grid by truncation toward zero.  An exact integer
// sum does not depend on the order of its terms, so the result is reproducible under
// any parallel order; the GPU library implements the same specification with pairs of
// 64-bit words, this file with the compiler's __int128 and libgcc's correctly rounded
// int128 -> double conversion, so agreement between the two is a real check.
#pragma once
#include <cmath>
#include <cstdint>

namespace oracle {

static const int FIX_FRAC = 84;
typedef __int128 fx_t;

// |x| must be below 2^42 (callers check with fx_ok).
static inline bool fx_ok(double x) { return std::fabs(x) < 4398046511104.0; }

static inline fx_t fx_from(double x) {
    if (x == 0.0 || !std::isfinite(x)) return 0;
    int e;
    double m = std::frexp(std::fabs(x), &e);                 // |x| = m * 2^e, m in [0.5,1)
    uint64_t M = (uint64_t)std::ldexp(m, 53);                 // exact 53-bit integer
    int s = e - 53 + FIX_FRAC;                                // |x| * 2^84 = M * 2^s
    unsigned __int128 mag;
    if (s >= 0) mag = ((unsigned __int128)M) << s;
    else if (s > -64) mag = (unsigned __int128)(M >> (-s));
    else mag = 0;
    fx_t v = (fx_t)mag;
    return x < 0 ? -v : v;
}

static inline double fx_to(fx_t v) {
    if (v == 0) return 0.0;
    bool neg = v < 0;
    unsigned __int128 mag = neg ? (unsigned __int128)(-v) : (unsigned __int128)v;
    double d = (double)mag;                                   // round-to-nearest-even
    d = std::ldexp(d, -FIX_FRAC);
    return neg ? -d : d;
}

}  // namespace oracle
