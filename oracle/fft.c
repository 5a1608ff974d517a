/*
 * fft.c -- oracle DFT. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The naive long-double DFT the reference's own FFT tests use as oracle
 * (tests/test_Fft_common.h:52-77: angle from (j*k) mod n, sign -1 forward),
 * with the numpy-style scale of fft/Fft.h:73-76 applied by the caller.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "oracle.h"

void orc_dft_c128(const double* in, double* out, uint64_t n, int forward, double scale)
{
    if (n == 0) return;
    const long double pi = acosl(-1.0L);
    long double* cs = (long double*)malloc(sizeof(long double) * 2 * n);
    for (uint64_t m = 0; m < n; ++m) {
        const long double ang = (forward ? -2.0L : 2.0L) * pi * (long double)m / (long double)n;
        cs[2 * m] = cosl(ang);
        cs[2 * m + 1] = sinl(ang);
    }
#pragma omp parallel for schedule(static)
    for (uint64_t k = 0; k < n; ++k) {
        long double re = 0.0L, im = 0.0L;
        uint64_t m = 0; /* (j*k) mod n, stepped */
        for (uint64_t j = 0; j < n; ++j) {
            const long double xr = in[2 * j], xi = in[2 * j + 1];
            const long double c = cs[2 * m], s = cs[2 * m + 1];
            re += xr * c - xi * s;
            im += xr * s + xi * c;
            m += k;
            if (m >= n) m -= n;
        }
        out[2 * k] = (double)(re * scale);
        out[2 * k + 1] = (double)(im * scale);
    }
    free(cs);
}
