/*
 * expr.c -- oracle restatement of aether's fused per-sample evaluation.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * One pass per sample, components unrolled, one store per component, no
 * temporaries (expr/Assign.h:124-160 RecursiveAssign, :71-91 storeWorking),
 * driven by an OpenMP loop over samples the way the reference's callers
 * drive it (tests/test_Random_common.h:62-68). Element (c, i) of an SoA
 * array lives at c*pitch + i (array/Array.h:80,100-110,279-295).
 */
#include <math.h>
#include <stdint.h>

#include "oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* Sum(Sum(a, Product(b, c)), Product(CWiseSin(a), CWiseExp(CWiseScale(-1, Product(b, b)))))
 * Arithmetic.h:62-70 (Sum), :100-104 (Product), :138-142 (CWiseScale);
 * Elementwise.h:291 (CWiseExp), :337 (CWiseSin). */
void orc_eval_cfg0_f64(const double* a, const double* b, const double* c, double* d, int ncomp, uint64_t n,
                       uint64_t pitch, int mathmode, int nthreads)
{
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        for (int k = 0; k < ncomp; ++k) {
            const uint64_t o = (uint64_t)k * pitch + i;
            const double av = a[o], bv = b[o], cv = c[o];
            const double lhs = av + bv * cv;
            const double nb2 = -1.0 * (bv * bv);
            const double sn = mathmode == 0 ? sin(av) : orc_sin(av);
            const double ex = mathmode == 0 ? exp(nb2) : orc_exp(nb2);
            d[o] = lhs + sn * ex;
        }
    }
}

void orc_eval_cfg0_f32(const float* a, const float* b, const float* c, float* d, int ncomp, uint64_t n,
                       uint64_t pitch, int mathmode, int nthreads)
{
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        for (int k = 0; k < ncomp; ++k) {
            const uint64_t o = (uint64_t)k * pitch + i;
            const float av = a[o], bv = b[o], cv = c[o];
            const float lhs = av + bv * cv;
            const float nb2 = -1.0f * (bv * bv);
            const float sn = mathmode == 0 ? sinf(av) : orc_sinf(av);
            const float ex = mathmode == 0 ? expf(nb2) : orc_expf(nb2);
            d[o] = lhs + sn * ex;
        }
    }
}

/* Cross: Geometric.h:55-71; dot: Expression.h:162-173 -> Reduce.h:172-188
 * (right-nested: (x2*y2) + ((x1*y1) + (x0*y0))); sqrt: math/math.h:132. */
void orc_eval_vec3geo_f64(const double* a, const double* b, const double* c, double* e, uint64_t n, uint64_t pitch,
                          int nthreads)
{
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        const double a0 = a[i], a1 = a[pitch + i], a2 = a[2 * pitch + i];
        const double b0 = b[i], b1 = b[pitch + i], b2 = b[2 * pitch + i];
        const double c0 = c[i], c1 = c[pitch + i], c2 = c[2 * pitch + i];
        const double x0 = a1 * b2 - a2 * b1;
        const double x1 = a2 * b0 - a0 * b2;
        const double x2 = a0 * b1 - a1 * b0;
        const double dxc = (x2 * c2) + ((x1 * c1) + (x0 * c0));
        const double daa = (a2 * a2) + ((a1 * a1) + (a0 * a0));
        e[i] = dxc + sqrt(daa + 1.0);
    }
}
