/*
 * Oracle: the Yee/CPML half steps of oracle/yee.py in C -- TEST INFRASTRUCTURE ONLY.
 *
 * A second restatement of torchfdtd/boundaries.py:363-438 with the reference's operation
 * order, for the tests that compare the engine with the oracle at configuration sizes
 * (the NumPy restatement needs about 0.12 s per step of a 768 k-cell grid).  Each product
 * and sum rounds to the field type on its own, exactly as the NumPy ufuncs of
 * oracle/yee.py do:
 *   d = f[high] - f[low]
 *   psi = psi*b;  psi = psi + c*d;  d = d*inv_k;  d = d + psi        (CPML rows)
 *   curl[out] = curl[out] + sign*d                                   (CURL_TERMS order, from +0)
 *   E = E + (S*inv_eps)*curl        H = H - (S*inv_mu)*curl
 * with the periodic wrap and the PEC upper ghost node of torchfdtd/boundaries.py:384-393.
 * Compiled with -ffp-contract=off and without fast-math (no FMA, no reassociation,
 * subnormals kept), so the result equals oracle/yee.py bit for bit;
 * tests/test_oracle_native.py checks that on every boundary layout.  Cells are independent
 * within a half step, so the OpenMP split over x planes does not change any value.
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>

typedef struct oracle_segment {
    int32_t axis, comp, lo, hi;   /* target rows [lo, hi) along axis */
    void* psi;                    /* grid shape with extent hi-lo along axis, C order */
    const void* b;
    const void* c;
    const void* inv_k;
} oracle_segment;

/* (derivative axis, input component): the first two entries of torchfdtd/boundaries.py:21-23 */
static const int TERM_AXIS[6] = {0, 0, 1, 1, 2, 2};
static const int TERM_COMP[6] = {2, 1, 2, 0, 1, 0};

#define REAL float
#define NAME(x) x##_f32
#include "yee_native_body.h"
#undef REAL
#undef NAME

#define REAL double
#define NAME(x) x##_f64
#include "yee_native_body.h"
#undef REAL
#undef NAME

/* One half step: forward = 0 is the E update (backward differences of src = H into dst = E),
 * forward = 1 the H update (forward differences of src = E into dst = H).  dtype 0 float, 1 double.
 * n = (nx, ny, nz); inverse_stride 0 (one value), 1 (per cell) or 3 (per cell and component);
 * wrap_mask / pec_mask: bit a for a periodic axis / a PEC upper face (H update only).
 * Returns 0, or 1 for invalid arguments. */
int oracle_half_step(int dtype, int forward, const int32_t* n, const void* src, void* dst, const void* inverse,
                     int inverse_stride, double courant, int wrap_mask, int pec_mask, const double* pec_factor, int nseg,
                     const oracle_segment* segs) {
    const int m[3] = {n[0], n[1], n[2]};
    if (m[0] < 1 || m[1] < 1 || m[2] < 1 || (inverse_stride != 0 && inverse_stride != 1 && inverse_stride != 3)) return 1;
    if (dtype != 0 && dtype != 1) return 1;
    if (!forward) pec_mask = 0;
    /* row_seg[t][i]: the segment of curl term t that covers target row i along its axis, or -1 */
    int* row_seg[6];
    for (int t = 0; t < 6; ++t) {
        const int a = TERM_AXIS[t];
        row_seg[t] = (int*)malloc(sizeof(int) * (size_t)m[a]);
        if (!row_seg[t]) {
            for (int k = 0; k < t; ++k) free(row_seg[k]);
            return 1;
        }
        for (int i = 0; i < m[a]; ++i) row_seg[t][i] = -1;
        for (int s = 0; s < nseg; ++s) {
            const oracle_segment* g = segs + s;
            if (g->axis != a || g->comp != TERM_COMP[t]) continue;
            for (int i = g->lo < 0 ? 0 : g->lo; i < g->hi && i < m[a]; ++i) row_seg[t][i] = s;
        }
    }
    if (dtype == 0)
        half_step_f32(forward, m, (const float*)src, (float*)dst, (const float*)inverse, inverse_stride, (float)courant,
                      wrap_mask, pec_mask, pec_factor, segs, row_seg);
    else
        half_step_f64(forward, m, (const double*)src, (double*)dst, (const double*)inverse, inverse_stride, courant,
                      wrap_mask, pec_mask, pec_factor, segs, row_seg);
    for (int t = 0; t < 6; ++t) free(row_seg[t]);
    return 0;
}
