/* The half step of oracle/yee_native.c for one field type; included once per type with
 * REAL (the type) and NAME(x) (a suffixing macro) defined.  TEST INFRASTRUCTURE ONLY. */

typedef struct NAME(ctx) {
    int forward;
    int n[3];
    long long stride[3];
    const REAL* src;
    REAL pec[3];
    int wrap_mask, pec_mask;
    const oracle_segment* segs;
    int* const* row_seg;
} NAME(ctx);

/* curl[co] += sign * d of curl term t (axis a, input component ci) at cell `cell`, whose coordinate
 * along a is i and whose coordinates are p; CPML memory of the term's segment updated in place. */
static inline __attribute__((always_inline)) void NAME(term)(const NAME(ctx)* k, int t, int a, int ci, int co, int sign,
                                                             long long cell, const int p[3], REAL curl[3]) {
    const int len = k->n[a], i = p[a];
    if (len == 1) return;
    const REAL* src = k->src;
    const long long s = k->stride[a];
    REAL d;
    int interior;
    if (k->forward) {
        interior = i < len - 1;
        if (interior) d = src[3 * (cell + s) + ci] - src[3 * cell + ci];
        else if (k->wrap_mask >> a & 1) d = src[3 * (cell - (long long)(len - 1) * s) + ci] - src[3 * cell + ci];
        else if (k->pec_mask >> a & 1) d = k->pec[a] * src[3 * cell + ci];
        else return;
    } else {
        interior = i >= 1;
        if (interior) d = src[3 * cell + ci] - src[3 * (cell - s) + ci];
        else if (k->wrap_mask >> a & 1) d = src[3 * cell + ci] - src[3 * (cell + (long long)(len - 1) * s) + ci];
        else return;
    }
    const int g_index = interior ? k->row_seg[t][i] : -1;
    if (g_index >= 0) {
        const oracle_segment* g = k->segs + g_index;
        const int w = g->hi - g->lo, q = i - g->lo;
        long long o;
        if (a == 0) o = ((long long)q * k->n[1] + p[1]) * k->n[2] + p[2];
        else if (a == 1) o = ((long long)p[0] * w + q) * k->n[2] + p[2];
        else o = ((long long)p[0] * k->n[1] + p[1]) * w + q;
        REAL* psi = (REAL*)g->psi;
        const REAL b = ((const REAL*)g->b)[q];
        const REAL c = ((const REAL*)g->c)[q];
        const REAL ik = ((const REAL*)g->inv_k)[q];
        REAL v = psi[o] * b;
        const REAL cd = c * d;
        v = v + cd;
        psi[o] = v;
        d = d * ik;
        d = d + v;
    }
    curl[co] = sign > 0 ? curl[co] + d : curl[co] - d;
}

static void NAME(half_step)(int forward, const int n[3], const REAL* src, REAL* dst, const REAL* inverse,
                            int inverse_stride, REAL courant, int wrap_mask, int pec_mask, const double* pec_factor,
                            const oracle_segment* segs, int* const row_seg[6]) {
    NAME(ctx) k;
    k.forward = forward;
    for (int a = 0; a < 3; ++a) {
        k.n[a] = n[a];
        k.pec[a] = (REAL)(-pec_factor[a]);
    }
    k.stride[0] = (long long)n[1] * n[2];
    k.stride[1] = n[2];
    k.stride[2] = 1;
    k.src = src;
    k.wrap_mask = wrap_mask;
    k.pec_mask = pec_mask;
    k.segs = segs;
    k.row_seg = row_seg;
#pragma omp parallel for schedule(static)
    for (int x = 0; x < n[0]; ++x)
        for (int y = 0; y < n[1]; ++y)
            for (int z = 0; z < n[2]; ++z) {
                const int p[3] = {x, y, z};
                const long long cell = x * k.stride[0] + y * k.stride[1] + z;
                REAL curl[3] = {(REAL)0, (REAL)0, (REAL)0};
                /* CURL_TERMS order, torchfdtd/boundaries.py:21-23 */
                NAME(term)(&k, 0, 0, 2, 1, -1, cell, p, curl);
                NAME(term)(&k, 1, 0, 1, 2, 1, cell, p, curl);
                NAME(term)(&k, 2, 1, 2, 0, 1, cell, p, curl);
                NAME(term)(&k, 3, 1, 0, 2, -1, cell, p, curl);
                NAME(term)(&k, 4, 2, 1, 0, -1, cell, p, curl);
                NAME(term)(&k, 5, 2, 0, 1, 1, cell, p, curl);
                for (int c = 0; c < 3; ++c) {
                    const REAL inv = inverse_stride == 0 ? inverse[0]
                                   : inverse_stride == 1 ? inverse[cell] : inverse[3 * cell + c];
                    const REAL f = courant * inv;
                    const REAL u = f * curl[c];
                    dst[3 * cell + c] = forward ? dst[3 * cell + c] - u : dst[3 * cell + c] + u;
                }
            }
}
