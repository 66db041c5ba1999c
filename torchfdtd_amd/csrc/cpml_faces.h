// CPML slabs of one field family in the compact layout of tfdtd_march_family: per axis and
// face, the rows [lo, hi), the element offsets of the two curl terms' memory arrays in one
// buffer, and one coefficient block b | c | inv_k.  Shared by the one-pass step and the
// adjoint step.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/tfdtd.h"
#include "common.h"

namespace {

struct Face {          // one face's slab along an axis: rows [lo, hi), shared by the axis's two terms
    int lo, hi;        // no slab: lo == hi == 0
    unsigned psi[2];   // element offsets of the two terms' psi arrays (curl order)
    unsigned coef;     // b at coef, c at coef + w, inv_k at coef + 2w (w = hi - lo)
};

template <typename real>
struct Family {
    const real* psi_in;
    real* psi_out;     // equals psi_in for the in-place H family
    const real* coef;
    Face face[3][2];   // [axis][low, high]
};

struct Hit {           // slab membership of a coordinate: face -1 = none
    int face, q, w;
    unsigned psi0, psi1, coef;
};

template <typename real>
__device__ __forceinline__ Hit find_hit(const Family<real>& F, int axis, int k) {
    const Face& f0 = F.face[axis][0];
    const Face& f1 = F.face[axis][1];
    if (k >= f0.lo && k < f0.hi) return Hit{0, k - f0.lo, f0.hi - f0.lo, f0.psi[0], f0.psi[1], f0.coef};
    if (k >= f1.lo && k < f1.hi) return Hit{1, k - f1.lo, f1.hi - f1.lo, f1.psi[0], f1.psi[1], f1.coef};
    return Hit{-1, 0, 1, 0u, 0u, 0u};
}

template <typename real>
int make_family(const tfdtd_march_family& in, bool split, const int dims[3], Family<real>* out) {
    std::memset(out, 0, sizeof *out);
    if (!in.psi_in || !in.psi_out || !in.coef) return tfdtd_fail(TFDTD_EINVAL, "cpml: null CPML buffer");
    if (split == (in.psi_in == in.psi_out))
        return tfdtd_fail(TFDTD_EINVAL, split ? "cpml: this family needs separate input and output memory buffers"
                                              : "cpml: this family updates its memory in place");
    out->psi_in = static_cast<const real*>(in.psi_in);
    out->psi_out = static_cast<real*>(in.psi_out);
    out->coef = static_cast<const real*>(in.coef);
    for (int ax = 0; ax < 3; ++ax) {
        for (int f = 0; f < 2; ++f) {
            const tfdtd_march_face& face = in.face[ax][f];
            if (face.lo == face.hi) continue;
            if (face.lo < 0 || face.hi < face.lo || face.hi > dims[ax])
                return tfdtd_fail(TFDTD_EINVAL, "cpml: face %d of axis %d has rows outside the axis", f, ax);
            if (face.psi[0] < 0 || face.psi[1] < 0 || face.coef < 0 || face.psi[0] > 0xffffffffLL ||
                face.psi[1] > 0xffffffffLL || face.coef > 0xffffffffLL)
                return tfdtd_fail(TFDTD_EINVAL, "cpml: CPML offsets must be non-negative 32-bit element offsets");
            out->face[ax][f] = Face{face.lo, face.hi, {(unsigned)face.psi[0], (unsigned)face.psi[1]}, (unsigned)face.coef};
        }
        const Face& lo = out->face[ax][0];
        const Face& hi = out->face[ax][1];
        if (lo.hi > lo.lo && hi.hi > hi.lo && lo.hi > hi.lo && hi.hi > lo.lo)
            return tfdtd_fail(TFDTD_EINVAL, "cpml: the two faces of axis %d overlap", ax);
    }
    return TFDTD_OK;
}

}  // namespace
