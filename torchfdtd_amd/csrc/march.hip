// One-pass Yee step for gfx950: E and H of every cell in a single launch.
//
// The step reads the state at time n from one buffer set (E, H and the E-family CPML
// memory) and writes time n+1 into the other, so every read sees time-n values no
// matter how far other workgroups have progressed; the caller alternates the sets
// between steps.  H-family CPML memory is touched only by the owner of its cell and
// stays in place.
//
// Work decomposition.  A wavefront owns TY rows (y) of a 62-cell z window and marches
// along x over a chunk of planes.  Lanes cover z0-1 .. z0+62: lane 0 only supplies
// H(z-1) to lane 1, lane 63 computes the E(z+1) that lane 62 needs, lanes 1..62 own
// their cells.  Neighbours along z move between lanes with shuffles, along y they are
// the previous/next row of the same thread, and along x they are carried in registers
// from the previous plane.  Rows y0-1 (H only) and y0+TY (E recomputed) are halos.
// For each plane x and row j: E(x, y0+j), then H(x-1, y0+j-1) from E(x-1) (carried),
// E(x) of the row just computed and the row above.  Four wavefronts of a workgroup
// take consecutive y tiles; the tile order keeps neighbouring tiles on one XCD so
// their shared halo rows are served from the same L2.
//
// CPML slab membership is uniform per plane along x, uniform per row along y and
// fixed per lane along z; the z-slab coefficients are loaded once per lane.
//
// Every product and sum rounds in the reference's order (torchfdtd/boundaries.py:363-438),
// so the result equals the split half-step kernels and the NumPy path bit for bit.
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

template <typename real>
struct MSource {
    int comp;
    int lo[3], hi[3];
    const real* values;
};

constexpr int kMarchSources = 8;

template <typename real>
struct March {
    int nx, ny, nz;
    const real* e_in;
    const real* h_in;
    real* e_out;
    real* h_out;
    const real* inv_eps;
    int eps_stride;
    const real* inv_mu;
    real courant;
    int pec_mask;
    real pec_factor[3];
    Family<real> ef;
    Family<real> hf;
    int nsrc_e, nsrc_h;
    int ebox_lo[3], ebox_hi[3], hbox_lo[3], hbox_hi[3];
    MSource<real> se[kMarchSources];
    MSource<real> sh[kMarchSources];
    long long source_samples;
    const long long* counter;
    int chunk, zwin, ygroups;
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

// psi <- psi*b + c*d (stored when `write`), d <- d*inv_k + psi, for slab row q of term k.
template <typename real>
__device__ __forceinline__ real cpml(real d, const Family<real>& F, const Hit& hit, int k, unsigned p, bool write) {
    const unsigned o = (k ? hit.psi1 : hit.psi0) + p;
    const real b = F.coef[hit.coef + hit.q];
    const real c = F.coef[hit.coef + hit.w + hit.q];
    const real ik = F.coef[hit.coef + 2 * hit.w + hit.q];
    const real m = F.psi_in[o] * b + c * d;
    if (write) F.psi_out[o] = m;
    return d * ik + m;
}

template <typename real>
struct ZCoef {         // per-lane z-slab coefficients (shared by the two z terms of a family)
    real b, c, ik;
};

template <typename real>
__device__ __forceinline__ real cpml_lane(real d, const Family<real>& F, const Hit& hit, const ZCoef<real>& zc, int k,
                                          unsigned p, bool write) {
    const unsigned o = (k ? hit.psi1 : hit.psi0) + p;
    const real m = F.psi_in[o] * zc.b + zc.c * d;
    if (write) F.psi_out[o] = m;
    return d * zc.ik + m;
}

template <typename real>
__device__ __forceinline__ ZCoef<real> lane_coefficients(const Family<real>& F, const Hit& hit) {
    ZCoef<real> zc{real(0), real(0), real(1)};
    if (hit.face >= 0) {
        zc.b = F.coef[hit.coef + hit.q];
        zc.c = F.coef[hit.coef + hit.w + hit.q];
        zc.ik = F.coef[hit.coef + 2 * hit.w + hit.q];
    }
    return zc;
}

template <typename real>
__device__ __forceinline__ void add_sources(const MSource<real>* table, int count, const int* lo, const int* hi,
                                            long long n, long long samples, int x, int y, int z, real& vx, real& vy, real& vz) {
    if (count == 0 || n >= samples) return;
    if (x < lo[0] || x >= hi[0] || y < lo[1] || y >= hi[1] || z < lo[2] || z >= hi[2]) return;
    for (int k = 0; k < count; ++k) {
        const MSource<real>& s = table[k];
        if (x >= s.lo[0] && x < s.hi[0] && y >= s.lo[1] && y < s.hi[1] && z >= s.lo[2] && z < s.hi[2]) {
            const real v = s.values[n];
            vx = s.comp == 0 ? vx + v : vx;
            vy = s.comp == 1 ? vy + v : vy;
            vz = s.comp == 2 ? vz + v : vz;
        }
    }
}

template <typename real>
__device__ __forceinline__ real shfl_up1(real v) { return __shfl_up(v, 1, 64); }
template <typename real>
__device__ __forceinline__ real shfl_down1(real v) { return __shfl_down(v, 1, 64); }

template <typename real>
struct V3 {
    real x, y, z;
};

template <typename real>
__device__ __forceinline__ V3<real> zero3() { return V3<real>{real(0), real(0), real(0)}; }

// Field offsets are 32-bit element indices (3 x cells < 2^32 is checked on the host).  A cell's
// three components are one 12-byte (24-byte) access per lane.
template <typename real>
__device__ __forceinline__ V3<real> load3(const real* __restrict__ base, unsigned i) {
    return *reinterpret_cast<const V3<real>*>(base + i);
}

template <typename real>
__device__ __forceinline__ void store3(real* __restrict__ base, unsigned i, V3<real> v) {
    *reinterpret_cast<V3<real>*>(base + i) = v;
}

// E_new at (x, y, z) from backward differences of H (terms in CURL_TERMS order).
// h: H(x,y,z); hxm: H(x-1,y,z); hym: H(x,y-1,z); hx_zm, hy_zm: Hx, Hy at (x,y,z-1).
template <typename real>
__device__ __forceinline__ V3<real> e_cell(const March<real>& a, int x, int y, int z, bool write, const Hit& hx, const Hit& hy,
                                           const Hit& hz, const ZCoef<real>& zc, V3<real> h, V3<real> hxm, V3<real> hym,
                                           real hx_zm, real hy_zm, V3<real> e, V3<real> inv) {
    const int ny = a.ny, nz = a.nz;
    real cx = real(0), cy = real(0), cz = real(0);
    const unsigned px = ((unsigned)hx.q * ny + y) * nz + z;
    const unsigned py = ((unsigned)x * hy.w + hy.q) * nz + z;
    const unsigned pz = ((unsigned)x * ny + y) * hz.w + hz.q;
    {   // t0: d/dx Hz -> Ey, sign -1
        real d = real(0);
        if (x > 0) {
            d = h.z - hxm.z;
            if (hx.face >= 0) d = cpml(d, a.ef, hx, 0, px, write);
        }
        cy = cy - d;
    }
    {   // t1: d/dx Hy -> Ez, +1
        real d = real(0);
        if (x > 0) {
            d = h.y - hxm.y;
            if (hx.face >= 0) d = cpml(d, a.ef, hx, 1, px, write);
        }
        cz = cz + d;
    }
    {   // t2: d/dy Hz -> Ex, +1
        real d = real(0);
        if (y > 0) {
            d = h.z - hym.z;
            if (hy.face >= 0) d = cpml(d, a.ef, hy, 0, py, write);
        }
        cx = cx + d;
    }
    {   // t3: d/dy Hx -> Ez, -1
        real d = real(0);
        if (y > 0) {
            d = h.x - hym.x;
            if (hy.face >= 0) d = cpml(d, a.ef, hy, 1, py, write);
        }
        cz = cz - d;
    }
    {   // t4: d/dz Hy -> Ex, -1
        real d = real(0);
        if (z > 0) {
            d = h.y - hy_zm;
            if (hz.face >= 0) d = cpml_lane(d, a.ef, hz, zc, 0, pz, write);
        }
        cx = cx - d;
    }
    {   // t5: d/dz Hx -> Ey, +1
        real d = real(0);
        if (z > 0) {
            d = h.x - hx_zm;
            if (hz.face >= 0) d = cpml_lane(d, a.ef, hz, zc, 1, pz, write);
        }
        cy = cy + d;
    }
    return V3<real>{e.x + (a.courant * inv.x) * cx, e.y + (a.courant * inv.y) * cy, e.z + (a.courant * inv.z) * cz};
}

// H_new at (x, y, z) from forward differences of E_new (terms in CURL_TERMS order).
// e: E(x,y,z); exn: E(x+1,y,z); eyn: E(x,y+1,z); ex_zp, ey_zp: Ex, Ey at (x,y,z+1).
template <typename real>
__device__ __forceinline__ V3<real> h_cell(const March<real>& a, int x, int y, int z, bool write, real coefficient, const Hit& hx,
                                           const Hit& hy, const Hit& hz, const ZCoef<real>& zc, V3<real> e, V3<real> exn,
                                           V3<real> eyn, real ex_zp, real ey_zp, V3<real> h) {
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    real cx = real(0), cy = real(0), cz = real(0);
    const unsigned px = ((unsigned)hx.q * ny + y) * nz + z;
    const unsigned py = ((unsigned)x * hy.w + hy.q) * nz + z;
    const unsigned pz = ((unsigned)x * ny + y) * hz.w + hz.q;
    {   // t0: d/dx Ez -> Hy, -1
        real d = real(0);
        if (x < nx - 1) {
            d = exn.z - e.z;
            if (hx.face >= 0) d = cpml(d, a.hf, hx, 0, px, write);
        } else if (a.pec_mask & 1) {
            d = -a.pec_factor[0] * e.z;
        }
        cy = cy - d;
    }
    {   // t1: d/dx Ey -> Hz, +1
        real d = real(0);
        if (x < nx - 1) {
            d = exn.y - e.y;
            if (hx.face >= 0) d = cpml(d, a.hf, hx, 1, px, write);
        } else if (a.pec_mask & 1) {
            d = -a.pec_factor[0] * e.y;
        }
        cz = cz + d;
    }
    {   // t2: d/dy Ez -> Hx, +1
        real d = real(0);
        if (y < ny - 1) {
            d = eyn.z - e.z;
            if (hy.face >= 0) d = cpml(d, a.hf, hy, 0, py, write);
        } else if (a.pec_mask & 2) {
            d = -a.pec_factor[1] * e.z;
        }
        cx = cx + d;
    }
    {   // t3: d/dy Ex -> Hz, -1
        real d = real(0);
        if (y < ny - 1) {
            d = eyn.x - e.x;
            if (hy.face >= 0) d = cpml(d, a.hf, hy, 1, py, write);
        } else if (a.pec_mask & 2) {
            d = -a.pec_factor[1] * e.x;
        }
        cz = cz - d;
    }
    {   // t4: d/dz Ey -> Hx, -1
        real d = real(0);
        if (z < nz - 1) {
            d = ey_zp - e.y;
            if (hz.face >= 0) d = cpml_lane(d, a.hf, hz, zc, 0, pz, write);
        } else if (a.pec_mask & 4) {
            d = -a.pec_factor[2] * e.y;
        }
        cx = cx - d;
    }
    {   // t5: d/dz Ex -> Hy, +1
        real d = real(0);
        if (z < nz - 1) {
            d = ex_zp - e.x;
            if (hz.face >= 0) d = cpml_lane(d, a.hf, hz, zc, 1, pz, write);
        } else if (a.pec_mask & 4) {
            d = -a.pec_factor[2] * e.x;
        }
        cy = cy + d;
    }
    return V3<real>{h.x - coefficient * cx, h.y - coefficient * cy, h.z - coefficient * cz};
}

template <typename real, int TY, bool EPS3>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TY <= 4 ? 3 : TY <= 8 ? 2 : 1, 8))) void march_kernel(const March<real> a) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // XCD-aware order: workgroup w runs on XCD w % 8; give each XCD a contiguous range of tiles.
    const int G = gridDim.x, w = blockIdx.x, full = (G / 8) * 8;
    int t = w < full ? (w % 8) * (G / 8) + (w / 8) : w;
    const int group = t % a.ygroups;
    t /= a.ygroups;
    const int zw = t % a.zwin;
    const int chunk = t / a.zwin;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int y0 = (group * 4 + wave) * TY;
    if (y0 >= ny) return;
    const int zr = zw * 62 - 1 + lane;
    const bool zin = zr >= 0 && zr < nz;
    const bool zown = lane >= 1 && lane <= 62 && zin;
    // Lanes and rows outside the grid read (and compute on) clamped in-grid cells; their
    // results are never stored and never reach an owned cell, so every load of a plane is
    // unconditional and can be issued at once.
    const int z = zin ? zr : 0;
    const int xa = chunk * a.chunk;
    const int xb = min(nx, xa + a.chunk);
    const long long n = *a.counter;
    const real* __restrict__ e_in = a.e_in;
    const real* __restrict__ h_in = a.h_in;
    real* __restrict__ e_out = a.e_out;
    real* __restrict__ h_out = a.h_out;
    const real* __restrict__ inv_eps = a.inv_eps;
    const real hcoef = a.courant * a.inv_mu[0];
    const unsigned plane = (unsigned)ny * nz;
    const Hit ez_hit = zin ? find_hit(a.ef, 2, z) : Hit{-1, 0, 1, 0u, 0u, 0u};
    const Hit hz_hit = zin ? find_hit(a.hf, 2, z) : Hit{-1, 0, 1, 0u, 0u, 0u};
    const ZCoef<real> ez_coef = lane_coefficients(a.ef, ez_hit);
    const ZCoef<real> hz_coef = lane_coefficients(a.hf, hz_hit);
    int row[TY + 2];   // clamped rows y0-1 .. y0+TY
#pragma unroll
    for (int j = 0; j < TY + 2; ++j) row[j] = min(max(y0 - 1 + j, 0), ny - 1);

    V3<real> Ec[TY + 1];   // E_new(x-1), rows y0..y0+TY
    V3<real> Hc[TY + 1];   // H_old(x-1), rows y0..y0+TY
#pragma unroll
    for (int j = 0; j <= TY; ++j) {
        Ec[j] = zero3<real>();
        Hc[j] = zero3<real>();
    }
    if (xa > 0) {
        const unsigned base = (unsigned)(xa - 1) * plane + (unsigned)z;
#pragma unroll
        for (int j = 0; j <= TY; ++j) Hc[j] = load3(h_in, 3u * (base + (unsigned)row[j + 1] * nz));
    }
    for (int x = xa; x <= xb; ++x) {
        const bool has_e = x < nx;
        const bool has_h = x > xa;
        const int xl = has_e ? x : nx - 1;
        const unsigned base = (unsigned)xl * plane + (unsigned)z;
        const unsigned hbase = (unsigned)(x - 1) * plane + (unsigned)z;   // plane x-1 (used only when has_h)
        const Hit ex_hit = find_hit(a.ef, 0, x);
        const Hit hx_hit = find_hit(a.hf, 0, x - 1);
        V3<real> hp[TY + 2], eo[TY + 1], inv[TY + 1];
#pragma unroll
        for (int j = 0; j < TY + 2; ++j) hp[j] = load3(h_in, 3u * (base + (unsigned)row[j] * nz));
#pragma unroll
        for (int j = 0; j <= TY; ++j) {
            const unsigned cell = base + (unsigned)row[j + 1] * nz;
            eo[j] = load3(e_in, 3u * cell);
            if (EPS3) {
                inv[j] = load3(inv_eps, 3u * cell);
            } else {
                const real v = inv_eps[cell];
                inv[j] = V3<real>{v, v, v};
            }
        }
        V3<real> prev_e = zero3<real>();       // E_new(x, y-1): the x+1 neighbour of the H row y-1
        // Row j: E_new(x, y0+j); then H_new(x-1, y0+j-1), which needs E_new(x-1) of rows j-1 and j
        // (carried) and E_new(x, y0+j-1) (previous row); then rotate row j-1 into the carried planes.
#pragma unroll
        for (int j = 0; j <= TY + 1; ++j) {
            V3<real> en = zero3<real>();
            if (j <= TY && has_e) {
                const int y = y0 + j;
                const V3<real> h = hp[j + 1];
                const real hx_zm = shfl_up1(h.x);
                const real hy_zm = shfl_up1(h.y);
                const bool owned = zown && j < TY && y < ny && x < xb;
                const Hit ey_hit = find_hit(a.ef, 1, y);
                en = e_cell(a, x, row[j + 1], z, owned, ex_hit, ey_hit, ez_hit, ez_coef, h, Hc[j], hp[j], hx_zm, hy_zm,
                            eo[j], inv[j]);
                add_sources(a.se, a.nsrc_e, a.ebox_lo, a.ebox_hi, n, a.source_samples, x, y, zr, en.x, en.y, en.z);
                if (owned) store3(e_out, 3u * (base + (unsigned)y * nz), en);
            }
            if (j >= 1) {
                const int r = j - 1;
                if (has_h && r < TY) {
                    const int y = y0 + r;
                    const real ex_zp = shfl_down1(Ec[r].x);
                    const real ey_zp = shfl_down1(Ec[r].y);
                    const bool owned = zown && y < ny;
                    const Hit hy_hit = find_hit(a.hf, 1, y);
                    V3<real> out = h_cell(a, x - 1, row[r + 1], z, owned, hcoef, hx_hit, hy_hit, hz_hit, hz_coef, Ec[r], prev_e,
                                          Ec[r + 1], ex_zp, ey_zp, Hc[r]);
                    add_sources(a.sh, a.nsrc_h, a.hbox_lo, a.hbox_hi, n, a.source_samples, x - 1, y, zr, out.x, out.y, out.z);
                    if (owned) store3(h_out, 3u * (hbase + (unsigned)y * nz), out);
                }
                if (has_e) {
                    Ec[r] = prev_e;
                    Hc[r] = hp[r + 1];
                }
            }
            prev_e = en;
        }
    }
}

template <typename real>
int make_family(const tfdtd_march_family& in, bool split, const int dims[3], Family<real>* out) {
    std::memset(out, 0, sizeof *out);
    if (!in.psi_in || !in.psi_out || !in.coef) return tfdtd_fail(TFDTD_EINVAL, "march: null CPML buffer");
    if (split == (in.psi_in == in.psi_out))
        return tfdtd_fail(TFDTD_EINVAL, split ? "march: E psi needs separate input and output buffers" : "march: H psi updates in place");
    out->psi_in = static_cast<const real*>(in.psi_in);
    out->psi_out = static_cast<real*>(in.psi_out);
    out->coef = static_cast<const real*>(in.coef);
    for (int ax = 0; ax < 3; ++ax) {
        for (int f = 0; f < 2; ++f) {
            const tfdtd_march_face& face = in.face[ax][f];
            if (face.lo == face.hi) continue;
            if (face.lo < 0 || face.hi < face.lo || face.hi > dims[ax])
                return tfdtd_fail(TFDTD_EINVAL, "march: face %d of axis %d has rows outside the axis", f, ax);
            if (face.psi[0] < 0 || face.psi[1] < 0 || face.coef < 0 || face.psi[0] > 0xffffffffLL ||
                face.psi[1] > 0xffffffffLL || face.coef > 0xffffffffLL)
                return tfdtd_fail(TFDTD_EINVAL, "march: CPML offsets must be non-negative 32-bit element offsets");
            out->face[ax][f] = Face{face.lo, face.hi, {(unsigned)face.psi[0], (unsigned)face.psi[1]}, (unsigned)face.coef};
        }
        const Face& lo = out->face[ax][0];
        const Face& hi = out->face[ax][1];
        if (lo.hi > lo.lo && hi.hi > hi.lo && lo.hi > hi.lo && hi.hi > lo.lo)
            return tfdtd_fail(TFDTD_EINVAL, "march: the two faces of axis %d overlap", ax);
    }
    return TFDTD_OK;
}

template <typename real>
int make_march(const tfdtd_march* d, March<real>* a) {
    std::memset(a, 0, sizeof *a);
    if (d->nx < 2 || d->ny < 2 || d->nz < 2) return tfdtd_fail(TFDTD_EINVAL, "march: every axis needs at least two cells");
    if (!d->e_in || !d->h_in || !d->e_out || !d->h_out || !d->inverse_permittivity || !d->inverse_permeability)
        return tfdtd_fail(TFDTD_EINVAL, "march: null field or coefficient pointer");
    if (d->e_in == d->e_out || d->h_in == d->h_out) return tfdtd_fail(TFDTD_EINVAL, "march: input and output fields must differ");
    if (d->eps_stride != 1 && d->eps_stride != 3) return tfdtd_fail(TFDTD_EINVAL, "march: eps_stride must be 1 or 3");
    if (3.0 * (double)d->nx * d->ny * d->nz >= 4294967296.0) return tfdtd_fail(TFDTD_EINVAL, "march: grid exceeds 32-bit offsets");
    a->nx = d->nx; a->ny = d->ny; a->nz = d->nz;
    a->e_in = static_cast<const real*>(d->e_in); a->h_in = static_cast<const real*>(d->h_in);
    a->e_out = static_cast<real*>(d->e_out); a->h_out = static_cast<real*>(d->h_out);
    a->inv_eps = static_cast<const real*>(d->inverse_permittivity); a->eps_stride = d->eps_stride;
    a->inv_mu = static_cast<const real*>(d->inverse_permeability);
    a->courant = (real)d->courant;
    a->pec_mask = d->pec_upper_mask & 7;
    for (int k = 0; k < 3; ++k) a->pec_factor[k] = (real)d->pec_upper_factor[k];
    const int dims[3] = {d->nx, d->ny, d->nz};
    int status = make_family(d->e, true, dims, &a->ef);
    if (!status) status = make_family(d->h, false, dims, &a->hf);
    if (status) return status;
    for (int family = 0; family < 2; ++family) {
        const int count = family ? d->nsrc_h : d->nsrc_e;
        const tfdtd_source* srcs = family ? d->srcs_h : d->srcs_e;
        MSource<real>* table = family ? a->sh : a->se;
        int* lo = family ? a->hbox_lo : a->ebox_lo;
        int* hi = family ? a->hbox_hi : a->ebox_hi;
        if (count < 0 || count > kMarchSources) return tfdtd_fail(TFDTD_EINVAL, "march: at most %d sources per family", kMarchSources);
        for (int ax = 0; ax < 3; ++ax) { lo[ax] = dims[ax]; hi[ax] = 0; }
        for (int k = 0; k < count; ++k) {
            const tfdtd_source& s = srcs[k];
            if (s.comp < 0 || s.comp > 2 || !s.values) return tfdtd_fail(TFDTD_EINVAL, "march: source %d is malformed", k);
            table[k].comp = s.comp;
            table[k].values = static_cast<const real*>(s.values);
            for (int ax = 0; ax < 3; ++ax) {
                if (s.lo[ax] < 0 || s.hi[ax] <= s.lo[ax] || s.hi[ax] > dims[ax]) return tfdtd_fail(TFDTD_EINVAL, "march: source %d box outside the grid", k);
                table[k].lo[ax] = s.lo[ax]; table[k].hi[ax] = s.hi[ax];
                lo[ax] = std::min(lo[ax], s.lo[ax]); hi[ax] = std::max(hi[ax], s.hi[ax]);
            }
        }
        if (family) a->nsrc_h = count; else a->nsrc_e = count;
    }
    if (!d->counter) return tfdtd_fail(TFDTD_EINVAL, "march: the device step counter is required");
    a->counter = reinterpret_cast<const long long*>(d->counter);
    a->source_samples = d->source_samples;
    return TFDTD_OK;
}

template <typename real, int TY>
int launch_march(const tfdtd_march* d, hipStream_t stream) {
    March<real> a;
    int status = make_march(d, &a);
    if (status) return status;
    const int chunk = d->chunk_x > 0 ? d->chunk_x : d->nx;
    a.chunk = chunk;
    a.zwin = (d->nz + 61) / 62;
    const int ytiles = (d->ny + TY - 1) / TY;
    a.ygroups = (ytiles + 3) / 4;
    const long long chunks = (d->nx + chunk - 1) / chunk;
    const long long blocks = chunks * a.zwin * a.ygroups;
    if (blocks > 0x7fffffffLL) return tfdtd_fail(TFDTD_EINVAL, "march: too many workgroups");
    if (d->eps_stride == 3)
        hipLaunchKernelGGL((march_kernel<real, TY, true>), dim3((unsigned)blocks), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL((march_kernel<real, TY, false>), dim3((unsigned)blocks), dim3(256), 0, stream, a);
    return tfdtd_check_launch("march");
}

template <typename real>
int dispatch_march(const tfdtd_march* d, hipStream_t stream) {
    switch (d->tile_y) {
        case 4: return launch_march<real, 4>(d, stream);
        case 8: return launch_march<real, 8>(d, stream);
        case 16: return launch_march<real, 16>(d, stream);
        default: return tfdtd_fail(TFDTD_EINVAL, "march: tile_y must be 4, 8 or 16");
    }
}

}  // namespace

extern "C" int tfdtd_yee_step_march(const tfdtd_march* d, void* stream) {
    if (!d) return tfdtd_fail(TFDTD_EINVAL, "null march descriptor");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (d->dtype == TFDTD_F32) return dispatch_march<float>(d, s);
    if (d->dtype == TFDTD_F64) return dispatch_march<double>(d, s);
    return tfdtd_fail(TFDTD_EINVAL, "march: dtype must be float32 or float64");
}
