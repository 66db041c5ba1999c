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
// The two-step pass (tile2_body, further down) advances two steps per pass over HBM: each tile
// recomputes step n+1 on a widened region and keeps it on chip.
//
// Every product and sum rounds in the reference's order (torchfdtd/boundaries.py:363-438),
// so the result equals the split half-step kernels and the NumPy path bit for bit.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <type_traits>

#include "../../include/tfdtd.h"
#include "common.h"
#include "cpml_faces.h"

namespace {

#ifndef TFDTD_STAGE_REGISTERS
constexpr bool kStageDMA = true;     // H planes staged into LDS by direct loads, double-buffered
#else
constexpr bool kStageDMA = false;
#endif

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
    const unsigned char* material;     // optional per-cell index into material_table (eps_stride 1)
    const real* material_table;
    int materials;
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
    int chunk, ytiles, ztiles;
    int box_lo[2], box_hi[2];   // x chunks and y tiles [lo, hi) free of x and y CPML rows
    int zbox_lo, zbox_hi;       // z tiles [lo, hi) free of z CPML rows
    int zshift;                 // z tile tz starts at tz*63 - zshift
    int nchunks;
    int blocks[3];              // workgroups of the interior, z-face and boundary variants
    // Two-step layout only: x is cut at x0 and x1 (the planes whose widened chunks avoid the low and
    // high x slabs) into three parts of xparts[0..2] chunks each, every part split evenly.
    int x0, x1;
    int xparts[3];
    int yshift;                 // y tile ty starts at ty*9 - yshift
    int iy0, iy1, iytiles;      // tall interior tiles: rows [iy0, iy1) in iytiles tiles of 13 rows
    int nxend;                  // boundary tiles in the x end chunks (the rest lie in x-middle chunks)
    // Two-step pass: E and H of step n+1 in snap_count boxes [sbox_lo[k], sbox_hi[k]) are also stored
    // here, box k from snapshot cell sbox_base[k] on; [snap_lo, snap_hi) is the boxes' x hull.
    real* snap_e;
    real* snap_h;
    int snap_lo, snap_hi;
    int snap_count;
    int sbox_lo[4][3], sbox_hi[4][3];
    unsigned sbox_base[4];
};

// The snapshot cell that holds cell (x, y, z), or -1 (read from the descriptor: only planes inside the
// boxes' x hull ask).
template <typename real>
__device__ __forceinline__ int snapshot_cell(const March<real>& a, int x, int y, int z) {
    for (int k = 0; k < a.snap_count; ++k) {
        const int* lo = a.sbox_lo[k];
        const int* hi = a.sbox_hi[k];
        if (x >= lo[0] && x < hi[0] && y >= lo[1] && y < hi[1] && z >= lo[2] && z < hi[2])
            return (int)(a.sbox_base[k] + ((unsigned)(x - lo[0]) * (unsigned)(hi[1] - lo[1]) + (unsigned)(y - lo[1])) *
                                              (unsigned)(hi[2] - lo[2]) + (unsigned)(z - lo[2]));
    }
    return -1;
}

// psi <- psi*b + c*d (stored when `write`), d <- d*inv_k + psi, for slab row q of term k.
template <typename real>
__device__ __forceinline__ real cpml(real d, const Family<real>& F, const Hit& hit, int k, unsigned p, bool write) {
    const unsigned o = (k ? hit.psi1 : hit.psi0) + p;
    const auto coef = global(F.coef);
    const real b = coef[hit.coef + hit.q];
    const real c = coef[hit.coef + hit.w + hit.q];
    const real ik = coef[hit.coef + 2 * hit.w + hit.q];
    const real m = global(F.psi_in)[o] * b + c * d;
    if (write) global(F.psi_out)[o] = m;
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
    const real m = global(F.psi_in)[o] * zc.b + zc.c * d;
    if (write) global(F.psi_out)[o] = m;
    return d * zc.ik + m;
}

template <typename real>
__device__ __forceinline__ ZCoef<real> lane_coefficients(const Family<real>& F, const Hit& hit) {
    ZCoef<real> zc{real(0), real(0), real(1)};
    if (hit.face >= 0) {
        const auto coef = global(F.coef);
        zc.b = coef[hit.coef + hit.q];
        zc.c = coef[hit.coef + hit.w + hit.q];
        zc.ik = coef[hit.coef + 2 * hit.w + hit.q];
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
            const real v = global(s.values)[n];
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
// three components are one 12-byte (24-byte) access per lane, through a global pointer (the
// batched kernel reads its field pointers from a device table).
template <typename real>
__device__ __forceinline__ V3<real> load3(const real* __restrict__ base, unsigned i) {
    const auto p = global(base + i);
    return V3<real>{p[0], p[1], p[2]};
}

template <typename real>
__device__ __forceinline__ void store3(real* __restrict__ base, unsigned i, V3<real> v) {
    const auto p = global(base + i);
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

// E_new at (x, y, z) from backward differences of H (terms in CURL_TERMS order).
// h: H(x,y,z); hxm: H(x-1,y,z); hym: H(x,y-1,z); hx_zm, hy_zm: Hx, Hy at (x,y,z-1).
template <typename real, int CPML = 7>
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
            if ((CPML & 1) && hx.face >= 0) d = cpml(d, a.ef, hx, 0, px, write);
        }
        cy = cy - d;
    }
    {   // t1: d/dx Hy -> Ez, +1
        real d = real(0);
        if (x > 0) {
            d = h.y - hxm.y;
            if ((CPML & 1) && hx.face >= 0) d = cpml(d, a.ef, hx, 1, px, write);
        }
        cz = cz + d;
    }
    {   // t2: d/dy Hz -> Ex, +1
        real d = real(0);
        if (y > 0) {
            d = h.z - hym.z;
            if ((CPML & 2) && hy.face >= 0) d = cpml(d, a.ef, hy, 0, py, write);
        }
        cx = cx + d;
    }
    {   // t3: d/dy Hx -> Ez, -1
        real d = real(0);
        if (y > 0) {
            d = h.x - hym.x;
            if ((CPML & 2) && hy.face >= 0) d = cpml(d, a.ef, hy, 1, py, write);
        }
        cz = cz - d;
    }
    {   // t4: d/dz Hy -> Ex, -1
        real d = real(0);
        if (z > 0) {
            d = h.y - hy_zm;
            if ((CPML & 4) && hz.face >= 0) d = cpml_lane(d, a.ef, hz, zc, 0, pz, write);
        }
        cx = cx - d;
    }
    {   // t5: d/dz Hx -> Ey, +1
        real d = real(0);
        if (z > 0) {
            d = h.x - hx_zm;
            if ((CPML & 4) && hz.face >= 0) d = cpml_lane(d, a.ef, hz, zc, 1, pz, write);
        }
        cy = cy + d;
    }
    return V3<real>{e.x + (a.courant * inv.x) * cx, e.y + (a.courant * inv.y) * cy, e.z + (a.courant * inv.z) * cz};
}

// H_new at (x, y, z) from forward differences of E_new (terms in CURL_TERMS order).
// e: E(x,y,z); exn: E(x+1,y,z); eyn: E(x,y+1,z); ex_zp, ey_zp: Ex, Ey at (x,y,z+1).
template <typename real, int CPML = 7>
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
            if ((CPML & 1) && hx.face >= 0) d = cpml(d, a.hf, hx, 0, px, write);
        } else if (a.pec_mask & 1) {
            d = -a.pec_factor[0] * e.z;
        }
        cy = cy - d;
    }
    {   // t1: d/dx Ey -> Hz, +1
        real d = real(0);
        if (x < nx - 1) {
            d = exn.y - e.y;
            if ((CPML & 1) && hx.face >= 0) d = cpml(d, a.hf, hx, 1, px, write);
        } else if (a.pec_mask & 1) {
            d = -a.pec_factor[0] * e.y;
        }
        cz = cz + d;
    }
    {   // t2: d/dy Ez -> Hx, +1
        real d = real(0);
        if (y < ny - 1) {
            d = eyn.z - e.z;
            if ((CPML & 2) && hy.face >= 0) d = cpml(d, a.hf, hy, 0, py, write);
        } else if (a.pec_mask & 2) {
            d = -a.pec_factor[1] * e.z;
        }
        cx = cx + d;
    }
    {   // t3: d/dy Ex -> Hz, -1
        real d = real(0);
        if (y < ny - 1) {
            d = eyn.x - e.x;
            if ((CPML & 2) && hy.face >= 0) d = cpml(d, a.hf, hy, 1, py, write);
        } else if (a.pec_mask & 2) {
            d = -a.pec_factor[1] * e.x;
        }
        cz = cz - d;
    }
    {   // t4: d/dz Ey -> Hx, -1
        real d = real(0);
        if (z < nz - 1) {
            d = ey_zp - e.y;
            if ((CPML & 4) && hz.face >= 0) d = cpml_lane(d, a.hf, hz, zc, 0, pz, write);
        } else if (a.pec_mask & 4) {
            d = -a.pec_factor[2] * e.y;
        }
        cx = cx - d;
    }
    {   // t5: d/dz Ex -> Hy, +1
        real d = real(0);
        if (z < nz - 1) {
            d = ex_zp - e.x;
            if ((CPML & 4) && hz.face >= 0) d = cpml_lane(d, a.hf, hz, zc, 1, pz, write);
        } else if (a.pec_mask & 4) {
            d = -a.pec_factor[2] * e.x;
        }
        cy = cy + d;
    }
    return V3<real>{h.x - coefficient * cx, h.y - coefficient * cy, h.z - coefficient * cz};
}

// Tile (y tile, z tile, x chunk) of workgroup w among the G workgroups of variant CPML of a grid.
// CPML 0 and 4 enumerate the box of x chunks [box_lo[0], box_hi[0]) and y tiles [box_lo[1], box_hi[1])
// with the z tiles free of z slabs (0) or the others (4); each XCD (workgroup w runs on XCD w % 8) takes a
// contiguous range, so neighbouring tiles share L2.  CPML 7 enumerates the complement of the box as four
// blocks, round robin over XCDs.
template <typename real, int CPML>
__device__ __forceinline__ void tile_coords(const March<real>& a, const int w, const int G, int& ty, int& tz, int& chunk) {
    if (CPML != 7) {
        const int full = (G / 8) * 8;
        int t = w < full ? (w % 8) * (G / 8) + (w / 8) : w;
        const int my = a.box_hi[1] - a.box_lo[1];
        const int zlo = a.zbox_lo, zhi = a.zbox_hi;
        const int mz = CPML == 0 ? zhi - zlo : a.ztiles - (zhi - zlo);
        ty = a.box_lo[1] + t % my;
        t /= my;
        const int k = t % mz;
        tz = CPML == 0 ? zlo + k : (k < zlo ? k : zhi + (k - zlo));
        chunk = a.box_lo[0] + t / mz;
    } else {
        int b = w;
        const int per_chunk = a.ytiles * a.ztiles, inner = a.box_hi[0] - a.box_lo[0];
        const int nA = a.box_lo[0] * per_chunk, nB = (a.nchunks - a.box_hi[0]) * per_chunk;
        const int nC = inner * a.box_lo[1] * a.ztiles;
        if (b < nA + nB) {
            chunk = b < nA ? b / per_chunk : a.box_hi[0] + (b - nA) / per_chunk;
            const int r = b < nA ? b % per_chunk : (b - nA) % per_chunk;
            ty = r % a.ytiles;
            tz = r / a.ytiles;
        } else if ((b -= nA + nB) < nC) {
            ty = b % a.box_lo[1];
            b /= a.box_lo[1];
            tz = b % a.ztiles;
            chunk = a.box_lo[0] + b / a.ztiles;
        } else {
            b -= nC;
            const int high = a.ytiles - a.box_hi[1];
            ty = a.box_hi[1] + b % high;
            b /= high;
            tz = b % a.ztiles;
            chunk = a.box_lo[0] + b / a.ztiles;
        }
    }
}

// One workgroup (W = 4 or 8 wavefronts) marches a tile of TY x 63 (y, z) cells, TY = 2W, over a chunk
// of x planes.  Wavefront w owns rows y0 + 2w, y0 + 2w + 1 (a wave-uniform y); lane l covers column
// z0 + l, lanes 0..62 own their cells and lane 63 recomputes E of the next tile's first column.
// Per plane x:
//   A. every lane loads E(x) and 1/eps of its cells into registers; H(x) on rows y0-1 .. y0+TY and
//      columns z0-1 .. z0+63 is already on its way into an LDS buffer (sH0 / sH1, alternating);
//   B. barrier (H(x) staged);  then the direct loads of H(x+1) into the other buffer start;
//   C. E(x) of the tile's rows (and, in wavefront 0, of the halo row y0+TY) from the staged H(x) and
//      the carried H(x-1);
//   D. H(x-1) of the owned cells from E(x-1) in LDS (sE), E(x) in registers and the carried H(x-1);
//   E. barrier (LDS reads done; the staging of H(x+1) continues);  F. E(x) into sE, H(x) becomes the
//      carried H(x-1).
// Cells outside the grid read clamped in-grid addresses and are never computed or stored.  z tiles
// start at tz*63 - zshift, chosen so a tile edge meets the end of the low z slab.
// CPML: bit 0 x, bit 1 y, bit 2 z slabs compiled in.  Variant 0 runs the tiles free of every slab,
// variant 4 the tiles that meet only z slabs, variant 7 the rest.
// w is the workgroup's index among the G workgroups of its variant and grid.
template <typename real, int W, bool EPS3, int CPML>
__device__ __forceinline__ void tile_body(const March<real>& a, const int w, const int G) {
    constexpr int RW = 2;                      // rows per wavefront
    constexpr int TY = W * RW, TZ = 63;
    constexpr int HR = TY + 2, HC = TZ + 2;    // sH: rows y0-1 .. y0+TY, columns z0-1 .. z0+63
    constexpr int EC = TZ + 1;                 // sE: rows y0 .. y0+TY, columns z0 .. z0+63
    constexpr int NC = RW + 1;                 // row slots per lane: RW owned + the halo row (wavefront 0)
    constexpr int SH = HR * HC * 3;            // reals of one staged H plane
    // Two staging buffers as two LDS objects: the compiler then knows that reading one does not wait
    // for the direct loads into the other.
    __shared__ real sH0[SH];
    __shared__ real sH1[kStageDMA ? SH : 1];
    __shared__ real sE[(TY + 1) * EC * 3];
    // The interior tiles (bandwidth-bound) read 1/eps through the grid's material index when it has
    // one; the z-face and boundary tiles (latency-bound) read the plain array, which the index
    // mirrors -- the dependent table lookup costs them more than the bytes it saves.
    constexpr bool INDEXED = !EPS3 && CPML == 0;
    __shared__ real sMat[INDEXED ? 256 : 1];   // 1/eps of each material
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    int ty, tz, chunk;
    tile_coords<real, CPML>(a, w, G, ty, tz, chunk);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int y0 = ty * TY, z0 = tz * TZ - a.zshift;
    const int xa = chunk * a.chunk, xb = min(nx, xa + a.chunk);
    const long long n = *global(a.counter);
    const real* __restrict__ e_in = a.e_in;
    const real* __restrict__ h_in = a.h_in;
    real* __restrict__ e_out = a.e_out;
    real* __restrict__ h_out = a.h_out;
    const real* __restrict__ inv_eps = a.inv_eps;
    const unsigned char* __restrict__ material = INDEXED ? a.material : nullptr;
    if (INDEXED) {
        if (material)
            for (int i = tid; i < a.materials; i += 64 * W) sMat[i] = global(a.material_table)[i];
        __syncthreads();   // the first plane's loads read the table
    }
    const real hcoef = a.courant * global(a.inv_mu)[0];
    const unsigned plane = (unsigned)ny * nz;

    const int z = z0 + lane;
    const bool zin = z >= 0 && z < nz;
    const bool zown = zin && lane < TZ;
    const unsigned zc = (unsigned)min(max(z, 0), nz - 1);
    // Row slots: k < RW the wavefront's rows; k == RW the halo row y0+TY, computed by wavefront 0 only.
    auto row = [&](int k) { return k < RW ? y0 + wave * RW + k : y0 + TY; };
    auto computed = [&](int k) { return zin && row(k) < ny && (k < RW || wave == 0); };
    // The lane's z-slab membership and coefficients, once for the whole march (z is fixed per lane).
    // Read inside the loop, the descriptor fields would be reloaded every row: the loop's stores may
    // alias the descriptor table as far as the compiler knows.
    const Hit none{-1, 0, 1, 0u, 0u, 0u};
    const Hit ez_hit = (CPML & 4) && zin ? find_hit(a.ef, 2, z) : none;
    const Hit hz_hit = (CPML & 4) && zin ? find_hit(a.hf, 2, z) : none;
    const ZCoef<real> ez_coef = lane_coefficients(a.ef, ez_hit);
    const ZCoef<real> hz_coef = lane_coefficients(a.hf, hz_hit);
    // Element tid of an sH row: component tid % 3 of column z0 - 1 + tid / 3 (clamped into the grid).
    const unsigned row_offset = 3u * (unsigned)min(max(z0 - 1 + tid / 3, 0), nz - 1) + (unsigned)(tid % 3);
    // Staging of H(x) rows y0-1 .. y0+TY straight into LDS buffer x & 1 (no registers), in 4-byte words:
    // word e of the flattened plane belongs to element e / WPE, which is row el / (3 HC) and column
    // element el % (3 HC), clamped into the grid like the register path.  Wavefront w issues segments
    // w, w + 4, ... of 64 consecutive words; a lane whose word lies past the plane stays inactive.
    constexpr int WPE = (int)(sizeof(real) / 4), SHW = SH * WPE;
    auto stage = [&](int x, real* into) {
        unsigned* buffer = reinterpret_cast<unsigned*>(into);
        const unsigned base = (unsigned)x * plane;
        for (int seg = wave; seg * 64 < SHW; seg += W) {
            const int e = seg * 64 + lane;
            if (e < SHW) {
                const int el = e / WPE, part = e - el * WPE;
                const int r = el / (3 * HC), c = el - r * (3 * HC);
                const int y = min(max(y0 - 1 + r, 0), ny - 1);
                const unsigned off = 3u * (unsigned)min(max(z0 - 1 + c / 3, 0), nz - 1) + (unsigned)(c % 3);
                const unsigned* word = reinterpret_cast<const unsigned*>(h_in + 3u * (base + (unsigned)y * nz) + off) + part;
                __builtin_amdgcn_global_load_lds(global(word), (__attribute__((address_space(3))) void*)(buffer + seg * 64),
                                                 4, 0, 0);
            }
        }
    };

    V3<real> Hc[NC];   // H(x-1) of each slot
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        Hc[k] = zero3<real>();
        if (xa > 0) Hc[k] = load3(h_in, 3u * ((unsigned)(xa - 1) * plane + (unsigned)min(row(k), ny - 1) * nz + zc));
    }
    // One plane x: sHx holds H(x) (staged by the previous plane, or before the march), sHn receives H(x+1).
    auto march_plane = [&](const int x, real* const sHx, real* const sHn) {
        const bool has_e = x < nx;
        const bool has_h = x > xa;
        V3<real> eo[NC], inv[NC];
        if (has_e) {
            const unsigned base = (unsigned)x * plane;
            // Row r of sH is 3*HC contiguous reals in memory (clamped at the grid edges): lane t < 3*HC
            // copies element t of every row.
            if (!kStageDMA && tid < 3 * HC) {
#pragma unroll
                for (int r = 0; r < HR; ++r) {
                    const int y = min(max(y0 - 1 + r, 0), ny - 1);
                    sHx[r * 3 * HC + tid] = global(h_in)[3u * (base + (unsigned)y * nz) + row_offset];
                }
            }
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                const unsigned cell = base + (unsigned)min(row(k), ny - 1) * nz + zc;
                eo[k] = load3(e_in, 3u * cell);
                if (EPS3) {
                    inv[k] = load3(inv_eps, 3u * cell);
                } else {
                    const real v = material ? sMat[global(material)[cell]] : global(inv_eps)[cell];
                    inv[k] = V3<real>{v, v, v};
                }
            }
        }
        __syncthreads();   // also completes the staging of H(x) into sHx
        // The next plane's H rows go to the other buffer while this plane computes; it was last read
        // before the barrier above.
        if (kStageDMA && x + 1 <= xb && x + 1 < nx) stage(x + 1, sHn);
        V3<real> en[NC], hn[NC];
        const Hit ex_hit = find_hit(a.ef, 0, x);
        const Hit hx_hit = find_hit(a.hf, 0, x - 1);
        if (has_e) {
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                en[k] = zero3<real>();
                const int y = row(k);
                const int r = y - y0 + 1;
                const real* p = sHx + 3 * (r * HC + lane + 1);
                hn[k] = V3<real>{p[0], p[1], p[2]};
                if (computed(k)) {
                    const real* pym = p - 3 * HC;
                    const real* pzm = p - 3;
                    const bool owned = zown && k < RW && x < xb;
                    const Hit ey_hit = find_hit(a.ef, 1, y);
                    en[k] = e_cell<real, CPML>(a, x, y, z, owned, ex_hit, ey_hit, ez_hit, ez_coef, hn[k], Hc[k],
                                               V3<real>{pym[0], pym[1], pym[2]}, pzm[0], pzm[1], eo[k], inv[k]);
                    add_sources(a.se, a.nsrc_e, a.ebox_lo, a.ebox_hi, n, a.source_samples, x, y, z, en[k].x, en[k].y, en[k].z);
                    if (owned) store3(e_out, 3u * ((unsigned)x * plane + (unsigned)y * nz + (unsigned)z), en[k]);
                }
            }
        }
        if (has_h) {
            const int xh = x - 1;
#pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int y = row(k);
                if (zown && y < ny) {
                    const real* p = sE + 3 * ((y - y0) * EC + lane);
                    const V3<real> e{p[0], p[1], p[2]};
                    const real* pyn = p + 3 * EC;
                    const real* pzn = p + 3;
                    const Hit hy_hit = find_hit(a.hf, 1, y);
                    V3<real> out = h_cell<real, CPML>(a, xh, y, z, true, hcoef, hx_hit, hy_hit, hz_hit, hz_coef, e,
                                                      has_e ? en[k] : e, V3<real>{pyn[0], pyn[1], pyn[2]}, pzn[0], pzn[1], Hc[k]);
                    add_sources(a.sh, a.nsrc_h, a.hbox_lo, a.hbox_hi, n, a.source_samples, xh, y, z, out.x, out.y, out.z);
                    store3(h_out, 3u * ((unsigned)xh * plane + (unsigned)y * nz + (unsigned)z), out);
                }
            }
        }
        if (kStageDMA) {
            // Every wavefront's reads of sE (and of sHx) are complete before sE is overwritten.  A plain
            // barrier: a __syncthreads here would also wait for the staging of the next plane.
            // The clobbers keep the compiler from moving LDS accesses across the barrier (the barrier
            // builtin itself is not a memory operation); lgkmcnt(0): this wavefront's LDS reads have returned.
            asm volatile("" ::: "memory");
            __builtin_amdgcn_s_waitcnt(0xc07f);
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_sched_barrier(0);
            asm volatile("" ::: "memory");
        } else {
            __syncthreads();
        }
        if (has_e) {
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                if (computed(k)) {
                    real* p = sE + 3 * ((row(k) - y0) * EC + lane);
                    p[0] = en[k].x; p[1] = en[k].y; p[2] = en[k].z;
                }
                Hc[k] = hn[k];
            }
        }
    };
    if constexpr (kStageDMA) {
        if (xa < nx) stage(xa, sH0);
        for (int x = xa; x <= xb; x += 2) {
            march_plane(x, sH0, sH1);
            if (x + 1 <= xb) march_plane(x + 1, sH1, sH0);
        }
    } else {
        for (int x = xa; x <= xb; ++x) march_plane(x, sH0, sH0);
    }
}

// ---------------------------------------------------------------------------------------------------
// Two steps per pass: time n -> n+2 with the intermediate step n+1 kept on chip.
//
// CPML memory of the two-step march: stage 1 (step n+1) reads psi at time n from memory and keeps psi
// at n+1 in a register slot of its cell; stage 2 (step n+2) continues from that slot and stores psi at
// n+2 (owned cells only).  Both families read psi_in and write psi_out, which must differ.
template <typename real, int STAGE>
__device__ __forceinline__ real cpml2(real d, const Family<real>& F, const Hit& hit, int k, unsigned p, bool write, real& slot) {
    const unsigned o = (k ? hit.psi1 : hit.psi0) + p;
    const auto coef = global(F.coef);
    const real b = coef[hit.coef + hit.q];
    const real c = coef[hit.coef + hit.w + hit.q];
    const real ik = coef[hit.coef + 2 * hit.w + hit.q];
    // STAGE 3: step n+1 with psi at time n already in the slot (loaded a plane ahead).
    const real m = (STAGE == 1 ? global(F.psi_in)[o] : slot) * b + c * d;
    if (STAGE == 1 || STAGE == 3) slot = m;
    else if (write) global(F.psi_out)[o] = m;
    return d * ik + m;
}

template <typename real, int STAGE>
__device__ __forceinline__ real cpml_lane2(real d, const Family<real>& F, const Hit& hit, const ZCoef<real>& zc, int k,
                                           unsigned p, bool write, real& slot) {
    const unsigned o = (k ? hit.psi1 : hit.psi0) + p;
    // STAGE 3: step n+1 with psi at time n already in the slot (loaded a plane ahead).
    const real m = (STAGE == 1 ? global(F.psi_in)[o] : slot) * zc.b + zc.c * d;
    if (STAGE == 1 || STAGE == 3) slot = m;
    else if (write) global(F.psi_out)[o] = m;
    return d * zc.ik + m;
}

// Loop-invariant scalars of the two-step pass, copied out of the descriptor once: read through the
// descriptor, the compiler reloads them after every store (the stores might alias the table).
template <typename real>
struct Consts2 {
    int nx, ny, nz, pec_mask;
    real courant, hcoef;
    real pec[3];
    int nse, nsh;
    long long samples;
    int ebl[3], ebh[3], hbl[3], hbh[3];
    int snap_lo, snap_hi;
};

// e_cell with the CPML memory of term t in slot[t].  ZPRE / XPRE: at stage 1 the z / x terms find
// their memory at time n already in the slot.
template <typename real, int CPML, int STAGE, bool ZPRE = false, bool XPRE = false>
__device__ __forceinline__ V3<real> e_cell2(const March<real>& a, const Consts2<real>& q, int x, int y, int z, bool write,
                                            const Hit& hx, const Hit& hy, const Hit& hz, const ZCoef<real>& zc, V3<real> h,
                                            V3<real> hxm, V3<real> hym, real hx_zm, real hy_zm, V3<real> e, V3<real> inv,
                                            real (&slot)[6]) {
    const int ny = q.ny, nz = q.nz;
    real cx = real(0), cy = real(0), cz = real(0);
    const unsigned px = ((unsigned)hx.q * ny + y) * nz + z;
    const unsigned py = ((unsigned)x * hy.w + hy.q) * nz + z;
    const unsigned pz = ((unsigned)x * ny + y) * hz.w + hz.q;
    {   // t0: d/dx Hz -> Ey, -1
        real d = real(0);
        if (x > 0) {
            d = h.z - hxm.z;
            if ((CPML & 1) && hx.face >= 0) d = cpml2<real, (XPRE && STAGE == 1) ? 3 : STAGE>(d, a.ef, hx, 0, px, write, slot[0]);
        }
        cy = cy - d;
    }
    {   // t1: d/dx Hy -> Ez, +1
        real d = real(0);
        if (x > 0) {
            d = h.y - hxm.y;
            if ((CPML & 1) && hx.face >= 0) d = cpml2<real, (XPRE && STAGE == 1) ? 3 : STAGE>(d, a.ef, hx, 1, px, write, slot[1]);
        }
        cz = cz + d;
    }
    {   // t2: d/dy Hz -> Ex, +1
        real d = real(0);
        if (y > 0) {
            d = h.z - hym.z;
            if ((CPML & 2) && hy.face >= 0) d = cpml2<real, STAGE>(d, a.ef, hy, 0, py, write, slot[2]);
        }
        cx = cx + d;
    }
    {   // t3: d/dy Hx -> Ez, -1
        real d = real(0);
        if (y > 0) {
            d = h.x - hym.x;
            if ((CPML & 2) && hy.face >= 0) d = cpml2<real, STAGE>(d, a.ef, hy, 1, py, write, slot[3]);
        }
        cz = cz - d;
    }
    {   // t4: d/dz Hy -> Ex, -1
        real d = real(0);
        if (z > 0) {
            d = h.y - hy_zm;
            if ((CPML & 4) && hz.face >= 0) d = cpml_lane2<real, (ZPRE && STAGE == 1) ? 3 : STAGE>(d, a.ef, hz, zc, 0, pz, write, slot[4]);
        }
        cx = cx - d;
    }
    {   // t5: d/dz Hx -> Ey, +1
        real d = real(0);
        if (z > 0) {
            d = h.x - hx_zm;
            if ((CPML & 4) && hz.face >= 0) d = cpml_lane2<real, (ZPRE && STAGE == 1) ? 3 : STAGE>(d, a.ef, hz, zc, 1, pz, write, slot[5]);
        }
        cy = cy + d;
    }
    return V3<real>{e.x + (q.courant * inv.x) * cx, e.y + (q.courant * inv.y) * cy, e.z + (q.courant * inv.z) * cz};
}

// h_cell with the CPML memory of term t in slot[t].
template <typename real, int CPML, int STAGE, bool ZPRE = false, bool XPRE = false>
__device__ __forceinline__ V3<real> h_cell2(const March<real>& a, const Consts2<real>& q, int x, int y, int z, bool write,
                                            real coefficient, const Hit& hx, const Hit& hy, const Hit& hz, const ZCoef<real>& zc,
                                            V3<real> e, V3<real> exn, V3<real> eyn, real ex_zp, real ey_zp, V3<real> h,
                                            real (&slot)[6]) {
    const int nx = q.nx, ny = q.ny, nz = q.nz;
    real cx = real(0), cy = real(0), cz = real(0);
    const unsigned px = ((unsigned)hx.q * ny + y) * nz + z;
    const unsigned py = ((unsigned)x * hy.w + hy.q) * nz + z;
    const unsigned pz = ((unsigned)x * ny + y) * hz.w + hz.q;
    {   // t0: d/dx Ez -> Hy, -1
        real d = real(0);
        if (x < nx - 1) {
            d = exn.z - e.z;
            if ((CPML & 1) && hx.face >= 0) d = cpml2<real, (XPRE && STAGE == 1) ? 3 : STAGE>(d, a.hf, hx, 0, px, write, slot[0]);
        } else if (q.pec_mask & 1) {
            d = -q.pec[0] * e.z;
        }
        cy = cy - d;
    }
    {   // t1: d/dx Ey -> Hz, +1
        real d = real(0);
        if (x < nx - 1) {
            d = exn.y - e.y;
            if ((CPML & 1) && hx.face >= 0) d = cpml2<real, (XPRE && STAGE == 1) ? 3 : STAGE>(d, a.hf, hx, 1, px, write, slot[1]);
        } else if (q.pec_mask & 1) {
            d = -q.pec[0] * e.y;
        }
        cz = cz + d;
    }
    {   // t2: d/dy Ez -> Hx, +1
        real d = real(0);
        if (y < ny - 1) {
            d = eyn.z - e.z;
            if ((CPML & 2) && hy.face >= 0) d = cpml2<real, STAGE>(d, a.hf, hy, 0, py, write, slot[2]);
        } else if (q.pec_mask & 2) {
            d = -q.pec[1] * e.z;
        }
        cx = cx + d;
    }
    {   // t3: d/dy Ex -> Hz, -1
        real d = real(0);
        if (y < ny - 1) {
            d = eyn.x - e.x;
            if ((CPML & 2) && hy.face >= 0) d = cpml2<real, STAGE>(d, a.hf, hy, 1, py, write, slot[3]);
        } else if (q.pec_mask & 2) {
            d = -q.pec[1] * e.x;
        }
        cz = cz - d;
    }
    {   // t4: d/dz Ey -> Hx, -1
        real d = real(0);
        if (z < nz - 1) {
            d = ey_zp - e.y;
            if ((CPML & 4) && hz.face >= 0) d = cpml_lane2<real, (ZPRE && STAGE == 1) ? 3 : STAGE>(d, a.hf, hz, zc, 0, pz, write, slot[4]);
        } else if (q.pec_mask & 4) {
            d = -q.pec[2] * e.y;
        }
        cx = cx - d;
    }
    {   // t5: d/dz Ex -> Hy, +1
        real d = real(0);
        if (z < nz - 1) {
            d = ex_zp - e.x;
            if ((CPML & 4) && hz.face >= 0) d = cpml_lane2<real, (ZPRE && STAGE == 1) ? 3 : STAGE>(d, a.hf, hz, zc, 1, pz, write, slot[5]);
        } else if (q.pec_mask & 4) {
            d = -q.pec[2] * e.x;
        }
        cy = cy + d;
    }
    return V3<real>{h.x - coefficient * cx, h.y - coefficient * cy, h.z - coefficient * cz};
}

// Planes [xa, xb) of chunk c in the two-step layout: part 0 = [0, x0), part 1 = [x0, x1), part 2 =
// [x1, nx), each split into its xparts[] chunks of near-equal length.
template <typename real>
__host__ __device__ __forceinline__ void chunk_range2(const March<real>& a, int c, int& xa, int& xb) {
    const int n0 = a.xparts[0], n1 = a.xparts[1];
    int lo = 0, hi = a.x0, i = c, n = n0;
    if (c >= n0 + n1) { lo = a.x1; hi = a.nx; i = c - n0 - n1; n = a.xparts[2]; }
    else if (c >= n0) { lo = a.x0; hi = a.x1; i = c - n0; n = n1; }
    const long long len = hi - lo;
    xa = lo + (int)(i * len / n);
    xb = lo + (int)((i + 1) * len / n);
}

constexpr int kTZ2 = 61;   // owned z columns of a two-step tile
constexpr int kTY2 = 9;    // owned y rows of a two-step tile

// One workgroup of W = 4, 6 or 12 wavefronts advances a tile of 9 rows x 61 columns (y, z) over a chunk
// of x planes by two steps.  Lane l covers column z0 - 1 + l; wavefront w covers the widened rows
// y0 - 1 + RW w + k (k < RW = 12 / W).  Step n+1 is computed on the widened tile, step n+2 on the owned cells:
//   stage 1   E1 = E(n+1) on rows y0-1 .. y0+10, lanes 0..63       (needs H(n) rows y0-2 .., staged)
//   stage 1b  H1 = H(n+1) on rows y0-1 .. y0+9,  lanes 0..62       (needs E1 of the row and column above)
//   stage 2   E2 = E(n+2) on rows y0 .. y0+9,    lanes 1..62       (needs H1 of the row and column below)
//   stage 2b  H2 = H(n+2) on rows y0 .. y0+8,    lanes 1..61 = the owned cells
// Along x the stages trail each other by a plane: iteration x computes E1(x), H1(x-1), E2(x-1) and
// H2(x-2), so a chunk [xa, xb) runs x = xa-1 .. xb+1.  H(n) planes are staged in LDS by direct loads one
// plane ahead; E1(x-1), H1(x-1) and E2(x-2) are exchanged between wavefronts through LDS, and every
// value a cell needs from its own previous plane (H(n), E1, H1, E2, 1/eps, CPML memory) is carried in
// registers.  Cells outside the grid are never computed or stored; their addresses are clamped.
// SNAP: the pass also stores step n+1 in the descriptor's snapshot boxes (compiled apart: the
// store code, even unused, slows the other kernels' loops).
template <typename real, bool EPS3, int CPML, int W, int NR = 12, bool SNAP = false>
__device__ __forceinline__ void tile2_body(const March<real>& a, const int w, const int G) {
    constexpr int RW = NR / W;   // NR widened rows (owned NR - 3), RW per wavefront
    static_assert(RW * W == NR, "the wavefronts split the widened rows evenly");
    static_assert(NR == 12 || CPML == 0, "tall tiles are interior tiles");
    // CPML 6 (y and z slabs) takes the boundary tiles of the x-middle chunks, CPML 7 and 1 (x slabs
    // only) those of the x ends: all enumerate the boundary tiles in one order.
    constexpr int SR = NR + 1, SC = kTZ2 + 4, SH = SR * SC * 3;   // staged H(n): rows y0-2 .. y0+10, columns z0-2 .. z0+62
    constexpr int PL = NR * 64 * 3 + 64 * 3 + 3;                 // one widened plane, padded past its last row
    __shared__ real sA[SH];
    __shared__ real sB[SH];
    __shared__ real sE1[PL];   // E1(x-1)
    __shared__ real sH1[PL];   // H1(x-1)
    __shared__ real sE2[PL];   // E2(x-2)
    constexpr bool INDEXED = !EPS3 && CPML == 0;
    __shared__ real sMat[INDEXED ? 256 : 1];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    int ty, tz, chunk, y0, ylim = a.ny;
    if constexpr (NR != 12) {
        // Tall interior tiles of NR - 3 rows cover the slab-free rows [iy0, iy1); the last one is cut at
        // iy1 (its rows past iy1 + 1 are computed without CPML and feed no owned cell).
        const int full = (G / 8) * 8;
        int t = w < full ? (w % 8) * (G / 8) + (w / 8) : w;
        const int my = a.iytiles, mz = a.zbox_hi - a.zbox_lo;
        ty = t % my;
        t /= my;
        tz = a.zbox_lo + t % mz;
        chunk = a.box_lo[0] + t / mz;
        y0 = a.iy0 + ty * (NR - 3);
        ylim = a.iy1;
    } else {
        tile_coords<real, (CPML == 6 || CPML == 1) ? 7 : CPML>(a, w, G, ty, tz, chunk);
        y0 = ty * kTY2 - a.yshift;
    }
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int z0 = tz * kTZ2 - a.zshift;
    int xa, xb;
    chunk_range2(a, chunk, xa, xb);
    const long long n = *global(a.counter);
    const real* __restrict__ e_in = a.e_in;
    const real* __restrict__ h_in = a.h_in;
    real* __restrict__ e_out = a.e_out;
    real* __restrict__ h_out = a.h_out;
    const real* __restrict__ inv_eps = a.inv_eps;
    const unsigned char* __restrict__ material = INDEXED ? a.material : nullptr;
    if (INDEXED) {
        if (material)
            for (int i = tid; i < a.materials; i += 64 * W) sMat[i] = global(a.material_table)[i];
        __syncthreads();   // the first plane's loads read the table
    }
    const real hcoef = a.courant * global(a.inv_mu)[0];
    const unsigned plane = (unsigned)ny * nz;
    Consts2<real> q;
    q.nx = nx; q.ny = ny; q.nz = nz; q.pec_mask = a.pec_mask;
    q.courant = a.courant; q.hcoef = hcoef;
    q.nse = a.nsrc_e; q.nsh = a.nsrc_h; q.samples = a.source_samples;
    q.snap_lo = a.snap_lo; q.snap_hi = a.snap_hi;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        q.pec[ax] = a.pec_factor[ax];
        q.ebl[ax] = a.ebox_lo[ax]; q.ebh[ax] = a.ebox_hi[ax];
        q.hbl[ax] = a.hbox_lo[ax]; q.hbh[ax] = a.hbox_hi[ax];
    }
    // Interior tiles compute every lane and discard the lanes outside the grid or the stage's columns
    // afterwards (no CPML memory to address, every LDS read is in bounds): per-lane branches cost more
    // scalar instructions than the discarded lanes cost vector ones.
    constexpr bool ALL_LANES = CPML == 0;
    auto sel3 = [](bool keep, V3<real> v) { return keep ? v : zero3<real>(); };

    const int z = z0 - 1 + lane;
    const bool zin = z >= 0 && z < nz;
    const unsigned zc = (unsigned)min(max(z, 0), nz - 1);
    auto rowof = [&](int k) { return wave * RW + k; };
    auto yof = [&](int k) { return y0 - 1 + wave * RW + k; };
    const Hit none{-1, 0, 1, 0u, 0u, 0u};
    const Hit ez_hit = (CPML & 4) && zin ? find_hit(a.ef, 2, z) : none;
    const Hit hz_hit = (CPML & 4) && zin ? find_hit(a.hf, 2, z) : none;
    const ZCoef<real> ez_coef = lane_coefficients(a.ef, ez_hit);
    const ZCoef<real> hz_coef = lane_coefficients(a.hf, hz_hit);

    // Staging of H(x) rows y0-2 .. y0+10, columns z0-2 .. z0+62 straight into LDS, in 4-byte words
    // (clamped into the grid).  Wavefront w issues segments w, w + 4, ... of 64 consecutive words.
    auto stage = [&](int x, real* into) {
        unsigned* buffer = reinterpret_cast<unsigned*>(into);
        const unsigned base = (unsigned)x * plane;
        for (int seg = wave; seg * 64 < SH; seg += W) {
            const int e = seg * 64 + lane;
            if (e < SH) {
                const int r = e / (3 * SC), c = e - r * (3 * SC);
                const int y = min(max(y0 - 2 + r, 0), ny - 1);
                const unsigned off = 3u * (unsigned)min(max(z0 - 2 + c / 3, 0), nz - 1) + (unsigned)(c % 3);
                const unsigned* word = reinterpret_cast<const unsigned*>(h_in + 3u * (base + (unsigned)y * nz) + off);
                __builtin_amdgcn_global_load_lds(global(word), (__attribute__((address_space(3))) void*)(buffer + seg * 64),
                                                 4, 0, 0);
            }
        }
    };
    auto lds3 = [](const real* p) { return V3<real>{p[0], p[1], p[2]}; };
    auto put3 = [](real* p, V3<real> v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; };
    // A plain barrier after this wavefront's LDS accesses have completed: it does not wait for the
    // direct loads of the next H plane.  The clobbers keep LDS accesses on their side of it.
    auto lds_barrier = []() {
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("" ::: "memory");
    };

    const int xs = max(xa - 1, 0), xe = xb + 1;
    V3<real> Hn_c[RW], H1_c[RW], E1_c[RW], E2_c[RW], inv_c[RW];
    real pe0[RW][6], pe1[RW][6], ph0[RW][6], ph1[RW][6];   // CPML memory at n+1: E1(x-1) / E1(x), H1(x-2) / H1(x-1)
#pragma unroll
    for (int k = 0; k < RW; ++k) {
        Hn_c[k] = zero3<real>(); H1_c[k] = zero3<real>(); E1_c[k] = zero3<real>(); E2_c[k] = zero3<real>();
        inv_c[k] = zero3<real>();
#pragma unroll
        for (int t = 0; t < 6; ++t) pe0[k][t] = pe1[k][t] = ph0[k][t] = ph1[k][t] = real(0);
        if (xs > 0) Hn_c[k] = load3(h_in, 3u * ((unsigned)(xs - 1) * plane + (unsigned)min(max(yof(k), 0), ny - 1) * nz + zc));
    }
    // The march is compiled twice for the interior tiles, with and without the material index, so that
    // the prefetched index and the prefetched 1/eps never meet in one register: merging them makes the
    // compiler wait for the loads (and for the staging issued before them) as soon as they are issued.
    auto run = [&](auto indexed) {
        constexpr bool MAT = decltype(indexed)::value;
        const unsigned char* __restrict__ mat = MAT ? material : nullptr;
        // E(n) and 1/eps (or its material index) of a plane's cells.  The lean variants load them one plane
        // ahead, right after the barrier, so that the loads overlap a plane of work instead of the barrier
        // that follows them; the boundary tiles (register-bound) load them just before that barrier.
        constexpr bool PREFETCH = CPML != 7;
        // The z-face tiles also load the z-slab CPML memory that step n+1 of the next plane reads (E at
        // plane x+1, H at plane x) with it.
        constexpr bool ZPRE = CPML == 4;
        // The x-end tiles without y or z slabs load the x-slab CPML memory of step n+1 a plane ahead
        // the same way (the x-slab rows of a plane share one slab row, so every lane reads one).
        constexpr bool XPRE = CPML == 1;
        V3<real> eo_n[RW], inv3_n[RW];
        real inv1_n[RW];
        unsigned idx_n[RW];
        real pzE[RW][2] = {}, pzH[RW][2] = {};
        real pxE[RW][2] = {}, pxH[RW][2] = {};
        // H-family x-slab memory of plane x-1, which stage 1b of plane x reads (x <= nx).
        auto load_xh = [&](int x) {
            if (!XPRE || x < 1) return;
            const Hit hx = find_hit(a.hf, 0, x - 1);
            if (hx.face < 0) return;
    #pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int yk = yof(k);
                if (zin && yk >= 0 && yk < ny) {
                    const unsigned px = ((unsigned)hx.q * ny + yk) * nz + z;
                    pxH[k][0] = global(a.hf.psi_in)[hx.psi0 + px];
                    pxH[k][1] = global(a.hf.psi_in)[hx.psi1 + px];
                }
            }
        };
        // H-family z-slab memory of plane x-1, which stage 1b of plane x reads (x <= nx).
        auto load_zh = [&](int x) {
    #pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int yk = yof(k);
                if (ZPRE && yk >= 0 && yk < ny && hz_hit.face >= 0 && x >= 1) {
                    const unsigned pz = ((unsigned)(x - 1) * ny + yk) * hz_hit.w + hz_hit.q;
                    pzH[k][0] = global(a.hf.psi_in)[hz_hit.psi0 + pz];
                    pzH[k][1] = global(a.hf.psi_in)[hz_hit.psi1 + pz];
                }
            }
        };
        auto load_plane = [&](int x) {
            const unsigned base = (unsigned)x * plane;
    #pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int yk = yof(k);
                const unsigned cell = base + (unsigned)min(max(yk, 0), ny - 1) * nz + zc;
                eo_n[k] = load3(e_in, 3u * cell);
                if (EPS3) inv3_n[k] = load3(inv_eps, 3u * cell);
                else if (mat) idx_n[k] = global(mat)[cell];
                else inv1_n[k] = global(inv_eps)[cell];
                if (ZPRE && yk >= 0 && yk < ny && ez_hit.face >= 0) {
                    const unsigned pz = ((unsigned)x * ny + yk) * ez_hit.w + ez_hit.q;
                    pzE[k][0] = global(a.ef.psi_in)[ez_hit.psi0 + pz];
                    pzE[k][1] = global(a.ef.psi_in)[ez_hit.psi1 + pz];
                }
            }
            if (XPRE) {
                const Hit hx = find_hit(a.ef, 0, x);
                if (hx.face >= 0) {
    #pragma unroll
                    for (int k = 0; k < RW; ++k) {
                        const int yk = yof(k);
                        if (zin && yk >= 0 && yk < ny) {
                            const unsigned px = ((unsigned)hx.q * ny + yk) * nz + z;
                            pxE[k][0] = global(a.ef.psi_in)[hx.psi0 + px];
                            pxE[k][1] = global(a.ef.psi_in)[hx.psi1 + px];
                        }
                    }
                }
            }
            load_zh(x);
            load_xh(x);
        };
        if (PREFETCH && xs < nx) load_plane(xs);
        auto march_plane = [&](const int x, real* const sHx, real* const sHn) {
            const bool pE1 = x < nx;
            const bool pH1 = x >= xa && x >= 1 && x - 1 < nx;
            const bool pE2 = x - 1 >= xa && x - 1 <= xb && x - 1 < nx;
            const bool pH2 = x - 2 >= xa && x - 2 < xb;
            if (!PREFETCH && pE1) load_plane(x);
            __syncthreads();   // H(n) of plane x staged; E1(x-1) and E2(x-2) written to LDS
            if (x + 1 <= xe && x + 1 < nx) stage(x + 1, sHn);
            V3<real> eo[RW], inv[RW];
            if (pE1) {
    #pragma unroll
                for (int k = 0; k < RW; ++k) {
                    eo[k] = eo_n[k];
                    if (EPS3) {
                        inv[k] = inv3_n[k];
                    } else {
                        const real v = mat ? sMat[idx_n[k]] : inv1_n[k];
                        inv[k] = V3<real>{v, v, v};
                    }
                }
            }
            real zE[RW][2], zH[RW][2];   // this plane's z-slab CPML memory, before the next plane's loads
            real xE[RW][2], xH[RW][2];   // and its x-slab memory
    #pragma unroll
            for (int k = 0; k < RW; ++k) {
                zE[k][0] = pzE[k][0]; zE[k][1] = pzE[k][1];
                zH[k][0] = pzH[k][0]; zH[k][1] = pzH[k][1];
                xE[k][0] = pxE[k][0]; xE[k][1] = pxE[k][1];
                xH[k][0] = pxH[k][0]; xH[k][1] = pxH[k][1];
            }
            if (PREFETCH && x + 1 <= xe && x + 1 < nx) load_plane(x + 1);
            // A chunk that ends at the grid's last plane (no high-x slab) still computes H1(nx-1) at
            // x = nx: its z- and x-slab memory is loaded without a plane nx of E to go with it.
            else if ((ZPRE || XPRE) && x + 1 <= xe && x + 1 == nx) { load_zh(nx); load_xh(nx); }
            V3<real> e1[RW], hn[RW], h1[RW], e2[RW];
            // Stage 1: E1(x).
            const Hit ex1 = find_hit(a.ef, 0, x);
    #pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int r = rowof(k), y = yof(k);
                const real* p = sHx + 3 * ((r + 1) * SC + lane + 1);
                hn[k] = lds3(p);
                e1[k] = zero3<real>();
                const bool lanes = zin;
                if (pE1 && (ALL_LANES || lanes) && y >= 0 && y < ny) {
                    const real* pym = p - 3 * SC;
                    const real* pzm = p - 3;
                    const Hit ey = find_hit(a.ef, 1, y);
                    if (ZPRE) { pe1[k][4] = zE[k][0]; pe1[k][5] = zE[k][1]; }
                    if (XPRE) { pe1[k][0] = xE[k][0]; pe1[k][1] = xE[k][1]; }
                    V3<real> v = e_cell2<real, CPML, 1, ZPRE, XPRE>(a, q, x, y, z, false, ex1, ey, ez_hit, ez_coef, hn[k], Hn_c[k],
                                                              lds3(pym), pzm[0], pzm[1], eo[k], inv[k], pe1[k]);
                    add_sources(a.se, q.nse, q.ebl, q.ebh, n, q.samples, x, y, z, v.x, v.y, v.z);
                    e1[k] = sel3(lanes, v);
                    // Snapshot of step n+1: the cells this tile owns (those whose E2 it stores).
                    if (SNAP && x >= q.snap_lo && x < q.snap_hi && x >= xa && x < xb && r >= 1 && r < NR - 2 && lane >= 1 &&
                        lane < 62 && zin && y < ylim) {
                        const int sc = snapshot_cell(a, x, y, z);
                        if (sc >= 0) store3(a.snap_e, 3u * (unsigned)sc, v);
                    }
                }
            }
            // Stage 1b: H1(x-1) from E1(x-1) (LDS, with the row and column above) and E1(x).
            const Hit hx1 = find_hit(a.hf, 0, x - 1);
    #pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int r = rowof(k), y = yof(k);
                h1[k] = zero3<real>();
                const bool lanes = lane < 63 && zin;
                if (pH1 && r < NR - 1 && (ALL_LANES || lanes) && y >= 0 && y < ny) {
                    const real* p = sE1 + 3 * (r * 64 + lane);
                    const real* pzn = p + 3;
                    const Hit hy = find_hit(a.hf, 1, y);
                    if (ZPRE) { ph1[k][4] = zH[k][0]; ph1[k][5] = zH[k][1]; }
                    if (XPRE) { ph1[k][0] = xH[k][0]; ph1[k][1] = xH[k][1]; }
                    V3<real> v = h_cell2<real, CPML, 1, ZPRE, XPRE>(a, q, x - 1, y, z, false, hcoef, hx1, hy, hz_hit, hz_coef, E1_c[k],
                                                              pE1 ? e1[k] : E1_c[k], lds3(p + 3 * 64), pzn[0], pzn[1], Hn_c[k],
                                                              ph1[k]);
                    add_sources(a.sh, q.nsh, q.hbl, q.hbh, n, q.samples, x - 1, y, z, v.x, v.y, v.z);
                    h1[k] = sel3(lanes, v);
                    if (SNAP && x - 1 >= q.snap_lo && x - 1 < q.snap_hi && x - 1 >= xa && x - 1 < xb && r >= 1 && r < NR - 2 &&
                        lane >= 1 && lane < 62 && zin && y < ylim) {
                        const int sc = snapshot_cell(a, x - 1, y, z);
                        if (sc >= 0) store3(a.snap_h, 3u * (unsigned)sc, v);
                    }
                }
                put3(sH1 + 3 * (r * 64 + lane), h1[k]);
            }
            lds_barrier();   // H1(x-1) of every row in LDS
            // Stage 2: E2(x-1); the owned cells (rows y0 .. y0+8, lanes 1..61, planes < xb) are stored.
            const Hit ex2 = find_hit(a.ef, 0, x - 1);
    #pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int r = rowof(k), y = yof(k);
                e2[k] = zero3<real>();
                const bool lanes = lane >= 1 && lane < 63 && zin;
                if (pE2 && r >= 1 && r < NR - 1 && (ALL_LANES || lanes) && y >= 0 && y < ny) {
                    const real* p = sH1 + 3 * (r * 64 + lane);
                    const real* pzm = p - 3;
                    const bool owned = lanes && r < NR - 2 && lane < 62 && x - 1 < xb && y < ylim;
                    const Hit ey = find_hit(a.ef, 1, y);
                    V3<real> v = e_cell2<real, CPML, 2>(a, q, x - 1, y, z, owned, ex2, ey, ez_hit, ez_coef, h1[k], H1_c[k],
                                                        lds3(p - 3 * 64), pzm[0], pzm[1], E1_c[k], inv_c[k], pe0[k]);
                    add_sources(a.se, q.nse, q.ebl, q.ebh, n + 1, q.samples, x - 1, y, z, v.x, v.y, v.z);
                    e2[k] = sel3(lanes, v);
                    if (owned) store3(e_out, 3u * ((unsigned)(x - 1) * plane + (unsigned)y * nz + (unsigned)z), e2[k]);
                }
            }
            // Stage 2b: H2(x-2) from E2(x-2) (LDS, with the row and column above) and E2(x-1).
            const Hit hx2 = find_hit(a.hf, 0, x - 2);
    #pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int r = rowof(k), y = yof(k);
                const bool lanes = lane >= 1 && lane < 62 && zin;
                if (pH2 && r >= 1 && r < NR - 2 && (ALL_LANES || lanes) && y >= 0 && y < ny && y < ylim) {
                    const real* p = sE2 + 3 * (r * 64 + lane);
                    const real* pzn = p + 3;
                    const Hit hy = find_hit(a.hf, 1, y);
                    V3<real> h2 = h_cell2<real, CPML, 2>(a, q, x - 2, y, z, lanes, hcoef, hx2, hy, hz_hit, hz_coef, E2_c[k],
                                                         pE2 ? e2[k] : E2_c[k], lds3(p + 3 * 64), pzn[0], pzn[1], H1_c[k], ph0[k]);
                    add_sources(a.sh, q.nsh, q.hbl, q.hbh, n + 1, q.samples, x - 2, y, z, h2.x, h2.y, h2.z);
                    if (lanes) store3(h_out, 3u * ((unsigned)(x - 2) * plane + (unsigned)y * nz + (unsigned)z), h2);
                }
            }
            lds_barrier();   // every read of sE1, sH1 and sE2 for this plane is done
    #pragma unroll
            for (int k = 0; k < RW; ++k) {
                const int r = rowof(k);
                put3(sE1 + 3 * (r * 64 + lane), e1[k]);
                put3(sE2 + 3 * (r * 64 + lane), e2[k]);
                Hn_c[k] = hn[k];
                H1_c[k] = h1[k];
                E1_c[k] = e1[k];
                E2_c[k] = e2[k];
                inv_c[k] = pE1 ? inv[k] : zero3<real>();
    #pragma unroll
                for (int t = 0; t < 6; ++t) {
                    pe0[k][t] = pe1[k][t];
                    ph0[k][t] = ph1[k][t];
                }
            }
        };
        if (xs < nx) stage(xs, sA);
        for (int x = xs; x <= xe; x += 2) {
            march_plane(x, sA, sB);
            if (x + 1 <= xe) march_plane(x + 1, sB, sA);
        }
    };
    if (INDEXED && material) run(std::true_type{});
    else run(std::false_type{});
}

#ifndef TFDTD_SERIAL_VARIANTS
constexpr bool kConcurrentVariants = true;
#else
constexpr bool kConcurrentVariants = false;
#endif

// Occupancy targets (waves per SIMD), measured at 512^3 with the H planes staged by direct loads:
// FP32 tiles of 8 rows ask 8 for the interior (it compiles to 73 VGPRs, i.e. 6) and 6 for the z-face
// and boundary tiles (80 VGPRs, a few spilled); FP64 and 16-row tiles hold twice the values per lane.
// Loading the next plane's E and 1/eps ahead in the z-face and boundary tiles, with 4 waves per SIMD
// so that it does not spill, measured slower: 2.44 against 2.25 ms per step at 512^3 and 0.173
// against 0.139 s per 100 steps on the 64-grid uneven batch.  Loading only the z-slab CPML memory a
// plane ahead at 6 waves (E after the plane's E, H after its H) spills more and was slower too: 2.29
// ms and 0.158 s.
// FP32 z-face and boundary tiles at 7 waves per SIMD: 2.36 against 2.24 ms per step at 512^3.
#ifndef TFDTD_TILE_EDGE_WAVES
#define TFDTD_TILE_EDGE_WAVES 6
#endif
#define TFDTD_TILE_BOUNDS                                                                          \
    __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(                                   \
        CPML == 0 ? 8 : sizeof(real) == 4 ? TFDTD_TILE_EDGE_WAVES : 3,                               \
        8)))

template <typename real, int W, bool EPS3, int CPML>
__global__ TFDTD_TILE_BOUNDS void tile_kernel(const March<real> a) {
    tile_body<real, W, EPS3, CPML>(a, blockIdx.x, gridDim.x);
}

// Many grids in one launch per variant.  Grid g owns workgroups [prefix[g], prefix[g+1]), a multiple
// of 8 so that its tiles keep the XCD assignment of a launch of their own; the padding exits at once.
template <typename real, int W, bool EPS3, int CPML>
__global__ TFDTD_TILE_BOUNDS void batch_tile_kernel(const March<real>* __restrict__ table, const int* __restrict__ prefix,
                                                    int grids) {
    const int b = blockIdx.x;
    int lo = 0, hi = grids;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= b) lo = mid;
        else hi = mid;
    }
    const March<real>& a = table[lo];
    const int count = a.blocks[CPML == 0 ? 0 : CPML == 4 ? 1 : 2];
    const int w = b - prefix[lo];
    if (w >= count) return;
    tile_body<real, W, EPS3, CPML>(a, w, count);
}

#undef TFDTD_TILE_BOUNDS

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
    if (d->material) {
        if (d->eps_stride != 1 || d->materials < 1 || d->materials > 256 || !d->material_table)
            return tfdtd_fail(TFDTD_EINVAL, "march: a material index needs eps_stride 1, 1 to 256 materials and a table");
        a->material = d->material;
        a->material_table = static_cast<const real*>(d->material_table);
        a->materials = d->materials;
    }
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

// Tiles [lo, hi) along one axis whose computed E and H cells avoid every CPML row of that axis.
// E cells of tile k cover [k*size, k*size + size + e_extra), H cells [k*size, k*size + size).
inline void free_range(int tiles, int size, int e_extra, int n, const Face* ef, const Face* hf, int* lo, int* hi, int shift = 0) {
    auto clear = [&](int k) {
        const int a0 = std::max(0, k * size - shift), e1 = std::min(n, k * size - shift + size + e_extra);
        const int h1 = std::min(n, k * size - shift + size);
        for (int f = 0; f < 2; ++f) {
            if (ef[f].hi > ef[f].lo && ef[f].lo < e1 && a0 < ef[f].hi) return false;
            if (hf[f].hi > hf[f].lo && hf[f].lo < h1 && a0 < hf[f].hi) return false;
        }
        return true;
    };
    int best_lo = 0, best_hi = 0;
    for (int k = 0; k < tiles;) {
        if (!clear(k)) { ++k; continue; }
        int j = k;
        while (j < tiles && clear(j)) ++j;
        if (j - k > best_hi - best_lo) { best_lo = k; best_hi = j; }
        k = j;
    }
    *lo = best_lo;
    *hi = best_hi;
}

// Tile layout of one grid: chunking, z shift, the slab-free boxes and the workgroups of each variant.
template <typename real, int W>
int layout_march(const tfdtd_march* d, March<real>* a) {
    if (d->snap_count != 0) return tfdtd_fail(TFDTD_EINVAL, "march: the step n+1 snapshot is an option of the two-step pass");
    const int status = make_march(d, a);
    if (status) return status;
    constexpr int TY = 2 * W;
    const int chunk = d->chunk_x > 0 ? d->chunk_x : d->nx;
    a->chunk = chunk;
    a->ytiles = (d->ny + TY - 1) / TY;
    // Shift the z tiles so one tile edge meets the end of the low z slabs (E and H), when that
    // lowers the cost of a z column: slab-free tiles weigh 10, tiles with z slab lanes 17 (their
    // measured relative durations at 512^3).  A short z axis keeps one unshifted tile.
    const int zend = std::max(a->ef.face[2][0].hi, a->hf.face[2][0].hi);
    auto column_cost = [&](int shift) {
        const int tiles = (d->nz + shift + 62) / 63;
        int lo, hi;
        free_range(tiles, 63, 1, d->nz, a->ef.face[2], a->hf.face[2], &lo, &hi, shift);
        return 10 * (hi - lo) + 17 * (tiles - (hi - lo));
    };
    const int aligned = zend > 0 ? (63 - zend % 63) % 63 : 0;
    a->zshift = column_cost(aligned) < column_cost(0) ? aligned : 0;
    a->ztiles = (d->nz + a->zshift + 62) / 63;
    const int chunks = (d->nx + chunk - 1) / chunk;
    const long long blocks = (long long)chunks * a->ytiles * a->ztiles;
    if (blocks > 0x7fffffffLL) return tfdtd_fail(TFDTD_EINVAL, "march: too many workgroups");
    // E planes of chunk c are xa .. xb (xb recomputed), so E extends one row/plane past each tile.
    free_range(chunks, chunk, 1, d->nx, a->ef.face[0], a->hf.face[0], &a->box_lo[0], &a->box_hi[0]);
    free_range(a->ytiles, TY, 1, d->ny, a->ef.face[1], a->hf.face[1], &a->box_lo[1], &a->box_hi[1]);
    a->nchunks = chunks;
    free_range(a->ztiles, 63, 1, d->nz, a->ef.face[2], a->hf.face[2], &a->zbox_lo, &a->zbox_hi, a->zshift);
    const long long xy = (long long)(a->box_hi[0] - a->box_lo[0]) * (a->box_hi[1] - a->box_lo[1]);
    const long long pure = xy * (a->zbox_hi - a->zbox_lo), inner = xy * a->ztiles;
    if (inner == 0) { a->box_lo[0] = a->box_hi[0] = 0; a->box_lo[1] = a->box_hi[1] = 0; }
    a->blocks[0] = (int)pure;
    a->blocks[1] = (int)(inner - pure);
    a->blocks[2] = (int)(blocks - inner);
    return TFDTD_OK;
}

// Two side streams per device for the concurrent variant launches, with the events that fork them
// from and join them back to the caller's stream.  Created on first use outside stream capture and
// kept for the life of the process; every stream has the default priority.  `enqueue` is held while
// a call records the fork, launches and records the join, so calls from several host threads keep
// their fork and join events apart.
struct Side {
    hipStream_t stream[2];
    hipEvent_t fork, join[2];
    std::mutex enqueue;
};

Side* side_streams(hipStream_t caller) {
    static std::mutex lock;
    static Side* sides[64] = {};
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= 64) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> guard(lock);
    if (sides[device]) return sides[device];
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(caller, &capture) != hipSuccess || capture != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        return nullptr;
    }
    Side* side = new Side;
    bool ok = hipEventCreateWithFlags(&side->fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < 2 && ok; ++k)
        ok = hipStreamCreateWithFlags(&side->stream[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&side->join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        delete side;     // a failed creation leaves the launches serial
        (void)hipGetLastError();
        return nullptr;
    }
    return sides[device] = side;
}

// The three variants write disjoint cells and read only time-n buffers, so they may run at once:
// the z-face and boundary tiles (latency-bound) overlap the interior tiles (bandwidth-bound).
// Without side streams (first use under capture) they run one after another on the caller's stream.
// launch(v, s) starts variant v (0 interior, 1 z faces, 2 boundary) on stream s.
template <typename Launch>
int launch_variants(const int blocks[3], hipStream_t stream, Launch&& launch) {
    int status = TFDTD_OK;
    auto run = [&](int v, hipStream_t s) {
        if (blocks[v] > 0 && !status) status = launch(v, s);
    };
    auto check = [&](hipError_t e, const char* what) {
        if (!status && e != hipSuccess) status = tfdtd_fail(TFDTD_EHIP, "march: %s: %s", what, hipGetErrorString(e));
    };
    Side* side = kConcurrentVariants && (blocks[1] > 0 || blocks[2] > 0) ? side_streams(stream) : nullptr;
    if (!side) {
        for (int v = 0; v < 3; ++v) run(v, stream);
        return status;
    }
    std::lock_guard<std::mutex> guard(side->enqueue);
    check(hipEventRecord(side->fork, stream), "fork");
    for (int k = 0; k < 2; ++k) check(hipStreamWaitEvent(side->stream[k], side->fork, 0), "fork wait");
    run(2, side->stream[0]);
    run(1, side->stream[1]);
    run(0, stream);
    for (int k = 0; k < 2; ++k) {
        check(hipEventRecord(side->join[k], side->stream[k]), "join");
        check(hipStreamWaitEvent(stream, side->join[k], 0), "join wait");
    }
    return status;
}

template <typename real, int W>
int launch_tile(const tfdtd_march* d, hipStream_t stream) {
    March<real> a;
    int status = layout_march<real, W>(d, &a);
    if (status) return status;
    const bool eps3 = d->eps_stride == 3;
    return launch_variants(a.blocks, stream, [&](int v, hipStream_t s) {
        auto go = [&](auto kernel, const char* what) {
            hipLaunchKernelGGL(kernel, dim3((unsigned)a.blocks[v]), dim3(64 * W), 0, s, a);
            return tfdtd_check_launch(what);
        };
        if (v == 0) return eps3 ? go(tile_kernel<real, W, true, 0>, "march interior") : go(tile_kernel<real, W, false, 0>, "march interior");
        if (v == 1) return eps3 ? go(tile_kernel<real, W, true, 4>, "march z faces") : go(tile_kernel<real, W, false, 4>, "march z faces");
        return eps3 ? go(tile_kernel<real, W, true, 7>, "march boundary") : go(tile_kernel<real, W, false, 7>, "march boundary");
    });
}

// ---------------------------------------------------------------------------------------------------
// Two-step launches (float32).

// The five LDS planes (51 KB) allow three workgroups per CU.  More wavefronts per workgroup hold fewer
// rows each, hence fewer registers: the boundary tiles carry the CPML memory of both stages of every
// axis in registers.  TFDTD_TILE2_BOUNDARY_WAVES sets the waves per SIMD the boundary variants
// (CPML 6 and 7, which compile in the y and z slabs) are held to; the other variants use 6 at W 12.
#ifndef TFDTD_TILE2_BOUNDARY_WAVES
#define TFDTD_TILE2_BOUNDARY_WAVES 6
#endif
// Waves per SIMD the 12-wavefront interior, z-face and x-only boundary variants ask for: at 7 the z-face
// variant compiles without scratch spills (81 VGPRs) and the 512^3 pass takes 1.344 instead of 1.381 ms
// per step (8: the same as 7).
#ifndef TFDTD_TILE2_INTERIOR_WAVES
#define TFDTD_TILE2_INTERIOR_WAVES 7
#endif
#define TFDTD_TILE2_WAVES(CPML, W) \
    ((CPML) == 6 || (CPML) == 7 ? TFDTD_TILE2_BOUNDARY_WAVES : (W) == 16 ? 8 : (W) == 12 ? TFDTD_TILE2_INTERIOR_WAVES : 1)

template <bool EPS3, int CPML, int W, bool SNAP = false>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(TFDTD_TILE2_WAVES(CPML, W), 8))) void tile2_kernel(const March<float> a) {
    tile2_body<float, EPS3, CPML, W, W == 16 ? 16 : 12, SNAP>(a, blockIdx.x, gridDim.x);
}

// Many grids (or one) from a device table, as batch_tile_kernel: the tile parameters are loaded on
// demand instead of occupying the kernel-argument registers.
template <bool EPS3, int CPML, int W, bool SNAP = false>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(TFDTD_TILE2_WAVES(CPML, W), 8))) void batch_tile2_kernel(
    const March<float>* __restrict__ table, const int* __restrict__ prefix, int grids) {
    const int b = blockIdx.x;
    int lo = 0, hi = grids;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= b) lo = mid;
        else hi = mid;
    }
    const March<float>& a = table[lo];
    const int count = a.blocks[CPML == 0 ? 0 : CPML == 4 ? 1 : 2];
    const int w = b - prefix[lo];
    if (w >= count) return;
    if ((CPML == 7 || CPML == 1) && w >= a.nxend) return;   // an x-middle boundary tile: the CPML 6 launch takes it
    if (CPML == 6 && w < a.nxend) return;
    if (CPML == 7 || CPML == 1) {
        // x-end tiles whose rows and columns avoid the y and z slabs need the x terms only (CPML 1).
        int ty, tz, chunk;
        tile_coords<float, 7>(a, w, count, ty, tz, chunk);
        const bool x_only = ty >= a.box_lo[1] && ty < a.box_hi[1] && tz >= a.zbox_lo && tz < a.zbox_hi;
        if ((CPML == 1) != x_only) return;
    }
    tile2_body<float, EPS3, CPML, W, W == 16 ? 16 : 12, SNAP>(a, w, count);
}

// Tiles [lo, hi) along one axis whose cells of both stages avoid every CPML row of that axis: a tile
// of `size` cells starting at s computes E on [s-1, s+size+2) and H on [s-1, s+size+1).
inline void free_range2(int tiles, int size, int n, const Face* ef, const Face* hf, int* lo, int* hi, int shift = 0) {
    auto clear = [&](int k) {
        const int s0 = k * size - shift;
        const int a0 = std::max(0, s0 - 1), e1 = std::min(n, s0 + size + 2), h1 = std::min(n, s0 + size + 1);
        for (int f = 0; f < 2; ++f) {
            if (ef[f].hi > ef[f].lo && ef[f].lo < e1 && a0 < ef[f].hi) return false;
            if (hf[f].hi > hf[f].lo && hf[f].lo < h1 && a0 < hf[f].hi) return false;
        }
        return true;
    };
    int best_lo = 0, best_hi = 0;
    for (int k = 0; k < tiles;) {
        if (!clear(k)) { ++k; continue; }
        int j = k;
        while (j < tiles && clear(j)) ++j;
        if (j - k > best_hi - best_lo) { best_lo = k; best_hi = j; }
        k = j;
    }
    *lo = best_lo;
    *hi = best_hi;
}

// Two-step tile layout: 9 x 61 tiles, chunks of chunk_x planes, the z shift and the slab-free boxes.
int layout_march2(const tfdtd_march* d, March<float>* a) {
    tfdtd_march in_place = *d;                  // make_march checks the one-step (in-place H memory) form
    in_place.h.psi_out = const_cast<void*>(d->h.psi_in);
    int status = make_march(&in_place, a);
    if (status) return status;
    bool h_slabs = false;
    for (int ax = 0; ax < 3; ++ax)
        for (int f = 0; f < 2; ++f) h_slabs |= a->hf.face[ax][f].hi > a->hf.face[ax][f].lo;
    if (h_slabs && (!d->h.psi_out || d->h.psi_out == d->h.psi_in))
        return tfdtd_fail(TFDTD_EINVAL, "march2: the H-family CPML memory needs separate input and output buffers");
    a->hf.psi_out = static_cast<float*>(h_slabs ? d->h.psi_out : const_cast<void*>(d->h.psi_in));
    if (d->snap_count != 0) {
        if (d->snap_count < 0 || d->snap_count > 4 || !d->snap_e || !d->snap_h)
            return tfdtd_fail(TFDTD_EINVAL, "march2: the step n+1 snapshot takes 1 to 4 boxes and an E and an H buffer");
        const int n[3] = {d->nx, d->ny, d->nz};
        long long base = 0;
        a->snap_lo = d->nx;
        a->snap_hi = 0;
        for (int k = 0; k < d->snap_count; ++k) {
            for (int ax = 0; ax < 3; ++ax) {
                const int lo = d->snap_lo[k][ax], hi = d->snap_hi[k][ax];
                if (lo < 0 || hi > n[ax] || hi <= lo)
                    return tfdtd_fail(TFDTD_EINVAL, "march2: snapshot box %d is empty or outside the grid along axis %d", k, ax);
                a->sbox_lo[k][ax] = lo;
                a->sbox_hi[k][ax] = hi;
            }
            for (int j = 0; j < k; ++j) {
                bool overlap = true;
                for (int ax = 0; ax < 3; ++ax)
                    overlap = overlap && d->snap_lo[k][ax] < d->snap_hi[j][ax] && d->snap_lo[j][ax] < d->snap_hi[k][ax];
                if (overlap) return tfdtd_fail(TFDTD_EINVAL, "march2: snapshot boxes %d and %d overlap", j, k);
            }
            a->sbox_base[k] = (unsigned)base;
            base += (long long)(d->snap_hi[k][0] - d->snap_lo[k][0]) * (d->snap_hi[k][1] - d->snap_lo[k][1]) *
                    (d->snap_hi[k][2] - d->snap_lo[k][2]);
            a->snap_lo = std::min(a->snap_lo, d->snap_lo[k][0]);
            a->snap_hi = std::max(a->snap_hi, d->snap_hi[k][0]);
        }
        if (3 * base > 0x7fffffffLL) return tfdtd_fail(TFDTD_EINVAL, "march2: the snapshot exceeds 31-bit offsets");
        a->snap_e = static_cast<float*>(d->snap_e);
        a->snap_h = static_cast<float*>(d->snap_h);
        a->snap_count = d->snap_count;
    }
    const int chunk = d->chunk_x > 0 ? d->chunk_x : d->nx;
    a->chunk = chunk;
    // Shift the y tiles to the cheapest of the nine phases (slab-free tiles weigh 10, the others 17).
    auto row_cost = [&](int shift) {
        const int tiles = (d->ny + shift + kTY2 - 1) / kTY2;
        int lo, hi;
        free_range2(tiles, kTY2, d->ny, a->ef.face[1], a->hf.face[1], &lo, &hi, shift);
        return 10 * (hi - lo) + 17 * (tiles - (hi - lo));
    };
    a->yshift = 0;
    for (int shift = 1; shift < kTY2; ++shift)
        if (row_cost(shift) < row_cost(a->yshift)) a->yshift = shift;
    a->ytiles = (d->ny + a->yshift + kTY2 - 1) / kTY2;
    // Shift the z tiles so one tile's widened columns start just past the low z slabs, when that lowers
    // the cost of a z column (slab-free tiles weigh 10, tiles with z slab columns 17).
    const int zend = std::max(a->ef.face[2][0].hi, a->hf.face[2][0].hi);
    auto column_cost = [&](int shift) {
        const int tiles = (d->nz + shift + kTZ2 - 1) / kTZ2;
        int lo, hi;
        free_range2(tiles, kTZ2, d->nz, a->ef.face[2], a->hf.face[2], &lo, &hi, shift);
        return 10 * (hi - lo) + 17 * (tiles - (hi - lo));
    };
    const int aligned = zend > 0 ? (kTZ2 - (zend + 1) % kTZ2) % kTZ2 : 0;
    a->zshift = column_cost(aligned) < column_cost(0) ? aligned : 0;
    a->ztiles = (d->nz + a->zshift + kTZ2 - 1) / kTZ2;
    // x: the slab-free middle [x0, x1) -- a chunk [xa, xb) computes E on [xa-1, xb+2) and H on
    // [xa-1, xb+1) -- and the two ends, each cut into chunks of at most chunk_x planes, so that the
    // boundary chunks are no longer than the slabs they cover.
    const Face& el = a->ef.face[0][0]; const Face& hl = a->hf.face[0][0];
    const Face& eh = a->ef.face[0][1]; const Face& hh = a->hf.face[0][1];
    const int low = std::max(el.hi > el.lo ? el.hi : 0, hl.hi > hl.lo ? hl.hi : 0);
    int x0 = low > 0 ? low + 1 : 0;
    int x1 = std::min(eh.hi > eh.lo ? eh.lo - 2 : d->nx, hh.hi > hh.lo ? hh.lo - 1 : d->nx);
    if (x1 <= x0) x0 = x1 = d->nx;     // no slab-free middle: every chunk is a boundary chunk
    a->x0 = x0;
    a->x1 = x1;
    const int part[3] = {x0, x1 - x0, d->nx - x1};
    for (int k = 0; k < 3; ++k) a->xparts[k] = (part[k] + chunk - 1) / chunk;
    const int chunks = a->xparts[0] + a->xparts[1] + a->xparts[2];
    const long long blocks = (long long)chunks * a->ytiles * a->ztiles;
    if (blocks > 0x7fffffffLL) return tfdtd_fail(TFDTD_EINVAL, "march2: too many workgroups");
    a->box_lo[0] = a->xparts[0];
    a->box_hi[0] = a->xparts[0] + a->xparts[1];
    free_range2(a->ytiles, kTY2, d->ny, a->ef.face[1], a->hf.face[1], &a->box_lo[1], &a->box_hi[1], a->yshift);
    a->nchunks = chunks;
    free_range2(a->ztiles, kTZ2, d->nz, a->ef.face[2], a->hf.face[2], &a->zbox_lo, &a->zbox_hi, a->zshift);
    const long long xy = (long long)(a->box_hi[0] - a->box_lo[0]) * (a->box_hi[1] - a->box_lo[1]);
    const long long pure = xy * (a->zbox_hi - a->zbox_lo), inner = xy * a->ztiles;
    if (inner == 0) { a->box_lo[0] = a->box_hi[0] = 0; a->box_lo[1] = a->box_hi[1] = 0; }
    a->blocks[0] = (int)pure;
    a->blocks[1] = (int)(inner - pure);
    a->blocks[2] = (int)(blocks - inner);
    // Tall interior tiles (tile_y 16: 16 wavefronts, 13 owned rows) retile the slab-free rows of the
    // 9-row tiles; the z-face and boundary tiles keep the 9-row tiles.
    a->iy0 = std::max(a->box_lo[1] * kTY2 - a->yshift, 0);
    a->iy1 = std::min(a->box_hi[1] * kTY2 - a->yshift, d->ny);
    a->iytiles = a->iy1 > a->iy0 ? (a->iy1 - a->iy0 + 12) / 13 : 0;
    if (d->tile_y == 16) a->blocks[0] = (a->box_hi[0] - a->box_lo[0]) * a->iytiles * (a->zbox_hi - a->zbox_lo);
    a->nxend = (chunks - (a->box_hi[0] - a->box_lo[0])) * a->ytiles * a->ztiles;   // blocks A and B of tile_coords
    return TFDTD_OK;
}

// W wavefronts per interior tile; the z-face and boundary tiles (more registers per row) use 12.
template <int W, bool SNAP = false>
int launch_tile2(const March<float>& a, bool eps3, hipStream_t stream) {
    return launch_variants(a.blocks, stream, [&](int v, hipStream_t s) {
        auto go = [&](auto kernel, int waves, const char* what) {
            hipLaunchKernelGGL(kernel, dim3((unsigned)a.blocks[v]), dim3(64 * waves), 0, s, a);
            return tfdtd_check_launch(what);
        };
        if (v == 0) return eps3 ? go(tile2_kernel<true, 0, W, SNAP>, W, "march2 interior") : go(tile2_kernel<false, 0, W, SNAP>, W, "march2 interior");
        if (v == 1) return eps3 ? go(tile2_kernel<true, 4, 12, SNAP>, 12, "march2 z faces") : go(tile2_kernel<false, 4, 12, SNAP>, 12, "march2 z faces");
        return eps3 ? go(tile2_kernel<true, 7, 12, SNAP>, 12, "march2 boundary") : go(tile2_kernel<false, 7, 12, SNAP>, 12, "march2 boundary");
    });
}

int pack_marches2(const tfdtd_march* d, int count, void* host_table, size_t bytes, int32_t* prefix) {
    if (!host_table || bytes < sizeof(March<float>) * (size_t)count) return tfdtd_fail(TFDTD_EINVAL, "batch march2: table buffer too small");
    March<float>* table = static_cast<March<float>*>(host_table);
    long long total[3] = {0, 0, 0};
    for (int k = 0; k < count; ++k) {
        if (d[k].dtype != TFDTD_F32) return tfdtd_fail(TFDTD_EINVAL, "march2: float32 only");
        if (d[k].tile_y != d[0].tile_y || d[k].eps_stride != d[0].eps_stride)
            return tfdtd_fail(TFDTD_EINVAL, "batch march2: descriptor %d differs in tile_y or eps_stride", k);
        const int status = layout_march2(d + k, table + k);
        if (status) return status;
        for (int v = 0; v < 3; ++v) {
            prefix[v * (count + 1) + k] = (int32_t)total[v];
            total[v] += (table[k].blocks[v] + 7) / 8 * 8;
            if (total[v] > 0x7fffffffLL) return tfdtd_fail(TFDTD_EINVAL, "batch march2: too many workgroups for one launch");
        }
    }
    for (int v = 0; v < 3; ++v) prefix[v * (count + 1) + count] = (int32_t)total[v];
    return TFDTD_OK;
}

template <int W, bool SNAP = false>
int launch_batch_tiles2(const void* table, const int32_t* prefix, int count, bool eps3, const int32_t* totals, hipStream_t stream) {
    const March<float>* t = static_cast<const March<float>*>(table);
    const int blocks[3] = {totals[0], totals[1], totals[2]};
    return launch_variants(blocks, stream, [&](int v, hipStream_t s) {
        auto go = [&](auto kernel, int waves, const char* what) {
            hipLaunchKernelGGL(kernel, dim3((unsigned)blocks[v]), dim3(64 * waves), 0, s, t, prefix + v * (count + 1), count);
            return tfdtd_check_launch(what);
        };
        if (v == 0) return eps3 ? go(batch_tile2_kernel<true, 0, W, SNAP>, W, "batch march2 interior")
                                : go(batch_tile2_kernel<false, 0, W, SNAP>, W, "batch march2 interior");
        if (v == 1) return eps3 ? go(batch_tile2_kernel<true, 4, 12, SNAP>, 12, "batch march2 z faces")
                                : go(batch_tile2_kernel<false, 4, 12, SNAP>, 12, "batch march2 z faces");
        // Boundary tiles, three launches over the same workgroup range (each skips the others' tiles), one
        // after the other: the x end chunks' tiles that meet y or z slabs (every slab), the other x end
        // tiles (x slabs only) and the x-middle chunks' tiles (y and z slabs only); the fewer slabs a
        // launch compiles in, the fewer registers it spills.
        int status = eps3 ? go(batch_tile2_kernel<true, 7, 12, SNAP>, 12, "batch march2 boundary")
                          : go(batch_tile2_kernel<false, 7, 12, SNAP>, 12, "batch march2 boundary");
        if (!status) status = eps3 ? go(batch_tile2_kernel<true, 1, 12, SNAP>, 12, "batch march2 x boundary")
                                   : go(batch_tile2_kernel<false, 1, 12, SNAP>, 12, "batch march2 x boundary");
        if (status) return status;
        return eps3 ? go(batch_tile2_kernel<true, 6, 12, SNAP>, 12, "batch march2 y boundary")
                    : go(batch_tile2_kernel<false, 6, 12, SNAP>, 12, "batch march2 y boundary");
    });
}

int dispatch_tile2(const tfdtd_march* d, hipStream_t stream) {
    March<float> a;
    int status = layout_march2(d, &a);
    if (status) return status;
    const bool eps3 = d->eps_stride == 3;
    if (a.snap_count > 0) {
        switch (d->tile_y) {
            case 0: case 12: return launch_tile2<12, true>(a, eps3, stream);
            case 16: return launch_tile2<16, true>(a, eps3, stream);
            default: return tfdtd_fail(TFDTD_EINVAL, "march2: a step n+1 snapshot takes 12 or 16 wavefronts per interior tile");
        }
    }
    switch (d->tile_y) {
        case 4: return launch_tile2<4>(a, eps3, stream);
        case 6: return launch_tile2<6>(a, eps3, stream);
        case 0: case 12: return launch_tile2<12>(a, eps3, stream);
        case 16: return launch_tile2<16>(a, eps3, stream);
        default: return tfdtd_fail(TFDTD_EINVAL, "march2: tile_y (wavefronts of an interior workgroup) must be 4, 6, 12, 16 or 0");
    }
}

template <typename real, int W>
int pack_marches(const tfdtd_march* d, int count, void* host_table, size_t bytes, int32_t* prefix) {
    if (!host_table || bytes < sizeof(March<real>) * (size_t)count) return tfdtd_fail(TFDTD_EINVAL, "batch march: table buffer too small");
    March<real>* table = static_cast<March<real>*>(host_table);
    long long total[3] = {0, 0, 0};
    for (int k = 0; k < count; ++k) {
        if (d[k].dtype != d[0].dtype || d[k].tile_y != d[0].tile_y || d[k].eps_stride != d[0].eps_stride)
            return tfdtd_fail(TFDTD_EINVAL, "batch march: descriptor %d differs in dtype, tile_y or eps_stride", k);
        const int status = layout_march<real, W>(d + k, table + k);
        if (status) return status;
        for (int v = 0; v < 3; ++v) {
            prefix[v * (count + 1) + k] = (int32_t)total[v];
            total[v] += (table[k].blocks[v] + 7) / 8 * 8;
            if (total[v] > 0x7fffffffLL) return tfdtd_fail(TFDTD_EINVAL, "batch march: too many workgroups for one launch");
        }
    }
    for (int v = 0; v < 3; ++v) prefix[v * (count + 1) + count] = (int32_t)total[v];
    return TFDTD_OK;
}

template <typename real, int W, bool EPS3>
int launch_batch_tiles(const void* table, const int32_t* prefix, int count, const int32_t* totals, hipStream_t stream) {
    const March<real>* t = static_cast<const March<real>*>(table);
    const int blocks[3] = {totals[0], totals[1], totals[2]};
    return launch_variants(blocks, stream, [&](int v, hipStream_t s) {
        auto go = [&](auto kernel, const char* what) {
            hipLaunchKernelGGL(kernel, dim3((unsigned)blocks[v]), dim3(64 * W), 0, s, t, prefix + v * (count + 1), count);
            return tfdtd_check_launch(what);
        };
        if (v == 0) return go(batch_tile_kernel<real, W, EPS3, 0>, "batch march interior");
        if (v == 1) return go(batch_tile_kernel<real, W, EPS3, 4>, "batch march z faces");
        return go(batch_tile_kernel<real, W, EPS3, 7>, "batch march boundary");
    });
}

template <typename real>
int batch_march(const void* table, const int32_t* prefix, int count, int tile_y, int eps_stride, const int32_t* totals,
                hipStream_t stream) {
    if (tile_y == 8) return eps_stride == 3 ? launch_batch_tiles<real, 4, true>(table, prefix, count, totals, stream)
                                            : launch_batch_tiles<real, 4, false>(table, prefix, count, totals, stream);
    return eps_stride == 3 ? launch_batch_tiles<real, 8, true>(table, prefix, count, totals, stream)
                           : launch_batch_tiles<real, 8, false>(table, prefix, count, totals, stream);
}

template <typename real>
int dispatch_march(const tfdtd_march* d, hipStream_t stream) {
    switch (d->tile_y) {
        case 8: return launch_tile<real, 4>(d, stream);    // 4 wavefronts x 2 rows
        case 16: return launch_tile<real, 8>(d, stream);   // 8 wavefronts x 2 rows
        default: return tfdtd_fail(TFDTD_EINVAL, "march: tile_y must be 8 or 16");
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

extern "C" int tfdtd_yee_step_march2(const tfdtd_march* d, void* stream) {
    if (!d) return tfdtd_fail(TFDTD_EINVAL, "null march descriptor");
    if (d->dtype != TFDTD_F32) return tfdtd_fail(TFDTD_EINVAL, "march2: float32 only");
    return dispatch_tile2(d, static_cast<hipStream_t>(stream));
}

extern "C" int tfdtd_batch_march2_pack(const tfdtd_march* steps, int32_t count, void* host_table, size_t bytes,
                                       int32_t* block_prefix) {
    if (!steps || count < 1 || !block_prefix) return tfdtd_fail(TFDTD_EINVAL, "batch march2: at least one descriptor and a prefix array");
    return pack_marches2(steps, count, host_table, bytes, block_prefix);
}

extern "C" int tfdtd_batch_march2(const void* device_table, const int32_t* device_prefix, int32_t count, int32_t tile_y,
                                  int32_t eps_stride, int32_t snapshot, const int32_t* totals, void* stream) {
    if (!device_table || !device_prefix || count < 1 || !totals) return tfdtd_fail(TFDTD_EINVAL, "batch march2: bad launch arguments");
    if (eps_stride != 1 && eps_stride != 3) return tfdtd_fail(TFDTD_EINVAL, "march: eps_stride must be 1 or 3");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool eps3 = eps_stride == 3;
    if (snapshot) {
        switch (tile_y) {
            case 0: case 12: return launch_batch_tiles2<12, true>(device_table, device_prefix, count, eps3, totals, s);
            case 16: return launch_batch_tiles2<16, true>(device_table, device_prefix, count, eps3, totals, s);
            default: return tfdtd_fail(TFDTD_EINVAL, "march2: a step n+1 snapshot takes 12 or 16 wavefronts per interior tile");
        }
    }
    switch (tile_y) {
        case 4: return launch_batch_tiles2<4>(device_table, device_prefix, count, eps3, totals, s);
        case 6: return launch_batch_tiles2<6>(device_table, device_prefix, count, eps3, totals, s);
        case 0: case 12: return launch_batch_tiles2<12>(device_table, device_prefix, count, eps3, totals, s);
        case 16: return launch_batch_tiles2<16>(device_table, device_prefix, count, eps3, totals, s);
        default: return tfdtd_fail(TFDTD_EINVAL, "march2: tile_y (wavefronts of an interior workgroup) must be 4, 6, 12, 16 or 0");
    }
}

extern "C" size_t tfdtd_batch_march_bytes(int32_t count, int32_t dtype) {
    if (count < 0) return 0;
    if (dtype == TFDTD_F32) return sizeof(March<float>) * (size_t)count;
    if (dtype == TFDTD_F64) return sizeof(March<double>) * (size_t)count;
    return 0;
}

extern "C" int tfdtd_batch_march_pack(const tfdtd_march* steps, int32_t count, void* host_table, size_t bytes, int32_t* block_prefix) {
    if (!steps || count < 1 || !block_prefix) return tfdtd_fail(TFDTD_EINVAL, "batch march: at least one descriptor and a prefix array");
    const int ty = steps[0].tile_y;
    if (ty != 8 && ty != 16) return tfdtd_fail(TFDTD_EINVAL, "march: tile_y must be 8 or 16");
    if (steps[0].dtype == TFDTD_F32)
        return ty == 8 ? pack_marches<float, 4>(steps, count, host_table, bytes, block_prefix)
                       : pack_marches<float, 8>(steps, count, host_table, bytes, block_prefix);
    if (steps[0].dtype == TFDTD_F64)
        return ty == 8 ? pack_marches<double, 4>(steps, count, host_table, bytes, block_prefix)
                       : pack_marches<double, 8>(steps, count, host_table, bytes, block_prefix);
    return tfdtd_fail(TFDTD_EINVAL, "march: dtype must be float32 or float64");
}

extern "C" int tfdtd_batch_march(const void* device_table, const int32_t* device_prefix, int32_t count, int32_t dtype,
                                 int32_t tile_y, int32_t eps_stride, const int32_t* totals, void* stream) {
    if (!device_table || !device_prefix || count < 1 || !totals) return tfdtd_fail(TFDTD_EINVAL, "batch march: bad launch arguments");
    if (tile_y != 8 && tile_y != 16) return tfdtd_fail(TFDTD_EINVAL, "march: tile_y must be 8 or 16");
    if (eps_stride != 1 && eps_stride != 3) return tfdtd_fail(TFDTD_EINVAL, "march: eps_stride must be 1 or 3");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == TFDTD_F32) return batch_march<float>(device_table, device_prefix, count, tile_y, eps_stride, totals, s);
    if (dtype == TFDTD_F64) return batch_march<double>(device_table, device_prefix, count, tile_y, eps_stride, totals, s);
    return tfdtd_fail(TFDTD_EINVAL, "march: dtype must be float32 or float64");
}
