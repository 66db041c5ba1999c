// Discrete adjoint of the Yee/CPML step for gfx950, and the forward record it needs.
//
// The exact transpose of one step (torchfdtd/differentiable.py:395-447 and 496-519; the
// reference's fused form is torchfdtd/cuda_adjoint.py:16-234) runs backwards in three
// launches per step, each thread owning one cell and writing only that cell (gather form,
// no atomics):
//   seeds:        e_bar/h_bar += the observation adjoints of step t (CSR over targets);
//   h_transpose:  e_bar += curl_T(forward)(-(S*inv_mu) h_bar), the H-family memory adjoint
//                 psi_bar_prev = b*(psi_bar + d) from one buffer into the other, and the
//                 permittivity gradient -S * e_bar . curl_H(t) * inv_eps^2 on the design box;
//   e_transpose:  h_bar += curl_T(backward)((S*inv_eps) e_bar), with the E-family memory
//                 adjoint likewise; block 0 then steps the counter back.
// The curl of H that step t's E update used is recorded during the forward run on the
// design box (record_curl) in slot t % modulus, so the backward sweep needs no forward
// state at all -- only the recorded box.
#include <hip/hip_runtime.h>

#include "../../include/tfdtd.h"
#include "common.h"
#include "cpml_faces.h"

namespace {

template <typename real>
struct Adjoint {
    int n[3];
    real* e_bar;
    real* h_bar;
    const real* inv_eps;
    int eps_stride;
    const real* inv_mu;           // one value
    real S;                       // courant, rounded as the forward step rounds it
    int periodic, pec;
    real pec_factor[3];
    Family<real> ef, hf;          // psi_in: adjoint of the memory after the step, psi_out: before it
    int nseed;
    const long long* seed_targets;
    const int* seed_offsets;
    const int* seed_columns;
    const real* seeds;
    int seed_width;
    long long seed_rows;
    int box_lo[3], box_hi[3];
    const real* record;
    long long modulus;
    double* gradient;
    int gradient_stride;
    long long* counter;
};

__device__ __forceinline__ long long cell_index(const int n[3], int x, int y, int z) {
    return ((long long)x * n[1] + y) * n[2] + z;
}

// Offset of cell (x, y, z) inside a slab memory array of `axis` (extent w, slab row q).
__device__ __forceinline__ unsigned slab_offset(const int n[3], int axis, int q, int w, int x, int y, int z) {
    int c[3] = {x, y, z}, d[3] = {n[0], n[1], n[2]};
    c[axis] = q;
    d[axis] = w;
    return (unsigned)(((long long)c[0] * d[1] + c[1]) * d[2] + c[2]);
}

// d_D of one term at one row: d*inv_k + c*(psi_bar + d); the owner also stores b*(psi_bar + d).
template <typename real>
__device__ __forceinline__ real transpose_row(real d, const Family<real>& F, int axis, int which, int row,
                                              const int n[3], int x, int y, int z, bool owner) {
    const Hit h = find_hit(F, axis, row);
    if (h.face < 0) return d;
    const unsigned o = (which ? h.psi1 : h.psi0) + slab_offset(n, axis, h.q, h.w, x, y, z);
    const real b = F.coef[h.coef + h.q];
    const real c = F.coef[h.coef + h.w + h.q];
    const real ik = F.coef[h.coef + 2 * h.w + h.q];
    const real m = F.psi_in[o] + d;
    if (owner) F.psi_out[o] = b * m;
    return d * ik + c * m;
}

// The cell of this thread in the transpose launches: lanes along z, the block's four wavefronts
// along y, blocks over (z tile of 64, y tile of 4, x) -- the block's coordinates come from uniform
// 32-bit arithmetic, so no lane divides 64-bit indices.
constexpr int kTileZ = 64, kTileY = 4;

struct Cell3 {
    int x, y, z;
    long long j;
    bool inside;
};

__device__ __forceinline__ Cell3 locate(const int n[3]) {
    const int zt = (n[2] + kTileZ - 1) / kTileZ, yt = (n[1] + kTileY - 1) / kTileY;
    const int b = blockIdx.x;
    const int r = b / zt;
    Cell3 c;
    c.z = (b - r * zt) * kTileZ + (int)(threadIdx.x & (kTileZ - 1));
    c.x = r / yt;
    c.y = (r - c.x * yt) * kTileY + (int)(threadIdx.x / kTileZ);
    c.inside = c.x < n[0] && c.y < n[1] && c.z < n[2];
    c.j = ((long long)c.x * n[1] + c.y) * n[2] + c.z;
    return c;
}

inline unsigned transpose_blocks(const int n[3]) {
    return (unsigned)n[0] * (unsigned)((n[1] + kTileY - 1) / kTileY) * (unsigned)((n[2] + kTileZ - 1) / kTileZ);
}

template <typename real>
__global__ __launch_bounds__(256) void seed_kernel(const Adjoint<real> a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nseed) return;
    const long long t = *a.counter;
    if (t < 0 || t >= a.seed_rows) return;
    real sum = real(0);
    for (int j = a.seed_offsets[k]; j < a.seed_offsets[k + 1]; ++j) sum = sum + a.seeds[t * a.seed_width + a.seed_columns[j]];
    real* field = a.seed_targets[2 * k] ? a.h_bar : a.e_bar;
    field[a.seed_targets[2 * k + 1]] += sum;
}

// Both transposes read every field value they need (the cell and its neighbour along each axis)
// into registers before the first memory-array store: stores through psi_out would otherwise keep
// the compiler from reusing or hoisting later field loads.
template <typename real>
__global__ __launch_bounds__(256) void h_transpose_kernel(const Adjoint<real> a) {
    const Cell3 cell = locate(a.n);
    if (!cell.inside) return;
    const int x = cell.x, y = cell.y, z = cell.z;
    const long long j = cell.j;
    const int p[3] = {x, y, z};
    const real hcoef = a.S * a.inv_mu[0];
    real ob[3], down[3][3], eb[3];
    for (int o = 0; o < 3; ++o) ob[o] = -hcoef * a.h_bar[3 * j + o];
    for (int c = 0; c < 3; ++c) eb[c] = a.e_bar[3 * j + c];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        // Row i-1 of the axis, or row n-1 across the periodic wrap at i = 0.
        const int na = a.n[ax], i = p[ax];
        const long long stride = ax == 0 ? (long long)a.n[1] * a.n[2] : ax == 1 ? a.n[2] : 1;
        const bool has = na > 1 && (i >= 1 || (a.periodic >> ax & 1));
        const long long nb = i >= 1 ? j - stride : j + (long long)(na - 1) * stride;
        for (int o = 0; o < 3; ++o) down[ax][o] = has ? -hcoef * a.h_bar[3 * nb + o] : real(0);
    }
    real acc[3] = {real(0), real(0), real(0)};
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int ax = term_axis(t), c = term_in(t), o = term_out(t), s = term_sign(t);
        const int na = a.n[ax], i = p[ax];
        if (na == 1) continue;
        // Forward differences: rows 0 .. n-2 carry a derivative; E_c[i+1] += d_D[i], E_c[i] -= d_D[i].
        if (i <= na - 2) {
            const real d = transpose_row(real(s) * ob[o], a.hf, ax, t & 1, i, a.n, x, y, z, true);
            acc[c] = acc[c] - d;
        }
        if (i >= 1) {
            int q[3] = {x, y, z};
            q[ax] = i - 1;
            const real d = transpose_row(real(s) * down[ax][o], a.hf, ax, t & 1, i - 1, a.n, q[0], q[1], q[2], false);
            acc[c] = acc[c] + d;
        }
        if (a.periodic >> ax & 1) {
            // Wrap term at row n-1: E_c[0] += s*ob[n-1], E_c[n-1] -= s*ob[n-1].
            if (i == 0) acc[c] = acc[c] + real(s) * down[ax][o];
            if (i == na - 1) acc[c] = acc[c] - real(s) * ob[o];
        } else if ((a.pec >> ax & 1) && i == na - 1) {
            acc[c] = acc[c] - real(s) * a.pec_factor[ax] * ob[o];
        }
    }
    real e[3];
    for (int c = 0; c < 3; ++c) {
        e[c] = eb[c] + acc[c];
        a.e_bar[3 * j + c] = e[c];
    }
    const long long t = *a.counter;
    if (t >= 0 && x >= a.box_lo[0] && x < a.box_hi[0] && y >= a.box_lo[1] && y < a.box_hi[1] && z >= a.box_lo[2] && z < a.box_hi[2]) {
        const int bx = a.box_hi[0] - a.box_lo[0], by = a.box_hi[1] - a.box_lo[1], bz = a.box_hi[2] - a.box_lo[2];
        const long long b = ((long long)(x - a.box_lo[0]) * by + (y - a.box_lo[1])) * bz + (z - a.box_lo[2]);
        const real* curl = a.record + ((t % a.modulus) * ((long long)bx * by * bz) + b) * 3;
        double sum = 0.;
        for (int c = 0; c < 3; ++c) {
            const real inv = a.inv_eps[a.eps_stride == 3 ? 3 * j + c : j];
            const real g = -a.S * e[c] * curl[c] * (inv * inv);
            if (a.gradient_stride == 3) a.gradient[3 * b + c] += (double)g;
            else sum += (double)g;
        }
        if (a.gradient_stride == 1) a.gradient[b] += sum;
    }
}

template <typename real>
__global__ __launch_bounds__(256) void e_transpose_kernel(const Adjoint<real> a) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.counter = *a.counter - 1;
    const Cell3 cell = locate(a.n);
    if (!cell.inside) return;
    const int x = cell.x, y = cell.y, z = cell.z;
    const long long j = cell.j;
    const int p[3] = {x, y, z};
    auto obar = [&](long long at, int o) {
        const real inv = a.inv_eps[a.eps_stride == 3 ? 3 * at + o : at];
        return (a.S * inv) * a.e_bar[3 * at + o];
    };
    real own[3], up[3][3], hb[3];
    for (int o = 0; o < 3; ++o) own[o] = obar(j, o);
    for (int c = 0; c < 3; ++c) hb[c] = a.h_bar[3 * j + c];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        // Row i+1 of the axis, or row 0 across the periodic wrap at i = n-1.
        const int na = a.n[ax], i = p[ax];
        const long long stride = ax == 0 ? (long long)a.n[1] * a.n[2] : ax == 1 ? a.n[2] : 1;
        const bool has = na > 1 && (i + 1 <= na - 1 || (a.periodic >> ax & 1));
        const long long nb = i + 1 <= na - 1 ? j + stride : j - (long long)(na - 1) * stride;
        for (int o = 0; o < 3; ++o) up[ax][o] = has ? obar(nb, o) : real(0);
    }
    real acc[3] = {real(0), real(0), real(0)};
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int ax = term_axis(t), c = term_in(t), o = term_out(t), s = term_sign(t);
        const int na = a.n[ax], i = p[ax];
        if (na == 1) continue;
        // Backward differences: rows 1 .. n-1 carry a derivative; H_c[i] += d_D[i], H_c[i-1] -= d_D[i].
        if (i >= 1) {
            const real d = transpose_row(real(s) * own[o], a.ef, ax, t & 1, i, a.n, x, y, z, true);
            acc[c] = acc[c] + d;
        }
        if (i + 1 <= na - 1) {
            int q[3] = {x, y, z};
            q[ax] = i + 1;
            const real d = transpose_row(real(s) * up[ax][o], a.ef, ax, t & 1, i + 1, a.n, q[0], q[1], q[2], false);
            acc[c] = acc[c] - d;
        }
        if (a.periodic >> ax & 1) {
            // Wrap term at row 0: H_c[0] += s*ob[0], H_c[n-1] -= s*ob[0].
            if (i == 0) acc[c] = acc[c] + real(s) * own[o];
            if (i == na - 1) acc[c] = acc[c] - real(s) * up[ax][o];
        }
    }
    for (int c = 0; c < 3; ++c) a.h_bar[3 * j + c] = hb[c] + acc[c];
}

template <typename real>
struct Record {
    int n[3];
    const real* H;
    int periodic;
    int box_lo[3], box_hi[3];
    real* record;
    long long modulus;
    const long long* counter;
    int cpml;                     // 1: apply the E-family slabs of ef (psi_in = memory before the step)
    Family<real> ef;
};

// curl_H of the E update (backward differences, terms in curl order) on the design box; in an
// E-family slab row a term is d inv_k + (b psi + c d), the value the update adds there.
template <typename real>
__global__ __launch_bounds__(256) void record_kernel(const Record<real> a) {
    const int bx = a.box_hi[0] - a.box_lo[0], by = a.box_hi[1] - a.box_lo[1], bz = a.box_hi[2] - a.box_lo[2];
    const long long count = (long long)bx * by * bz;
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= count) return;
    const int z = a.box_lo[2] + (int)(b % bz), y = a.box_lo[1] + (int)((b / bz) % by), x = a.box_lo[0] + (int)(b / ((long long)by * bz));
    const int p[3] = {x, y, z};
    const long long j = cell_index(a.n, x, y, z);
    real curl[3] = {real(0), real(0), real(0)};
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int ax = term_axis(t), c = term_in(t), o = term_out(t), s = term_sign(t);
        const int na = a.n[ax];
        if (na == 1) continue;
        const long long stride = ax == 0 ? (long long)a.n[1] * a.n[2] : ax == 1 ? a.n[2] : 1;
        real d;
        if (p[ax] >= 1) d = a.H[3 * j + c] - a.H[3 * (j - stride) + c];
        else if (a.periodic >> ax & 1) d = a.H[3 * j + c] - a.H[3 * (j + (long long)(na - 1) * stride) + c];
        else continue;
        if (a.cpml) {
            const Hit h = find_hit(a.ef, ax, p[ax]);
            if (h.face >= 0) {
                const unsigned m = ((t & 1) ? h.psi1 : h.psi0) + slab_offset(a.n, ax, h.q, h.w, x, y, z);
                const real b = a.ef.coef[h.coef + h.q];
                const real cc = a.ef.coef[h.coef + h.w + h.q];
                const real ik = a.ef.coef[h.coef + 2 * h.w + h.q];
                const real psi = b * a.ef.psi_in[m] + cc * d;
                d = d * ik + psi;
            }
        }
        curl[o] = curl[o] + real(s) * d;
    }
    real* out = a.record + ((*a.counter % a.modulus) * count + b) * 3;
    for (int o = 0; o < 3; ++o) out[o] = curl[o];
}

int check_box(const int32_t lo[3], const int32_t hi[3], const int n[3]) {
    for (int ax = 0; ax < 3; ++ax)
        if (lo[ax] < 0 || hi[ax] <= lo[ax] || hi[ax] > n[ax]) return tfdtd_fail(TFDTD_EINVAL, "adjoint: design box outside the grid");
    return TFDTD_OK;
}

template <typename real>
int make_adjoint(const tfdtd_adjoint* d, Adjoint<real>* a) {
    std::memset(a, 0, sizeof *a);
    const int n[3] = {d->nx, d->ny, d->nz};
    if (n[0] < 1 || n[1] < 1 || n[2] < 1) return tfdtd_fail(TFDTD_EINVAL, "adjoint: grid dimensions must be positive");
    if (!d->e_bar || !d->h_bar || !d->inverse_permittivity || !d->inverse_permeability || !d->counter)
        return tfdtd_fail(TFDTD_EINVAL, "adjoint: null field, material or counter pointer");
    if (d->eps_stride != 1 && d->eps_stride != 3) return tfdtd_fail(TFDTD_EINVAL, "adjoint: eps_stride must be 1 or 3");
    for (int ax = 0; ax < 3; ++ax) a->n[ax] = n[ax];
    a->e_bar = static_cast<real*>(d->e_bar);
    a->h_bar = static_cast<real*>(d->h_bar);
    a->inv_eps = static_cast<const real*>(d->inverse_permittivity);
    a->eps_stride = d->eps_stride;
    return TFDTD_OK;
}

}  // namespace

extern "C" int tfdtd_record_curl(const tfdtd_curl_record* d, void* stream) {
    if (!d || !d->H || !d->record || !d->counter || d->modulus < 1) return tfdtd_fail(TFDTD_EINVAL, "record: null pointer or modulus < 1");
    const int n[3] = {d->nx, d->ny, d->nz};
    if (n[0] < 1 || n[1] < 1 || n[2] < 1) return tfdtd_fail(TFDTD_EINVAL, "record: grid dimensions must be positive");
    int status = check_box(d->box_lo, d->box_hi, n);
    if (status) return status;
    const long long count = (long long)(d->box_hi[0] - d->box_lo[0]) * (d->box_hi[1] - d->box_lo[1]) * (d->box_hi[2] - d->box_lo[2]);
    const unsigned blocks = (unsigned)((count + 255) / 256);
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto fill = [&](auto* a) -> int {
        using real = std::remove_pointer_t<decltype(a->record)>;
        for (int ax = 0; ax < 3; ++ax) { a->n[ax] = n[ax]; a->box_lo[ax] = d->box_lo[ax]; a->box_hi[ax] = d->box_hi[ax]; }
        a->H = static_cast<const real*>(d->H);
        a->periodic = d->periodic_mask & 7;
        a->record = static_cast<real*>(d->record);
        a->modulus = d->modulus;
        a->counter = reinterpret_cast<const long long*>(d->counter);
        a->cpml = d->e.psi_in != nullptr;
        if (!a->cpml) {
            std::memset(&a->ef, 0, sizeof a->ef);
            return TFDTD_OK;
        }
        // The record reads psi_in only: in-place (split step) and double-buffered layouts both serve.
        return make_family(d->e, d->e.psi_in != d->e.psi_out, n, &a->ef);
    };
    if (d->dtype == TFDTD_F32) {
        Record<float> a;
        if ((status = fill(&a))) return status;
        hipLaunchKernelGGL(record_kernel<float>, dim3(blocks), dim3(256), 0, s, a);
    } else if (d->dtype == TFDTD_F64) {
        Record<double> a;
        if ((status = fill(&a))) return status;
        hipLaunchKernelGGL(record_kernel<double>, dim3(blocks), dim3(256), 0, s, a);
    } else {
        return tfdtd_fail(TFDTD_EINVAL, "record: dtype must be float32 or float64");
    }
    return tfdtd_check_launch("record_curl");
}

template <typename real>
static int launch_adjoint(const tfdtd_adjoint* d, hipStream_t s) {
    Adjoint<real> a;
    int status = make_adjoint(d, &a);
    if (status) return status;
    const int dims[3] = {d->nx, d->ny, d->nz};
    status = make_family(d->e, true, dims, &a.ef);
    if (!status) status = make_family(d->h, true, dims, &a.hf);
    if (status) return status;
    a.inv_mu = static_cast<const real*>(d->inverse_permeability);
    a.S = (real)d->courant;
    a.periodic = d->periodic_mask & 7;
    a.pec = d->pec_upper_mask & 7 & ~a.periodic;
    for (int ax = 0; ax < 3; ++ax) a.pec_factor[ax] = (real)d->pec_upper_factor[ax];
    a.nseed = d->nseed;
    if (d->nseed < 0 || (d->nseed > 0 && (!d->seed_targets || !d->seed_offsets || !d->seed_columns || !d->seeds || d->seed_width < 1)))
        return tfdtd_fail(TFDTD_EINVAL, "adjoint: malformed seed table");
    a.seed_targets = reinterpret_cast<const long long*>(d->seed_targets);
    a.seed_offsets = d->seed_offsets;
    a.seed_columns = d->seed_columns;
    a.seeds = static_cast<const real*>(d->seeds);
    a.seed_width = d->seed_width;
    a.seed_rows = d->seed_rows;
    if (d->record) {
        status = check_box(d->box_lo, d->box_hi, dims);
        if (status) return status;
        if (!d->gradient || d->record_modulus < 1 || (d->gradient_stride != 1 && d->gradient_stride != 3))
            return tfdtd_fail(TFDTD_EINVAL, "adjoint: a record needs a gradient buffer, a modulus >= 1 and gradient_stride 1 or 3");
        // A plain difference record serves only boxes that stay out of the E-family slabs.
        for (int ax = 0; ax < 3 && !d->record_cpml; ++ax)
            for (int f = 0; f < 2; ++f) {
                const Face& face = a.ef.face[ax][f];
                if (face.hi > face.lo && face.lo < d->box_hi[ax] && d->box_lo[ax] < face.hi)
                    return tfdtd_fail(TFDTD_EINVAL, "adjoint: the design box meets the CPML slab of axis %d and the "
                                                    "record has no CPML terms", ax);
            }
        for (int ax = 0; ax < 3; ++ax) { a.box_lo[ax] = d->box_lo[ax]; a.box_hi[ax] = d->box_hi[ax]; }
        a.record = static_cast<const real*>(d->record);
        a.modulus = d->record_modulus;
        a.gradient = d->gradient;
        a.gradient_stride = d->gradient_stride;
    } else {
        for (int ax = 0; ax < 3; ++ax) { a.box_lo[ax] = 0; a.box_hi[ax] = 0; }   // no gradient cells
        a.modulus = 1;
    }
    a.counter = reinterpret_cast<long long*>(d->counter);
    const unsigned long long tiles = (unsigned long long)dims[0] * ((dims[1] + kTileY - 1) / kTileY) *
                                     ((dims[2] + kTileZ - 1) / kTileZ);
    if (tiles > 0x7fffffffULL) return tfdtd_fail(TFDTD_EINVAL, "adjoint: grid too large for one launch");
    const unsigned blocks = transpose_blocks(dims);
    if (a.nseed > 0) {
        hipLaunchKernelGGL(seed_kernel<real>, dim3((unsigned)((a.nseed + 255) / 256)), dim3(256), 0, s, a);
        status = tfdtd_check_launch("adjoint seeds");
        if (status) return status;
    }
    hipLaunchKernelGGL(h_transpose_kernel<real>, dim3(blocks), dim3(256), 0, s, a);
    status = tfdtd_check_launch("adjoint h_transpose");
    if (status) return status;
    hipLaunchKernelGGL(e_transpose_kernel<real>, dim3(blocks), dim3(256), 0, s, a);
    return tfdtd_check_launch("adjoint e_transpose");
}

extern "C" int tfdtd_adjoint_step(const tfdtd_adjoint* d, void* stream) {
    if (!d) return tfdtd_fail(TFDTD_EINVAL, "null adjoint descriptor");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (d->dtype == TFDTD_F32) return launch_adjoint<float>(d, s);
    if (d->dtype == TFDTD_F64) return launch_adjoint<double>(d, s);
    return tfdtd_fail(TFDTD_EINVAL, "adjoint: dtype must be float32 or float64");
}
