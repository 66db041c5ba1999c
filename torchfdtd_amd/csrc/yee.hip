// Yee/CPML half steps, point traces and plane DFTs for gfx950, behind the C ABI of
// include/tfdtd.h.
//
// Arithmetic follows the reference update (torchfdtd/boundaries.py:363-438) term by
// term: every product and sum is rounded to the field precision in the same order,
// and the library is built with -ffp-contract=off and denormals kept, so a step is
// reproducible bit for bit against the NumPy/Torch path of the same precision.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <string>

#include "../../include/tfdtd.h"
#include "common.h"

namespace {

using ::tfdtd_fail;

int fail(int code, const char* msg) { return tfdtd_fail(code, "%s", msg); }

template <typename... A>
int fail(int code, const char* fmt, A... args) { return tfdtd_fail(code, fmt, args...); }

int check_launch(const char* what) { return tfdtd_check_launch(what); }

template <typename real>
struct Slab {          // one CPML segment in its fixed slot 2*term+side
    int lo, hi;        // empty slot: lo == hi == 0
    real* psi;
    const real* b;
    const real* c;
    const real* inv_k;
};

template <typename real>
struct Source {
    int comp;
    int lo[3], hi[3];
    const real* values;
};

template <typename real>
struct HalfStep {
    int nx, ny, nz;
    const real* src;
    real* dst;
    const real* inverse;
    int inverse_stride;
    real courant;
    int periodic_mask, pec_mask;
    real pec_factor[3];
    Slab<real> slab[12];
    int nsrc;
    int box_lo[3], box_hi[3];   // union of the source boxes
    Source<real> srcs[TFDTD_MAX_SOURCES];
    long long source_samples;
    const long long* counter;
};

// The three components of one cell, read as one 12-byte (24-byte) access through a global pointer.
template <typename real>
struct Cell {
    real v[3];
};

template <typename real>
__device__ __forceinline__ Cell<real> cell_at(const real* base, long long i) {
    const auto p = global(base + 3 * i);
    return Cell<real>{{p[0], p[1], p[2]}};
}

// One cell.  forward=false: E update from backward differences of H;
// forward=true: H update from forward differences of E.  The cell and its three neighbours
// (x, y, z, one row back or ahead) are read whole; the terms then follow CURL_TERMS order.
template <typename real, bool forward>
__device__ __forceinline__ void half_step_cell(const HalfStep<real>& a, const long long i) {
    const int nyz = a.ny * a.nz;
    const int x = (int)(i / nyz);
    const int r = (int)(i - (long long)x * nyz);
    const int y = r / a.nz;
    const int z = r - y * a.nz;
    const int coord[3] = {x, y, z};
    const int dims[3] = {a.nx, a.ny, a.nz};
    const long long stride[3] = {(long long)nyz, (long long)a.nz, 1};
    const Cell<real> c = cell_at(a.src, i);
    Cell<real> nb[3];
    bool inside[3];
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        const int n = dims[axis], k = coord[axis];
        const long long s = stride[axis];
        inside[axis] = n > 1 && (forward ? (k < n - 1) : (k > 0));
        nb[axis] = Cell<real>{{real(0), real(0), real(0)}};
        if (inside[axis]) nb[axis] = cell_at(a.src, forward ? i + s : i - s);
        else if (n > 1 && ((a.periodic_mask >> axis) & 1)) {
            const long long span = (long long)(n - 1) * s;
            nb[axis] = cell_at(a.src, forward ? i - span : i + span);
        }
    }
    // CPML memory.  The two slabs of a term never overlap (tfdtd_assign_slot), so a cell meets at most
    // one slab per term.  Every memory value and coefficient the cell needs is read here and the new
    // memory is stored after the last load: a load placed after a store would wait for it (the
    // compiler cannot tell the arrays apart), one round trip per term.
    bool hit[6];
    real* at[6];
    real old_psi[6], cb[6], cc[6], cik[6], memory[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int axis = term_axis(t);
        hit[t] = false;
        at[t] = nullptr;
        old_psi[t] = cb[t] = cc[t] = cik[t] = memory[t] = real(0);
        if (dims[axis] == 1 || !inside[axis]) continue;
        const int k = coord[axis];
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const Slab<real>& g = a.slab[2 * t + side];
            if (k >= g.lo && k < g.hi) {
                const int q = k - g.lo;
                const int w = g.hi - g.lo;
                long long p;
                if (axis == 0) p = ((long long)q * a.ny + y) * a.nz + z;
                else if (axis == 1) p = ((long long)x * w + q) * a.nz + z;
                else p = ((long long)x * a.ny + y) * w + q;
                hit[t] = true;
                at[t] = g.psi + p;
                old_psi[t] = *global(at[t]);
                cb[t] = global(g.b)[q];
                cc[t] = global(g.c)[q];
                cik[t] = global(g.inv_k)[q];
            }
        }
    }
    real curl[3] = {real(0), real(0), real(0)};
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int axis = term_axis(t), ic = term_in(t), oc = term_out(t);
        const int n = dims[axis];
        if (n == 1) continue;
        real d = real(0);
        if (inside[axis]) {
            d = forward ? (nb[axis].v[ic] - c.v[ic]) : (c.v[ic] - nb[axis].v[ic]);
            if (hit[t]) {
                memory[t] = old_psi[t] * cb[t] + cc[t] * d;
                d = d * cik[t] + memory[t];
            }
        } else if ((a.periodic_mask >> axis) & 1) {
            d = forward ? (nb[axis].v[ic] - c.v[ic]) : (c.v[ic] - nb[axis].v[ic]);
        } else if (forward && ((a.pec_mask >> axis) & 1)) {
            d = -a.pec_factor[axis] * c.v[ic];
        }
        curl[oc] = term_sign(t) > 0 ? curl[oc] + d : curl[oc] - d;
    }
    const Cell<real> old = cell_at(a.dst, i);
    Cell<real> inv;
    if (a.inverse_stride == 3) inv = cell_at(a.inverse, i);
    else {
        const real v = global(a.inverse)[a.inverse_stride == 1 ? i : 0];
        inv = Cell<real>{{v, v, v}};
    }
    real out[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const real coefficient = a.courant * inv.v[m];
        out[m] = forward ? old.v[m] - coefficient * curl[m] : old.v[m] + coefficient * curl[m];
    }
    if (a.nsrc > 0 && x >= a.box_lo[0] && x < a.box_hi[0] && y >= a.box_lo[1] && y < a.box_hi[1] &&
        z >= a.box_lo[2] && z < a.box_hi[2]) {
        const long long n = *global(a.counter);
        for (int k = 0; k < a.nsrc && n < a.source_samples; ++k) {
            const Source<real>& box = a.srcs[k];
            if (x >= box.lo[0] && x < box.hi[0] && y >= box.lo[1] && y < box.hi[1] && z >= box.lo[2] &&
                z < box.hi[2]) {
                const real v = global(box.values)[n];
                if (box.comp == 0) out[0] = out[0] + v;
                else if (box.comp == 1) out[1] = out[1] + v;
                else out[2] = out[2] + v;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 6; ++t)
        if (hit[t]) *global(at[t]) = memory[t];
    const auto cell = global(a.dst + 3 * i);
    cell[0] = out[0];
    cell[1] = out[1];
    cell[2] = out[2];
}

template <typename real, bool forward>
__global__ __launch_bounds__(256) void yee_half_step_kernel(const HalfStep<real> a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)a.nx * a.ny * a.nz) return;
    half_step_cell<real, forward>(a, i);
}

// Cells per thread of the batched half step: a workgroup covers 256 * kBatchSpan cells of one grid, so
// the search for its grid and the descriptor loads are shared by more cells (2: 2% faster on the
// 64-grid uneven batch than 1 or 4).
constexpr int kBatchSpan = 2;

// Many grids in one launch: block b belongs to the last grid g with prefix[g] <= b; the grids'
// descriptors are a table in device memory, read with wave-uniform loads.
template <typename real, bool forward>
__global__ __launch_bounds__(256) void batch_half_step_kernel(const HalfStep<real>* __restrict__ table,
                                                              const int* __restrict__ prefix, int grids) {
    const int b = blockIdx.x;
    int lo = 0, hi = grids;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= b) lo = mid;
        else hi = mid;
    }
    const HalfStep<real>& a = table[lo];
    const long long cells = (long long)a.nx * a.ny * a.nz;
    const long long base = (long long)(b - prefix[lo]) * (256 * kBatchSpan) + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kBatchSpan; ++j) {
        const long long i = base + 256 * j;
        if (i < cells) half_step_cell<real, forward>(a, i);
    }
}

template <typename real>
struct Observe {
    int count;
    const long long* table;
    real* traces;
    long long steps;
    long long* counter;
};

// Single block: record the traces of this step, then advance the counter.
template <typename real>
__global__ __launch_bounds__(256) void observe_kernel(const Observe<real> a) {
    const long long n = *a.counter;
    __syncthreads();
    if (n < a.steps) {
        for (int k = threadIdx.x; k < a.count; k += blockDim.x) {
            const auto field = global(reinterpret_cast<const real*>(a.table[2 * k]));
            a.traces[n * a.count + k] = field[a.table[2 * k + 1]];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) a.counter[0] = n + 1;
}

// Many traces: one trace per thread over the grid (the counter advances in a launch of its own).
template <typename real>
__global__ __launch_bounds__(256) void observe_wide_kernel(const Observe<real> a) {
    const long long n = *a.counter;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.steps || k >= a.count) return;
    const auto field = global(reinterpret_cast<const real*>(a.table[2 * k]));
    a.traces[n * a.count + k] = field[a.table[2 * k + 1]];
}

__global__ void advance_counter_kernel(long long* counter) { counter[0] = counter[0] + 1; }

template <typename real, typename spec>
struct Plane {
    long long points;
    int nf, nc, ne, corners, downsample;
    const real* E;
    const real* H;
    const long long* indices[6];
    const real* weights[6];
    const real* window_e;
    const real* window_h;
    const double* omega;
    const spec* half;   // interleaved re, im
    spec* value;        // interleaved re, im
    long long samples;
    const long long* counter;
    int block;          // > 1: samples of `block` steps are staged and accumulated together
    real* stage;        // [block][points * nc]
    const long long* flushed;
};

// exp(i omega (step+1)) for E and the same times exp(i pi f dt) for H, from the absolute step in
// FP64 (cuda_monitors.py:68-84): p[0..1] E (re, im), p[2..3] H.
template <typename real, typename spec>
__device__ __forceinline__ void phases(const Plane<real, spec>& a, long long step, int f, spec* p) {
    double si, co;
    sincos(global(a.omega)[f] * (double)(step + 1), &si, &co);
    const spec er = (spec)co, ei = (spec)si;
    const spec hr = global(a.half)[2 * f], hi = global(a.half)[2 * f + 1];
    p[0] = er;
    p[1] = ei;
    p[2] = er * hr - ei * hi;
    p[3] = er * hi + ei * hr;
}

// Interpolated sample of (point, component) i at step n, times the window.
template <typename real, typename spec>
__device__ __forceinline__ real plane_sample(const Plane<real, spec>& a, long long i, long long n) {
    const int c = (int)(i % a.nc);
    const long long p = i / a.nc;
    const bool electric = c < a.ne;
    const auto field = global(electric ? a.E : a.H);
    const auto idx = global(a.indices[c]);
    const auto w = global(a.weights[c]);
    real v = real(0);
    if (a.corners == 8 || a.corners == 4) {
        // Every corner's index and weight first, then every corner's field value, then the sum in
        // corner order: the loads overlap instead of forming one round trip per corner.
        long long at[8];
        real wt[8], f[8];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (j < a.corners) {
                at[j] = idx[j * a.points + p];
                wt[j] = w[j * a.points + p];
            }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (j < a.corners) f[j] = field[at[j]];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (j < a.corners) v = v + f[j] * wt[j];
    } else {
        for (int j = 0; j < a.corners; ++j) v = v + field[idx[j * a.points + p]] * w[j * a.points + p];
    }
    return v * global(electric ? a.window_e : a.window_h)[n];
}

// Accumulate the staged samples of steps [lo, hi) (one block at most): the phases of those steps go
// to LDS, each thread sums its sample's products in FP64 and adds the sum to its accumulators.
template <typename real, typename spec>
__device__ __forceinline__ void flush_range(const Plane<real, spec>& a, long long i, long long lo, long long hi) {
    extern __shared__ unsigned char lds_raw[];
    spec* phase = reinterpret_cast<spec*>(lds_raw);
    const int count = (int)(hi - lo);
    for (int e = threadIdx.x; e < count * a.nf; e += blockDim.x) phases(a, lo + e / a.nf, e % a.nf, phase + 4 * e);
    __syncthreads();
    const long long samples = a.points * a.nc;
    if (i >= samples) return;
    const int o = (int)(i % a.nc) < a.ne ? 0 : 2;
    for (int f = 0; f < a.nf; ++f) {
        double sr = 0., si = 0.;
        for (int j = 0; j < count; ++j) {
            const long long step = lo + j;
            if (step % a.downsample) continue;
            const spec v = (spec)global(a.stage)[(step % a.block) * samples + i];
            const spec* p = phase + 4 * (j * a.nf + f) + o;
            sr += (double)(p[0] * v);
            si += (double)(p[1] * v);
        }
        const auto out = global(a.value + 2 * ((long long)f * samples + i));
        out[0] = out[0] + (spec)sr;
        out[1] = out[1] + (spec)si;
    }
}

// One thread per (point, component) sample.  Per-step accumulation (block <= 1): the block first
// evaluates the nf phases of this step into LDS.  Blocked accumulation: the sample is staged and
// the block's last step (or the window's last sample) accumulates the staged block.
template <typename real, typename spec>
__device__ __forceinline__ void dft_plane_block(const Plane<real, spec>& a, long long block) {
    extern __shared__ unsigned char lds_raw[];
    spec* phase = reinterpret_cast<spec*>(lds_raw);
    const long long n = *global(a.counter);
    if (n >= a.samples) return;
    const long long samples = a.points * a.nc;
    const long long i = block * blockDim.x + threadIdx.x;
    if (a.block > 1) {
        if (n % a.downsample == 0 && i < samples) global(a.stage)[(n % a.block) * samples + i] = plane_sample(a, i, n);
        if ((n + 1) % a.block == 0 || n + 1 == a.samples) {
            const long long first = n - n % a.block;
            const long long flushed = *global(a.flushed);
            flush_range(a, i, first > flushed ? first : flushed, n + 1);
        }
        return;
    }
    if (n % a.downsample) return;
    for (int f = threadIdx.x; f < a.nf; f += blockDim.x) phases(a, n, f, phase + 4 * f);
    __syncthreads();
    if (i >= samples) return;
    const spec sv = (spec)plane_sample(a, i, n);
    const int o = (int)(i % a.nc) < a.ne ? 0 : 2;
    for (int f = 0; f < a.nf; ++f) {
        const auto out = global(a.value + 2 * ((long long)f * samples + i));
        out[0] = out[0] + phase[4 * f + o] * sv;
        out[1] = out[1] + phase[4 * f + o + 1] * sv;
    }
}

// Accumulate the staged steps of the unfinished block: [max(flushed, first of block), counter).
template <typename real, typename spec>
__global__ __launch_bounds__(256) void dft_flush_kernel(const Plane<real, spec> a) {
    const long long n = *a.counter < a.samples ? *a.counter : a.samples;
    if (a.block <= 1 || n == a.samples) return;     // a finished window was accumulated by its last step
    const long long first = n - n % a.block;
    const long long lo = first > *a.flushed ? first : *a.flushed;
    if (lo >= n) return;
    flush_range(a, (long long)blockIdx.x * blockDim.x + threadIdx.x, lo, n);
}

__global__ void set_flushed_kernel(const long long* counter, long long samples, long long* flushed) {
    flushed[0] = counter[0] < samples ? counter[0] : samples;
}

// 8 waves per SIMD (at most 64 VGPRs): the sampling is latency-bound, and the 64-grid batch and the
// observed 512^3 cube ran 0.3-0.4% faster than at the 75 VGPRs the flush path asks for otherwise.
template <typename real, typename spec>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void dft_plane_kernel(const Plane<real, spec> a) {
    dft_plane_block(a, blockIdx.x);
}

// Many planes in one launch: blockIdx.y selects the plane of a device table.
template <typename real, typename spec>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void batch_dft_kernel(const Plane<real, spec>* __restrict__ table) {
    const Plane<real, spec>& a = table[blockIdx.y];
    if ((long long)blockIdx.x * blockDim.x >= a.points * a.nc) return;
    dft_plane_block(a, blockIdx.x);
}

// ---------------------------------------------------------------- host side

template <typename real>
int make_half_step(const tfdtd_half_step* d, HalfStep<real>* a) {
    if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(TFDTD_EINVAL, "grid dimensions must be positive");
    if (!d->src || !d->dst || !d->inverse) return fail(TFDTD_EINVAL, "field and coefficient pointers are required");
    if (d->inverse_stride != 0 && d->inverse_stride != 1 && d->inverse_stride != 3)
        return fail(TFDTD_EINVAL, "inverse_stride must be 0, 1 or 3");
    if (d->nseg < 0 || d->nseg > TFDTD_MAX_SEGMENTS) return fail(TFDTD_EINVAL, "too many CPML segments");
    if (d->nsrc < 0 || d->nsrc > TFDTD_MAX_SOURCES) return fail(TFDTD_EINVAL, "too many sources");
    std::memset(a, 0, sizeof *a);
    a->nx = d->nx; a->ny = d->ny; a->nz = d->nz;
    a->src = static_cast<const real*>(d->src);
    a->dst = static_cast<real*>(d->dst);
    a->inverse = static_cast<const real*>(d->inverse);
    a->inverse_stride = d->inverse_stride;
    a->courant = (real)d->courant;
    a->periodic_mask = d->periodic_mask & 7;
    a->pec_mask = d->pec_upper_mask & 7;
    for (int k = 0; k < 3; ++k) a->pec_factor[k] = (real)d->pec_upper_factor[k];
    const int dims[3] = {d->nx, d->ny, d->nz};
    int slot_lo[12] = {0}, slot_hi[12] = {0};
    for (int k = 0; k < d->nseg; ++k) {
        const tfdtd_cpml_segment& s = d->seg[k];
        const int slot = tfdtd_assign_slot(s.axis, s.comp, s.lo, s.hi, dims, k, slot_lo, slot_hi);
        if (slot < 0) return TFDTD_EINVAL;
        if (!s.psi || !s.b || !s.c || !s.inv_k) return fail(TFDTD_EINVAL, "segment %d: null pointer", k);
        a->slab[slot] = Slab<real>{s.lo, s.hi, static_cast<real*>(s.psi), static_cast<const real*>(s.b),
                                   static_cast<const real*>(s.c), static_cast<const real*>(s.inv_k)};
    }
    a->nsrc = d->nsrc;
    for (int k = 0; k < 3; ++k) { a->box_lo[k] = dims[k]; a->box_hi[k] = 0; }
    for (int k = 0; k < d->nsrc; ++k) {
        const tfdtd_source& s = d->srcs[k];
        if (s.comp < 0 || s.comp > 2 || !s.values) return fail(TFDTD_EINVAL, "source %d: bad component or values", k);
        Source<real>& o = a->srcs[k];
        o.comp = s.comp;
        o.values = static_cast<const real*>(s.values);
        for (int ax = 0; ax < 3; ++ax) {
            if (s.lo[ax] < 0 || s.hi[ax] <= s.lo[ax] || s.hi[ax] > dims[ax]) return fail(TFDTD_EINVAL, "source %d: box outside the grid", k);
            o.lo[ax] = s.lo[ax]; o.hi[ax] = s.hi[ax];
            if (s.lo[ax] < a->box_lo[ax]) a->box_lo[ax] = s.lo[ax];
            if (s.hi[ax] > a->box_hi[ax]) a->box_hi[ax] = s.hi[ax];
        }
    }
    if (d->nsrc > 0 && (!d->counter || d->source_samples < 0))
        return fail(TFDTD_EINVAL, "sources need the device step counter and their sample count");
    a->source_samples = d->source_samples;
    a->counter = reinterpret_cast<const long long*>(d->counter);
    return TFDTD_OK;
}

template <typename real>
int launch_half_step(const tfdtd_half_step* d, hipStream_t stream) {
    HalfStep<real> a;
    int status = make_half_step(d, &a);
    if (status) return status;
    const long long count = (long long)d->nx * d->ny * d->nz;
    const long long blocks = (count + 255) / 256;
    if (blocks > 0x7fffffffLL) return fail(TFDTD_EINVAL, "grid too large for one launch");
    if (d->family == TFDTD_FAMILY_E)
        hipLaunchKernelGGL((yee_half_step_kernel<real, false>), dim3((unsigned)blocks), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL((yee_half_step_kernel<real, true>), dim3((unsigned)blocks), dim3(256), 0, stream, a);
    return check_launch("yee_half_step");
}

template <typename real>
int launch_observe(const tfdtd_observe* d, hipStream_t stream) {
    if (!d->counter) return fail(TFDTD_EINVAL, "observe needs the device step counter");
    if (d->count < 0 || (d->count > 0 && (!d->table || !d->traces)))
        return fail(TFDTD_EINVAL, "observe: trace table and traces are required");
    Observe<real> a{d->count, reinterpret_cast<const long long*>(d->table), static_cast<real*>(d->traces),
                    (long long)d->steps, reinterpret_cast<long long*>(d->counter)};
    constexpr int kOneBlock = 4096;
    if (d->count <= kOneBlock) {
        hipLaunchKernelGGL((observe_kernel<real>), dim3(1), dim3(256), 0, stream, a);
        return check_launch("observe");
    }
    hipLaunchKernelGGL((observe_wide_kernel<real>), dim3((unsigned)((d->count + 255) / 256)), dim3(256), 0, stream, a);
    const int status = check_launch("observe");
    if (status) return status;
    hipLaunchKernelGGL(advance_counter_kernel, dim3(1), dim3(1), 0, stream, a.counter);
    return check_launch("observe counter");
}

template <typename real, typename spec>
int make_plane(const tfdtd_plane* d, Plane<real, spec>* out);

// LDS of one plane launch: the phases of one step, or of every step of a block.
template <typename spec>
size_t plane_lds(const tfdtd_plane* d) {
    return 4 * sizeof(spec) * (size_t)d->frequencies * (size_t)(d->block > 1 ? d->block : 1);
}

template <typename real, typename spec>
int launch_plane(const tfdtd_plane* d, hipStream_t stream) {
    Plane<real, spec> a;
    int status = make_plane(d, &a);
    if (status) return status;
    const size_t lds = plane_lds<spec>(d);
    const long long blocks = (d->points * d->components + 255) / 256;
    hipLaunchKernelGGL((dft_plane_kernel<real, spec>), dim3((unsigned)blocks), dim3(256), lds, stream, a);
    return check_launch("dft_plane");
}

template <typename real, typename spec>
int launch_flush(const tfdtd_plane* d, hipStream_t stream) {
    Plane<real, spec> a;
    int status = make_plane(d, &a);
    if (status) return status;
    const long long blocks = (d->points * d->components + 255) / 256;
    hipLaunchKernelGGL((dft_flush_kernel<real, spec>), dim3((unsigned)blocks), dim3(256), plane_lds<spec>(d), stream, a);
    status = check_launch("dft_flush");
    if (status) return status;
    hipLaunchKernelGGL(set_flushed_kernel, dim3(1), dim3(1), 0, stream, a.counter, a.samples, const_cast<long long*>(a.flushed));
    return check_launch("dft_flush counter");
}

template <typename real, typename spec>
int make_plane(const tfdtd_plane* d, Plane<real, spec>* out) {
    Plane<real, spec>& a = *out;
    if (d->points < 1 || d->frequencies < 1 || d->components < 1 || d->components > 6 || d->electric < 0 ||
        d->electric > d->components || d->corners < 1 || d->corners > 8 || d->time_downsample < 1)
        return fail(TFDTD_EINVAL, "plane: bad sizes");
    if (!d->E || !d->H || !d->window_e || !d->window_h || !d->omega || !d->half || !d->value || !d->counter)
        return fail(TFDTD_EINVAL, "plane: null pointer");
    if (d->block < 0) return fail(TFDTD_EINVAL, "plane: negative block");
    if (plane_lds<spec>(d) > 64 * 1024)
        return fail(TFDTD_EINVAL, "plane: %d frequencies x block %d exceed one launch", d->frequencies, d->block);
    if (d->block > 1 && (!d->stage || !d->flushed)) return fail(TFDTD_EINVAL, "plane: a block needs stage and flushed buffers");
    std::memset(&a, 0, sizeof a);
    a.points = d->points; a.nf = d->frequencies; a.nc = d->components; a.ne = d->electric;
    a.corners = d->corners; a.downsample = d->time_downsample;
    a.E = static_cast<const real*>(d->E); a.H = static_cast<const real*>(d->H);
    for (int c = 0; c < d->components; ++c) {
        if (!d->indices[c] || !d->weights[c]) return fail(TFDTD_EINVAL, "plane: component %d map missing", c);
        a.indices[c] = reinterpret_cast<const long long*>(d->indices[c]);
        a.weights[c] = static_cast<const real*>(d->weights[c]);
    }
    a.window_e = static_cast<const real*>(d->window_e); a.window_h = static_cast<const real*>(d->window_h);
    a.omega = d->omega; a.half = static_cast<const spec*>(d->half); a.value = static_cast<spec*>(d->value);
    a.counter = reinterpret_cast<const long long*>(d->counter);
    if (d->samples < 0) return fail(TFDTD_EINVAL, "plane: negative window length");
    a.samples = d->samples;
    a.block = d->block > 1 ? d->block : 1;
    a.stage = static_cast<real*>(d->stage);
    a.flushed = reinterpret_cast<const long long*>(d->flushed);
    if ((d->points * d->components + 255) / 256 > 0x7fffffffLL) return fail(TFDTD_EINVAL, "plane: too many samples");
    return TFDTD_OK;
}

// The source boxes of a half step alone: field[cell][comp] += values[counter] in source order, over
// the union box of the sources (the step's own update is not applied).
template <typename real>
__global__ __launch_bounds__(256) void inject_kernel(const HalfStep<real> a) {
    const int bx = a.box_hi[0] - a.box_lo[0], by = a.box_hi[1] - a.box_lo[1], bz = a.box_hi[2] - a.box_lo[2];
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (long long)bx * by * bz) return;
    const int z = a.box_lo[2] + (int)(j % bz), y = a.box_lo[1] + (int)((j / bz) % by), x = a.box_lo[0] + (int)(j / ((long long)by * bz));
    const long long n = *global(a.counter);
    if (n >= a.source_samples) return;
    const long long i = ((long long)x * a.ny + y) * a.nz + z;
    real out[3] = {a.dst[3 * i], a.dst[3 * i + 1], a.dst[3 * i + 2]};
    bool hit = false;
    for (int k = 0; k < a.nsrc; ++k) {
        const Source<real>& box = a.srcs[k];
        if (x >= box.lo[0] && x < box.hi[0] && y >= box.lo[1] && y < box.hi[1] && z >= box.lo[2] && z < box.hi[2]) {
            const real v = global(box.values)[n];
            out[box.comp] = out[box.comp] + v;
            hit = true;
        }
    }
    if (hit) {
        a.dst[3 * i] = out[0];
        a.dst[3 * i + 1] = out[1];
        a.dst[3 * i + 2] = out[2];
    }
}

template <typename real>
int launch_inject(const tfdtd_half_step* d, hipStream_t stream) {
    HalfStep<real> a;
    int status = make_half_step(d, &a);
    if (status) return status;
    if (a.nsrc == 0) return TFDTD_OK;
    const long long count = (long long)(a.box_hi[0] - a.box_lo[0]) * (a.box_hi[1] - a.box_lo[1]) * (a.box_hi[2] - a.box_lo[2]);
    hipLaunchKernelGGL(inject_kernel<real>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, a);
    return check_launch("inject");
}

int half_step(const tfdtd_half_step* d, hipStream_t stream) {
    if (!d) return fail(TFDTD_EINVAL, "null half-step descriptor");
    if (d->family != TFDTD_FAMILY_E && d->family != TFDTD_FAMILY_H) return fail(TFDTD_EINVAL, "family must be E or H");
    if (d->dtype == TFDTD_F32) return launch_half_step<float>(d, stream);
    if (d->dtype == TFDTD_F64) return launch_half_step<double>(d, stream);
    return fail(TFDTD_EINVAL, "dtype must be float32 or float64");
}

int planes(const tfdtd_plane* p, int count, hipStream_t stream) {
    if (count < 0 || (count > 0 && !p)) return fail(TFDTD_EINVAL, "bad plane table");
    for (int k = 0; k < count; ++k) {
        int status;
        const tfdtd_plane* d = p + k;
        if (d->dtype == TFDTD_F32 && d->spectral == 0) status = launch_plane<float, float>(d, stream);
        else if (d->dtype == TFDTD_F32 && d->spectral == 1) status = launch_plane<float, double>(d, stream);
        else if (d->dtype == TFDTD_F64 && d->spectral == 1) status = launch_plane<double, double>(d, stream);
        else status = fail(TFDTD_EINVAL, "plane %d: unsupported dtype/spectral pair", k);
        if (status) return status;
    }
    return TFDTD_OK;
}

int observe(const tfdtd_observe* d, hipStream_t stream) {
    if (!d) return fail(TFDTD_EINVAL, "null observe descriptor");
    if (d->dtype == TFDTD_F32) return launch_observe<float>(d, stream);
    if (d->dtype == TFDTD_F64) return launch_observe<double>(d, stream);
    return fail(TFDTD_EINVAL, "dtype must be float32 or float64");
}

template <typename real>
int pack_half_steps(const tfdtd_half_step* steps, int count, void* host_table, size_t bytes, int32_t* prefix) {
    if (count < 1 || !steps || !prefix) return fail(TFDTD_EINVAL, "batch: at least one descriptor and a prefix array");
    if (!host_table || bytes < sizeof(HalfStep<real>) * (size_t)count) return fail(TFDTD_EINVAL, "batch: table buffer too small");
    HalfStep<real>* table = static_cast<HalfStep<real>*>(host_table);
    long long total = 0;
    for (int k = 0; k < count; ++k) {
        if (steps[k].family != steps[0].family || steps[k].dtype != steps[0].dtype)
            return fail(TFDTD_EINVAL, "batch: descriptor %d differs in family or dtype", k);
        const int status = make_half_step(steps + k, table + k);
        if (status) return status;
        prefix[k] = (int32_t)total;
        total += ((long long)steps[k].nx * steps[k].ny * steps[k].nz + 256 * kBatchSpan - 1) / (256 * kBatchSpan);
        if (total > 0x7fffffffLL) return fail(TFDTD_EINVAL, "batch: too many cells for one launch");
    }
    prefix[count] = (int32_t)total;
    return TFDTD_OK;
}

template <typename real, typename spec>
int pack_planes(const tfdtd_plane* planes, int count, void* host_table, size_t bytes, int32_t* max_blocks, int32_t* max_phases) {
    if (count < 1 || !planes || !max_blocks || !max_phases) return fail(TFDTD_EINVAL, "batch planes: bad arguments");
    if (!host_table || bytes < sizeof(Plane<real, spec>) * (size_t)count) return fail(TFDTD_EINVAL, "batch planes: table buffer too small");
    Plane<real, spec>* table = static_cast<Plane<real, spec>*>(host_table);
    *max_blocks = 0;
    *max_phases = 0;
    for (int k = 0; k < count; ++k) {
        const int status = make_plane(planes + k, table + k);
        if (status) return status;
        *max_blocks = std::max<int32_t>(*max_blocks, (int32_t)((planes[k].points * planes[k].components + 255) / 256));
        *max_phases = std::max(*max_phases, planes[k].frequencies * std::max(planes[k].block, 1));
    }
    return TFDTD_OK;
}

}  // namespace

extern "C" {

size_t tfdtd_batch_table_bytes(int32_t count, int32_t dtype) {
    if (count < 0) return 0;
    return (size_t)count * (dtype == TFDTD_F64 ? sizeof(HalfStep<double>) : sizeof(HalfStep<float>));
}

int tfdtd_batch_pack(const tfdtd_half_step* steps, int32_t count, void* host_table, size_t bytes, int32_t* block_prefix) {
    if (count < 1 || !steps) return fail(TFDTD_EINVAL, "batch: at least one descriptor");
    if (steps[0].dtype == TFDTD_F32) return pack_half_steps<float>(steps, count, host_table, bytes, block_prefix);
    if (steps[0].dtype == TFDTD_F64) return pack_half_steps<double>(steps, count, host_table, bytes, block_prefix);
    return fail(TFDTD_EINVAL, "batch: dtype must be float32 or float64");
}

int tfdtd_batch_half_step(const void* device_table, const int32_t* device_prefix, int32_t count, int32_t total_blocks,
                          int32_t dtype, int32_t family, void* stream) {
    if (!device_table || !device_prefix || count < 1 || total_blocks < 0) return fail(TFDTD_EINVAL, "batch: bad launch arguments");
    if (total_blocks == 0) return TFDTD_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int* prefix = reinterpret_cast<const int*>(device_prefix);
    if (dtype == TFDTD_F32) {
        const HalfStep<float>* t = static_cast<const HalfStep<float>*>(device_table);
        if (family == TFDTD_FAMILY_E) hipLaunchKernelGGL((batch_half_step_kernel<float, false>), dim3(total_blocks), dim3(256), 0, s, t, prefix, count);
        else hipLaunchKernelGGL((batch_half_step_kernel<float, true>), dim3(total_blocks), dim3(256), 0, s, t, prefix, count);
    } else if (dtype == TFDTD_F64) {
        const HalfStep<double>* t = static_cast<const HalfStep<double>*>(device_table);
        if (family == TFDTD_FAMILY_E) hipLaunchKernelGGL((batch_half_step_kernel<double, false>), dim3(total_blocks), dim3(256), 0, s, t, prefix, count);
        else hipLaunchKernelGGL((batch_half_step_kernel<double, true>), dim3(total_blocks), dim3(256), 0, s, t, prefix, count);
    } else {
        return fail(TFDTD_EINVAL, "batch: dtype must be float32 or float64");
    }
    return check_launch("batch_half_step");
}

size_t tfdtd_batch_planes_bytes(int32_t count, int32_t dtype, int32_t spectral) {
    if (count < 0) return 0;
    if (dtype == TFDTD_F32 && spectral == 0) return (size_t)count * sizeof(Plane<float, float>);
    if (dtype == TFDTD_F32 && spectral == 1) return (size_t)count * sizeof(Plane<float, double>);
    return (size_t)count * sizeof(Plane<double, double>);
}

int tfdtd_batch_planes_pack(const tfdtd_plane* p, int32_t count, void* host_table, size_t bytes, int32_t* max_blocks,
                            int32_t* max_phases) {
    if (count < 1 || !p) return fail(TFDTD_EINVAL, "batch planes: at least one plane");
    for (int k = 1; k < count; ++k)
        if (p[k].dtype != p[0].dtype || p[k].spectral != p[0].spectral) return fail(TFDTD_EINVAL, "batch planes: mixed precision");
    if (p[0].dtype == TFDTD_F32 && p[0].spectral == 0) return pack_planes<float, float>(p, count, host_table, bytes, max_blocks, max_phases);
    if (p[0].dtype == TFDTD_F32 && p[0].spectral == 1) return pack_planes<float, double>(p, count, host_table, bytes, max_blocks, max_phases);
    if (p[0].dtype == TFDTD_F64 && p[0].spectral == 1) return pack_planes<double, double>(p, count, host_table, bytes, max_blocks, max_phases);
    return fail(TFDTD_EINVAL, "batch planes: unsupported dtype/spectral pair");
}

int tfdtd_batch_dft_update(const void* device_table, int32_t count, int32_t max_blocks, int32_t max_phases,
                           int32_t dtype, int32_t spectral, void* stream) {
    if (!device_table || count < 1 || max_blocks < 1 || max_phases < 1) return fail(TFDTD_EINVAL, "batch planes: bad launch arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid((unsigned)max_blocks, (unsigned)count);
    if (dtype == TFDTD_F32 && spectral == 0) {
        const size_t lds = 4 * sizeof(float) * (size_t)max_phases;
        hipLaunchKernelGGL((batch_dft_kernel<float, float>), grid, dim3(256), lds, s, static_cast<const Plane<float, float>*>(device_table));
    } else if (dtype == TFDTD_F32 && spectral == 1) {
        const size_t lds = 4 * sizeof(double) * (size_t)max_phases;
        hipLaunchKernelGGL((batch_dft_kernel<float, double>), grid, dim3(256), lds, s, static_cast<const Plane<float, double>*>(device_table));
    } else if (dtype == TFDTD_F64 && spectral == 1) {
        const size_t lds = 4 * sizeof(double) * (size_t)max_phases;
        hipLaunchKernelGGL((batch_dft_kernel<double, double>), grid, dim3(256), lds, s, static_cast<const Plane<double, double>*>(device_table));
    } else {
        return fail(TFDTD_EINVAL, "batch planes: unsupported dtype/spectral pair");
    }
    return check_launch("batch_dft");
}


int tfdtd_abi_version(void) { return TFDTD_ABI_VERSION; }

#ifndef TFDTD_SOURCE_DIGEST
#define TFDTD_SOURCE_DIGEST "unknown"
#endif
const char* tfdtd_source_digest(void) { return TFDTD_SOURCE_DIGEST; }

size_t tfdtd_last_error(char* buffer, size_t size) {
    const std::string& message = tfdtd_detail::last_error;
    if (buffer && size) {
        std::strncpy(buffer, message.c_str(), size - 1);
        buffer[size - 1] = 0;
    }
    return message.size();
}

int tfdtd_yee_half_step(const tfdtd_half_step* step, void* stream) {
    return half_step(step, static_cast<hipStream_t>(stream));
}

int tfdtd_inject(const tfdtd_half_step* step, void* stream) {
    if (!step) return fail(TFDTD_EINVAL, "null half-step descriptor");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (step->dtype == TFDTD_F32) return launch_inject<float>(step, s);
    if (step->dtype == TFDTD_F64) return launch_inject<double>(step, s);
    return fail(TFDTD_EINVAL, "dtype must be float32 or float64");
}

int tfdtd_dft_update(const tfdtd_plane* p, int32_t count, void* stream) {
    return planes(p, count, static_cast<hipStream_t>(stream));
}

int tfdtd_observe_step(const tfdtd_observe* d, void* stream) { return observe(d, static_cast<hipStream_t>(stream)); }

int tfdtd_dft_flush(const tfdtd_plane* p, int32_t count, void* stream) {
    if (count < 0 || (count > 0 && !p)) return fail(TFDTD_EINVAL, "bad plane table");
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (int k = 0; k < count; ++k) {
        const tfdtd_plane* d = p + k;
        if (d->block <= 1) continue;
        int status;
        if (d->dtype == TFDTD_F32 && d->spectral == 0) status = launch_flush<float, float>(d, s);
        else if (d->dtype == TFDTD_F32 && d->spectral == 1) status = launch_flush<float, double>(d, s);
        else if (d->dtype == TFDTD_F64 && d->spectral == 1) status = launch_flush<double, double>(d, s);
        else status = fail(TFDTD_EINVAL, "plane %d: unsupported dtype/spectral pair", k);
        if (status) return status;
    }
    return TFDTD_OK;
}

int tfdtd_step_fused(const tfdtd_half_step* e_step, const tfdtd_half_step* h_step, const tfdtd_plane* p,
                     int32_t plane_count, const tfdtd_observe* obs, int32_t n_steps, void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n_steps < 0) return fail(TFDTD_EINVAL, "n_steps must be non-negative");
    if (!e_step || !h_step || e_step->family != TFDTD_FAMILY_E || h_step->family != TFDTD_FAMILY_H)
        return fail(TFDTD_EINVAL, "step_fused needs an E and an H descriptor");
    if (!obs) return fail(TFDTD_EINVAL, "step_fused needs the observe descriptor (it advances the counter)");
    for (int k = 0; k < n_steps; ++k) {
        int status = half_step(e_step, s);
        if (!status) status = half_step(h_step, s);
        if (!status) status = planes(p, plane_count, s);
        if (!status) status = observe(obs, s);
        if (status) return status;
    }
    return TFDTD_OK;
}

}  // extern "C"
