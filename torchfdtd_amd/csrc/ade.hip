// Passive Drude/Lorentz materials: the coupled trapezoidal auxiliary-differential-equation update
// of torchfdtd/materials.py:29-87 (MaterialADE.prepare / correct) around the E half step, for the
// samples one dispersive material owns.
//
// tfdtd_ade_save (before the E update) keeps E at the owned samples; tfdtd_ade_correct (after it,
// before the E sources) solves the field and the oscillators together, per sample and pole j:
//   response_j = (Q_j - a_j P_j) / d_j
//   new        = (eps_inf free - k_sum old - sum_j response_j) / (eps_inf + k_sum)
//   delta_j    = response_j + k_j (new + old);  P_j += delta_j;  Q_j = -Q_j + 2 delta_j;  E = new
// with every product, quotient and sum rounded in the field precision in this order.  One pole
// keeps the reference's scalar coefficients: eps_inf + k is formed in double and rounded once;
// several poles hold them in the field precision (k_sum summed pole by pole, eps_inf + k_sum
// rounded in the field precision), as the reference's array coefficients do.
#include "common.h"

namespace {

template <typename real>
struct Ade {
    long long samples;        // owned samples (field elements)
    int per_cell;             // 1: index[i / 3] is a cell and the sample is its component i % 3
    int poles;
    const long long* index;
    real* field;
    real* old;
    real* P;
    real* Q;
    real a[TFDTD_ADE_MAX_POLES], d[TFDTD_ADE_MAX_POLES], k[TFDTD_ADE_MAX_POLES];
    real eps, k_sum, denominator;
};

template <typename real>
__device__ __forceinline__ long long element(const Ade<real>& a, long long i) {
    return a.per_cell ? 3 * a.index[i / 3] + i % 3 : a.index[i];
}

template <typename real>
__global__ __launch_bounds__(256) void ade_save_kernel(const Ade<real> a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.samples) return;
    a.old[i] = a.field[element(a, i)];
}

template <typename real>
__global__ __launch_bounds__(256) void ade_correct_kernel(const Ade<real> a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.samples) return;
    const long long e = element(a, i);
    const real old = a.old[i], free = a.field[e];
    real response[TFDTD_ADE_MAX_POLES];
    real total = real(0);
    for (int j = 0; j < a.poles; ++j) {
        const long long s = (long long)j * a.samples + i;
        response[j] = (a.Q[s] - a.a[j] * a.P[s]) / a.d[j];
        total = j == 0 ? response[0] : total + response[j];
    }
    const real fresh = (a.eps * free - a.k_sum * old - total) / a.denominator;
    for (int j = 0; j < a.poles; ++j) {
        const long long s = (long long)j * a.samples + i;
        const real delta = response[j] + a.k[j] * (fresh + old);
        a.P[s] = a.P[s] + delta;
        a.Q[s] = -a.Q[s] + real(2) * delta;
    }
    a.field[e] = fresh;
}

template <typename real>
int make_ade(const tfdtd_ade* d, Ade<real>* a) {
    if (d->poles < 1 || d->poles > TFDTD_ADE_MAX_POLES) return tfdtd_fail(TFDTD_EINVAL, "ade: 1 to %d poles", TFDTD_ADE_MAX_POLES);
    if (d->samples < 0 || (d->samples > 0 && (!d->index || !d->field || !d->old || !d->P || !d->Q)))
        return tfdtd_fail(TFDTD_EINVAL, "ade: null index, field, old, P or Q");
    a->samples = d->samples;
    a->per_cell = d->per_cell != 0;
    a->poles = d->poles;
    a->index = reinterpret_cast<const long long*>(d->index);
    a->field = static_cast<real*>(d->field);
    a->old = static_cast<real*>(d->old);
    a->P = static_cast<real*>(d->P);
    a->Q = static_cast<real*>(d->Q);
    real k_sum = real(0);
    for (int j = 0; j < d->poles; ++j) {
        a->a[j] = (real)d->a[j];
        a->d[j] = (real)d->d[j];
        a->k[j] = (real)d->k[j];
        k_sum = j == 0 ? a->k[0] : k_sum + a->k[j];
    }
    a->eps = (real)d->eps_inf;
    a->k_sum = k_sum;
    a->denominator = d->poles == 1 ? (real)(d->eps_inf + d->k[0]) : a->eps + k_sum;
    return TFDTD_OK;
}

template <typename real>
int launch_ade(const tfdtd_ade* d, bool correct, hipStream_t stream) {
    Ade<real> a;
    const int status = make_ade(d, &a);
    if (status || a.samples == 0) return status;
    const unsigned blocks = (unsigned)((a.samples + 255) / 256);
    if (correct) hipLaunchKernelGGL(ade_correct_kernel<real>, dim3(blocks), dim3(256), 0, stream, a);
    else hipLaunchKernelGGL(ade_save_kernel<real>, dim3(blocks), dim3(256), 0, stream, a);
    return tfdtd_check_launch(correct ? "ade_correct" : "ade_save");
}

int ade(const tfdtd_ade* d, bool correct, void* stream) {
    if (!d) return tfdtd_fail(TFDTD_EINVAL, "ade: null descriptor");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (d->dtype == TFDTD_F32) return launch_ade<float>(d, correct, s);
    if (d->dtype == TFDTD_F64) return launch_ade<double>(d, correct, s);
    return tfdtd_fail(TFDTD_EINVAL, "ade: dtype must be float32 or float64");
}

}  // namespace

extern "C" {

int tfdtd_ade_save(const tfdtd_ade* d, void* stream) { return ade(d, false, stream); }

int tfdtd_ade_correct(const tfdtd_ade* d, void* stream) { return ade(d, true, stream); }

}  // extern "C"
