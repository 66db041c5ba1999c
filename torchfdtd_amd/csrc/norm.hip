// State norms for the run-control diagnostics (torchfdtd/run_control.py:127-159,
// torchfdtd/cuda_diagnostics.py:63-135): per grid, in FP64,
//   energy = sum_c vol E_c^2 / inv_eps_c + vol H_c^2,   peak = max |E|, |H|,
// and a flag for any non-finite value of E, H or the CPML memory.
//
// Two launches for any number of grids: norm_partial reduces one 4096-element task of one
// array per workgroup; norm_final reduces each grid's task range with one workgroup per
// grid.  Every grid's result lands in out[grid][3] (energy, peak, non-finite flag), so a
// whole batch is measured with one pair of launches and one 24*grids byte copy.

#include "common.h"

#include <mutex>

namespace {

constexpr int kTask = 4096;
constexpr int kThreads = 256;

struct NormGrid {
    int nx, ny, nz, eps_stride;
    const void* E;
    const void* H;
    const void* inverse;
    const double* widths[3];
    const void* psi[2];
    long long psi_count[2];
    int task_lo, task_hi;
};

struct NormTask {
    int grid, part;        // part 0 E, 1 H, 2 and 3 the CPML memory buffers
    long long start;
};

__device__ __forceinline__ void block_reduce(double sum, double peak, double bad, double* out) {
    __shared__ double s[3][kThreads];
    const int t = threadIdx.x;
    s[0][t] = sum;
    s[1][t] = peak;
    s[2][t] = bad;
    __syncthreads();
    for (int k = kThreads / 2; k > 0; k >>= 1) {
        if (t < k) {
            s[0][t] += s[0][t + k];
            s[1][t] = fmax(s[1][t], s[1][t + k]);
            s[2][t] = fmax(s[2][t], s[2][t + k]);
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = s[0][0];
        out[1] = s[1][0];
        out[2] = s[2][0];
    }
}

template <typename real>
__global__ __launch_bounds__(kThreads) void norm_partial(const NormGrid* __restrict__ grids, const NormTask* __restrict__ tasks,
                                                        double* __restrict__ partial) {
    const NormTask task = tasks[blockIdx.x];
    const NormGrid& g = grids[task.grid];
    double sum = 0., peak = 0., bad = 0.;
    if (task.part < 2) {
        const auto data = global(static_cast<const real*>(task.part == 0 ? g.E : g.H));
        const auto inverse = global(static_cast<const real*>(g.inverse));
        const long long n = 3LL * g.nx * g.ny * g.nz;
        const long long end = min(n, task.start + kTask);
        for (long long i = task.start + threadIdx.x; i < end; i += kThreads) {
            const double v = (double)data[i];
            if (!isfinite(v)) bad = 1.;
            const long long cell = i / 3;
            double vol = 1.;
            if (g.widths[0]) {
                const int z = (int)(cell % g.nz);
                const int y = (int)((cell / g.nz) % g.ny);
                const int x = (int)(cell / ((long long)g.nz * g.ny));
                vol = global(g.widths[0])[x] * global(g.widths[1])[y] * global(g.widths[2])[z];
            }
            double w = vol;
            if (task.part == 0) w = vol / (double)inverse[g.eps_stride == 3 ? i : cell];
            sum += v * v * w;
            peak = fmax(peak, fabs(v));
        }
    } else {
        const int k = task.part - 2;
        const auto data = global(static_cast<const real*>(g.psi[k]));
        const long long end = min(g.psi_count[k], task.start + kTask);
        for (long long i = task.start + threadIdx.x; i < end; i += kThreads)
            if (!isfinite((double)data[i])) bad = 1.;
    }
    block_reduce(sum, peak, bad, partial + 3LL * blockIdx.x);
}

__global__ __launch_bounds__(kThreads) void norm_final(const NormGrid* __restrict__ grids, const double* __restrict__ partial,
                                                      double* __restrict__ out) {
    const NormGrid& g = grids[blockIdx.x];
    double sum = 0., peak = 0., bad = 0.;
    for (int t = g.task_lo + threadIdx.x; t < g.task_hi; t += kThreads) {
        sum += partial[3LL * t];
        peak = fmax(peak, partial[3LL * t + 1]);
        bad = fmax(bad, partial[3LL * t + 2]);
    }
    block_reduce(sum, peak, bad, out + 3LL * blockIdx.x);
}

// The run-control rule of one grid at one barrier (one lane per grid).  Every expression is
// the FP64 operation a host evaluation makes, so the verdicts and history equal a host rule's.
__global__ __launch_bounds__(64) void decay_update(const tfdtd_decay_rule* __restrict__ rules,
                                                   tfdtd_decay_state* __restrict__ states,
                                                   const double* __restrict__ norms, int count, long long step,
                                                   double* __restrict__ history, int rows) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= count) return;
    const tfdtd_decay_rule rule = rules[g];
    tfdtd_decay_state s = states[g];
    if (s.verdict != TFDTD_RUN_GOING) return;
    const double energy = norms[3LL * g], field_peak = norms[3LL * g + 1], flagged = norms[3LL * g + 2];
    const double when = (double)step * rule.time_step;
    if (flagged != 0. || !isfinite(energy) || !isfinite(field_peak)) {
        s.verdict = TFDTD_RUN_NONFINITE;
        s.verdict_step = step;
        s.verdict_value = isfinite(energy) && isfinite(field_peak) ? 1. : 0.;   // 1: only the CPML memory
    } else if (rule.divergence_check && rule.field_limit > 0. && field_peak > rule.field_limit) {
        s.verdict = TFDTD_RUN_FIELD_LIMIT;
        s.verdict_step = step;
        s.verdict_value = field_peak;
    } else {
        const bool eligible = step >= rule.min_steps && when >= rule.eligible_time;
        s.peak_energy = fmax(s.peak_energy, energy);
        if (eligible && !s.frozen) {
            s.frozen = 1;
            s.baseline = s.peak_energy;
        }
        const double base = s.frozen ? s.baseline : s.peak_energy;
        const double ratio = base != 0. ? energy / base : 0.;
        if (rule.divergence_check && eligible && ratio > rule.growth_limit) {
            s.verdict = TFDTD_RUN_GROWTH;
            s.verdict_step = step;
            s.verdict_value = ratio;
        } else {
            s.below = eligible && ratio <= rule.decay_threshold ? s.below + 1 : 0;
            if (s.checks < rows) {
                double* row = history + 6LL * ((long long)g * rows + s.checks);
                row[0] = (double)step;
                row[1] = when;
                row[2] = energy;
                row[3] = ratio;
                row[4] = field_peak;
                row[5] = eligible ? 1. : 0.;
            }
            s.checks += 1;
            if (rule.auto_shutoff && s.below >= rule.consecutive_checks) {
                s.verdict = TFDTD_RUN_DECAYED;
                s.verdict_step = step;
            }
        }
    }
    states[g] = s;
}

// Partial sums of tfdtd_state_norm_noalloc: one copy per device, held until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host it names a stand-in, not
// device memory); it uses the per-device address from hipGetSymbolAddress below.
static __device__ __align__(256) double scratch[3 * TFDTD_NORM_SCRATCH_TASKS];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double* scratch_address[kMaxDevices] = {};

// The address of the current device's copy of `scratch`, looked up once per device.
int scratch_for_current_device(double** out) {
    int device = -1;
    hipError_t status = hipGetDevice(&device);
    if (status != hipSuccess) return tfdtd_fail(TFDTD_EHIP, "norm: hipGetDevice: %s", hipGetErrorString(status));
    if (device < 0 || device >= kMaxDevices) return tfdtd_fail(TFDTD_EINVAL, "norm: device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_address[device]) {
        void* address = nullptr;
        status = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch));
        if (status != hipSuccess || !address)
            return tfdtd_fail(TFDTD_EHIP, "norm: hipGetSymbolAddress(scratch): %s", hipGetErrorString(status));
        scratch_address[device] = static_cast<double*>(address);
    }
    *out = scratch_address[device];
    return TFDTD_OK;
}

long long tasks_of(long long n) { return (n + kTask - 1) / kTask; }

long long grid_tasks(const tfdtd_norm_grid& d) {
    const long long n = 3LL * d.nx * d.ny * d.nz;
    return 2 * tasks_of(n) + tasks_of(d.psi[0] ? d.psi_count[0] : 0) + tasks_of(d.psi[1] ? d.psi_count[1] : 0);
}

size_t table_bytes(int count, long long tasks) {
    return sizeof(NormGrid) * (size_t)count + sizeof(NormTask) * (size_t)tasks;
}

}  // namespace

extern "C" {

int64_t tfdtd_norm_tasks(const tfdtd_norm_grid* grids, int32_t count) {
    if (!grids || count < 1) return -1;
    long long tasks = 0;
    for (int k = 0; k < count; ++k) tasks += grid_tasks(grids[k]);
    return tasks;
}

size_t tfdtd_norm_table_bytes(const tfdtd_norm_grid* grids, int32_t count) {
    const int64_t tasks = tfdtd_norm_tasks(grids, count);
    return tasks < 0 ? 0 : table_bytes(count, tasks);
}

int tfdtd_norm_pack(const tfdtd_norm_grid* grids, int32_t count, void* host_table, size_t bytes) {
    const int64_t total = tfdtd_norm_tasks(grids, count);
    if (total < 0) return tfdtd_fail(TFDTD_EINVAL, "norm: at least one grid");
    if (total > 0x7fffffffLL) return tfdtd_fail(TFDTD_EINVAL, "norm: too many tasks for one launch");
    if (!host_table || bytes < table_bytes(count, total)) return tfdtd_fail(TFDTD_EINVAL, "norm: table buffer too small");
    NormGrid* table = static_cast<NormGrid*>(host_table);
    NormTask* tasks = reinterpret_cast<NormTask*>(table + count);
    long long t = 0;
    for (int k = 0; k < count; ++k) {
        const tfdtd_norm_grid& d = grids[k];
        if (d.dtype != grids[0].dtype) return tfdtd_fail(TFDTD_EINVAL, "norm: grid %d differs in dtype", k);
        if (d.nx < 1 || d.ny < 1 || d.nz < 1 || !d.E || !d.H || !d.inverse_permittivity)
            return tfdtd_fail(TFDTD_EINVAL, "norm: grid %d lacks a shape, a field or the inverse permittivity", k);
        if (d.eps_stride != 1 && d.eps_stride != 3) return tfdtd_fail(TFDTD_EINVAL, "norm: eps_stride must be 1 or 3");
        const bool weighted = d.widths[0] || d.widths[1] || d.widths[2];
        if (weighted && !(d.widths[0] && d.widths[1] && d.widths[2]))
            return tfdtd_fail(TFDTD_EINVAL, "norm: grid %d: give all three width arrays or none", k);
        NormGrid& g = table[k];
        g.nx = d.nx; g.ny = d.ny; g.nz = d.nz; g.eps_stride = d.eps_stride;
        g.E = d.E; g.H = d.H; g.inverse = d.inverse_permittivity;
        for (int a = 0; a < 3; ++a) g.widths[a] = weighted ? d.widths[a] : nullptr;
        for (int j = 0; j < 2; ++j) {
            g.psi[j] = d.psi[j];
            g.psi_count[j] = d.psi[j] ? d.psi_count[j] : 0;
            if (g.psi_count[j] < 0) return tfdtd_fail(TFDTD_EINVAL, "norm: negative CPML memory size");
        }
        g.task_lo = (int)t;
        const long long n = 3LL * d.nx * d.ny * d.nz;
        for (int part = 0; part < 4; ++part) {
            const long long size = part < 2 ? n : g.psi_count[part - 2];
            for (long long s = 0; s < size; s += kTask) tasks[t++] = NormTask{k, part, s};
        }
        g.task_hi = (int)t;
    }
    return TFDTD_OK;
}

int tfdtd_state_norm(const void* device_table, int32_t count, int64_t tasks, int32_t dtype, double* partial, double* out,
                     void* stream) {
    if (!device_table || count < 1 || tasks < 1 || tasks > 0x7fffffffLL || !partial || !out)
        return tfdtd_fail(TFDTD_EINVAL, "norm: bad launch arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const NormGrid* grids = static_cast<const NormGrid*>(device_table);
    const NormTask* list = reinterpret_cast<const NormTask*>(grids + count);
    if (dtype == TFDTD_F32) hipLaunchKernelGGL(norm_partial<float>, dim3((unsigned)tasks), dim3(kThreads), 0, s, grids, list, partial);
    else if (dtype == TFDTD_F64) hipLaunchKernelGGL(norm_partial<double>, dim3((unsigned)tasks), dim3(kThreads), 0, s, grids, list, partial);
    else return tfdtd_fail(TFDTD_EINVAL, "norm: dtype must be float32 or float64");
    int status = tfdtd_check_launch("norm_partial");
    if (status) return status;
    hipLaunchKernelGGL(norm_final, dim3((unsigned)count), dim3(kThreads), 0, s, grids, partial, out);
    return tfdtd_check_launch("norm_final");
}

int tfdtd_decay_update(const tfdtd_decay_rule* rules, tfdtd_decay_state* states, const double* norms, int32_t count,
                       int64_t step, double* history, int32_t rows_per_grid, void* stream) {
    if (!rules || !states || !norms || count < 1 || step < 0 || rows_per_grid < 0 || (rows_per_grid > 0 && !history))
        return tfdtd_fail(TFDTD_EINVAL, "decay: bad arguments");
    hipLaunchKernelGGL(decay_update, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, static_cast<hipStream_t>(stream),
                       rules, states, norms, count, (long long)step, history, rows_per_grid);
    return tfdtd_check_launch("decay_update");
}

int tfdtd_state_norm_noalloc(const void* device_table, int32_t count, int64_t tasks, int32_t dtype, double* out,
                             void* stream) {
    if (tasks > TFDTD_NORM_SCRATCH_TASKS)
        return tfdtd_fail(TFDTD_EINVAL, "norm: %lld tasks exceed the %d of the library's partial-sum buffer",
                          (long long)tasks, TFDTD_NORM_SCRATCH_TASKS);
    double* partial = nullptr;
    const int status = scratch_for_current_device(&partial);
    if (status) return status;
    // Every task writes its own row of the buffer before norm_final reads it: nothing to clear.
    return tfdtd_state_norm(device_table, count, tasks, dtype, partial, out, stream);
}

}  // extern "C"
