/*
 * tfdtd.h -- C ABI of the MI355X (gfx950) Yee/CPML time-stepping engine.
 *
 * Plain C: descriptors of raw device pointers and sizes, a hipStream_t passed as
 * void*, an int status (0 = success).  No allocation and no host synchronisation
 * happens inside any entry point, so every call may be recorded into a HIP graph.
 * The caller owns every buffer.
 *
 * Layout of field arrays: real[nx][ny][nz][3], z fastest, components interleaved
 * (the reference YeeGrid layout, torchfdtd/boundaries.py:238).  real is float
 * (dtype 0) or double (dtype 1).
 */
#ifndef TORCHFDTD_AMD_TFDTD_H_
#define TORCHFDTD_AMD_TFDTD_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TFDTD_ABI_VERSION 1
#define TFDTD_MAX_SEGMENTS 12   /* 6 curl terms x 2 PML faces, per family */
#define TFDTD_MAX_SOURCES 16    /* soft source boxes per family */

enum { TFDTD_OK = 0, TFDTD_EINVAL = 1, TFDTD_EHIP = 2 };
enum { TFDTD_F32 = 0, TFDTD_F64 = 1 };
enum { TFDTD_FAMILY_E = 0, TFDTD_FAMILY_H = 1 };

/* One CPML auxiliary slab of one curl term (torchfdtd/boundaries.py:307-355).
 * psi has the grid's shape with extent hi-lo along `axis` (x,y,z order, z fastest);
 * b, c and inv_k hold hi-lo coefficients of the target rows lo..hi-1. */
typedef struct tfdtd_cpml_segment {
    int32_t axis;        /* derivative axis 0..2 */
    int32_t comp;        /* input component differentiated */
    int32_t lo, hi;      /* target rows [lo, hi) along axis */
    void* psi;
    const void* b;
    const void* c;
    const void* inv_k;
} tfdtd_cpml_segment;

/* A soft additive source on a box of cells (torchfdtd/injection.py:117-141):
 * field[box, comp] += values[counter] after the family's update. */
typedef struct tfdtd_source {
    int32_t comp;
    int32_t lo[3], hi[3];
    const void* values;  /* real[steps]: waveform x polarisation weight */
} tfdtd_source;

/* One half step of one family: E (backward differences of H) or H (forward
 * differences of E), torchfdtd/boundaries.py:363-438. */
typedef struct tfdtd_half_step {
    int32_t dtype;
    int32_t family;                 /* TFDTD_FAMILY_E or TFDTD_FAMILY_H */
    int32_t nx, ny, nz;
    const void* src;                /* H for the E update, E for the H update */
    void* dst;                      /* E for the E update, H for the H update */
    const void* inverse;            /* inverse permittivity (E) or permeability (H) */
    int32_t inverse_stride;         /* 0 one value, 1 per cell, 3 per cell and component */
    double courant;                 /* S, rounded to real inside the kernel */
    int32_t periodic_mask;          /* bit a: axis a wraps (real fields) */
    int32_t pec_upper_mask;         /* bit a: PEC ghost node above axis a (H update) */
    double pec_upper_factor[3];
    int32_t nseg;
    tfdtd_cpml_segment seg[TFDTD_MAX_SEGMENTS];
    int32_t nsrc;
    tfdtd_source srcs[TFDTD_MAX_SOURCES];
    int64_t source_samples;         /* length of every values array; no drive at or beyond it */
    const int64_t* counter;         /* device step counter read for source samples */
} tfdtd_half_step;

/* Point traces and the step counter (torchfdtd/solver.py:791-797). */
typedef struct tfdtd_observe {
    int32_t dtype;
    int32_t count;                  /* number of traces */
    const int64_t* table;           /* device int64[count][2]: field base address, element index
                                       ((x*ny+y)*nz+z)*3+c; traces of many grids may share one table */
    void* traces;                   /* real[steps][count] */
    int64_t steps;                  /* rows of traces; records stop beyond */
    int64_t* counter;               /* incremented by one after recording */
} tfdtd_observe;

/* One frequency plane (torchfdtd/field_monitors.py:121-164, cuda_monitors.py:65-121). */
typedef struct tfdtd_plane {
    int32_t dtype;                  /* field real type */
    int32_t spectral;               /* accumulator: 0 complex64, 1 complex128 */
    int64_t points;                 /* n */
    int32_t frequencies;            /* nf */
    int32_t components;             /* nc <= 6: E components first, then H */
    int32_t electric;               /* ne */
    int32_t corners;                /* interpolation terms per point (1..8) */
    int32_t time_downsample;
    const void* E;
    const void* H;
    const int64_t* indices[6];      /* per component: int64[corners][n] flat field indices */
    const void* weights[6];         /* per component: real[corners][n] */
    const void* window_e;           /* real[steps]: apodisation x dt x downsample mask */
    const void* window_h;           /* same at the half-step times */
    const double* omega;            /* double[nf]: 2 pi f dt */
    const void* half;               /* spectral complex[nf]: exp(i pi f dt) */
    void* value;                    /* spectral complex[nf][n][nc], accumulated */
    int64_t samples;                /* length of the windows; no accumulation at or beyond it */
    const int64_t* counter;
    int32_t block;                  /* > 1: stage the samples of `block` steps and accumulate them together
                                       (FP64 sums of the block's products); 0 or 1: every step */
    void* stage;                    /* real[block][n][nc] */
    int64_t* flushed;               /* step up to which tfdtd_dft_flush accumulated (zero it with the state) */
} tfdtd_plane;

/* CPML slabs of one family for the one-pass step.  Every psi array of the family lies
 * in one buffer and every coefficient array in another; the two curl terms that
 * differentiate along one axis share their rows and coefficients on each face. */
typedef struct tfdtd_march_face {
    int32_t lo, hi;                 /* target rows [lo, hi) along the axis; lo == hi: no slab */
    int64_t psi[2];                 /* element offsets of the two terms' psi arrays, curl order */
    int64_t coef;                   /* element offset of b[hi-lo], then c[hi-lo], then inv_k[hi-lo] */
} tfdtd_march_face;

typedef struct tfdtd_march_family {
    const void* psi_in;             /* psi at time n */
    void* psi_out;                  /* psi at time n+1: a separate buffer for E, psi_in for H */
    const void* coef;
    tfdtd_march_face face[3][2];    /* [axis][low face, high face] */
} tfdtd_march_family;

/* One whole Yee step -- E, then H, of every cell -- in a single launch.  Time n is read
 * from the *_in buffers and time n+1 written to the *_out buffers, which may not alias
 * them; the caller alternates the two sets (and the E psi buffers) between steps.  The
 * H psi buffer updates in place.  Periodic axes are not supported here (use the half
 * steps), and 3 x cells must stay below 2^32. */
typedef struct tfdtd_march {
    int32_t dtype;
    int32_t nx, ny, nz;             /* every axis >= 2 */
    const void* e_in;
    const void* h_in;
    void* e_out;
    void* h_out;
    const void* inverse_permittivity;
    int32_t eps_stride;             /* 1 per cell, 3 per cell and component */
    const void* inverse_permeability; /* one value */
    double courant;
    int32_t pec_upper_mask;
    double pec_upper_factor[3];
    tfdtd_march_family e;
    tfdtd_march_family h;
    int32_t nsrc_e;                 /* at most 8 per family */
    tfdtd_source srcs_e[TFDTD_MAX_SOURCES];
    int32_t nsrc_h;
    tfdtd_source srcs_h[TFDTD_MAX_SOURCES];
    int64_t source_samples;
    const int64_t* counter;
    int32_t tile_y;                 /* rows per tile: 8 (4 wavefronts) or 16 (8 wavefronts) */
    int32_t chunk_x;                /* planes per workgroup; 0 = the whole x axis */
    /* Optional per-cell material index, eps_stride 1 only: when non-null, 1/eps of cell i is
     * material_table[material[i]] (`materials` entries, 1 to 256, of dtype) and the step does not
     * read inverse_permittivity.  The caller keeps index and table equal to the permittivity. */
    const uint8_t* material;
    const void* material_table;
    int32_t materials;
    /* Two-step pass only (the one-step kernel refuses them): E and H of the intermediate step n+1
     * in snap_count (at most 4) disjoint boxes of cells [snap_lo[k][a], snap_hi[k][a]) are also
     * written to snap_e and snap_h: box k's cells after box k-1's, each box x-major like the
     * fields, 3 values per cell, so that plane monitors and traces inside the boxes can sample step
     * n+1 through indices into these buffers.  snap_count 0 when unused. */
    void* snap_e;
    void* snap_h;
    int32_t snap_count;
    int32_t snap_lo[4][3], snap_hi[4][3];
} tfdtd_march;

/* torchfdtd/reversible_cuda_fused.py:259-325 (FusedEHMarching) generalised to the
 * CPML, PEC and source layouts of torchfdtd/cuda_kernels.py:150-259: E and H of one
 * step in one pass over HBM.
 *
 * The step's three tile launches (interior, z-face and boundary tiles: disjoint cells, time-n
 * inputs only) run concurrently: the library forks two side streams of its own from `stream`
 * with events and joins them back, so the work is ordered on `stream` as one launch would be
 * and can be captured into a graph.  The side streams (default priority, one pair per device)
 * are created by the first call on a device made outside stream capture; calls captured
 * before that run the three launches one after another on `stream`.  Eager calls on one
 * device from several host threads at once are safe: the library holds a per-device lock while
 * it enqueues the fork, the launches and the join.  While a thread captures such a call into a
 * graph, the side streams belong to that capture until it ends: other threads must not call
 * these entry points on that device meanwhile.  The same holds for tfdtd_batch_march and the
 * two-step entry points. */
int tfdtd_yee_step_march(const tfdtd_march* step, void* stream);

/* Two whole Yee steps, n -> n+2, in one pass over HBM (float32): the step closure of
 * torchfdtd/solver.py:785-797 twice without the observers, with the intermediate step kept
 * on chip (observers of step n+1 read the snapshot of tfdtd_march.snap_*, launched by the caller
 * after the pass, then those of step n+2).  The result -- E, H and both CPML memories -- equals two calls of
 * tfdtd_yee_step_march (*_in -> *_out -> *_in) bit for bit, but lands in the *_out set.
 * Time n is read from the *_in buffers and n+2 written to the *_out buffers; unlike the one-step
 * call, the H family's CPML memory is double-buffered too (h.psi_out must differ from h.psi_in
 * when the grid has H slabs), because a tile recomputes step n+1 on cells its neighbours own.
 * Sources are added at both steps (samples n and n+1 of the device counter, which the call reads
 * and does not advance).  Tiles are 9 rows x 61 columns; chunk_x is the planes per workgroup
 * (about 20 suits a 512^3 grid) and tile_y the wavefronts of an interior workgroup, 4, 6 or 12
 * (0 = 12) splitting the 12 widened rows, or 16 for tall interior tiles of 13 owned rows
 * (z-face and boundary tiles use 12).
 * Concurrency as tfdtd_yee_step_march. */
int tfdtd_yee_step_march2(const tfdtd_march* step, void* stream);

/* The two-step pass of one or many grids from a device table, as tfdtd_batch_march_pack /
 * tfdtd_batch_march for the one-step kernel: tfdtd_batch_march2_pack validates `count` float32
 * descriptors of one tile_y and eps_stride (the layout of tfdtd_yee_step_march2), writes the table
 * (tfdtd_batch_march_bytes(count, TFDTD_F32) bytes) and block_prefix[3][count+1]; the caller copies
 * both to device memory once and launches tfdtd_batch_march2 for every pass with totals[v] =
 * block_prefix[v][count].  The tile parameters are then read from device memory instead of being
 * held in kernel-argument registers.  `snapshot` is nonzero exactly when a packed descriptor has a
 * step n+1 snapshot (snap_count > 0; tile_y 12 or 16): the kernels that store it are compiled apart,
 * and the others write none. */
int tfdtd_batch_march2_pack(const tfdtd_march* steps, int32_t count, void* host_table, size_t bytes, int32_t* block_prefix);
int tfdtd_batch_march2(const void* device_table, const int32_t* device_prefix, int32_t count, int32_t tile_y,
                       int32_t eps_stride, int32_t snapshot, const int32_t* totals, void* stream);

/* torchfdtd/cuda_kernels.py:349-365 (FusedYeeCUDA.update_E / update_H), with the
 * family's sources of torchfdtd/solver.py:778-783 folded in when nsrc > 0. */
int tfdtd_yee_half_step(const tfdtd_half_step* step, void* stream);

/* torchfdtd/cuda_monitors.py:135-155 (FusedFrequencyPlanes.update): phases from the
 * absolute counter in FP64, eight-corner sampling and accumulation, one launch per plane. */
int tfdtd_dft_update(const tfdtd_plane* planes, int32_t count, void* stream);

/* Accumulate the staged samples of an unfinished block (blocked planes; nothing for block <= 1):
 * call before reading `value` mid-window or after a run that stopped early. */
int tfdtd_dft_flush(const tfdtd_plane* planes, int32_t count, void* stream);

/* torchfdtd/solver.py:791-797: trace record (traces.index_copy_) and counter.add_(1). */
int tfdtd_observe_step(const tfdtd_observe* observe, void* stream);

/* torchfdtd/solver.py:785-797 (the step closure) repeated n_steps times:
 * E update + E sources, H update + H sources, plane DFTs, traces, counter += 1. */
int tfdtd_step_fused(const tfdtd_half_step* e_step, const tfdtd_half_step* h_step,
                     const tfdtd_plane* planes, int32_t plane_count,
                     const tfdtd_observe* observe, int32_t n_steps, void* stream);

/* Uneven batches: many grids of any shape advance together, one launch per family.
 * torchfdtd/cuda_batch.py:19-71 (FusedBatchYeeCUDA, identical-topology cohorts) generalised to
 * grids of different shapes, CPML layouts and sources.  tfdtd_batch_pack validates `count`
 * half-step descriptors of one family and dtype, writes the engine's internal table into
 * host_table (tfdtd_batch_table_bytes bytes) and block_prefix[count+1]; the caller copies the table
 * and the prefix to device memory once, then launches tfdtd_batch_half_step every step with
 * total_blocks = block_prefix[count]. */
size_t tfdtd_batch_table_bytes(int32_t count, int32_t dtype);
int tfdtd_batch_pack(const tfdtd_half_step* steps, int32_t count, void* host_table, size_t bytes, int32_t* block_prefix);
int tfdtd_batch_half_step(const void* device_table, const int32_t* device_prefix, int32_t count, int32_t total_blocks,
                          int32_t dtype, int32_t family, void* stream);

/* The one-pass step of many grids (tfdtd_yee_step_march descriptors of one dtype, tile_y and
 * eps_stride): three launches per step -- interior, z-face and boundary tiles -- each covering
 * every grid of the batch.  tfdtd_batch_march_pack writes the internal table (tfdtd_batch_march_bytes
 * bytes) and block_prefix int32[3][count+1]; the caller copies both to device memory once and passes
 * totals[v] = block_prefix[v][count] (host memory) to every tfdtd_batch_march call.  A batch of
 * one grid is also valid and runs the single-grid step from a device-resident descriptor. */
size_t tfdtd_batch_march_bytes(int32_t count, int32_t dtype);
int tfdtd_batch_march_pack(const tfdtd_march* steps, int32_t count, void* host_table, size_t bytes, int32_t* block_prefix);
int tfdtd_batch_march(const void* device_table, const int32_t* device_prefix, int32_t count, int32_t dtype,
                      int32_t tile_y, int32_t eps_stride, const int32_t* totals, void* stream);

/* The same for frequency planes of many grids (torchfdtd/cuda_batch.py:74-169, FusedBatchIO, and
 * torchfdtd/cuda_monitors.py:14-155): one launch for every plane of the batch. */
size_t tfdtd_batch_planes_bytes(int32_t count, int32_t dtype, int32_t spectral);
int tfdtd_batch_planes_pack(const tfdtd_plane* planes, int32_t count, void* host_table, size_t bytes,
                            int32_t* max_blocks, int32_t* max_phases);
/* max_phases: the largest frequencies x max(block, 1) of the packed planes (the launch's LDS). */
int tfdtd_batch_dft_update(const void* device_table, int32_t count, int32_t max_blocks, int32_t max_phases,
                           int32_t dtype, int32_t spectral, void* stream);

/* State norms of the run-control diagnostics (torchfdtd/run_control.py:127-159 and the fused
 * reduction of torchfdtd/cuda_diagnostics.py:63-135), per grid in FP64:
 *   out[g] = { sum_c vol E_c^2 / inv_eps_c + vol H_c^2,  max |E|,|H|,  1 if any value of E, H
 *             or the CPML memory is not finite else 0 },
 * vol = widths[0][x] * widths[1][y] * widths[2][z] (cell widths over the reference step; all
 * three null: 1).  tfdtd_norm_pack writes the table (tfdtd_norm_table_bytes bytes) for the caller
 * to copy to device memory once; tfdtd_state_norm then needs partial double[tasks][3] and
 * out double[count][3] in device memory, tasks = tfdtd_norm_tasks. */
typedef struct tfdtd_norm_grid {
    int32_t dtype;
    int32_t nx, ny, nz;
    const void* E;
    const void* H;
    const void* inverse_permittivity;
    int32_t eps_stride;             /* 1 per cell, 3 per cell and component */
    const double* widths[3];
    const void* psi[2];             /* CPML memory buffers (null: none) */
    int64_t psi_count[2];           /* elements of each */
} tfdtd_norm_grid;

int64_t tfdtd_norm_tasks(const tfdtd_norm_grid* grids, int32_t count);
size_t tfdtd_norm_table_bytes(const tfdtd_norm_grid* grids, int32_t count);
int tfdtd_norm_pack(const tfdtd_norm_grid* grids, int32_t count, void* host_table, size_t bytes);
int tfdtd_state_norm(const void* device_table, int32_t count, int64_t tasks, int32_t dtype, double* partial, double* out,
                     void* stream);

/* tfdtd_state_norm without the caller's partial buffer, and without any allocation: the partial
 * sums live in a buffer of TFDTD_NORM_SCRATCH_TASKS tasks (6 MiB) that the library holds in each
 * device's memory until the process ends.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under graph capture, serving loops, harnesses that forbid
 * hipMalloc).  tasks beyond the limit return TFDTD_EINVAL.  Every call on one device shares that
 * device's buffer: calls that overlap on one device (two streams, or two host threads) race and
 * must be ordered by the caller.  The result equals tfdtd_state_norm's bit for bit. */
#define TFDTD_NORM_SCRATCH_TASKS 262144
int tfdtd_state_norm_noalloc(const void* device_table, int32_t count, int64_t tasks, int32_t dtype, double* out,
                             void* stream);

/* Run-control decision at a diagnostic barrier, on the device (replaces the host-side
 * DecayDecision.update of torchfdtd/run_control.py:58-85 and the host check of
 * torchfdtd/run_control.py:127-159), so that a barrier enqueues work and never waits for the GPU.
 * Per grid g it reads the norms out[g] = {energy, field peak, non-finite flag} that
 * tfdtd_state_norm wrote and applies, in order:
 *   non-finite flag, energy or peak                          -> verdict NONFINITE (value 1 when
 *                                                               energy and peak are finite: the
 *                                                               CPML memory alone, else 0)
 *   divergence_check and field_limit > 0 and peak > limit     -> verdict FIELD_LIMIT (value = peak)
 *   eligible = step >= min_steps and step * time_step >= eligible_time
 *   peak_energy = max(peak_energy, energy); at the first eligible check the baseline freezes at it
 *   ratio = energy / baseline (baseline = peak_energy until frozen; 0 when the baseline is 0)
 *   divergence_check and eligible and ratio > growth_limit    -> verdict GROWTH (value = ratio)
 *   below = eligible and ratio <= decay_threshold ? below + 1 : 0
 *   history[g][checks++] = {step, step * time_step, energy, ratio, peak, eligible}
 *   auto_shutoff and below >= consecutive_checks              -> verdict DECAYED
 * A grid with a verdict other than GOING is left unchanged by later calls.  rules and states are
 * device arrays of `count`; zero the states before the run.  history holds rows_per_grid rows of 6
 * doubles per grid (rows beyond it are dropped; the states still count them). */
enum { TFDTD_RUN_GOING = 0, TFDTD_RUN_DECAYED = 1, TFDTD_RUN_NONFINITE = 2, TFDTD_RUN_FIELD_LIMIT = 3,
       TFDTD_RUN_GROWTH = 4 };

typedef struct tfdtd_decay_rule {
    double eligible_time;           /* end of the sources plus after_source_s, in s */
    double time_step;               /* s */
    double decay_threshold;
    double growth_limit;
    double field_limit;             /* 0: no limit */
    int32_t min_steps;
    int32_t consecutive_checks;
    int32_t divergence_check;
    int32_t auto_shutoff;
} tfdtd_decay_rule;

typedef struct tfdtd_decay_state {
    double peak_energy;             /* largest energy so far */
    double baseline;                /* frozen baseline (valid once frozen != 0) */
    double verdict_value;           /* field peak (FIELD_LIMIT), ratio (GROWTH), CPML-only flag (NONFINITE) */
    int64_t verdict_step;
    int32_t frozen;
    int32_t below;                  /* eligible checks in a row at or under the threshold */
    int32_t verdict;                /* TFDTD_RUN_* */
    int32_t checks;                 /* checks evaluated (history rows written, up to rows_per_grid) */
} tfdtd_decay_state;

int tfdtd_decay_update(const tfdtd_decay_rule* rules, tfdtd_decay_state* states, const double* norms, int32_t count,
                       int64_t step, double* history, int32_t rows_per_grid, void* stream);

/* The source boxes of a half step alone (torchfdtd/solver.py:778-783): dst[box][comp] +=
 * values[*counter] in source order, without the step's curl update.  A grid whose E update is
 * followed by a dispersive-material correction runs its E half step with nsrc = 0, the correction,
 * then this call, in the reference's order. */
int tfdtd_inject(const tfdtd_half_step* step, void* stream);

/* Passive Drude/Lorentz materials (torchfdtd/materials.py:29-87, MaterialADE): the coupled
 * trapezoidal update of one material's owned samples around the E half step.
 *   tfdtd_ade_save     before the E update: old = E at the samples;
 *   tfdtd_ade_correct  after it, before the E sources, per sample and pole j:
 *     response_j = (Q_j - a_j P_j)/d_j
 *     new = (eps_inf E - k_sum old - sum_j response_j)/(eps_inf + k_sum)
 *     delta_j = response_j + k_j (new + old);  P_j += delta_j;  Q_j = -Q_j + 2 delta_j;  E = new
 * with a_j = (w0 dt)^2/2, d_j = 1 + g dt/2 + (w0 dt)^2/4, k_j = s dt^2/(4 d_j) (computed by the
 * caller in double, rounded to the field precision here).  One pole forms eps_inf + k in double
 * and rounds it once; several poles sum k_j and add eps_inf in the field precision (the
 * reference's scalar and array coefficients).  Samples are field elements: per_cell = 1 takes
 * index[i/3] as a cell and component i % 3 ('cell' sampling), per_cell = 0 takes index[i] as an
 * element index ((x*ny+y)*nz+z)*3+c ('yee' sampling).  P and Q are real[poles][samples]; old is
 * real[samples], written by save and read by correct. */
#define TFDTD_ADE_MAX_POLES 16
typedef struct tfdtd_ade {
    int32_t dtype;
    int32_t poles;
    int64_t samples;
    int32_t per_cell;
    const int64_t* index;
    void* field;
    void* old;
    void* P;
    void* Q;
    double a[TFDTD_ADE_MAX_POLES], d[TFDTD_ADE_MAX_POLES], k[TFDTD_ADE_MAX_POLES];
    double eps_inf;
} tfdtd_ade;

int tfdtd_ade_save(const tfdtd_ade* ade, void* stream);
int tfdtd_ade_correct(const tfdtd_ade* ade, void* stream);

/* Forward record for the adjoint: curl H of the E update (backward differences, periodic wrap
 * included) on the design box cells [box_lo, box_hi), written to
 * record[counter % modulus][box cell][3] -- read before the step's E update.  With the E-family
 * CPML `e` (psi_in = the memory before this step's E update; psi_out is not written) a term in a
 * slab row records the curl the update applies there, d inv_k + (b psi + c d); with e.psi_in
 * NULL every term records the plain difference d.  No reference
 * counterpart: the reference recomputes this curl from replayed states
 * (torchfdtd/differentiable.py:508-509, torchfdtd/cuda_adjoint.py:108-216). */
typedef struct tfdtd_curl_record {
    int32_t dtype;
    int32_t nx, ny, nz;
    const void* H;
    int32_t periodic_mask;
    int32_t box_lo[3], box_hi[3];
    void* record;
    int64_t modulus;
    const int64_t* counter;
    tfdtd_march_family e;
} tfdtd_curl_record;

int tfdtd_record_curl(const tfdtd_curl_record* record, void* stream);

/* One step of the discrete adjoint, backwards (torchfdtd/differentiable.py:496-519 and
 * torchfdtd/cuda_adjoint.py:218-234, FusedAdjointCUDA.step): for t = *counter,
 *   e_bar/h_bar[target k] += sum of seeds[t][seed_columns[seed_offsets[k] .. seed_offsets[k+1])]
 *   e_bar += curl_T(forward)(-(S inv_mu) h_bar)            (H-family memory adjoints e.psi/h.psi)
 *   gradient[box cell] += -S e_bar . curl_H(t) inv_eps^2   (curl from record slot t % record_modulus)
 *   h_bar += curl_T(backward)((S inv_eps) e_bar)
 *   *counter = t - 1.
 * The memory adjoints use the forward psi layouts of tfdtd_march_family: psi_in holds the
 * adjoint after the step and psi_out receives the adjoint before it (separate buffers, which
 * the caller alternates); coef is the forward coefficient block.  seed_targets are
 * int64[nseed][2] = (0 E | 1 H, element index).  The gradient is double[box cells][stride]
 * (stride 1 sums the components).  With record == NULL no gradient is formed; a step with t < 0
 * adds neither seeds nor gradient.  A box that meets an E-family CPML slab needs a record made
 * with that CPML (record_cpml = 1): the gradient there uses the curl the update applied. */
typedef struct tfdtd_adjoint {
    int32_t dtype;
    int32_t nx, ny, nz;
    void* e_bar;
    void* h_bar;
    const void* inverse_permittivity;
    int32_t eps_stride;
    const void* inverse_permeability;   /* one value */
    double courant;
    int32_t periodic_mask;
    int32_t pec_upper_mask;
    double pec_upper_factor[3];
    tfdtd_march_family e;
    tfdtd_march_family h;
    int32_t nseed;
    const int64_t* seed_targets;
    const int32_t* seed_offsets;
    const int32_t* seed_columns;
    const void* seeds;                  /* real[seed_rows][seed_width] */
    int32_t seed_width;
    int64_t seed_rows;                  /* steps t outside [0, seed_rows) add no seed */
    int32_t box_lo[3], box_hi[3];
    const void* record;
    int64_t record_modulus;
    double* gradient;
    int32_t gradient_stride;
    int64_t* counter;
    int32_t record_cpml;                /* 1: the record carries the E-family CPML terms */
} tfdtd_adjoint;

int tfdtd_adjoint_step(const tfdtd_adjoint* step, void* stream);

/* BF16 implicit-GEMM 3D convolution on MFMA with FP32 accumulation (the learned screening
 * model of the inverse-design loop; no counterpart in the reference, which has no neural
 * surrogate).  Activations are channels-last bf16 [n][d][h][w][c]; a kernel of ks^3 taps
 * (ks 1 or 3), stride 1 or 2, zero padding `pad`.  Forward weights are bf16 [co][kp] with
 * k = ((kd*ks + kh)*ks + kw)*ci + c, zero beyond ks^3*ci up to kp (a multiple of 32). */
typedef struct tfdtd_conv {
    int32_t n, di, hi, wi, ci;
    int32_t co;
    int32_t ks, stride, pad;
    int32_t kp;
} tfdtd_conv;

int tfdtd_conv_output(const tfdtd_conv* conv, int32_t* dout, int32_t* hout, int32_t* wout);
/* pre[m][co] = sum_k im2col(x)[m][k] W[co][k] + bias[co] (FP32); act = bf16(GELU(pre)) or bf16(pre). */
int tfdtd_conv_forward(const tfdtd_conv* conv, const void* x, const void* weight, const float* bias, int32_t activation,
                       float* pre, void* act, void* stream);
/* dx[m_in][ci] (FP32) from dy bf16 [m_out][co] and weight_t bf16 [ci][ks^3*co] (k = tap*co + c). */
int tfdtd_conv_dgrad(const tfdtd_conv* conv, const void* dy, const void* weight_t, float* dx, void* stream);
/* dw[co][kp] (FP32) from dy_t bf16 [co][m_out] and x; K split over `splits` workgroup layers into
 * partial float[splits][co][kp], then summed in split order. */
int tfdtd_conv_wgrad(const tfdtd_conv* conv, const void* dy_t, const void* x, int32_t splits, float* partial, float* dw,
                     void* stream);
/* dpre = dy * GELU'(pre) (activation 1) or dy (0), in FP32 and bf16. */
int tfdtd_gelu_backward(const float* dy, const float* pre, int64_t count, int32_t activation, float* dpre, void* dpre16,
                        void* stream);

/* Copies the calling thread's last error message; returns its length. */
size_t tfdtd_last_error(char* buffer, size_t size);

int tfdtd_abi_version(void);

/* SHA-256 (hex) of the sources the library was compiled from (csrc and this header), set by the
 * build; a loader compares it with the sources beside it and refuses a stale library.  No
 * reference counterpart (the reference compiles its kernels at run time). */
const char* tfdtd_source_digest(void);

#ifdef __cplusplus
}
#endif

#endif /* TORCHFDTD_AMD_TFDTD_H_ */
