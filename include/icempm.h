/*
 * icempm.h -- C ABI of libicempm_hip.so, the MI355X (gfx950) backend of the
 * iceMPM 2D MLS-MPM substep.
 *
 * Each entry point replaces one member of the reference's device backend
 * class GPU_Implementation5 (simulation/gpu_implementation5.h:40-72,
 * simulation/gpu_implementation5.cu:434-593) or one step of the host loop
 * icy::Model::Step (simulation/model.cpp:11-47).  The C++ drop-in surface in
 * icempm_amd/host/ (same class names and members as the reference) calls
 * only these functions.
 *
 * Conventions
 *   - Every function returns 0 (ICEMPM_OK) or a nonzero status; the text of
 *     the last failure is returned by icempm_last_error().  The C++ surface
 *     rethrows with the reference's std::runtime_error tags.
 *   - All device work of a context is queued on one HIP stream, in call
 *     order.  Calls on one context must not overlap (the reference calls
 *     every method from one simulation thread).
 *   - Parameters are the reference's icy::SimParams POD, passed verbatim
 *     (parameters_sim.h:19-90, byte layout in icempm_amd/host/parameters_sim.h).
 *   - Particle and grid buffers keep the reference layout: points are 16
 *     rows x nPtsPitch doubles (row 0 utility bytes: crushed u8 at byte i,
 *     grain i16 at byte pitch+2i; rows 1..15 P, Q, Jp_inv, pos, vel, Fe, Bp),
 *     the grid is 3 rows x nGridPitch doubles (mass, x, y).
 */
#ifndef ICEMPM_H
#define ICEMPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct icempm_ctx icempm_ctx;

enum {
    ICEMPM_OK = 0,
    ICEMPM_E_HIP = 1,      /* a HIP runtime call failed */
    ICEMPM_E_ARG = 2,      /* bad argument (null, size mismatch, out of range) */
    ICEMPM_E_STATE = 3,    /* call out of order (e.g. step before allocate) */
    ICEMPM_E_NODEV = 4     /* no HIP device */
};

/* P2G summation modes (icempm_set_mode).  FAST accumulates in LDS and then
 * with hardware atomics (summation order varies run to run, as in the
 * reference).  ORDERED sums every grid node's contributions in increasing
 * particle index, and the indenter force in increasing node index, so the
 * result is bit-reproducible and equal to a sequential restatement. */
enum { ICEMPM_MODE_FAST = 0, ICEMPM_MODE_ORDERED = 1 };

/* -- lifecycle ---------------------------------------------------------- */

/* GPU_Implementation5::initialize (gpu_implementation5.cu:452-469):
 * selects device 0, creates the stream and the cycle events. */
int icempm_create(icempm_ctx **out);
/* No counterpart (the reference never frees); releases every resource. */
void icempm_destroy(icempm_ctx *ctx);
/* Text of the last failure on this context ("" if none).  ctx may be NULL
 * for failures of icempm_create. */
const char *icempm_last_error(const icempm_ctx *ctx);
/* sizeof(icy::SimParams) that this library was built against (432 on x86-64). */
size_t icempm_simparams_size(void);
/* GPU_Implementation5::test (gpu_implementation5.cu:434-440): launches a
 * trivial kernel and waits for it. */
int icempm_test(icempm_ctx *ctx);
/* GPU_Implementation5::synchronize (gpu_implementation5.cu:442-446). */
int icempm_synchronize(icempm_ctx *ctx);

/* -- setup -------------------------------------------------------------- */

/* GPU_Implementation5::cuda_update_constants (gpu_implementation5.cu:471-479):
 * copies the SimParams POD (nbytes must equal icempm_simparams_size()) and
 * resets the device error flag to the host error_code. */
int icempm_update_constants(icempm_ctx *ctx, const void *simparams, size_t nbytes);

/* GPU_Implementation5::cuda_allocate_arrays (gpu_implementation5.cu:481-513):
 * (re)allocates the grid, the point SoA, the force bins and the pinned host
 * buffers.  Returns the pitches (in doubles) and the pinned host pointers
 * (tmp_transfer_buffer, host_side_indenter_force_accumulator) that the
 * library owns.  nBins is 2*n_indenter_subdivisions doubles. */
int icempm_allocate_arrays(icempm_ctx *ctx, size_t nGridNodes, size_t nPoints, int nBins,
                           size_t *nGridPitch, size_t *nPtsPitch, double **host_points,
                           double **host_force);

/* GPU_Implementation5::transfer_ponts_to_device (gpu_implementation5.cu:515-526):
 * blocking copy of the pinned point buffer to the device; zeroes the host
 * force buffer. */
int icempm_transfer_points_to_device(icempm_ctx *ctx);

/* -- the substep, one operation at a time (model.cpp:25-28) ------------- */

/* cuda_reset_grid (gpu_implementation5.cu:554-559) */
int icempm_reset_grid(icempm_ctx *ctx);
/* cuda_reset_indenter_force_accumulator (gpu_implementation5.cu:561-565) */
int icempm_reset_indenter_force_accumulator(icempm_ctx *ctx);
/* cuda_p2g (gpu_implementation5.cu:568-575): v2_kernel_p2g */
int icempm_p2g(icempm_ctx *ctx);
/* cuda_update_nodes (gpu_implementation5.cu:577-584): v2_kernel_update_nodes */
int icempm_update_nodes(icempm_ctx *ctx, double indenter_x, double indenter_y);
/* cuda_g2p (gpu_implementation5.cu:586-593): v2_kernel_g2p */
int icempm_g2p(icempm_ctx *ctx, int recordPQ);

/* -- the substep loop on the device (model.cpp:20-29) -------------------
 * Runs n substeps starting at global step `first_step` and time *t:
 *   for each substep s = first_step .. first_step+n-1:
 *     if (s % UpdateEveryNthStep == 0 && reset_bins_at_cycle) zero the bins
 *     *t += dt; indenter_x = indenter_x_initial + *t * IndVelocity
 *     zero grid; P2G; update nodes (indenter_x, indenter_y); G2P with
 *     recordPQ = ((s+1) % UpdateEveryNthStep == 0)
 * *t is accumulated by repeated addition exactly as Model::Step does.  When
 * force_history is non-NULL (device or host pointer is not accepted: pass
 * NULL here and read icempm_force_history), the total indenter force of each
 * substep is recorded.  Kernels are fused where the result is unchanged. */
int icempm_run_substeps(icempm_ctx *ctx, int first_step, int n, double *t,
                        double indenter_x_initial, double indenter_y, int reset_bins_at_cycle);

/* Per-substep total indenter force (fx, fy) of the last icempm_run_substeps
 * call, copied to host: out receives 2*n doubles (n <= that call's n). */
int icempm_force_history(icempm_ctx *ctx, double *out, int n);

/* -- readback (gpu_implementation5.cu:528-552) -------------------------- */

typedef void (*icempm_callback)(void *user);
/* cuda_transfer_from_device: queues async copies of the point SoA, the
 * force bins and the error flag into the pinned host buffers, then calls
 * cb(user) on the runtime's callback thread when they have landed (cb may
 * be NULL).  The callback must not call into this library. */
int icempm_transfer_from_device(icempm_ctx *ctx, icempm_callback cb, void *user);
/* GPU_Implementation5::error_code: the host copy written by the last
 * completed icempm_transfer_from_device (lags like the reference). */
int icempm_error_code(const icempm_ctx *ctx);
/* Blocking read of the device error flag (no counterpart; for tests). */
int icempm_read_error_flag(icempm_ctx *ctx, int *flag);

/* -- cycle timing events (model.cpp:19,30,38-39) ------------------------ */
int icempm_event_record_start(icempm_ctx *ctx);
int icempm_event_record_stop(icempm_ctx *ctx);
/* waits for the stop event; milliseconds between start and stop */
int icempm_event_elapsed_ms(icempm_ctx *ctx, float *ms);

/* -- cell-key sort (snapshotmanager.cpp:288-304; keys as in
 *    gpu_implementation5.cu:185,325) --------------------------------------
 * Computes key_i = (int)(x_i*hinv - 0.5) + (int)(y_i*hinv - 0.5)*gridX on
 * the device with no contraction, sorts stably by key (ties keep the
 * previous order) with an LSD radix sort, and permutes the whole point SoA
 * in place.  perm_out (host, may be NULL) receives the previous index of
 * each new slot; keys_out (host, may be NULL) the sorted keys. */
int icempm_sort_points(icempm_ctx *ctx, int *perm_out, int *keys_out);

/* Stable sort of raw float points by the reference's load-time key
 * (SnapshotManager::SortPoints): xy is n interleaved float pairs, grains n
 * shorts; both are permuted in place (host buffers; the sort runs on the
 * device).  hinv = 1/cellsize, gridX = GridX. */
int icempm_sort_raw_points(icempm_ctx *ctx, float *xy, int16_t *grains, int n, double hinv,
                           int gridX);

/* -- configuration (no counterpart) ------------------------------------- */
int icempm_set_mode(icempm_ctx *ctx, int mode);
int icempm_get_mode(const icempm_ctx *ctx);

/* Tuning knobs (no counterpart).  ICEMPM_TUNE_PASS_KIND selects the kernel
 * of the particle pass: 0 workgroup-wide shared-memory windows for gather and
 * scatter, 1 wavefront windows (direct gather, per-wavefront scatter), 2 as 1
 * with the per-lane stencil rotation (default).  Results differ only in the
 * order of the FAST-mode P2G sums. */
enum { ICEMPM_TUNE_PASS_KIND = 0 };
int icempm_set_tuning(icempm_ctx *ctx, int knob, int value);
/* current value, or -1 for an unknown knob */
int icempm_get_tuning(const icempm_ctx *ctx, int knob);

/* Device-side timing of the dominant kernel (the fused G2P+P2G pass, or
 * G2P in ORDERED mode) over icempm_run_substeps calls: when enabled, each
 * launch is bracketed by events on the context's stream and the durations
 * are summed.  Returns total milliseconds and the launch count. */
int icempm_kernel_timing(icempm_ctx *ctx, int enable);
int icempm_kernel_timing_read(icempm_ctx *ctx, double *total_ms, int *launches);

/* Device copy bandwidth (no counterpart): reps launches of a plain streaming
 * copy kernel (k_copy) over nbytes; *gbps = (read + written bytes) / time. */
int icempm_measure_copy(icempm_ctx *ctx, size_t nbytes, int reps, double *gbps);

/* Raw device buffers for tests: copies the device grid (3*nGridPitch
 * doubles) or the point SoA (16*nPtsPitch doubles) to host memory. */
int icempm_read_grid(icempm_ctx *ctx, double *out);
int icempm_read_points(icempm_ctx *ctx, double *out);
/* and writes a host grid to the device (3*nGridPitch doubles) */
int icempm_write_grid(icempm_ctx *ctx, const double *in);
/* device bins (2*n_indenter_subdivisions doubles) to host */
int icempm_read_bins(icempm_ctx *ctx, double *out);

#ifdef __cplusplus
}
#endif
#endif
