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
 *     the last failure is returned by icempm_last_error().  No C++ exception
 *     leaves a function of this interface (a failed host allocation returns
 *     ICEMPM_E_HOST).  The C++ surface rethrows with the reference's
 *     std::runtime_error tags.
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
    ICEMPM_E_NODEV = 4,    /* no HIP device */
    ICEMPM_E_IO = 5,       /* file input/output (libicempm_h5.so) */
    ICEMPM_E_HOST = 6      /* host memory (or a thread) could not be allocated; the handle's
                              state is then unspecified: destroy it */
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
/* Waits until the copies queued by the last icempm_transfer_from_device
 * (points, bins, error flag) have landed in the pinned buffers -- not for
 * its callback -- and refreshes error_code from the copied flag.  The
 * reference's cuda_transfer_from_device copies the flag into the pageable
 * member error_code (gpu_implementation5.cu:539-540), a copy that returns
 * only once it has completed, so Model::Step (model.cpp:46) tests the flag
 * of the cycle it has just run; GPU_Implementation5 calls this for the same
 * effect.  No-op when nothing is pending. */
int icempm_transfer_wait(icempm_ctx *ctx);
/* GPU_Implementation5::error_code: the flag of the last readback whose
 * copies have landed (icempm_transfer_wait or the readback callback). */
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

/* -- allocation-free variant (no counterpart in the reference) ----------
 * For callers that cannot allocate device memory (serving loops, harnesses
 * that forbid hipMalloc): icempm_sort_raw_points with the same arguments,
 * results and blocking behaviour, but every device buffer it needs -- the
 * staged points and grains, the keys, the values and the radix scratch --
 * lives in a static device array of the library (about 36 MB, one copy on
 * each GPU the library runs on, held until the process ends) whose address
 * icempm_create looks up.  It calls no hipMalloc or hipFree.  n may be at
 * most ICEMPM_NOALLOC_MAX_POINTS (ICEMPM_E_ARG with a message otherwise).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from several threads or contexts) race. */
#define ICEMPM_NOALLOC_MAX_POINTS 1048576
int icempm_sort_raw_points_noalloc(icempm_ctx *ctx, float *xy, int16_t *grains, int n, double hinv,
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

/* -- batches of independent cases (no counterpart) ----------------------
 * Many indentation cases of uneven size advanced together: every substep is
 * one grid-update launch and one particle-pass launch over all cases, and
 * the substep loop can be captured in a HIP graph and replayed.  Cases are
 * added from the reference SoA layout (16 rows x pitch doubles, as read back
 * by icempm_read_points after LoadRawPoints); all cases share the time step.
 * A case with GrainVariability != 0 takes each point's NACC strengths from
 * its grain id (row 0 bytes [pitch, pitch + 2n), GetParametersForGrain,
 * gpu_implementation5.cu:97-111); otherwise its material is uniform.
 * Precision: ICEMPM_BATCH_F64 stores the point state in FP64;
 * ICEMPM_BATCH_F32 stores velocity, Fe - I, Bp and Jp_inv - 1 in FP32 and the
 * positions as 48-bit fixed point in units of cellsize * 2^-F, F = 48 - the
 * bits of max(GridX, GridY) (grids up to 65536 nodes a side; a quantum of at
 * most 2 * GridX * cellsize / 2^48), each point's position (with its crushed flag) and
 * its Fe - I as one 16-byte record apiece; the grid sums stay FP64, the grid velocities
 * reach the gathers as FP32 pairs, and all arithmetic is FP64.  The
 * per-substep total indenter force of every case is recorded, and its
 * angular bins (Indenter_2D, gpu_implementation5.cu:262-269) summed over
 * each icempm_batch_run call. */
typedef struct icempm_batch icempm_batch;
enum { ICEMPM_BATCH_F64 = 0, ICEMPM_BATCH_F32 = 1 };
int icempm_batch_create(icempm_batch **out, int precision);
void icempm_batch_destroy(icempm_batch *b);
const char *icempm_batch_last_error(const icempm_batch *b);
/* returns the case index (>= 0) or -status */
int icempm_batch_add_case(icempm_batch *b, const void *simparams, size_t nbytes, const double *soa,
                          size_t pitch);
int icempm_batch_finalize(icempm_batch *b);
/* n substeps of every case from the batch's current time; use_graph != 0
 * captures the loop once per n (and start, below) and replays it, keeping
 * the graphs of the two most recently used step counts per start.  The last
 * pass of a run leaves the next substep's P2G in the batch's grids, made
 * from the points as the run leaves them, and the next run starts from it;
 * after icempm_batch_restore or a re-sort the next run makes its own. */
int icempm_batch_run(icempm_batch *b, int n, int use_graph);
/* Host set-up of graph runs of n substeps ahead of time: the time-table and
 * history buffers, and the graphs of such a run from each of its possible
 * starts (fresh, or continuing a previous run), captured, instantiated and
 * uploaded.  A run of n substeps then launches at once, whatever runs come
 * between (shorter runs keep the buffers; runs of two other lengths from the
 * same start replace that start's graph).  Preparing the longer of two run
 * lengths first keeps both. */
int icempm_batch_prepare(icempm_batch *b, int n);
/* device time of the last icempm_batch_run (HIP events on the batch stream;
 * a graph run is timed from its launch, after capture, instantiation and
 * hipGraphUpload of the graph) */
int icempm_batch_last_ms(const icempm_batch *b, float *ms);
int icempm_batch_num_points(const icempm_batch *b);
long icempm_batch_num_nodes(const icempm_batch *b);
/* Device timing of the fused particle pass: with stride > 0, every
 * stride-th fused particle kernel of a stream run (use_graph = 0; a graph
 * run then fails) is bracketed by events and the durations are summed (the
 * learned map's kernel that follows it is outside the bracket); stride 0
 * switches it off.  Enabling resets the sums. */
int icempm_batch_kernel_timing(icempm_batch *b, int stride);
int icempm_batch_kernel_timing_read(icempm_batch *b, double *total_ms, int *launches);
/* total indenter force (fx, fy) of case c for each substep of the last run */
int icempm_batch_force_history(icempm_batch *b, int c, double *out, int n);
/* indenter force bins of case c summed over the last icempm_batch_run call:
 * out receives 2*nbins doubles (fx, fy interleaved per angular bin, the
 * layout of host_side_indenter_force_accumulator); nbins must be the case's
 * n_indenter_subdivisions */
int icempm_batch_force_bins(icempm_batch *b, int c, double *out, int nbins);
/* the state of case c in the reference SoA layout (P and Q zero) */
int icempm_batch_read_points(icempm_batch *b, int c, double *soa, size_t pitch);
int icempm_batch_error_flag(icempm_batch *b, int *flag);
/* Tuning knobs: ICEMPM_BATCH_TUNE_CHUNK, points per chunk = threads per
 * workgroup of the particle pass (64, 128 or 256; default 256);
 * ICEMPM_BATCH_TUNE_RESORT, re-sort the points by cell every that many
 * substeps within a run (0, the default: never);
 * ICEMPM_BATCH_TUNE_MAP_MIN_ACCEPT, the learned map's acceptance threshold
 * in percent (default 50) below which icempm_batch_autotune switches the
 * network off for a branch.  get returns -1 for an unknown knob. */
enum { ICEMPM_BATCH_TUNE_CHUNK = 0, ICEMPM_BATCH_TUNE_RESORT = 1, ICEMPM_BATCH_TUNE_MAP_MIN_ACCEPT = 2 };
int icempm_batch_set_tuning(icempm_batch *b, int knob, int value);
int icempm_batch_get_tuning(const icempm_batch *b, int knob);
/* stable re-sort of every case's points by current cell (reads keep returning
 * the points in their original order) and a new chunk table */
int icempm_batch_resort(icempm_batch *b);
/* device copy of the point state, time and step (one checkpoint per batch);
 * restore returns to it */
int icempm_batch_checkpoint(icempm_batch *b);
int icempm_batch_restore(icempm_batch *b);
/* Autotuner: from a checkpoint of the current state, times probe_steps
 * substeps for every points-per-chunk value and the re-sort cadences
 * {never, probe/2, probe/4}, each on the host clock over a second probe run
 * (re-sorts and graph captures included; without re-sorting the second run
 * replays the graph the first captured); report receives rows (chunk,
 * re-sort, ms per substep), *rows their count.  The point state, the map
 * statistics and the error flag are restored and the fastest setting is
 * left in place.  Refused in harvest mode (the probes would append pairs). */
int icempm_batch_autotune(icempm_batch *b, int probe_steps, double *report, int cap, int *rows);
/* Return map of crushed points (gpu_implementation5.cu:48-94 is the exact
 * one).  ICEMPM_MAP_EXACT: the exact FP64 map.  ICEMPM_MAP_HARVEST: the
 * exact map, and every crushed point with positive singular values also
 * records one (features, targets) pair, up to harvest_capacity pairs.
 * ICEMPM_MAP_LEARNED: the branch of the map is decided exactly (tension,
 * shear projection, or inside the surface, where the map is the identity);
 * for the two projecting branches the MLP below (params:
 * icempm_mlp_num_params() floats; norm: 16 floats = input mean[5], input
 * scale[5], output mean[3], output scale[3]) proposes the projected state
 * and the exact map replaces it wherever the yield residual (stress error
 * relative to the case's strengths) exceeds tol.
 * Pairs are harvested from the projecting branches only.  Features: log s0,
 * log s1, log Jp_inv of the trial state, 0 (no hardening variable), 1 in
 * the tension branch else 0.  Targets: the change of log s0, log s1 and
 * log Jp_inv made by the map. */
enum { ICEMPM_MAP_EXACT = 0, ICEMPM_MAP_HARVEST = 1, ICEMPM_MAP_LEARNED = 2 };
int icempm_batch_set_return_map(icempm_batch *b, int mode, const float *params, const float *norm,
                                double tol, int harvest_capacity);
/* pairs recorded since harvest mode was set: *total = pairs offered (may
 * exceed the capacity); copies min(total, capacity, cap) rows of X (n x 5)
 * and T (n x 3) */
int icempm_batch_read_harvest(icempm_batch *b, float *X, float *T, int cap,
                              unsigned long long *total);
/* out[0] points accepted from the MLP, out[1] points that fell back to the
 * exact map, out[2] crushed points inside the surface (no projection), since
 * the last reset; reset != 0 zeroes the counters after reading (these and the
 * per-branch ones) */
int icempm_batch_map_stats(icempm_batch *b, unsigned long long *out, int reset);
/* The network's proposals per branch of the map since the last reset:
 * out[0] tension accepted, out[1] tension fell back, out[2] shear accepted,
 * out[3] shear fell back (points the network was asked about) */
int icempm_batch_map_branch_stats(icempm_batch *b, unsigned long long *out, int reset);
/* Branches the network proposes for (ICEMPM_MAP_TENSION | ICEMPM_MAP_SHEAR,
 * both by default); a switched-off branch takes the exact map in the particle
 * pass, with no queue, and with both off the pass is the exact map's (no map
 * kernel, and the learned map's counters stay as they are).  In learned mode
 * icempm_batch_autotune switches off each branch whose proposals passed the
 * yield gate less often than ICEMPM_BATCH_TUNE_MAP_MIN_ACCEPT percent over
 * its probes; autotune_branch_stats returns those probes' counts (as
 * map_branch_stats). */
enum { ICEMPM_MAP_TENSION = 1, ICEMPM_MAP_SHEAR = 2 };
int icempm_batch_set_map_branches(icempm_batch *b, int mask);
int icempm_batch_get_map_branches(const icempm_batch *b);
int icempm_batch_autotune_branch_stats(const icempm_batch *b, unsigned long long *out);

/* -- the learned return map (no counterpart) -----------------------------
 * A 5 -> 64 -> 64 -> 3 MLP (tanh hidden layers) with BF16 weights and
 * activations on MFMA, FP32 accumulation, trained on the GPU with Adam.
 * Parameters: icempm_mlp_num_params() floats, layout in csrc/mlp.hip.
 * Inputs and targets are normalised by the caller; X is n x 5, T and Y are
 * n x 3, all FP32 device pointers. */
typedef struct icempm_mlp icempm_mlp;
int icempm_mlp_num_params(void);
int icempm_mlp_create(icempm_mlp **out, const float *init_params);
void icempm_mlp_destroy(icempm_mlp *m);
const char *icempm_mlp_last_error(const icempm_mlp *m);
int icempm_mlp_get_params(icempm_mlp *m, float *out);
const float *icempm_mlp_device_params(const icempm_mlp *m);
/* gradient of 1/(2B) sum |MLP(X) - T|^2 over rows [r0, r0+B) (host copy in
 * grad_out if not NULL) and the mean loss; lr > 0 also applies one Adam
 * step (beta 0.9, 0.999, eps 1e-8) */
int icempm_mlp_step(icempm_mlp *m, const float *X, const float *T, int r0, int B, float lr,
                    float *grad_out, float *loss_out);
int icempm_mlp_infer(icempm_mlp *m, const float *X, int n, float *Y);
/* The whole training of the learned map from harvested pairs (host arrays:
 * X n x 5 features, T n x 3 targets, as icempm_batch_read_harvest returns
 * them): fits the normalisation (norm_out: 16 floats, the layout
 * icempm_batch_set_return_map takes), shuffles the rows once (seeded),
 * draws Glorot-scaled initial weights (seeded), runs `steps` Adam steps on
 * consecutive windows of `batch` rows with the learning rate decaying from
 * lr to lr / 20 along a cosine, and writes the parameters to params_out
 * (icempm_mlp_num_params() floats).  trace (may be NULL) receives the window
 * loss of every trace_every-th step; final_loss (may be NULL) the loss over
 * all rows. */
int icempm_mlp_train(const float *X, const float *T, int n, int steps, int batch, float lr,
                     unsigned long long seed, float *params_out, float *norm_out, float *trace,
                     int trace_every, float *final_loss);
/* plain device memory for training data and tests */
int icempm_device_alloc(size_t nbytes, void **ptr);
void icempm_device_free(void *ptr);
int icempm_copy_to_device(void *dst, const void *src, size_t nbytes);
int icempm_copy_to_host(void *dst, const void *src, size_t nbytes);
/* one 16x16x16 BF16 MFMA tile on host matrices (row-major), for layout tests */
int icempm_mfma_probe(const float *A, const float *B, float *D);
/* D (16 x 16) = A (16 x K) * B (K x 16), row-major host matrices, through the
 * training kernels' tile routine: K = 16 the 16x16x16 form, K = 32 or 64 the
 * 16x16x32 form (one or two steps); for layout tests */
int icempm_mfma_probe_k(const float *A, const float *B, float *D, int K);

#ifdef __cplusplus
}
#endif
#endif
