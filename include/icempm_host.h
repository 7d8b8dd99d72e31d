/*
 * icempm_host.h -- C entry points of libicempm_host.so, the C++ drop-in
 * surface (icy::SimParams, icy::Model, GPU_Implementation5,
 * icy::SnapshotManager ingest) built over libicempm_hip.so.
 *
 * These let non-C++ callers (the Python tests and bench) drive exactly the
 * code path a C++ frontend uses.  Each function cites the reference routine
 * whose behaviour it exposes.  Status codes as in icempm.h; the text of the
 * last failure (the std::runtime_error message) is returned by
 * icempm_model_last_error().
 */
#ifndef ICEMPM_HOST_H
#define ICEMPM_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "icempm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct icempm_model icempm_model;

/* icy::SimParams methods on a caller-owned 432-byte buffer
 * (parameters_sim.cpp:5-56, 121-147) */
size_t icempm_host_simparams_size(void);
void icempm_simparams_reset(void *prms);                  /* Reset() */
void icempm_simparams_derive(void *prms);                 /* ComputeCamClayParams2 + ComputeHelperVariables */
/* ParseFile (parameters_sim.cpp:59-119); raw_points receives InputRawPoints */
int icempm_simparams_parse_file(void *prms, const char *path, char *raw_points, size_t cap,
                                char *err, size_t errcap);

/* icy::Model (model.cpp:4-9) with GPU_Implementation5::initialize */
int icempm_model_create(icempm_model **out);
void icempm_model_destroy(icempm_model *m);
const char *icempm_model_last_error(const icempm_model *m);
/* the model's icy::SimParams (read/write) */
void *icempm_model_params(icempm_model *m);
icempm_ctx *icempm_model_context(icempm_model *m);
/* SnapshotManager::LoadRawPoints on arrays (snapshotmanager.cpp:186-286):
 * xy = n interleaved f32 pairs, grains = n i16, volume = body area */
int icempm_model_load_raw_points(icempm_model *m, const float *xy, const int16_t *grains, int n,
                                 float volume);
/* Model::Prepare (model.cpp:67-72): uploads the current parameters */
int icempm_model_prepare(icempm_model *m);
/* Model::Step (model.cpp:11-47): one output cycle; *cont = its return value */
int icempm_model_step(icempm_model *m, int *cont);
/* Waits until the readback queued by the last Step has been consumed */
int icempm_model_wait_cycle(icempm_model *m);
float icempm_model_cycle_time(const icempm_model *m);
/* pinned host buffers of the readback (tmp_transfer_buffer and
 * host_side_indenter_force_accumulator) */
double *icempm_model_host_points(icempm_model *m);
double *icempm_model_host_force(icempm_model *m);
/* GPU_Implementation5::set_mode (FAST / ORDERED) */
int icempm_model_set_mode(icempm_model *m, int mode);

/* Synthetic jittered lattice in the raw-point format: for cy < ncy,
 * cx < ncx, b < k, a < k one point at ((cx + (a+0.5)/k) h, (cy + (b+0.5)/k) h)
 * plus a uniform jitter in [-j h, j h) per axis drawn from splitmix64
 * (53-bit uniforms, x then y).  xy receives 2*ncx*ncy*k*k floats, grains
 * zeros.  Returns the point count. */
long icempm_scene_lattice(int ncx, int ncy, int k, double h, double j, uint64_t seed, float *xy,
                          int16_t *grains);

#ifdef __cplusplus
}
#endif
#endif
