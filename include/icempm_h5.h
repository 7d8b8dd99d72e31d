/*
 * icempm_h5.h -- C entry points of libicempm_h5.so: the HDF5 files of
 * icy::SnapshotManager (reference: simulation/snapshotmanager.cpp), built on
 * the HDF5 C library.  Status codes as in icempm.h; a failure's text is
 * returned by icempm_model_last_error().
 */
#ifndef ICEMPM_H5_H
#define ICEMPM_H5_H

#include <stddef.h>
#include <stdint.h>

#include "icempm_host.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SnapshotManager::LoadRawPoints(fileName) (snapshotmanager.cpp:186-286):
 * Points_Raw_2D f32[N][2], GrainIDs i16[N] with a f32 attribute "volume";
 * then the ingest of icempm_model_load_raw_points */
int icempm_model_load_raw_points_file(icempm_model *m, const char *path);
/* SnapshotManager::SaveSnapshot (snapshotmanager.cpp:60-105): writes
 * <dir>/snapshots/dNNNNN.h5 (NNNNN = SimulationStep / UpdateEveryNthStep)
 * from the readback buffers -- Indenter_2D f64[2 n_indenter_subdivisions],
 * Params u8[432] (the raw icy::SimParams, with the 14 attributes of
 * SaveParametersAsAttributes, snapshotmanager.cpp:107-145), Points
 * f64[16 nPtsPitch] */
int icempm_model_save_snapshot(icempm_model *m, const char *dir);
/* The same snapshot file from caller-held data (no counterpart in the
 * reference; for cm's batch mode, which has no icy::Model per case):
 * simparams is a full icy::SimParams (nbytes == its size) whose
 * SimulationStep, SimulationTime and indenter_x/y describe the state;
 * indenter_force holds 2 n_indenter_subdivisions doubles and points the
 * pitched SoA, 16 nPtsPitch doubles.  Writes <dir>/snapshots/dNNNNN.h5. */
int icempm_h5_write_snapshot(const char *dir, const void *simparams, size_t nbytes,
                             const double *indenter_force, const double *points);
/* Resume from a snapshot written by icempm_model_save_snapshot (the
 * reference's ReadSnapshot, snapshotmanager.cpp:147-184, is commented out):
 * parameters, arrays and points restored and uploaded; stepping continues
 * from the snapshot's SimulationStep and SimulationTime */
int icempm_model_read_snapshot(icempm_model *m, const char *path);
/* SnapshotManager::SavePQ (snapshotmanager.cpp:14-56): <dir>/pq/pqNNNNN.h5
 * with PQ_intact f64[n][2] (with the attributes) and PQ_crushed f64[n][2] */
int icempm_model_save_pq(icempm_model *m, const char *dir);
/* a raw point file in the format LoadRawPoints reads (no counterpart: the
 * reference's inputs were made elsewhere); deflate 0..9 (0: contiguous,
 * otherwise chunked and compressed, like the reference's input files) */
int icempm_h5_write_raw_points(const char *path, const float *xy, const int16_t *grains, int n,
                               float volume, int deflate);

#ifdef __cplusplus
}
#endif
#endif
