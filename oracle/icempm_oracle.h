/*
 * icempm_oracle.h -- CPU restatement of iceMPM's 2D MLS-MPM substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path in icempm_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to check or to time the CPU
 * baseline.  Nothing in the product links it.
 *
 * Every function restates a reference routine (paths relative to the
 * reference tree); the operation order is written out explicitly and the
 * library is built with -ffp-contract=off, so each result is a defined IEEE
 * double sequence.  The HIP kernels follow the same sequence.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   oracle_svd2x2           numpy/LAPACK singular values, reconstruction
 *                           U S V^T = F, det U = det V = +1, special cases
 *   oracle_kirchhoff        closed forms F=I -> 0, F=sI -> kappa(s^4-1)I,
 *                           rotation invariance, a hand-expanded case
 *   oracle_compute_pq       closed forms from sigma; F=I -> p=q=0
 *   oracle_grain_params     table for grains 0..29 computed by hand formula
 *   oracle_failure_check    points on both sides of each boundary
 *   oracle_dp_return        invariants: q(sigma_new)=q_lim, sum sigma^2 kept,
 *                           exact J in the tension branch, elastic no-op
 *   oracle_p2g              mass/momentum conservation, brute-force stencil
 *                           sums for 1-4 particles, partition of unity
 *   oracle_update_nodes     per-node closed forms (free, gravity, clamp,
 *                           indenter projection, each BC region), bin KATs
 *   oracle_g2p              uniform field -> Bp=0, linear field -> Bp=A
 *   oracle_run              t / indenter_x accumulation, cycle cadence,
 *                           force history sum == bins sum
 *   oracle_cell_keys_*      brute force on lattice points, no-FMA formula
 */
#ifndef ICEMPM_ORACLE_H
#define ICEMPM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* sizeof and member offsets of the SimParams restatement (for layout tests) */
int oracle_simparams_size(void);
int oracle_simparams_offsets(long *out, int max_n);

/* ---- 2x2 linear algebra (helper_math.cuh:26-140, givens.cuh:30-176) ---- */
/* F row-major {F00,F01,F10,F11}; U,V row-major; S = signed singular values */
void oracle_svd2x2(const double F[4], double U[4], double S[2], double V[4]);

/* ---- constitutive (gpu_implementation5.cu:24-141,386-394) ---- */
void oracle_kirchhoff(const double F[4], double mu, double kappa, double tau[4]);
/* out: U[4] S[2] V[4] Je p q shat[2] SS[2]  (16 doubles) */
void oracle_compute_pq(const double F[4], double mu, double kappa, double out[16]);
/* out: pmin pmax qmax beta Msq */
void oracle_grain_params(const void *prms, int grain, double out[5]);
int oracle_failure_check(const void *prms, int grain, double p, double q);
/* Fe[4] in/out, Jp_inv in/out; runs ComputePQ then the DP return map */
void oracle_dp_return(const void *prms, double Fe[4], double *Jp_inv);

/* ---- kernels (gpu_implementation5.cu:148-379), SoA layout of point.cpp ---- */
void oracle_reset_grid(const void *prms, double *grid);
int oracle_p2g(const void *prms, const double *pts, double *grid);
/* fsum[2] receives the node-ordered sum of contact forces of this call */
void oracle_update_nodes(const void *prms, double *grid, double *bins,
                         double indenter_x, double indenter_y, double fsum[2]);
int oracle_g2p(const void *prms, double *pts, const double *grid, int recordPQ);

/* ---- Model::Step (model.cpp:11-47), generalised to n substeps ----
 * prms is updated (SimulationStep, SimulationTime, indenter_x).  hist gets
 * 2 doubles per substep (total indenter force of that substep) or may be
 * NULL.  Bins are zeroed at each cycle start.  Returns the OR of error flags.
 * nthreads > 1 runs the OpenMP CPU baseline (P2G into per-thread row bands
 * reduced in thread order; deterministic per thread count). */
int oracle_run(void *prms, double *pts, double *grid, double *bins,
               int nsub, double *hist, int nthreads);

/* ---- cell keys and sort (snapshotmanager.cpp:288-304; kernel :185,325) ---- */
void oracle_cell_keys_f32(const float *xy, int n, double hinv, int gridx, int *keys);
void oracle_cell_keys_soa(const double *pts, size_t pitch, int n, double hinv,
                          int gridx, int *keys);
/* stable ascending order of keys: perm[k] = index of k-th smallest */
void oracle_stable_order(const int *keys, int n, int *perm);

/* ---- learned return map: FP64 forward/backward of the MLP of csrc/mlp.hip */
double oracle_mlp_loss_grad(const double *params, const double *X, const double *T, int B,
                            double *grad, double *Y);

#ifdef __cplusplus
}
#endif
#endif
