// mpm_device.h -- per-particle and per-node arithmetic of the MLS-MPM
// substep for gfx950, in FP64.
//
// Every expression is written in the evaluation order of the reference
// routine it replaces (cited per function, paths relative to the reference
// tree).  Translation units built with -ffp-contract=off get defined IEEE
// sequences (sqrt and division are correctly rounded on gfx950); the FAST
// particle pass includes this header after `#pragma clang fp contract(fast)`.
// All functions have internal linkage, so the two builds never collide.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace icempm {

// SoA row indices of the point buffer (parameters_sim.h:26-35)
constexpr int kRowUtil = 0, kRowP = 1, kRowQ = 2, kRowJp = 3, kRowPosX = 4, kRowVelX = 6,
              kRowFe00 = 8, kRowBp00 = 12, kPtsRows = 16, kGridRows = 3;
constexpr double kPi = 3.14159265358979323846;   // parameters_sim.h:21
constexpr double kSqrt2 = 1.4142135623730950;    // gpu_implementation5.cu:22 (coeff1)

// Everything a kernel needs, passed by value (lands in the kernarg segment
// and is read into SGPRs).
struct KParams {
    double *pts;        // 16 rows x pp doubles
    const double *grainTab;   // kGrainEntries x kGrainFields per-grain strengths
    size_t pp;          // point pitch (doubles)
    size_t gp;          // grid pitch (doubles)
    double *bins;       // 2 * nbins doubles
    int *err;           // out-of-grid flag (gpu_error_indicator)
    int nPts, gridX, gridY, gridTotal, nbins;
    double dt, vol, mass, h, h_inv, Dinv, mu, kappa, gravity, indRsq, indV, vmax;
    double ics, its, iss, grainVar, tanPhi, dpThr;
    // NACC constants of grain 0 (kGPmax.. order): every grain's when
    // GrainVariability == 0
    double g0[5];
};

// ----------------------------------------------------------------------------
// Signed 2x2 SVD with rotation-only U, V.  helper_math.cuh:26-46 (polar
// step), :80-140 (symmetric eigen step and sort); givens.cuh:99-128,161-166.
// F, U, V row-major; U = [[uc, us], [-us, uc]], likewise V.
struct Svd2 {
    double uc, us, vc, vs, s0, s1;
};

static __device__ __forceinline__ Svd2 svd2x2(double F00, double F01, double F10, double F11) {
    // polar step on the column-major copy {F00, F10, F01, F11}
    double x0 = F00 + F11;
    double x1 = F10 - F01;
    double den = sqrt(x0 * x0 + x1 * x1);
    double rc = 1.0, rs = 0.0;
    if (den != 0) {
        rc = x0 / den;
        rs = -x1 / den;
    }
    // S = G * A (rows rotated); only S00, S01, S11 are needed
    double S00 = rc * F00 - rs * F10;
    double S01 = rc * F01 - rs * F11;
    double S11 = rs * F01 + rc * F11;
    double cosine, sine, s0, s1;
    const double y2 = S01 * S01;
    if (y2 == 0) {
        cosine = 1;
        sine = 0;
        s0 = S00;
        s1 = S11;
    } else {
        double tau = 0.5 * (S00 - S11);
        double w = sqrt(tau * tau + y2);
        double t = (tau > 0) ? S01 / (tau + w) : S01 / (tau - w);
        cosine = 1.0 / sqrt(t * t + 1.0);
        sine = -t * cosine;
        double c2 = cosine * cosine;
        double csy = 2 * cosine * sine * S01;
        double s2 = sine * sine;
        s0 = c2 * S00 - csy + s2 * S11;
        s1 = s2 * S00 + csy + c2 * S11;
    }
    Svd2 r;
    if (s0 < s1) {
        r.s0 = s1;
        r.s1 = s0;
        r.vc = -sine;
        r.vs = cosine;
    } else {
        r.s0 = s0;
        r.s1 = s1;
        r.vc = cosine;
        r.vs = sine;
    }
    r.uc = rc * r.vc - rs * r.vs;
    r.us = rs * r.vc + rc * r.vs;
    return r;
}

// F = U diag(a, b) V^T for the rotation pair of an Svd2.
static __device__ __forceinline__ void usvt(const Svd2 &r, double a, double b, double &F00, double &F01,
                                     double &F10, double &F11) {
    // U = [[uc, us], [-us, uc]], V = [[vc, vs], [-vs, vc]]
    double U00 = r.uc, U01 = r.us, U10 = -r.us, U11 = r.uc;
    double V00 = r.vc, V01 = r.vs, V10 = -r.vs, V11 = r.vc;
    F00 = (U00 * a) * V00 + (U01 * b) * V01;
    F01 = (U00 * a) * V10 + (U01 * b) * V11;
    F10 = (U10 * a) * V00 + (U11 * b) * V01;
    F11 = (U10 * a) * V10 + (U11 * b) * V11;
}

// ----------------------------------------------------------------------------
// Kirchhoff stress tau = mu/J dev(F F^T) + kappa (J^2 - 1) I
// (KirchhoffStress_Wolper, gpu_implementation5.cu:24-34; dev :391-394)
static __device__ __forceinline__ void kirchhoff(double F00, double F01, double F10, double F11,
                                          double mu, double kappa, double &t00, double &t01,
                                          double &t11) {
    double Je = F00 * F11 - F10 * F01;
    double b00 = F00 * F00 + F01 * F01;
    double b01 = F00 * F10 + F01 * F11;
    double b11 = F10 * F10 + F11 * F11;
    double half_tr = (b00 + b11) / 2;
    double c1 = mu * (1 / Je);
    double c2 = kappa * (Je * Je - 1.);
    t00 = c1 * (b00 - half_tr) + c2;
    t01 = c1 * b01;
    t11 = c1 * (b11 - half_tr) + c2;
}

// ----------------------------------------------------------------------------
// Trial state of the return map (ComputePQ, gpu_implementation5.cu:36-44;
// dev_d :386-389)
struct Trial {
    Svd2 svd;
    double Je, p, q, sh0, sh1, ss0, ss1;
};

static __device__ __forceinline__ Trial compute_pq(double F00, double F01, double F10, double F11,
                                            double mu, double kappa) {
    Trial r;
    r.svd = svd2x2(F00, F01, F10, F11);
    r.Je = r.svd.s0 * r.svd.s1;
    r.p = -(kappa / 2.) * (r.Je * r.Je - 1.);
    r.ss0 = r.svd.s0 * r.svd.s0;
    r.ss1 = r.svd.s1 * r.svd.s1;
    double half_sum = (r.ss0 + r.ss1) / 2;
    double c = mu / r.Je;
    r.sh0 = c * (r.ss0 - half_sum);
    r.sh1 = c * (r.ss1 - half_sum);
    r.q = kSqrt2 * sqrt(r.sh0 * r.sh0 + r.sh1 * r.sh1);
    return r;
}

// FAST-mode forms of the same quantities (not bit-identical to the
// reference's operation order): one reciprocal for the polar rotation, and
// q = 2 |sh0| (in 2D the deviator has sh1 = -sh0, so sqrt2 |sh| = 2 |sh0|).
static __device__ __forceinline__ Svd2 svd2x2_fast(double F00, double F01, double F10, double F11) {
    double x0 = F00 + F11;
    double x1 = F10 - F01;
    double den2 = x0 * x0 + x1 * x1;
    double rc = 1.0, rs = 0.0;
    if (den2 != 0) {
        const double inv = 1.0 / sqrt(den2);
        rc = x0 * inv;
        rs = -x1 * inv;
    }
    double S00 = rc * F00 - rs * F10;
    double S01 = rc * F01 - rs * F11;
    double S11 = rs * F01 + rc * F11;
    double cosine, sine, s0, s1;
    const double y2 = S01 * S01;
    if (y2 == 0) {
        cosine = 1;
        sine = 0;
        s0 = S00;
        s1 = S11;
    } else {
        double tau = 0.5 * (S00 - S11);
        double w = sqrt(tau * tau + y2);
        double t = S01 / ((tau > 0) ? (tau + w) : (tau - w));
        cosine = 1.0 / sqrt(t * t + 1.0);
        sine = -t * cosine;
        double c2 = cosine * cosine;
        double csy = 2 * cosine * sine * S01;
        double s2 = sine * sine;
        s0 = c2 * S00 - csy + s2 * S11;
        s1 = s2 * S00 + csy + c2 * S11;
    }
    Svd2 r;
    if (s0 < s1) {
        r.s0 = s1;
        r.s1 = s0;
        r.vc = -sine;
        r.vs = cosine;
    } else {
        r.s0 = s0;
        r.s1 = s1;
        r.vc = cosine;
        r.vs = sine;
    }
    r.uc = rc * r.vc - rs * r.vs;
    r.us = rs * r.vc + rc * r.vs;
    return r;
}

static __device__ __forceinline__ Trial compute_pq_fast(double F00, double F01, double F10, double F11,
                                                        double mu, double kappa) {
    Trial r;
    r.svd = svd2x2_fast(F00, F01, F10, F11);
    r.Je = r.svd.s0 * r.svd.s1;
    r.p = -(kappa / 2.) * (r.Je * r.Je - 1.);
    r.ss0 = r.svd.s0 * r.svd.s0;
    r.ss1 = r.svd.s1 * r.svd.s1;
    const double c = mu / r.Je;
    r.sh0 = c * (0.5 * (r.ss0 - r.ss1));
    r.sh1 = -r.sh0;
    r.q = 2.0 * fabs(r.sh0);
    return r;
}

// Kirchhoff stress of the state a Trial describes (its F unchanged since):
// tau = U diag(sh_i + kappa (J^2 - 1)) U^T = U diag(sh_i - 2 p) U^T.
static __device__ __forceinline__ void kirchhoff_from_trial(const Trial &r, double &t00, double &t01,
                                                            double &t11) {
    const double d0 = r.sh0 - 2. * r.p, d1 = r.sh1 - 2. * r.p;
    const double cc = r.svd.uc * r.svd.uc, ss = r.svd.us * r.svd.us, cs = r.svd.uc * r.svd.us;
    t00 = cc * d0 + ss * d1;
    t01 = cs * (d1 - d0);
    t11 = ss * d0 + cc * d1;
}

// FAST-mode trial invariants without the SVD.  With J = det F and
// b = F F^T, the squared singular values are b's eigenvalues, so
// ss0 - ss1 = sqrt((b00 - b11)^2 + 4 b01^2) and
//   p = -kappa/2 (J^2 - 1),  q = |mu / J| (ss0 - ss1)
// (the same p and q as ComputePQ, gpu_implementation5.cu:36-44).  The NACC
// test and the Kirchhoff stress need nothing else; only a point the
// return map projects needs the SVD.
struct Inv {
    double c, d, b01, p, q;   // c = mu / J, d = b00 - b11
};

static __device__ __forceinline__ Inv invariants_fast(double F00, double F01, double F10, double F11,
                                                      double mu, double kappa) {
    Inv v;
    const double J = F00 * F11 - F10 * F01;
    const double b00 = F00 * F00 + F01 * F01, b11 = F10 * F10 + F11 * F11;
    v.b01 = F00 * F10 + F01 * F11;
    v.d = b00 - b11;
    v.c = mu / J;
    v.p = -(kappa / 2.) * (J * J - 1.);
    v.q = fabs(v.c) * sqrt(v.d * v.d + 4. * (v.b01 * v.b01));
    return v;
}

// tau = mu/J dev(b) + kappa (J^2 - 1) I from the invariants of the same F
static __device__ __forceinline__ void kirchhoff_from_inv(const Inv &v, double &t00, double &t01,
                                                          double &t11) {
    const double a = 0.5 * (v.c * v.d);
    t00 = a - 2. * v.p;
    t01 = v.c * v.b01;
    t11 = -a - 2. * v.p;
}

// 1/x for a normal x: the hardware reciprocal (v_rcp_f64) refined by two
// Newton-Raphson steps, r <- r + r (1 - x r), to within an ulp or two of the
// correctly rounded quotient.  FAST mode only (not the IEEE division).
static __device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// The FAST trial invariants with q^2 in place of q: every test on q in the
// fused passes compares it with a non-negative limit (the NACC ellipse uses
// q^2 itself), so no square root is taken, and c = mu / J comes from rcp_nr.
//   q^2 = (mu / J)^2 ((b00 - b11)^2 + 4 b01^2)
struct InvSq {
    double c, d, b01, p, q2;   // c = mu / J, d = b00 - b11
};

static __device__ __forceinline__ InvSq invariants_sq(double F00, double F01, double F10, double F11,
                                                      double mu, double kappa) {
    InvSq v;
    const double J = F00 * F11 - F10 * F01;
    const double b00 = F00 * F00 + F01 * F01, b11 = F10 * F10 + F11 * F11;
    v.b01 = F00 * F10 + F01 * F11;
    v.d = b00 - b11;
    v.c = mu * rcp_nr(J);
    v.p = -(kappa / 2.) * (J * J - 1.);
    v.q2 = (v.c * v.c) * (v.d * v.d + 4. * (v.b01 * v.b01));
    return v;
}

static __device__ __forceinline__ void kirchhoff_from_inv(const InvSq &v, double &t00, double &t01,
                                                          double &t11) {
    const double a = 0.5 * (v.c * v.d);
    t00 = a - 2. * v.p;
    t01 = v.c * v.b01;
    t11 = -a - 2. * v.p;
}

// KirchhoffStress_Wolper (gpu_implementation5.cu:24-34) with 1/J from rcp_nr
static __device__ __forceinline__ void kirchhoff_rcp(double F00, double F01, double F10, double F11, double mu,
                                                     double kappa, double &t00, double &t01, double &t11) {
    const double Je = F00 * F11 - F10 * F01;
    const double b00 = F00 * F00 + F01 * F01;
    const double b01 = F00 * F10 + F01 * F11;
    const double b11 = F10 * F10 + F11 * F11;
    const double half_tr = (b00 + b11) / 2;
    const double c1 = mu * rcp_nr(Je);
    const double c2 = kappa * (Je * Je - 1.);
    t00 = c1 * (b00 - half_tr) + c2;
    t01 = c1 * b01;
    t11 = c1 * (b11 - half_tr) + c2;
}

// Whether the Drucker-Prager map (dp_return) changes a crushed trial state
// with these p and q: the tension branch, or q at or above the cone.
static __device__ __forceinline__ bool dp_projects(const KParams &k, double p, double q, double Jp_inv) {
    if (p < -k.dpThr || Jp_inv < 1) return true;
    const double q_n_1 = (p > k.ics) ? 0.0 : fmin((p + k.dpThr) * k.tanPhi, k.iss);
    return q >= q_n_1;
}

// the same from q^2: past the tension test p + p_thr >= 0, so the cone's
// q_n_1 is >= 0 (for tan(phi), qmax >= 0) and q >= q_n_1 iff q^2 >= q_n_1^2
static __device__ __forceinline__ bool dp_projects_sq(const KParams &k, double p, double q2, double Jp_inv) {
    if (p < -k.dpThr || Jp_inv < 1) return true;
    const double q_n_1 = (p > k.ics) ? 0.0 : fmin((p + k.dpThr) * k.tanPhi, k.iss);
    return q_n_1 <= 0.0 || q2 >= q_n_1 * q_n_1;
}

// NACC test with the constants of grain 0 (GrainVariability == 0)
static __device__ __forceinline__ bool fails_g0(const KParams &k, double p, double q) {
    if (p < 0) return (p < -3e6) || (q > k.g0[4] * (p - -3e6));
    return k.g0[3] * q * q + k.g0[1] * (p + k.g0[2]) * (p - k.g0[0]) > 0;
}

// Per-grain constants of the NACC test, one row per short grain id
// (index grain + 32768), computed on the host in the order of
// GetParametersForGrain (gpu_implementation5.cu:97-111) and
// CheckIfPointIsInsideFailureSurface (:114-141): pmax, mSq, beta*pmax,
// (1 + 2 beta), and the slope of the tension line -q0/pmin2.
constexpr int kGrainEntries = 65536;
constexpr int kGrainFields = 5;
enum { kGPmax = 0, kGMsq = 1, kGBetaPmax = 2, kGOnePlus2Beta = 3, kGSlope = 4 };

// the same test as fails(), reading the per-grain constants from the table
static __device__ __forceinline__ bool fails_tab(const KParams &k, short grain, double p, double q) {
    const double *g = k.grainTab + ((int)grain + 32768) * kGrainFields;
    const double pmin2 = -3e6;
    if (p < 0) {
        if (p < pmin2) return true;
        const double q_limit = g[kGSlope] * (p - pmin2);
        return q > q_limit;
    }
    const double pmax = g[kGPmax];
    const double y = g[kGOnePlus2Beta] * q * q + g[kGMsq] * (p + g[kGBetaPmax]) * (p - pmax);
    return y > 0;
}

// NACC failure test for intact points with per-grain strengths
// (CheckIfPointIsInsideFailureSurface :114-141, GetParametersForGrain :97-111)
static __device__ __forceinline__ bool fails(const KParams &k, short grain, double p, double q) {
    double var1 = 1.0 + k.grainVar * 0.05 * (-10 + grain % 21);
    double var2 = 1.0 + k.grainVar * 0.033 * (-15 + (grain + 3) % 30);
    double var3 = 1.0 + k.grainVar * 0.1 * (-10 + (grain + 4) % 11);
    double pmax = k.ics * var1;
    double pmin = -k.its * var2;
    double qmax = k.iss * var3;
    double beta = -pmin / pmax;
    double a = pmax * (1 + beta);
    double mSq = (4 * qmax * qmax * (1 + 2 * beta)) / (a * a);
    const double pmin2 = -3e6;
    if (p < 0) {
        if (p < pmin2) return true;
        double q0 = 2 * sqrt(-pmax * pmin) * qmax / (pmax - pmin);
        double kk = -q0 / pmin2;
        double q_limit = kk * (p - pmin2);
        return q > q_limit;
    }
    double y = (1. + 2. * beta) * q * q + mSq * (p + beta * pmax) * (p - pmax);
    return y > 0;
}

// fails() from q^2: the tension line's q_limit is >= 0 for qmax >= 0, and
// q > q_limit iff q^2 > q_limit^2 there; a negative q_limit (qmax < 0, a
// GrainVariability above 1) is exceeded by every q >= 0
static __device__ __forceinline__ bool fails_sq(const KParams &k, short grain, double p, double q2) {
    double var1 = 1.0 + k.grainVar * 0.05 * (-10 + grain % 21);
    double var2 = 1.0 + k.grainVar * 0.033 * (-15 + (grain + 3) % 30);
    double var3 = 1.0 + k.grainVar * 0.1 * (-10 + (grain + 4) % 11);
    double pmax = k.ics * var1;
    double pmin = -k.its * var2;
    double qmax = k.iss * var3;
    double beta = -pmin / pmax;
    double a = pmax * (1 + beta);
    double mSq = (4 * qmax * qmax * (1 + 2 * beta)) / (a * a);
    const double pmin2 = -3e6;
    if (p < 0) {
        if (p < pmin2) return true;
        double q0 = 2 * sqrt(-pmax * pmin) * qmax / (pmax - pmin);
        double kk = -q0 / pmin2;
        double q_limit = kk * (p - pmin2);
        return q_limit < 0 || q2 > q_limit * q_limit;
    }
    double y = (1. + 2. * beta) * q2 + mSq * (p + beta * pmax) * (p - pmax);
    return y > 0;
}

// Drucker-Prager return map for crushed points (Wolper_Drucker_Prager
// :48-94).  Updates Fe and Jp_inv in place.
static __device__ __forceinline__ void dp_return(const KParams &k, const Trial &r, double &F00,
                                          double &F01, double &F10, double &F11,
                                          double &Jp_inv) {
    if (r.p < -k.dpThr || Jp_inv < 1) {
        double p_new = -k.dpThr;
        double Je_new = sqrt(-2. * p_new / k.kappa + 1.);
        double sq = sqrt(Je_new);
        usvt(r.svd, sq, sq, F00, F01, F10, F11);
        Jp_inv *= Je_new / r.Je;
    } else {
        double q_n_1;
        if (r.p > k.ics) {
            q_n_1 = 0;
        } else {
            double q_from_dp = (r.p + k.dpThr) * k.tanPhi;
            q_n_1 = fmin(q_from_dp, k.iss);
        }
        if (r.q >= q_n_1) {
            double s_norm = q_n_1 / kSqrt2;
            double z = r.sh0 * r.sh0 + r.sh1 * r.sh1;
            double n0 = r.sh0, n1 = r.sh1;
            if (z > 0) {   // normalising a zero vector leaves it zero
                double nz = sqrt(z);
                n0 = n0 / nz;
                n1 = n1 / nz;
            }
            double c = s_norm * (r.Je / k.mu);
            double m = (r.ss0 + r.ss1) / 2;
            usvt(r.svd, sqrt(c * n0 + m), sqrt(c * n1 + m), F00, F01, F10, F11);
        }
    }
}

// ----------------------------------------------------------------------------
// Quadratic B-spline stencil along one axis (gpu_implementation5.cu:185-193)
// Returns false for a position whose base index cannot be represented.
static __device__ __forceinline__ bool stencil(double x, double h_inv, int &base, double &fx,
                                        double &w0, double &w1, double &w2) {
    double bx = x * h_inv - 0.5;
    if (!(bx > -1.0 && bx < 2147483000.0)) return false;
    base = (int)bx;
    fx = x * h_inv - (double)base;
    double v0 = 1.5 - fx;
    double v1 = fx - 1.0;
    double v2 = fx - 0.5;
    w0 = 0.5 * v0 * v0;
    w1 = 0.75 - v1 * v1;
    w2 = 0.5 * v2 * v2;
    return true;
}

static __device__ __forceinline__ double sel3(int i, double a, double b, double c) {
    return i == 0 ? a : (i == 1 ? b : c);
}

// Cell key of a position (SnapshotManager::SortPoints key,
// snapshotmanager.cpp:300-301, on the kernel's base index)
static __device__ __forceinline__ int cell_key(double x, double y, double hinv, int gridX) {
    return (int)(x * hinv - 0.5) + ((int)(y * hinv - 0.5)) * gridX;
}

static __device__ __forceinline__ void flag_error(const KParams &k) { k.err[0] = 1; }

// Wavefront reductions through DPP: permutations within each row of 16
// lanes (pairs, quads, half-row mirror, row mirror), then the four row
// results through readlane, so the result is wavefront-uniform.  Every lane
// of the wavefront must be active.
template <bool MAX>
static __device__ __forceinline__ int wreduce_i32(int v) {
    const auto op = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));    // quad_perm 1,0,3,2
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));    // quad_perm 2,3,0,1
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false));   // row_half_mirror
    v = op(v, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false));   // row_mirror
    return op(op(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
              op(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// Wavefront minimum of an unsigned value, in lane 63's register read back
// as a scalar: four DPP steps within each row of 16 lanes (quad_perm 1032,
// 2301, half-row mirror, row mirror), then row_bcast:15 into rows 1 and 3
// and row_bcast:31 into rows 2 and 3 (GFX9 DPP).  With the identity
// (0xFFFFFFFF) as the DPP "old" value every step folds into one
// v_min_u32_dpp.  Every lane of the wavefront must be active.
static __device__ __forceinline__ unsigned wave_min_u32(unsigned v) {
    const int I = (int)0xFFFFFFFFu;
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(I, (int)v, 0xB1, 0xF, 0xF, false));
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(I, (int)v, 0x4E, 0xF, 0xF, false));
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(I, (int)v, 0x141, 0xF, 0xF, false));
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(I, (int)v, 0x140, 0xF, 0xF, false));
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(I, (int)v, 0x142, 0xA, 0xF, false));
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(I, (int)v, 0x143, 0xC, 0xF, false));
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}
// maximum, as the complement of the minimum of complements
static __device__ __forceinline__ unsigned wave_max_u32(unsigned v) { return ~wave_min_u32(~v); }

// The same for the two 16-bit halves of a word at once.
typedef unsigned short U16x2 __attribute__((ext_vector_type(2)));
template <bool MAX>
static __device__ __forceinline__ unsigned pk_sel(unsigned a, unsigned b) {
    const U16x2 x = __builtin_bit_cast(U16x2, a), y = __builtin_bit_cast(U16x2, b);
    return __builtin_bit_cast(unsigned, MAX ? __builtin_elementwise_max(x, y) : __builtin_elementwise_min(x, y));
}
template <bool MAX>
static __device__ __forceinline__ unsigned wreduce_u16x2(unsigned v) {
    v = pk_sel<MAX>(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));
    v = pk_sel<MAX>(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));
    v = pk_sel<MAX>(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));
    v = pk_sel<MAX>(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));
    const unsigned r0 = (unsigned)__builtin_amdgcn_readlane((int)v, 0), r1 = (unsigned)__builtin_amdgcn_readlane((int)v, 16);
    const unsigned r2 = (unsigned)__builtin_amdgcn_readlane((int)v, 32), r3 = (unsigned)__builtin_amdgcn_readlane((int)v, 48);
    return pk_sel<MAX>(pk_sel<MAX>(r0, r1), pk_sel<MAX>(r2, r3));
}

}  // namespace icempm
