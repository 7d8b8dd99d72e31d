// icempm_oracle.cpp -- CPU restatement of iceMPM's MLS-MPM substep.
//
// TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline).  The
// product in icempm_amd/ never links or calls this file.  Build with
// -ffp-contract=off: every expression below is evaluated exactly in the
// order written, and the HIP kernels follow the same order.
//
// References are to the reference tree (simulation/...).

#include "icempm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// SimParams restatement: member order and types of parameters_sim.h:38-80.
// The struct is uploaded verbatim and dumped to snapshots by the reference,
// so the byte layout is part of the interface.
struct SimParams {
    double *grid_array;
    double *indenter_force_accumulator;
    double *pts_array;
    size_t nPtsPitch, nGridPitch;
    int n_indenter_subdivisions;
    int tpb_P2G, tpb_Upd, tpb_G2P;
    int nPts;
    int GridX, GridY, GridTotal;
    double GridXDimension;
    double InitialTimeStep, SimulationEndTime;
    double Gravity, Density, PoissonsRatio, YoungsModulus;
    double lambda, mu, kappa;
    double IceCompressiveStrength, IceTensileStrength, IceShearStrength;
    double NACC_beta, NACC_M, NACC_Msq;
    double DP_tan_phi, DP_threshold_p;
    double cellsize, cellsize_inv, Dp_inv;
    int UpdateEveryNthStep;
    double IndDiameter, IndRSq, IndVelocity, IndDepth;
    double xmin, xmax, ymin, ymax;
    int nxmin, nxmax, nymin, nymax;
    double ParticleVolume, ParticleMass, ParticleViewSize;
    int SimulationStep;
    double SimulationTime;
    double indenter_x, indenter_x_initial, indenter_y, indenter_y_initial;
    double Volume;
    int SetupType;
    double GrainVariability;
    int UseWarpLevelReduction;
    int SortPointsByGridCell;
};

// SoA row indices, parameters_sim.h:26-35
constexpr size_t kUtil = 0, kP = 1, kQ = 2, kJp = 3, kPosX = 4, kVelX = 6, kFe00 = 8,
                 kBp00 = 12;
constexpr double kPi = 3.14159265358979323846;     // parameters_sim.h:21
constexpr double kCoeff1 = 1.4142135623730950;     // gpu_implementation5.cu:22
constexpr double kD = 2;                           // gpu_implementation5.cu:21

// ---------------------------------------------------------------------------
// Signed 2x2 SVD: helper_math.cuh:26-46 (polar step), :80-140 (symmetric
// eigen step + sort), givens.cuh:99-128,161-166 (fill, mat_rotation, *=).
struct Rot { double c, s; };

void polar(const double a[4] /*column-major*/, Rot &r, double s[4]) {
    double x0 = a[0] + a[3];
    double x1 = a[1] - a[2];
    double den = std::sqrt(x0 * x0 + x1 * x1);
    r.c = 1.0;
    r.s = 0.0;
    if (den != 0) {
        r.c = x0 / den;
        r.s = -x1 / den;
    }
    for (int i = 0; i < 4; ++i) s[i] = a[i];
    // row rotation of a column-major 2x2 (rowi=0, rowk=1)
    for (int d = 0; d < 2; ++d) {
        double t1 = s[0 + d * 2];
        double t2 = s[1 + d * 2];
        s[0 + d * 2] = r.c * t1 - r.s * t2;
        s[1 + d * 2] = r.s * t1 + r.c * t2;
    }
}

// F row-major; U, V row-major rotations [[c, s], [-s, c]] as produced by
// GivensRotation::fill followed by the row-major comma initialiser
// (gpu_implementation5.cu:399-416).
void svd2x2(const double F[4], double U[4], double S[2], double V[4]) {
    // singular_value_decomposition receives aa = {F00,F01,F10,F11} and
    // builds the column-major copy {aa0, aa2, aa1, aa3}.
    const double a[4] = {F[0], F[2], F[1], F[3]};
    Rot u, v;
    double ss[4];
    polar(a, u, ss);
    double cosine, sine;
    const double x = ss[0], y = ss[2], z = ss[3];
    const double y2 = y * y;
    double s0, s1;
    if (y2 == 0) {
        cosine = 1;
        sine = 0;
        s0 = x;
        s1 = z;
    } else {
        double tau = 0.5 * (x - z);
        double w = std::sqrt(tau * tau + y2);
        double t;
        if (tau > 0)
            t = y / (tau + w);
        else
            t = y / (tau - w);
        cosine = 1.0 / std::sqrt(t * t + 1.0);
        sine = -t * cosine;
        double c2 = cosine * cosine;
        double csy = 2 * cosine * sine * y;
        double s2 = sine * sine;
        s0 = c2 * x - csy + s2 * z;
        s1 = s2 * x + csy + c2 * z;
    }
    if (s0 < s1) {
        std::swap(s0, s1);
        v.c = -sine;
        v.s = cosine;
    } else {
        v.c = cosine;
        v.s = sine;
    }
    // u *= v
    double nc = u.c * v.c - u.s * v.s;
    double ns = u.s * v.c + u.c * v.s;
    u.c = nc;
    u.s = ns;
    U[0] = u.c; U[1] = u.s; U[2] = -u.s; U[3] = u.c;
    V[0] = v.c; V[1] = v.s; V[2] = -v.s; V[3] = v.c;
    S[0] = s0;
    S[1] = s1;
}

// ---------------------------------------------------------------------------
// KirchhoffStress_Wolper, gpu_implementation5.cu:24-34 with dev() :391-394.
// tau = mu/J dev(F F^T) + kappa (J^2 - 1) I
void kirchhoff(const double F[4], double mu, double kappa, double tau[4]) {
    double Je = F[0] * F[3] - F[2] * F[1];
    double b00 = F[0] * F[0] + F[1] * F[1];
    double b01 = F[0] * F[2] + F[1] * F[3];
    double b11 = F[2] * F[2] + F[3] * F[3];
    double half_tr = (b00 + b11) / 2;
    double c1 = mu * (1 / Je);
    double c2 = kappa * (Je * Je - 1.);
    tau[0] = c1 * (b00 - half_tr) + c2;
    tau[1] = c1 * b01;
    tau[2] = c1 * b01;
    tau[3] = c1 * (b11 - half_tr) + c2;
}

struct PQ {
    double U[4], S[2], V[4];
    double Je, p, q;
    double shat[2], SS[2];
};

// ComputePQ, gpu_implementation5.cu:36-44, dev_d :386-389
void compute_pq(const double F[4], double mu, double kappa, PQ &r) {
    svd2x2(F, r.U, r.S, r.V);
    r.Je = r.S[0] * r.S[1];
    r.p = -(kappa / 2.) * (r.Je * r.Je - 1.);
    r.SS[0] = r.S[0] * r.S[0];
    r.SS[1] = r.S[1] * r.S[1];
    double half_sum = (r.SS[0] + r.SS[1]) / 2;
    double c = mu / r.Je;
    r.shat[0] = c * (r.SS[0] - half_sum);
    r.shat[1] = c * (r.SS[1] - half_sum);
    r.q = kCoeff1 * std::sqrt(r.shat[0] * r.shat[0] + r.shat[1] * r.shat[1]);
}

// Fe = U diag(sig) V^T
void usvt(const double U[4], const double sig[2], const double V[4], double Fe[4]) {
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            Fe[i * 2 + j] = (U[i * 2 + 0] * sig[0]) * V[j * 2 + 0] +
                            (U[i * 2 + 1] * sig[1]) * V[j * 2 + 1];
}

// GetParametersForGrain, gpu_implementation5.cu:97-111
void grain_params(const SimParams &P, int grain, double &pmin, double &pmax, double &qmax,
                  double &beta, double &mSq) {
    short g = (short)grain;
    double var1 = 1.0 + P.GrainVariability * 0.05 * (-10 + g % 21);
    double var2 = 1.0 + P.GrainVariability * 0.033 * (-15 + (g + 3) % 30);
    double var3 = 1.0 + P.GrainVariability * 0.1 * (-10 + (g + 4) % 11);
    pmax = P.IceCompressiveStrength * var1;
    pmin = -P.IceTensileStrength * var2;
    qmax = P.IceShearStrength * var3;
    beta = -pmin / pmax;
    double a = pmax * (1 + beta);
    mSq = (4 * qmax * qmax * (1 + 2 * beta)) / (a * a);
}

// CheckIfPointIsInsideFailureSurface, gpu_implementation5.cu:114-141
int failure_check(const SimParams &P, int grain, double p, double q) {
    double pmin, pmax, qmax, beta, M_sq;
    grain_params(P, grain, pmin, pmax, qmax, beta, M_sq);
    const double pmin2 = -3e6;
    if (p < 0) {
        if (p < pmin2) return 1;
        double q0 = 2 * std::sqrt(-pmax * pmin) * qmax / (pmax - pmin);
        double k = -q0 / pmin2;
        double q_limit = k * (p - pmin2);
        if (q > q_limit) return 1;
    } else {
        double y = (1. + 2. * beta) * q * q + M_sq * (p + beta * pmax) * (p - pmax);
        if (y > 0) return 1;
    }
    return 0;
}

// Wolper_Drucker_Prager, gpu_implementation5.cu:48-94
void dp_return(const SimParams &P, const PQ &r, double Fe[4], double &Jp_inv) {
    const double mu = P.mu, kappa = P.kappa;
    const double tan_phi = P.DP_tan_phi, thr = P.DP_threshold_p;
    const double pmax = P.IceCompressiveStrength, qmax = P.IceShearStrength;
    if (r.p < -thr || Jp_inv < 1) {
        double p_new = -thr;
        double Je_new = std::sqrt(-2. * p_new / kappa + 1.);
        double sq = std::sqrt(Je_new);
        double sig[2] = {sq, sq};
        usvt(r.U, sig, r.V, Fe);
        Jp_inv *= Je_new / r.Je;
    } else {
        double q_n_1;
        if (r.p > pmax) {
            q_n_1 = 0;
        } else {
            double q_from_dp = (r.p + thr) * tan_phi;
            q_n_1 = std::min(q_from_dp, qmax);
        }
        if (r.q >= q_n_1) {
            double s_norm = q_n_1 / kCoeff1;
            double z = r.shat[0] * r.shat[0] + r.shat[1] * r.shat[1];
            double n0 = r.shat[0], n1 = r.shat[1];
            if (z > 0) {   // Eigen normalized(): zero vector stays zero
                double nz = std::sqrt(z);
                n0 = n0 / nz;
                n1 = n1 / nz;
            }
            double c = s_norm * (r.Je / mu);
            double m = (r.SS[0] + r.SS[1]) / kD;
            double sig[2] = {std::sqrt(c * n0 + m), std::sqrt(c * n1 + m)};
            usvt(r.U, sig, r.V, Fe);
        }
    }
}

// Quadratic B-spline weights, gpu_implementation5.cu:185-193 (:325-333)
// Returns false when the base index would leave [-1, grid) (the reference
// casts without a check; the out-of-grid case is flagged, :206).
inline bool weights(double x, double h_inv, int &base, double &fx, double w[3]) {
    double bx = x * h_inv - 0.5;
    if (!(bx > -1.0 && bx < 2147483000.0)) return false;
    base = (int)bx;                     // truncation toward zero
    fx = x * h_inv - (double)base;
    double v0 = 1.5 - fx;
    double v1 = fx - 1.0;
    double v2 = fx - 0.5;
    w[0] = 0.5 * v0 * v0;
    w[1] = 0.75 - v1 * v1;
    w[2] = 0.5 * v2 * v2;
    return true;
}

inline double *row(double *pts, size_t pitch, size_t r) { return pts + pitch * r; }
inline const double *row(const double *pts, size_t pitch, size_t r) { return pts + pitch * r; }

// One particle's P2G: gpu_implementation5.cu:148-213.  Contributions are
// handed to `sink(node, m, px, py)` in the reference's stencil order
// (i outer, j inner).  Out-of-grid nodes set the error flag and are skipped.
template <class Sink>
inline int p2g_particle(const SimParams &P, const double *pts, int pt, Sink &&sink) {
    const size_t pitch = P.nPtsPitch;
    const double dt = P.InitialTimeStep, vol = P.ParticleVolume, h = P.cellsize,
                 h_inv = P.cellsize_inv, Dinv = P.Dp_inv, m = P.ParticleMass;
    const int gx = P.GridX, gy = P.GridY;
    double pos[2], vel[2], Fe[4], Bp[4];
    for (int i = 0; i < 2; ++i) {
        pos[i] = row(pts, pitch, kPosX + i)[pt];
        vel[i] = row(pts, pitch, kVelX + i)[pt];
        for (int j = 0; j < 2; ++j) {
            Fe[i * 2 + j] = row(pts, pitch, kFe00 + i * 2 + j)[pt];
            Bp[i * 2 + j] = row(pts, pitch, kBp00 + i * 2 + j)[pt];
        }
    }
    double tau[4];
    kirchhoff(Fe, P.mu, P.kappa, tau);
    double s = dt * vol * Dinv;
    double A[4];
    for (int k = 0; k < 4; ++k) A[k] = m * Bp[k] - s * tau[k];

    int bx, by;
    double fx0, fx1, wx[3], wy[3];
    if (!weights(pos[0], h_inv, bx, fx0, wx) || !weights(pos[1], h_inv, by, fx1, wy)) return 1;
    double mv0 = vel[0] * m, mv1 = vel[1] * m;
    int err = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double Wip = wx[i] * wy[j];
            double d0 = (i - fx0) * h;
            double d1 = (j - fx1) * h;
            double incV0 = Wip * (mv0 + (A[0] * d0 + A[1] * d1));
            double incV1 = Wip * (mv1 + (A[2] * d0 + A[3] * d1));
            double incM = Wip * m;
            int i2 = bx + i, j2 = by + j;
            if (i2 < 0 || j2 < 0 || i2 >= gx || j2 >= gy) {
                err = 1;
                continue;
            }
            sink(i2 + j2 * gx, incM, incV0, incV1);
        }
    return err;
}

// One grid node: v2_kernel_update_nodes, gpu_implementation5.cu:215-290.
// Returns the bin index (>=0) and the force when the node is in contact.
inline int update_node(const SimParams &P, double *grid, int idx, double ind_x, double ind_y,
                       double &f0, double &f1) {
    const size_t gp = P.nGridPitch;
    double mass = grid[idx];
    if (mass == 0) return -1;
    const double dt = P.InitialTimeStep, h = P.cellsize, V = P.IndVelocity;
    const int gx = P.GridX, gy = P.GridY;
    double v0 = grid[gp + idx], v1 = grid[2 * gp + idx];
    v0 = v0 / mass;
    v1 = v1 / mass;
    v1 = v1 - dt * P.Gravity;
    double vmax = 0.5 * h / dt;
    double vn2 = std::sqrt(v0 * v0 + v1 * v1);
    if (vn2 > vmax) {
        double nrm = std::sqrt(v0 * v0 + v1 * v1);
        v0 = v0 / nrm * vmax;
        v1 = v1 / nrm * vmax;
    }
    int ix = idx % gx;
    int iy = idx / gx;
    double n0 = ix * h - ind_x;
    double n1 = iy * h - ind_y;
    int bin = -1;
    if (n0 * n0 + n1 * n1 < P.IndRSq) {
        double r0 = v0 - V;
        double r1 = v1 - 0.0;
        double z = n0 * n0 + n1 * n1;
        if (z > 0) {
            double nz = std::sqrt(z);
            n0 = n0 / nz;
            n1 = n1 / nz;
        }
        double vn = r0 * n0 + r1 * n1;
        if (vn < 0) {
            double t0 = r0 - n0 * vn;
            double t1 = r1 - n1 * vn;
            double p0 = v0, p1 = v1;
            v0 = V + t0;
            v1 = 0.0 + t1;
            f0 = (p0 - v0) * mass / dt;
            f1 = (p1 - v1) * mass / dt;
            double angle = std::atan2(n0, n1);
            angle += kPi;
            angle *= P.n_indenter_subdivisions / (2 * kPi);
            int index = (int)angle;
            index = std::max(index, 0);
            index = std::min(index, P.n_indenter_subdivisions - 1);
            bin = index;
        }
    }
    if (iy <= 2) {
        v0 = 0;
        v1 = 0;
    } else if (iy >= gy - 3 && v1 > 0)
        v1 = 0;
    if (ix <= 2 && v0 < 0)
        v0 = 0;
    else if (ix >= gx - 3 && v0 > 0)
        v0 = 0;
    grid[gp + idx] = v0;
    grid[2 * gp + idx] = v1;
    return bin;
}

// One particle's G2P: v2_kernel_g2p, gpu_implementation5.cu:294-379.
inline int g2p_particle(const SimParams &P, double *pts, const double *grid, int pt,
                        int recordPQ) {
    const size_t pitch = P.nPtsPitch, gp = P.nGridPitch;
    const double h_inv = P.cellsize_inv, dt = P.InitialTimeStep;
    const int gx = P.GridX, gy = P.GridY;
    double pos[2], Fe[4];
    for (int i = 0; i < 2; ++i) {
        pos[i] = row(pts, pitch, kPosX + i)[pt];
        for (int j = 0; j < 2; ++j) Fe[i * 2 + j] = row(pts, pitch, kFe00 + i * 2 + j)[pt];
    }
    unsigned char *util = reinterpret_cast<unsigned char *>(row(pts, pitch, kUtil));
    unsigned char crushed = util[pt];
    short grain;
    std::memcpy(&grain, util + pitch + 2 * (size_t)pt, sizeof(short));
    double Jp_inv = row(pts, pitch, kJp)[pt];

    int bx, by;
    double fx0, fx1, wx[3], wy[3];
    int err = 0;
    if (!weights(pos[0], h_inv, bx, fx0, wx) || !weights(pos[1], h_inv, by, fx1, wy)) {
        // unusable position: flag and leave the particle unchanged
        return 1;
    }
    double v[2] = {0, 0}, Bp[4] = {0, 0, 0, 0};
    const double c4 = 4. * h_inv;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double d0 = i - fx0;
            double d1 = j - fx1;
            double w = wx[i] * wy[j];
            int i2 = bx + i, j2 = by + j;
            double nv0 = 0, nv1 = 0;
            if (i2 < 0 || j2 < 0 || i2 >= gx || j2 >= gy) {
                err = 1;
            } else {
                int node = i2 + j2 * gx;
                nv0 = grid[gp + node];
                nv1 = grid[2 * gp + node];
            }
            v[0] += w * nv0;
            v[1] += w * nv1;
            double cw = c4 * w;
            Bp[0] += cw * (nv0 * d0);
            Bp[1] += cw * (nv0 * d1);
            Bp[2] += cw * (nv1 * d0);
            Bp[3] += cw * (nv1 * d1);
        }
    pos[0] += v[0] * dt;
    pos[1] += v[1] * dt;
    double M[4] = {1.0 + dt * Bp[0], dt * Bp[1], dt * Bp[2], 1.0 + dt * Bp[3]};
    double Fn[4];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) Fn[i * 2 + j] = M[i * 2 + 0] * Fe[0 + j] + M[i * 2 + 1] * Fe[2 + j];
    std::memcpy(Fe, Fn, sizeof(Fe));

    PQ r;
    compute_pq(Fe, P.mu, P.kappa, r);
    if (crushed == 0) crushed = (unsigned char)failure_check(P, grain, r.p, r.q);
    if (crushed == 1) dp_return(P, r, Fe, Jp_inv);

    row(pts, pitch, kJp)[pt] = Jp_inv;
    for (int i = 0; i < 2; ++i) {
        row(pts, pitch, kPosX + i)[pt] = pos[i];
        row(pts, pitch, kVelX + i)[pt] = v[i];
        for (int j = 0; j < 2; ++j) {
            row(pts, pitch, kFe00 + i * 2 + j)[pt] = Fe[i * 2 + j];
            row(pts, pitch, kBp00 + i * 2 + j)[pt] = Bp[i * 2 + j];
        }
    }
    util[pt] = crushed;
    if (recordPQ) {
        row(pts, pitch, kP)[pt] = r.p;
        row(pts, pitch, kQ)[pt] = r.q;
    }
    return err;
}

const SimParams &cp(const void *p) { return *static_cast<const SimParams *>(p); }

// ---------------------------------------------------------------------------
// OpenMP CPU baseline for P2G: thread t takes a contiguous particle range
// and accumulates into a private band of grid rows; bands are then summed
// into the grid in thread order.
int p2g_parallel(const SimParams &P, const double *pts, double *grid, int nthreads) {
#ifdef _OPENMP
    const int n = P.nPts, gx = P.GridX, gy = P.GridY;
    const size_t gp = P.nGridPitch;
    struct Band { int y0 = 0, y1 = -1; std::vector<double> buf; int err = 0; };
    static std::vector<Band> bands;   // one oracle_run at a time
    bands.resize(nthreads);
#pragma omp parallel num_threads(nthreads)
    {
        int t = omp_get_thread_num();
        int a = (int)((long)n * t / nthreads), b = (int)((long)n * (t + 1) / nthreads);
        Band &B = bands[t];
        int y0 = gy, y1 = -1;
        for (int p = a; p < b; ++p) {
            double y = row(pts, P.nPtsPitch, kPosX + 1)[p] * P.cellsize_inv - 0.5;
            int by = (y > -1.0 && y < (double)gy) ? (int)y : 0;
            y0 = std::min(y0, by);
            y1 = std::max(y1, std::min(by + 2, gy - 1));
        }
        if (y1 < y0) { y0 = 0; y1 = -1; }
        B.y0 = y0; B.y1 = y1; B.err = 0;
        size_t nrows = (size_t)(y1 - y0 + 1);
        B.buf.assign(3 * nrows * gx, 0.0);
        double *m = B.buf.data(), *px = m + nrows * gx, *py = px + nrows * gx;
        for (int p = a; p < b; ++p) {
            B.err |= p2g_particle(P, pts, p, [&](int node, double dm, double d0, double d1) {
                int ry = node / gx - y0;
                if (ry < 0 || ry >= (int)nrows) return;   // cannot happen for in-band rows
                size_t k = (size_t)ry * gx + node % gx;
                m[k] += dm; px[k] += d0; py[k] += d1;
            });
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (int y = 0; y < gy; ++y) {
            for (int tt = 0; tt < nthreads; ++tt) {
                Band &C = bands[tt];
                if (y < C.y0 || y > C.y1) continue;
                size_t nr = (size_t)(C.y1 - C.y0 + 1);
                const double *cm = C.buf.data() + (size_t)(y - C.y0) * gx;
                const double *cx = cm + nr * gx, *cy = cx + nr * gx;
                double *gm = grid + (size_t)y * gx, *gxv = gm + gp, *gyv = gm + 2 * gp;
                for (int x = 0; x < gx; ++x) { gm[x] += cm[x]; gxv[x] += cx[x]; gyv[x] += cy[x]; }
            }
        }
    }
    int err = 0;
    for (auto &B : bands) err |= B.err;
    return err;
#else
    (void)nthreads;
    return oracle_p2g(&P, pts, grid);
#endif
}

}  // namespace

// ===========================================================================
extern "C" {

int oracle_simparams_size(void) { return (int)sizeof(SimParams); }

int oracle_simparams_offsets(long *out, int max_n) {
#define OFF(f) offsetof(SimParams, f)
    const long offs[] = {
        OFF(grid_array), OFF(indenter_force_accumulator), OFF(pts_array), OFF(nPtsPitch),
        OFF(nGridPitch), OFF(n_indenter_subdivisions), OFF(tpb_P2G), OFF(tpb_Upd), OFF(tpb_G2P),
        OFF(nPts), OFF(GridX), OFF(GridY), OFF(GridTotal), OFF(GridXDimension),
        OFF(InitialTimeStep), OFF(SimulationEndTime), OFF(Gravity), OFF(Density),
        OFF(PoissonsRatio), OFF(YoungsModulus), OFF(lambda), OFF(mu), OFF(kappa),
        OFF(IceCompressiveStrength), OFF(IceTensileStrength), OFF(IceShearStrength),
        OFF(NACC_beta), OFF(NACC_M), OFF(NACC_Msq), OFF(DP_tan_phi), OFF(DP_threshold_p),
        OFF(cellsize), OFF(cellsize_inv), OFF(Dp_inv), OFF(UpdateEveryNthStep),
        OFF(IndDiameter), OFF(IndRSq), OFF(IndVelocity), OFF(IndDepth), OFF(xmin), OFF(xmax),
        OFF(ymin), OFF(ymax), OFF(nxmin), OFF(nxmax), OFF(nymin), OFF(nymax),
        OFF(ParticleVolume), OFF(ParticleMass), OFF(ParticleViewSize), OFF(SimulationStep),
        OFF(SimulationTime), OFF(indenter_x), OFF(indenter_x_initial), OFF(indenter_y),
        OFF(indenter_y_initial), OFF(Volume), OFF(SetupType), OFF(GrainVariability),
        OFF(UseWarpLevelReduction), OFF(SortPointsByGridCell)};
#undef OFF
    int n = (int)(sizeof(offs) / sizeof(offs[0]));
    for (int i = 0; i < n && i < max_n; ++i) out[i] = offs[i];
    return n;
}

void oracle_svd2x2(const double F[4], double U[4], double S[2], double V[4]) {
    svd2x2(F, U, S, V);
}

void oracle_kirchhoff(const double F[4], double mu, double kappa, double tau[4]) {
    kirchhoff(F, mu, kappa, tau);
}

void oracle_compute_pq(const double F[4], double mu, double kappa, double out[16]) {
    PQ r;
    compute_pq(F, mu, kappa, r);
    std::memcpy(out, r.U, 4 * sizeof(double));
    std::memcpy(out + 4, r.S, 2 * sizeof(double));
    std::memcpy(out + 6, r.V, 4 * sizeof(double));
    out[10] = r.Je;
    out[11] = r.p;
    out[12] = r.q;
    out[13] = r.shat[0];
    out[14] = r.shat[1];
    out[15] = r.SS[0];
    // SS[1] is S[1]^2; callers recompute it if needed
}

void oracle_grain_params(const void *prms, int grain, double out[5]) {
    grain_params(cp(prms), grain, out[0], out[1], out[2], out[3], out[4]);
}

int oracle_failure_check(const void *prms, int grain, double p, double q) {
    return failure_check(cp(prms), grain, p, q);
}

void oracle_dp_return(const void *prms, double Fe[4], double *Jp_inv) {
    const SimParams &P = cp(prms);
    PQ r;
    compute_pq(Fe, P.mu, P.kappa, r);
    dp_return(P, r, Fe, *Jp_inv);
}

void oracle_reset_grid(const void *prms, double *grid) {
    const SimParams &P = cp(prms);
    std::fill(grid, grid + 3 * P.nGridPitch, 0.0);
}

int oracle_p2g(const void *prms, const double *pts, double *grid) {
    const SimParams &P = cp(prms);
    const size_t gp = P.nGridPitch;
    int err = 0;
    for (int p = 0; p < P.nPts; ++p)
        err |= p2g_particle(P, pts, p, [&](int node, double dm, double d0, double d1) {
            grid[node] += dm;
            grid[gp + node] += d0;
            grid[2 * gp + node] += d1;
        });
    return err;
}

void oracle_update_nodes(const void *prms, double *grid, double *bins, double indenter_x,
                         double indenter_y, double fsum[2]) {
    const SimParams &P = cp(prms);
    double s0 = 0, s1 = 0;
    for (int idx = 0; idx < P.GridTotal; ++idx) {
        double f0 = 0, f1 = 0;
        int b = update_node(P, grid, idx, indenter_x, indenter_y, f0, f1);
        if (b >= 0) {
            bins[2 * b] += f0;
            bins[2 * b + 1] += f1;
            s0 += f0;
            s1 += f1;
        }
    }
    if (fsum) { fsum[0] = s0; fsum[1] = s1; }
}

int oracle_g2p(const void *prms, double *pts, const double *grid, int recordPQ) {
    const SimParams &P = cp(prms);
    int err = 0;
    for (int p = 0; p < P.nPts; ++p) err |= g2p_particle(P, pts, grid, p, recordPQ);
    return err;
}

int oracle_run(void *prms, double *pts, double *grid, double *bins, int nsub, double *hist,
               int nthreads) {
    SimParams &P = *static_cast<SimParams *>(prms);
    const int N = P.UpdateEveryNthStep;
    double t = P.SimulationTime;
    int err = 0;
    std::vector<double> tb;   // per-thread bin partials (parallel mode)
    for (int k = 0; k < nsub; ++k) {
        const int s = P.SimulationStep;
        if (s % N == 0) std::fill(bins, bins + 2 * P.n_indenter_subdivisions, 0.0);
        t += P.InitialTimeStep;
        P.indenter_x = P.indenter_x_initial + t * P.IndVelocity;
        std::fill(grid, grid + 3 * P.nGridPitch, 0.0);
        const int record = ((s + 1) % N == 0) ? 1 : 0;
        double fs[2] = {0, 0};
        if (nthreads <= 1) {
            err |= oracle_p2g(&P, pts, grid);
            oracle_update_nodes(&P, grid, bins, P.indenter_x, P.indenter_y, fs);
            err |= oracle_g2p(&P, pts, grid, record);
        } else {
#ifdef _OPENMP
            err |= p2g_parallel(P, pts, grid, nthreads);
            const int nb = P.n_indenter_subdivisions;
            tb.assign((size_t)nthreads * (2 * nb + 2), 0.0);
#pragma omp parallel num_threads(nthreads)
            {
                int th = omp_get_thread_num();
                double *mb = tb.data() + (size_t)th * (2 * nb + 2);
#pragma omp for schedule(static)
                for (int idx = 0; idx < P.GridTotal; ++idx) {
                    double f0 = 0, f1 = 0;
                    int b = update_node(P, grid, idx, P.indenter_x, P.indenter_y, f0, f1);
                    if (b >= 0) {
                        mb[2 * b] += f0; mb[2 * b + 1] += f1;
                        mb[2 * nb] += f0; mb[2 * nb + 1] += f1;
                    }
                }
            }
            for (int th = 0; th < nthreads; ++th) {
                const double *mb = tb.data() + (size_t)th * (2 * nb + 2);
                for (int b = 0; b < 2 * nb; ++b) bins[b] += mb[b];
                fs[0] += mb[2 * nb]; fs[1] += mb[2 * nb + 1];
            }
            int gerr = 0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(| : gerr)
            for (int p = 0; p < P.nPts; ++p) gerr |= g2p_particle(P, pts, grid, p, record);
            err |= gerr;
#else
            err |= oracle_p2g(&P, pts, grid);
            oracle_update_nodes(&P, grid, bins, P.indenter_x, P.indenter_y, fs);
            err |= oracle_g2p(&P, pts, grid, record);
#endif
        }
        if (hist) { hist[2 * k] = fs[0]; hist[2 * k + 1] = fs[1]; }
        P.SimulationStep = s + 1;
        P.SimulationTime = t;
    }
    return err;
}

// ---- learned return map: FP64 restatement of the MLP of csrc/mlp.hip ----
// params: W1[64][16] b1[64] W2[64][64] b2[64] W3[16][64] b3[16] (6288);
// X is B x 5, T is B x 3.  Returns L = 1/(2B) sum (y - t)^2 over the 3
// outputs; grad (6288, may be NULL) receives dL/dparams; Y (B x 3, may be
// NULL) the outputs.
double oracle_mlp_loss_grad(const double *P, const double *X, const double *T, int B,
                            double *grad, double *Y) {
    const int oW1 = 0, oB1 = 1024, oW2 = 1088, oB2 = 5184, oW3 = 5248, oB3 = 6272, NP = 6288;
    if (grad) std::fill(grad, grad + NP, 0.0);
    double L = 0;
    double x[16], h1[64], h2[64], y[16], d2[64], d1[64], dy[16];
    for (int r = 0; r < B; ++r) {
        for (int k = 0; k < 16; ++k) x[k] = k < 5 ? X[5 * r + k] : 0.0;
        for (int n = 0; n < 64; ++n) {
            double a = P[oB1 + n];
            for (int k = 0; k < 16; ++k) a += P[oW1 + n * 16 + k] * x[k];
            h1[n] = std::tanh(a);
        }
        for (int n = 0; n < 64; ++n) {
            double a = P[oB2 + n];
            for (int k = 0; k < 64; ++k) a += P[oW2 + n * 64 + k] * h1[k];
            h2[n] = std::tanh(a);
        }
        for (int o = 0; o < 16; ++o) {
            double a = P[oB3 + o];
            for (int k = 0; k < 64; ++k) a += P[oW3 + o * 64 + k] * h2[k];
            y[o] = a;
        }
        for (int o = 0; o < 16; ++o) {
            dy[o] = 0;
            if (o < 3) {
                const double d = y[o] - T[3 * r + o];
                L += 0.5 * d * d / B;
                dy[o] = d / B;
                if (Y) Y[3 * r + o] = y[o];
            }
        }
        if (!grad) continue;
        for (int o = 0; o < 16; ++o) {
            grad[oB3 + o] += dy[o];
            for (int k = 0; k < 64; ++k) grad[oW3 + o * 64 + k] += dy[o] * h2[k];
        }
        for (int k = 0; k < 64; ++k) {
            double a = 0;
            for (int o = 0; o < 16; ++o) a += dy[o] * P[oW3 + o * 64 + k];
            d2[k] = a * (1 - h2[k] * h2[k]);
        }
        for (int n = 0; n < 64; ++n) {
            grad[oB2 + n] += d2[n];
            for (int k = 0; k < 64; ++k) grad[oW2 + n * 64 + k] += d2[n] * h1[k];
        }
        for (int k = 0; k < 64; ++k) {
            double a = 0;
            for (int n = 0; n < 64; ++n) a += d2[n] * P[oW2 + n * 64 + k];
            d1[k] = a * (1 - h1[k] * h1[k]);
        }
        for (int n = 0; n < 64; ++n) {
            grad[oB1 + n] += d1[n];
            for (int k = 0; k < 16; ++k) grad[oW1 + n * 16 + k] += d1[n] * x[k];
        }
    }
    return L;
}

void oracle_cell_keys_f32(const float *xy, int n, double hinv, int gridx, int *keys) {
    for (int i = 0; i < n; ++i) {
        float ax = xy[2 * i], ay = xy[2 * i + 1];
        keys[i] = (int)(ax * hinv - 0.5) + ((int)(ay * hinv - 0.5)) * gridx;
    }
}

void oracle_cell_keys_soa(const double *pts, size_t pitch, int n, double hinv, int gridx,
                          int *keys) {
    const double *px = pts + pitch * kPosX, *py = pts + pitch * (kPosX + 1);
    for (int i = 0; i < n; ++i)
        keys[i] = (int)(px[i] * hinv - 0.5) + ((int)(py[i] * hinv - 0.5)) * gridx;
}

void oracle_stable_order(const int *keys, int n, int *perm) {
    std::iota(perm, perm + n, 0);
    std::stable_sort(perm, perm + n, [keys](int a, int b) { return keys[a] < keys[b]; });
}

}  // extern "C"
