// api.hip -- the C ABI of libicempm_hip.so (declared in include/icempm.h).
//
// One icempm_ctx replaces one GPU_Implementation5 object
// (simulation/gpu_implementation5.cu:434-593): it owns the HIP stream, the
// device arrays, the pinned host buffers and the device error flag.  All
// device work is queued on the context's stream.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/icempm.h"
#include "../host/parameters_sim.h"
#include "internal.h"

using namespace icempm;

struct icempm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t evStart = nullptr, evStop = nullptr;
    hipEvent_t evXfer = nullptr;   // recorded after the readback's copies
    bool xferPending = false;
    std::string lastError;

    icy::SimParams prms{};
    bool haveParams = false;
    int mode = ICEMPM_MODE_FAST;
    int passKind = 2;     // particle-pass kernel variant (see internal.h)

    // device arrays
    double *pts = nullptr, *ptsTmp = nullptr;
    size_t pp = 0, nPtsAlloc = 0;
    double *gridA = nullptr, *gridB = nullptr;
    size_t gp = 0, nGridAlloc = 0;
    double *bins = nullptr;
    int nBinsAlloc = 0;   // doubles
    int *dErr = nullptr;
    double *grainTab = nullptr;   // kGrainEntries x kGrainFields
    double *hist = nullptr;
    size_t histCap = 0;   // doubles
    int lastRunSubsteps = 0;

    // chunk table of the particle pass
    int *chunkStart = nullptr;
    int nChunks = 0;
    size_t chunkCap = 0;
    bool chunksDirty = true;

    // sort scratch
    int *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr, *sortTmp = nullptr;
    size_t sortCap = 0, sortTmpCap = 0;

    // ORDERED-mode scratch
    OrderedScratch ord;

    // pinned host buffers
    double *hostPts = nullptr, *hostForce = nullptr;
    int *hostErr = nullptr;
    std::atomic<int> errorCode{0};

    // kernel timing
    bool timing = false;
    int timingStride = 1;
    std::vector<hipEvent_t> tev;
    std::vector<int> tevUsed;
    double timedMs = 0;
    int timedLaunches = 0;
};

namespace {

thread_local std::string g_createError;

int fail(icempm_ctx *c, int code, const std::string &msg) {
    if (c) c->lastError = msg;
    return code;
}

int hipfail(icempm_ctx *c, hipError_t e, const char *tag) {
    return fail(c, ICEMPM_E_HIP, std::string(tag) + ": " + hipGetErrorString(e));
}

#define HIPCHK(c, call, tag)                           \
    do {                                               \
        hipError_t _e = (call);                        \
        if (_e != hipSuccess) return hipfail(c, _e, tag); \
    } while (0)

size_t round_pitch(size_t n) {
    size_t p = (n + 63) / 64 * 64;
    return p < 64 ? 64 : p;
}

template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <class T>
void hfree(T *&p) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
}

void grain_row(const icy::SimParams &P, int gi, double *row);

KParams make_kparams(const icempm_ctx *c) {
    const icy::SimParams &P = c->prms;
    KParams k{};
    k.pts = c->pts;
    k.grainTab = c->grainTab;
    k.pp = c->pp;
    k.gp = c->gp;
    k.bins = c->bins;
    k.err = c->dErr;
    k.nPts = P.nPts;
    k.gridX = P.GridX;
    k.gridY = P.GridY;
    k.gridTotal = P.GridTotal;
    k.nbins = P.n_indenter_subdivisions;
    k.dt = P.InitialTimeStep;
    k.vol = P.ParticleVolume;
    k.mass = P.ParticleMass;
    k.h = P.cellsize;
    k.h_inv = P.cellsize_inv;
    k.Dinv = P.Dp_inv;
    k.mu = P.mu;
    k.kappa = P.kappa;
    k.gravity = P.Gravity;
    k.indRsq = P.IndRSq;
    k.indV = P.IndVelocity;
    k.vmax = 0.5 * P.cellsize / P.InitialTimeStep;
    k.ics = P.IceCompressiveStrength;
    k.its = P.IceTensileStrength;
    k.iss = P.IceShearStrength;
    k.grainVar = P.GrainVariability;
    k.tanPhi = P.DP_tan_phi;
    k.dpThr = P.DP_threshold_p;
    grain_row(P, 0, k.g0);
    return k;
}

// Per-grain constants of the NACC test for every short grain id, evaluated in
// the order of GetParametersForGrain / CheckIfPointIsInsideFailureSurface
// (gpu_implementation5.cu:97-141), so that the table lookup returns exactly
// what the per-particle formula would (this file is built without FMA
// contraction; sqrt and division are IEEE on host and device alike).
void grain_row(const icy::SimParams &P, int gi, double *row) {
    const short grain = (short)gi;
    const double var1 = 1.0 + P.GrainVariability * 0.05 * (-10 + grain % 21);
    const double var2 = 1.0 + P.GrainVariability * 0.033 * (-15 + (grain + 3) % 30);
    const double var3 = 1.0 + P.GrainVariability * 0.1 * (-10 + (grain + 4) % 11);
    const double pmax = P.IceCompressiveStrength * var1;
    const double pmin = -P.IceTensileStrength * var2;
    const double qmax = P.IceShearStrength * var3;
    const double beta = -pmin / pmax;
    const double a = pmax * (1 + beta);
    const double mSq = (4 * qmax * qmax * (1 + 2 * beta)) / (a * a);
    const double pmin2 = -3e6;
    const double q0 = 2 * std::sqrt(-pmax * pmin) * qmax / (pmax - pmin);
    const double slope = -q0 / pmin2;
    row[kGPmax] = pmax;
    row[kGMsq] = mSq;
    row[kGBetaPmax] = beta * pmax;
    row[kGOnePlus2Beta] = 1. + 2. * beta;
    row[kGSlope] = slope;
}

void grain_table(const icy::SimParams &P, std::vector<double> &t) {
    t.assign((size_t)kGrainEntries * kGrainFields, 0.0);
    for (int gi = -32768; gi < 32768; ++gi) grain_row(P, gi, t.data() + (size_t)(gi + 32768) * kGrainFields);
}

// Launch preconditions shared by every substep operation.
int check_ready(icempm_ctx *c, const char *tag) {
    if (!c) return ICEMPM_E_ARG;
    if (!c->haveParams) return fail(c, ICEMPM_E_STATE, std::string(tag) + ": constants not uploaded");
    if (!c->pts || !c->gridA) return fail(c, ICEMPM_E_STATE, std::string(tag) + ": arrays not allocated");
    const icy::SimParams &P = c->prms;
    if (P.nPts < 0 || (size_t)P.nPts > c->nPtsAlloc)
        return fail(c, ICEMPM_E_STATE, std::string(tag) + ": nPts exceeds the allocation");
    if (P.GridX <= 0 || P.GridY <= 0 || P.GridTotal != P.GridX * P.GridY ||
        (size_t)P.GridTotal > c->nGridAlloc)
        return fail(c, ICEMPM_E_STATE, std::string(tag) + ": grid size does not match the allocation");
    if (P.n_indenter_subdivisions <= 0 || 2 * P.n_indenter_subdivisions > c->nBinsAlloc)
        return fail(c, ICEMPM_E_STATE, std::string(tag) + ": indenter bins exceed the allocation");
    if (!(P.cellsize > 0) || !(P.InitialTimeStep > 0))
        return fail(c, ICEMPM_E_ARG, std::string(tag) + ": cellsize and time step must be positive");
    return ICEMPM_OK;
}

int ensure_sort_scratch(icempm_ctx *c, size_t n) {
    const size_t need = n < 1 ? 1 : n;
    if (c->sortCap < need) {
        dfree(c->keys); dfree(c->vals); dfree(c->keys2); dfree(c->vals2);
        HIPCHK(c, hipMalloc(&c->keys, need * sizeof(int)), "sort scratch");
        HIPCHK(c, hipMalloc(&c->vals, need * sizeof(int)), "sort scratch");
        HIPCHK(c, hipMalloc(&c->keys2, need * sizeof(int)), "sort scratch");
        HIPCHK(c, hipMalloc(&c->vals2, need * sizeof(int)), "sort scratch");
        c->sortCap = need;
    }
    const size_t tneed = radix_tmp_ints(need);
    if (c->sortTmpCap < tneed) {
        dfree(c->sortTmp);
        HIPCHK(c, hipMalloc(&c->sortTmp, tneed * sizeof(int)), "sort scratch");
        c->sortTmpCap = tneed;
    }
    return ICEMPM_OK;
}

// Chunk table of the particle pass: runs of consecutive particles whose
// base cells share one cell row and span fewer than kWinX-2 columns, at
// most kChunk long.  Built from the keys of the current positions; any
// partition is correct, a compact one keeps P2G in shared memory.
int build_chunks(icempm_ctx *c) {
    const int n = c->prms.nPts;
    const int gx = c->prms.GridX;
    std::vector<int> starts;
    if (n > 0) {
        int rc = ensure_sort_scratch(c, (size_t)n);
        if (rc) return rc;
        HIPCHK(c, launch_point_keys(c->pts, c->pp, n, c->prms.cellsize_inv, gx, c->keys, nullptr,
                                    c->stream), "chunk keys");
        std::vector<int> hk((size_t)n);
        HIPCHK(c, hipMemcpyAsync(hk.data(), c->keys, (size_t)n * sizeof(int),
                                 hipMemcpyDeviceToHost, c->stream), "chunk keys");
        HIPCHK(c, hipStreamSynchronize(c->stream), "chunk keys");
        starts.reserve((size_t)n / 64 + 2);
        int cs = 0;
        starts.push_back(0);
        auto rowof = [gx](int key) { return key >= 0 ? key / gx : -1; };
        auto colof = [gx](int key) { return key >= 0 ? key % gx : -1; };
        if (c->passKind == 0) {
            // workgroup windows: one cell row, columns ascending within the window
            for (int i = 1; i < n; ++i) {
                const int kc = hk[(size_t)cs], ki = hk[(size_t)i];
                const bool brk = (i - cs >= kChunk) || rowof(ki) != rowof(kc) ||
                                 colof(ki) < colof(kc) || colof(ki) > colof(kc) + (kWinX - 3);
                if (brk) {
                    starts.push_back(i);
                    cs = i;
                }
            }
        } else {
            // wavefront windows (4 rows): at most two cell rows and 126 columns per
            // chunk.  The ingest sorts by the cell of the raw coordinates; after
            // the block offset the points of one raw row can interleave two rows.
            int r0 = rowof(hk[0]), r1 = r0, c0 = colof(hk[0]), c1 = c0;
            for (int i = 1; i < n; ++i) {
                const int ki = hk[(size_t)i], r = rowof(ki), cc = colof(ki);
                const int nr0 = std::min(r0, r), nr1 = std::max(r1, r);
                const int nc0 = std::min(c0, cc), nc1 = std::max(c1, cc);
                const bool brk = (i - cs >= kChunk) || nr1 - nr0 > 1 || nc1 - nc0 > 125 || ki < 0 ||
                                 hk[(size_t)cs] < 0;
                if (brk) {
                    starts.push_back(i);
                    cs = i;
                    r0 = r1 = r;
                    c0 = c1 = cc;
                } else {
                    r0 = nr0; r1 = nr1; c0 = nc0; c1 = nc1;
                }
            }
        }
        // unsorted input: fixed chunks (correct, P2G then falls back to
        // global atomics for the particles outside a window)
        if (starts.size() > (size_t)n / 32 + 64) {
            starts.clear();
            for (int i = 0; i < n; i += kChunk) starts.push_back(i);
        }
        starts.push_back(n);
    } else {
        starts.push_back(0);
    }
    const size_t need = starts.size();
    if (c->chunkCap < need) {
        dfree(c->chunkStart);
        HIPCHK(c, hipMalloc(&c->chunkStart, need * sizeof(int)), "chunk table");
        c->chunkCap = need;
    }
    HIPCHK(c, hipMemcpyAsync(c->chunkStart, starts.data(), need * sizeof(int),
                             hipMemcpyHostToDevice, c->stream), "chunk table");
    HIPCHK(c, hipStreamSynchronize(c->stream), "chunk table");
    c->nChunks = (int)need - 1;
    c->chunksDirty = false;
    return ICEMPM_OK;
}

int ensure_chunks(icempm_ctx *c) { return c->chunksDirty ? build_chunks(c) : ICEMPM_OK; }

int ensure_ordered(icempm_ctx *c) {
    OrderedScratch &o = c->ord;
    const size_t N = (size_t)c->prms.nPts, G = (size_t)c->prms.GridTotal;
    if (o.nPts < N || !o.contrib) {
        dfree(o.contrib); dfree(o.cnode); dfree(o.list);
        const size_t n = N < 1 ? 1 : N;
        HIPCHK(c, hipMalloc(&o.contrib, 27 * n * sizeof(double)), "ordered scratch");
        HIPCHK(c, hipMalloc(&o.cnode, 9 * n * sizeof(int)), "ordered scratch");
        HIPCHK(c, hipMalloc(&o.list, 9 * n * sizeof(int)), "ordered scratch");
        o.nPts = n;
    }
    if (o.nGrid < G || !o.count) {
        dfree(o.count); dfree(o.start); dfree(o.cursor); dfree(o.nodeBin); dfree(o.nodeF);
        dfree(o.flags); dfree(o.cpos); dfree(o.contacts); dfree(o.scanTmp);
        HIPCHK(c, hipMalloc(&o.count, (G + 1) * sizeof(int)), "ordered scratch");
        HIPCHK(c, hipMalloc(&o.start, (G + 1) * sizeof(int)), "ordered scratch");
        HIPCHK(c, hipMalloc(&o.cursor, (G + 1) * sizeof(int)), "ordered scratch");
        HIPCHK(c, hipMalloc(&o.nodeBin, (G + 1) * sizeof(int)), "ordered scratch");
        HIPCHK(c, hipMalloc(&o.nodeF, 2 * (G + 1) * sizeof(double)), "ordered scratch");
        HIPCHK(c, hipMalloc(&o.flags, (G + 1) * sizeof(int)), "ordered scratch");
        HIPCHK(c, hipMalloc(&o.cpos, (G + 1) * sizeof(int)), "ordered scratch");
        HIPCHK(c, hipMalloc(&o.contacts, (G + 1) * sizeof(int)), "ordered scratch");
        HIPCHK(c, hipMalloc(&o.scanTmp, scan_tmp_ints(G + 1) * sizeof(int)), "ordered scratch");
        o.nGrid = G;
    }
    return ICEMPM_OK;
}

void free_all(icempm_ctx *c) {
    dfree(c->pts); dfree(c->ptsTmp); dfree(c->gridA); dfree(c->gridB); dfree(c->bins);
    dfree(c->hist); dfree(c->chunkStart);
    dfree(c->keys); dfree(c->vals); dfree(c->keys2); dfree(c->vals2); dfree(c->sortTmp);
    OrderedScratch &o = c->ord;
    dfree(o.contrib); dfree(o.cnode); dfree(o.count); dfree(o.start); dfree(o.cursor);
    dfree(o.list); dfree(o.nodeBin); dfree(o.nodeF); dfree(o.flags); dfree(o.cpos);
    dfree(o.contacts); dfree(o.scanTmp);
    o = OrderedScratch{};
    hfree(c->hostPts); hfree(c->hostForce);
    c->histCap = 0; c->chunkCap = 0; c->sortCap = 0; c->sortTmpCap = 0;
    c->nPtsAlloc = 0; c->nGridAlloc = 0; c->nBinsAlloc = 0; c->nChunks = 0;
    c->chunksDirty = true;
}

// the particle pass of the context's mode: FMA-contracted in FAST mode,
// exact in ORDERED mode
hipError_t particles(icempm_ctx *c, const KParams &k, bool g2p, bool p2g, const double *gridV,
                     double *gridOut, int recordPQ) {
    if (c->mode == ICEMPM_MODE_ORDERED)
        return exact::launch_particles(k, g2p, p2g, gridV, gridOut, c->chunkStart, c->nChunks,
                                       recordPQ, c->stream, c->passKind);
    return fast::launch_particles(k, g2p, p2g, gridV, gridOut, c->chunkStart, c->nChunks,
                                  recordPQ, c->stream, c->passKind);
}

// the dominant kernel of a substep, optionally bracketed by timing events
int launch_particles_timed(icempm_ctx *c, const KParams &k, bool g2p, bool p2g,
                           const double *gridV, double *gridOut, int recordPQ, int timeIt) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timeIt) {
        size_t slot = 0;
        for (; slot < c->tevUsed.size(); ++slot)
            if (!c->tevUsed[slot]) break;
        if (slot == c->tevUsed.size()) {
            hipEvent_t a, b;
            HIPCHK(c, hipEventCreate(&a), "timing events");
            HIPCHK(c, hipEventCreate(&b), "timing events");
            c->tev.push_back(a);
            c->tev.push_back(b);
            c->tevUsed.push_back(0);
        }
        c->tevUsed[slot] = 1;
        e0 = c->tev[2 * slot];
        e1 = c->tev[2 * slot + 1];
        HIPCHK(c, hipEventRecord(e0, c->stream), "timing events");
    }
    HIPCHK(c, particles(c, k, g2p, p2g, gridV, gridOut, recordPQ), "particle pass");
    if (timeIt) HIPCHK(c, hipEventRecord(e1, c->stream), "timing events");
    return ICEMPM_OK;
}

int collect_timing(icempm_ctx *c) {
    for (size_t s = 0; s < c->tevUsed.size(); ++s) {
        if (!c->tevUsed[s]) continue;
        HIPCHK(c, hipEventSynchronize(c->tev[2 * s + 1]), "timing events");
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, c->tev[2 * s], c->tev[2 * s + 1]), "timing events");
        c->timedMs += ms;
        c->timedLaunches += 1;
        c->tevUsed[s] = 0;
    }
    return ICEMPM_OK;
}

struct ReadbackJob {
    icempm_ctx *ctx;
    icempm_callback cb;
    void *user;
};

// ---------------------------------------------------------------------------
// Scratch of the allocation-free sort (icempm_sort_raw_points_noalloc): one
// static device array, a copy on every GPU the library runs on, held until
// the process ends.  In host code `scratch` names a host-side stand-in, not
// a device address: icempm_create looks the array up on its device with
// hipGetSymbolAddress and keeps that address per device, and only that
// address reaches kernels and copies.  Layout for n <= kNoAllocMaxPoints, in
// 256-byte aligned pieces: staged xy in and out (8 B a point each), grains in
// and out (2 B), keys, values and their ping-pong copies (4 B each), radix
// histograms and scan sums.
constexpr size_t kNoAllocMaxPoints = ICEMPM_NOALLOC_MAX_POINTS;
constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }
// radix_tmp_ints(n) = 2 h + scan_tmp_ints(h) + 64 with h = 16 ceil(n / 2048);
// scan_tmp_ints(h) < h / 2048 + 64 per level: 3 h + 1024 ints bound it
constexpr size_t kRadixTmpInts = 3 * (16 * ((kNoAllocMaxPoints + 2047) / 2048)) + 1024;
constexpr size_t kScratchBytes = 2 * al256(kNoAllocMaxPoints * 8) + 2 * al256(kNoAllocMaxPoints * 2) +
                                 4 * al256(kNoAllocMaxPoints * 4) + al256(kRadixTmpInts * 4);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
unsigned char *g_scratchPtr[kMaxDevices] = {};

// the device address of `scratch` on `dev` (current), looked up once per device
hipError_t scratch_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratchPtr[dev]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratchPtr[dev] = static_cast<unsigned char *>(ptr);
    return hipSuccess;
}

unsigned char *scratch_of(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    return g_scratchPtr[dev];
}

void readback_done(void *p) {
    ReadbackJob *j = static_cast<ReadbackJob *>(p);
    j->ctx->errorCode.store(*j->ctx->hostErr);
    if (j->cb) j->cb(j->user);
    delete j;
}

}  // namespace

// ===========================================================================
extern "C" {

int icempm_create(icempm_ctx **out) {
    return host_guard(&g_createError, [&]() -> int {
        if (!out) return ICEMPM_E_ARG;
        *out = nullptr;
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) {
            g_createError = std::string("initialize() hip error: ") + hipGetErrorString(e);
            return ICEMPM_E_NODEV;
        }
        if (n == 0) {
            g_createError = "No available HIP devices";
            return ICEMPM_E_NODEV;
        }
        icempm_ctx *c = new icempm_ctx();
        c->device = 0;
        e = hipSetDevice(0);
        if (e == hipSuccess) e = scratch_setup(c->device);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&c->evStart);
        if (e == hipSuccess) e = hipEventCreate(&c->evStop);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->evXfer, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc(&c->dErr, sizeof(int));
        if (e == hipSuccess) e = hipHostMalloc(&c->hostErr, sizeof(int), hipHostMallocDefault);
        if (e == hipSuccess) e = launch_fill_i32(c->dErr, 1, 0, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            g_createError = std::string("initialize(): ") + hipGetErrorString(e);
            icempm_destroy(c);
            return ICEMPM_E_HIP;
        }
        *c->hostErr = 0;
        *out = c;
        return ICEMPM_OK;
    });
}

void icempm_destroy(icempm_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    for (hipEvent_t ev : c->tev) (void)hipEventDestroy(ev);
    if (c->evStart) (void)hipEventDestroy(c->evStart);
    if (c->evStop) (void)hipEventDestroy(c->evStop);
    if (c->evXfer) (void)hipEventDestroy(c->evXfer);
    dfree(c->dErr);
    dfree(c->grainTab);
    hfree(c->hostErr);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *icempm_last_error(const icempm_ctx *c) {
    return c ? c->lastError.c_str() : g_createError.c_str();
}

size_t icempm_simparams_size(void) { return sizeof(icy::SimParams); }

int icempm_test(icempm_ctx *c) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c) return ICEMPM_E_ARG;
        int *flag = nullptr;
        HIPCHK(c, hipMalloc(&flag, sizeof(int)), "cuda test");
        int h = 0;
        hipError_t e = launch_fill_i32(flag, 1, 0, c->stream);
        if (e == hipSuccess) e = launch_hello(flag, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        (void)hipFree(flag);
        if (e != hipSuccess) return hipfail(c, e, "cuda test");
        if (h != 1) return fail(c, ICEMPM_E_HIP, "cuda test: kernel did not run");
        return ICEMPM_OK;
    });
}

int icempm_synchronize(icempm_ctx *c) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c) return ICEMPM_E_ARG;
        HIPCHK(c, hipStreamSynchronize(c->stream), "synchronize");
        HIPCHK(c, hipDeviceSynchronize(), "synchronize");
        return ICEMPM_OK;
    });
}

int icempm_update_constants(icempm_ctx *c, const void *simparams, size_t nbytes) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || !simparams) return ICEMPM_E_ARG;
        if (nbytes != sizeof(icy::SimParams))
            return fail(c, ICEMPM_E_ARG, "cuda_update_constants: SimParams size mismatch");
        std::memcpy(&c->prms, simparams, sizeof(icy::SimParams));
        if (!c->grainTab)
            HIPCHK(c, hipMalloc(&c->grainTab, (size_t)kGrainEntries * kGrainFields * sizeof(double)),
                   "cuda_update_constants: gprms");
        {
            std::vector<double> t;
            grain_table(c->prms, t);
            HIPCHK(c, hipMemcpyAsync(c->grainTab, t.data(), t.size() * sizeof(double),
                                     hipMemcpyHostToDevice, c->stream), "cuda_update_constants: gprms");
            HIPCHK(c, hipStreamSynchronize(c->stream), "cuda_update_constants: gprms");
        }
        c->haveParams = true;
        c->chunksDirty = true;
        // gpu_error_indicator <- error_code (gpu_implementation5.cu:474)
        HIPCHK(c, launch_fill_i32(c->dErr, 1, c->errorCode.load(), c->stream),
               "gpu_error_indicator initialization");
        HIPCHK(c, hipStreamSynchronize(c->stream), "cuda_update_constants: gprms");
        return ICEMPM_OK;
    });
}

int icempm_allocate_arrays(icempm_ctx *c, size_t nGridNodes, size_t nPoints, int nBins,
                           size_t *nGridPitch, size_t *nPtsPitch, double **host_points,
                           double **host_force) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || nBins <= 0) return ICEMPM_E_ARG;
        HIPCHK(c, hipStreamSynchronize(c->stream), "cuda_allocate_arrays");
        free_all(c);
        c->gp = round_pitch(nGridNodes);
        c->pp = round_pitch(nPoints);
        HIPCHK(c, hipMalloc(&c->gridA, c->gp * kGridRows * sizeof(double)), "cuda_allocate_arrays");
        HIPCHK(c, hipMalloc(&c->gridB, c->gp * kGridRows * sizeof(double)), "cuda_allocate_arrays");
        HIPCHK(c, hipMalloc(&c->pts, c->pp * kPtsRows * sizeof(double)), "cuda_allocate_arrays");
        HIPCHK(c, hipMalloc(&c->bins, (size_t)nBins * sizeof(double)), "cuda_allocate_arrays");
        HIPCHK(c, hipHostMalloc(&c->hostPts, c->pp * kPtsRows * sizeof(double), hipHostMallocDefault),
               "cuda_allocate_arrays");
        HIPCHK(c, hipHostMalloc(&c->hostForce, (size_t)nBins * sizeof(double), hipHostMallocDefault),
               "cuda_allocate_arrays");
        std::memset(c->hostPts, 0, c->pp * kPtsRows * sizeof(double));
        std::memset(c->hostForce, 0, (size_t)nBins * sizeof(double));
        HIPCHK(c, launch_fill_f64(c->gridA, c->gp * kGridRows, 0.0, c->stream), "cuda_allocate_arrays");
        HIPCHK(c, launch_fill_f64(c->gridB, c->gp * kGridRows, 0.0, c->stream), "cuda_allocate_arrays");
        HIPCHK(c, launch_fill_f64(c->pts, c->pp * kPtsRows, 0.0, c->stream), "cuda_allocate_arrays");
        HIPCHK(c, launch_fill_f64(c->bins, (size_t)nBins, 0.0, c->stream), "cuda_allocate_arrays");
        HIPCHK(c, hipStreamSynchronize(c->stream), "cuda_allocate_arrays");
        c->nPtsAlloc = nPoints;
        c->nGridAlloc = nGridNodes;
        c->nBinsAlloc = nBins;
        c->errorCode.store(0);
        *c->hostErr = 0;
        HIPCHK(c, launch_fill_i32(c->dErr, 1, 0, c->stream), "cuda_allocate_arrays");
        if (nGridPitch) *nGridPitch = c->gp;
        if (nPtsPitch) *nPtsPitch = c->pp;
        if (host_points) *host_points = c->hostPts;
        if (host_force) *host_force = c->hostForce;
        return ICEMPM_OK;
    });
}

int icempm_transfer_points_to_device(icempm_ctx *c) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c) return ICEMPM_E_ARG;
        if (!c->pts) return fail(c, ICEMPM_E_STATE, "transfer_points_to_device: arrays not allocated");
        HIPCHK(c, hipMemcpyAsync(c->pts, c->hostPts, c->pp * kPtsRows * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream), "transfer_points_to_device");
        HIPCHK(c, hipStreamSynchronize(c->stream), "transfer_points_to_device");
        std::memset(c->hostForce, 0, (size_t)c->nBinsAlloc * sizeof(double));
        c->chunksDirty = true;
        return ICEMPM_OK;
    });
}

int icempm_reset_grid(icempm_ctx *c) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c) return ICEMPM_E_ARG;
        if (!c->gridA) return fail(c, ICEMPM_E_STATE, "cuda_reset_grid error: arrays not allocated");
        HIPCHK(c, launch_fill_f64(c->gridA, c->gp * kGridRows, 0.0, c->stream), "cuda_reset_grid error");
        return ICEMPM_OK;
    });
}

int icempm_reset_indenter_force_accumulator(icempm_ctx *c) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c) return ICEMPM_E_ARG;
        if (!c->bins) return fail(c, ICEMPM_E_STATE, "cuda_reset_grid error: arrays not allocated");
        HIPCHK(c, launch_fill_f64(c->bins, (size_t)c->nBinsAlloc, 0.0, c->stream),
               "cuda_reset_grid error");
        return ICEMPM_OK;
    });
}

int icempm_p2g(icempm_ctx *c) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        int rc = check_ready(c, "cuda_p2g");
        if (rc) return rc;
        const KParams k = make_kparams(c);
        if (c->mode == ICEMPM_MODE_ORDERED) {
            if ((rc = ensure_ordered(c))) return rc;
            HIPCHK(c, ordered_p2g(k, c->gridA, c->ord, c->stream), "cuda_p2g");
            return ICEMPM_OK;
        }
        if ((rc = ensure_chunks(c))) return rc;
        HIPCHK(c, particles(c, k, false, true, c->gridA, c->gridA, 0), "cuda_p2g");
        return ICEMPM_OK;
    });
}

int icempm_update_nodes(icempm_ctx *c, double indenter_x, double indenter_y) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        int rc = check_ready(c, "cuda_update_nodes");
        if (rc) return rc;
        const KParams k = make_kparams(c);
        if (c->mode == ICEMPM_MODE_ORDERED) {
            if ((rc = ensure_ordered(c))) return rc;
            HIPCHK(c, launch_update(k, c->gridA, indenter_x, indenter_y, nullptr, nullptr, true,
                                    c->ord.nodeBin, c->ord.nodeF, c->stream), "cuda_update_nodes");
            HIPCHK(c, ordered_fold_forces(k, c->ord, nullptr, c->stream), "cuda_update_nodes");
            return ICEMPM_OK;
        }
        HIPCHK(c, launch_update(k, c->gridA, indenter_x, indenter_y, nullptr, nullptr, false, nullptr,
                                nullptr, c->stream), "cuda_update_nodes");
        return ICEMPM_OK;
    });
}

int icempm_g2p(icempm_ctx *c, int recordPQ) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        int rc = check_ready(c, "cuda_g2p");
        if (rc) return rc;
        if ((rc = ensure_chunks(c))) return rc;
        const KParams k = make_kparams(c);
        HIPCHK(c, particles(c, k, true, false, c->gridA, nullptr, recordPQ ? 1 : 0), "cuda_g2p");
        return ICEMPM_OK;
    });
}

int icempm_run_substeps(icempm_ctx *c, int first_step, int n, double *t,
                        double indenter_x_initial, double indenter_y, int reset_bins_at_cycle) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        int rc = check_ready(c, "run_substeps");
        if (rc) return rc;
        if (!t || n < 0 || first_step < 0) return fail(c, ICEMPM_E_ARG, "run_substeps: bad arguments");
        const int N = c->prms.UpdateEveryNthStep;
        if (N <= 0) return fail(c, ICEMPM_E_ARG, "run_substeps: UpdateEveryNthStep must be positive");
        const size_t histNeed = 2 * (size_t)n;
        if (histNeed > c->histCap) {
            dfree(c->hist);
            const size_t cap = histNeed < 64 ? 64 : histNeed;
            HIPCHK(c, hipMalloc(&c->hist, cap * sizeof(double)), "run_substeps");
            c->histCap = cap;
        }
        HIPCHK(c, launch_fill_f64(c->hist, histNeed, 0.0, c->stream), "run_substeps");
        c->lastRunSubsteps = n;
        if (n == 0) return ICEMPM_OK;
        const double dt = c->prms.InitialTimeStep, V = c->prms.IndVelocity;
        double tt = *t;
        KParams k = make_kparams(c);
        const bool ordered = c->mode == ICEMPM_MODE_ORDERED;
        if (ordered) {
            if ((rc = ensure_ordered(c))) return rc;
        }
        if ((rc = ensure_chunks(c))) return rc;
        if (!ordered) {
            HIPCHK(c, launch_fill_f64(c->gridA, c->gp * kGridRows, 0.0, c->stream), "run_substeps");
            HIPCHK(c, particles(c, k, false, true, c->gridA, c->gridA, 0), "run_substeps");
        }
        for (int i = 0; i < n; ++i) {
            const int s = first_step + i;
            if (reset_bins_at_cycle && s % N == 0)
                HIPCHK(c, launch_fill_f64(c->bins, (size_t)c->nBinsAlloc, 0.0, c->stream), "run_substeps");
            tt += dt;
            const double ix = indenter_x_initial + tt * V;
            const int rec = ((s + 1) % N == 0) ? 1 : 0;
            double *h2 = c->hist + 2 * (size_t)i;
            const int timeIt = c->timing && (i % c->timingStride == 0);
            if (ordered) {
                HIPCHK(c, launch_fill_f64(c->gridA, c->gp * kGridRows, 0.0, c->stream), "run_substeps");
                k.pts = c->pts;
                HIPCHK(c, ordered_p2g(k, c->gridA, c->ord, c->stream), "run_substeps");
                HIPCHK(c, launch_update(k, c->gridA, ix, indenter_y, nullptr, nullptr, true,
                                        c->ord.nodeBin, c->ord.nodeF, c->stream), "run_substeps");
                HIPCHK(c, ordered_fold_forces(k, c->ord, h2, c->stream), "run_substeps");
                if ((rc = launch_particles_timed(c, k, true, false, c->gridA, nullptr, rec, timeIt)))
                    return rc;
            } else {
                const bool last = (i == n - 1);
                HIPCHK(c, launch_update(k, c->gridA, ix, indenter_y, h2, last ? nullptr : c->gridB,
                                        false, nullptr, nullptr, c->stream), "run_substeps");
                if (!last) {
                    if ((rc = launch_particles_timed(c, k, true, true, c->gridA, c->gridB, rec, timeIt)))
                        return rc;
                    std::swap(c->gridA, c->gridB);
                } else {
                    if ((rc = launch_particles_timed(c, k, true, false, c->gridA, nullptr, rec, timeIt)))
                        return rc;
                }
            }
        }
        *t = tt;
        if (c->timing) {
            if ((rc = collect_timing(c))) return rc;
        }
        return ICEMPM_OK;
    });
}

int icempm_force_history(icempm_ctx *c, double *out, int n) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || !out || n < 0) return ICEMPM_E_ARG;
        if (n > c->lastRunSubsteps) return fail(c, ICEMPM_E_ARG, "force_history: n exceeds the last run");
        if (n == 0) return ICEMPM_OK;
        HIPCHK(c, hipMemcpyAsync(out, c->hist, 2 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream), "force_history");
        HIPCHK(c, hipStreamSynchronize(c->stream), "force_history");
        return ICEMPM_OK;
    });
}

int icempm_transfer_from_device(icempm_ctx *c, icempm_callback cb, void *user) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c) return ICEMPM_E_ARG;
        if (!c->pts) return fail(c, ICEMPM_E_STATE, "cuda_transfer_from_device: arrays not allocated");
        HIPCHK(c, hipMemcpyAsync(c->hostPts, c->pts, c->pp * kPtsRows * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream), "cuda_transfer_from_device");
        HIPCHK(c, hipMemcpyAsync(c->hostForce, c->bins, (size_t)c->nBinsAlloc * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream), "cuda_transfer_from_device");
        HIPCHK(c, hipMemcpyAsync(c->hostErr, c->dErr, sizeof(int), hipMemcpyDeviceToHost, c->stream),
               "cuda_transfer_from_device");
        HIPCHK(c, hipEventRecord(c->evXfer, c->stream), "cuda_transfer_from_device");
        c->xferPending = true;
        ReadbackJob *j = new ReadbackJob{c, cb, user};
        hipError_t e = hipLaunchHostFunc(c->stream, readback_done, j);
        if (e != hipSuccess) {
            delete j;
            return hipfail(c, e, "cuda_transfer_from_device");
        }
        return ICEMPM_OK;
    });
}

int icempm_transfer_wait(icempm_ctx *c) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c) return ICEMPM_E_ARG;
        if (!c->xferPending) return ICEMPM_OK;
        // the copies only: the completion callback queued after them may still run
        HIPCHK(c, hipEventSynchronize(c->evXfer), "cuda_transfer_from_device");
        c->xferPending = false;
        c->errorCode.store(*c->hostErr);
        return ICEMPM_OK;
    });
}

int icempm_error_code(const icempm_ctx *c) { return c ? c->errorCode.load() : 0; }

int icempm_read_error_flag(icempm_ctx *c, int *flag) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || !flag) return ICEMPM_E_ARG;
        HIPCHK(c, hipMemcpyAsync(c->hostErr, c->dErr, sizeof(int), hipMemcpyDeviceToHost, c->stream),
               "read_error_flag");
        HIPCHK(c, hipStreamSynchronize(c->stream), "read_error_flag");
        *flag = *c->hostErr;
        return ICEMPM_OK;
    });
}

int icempm_event_record_start(icempm_ctx *c) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c) return ICEMPM_E_ARG;
        HIPCHK(c, hipEventRecord(c->evStart, c->stream), "event record");
        return ICEMPM_OK;
    });
}

int icempm_event_record_stop(icempm_ctx *c) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c) return ICEMPM_E_ARG;
        HIPCHK(c, hipEventRecord(c->evStop, c->stream), "event record");
        return ICEMPM_OK;
    });
}

int icempm_event_elapsed_ms(icempm_ctx *c, float *ms) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || !ms) return ICEMPM_E_ARG;
        HIPCHK(c, hipEventSynchronize(c->evStop), "event elapsed");
        HIPCHK(c, hipEventElapsedTime(ms, c->evStart, c->evStop), "event elapsed");
        return ICEMPM_OK;
    });
}

int icempm_sort_points(icempm_ctx *c, int *perm_out, int *keys_out) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        int rc = check_ready(c, "sort_points");
        if (rc) return rc;
        const int n = c->prms.nPts;
        if ((rc = ensure_sort_scratch(c, (size_t)n))) return rc;
        if (!c->ptsTmp)
            HIPCHK(c, hipMalloc(&c->ptsTmp, c->pp * kPtsRows * sizeof(double)), "sort_points");
        HIPCHK(c, launch_point_keys(c->pts, c->pp, n, c->prms.cellsize_inv, c->prms.GridX, c->keys,
                                    c->vals, c->stream), "sort_points");
        HIPCHK(c, radix_sort_pairs(c->keys, c->vals, c->keys2, c->vals2, (size_t)n, c->sortTmp,
                                   c->stream), "sort_points");
        HIPCHK(c, launch_permute_points(c->pts, c->ptsTmp, c->pp, n, c->vals, c->stream), "sort_points");
        std::swap(c->pts, c->ptsTmp);
        if (perm_out && n > 0)
            HIPCHK(c, hipMemcpyAsync(perm_out, c->vals, (size_t)n * sizeof(int), hipMemcpyDeviceToHost,
                                     c->stream), "sort_points");
        if (keys_out && n > 0)
            HIPCHK(c, hipMemcpyAsync(keys_out, c->keys, (size_t)n * sizeof(int), hipMemcpyDeviceToHost,
                                     c->stream), "sort_points");
        HIPCHK(c, hipStreamSynchronize(c->stream), "sort_points");
        c->chunksDirty = true;
        return ICEMPM_OK;
    });
}

int icempm_sort_raw_points(icempm_ctx *c, float *xy, int16_t *grains, int n, double hinv,
                           int gridX) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || n < 0 || (n > 0 && (!xy || !grains))) return ICEMPM_E_ARG;
        if (n <= 1) return ICEMPM_OK;
        int rc = ensure_sort_scratch(c, (size_t)n);
        if (rc) return rc;
        float *dxy = nullptr, *dxy2 = nullptr;
        int16_t *dg = nullptr, *dg2 = nullptr;
        hipError_t e = hipMalloc(&dxy, 2 * (size_t)n * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&dxy2, 2 * (size_t)n * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&dg, (size_t)n * sizeof(int16_t));
        if (e == hipSuccess) e = hipMalloc(&dg2, (size_t)n * sizeof(int16_t));
        if (e == hipSuccess)
            e = hipMemcpyAsync(dxy, xy, 2 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dg, grains, (size_t)n * sizeof(int16_t), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = launch_raw_keys(dxy, n, hinv, gridX, c->keys, c->vals, c->stream);
        if (e == hipSuccess)
            e = radix_sort_pairs(c->keys, c->vals, c->keys2, c->vals2, (size_t)n, c->sortTmp, c->stream);
        if (e == hipSuccess) e = launch_permute_raw(dxy, dg, dxy2, dg2, n, c->vals, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(xy, dxy2, 2 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(grains, dg2, (size_t)n * sizeof(int16_t), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->stream);
        dfree(dxy); dfree(dxy2); dfree(dg); dfree(dg2);
        if (e != hipSuccess) return hipfail(c, e, "sort_raw_points");
        return ICEMPM_OK;
    });
}

int icempm_sort_raw_points_noalloc(icempm_ctx *c, float *xy, int16_t *grains, int n, double hinv,
                                   int gridX) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || n < 0 || (n > 0 && (!xy || !grains))) return ICEMPM_E_ARG;
        if ((size_t)n > kNoAllocMaxPoints)
            return fail(c, ICEMPM_E_ARG, "sort_raw_points_noalloc: n = " + std::to_string(n) +
                                             " exceeds ICEMPM_NOALLOC_MAX_POINTS = " + std::to_string(kNoAllocMaxPoints));
        if (n <= 1) return ICEMPM_OK;
        if (radix_tmp_ints((size_t)n) > kRadixTmpInts)
            return fail(c, ICEMPM_E_STATE, "sort_raw_points_noalloc: radix scratch bound too small");
        unsigned char *base = scratch_of(c->device);
        if (!base) return fail(c, ICEMPM_E_STATE, "sort_raw_points_noalloc: scratch not set up on this device");
        unsigned char *q = base;
        auto take = [&q](size_t bytes) {
            unsigned char *r = q;
            q += al256(bytes);
            return r;
        };
        float *dxy = reinterpret_cast<float *>(take(kNoAllocMaxPoints * 8));
        float *dxy2 = reinterpret_cast<float *>(take(kNoAllocMaxPoints * 8));
        int16_t *dg = reinterpret_cast<int16_t *>(take(kNoAllocMaxPoints * 2));
        int16_t *dg2 = reinterpret_cast<int16_t *>(take(kNoAllocMaxPoints * 2));
        int *keys = reinterpret_cast<int *>(take(kNoAllocMaxPoints * 4));
        int *vals = reinterpret_cast<int *>(take(kNoAllocMaxPoints * 4));
        int *keys2 = reinterpret_cast<int *>(take(kNoAllocMaxPoints * 4));
        int *vals2 = reinterpret_cast<int *>(take(kNoAllocMaxPoints * 4));
        int *tmp = reinterpret_cast<int *>(take(kRadixTmpInts * 4));
        if ((size_t)(q - base) > kScratchBytes)
            return fail(c, ICEMPM_E_STATE, "sort_raw_points_noalloc: scratch layout exceeds the array");
        // every piece is written before it is read (keys and values by k_raw_keys,
        // histograms and scan sums by their kernels each pass): nothing to clear
        hipError_t e = hipMemcpyAsync(dxy, xy, 2 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dg, grains, (size_t)n * sizeof(int16_t), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = launch_raw_keys(dxy, n, hinv, gridX, keys, vals, c->stream);
        if (e == hipSuccess) e = radix_sort_pairs(keys, vals, keys2, vals2, (size_t)n, tmp, c->stream);
        if (e == hipSuccess) e = launch_permute_raw(dxy, dg, dxy2, dg2, n, vals, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(xy, dxy2, 2 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(grains, dg2, (size_t)n * sizeof(int16_t), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(c->stream);
            return hipfail(c, e, "sort_raw_points_noalloc");
        }
        return ICEMPM_OK;
    });
}

int icempm_set_mode(icempm_ctx *c, int mode) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || (mode != ICEMPM_MODE_FAST && mode != ICEMPM_MODE_ORDERED)) return ICEMPM_E_ARG;
        c->mode = mode;
        return ICEMPM_OK;
    });
}

int icempm_get_mode(const icempm_ctx *c) { return c ? c->mode : -1; }

int icempm_set_tuning(icempm_ctx *c, int knob, int value) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c) return ICEMPM_E_ARG;
        switch (knob) {
        case ICEMPM_TUNE_PASS_KIND:
            if (value < 0 || value > 2) return fail(c, ICEMPM_E_ARG, "set_tuning: pass kind is 0, 1 or 2");
            if ((value == 0) != (c->passKind == 0)) c->chunksDirty = true;   // the chunk rule differs
            c->passKind = value;
            return ICEMPM_OK;
        default:
            return fail(c, ICEMPM_E_ARG, "set_tuning: unknown knob");
        }
    });
}

int icempm_get_tuning(const icempm_ctx *c, int knob) {
    if (!c) return -1;
    switch (knob) {
    case ICEMPM_TUNE_PASS_KIND: return c->passKind;
    default: return -1;
    }
}

int icempm_kernel_timing(icempm_ctx *c, int enable) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || enable < 0) return ICEMPM_E_ARG;
        c->timing = enable > 0;
        c->timingStride = enable > 0 ? enable : 1;
        c->timedMs = 0;
        c->timedLaunches = 0;
        return ICEMPM_OK;
    });
}

int icempm_kernel_timing_read(icempm_ctx *c, double *total_ms, int *launches) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || !total_ms || !launches) return ICEMPM_E_ARG;
        *total_ms = c->timedMs;
        *launches = c->timedLaunches;
        return ICEMPM_OK;
    });
}

int icempm_measure_copy(icempm_ctx *c, size_t nbytes, int reps, double *gbps) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || !gbps || reps <= 0 || nbytes < 1024) return ICEMPM_E_ARG;
        const size_t n = nbytes / sizeof(double) / 2 * 2;
        double *a = nullptr, *b = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipMalloc(&a, n * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&b, n * sizeof(double));
        if (e == hipSuccess) e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e == hipSuccess) e = launch_fill_f64(a, n, 1.0, c->stream);
        if (e == hipSuccess) e = launch_copy(a, b, n, c->stream);   // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy(a, b, n, c->stream);
        if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
        float ms = 0;
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        (void)hipStreamSynchronize(c->stream);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        dfree(a);
        dfree(b);
        if (e != hipSuccess) return hipfail(c, e, "measure_copy");
        // bytes read + bytes written per copy
        *gbps = 2.0 * (double)n * sizeof(double) * reps / (ms * 1e-3) / 1e9;
        return ICEMPM_OK;
    });
}

int icempm_read_grid(icempm_ctx *c, double *out) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || !out) return ICEMPM_E_ARG;
        if (!c->gridA) return fail(c, ICEMPM_E_STATE, "read_grid: arrays not allocated");
        HIPCHK(c, hipMemcpyAsync(out, c->gridA, c->gp * kGridRows * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream), "read_grid");
        HIPCHK(c, hipStreamSynchronize(c->stream), "read_grid");
        return ICEMPM_OK;
    });
}

int icempm_write_grid(icempm_ctx *c, const double *in) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || !in) return ICEMPM_E_ARG;
        if (!c->gridA) return fail(c, ICEMPM_E_STATE, "write_grid: arrays not allocated");
        HIPCHK(c, hipMemcpyAsync(c->gridA, in, c->gp * kGridRows * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream), "write_grid");
        HIPCHK(c, hipStreamSynchronize(c->stream), "write_grid");
        return ICEMPM_OK;
    });
}

int icempm_read_points(icempm_ctx *c, double *out) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || !out) return ICEMPM_E_ARG;
        if (!c->pts) return fail(c, ICEMPM_E_STATE, "read_points: arrays not allocated");
        HIPCHK(c, hipMemcpyAsync(out, c->pts, c->pp * kPtsRows * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream), "read_points");
        HIPCHK(c, hipStreamSynchronize(c->stream), "read_points");
        return ICEMPM_OK;
    });
}

int icempm_read_bins(icempm_ctx *c, double *out) {
    return host_guard(c ? &c->lastError : nullptr, [&]() -> int {
        if (!c || !out) return ICEMPM_E_ARG;
        if (!c->bins) return fail(c, ICEMPM_E_STATE, "read_bins: arrays not allocated");
        HIPCHK(c, hipMemcpyAsync(out, c->bins, (size_t)c->nBinsAlloc * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream), "read_bins");
        HIPCHK(c, hipStreamSynchronize(c->stream), "read_bins");
        return ICEMPM_OK;
    });
}

}  // extern "C"
