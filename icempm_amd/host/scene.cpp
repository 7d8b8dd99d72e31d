// scene.cpp -- seeded synthetic ice blocks in the raw-point format that
// SnapshotManager::LoadRawPoints ingests (f32 x/y pairs + i16 grain ids).
#include <cstdint>

#include "icempm_host.h"

namespace {

// splitmix64 (Steele, Lea, Flood 2014): a 64-bit counter-based generator.
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0, 1) from the top 53 bits
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

extern "C" long icempm_scene_lattice(int ncx, int ncy, int k, double h, double j, uint64_t seed,
                                     float *xy, int16_t *grains) {
    if (ncx <= 0 || ncy <= 0 || k <= 0) return 0;
    SplitMix64 rng{seed};
    long n = 0;
    for (int cy = 0; cy < ncy; ++cy)
        for (int cx = 0; cx < ncx; ++cx)
            for (int b = 0; b < k; ++b)
                for (int a = 0; a < k; ++a) {
                    const double ux = rng.uniform();
                    const double uy = rng.uniform();
                    const double x = (cx + (a + 0.5) / k) * h + (2.0 * ux - 1.0) * j * h;
                    const double y = (cy + (b + 0.5) / k) * h + (2.0 * uy - 1.0) * j * h;
                    if (xy) {
                        xy[2 * n] = (float)x;
                        xy[2 * n + 1] = (float)y;
                    }
                    if (grains) grains[n] = 0;
                    ++n;
                }
    return n;
}
