// cm.cpp -- the command-line frontend on this backend (reference:
// cli/main.cpp:17-100): parse a JSON configuration, load its raw point file
// (HDF5), step icy::Model on a simulation thread, and write a snapshot
// (<dir>/snapshots/dNNNNN.h5, plus <dir>/pq/pqNNNNN.h5 with --pq) from each
// output cycle's readback on a snapshot thread.  Console input as in the
// reference: 's' requests a snapshot (every cycle is saved anyway), 'q' stops
// after the current cycle.  The run also ends when Model::Step reports the
// end time or an out-of-grid point.
//
// usage: cm <config.json> [--snapshots DIR] [--ordered] [--pq] [--cycles N]
//        cm --batch <manifest.json> [--out DIR] [--f32 | --ordered] [--cycles N]
//
// The batch mode (no counterpart in the reference) runs many configurations
// as one batch on the batched engine: the manifest is a JSON object whose
// "cases" array lists configuration files (paths relative to the manifest),
// with optional "precision" ("f64", "f32" or "ordered") and "output"
// (directory).
// Each configuration's raw point file is ingested as the single-case path
// does (LoadRawPoints), grains included; all cases share one time step.
// Output cycles are UpdateEveryNthStep substeps of the first case; the run
// ends after --cycles cycles or when the first case reaches its
// SimulationEndTime.  Each cycle writes every case's snapshot in the
// single-case format, DIR/case_K/snapshots/dNNNNN.h5 (Indenter_2D: the
// cycle's force bins; Params with SimulationStep, SimulationTime and the
// indenter position; Points in the ingest order), appends the case's
// per-substep total indenter force to DIR/case_K.csv (substep, time, fx, fy)
// and one line to DIR/batch.csv.  With --ordered (or "precision": "ordered")
// the cases run one after another on the single-case engine in ORDERED mode
// instead (the batched engine sums with hardware atomics and has no such
// mode): the same files, each snapshot bit for bit the one `cm --ordered`
// writes for that configuration; every case then needs the first case's
// time step and UpdateEveryNthStep.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "icempm.h"
#include "icempm_h5.h"
#include "model.h"
#include "snapshotmanager.h"

namespace {

const char *kUsage =
    "usage: cm <config.json> [--snapshots DIR] [--ordered] [--pq] [--cycles N]\n"
    "       cm --batch <manifest.json> [--out DIR] [--f32 | --ordered] [--cycles N]\n";

// The "cases" string array and the "precision" / "output" strings of a
// manifest (a JSON object; other keys are ignored).
struct Manifest {
    std::vector<std::string> cases;
    std::string precision = "f64", output;
};

bool read_string(const std::string &s, size_t &i, std::string &out) {
    if (i >= s.size() || s[i] != '"') return false;
    out.clear();
    for (++i; i < s.size() && s[i] != '"'; ++i) {
        if (s[i] == '\\' && i + 1 < s.size()) ++i;
        out.push_back(s[i]);
    }
    if (i >= s.size()) return false;
    ++i;
    return true;
}

void skip_ws(const std::string &s, size_t &i) {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
}

Manifest parse_manifest(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("manifest " + path + " not found");
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    Manifest m;
    size_t i = 0;
    skip_ws(s, i);
    if (i >= s.size() || s[i] != '{') throw std::runtime_error("manifest is not a JSON object");
    ++i;
    while (true) {
        skip_ws(s, i);
        if (i < s.size() && s[i] == '}') break;
        std::string key;
        if (!read_string(s, i, key)) throw std::runtime_error("manifest: expected a key");
        skip_ws(s, i);
        if (i >= s.size() || s[i] != ':') throw std::runtime_error("manifest: expected ':'");
        ++i;
        skip_ws(s, i);
        if (key == "cases") {
            if (i >= s.size() || s[i] != '[') throw std::runtime_error("manifest: \"cases\" must be an array");
            ++i;
            while (true) {
                skip_ws(s, i);
                if (i < s.size() && s[i] == ']') {
                    ++i;
                    break;
                }
                std::string c;
                if (!read_string(s, i, c)) throw std::runtime_error("manifest: a case must be a string");
                m.cases.push_back(c);
                skip_ws(s, i);
                if (i < s.size() && s[i] == ',') ++i;
            }
        } else if (i < s.size() && s[i] == '"') {
            std::string v;
            if (!read_string(s, i, v)) throw std::runtime_error("manifest: bad string");
            if (key == "precision") m.precision = v;
            if (key == "output") m.output = v;
        } else {   // a number or literal: skip to the next ',' or '}'
            while (i < s.size() && s[i] != ',' && s[i] != '}') ++i;
        }
        skip_ws(s, i);
        if (i < s.size() && s[i] == ',') {
            ++i;
            continue;
        }
        if (i < s.size() && s[i] == '}') break;
        throw std::runtime_error("manifest: expected ',' or '}'");
    }
    if (m.cases.empty()) throw std::runtime_error("manifest: no cases");
    if (m.precision != "f64" && m.precision != "f32" && m.precision != "ordered")
        throw std::runtime_error("manifest: precision must be \"f64\", \"f32\" or \"ordered\"");
    return m;
}

void check(icempm_batch *b, int rc, const char *what) {
    if (rc != 0)
        throw std::runtime_error(std::string(what) + ": " + (b ? icempm_batch_last_error(b) : "") +
                                 " (code " + std::to_string(rc) + ")");
}

std::filesystem::path case_config(const std::filesystem::path &base, const std::string &c) {
    return std::filesystem::path(c).is_absolute() ? std::filesystem::path(c) : base / c;
}

// The manifest's cases on the single-case engine in ORDERED mode, one model
// per case, stepped case after case each cycle by Model::Step.
int run_batch_ordered(const Manifest &man, const std::filesystem::path &base, const std::string &out,
                      long maxCycles) {
    std::vector<std::unique_ptr<icy::Model>> models;
    std::vector<std::unique_ptr<icy::SnapshotManager>> snaps;
    long total = 0;
    for (const std::string &c : man.cases) {
        const std::filesystem::path cfg = case_config(base, c);
        models.push_back(std::make_unique<icy::Model>());
        snaps.push_back(std::make_unique<icy::SnapshotManager>());
        icy::Model &model = *models.back();
        model.gpu.initialize();
        model.Reset();
        model.gpu.set_mode(ICEMPM_MODE_ORDERED);
        snaps.back()->model = &model;
        const std::string raw = model.prms.ParseFile(cfg.string());
        snaps.back()->LoadRawPoints(raw);
        const icy::SimParams &P0 = models[0]->prms;
        if (model.prms.InitialTimeStep != P0.InitialTimeStep || model.prms.UpdateEveryNthStep != P0.UpdateEveryNthStep)
            throw std::runtime_error(cfg.string() + ": time step or UpdateEveryNthStep differs from the first case's");
        total += model.prms.nPts;
        std::cerr << "cm: case " << models.size() - 1 << " " << cfg.string() << ": " << model.prms.nPts
                  << " points, grid " << model.prms.GridX << " x " << model.prms.GridY << "\n";
    }
    const int N = models[0]->prms.UpdateEveryNthStep;
    if (N <= 0) throw std::runtime_error("UpdateEveryNthStep must be positive");
    long cycles = maxCycles;
    if (cycles < 0) cycles = (long)std::ceil(models[0]->prms.SimulationEndTime / (models[0]->prms.InitialTimeStep * N));
    std::filesystem::create_directories(out);
    std::vector<std::ofstream> csv;
    for (size_t c = 0; c < models.size(); ++c) {
        csv.emplace_back(std::filesystem::path(out) / ("case_" + std::to_string(c) + ".csv"));
        csv.back() << "substep,time,fx,fy\n";
        csv.back().precision(17);
    }
    std::ofstream summary(std::filesystem::path(out) / "batch.csv");
    summary << "cycle,substeps,ms,particle_updates_per_s\n";
    std::cerr << "cm: " << models.size() << " cases, " << total << " points, ORDERED single-case engine, " << cycles
              << " cycles of " << N << " substeps\n";
    std::vector<double> hist(2 * (size_t)N);
    for (long cy = 1; cy <= cycles; ++cy) {
        const auto t0 = std::chrono::steady_clock::now();
        bool bad = false;
        for (size_t c = 0; c < models.size(); ++c) {
            icy::Model &model = *models[c];
            const double tStart = model.prms.SimulationTime;
            const int step0 = model.prms.SimulationStep;
            model.Step();
            // no transfer callback: the readback's completion releases the cycle
            model.processing_current_cycle_data.lock();
            model.processing_current_cycle_data.unlock();
            if (icempm_force_history(model.gpu.context(), hist.data(), N) != 0)
                throw std::runtime_error("force_history of case " + std::to_string(c) + " failed");
            double t = tStart;
            for (int i = 0; i < N; ++i) {
                t += model.prms.InitialTimeStep;
                csv[c] << step0 + i + 1 << "," << t << "," << hist[2 * i] << "," << hist[2 * i + 1] << "\n";
            }
            snaps[c]->SaveSnapshot((std::filesystem::path(out) / ("case_" + std::to_string(c))).string());
            bad = bad || model.gpu.error_code != 0;
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        summary << cy << "," << N << "," << ms << "," << (ms > 0 ? total * (double)N / (ms * 1e-3) : 0.0) << "\n";
        std::cerr << "cm: cycle " << cy << " of " << cycles << ", " << ms / N << " ms per substep (host clock)\n";
        if (bad) {
            std::cerr << "cm: a point left the grid\n";
            return 3;
        }
    }
    return 0;
}

int run_batch(const std::string &manifestPath, std::string out, bool f32, bool ordered, long maxCycles) {
    const Manifest man = parse_manifest(manifestPath);
    if (out.empty()) out = man.output.empty() ? "cm_batch" : man.output;
    if (ordered || man.precision == "ordered") {
        if (f32) throw std::runtime_error("--ordered runs FP64 storage; --f32 does not apply");
        return run_batch_ordered(man, std::filesystem::path(manifestPath).parent_path(), out, maxCycles);
    }
    const bool useF32 = f32 || man.precision == "f32";
    const std::filesystem::path base = std::filesystem::path(manifestPath).parent_path();
    icempm_batch *b = nullptr;
    if (icempm_batch_create(&b, useF32 ? ICEMPM_BATCH_F32 : ICEMPM_BATCH_F64) != 0 || !b)
        throw std::runtime_error("batch_create failed");
    struct Guard {
        icempm_batch *b;
        ~Guard() { icempm_batch_destroy(b); }
    } guard{b};
    std::vector<icy::SimParams> prms;
    long total = 0;
    for (const std::string &c : man.cases) {
        const std::filesystem::path cfg = case_config(base, c);
        // the single-case ingest: parse, load and sort the raw points on the device, read them back
        icy::Model model;
        icy::SnapshotManager snapshot;
        model.gpu.initialize();
        model.Reset();
        snapshot.model = &model;
        const std::string raw = model.prms.ParseFile(cfg.string());
        snapshot.LoadRawPoints(raw);
        std::vector<double> soa(16 * model.prms.nPtsPitch);
        if (icempm_read_points(model.gpu.context(), soa.data()) != 0)
            throw std::runtime_error("reading the ingested points of " + cfg.string() + " failed");
        const int k = icempm_batch_add_case(b, &model.prms, sizeof(icy::SimParams), soa.data(), model.prms.nPtsPitch);
        if (k < 0) check(b, -k, ("add_case " + cfg.string()).c_str());
        prms.push_back(model.prms);
        total += model.prms.nPts;
        std::cerr << "cm: case " << k << " " << cfg.string() << ": " << model.prms.nPts << " points, grid "
                  << model.prms.GridX << " x " << model.prms.GridY << "\n";
    }
    check(b, icempm_batch_finalize(b), "finalize");
    const int N = prms[0].UpdateEveryNthStep;
    const double dt = prms[0].InitialTimeStep;
    if (N <= 0) throw std::runtime_error("UpdateEveryNthStep must be positive");
    long cycles = maxCycles;
    if (cycles < 0) cycles = (long)std::ceil(prms[0].SimulationEndTime / (dt * N));
    std::filesystem::create_directories(out);
    std::vector<std::ofstream> csv;
    for (size_t c = 0; c < prms.size(); ++c) {
        csv.emplace_back(std::filesystem::path(out) / ("case_" + std::to_string(c) + ".csv"));
        csv.back() << "substep,time,fx,fy\n";
        csv.back().precision(17);
    }
    std::ofstream summary(std::filesystem::path(out) / "batch.csv");
    summary << "cycle,substeps,ms,particle_updates_per_s\n";
    std::cerr << "cm: batch of " << prms.size() << " cases, " << total << " points, " << (useF32 ? "f32" : "f64")
              << " storage, " << cycles << " cycles of " << N << " substeps\n";
    std::vector<double> hist(2 * (size_t)N), bins, soa;
    double t = 0;   // by repeated addition, as Model::Step does
    long step = 0;
    for (long cy = 1; cy <= cycles; ++cy) {
        check(b, icempm_batch_run(b, N, 1), "run");
        float ms = 0;
        check(b, icempm_batch_last_ms(b, &ms), "last_ms");
        int flag = 0;
        check(b, icempm_batch_error_flag(b, &flag), "error_flag");
        std::vector<double> times(N);
        for (int i = 0; i < N; ++i) {
            t += dt;
            times[i] = t;
        }
        for (size_t c = 0; c < prms.size(); ++c) {
            check(b, icempm_batch_force_history(b, (int)c, hist.data(), N), "force_history");
            for (int i = 0; i < N; ++i)
                csv[c] << step + i + 1 << "," << times[i] << "," << hist[2 * i] << "," << hist[2 * i + 1] << "\n";
            icy::SimParams &P = prms[c];
            P.SimulationStep = (int)(step + N);
            P.SimulationTime = t;
            P.indenter_x = P.indenter_x_initial + t * P.IndVelocity;
            bins.assign(2 * (size_t)P.n_indenter_subdivisions, 0.0);
            soa.assign(icy::SimParams::nPtsArrays * (size_t)P.nPtsPitch, 0.0);
            check(b, icempm_batch_force_bins(b, (int)c, bins.data(), P.n_indenter_subdivisions), "force_bins");
            check(b, icempm_batch_read_points(b, (int)c, soa.data(), P.nPtsPitch), "read_points");
            const std::string dir = (std::filesystem::path(out) / ("case_" + std::to_string(c))).string();
            if (icempm_h5_write_snapshot(dir.c_str(), &P, sizeof(P), bins.data(), soa.data()) != 0)
                throw std::runtime_error("writing the snapshot of case " + std::to_string(c) + " failed");
        }
        step += N;
        summary << cy << "," << N << "," << ms << "," << (ms > 0 ? total * (double)N / (ms * 1e-3) : 0.0) << "\n";
        std::cerr << "cm: cycle " << cy << " of " << cycles << ", " << ms / N << " ms per substep\n";
        if (flag) {
            std::cerr << "cm: a point left the grid\n";
            return 3;
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    std::string config, dir = "cm_snapshots", manifest, out;
    bool ordered = false, pq = false, f32 = false;
    long maxCycles = -1;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--snapshots" && i + 1 < argc) dir = argv[++i];
        else if (a == "--ordered") ordered = true;
        else if (a == "--pq") pq = true;
        else if (a == "--cycles" && i + 1 < argc) maxCycles = std::atol(argv[++i]);
        else if (a == "--batch" && i + 1 < argc) manifest = argv[++i];
        else if (a == "--out" && i + 1 < argc) out = argv[++i];
        else if (a == "--f32") f32 = true;
        else if (!a.empty() && a[0] != '-' && config.empty()) config = a;
        else {
            std::cerr << kUsage;
            return 2;
        }
    }
    if (!manifest.empty()) {
        if (!config.empty() || pq) {
            std::cerr << kUsage;
            return 2;
        }
        try {
            const int rc = run_batch(manifest, out, f32, ordered, maxCycles);
            if (rc) return rc;
        } catch (const std::exception &e) {
            std::cerr << "cm: " << e.what() << "\n";
            return 1;
        }
        std::cout << "cm done\n";
        return 0;
    }
    if (config.empty()) {
        std::cerr << kUsage;
        return 2;
    }

    icy::Model model;
    icy::SnapshotManager snapshot;
    std::atomic<bool> request_terminate{false};
    std::thread snapshot_thread;
    std::mutex snapshot_lock;   // the callback thread and main both touch snapshot_thread
    try {
        model.gpu.initialize();
        model.Reset();
        snapshot.model = &model;
        if (ordered) model.gpu.set_mode(ICEMPM_MODE_ORDERED);
        const std::string pointCloudFile = model.prms.ParseFile(config);
        snapshot.LoadRawPoints(pointCloudFile);
        std::cerr << "cm: " << model.prms.nPts << " points, grid " << model.prms.GridX << " x "
                  << model.prms.GridY << ", " << model.prms.UpdateEveryNthStep << " substeps per cycle\n";

        // what to do once a cycle's data is available (cli/main.cpp:52-63)
        model.gpu.transfer_completion_callback = [&]() {
            std::lock_guard<std::mutex> g(snapshot_lock);
            if (snapshot_thread.joinable()) snapshot_thread.join();
            snapshot_thread = std::thread([&]() {
                const int frame = model.prms.AnimationFrameNumber();
                if (request_terminate) {
                    model.UnlockCycleMutex();
                    std::cerr << "cm: snapshot aborted\n";
                    return;
                }
                try {
                    snapshot.SaveSnapshot(dir);
                    if (pq) snapshot.SavePQ(dir);
                } catch (const std::exception &e) {
                    std::cerr << "cm: snapshot " << frame << " failed: " << e.what() << "\n";
                    request_terminate = true;
                }
                model.UnlockCycleMutex();
                std::cerr << "cm: cycle " << frame << " saved; compute time per substep "
                          << model.compute_time_per_cycle << " ms\n";
            });
        };
        std::filesystem::create_directories(dir);

        std::thread simulation_thread([&]() {
            bool result;
            long cycles = 0;
            do {
                result = model.Step();
                ++cycles;
            } while (!request_terminate && result && (maxCycles < 0 || cycles < maxCycles));
            request_terminate = request_terminate || !result;
        });
        // console input; detached, since stdin may stay open after the run ends
        std::thread([&]() {
            std::string in;
            while (std::cin >> in) {
                if (in[0] == 's') std::cerr << "cm: every cycle is saved\n";
                else if (in[0] == 'q') {
                    request_terminate = true;
                    std::cerr << "cm: stopping after the current cycle\n";
                    break;
                }
            }
        }).detach();

        simulation_thread.join();
        model.gpu.synchronize();
        // the last cycle's snapshot thread, started by the readback callback
        model.processing_current_cycle_data.lock();
        model.processing_current_cycle_data.unlock();
        {
            std::lock_guard<std::mutex> g(snapshot_lock);
            if (snapshot_thread.joinable()) snapshot_thread.join();
        }
        if (model.gpu.error_code) std::cerr << "cm: a point left the grid (error code " << model.gpu.error_code << ")\n";
    } catch (const std::exception &e) {
        std::cerr << "cm: " << e.what() << "\n";
        return 1;
    }
    std::cout << "cm done\n";
    return 0;
}
