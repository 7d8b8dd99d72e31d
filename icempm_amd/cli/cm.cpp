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
#include <atomic>
#include <cstdlib>
#include <filesystem>
#include <iostream>
#include <mutex>
#include <string>
#include <thread>

#include "icempm.h"
#include "model.h"
#include "snapshotmanager.h"

int main(int argc, char **argv)
{
    std::string config, dir = "cm_snapshots";
    bool ordered = false, pq = false;
    long maxCycles = -1;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--snapshots" && i + 1 < argc) dir = argv[++i];
        else if (a == "--ordered") ordered = true;
        else if (a == "--pq") pq = true;
        else if (a == "--cycles" && i + 1 < argc) maxCycles = std::atol(argv[++i]);
        else if (!a.empty() && a[0] != '-' && config.empty()) config = a;
        else {
            std::cerr << "usage: cm <config.json> [--snapshots DIR] [--ordered] [--pq] [--cycles N]\n";
            return 2;
        }
    }
    if (config.empty()) {
        std::cerr << "usage: cm <config.json> [--snapshots DIR] [--ordered] [--pq] [--cycles N]\n";
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
