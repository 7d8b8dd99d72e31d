// drop_in_driver.cpp -- a frontend written against the drop-in surface the
// way the reference's CLI (cli/main.cpp:17-100) uses it: parse the JSON
// config, ingest raw points, install the readback callback, loop
// Model::Step on a simulation thread, consume each cycle's readback on
// another thread, then synchronize.  Raw points come from the seeded lattice
// generator instead of an HDF5 file.
//
// usage: drop_in_driver <config.json> <ncx> <ncy> <per_side> <h> <seed> <cycles>
//                       <mode 0=FAST 1=ORDERED> <out.bin>
// out.bin: the last readback (tmp_transfer_buffer, 16*nPtsPitch doubles),
// then the force bins of that cycle, then SimulationStep and SimulationTime.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "icempm_host.h"
#include "model.h"
#include "snapshotmanager.h"

int main(int argc, char **argv)
{
    if (argc != 10) {
        std::cerr << "usage: drop_in_driver config.json ncx ncy per_side h seed cycles mode out.bin\n";
        return 2;
    }
    icy::Model model;
    icy::SnapshotManager snapshot;
    snapshot.model = &model;
    try {
        model.gpu.initialize();
        model.Reset();
        model.gpu.set_mode(std::atoi(argv[8]));
        std::string raw = model.prms.ParseFile(argv[1]);
        std::cout << "config " << argv[1] << " raw points: " << raw << "\n";

        const int ncx = std::atoi(argv[2]), ncy = std::atoi(argv[3]), k = std::atoi(argv[4]);
        const double h = std::atof(argv[5]);
        const uint64_t seed = std::strtoull(argv[6], nullptr, 0);
        const int cycles = std::atoi(argv[7]);
        const long n = icempm_scene_lattice(ncx, ncy, k, h, 0.25, seed, nullptr, nullptr);
        std::vector<float> xy(2 * n);
        std::vector<int16_t> grains(n);
        icempm_scene_lattice(ncx, ncy, k, h, 0.25, seed, xy.data(), grains.data());
        snapshot.LoadRawPoints(xy.data(), grains.data(), (int)n, (float)(ncx * h * ncy * h));

        std::vector<double> last;
        std::vector<double> bins;
        std::atomic<int> consumed{0};
        std::thread consumer;
        std::mutex consumerLock;   // the callback thread and main both touch `consumer`
        model.gpu.transfer_completion_callback = [&]() {
            std::lock_guard<std::mutex> g(consumerLock);
            if (consumer.joinable()) consumer.join();
            consumer = std::thread([&]() {
                const icy::SimParams &P = model.prms;
                last.assign(model.gpu.tmp_transfer_buffer,
                            model.gpu.tmp_transfer_buffer + P.nPtsPitch * icy::SimParams::nPtsArrays);
                bins.assign(model.gpu.host_side_indenter_force_accumulator,
                            model.gpu.host_side_indenter_force_accumulator + 2 * P.n_indenter_subdivisions);
                consumed++;
                model.UnlockCycleMutex();
            });
        };

        std::thread sim([&]() {
            for (int c = 0; c < cycles; ++c)
                if (!model.Step()) break;
        });
        sim.join();
        model.gpu.synchronize();
        while (consumed.load() < cycles) std::this_thread::yield();
        {
            std::lock_guard<std::mutex> g(consumerLock);
            if (consumer.joinable()) consumer.join();
        }

        std::ofstream out(argv[9], std::ios::binary);
        out.write(reinterpret_cast<const char *>(last.data()), last.size() * sizeof(double));
        out.write(reinterpret_cast<const char *>(bins.data()), bins.size() * sizeof(double));
        const double tail[2] = {(double)model.prms.SimulationStep, model.prms.SimulationTime};
        out.write(reinterpret_cast<const char *>(tail), sizeof(tail));
        std::cout << "cycles " << consumed.load() << " step " << model.prms.SimulationStep
                  << " time " << model.prms.SimulationTime << " cycle time "
                  << model.compute_time_per_cycle << " ms/substep\n";
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
