// model.cpp -- icy::Model::Step and friends (reference: simulation/model.cpp).
#include "model.h"

#include <iostream>

icy::Model::Model()
{
    prms.Reset();
    gpu.model = this;
    compute_time_per_cycle = 0;
    abortRequested = false;
}

// One output cycle: the substeps up to the next multiple of
// UpdateEveryNthStep (model.cpp:11-47).  The per-substep work
// (grid reset, P2G, node update with the indenter at
// indenter_x_initial + t*IndVelocity, G2P with recordPQ on the cycle's last
// substep, t advanced by repeated addition of the time step) runs as one
// device loop in GPU_Implementation5::run_substeps.
bool icy::Model::Step()
{
    const int N = prms.UpdateEveryNthStep;
    double simulation_time = prms.SimulationTime;
    const int count = N - (prms.SimulationStep % N);
    const bool timed = prms.SimulationStep % (N * 2) == 0;

    gpu.cuda_reset_indenter_force_accumulator();
    if (timed) gpu.record_event(gpu.eventCycleStart);
    gpu.run_substeps(prms.SimulationStep, count, simulation_time, prms.indenter_x_initial,
                     prms.indenter_y, false);
    prms.indenter_x = prms.indenter_x_initial + simulation_time * prms.IndVelocity;
    if (timed) gpu.record_event(gpu.eventCycleStop);

    processing_current_cycle_data.lock();   // previous cycle's data still being consumed?

    gpu.cuda_transfer_from_device();

    if (timed) cycleTimed = true;
    else if (cycleTimed)
    {
        compute_time_per_cycle = gpu.elapsed_ms() / N;
    }

    prms.SimulationTime = simulation_time;
    prms.SimulationStep += count;
    gpu.refresh_error_code();
    return (prms.SimulationTime < prms.SimulationEndTime && !gpu.error_code);
}

void icy::Model::UnlockCycleMutex()
{
    processing_current_cycle_data.unlock();
}

void icy::Model::Reset()
{
    prms.SimulationStep = 0;
    prms.SimulationTime = 0;
    compute_time_per_cycle = 0;
}

void icy::Model::Prepare()
{
    abortRequested = false;
    gpu.cuda_update_constants();
}
