// gpu_implementation5.cpp -- GPU_Implementation5 over the C ABI.
// Method-by-method counterpart of simulation/gpu_implementation5.cu:434-593.
#include "gpu_implementation5.h"

#include <iostream>
#include <stdexcept>
#include <string>

#include "model.h"

void GPU_Implementation5::check(int rc, const char *tag)
{
    if (rc == ICEMPM_OK) return;
    std::string msg = tag;
    const char *detail = icempm_last_error(ctx);
    if (detail && *detail) msg += std::string(" (") + detail + ")";
    throw std::runtime_error(msg);
}

GPU_Implementation5::~GPU_Implementation5()
{
    if (ctx) icempm_destroy(ctx);
    ctx = nullptr;
}

void GPU_Implementation5::test()
{
    check(icempm_test(ctx), "cuda test");
    std::cout << "hello kernel executed successfully\n";
}

void GPU_Implementation5::synchronize()
{
    if (!initialized) return;
    check(icempm_synchronize(ctx), "synchronize");
}

void GPU_Implementation5::initialize()
{
    if (initialized) return;
    int rc = icempm_create(&ctx);
    if (rc == ICEMPM_E_NODEV) throw std::runtime_error("No avaialble CUDA devices");
    if (rc != ICEMPM_OK) throw std::runtime_error("initialize() cuda error");
    initialized = true;
}

void GPU_Implementation5::cuda_update_constants()
{
    if (!initialized) initialize();
    check(icempm_update_constants(ctx, &model->prms, sizeof(icy::SimParams)),
          "cuda_update_constants: gprms");
}

void GPU_Implementation5::cuda_allocate_arrays(size_t nGridNodes, size_t nPoints)
{
    if (!initialized) initialize();
    icy::SimParams &P = model->prms;
    size_t gridPitch = 0, ptsPitch = 0;
    check(icempm_allocate_arrays(ctx, nGridNodes, nPoints, 2 * P.n_indenter_subdivisions,
                                 &gridPitch, &ptsPitch, &tmp_transfer_buffer,
                                 &host_side_indenter_force_accumulator),
          "cuda_allocate_arrays");
    P.nGridPitch = gridPitch;
    P.nPtsPitch = ptsPitch;
    // the device arrays are owned by the context; the POD keeps no device
    // addresses (it is dumped into snapshots)
    P.grid_array = nullptr;
    P.pts_array = nullptr;
    P.indenter_force_accumulator = nullptr;
    error_code = 0;
}

void GPU_Implementation5::transfer_ponts_to_device()
{
    check(icempm_transfer_points_to_device(ctx), "transfer_points_to_device");
}

void GPU_Implementation5::cuda_transfer_from_device()
{
    check(icempm_transfer_from_device(ctx, &GPU_Implementation5::callback_from_stream, this),
          "cuda_transfer_from_device");
    // the reference's flag copy into pageable memory returns once it has
    // landed (gpu_implementation5.cu:539-540): error_code is this cycle's
    check(icempm_transfer_wait(ctx), "cuda_transfer_from_device");
    error_code = icempm_error_code(ctx);
}

void GPU_Implementation5::callback_from_stream(void *userData)
{
    // runs on the runtime's callback thread: no library calls here
    GPU_Implementation5 *gpu = reinterpret_cast<GPU_Implementation5 *>(userData);
    if (gpu->transfer_completion_callback)
        gpu->transfer_completion_callback();
    else if (gpu->model)
        gpu->model->UnlockCycleMutex();   // no consumer: the cycle counts as processed
}

void GPU_Implementation5::cuda_reset_grid()
{
    check(icempm_reset_grid(ctx), "cuda_reset_grid error");
}

void GPU_Implementation5::cuda_reset_indenter_force_accumulator()
{
    check(icempm_reset_indenter_force_accumulator(ctx), "cuda_reset_grid error");
}

void GPU_Implementation5::cuda_p2g()
{
    check(icempm_p2g(ctx), "cuda_p2g");
}

void GPU_Implementation5::cuda_update_nodes(double indenter_x, double indenter_y)
{
    check(icempm_update_nodes(ctx, indenter_x, indenter_y), "cuda_update_nodes");
}

void GPU_Implementation5::cuda_g2p(bool recordPQ)
{
    check(icempm_g2p(ctx, recordPQ ? 1 : 0), "cuda_g2p");
}

void GPU_Implementation5::run_substeps(int first_step, int n, double &t,
                                       double indenter_x_initial, double indenter_y,
                                       bool reset_bins_at_cycle)
{
    check(icempm_run_substeps(ctx, first_step, n, &t, indenter_x_initial, indenter_y,
                              reset_bins_at_cycle ? 1 : 0),
          "cuda_substeps");
}

void GPU_Implementation5::record_event(const icempm_cycle_event &ev)
{
    check(ev.which == 0 ? icempm_event_record_start(ctx) : icempm_event_record_stop(ctx),
          "event record");
}

float GPU_Implementation5::elapsed_ms()
{
    float ms = 0;
    check(icempm_event_elapsed_ms(ctx, &ms), "event elapsed");
    return ms;
}

void GPU_Implementation5::sort_points_by_cell()
{
    check(icempm_sort_points(ctx, nullptr, nullptr), "sort_points");
}

void GPU_Implementation5::set_mode(int mode)
{
    if (!initialized) initialize();
    check(icempm_set_mode(ctx, mode), "set_mode");
}

int GPU_Implementation5::refresh_error_code()
{
    error_code = icempm_error_code(ctx);
    return error_code;
}
