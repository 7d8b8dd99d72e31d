// gpu_implementation5.h -- the device backend class of the drop-in surface.
//
// Same public members as the reference's GPU_Implementation5
// (simulation/gpu_implementation5.h:40-72), which icy::Model, the snapshot
// manager and the CLI/GUI frontends use.  Every method calls the C ABI of
// libicempm_hip.so (include/icempm.h); failures are rethrown as
// std::runtime_error with the reference's tag strings.  No HIP or CUDA
// header is needed by code that includes this file.
#ifndef GPU_IMPLEMENTATION5_H
#define GPU_IMPLEMENTATION5_H

#include <functional>

#include "icempm.h"
#include "parameters_sim.h"

namespace icy { class Model; }

// Stands in for the reference's cudaEvent_t members: a tag naming one of
// the two cycle events that the context owns.
struct icempm_cycle_event { int which; };

class GPU_Implementation5
{
public:
    icy::Model *model = nullptr;
    int error_code = 0;
    std::function<void()> transfer_completion_callback;

    void initialize();
    void test();
    void synchronize(); // call before terminating the main thread
    void cuda_update_constants();
    void cuda_allocate_arrays(size_t nGridNodes, size_t nPoints);
    void cuda_reset_grid();
    void transfer_ponts_to_device();
    void cuda_p2g();
    void cuda_g2p(bool recordPQ);
    void cuda_update_nodes(double indenter_x, double indenter_y);
    void cuda_reset_indenter_force_accumulator();

    void cuda_transfer_from_device();

    icempm_cycle_event eventCycleStart{0}, eventCycleStop{1};

    double *tmp_transfer_buffer = nullptr; // pinned host copy of the point SoA
    double *host_side_indenter_force_accumulator = nullptr;

    // --- additions (no counterpart in the reference) ---
    // the whole substep loop of Model::Step on the device
    void run_substeps(int first_step, int n, double &t, double indenter_x_initial,
                      double indenter_y, bool reset_bins_at_cycle);
    void record_event(const icempm_cycle_event &ev);
    float elapsed_ms();               // waits for eventCycleStop
    void sort_points_by_cell();       // stable device sort of the point SoA by cell key
    void set_mode(int mode);          // ICEMPM_MODE_FAST / ICEMPM_MODE_ORDERED
    int refresh_error_code();         // error_code <- last completed readback
    icempm_ctx *context() const { return ctx; }

    ~GPU_Implementation5();

private:
    icempm_ctx *ctx = nullptr;
    bool initialized = false;

    void check(int rc, const char *tag);
    static void callback_from_stream(void *userData);
};

#endif
