// model.h -- icy::Model, the host loop of the substep (reference:
// simulation/model.h:26-52, simulation/model.cpp:11-72).
#ifndef MODEL_H
#define MODEL_H

#include <condition_variable>
#include <mutex>

#include "gpu_implementation5.h"
#include "parameters_sim.h"

namespace icy { class Model; }

// The cycle handshake: Step() takes it before queueing the readback and the
// consumer of the readback (another thread) releases it.  The reference
// uses a std::mutex for this, which may not be unlocked by a thread that
// does not own it; this type keeps the lock()/unlock() interface and allows
// the release from any thread.
class CycleHandshake
{
public:
    void lock()
    {
        std::unique_lock<std::mutex> g(m);
        cv.wait(g, [this] { return !busy; });
        busy = true;
    }
    void unlock()
    {
        {
            std::lock_guard<std::mutex> g(m);
            busy = false;
        }
        cv.notify_all();
    }
    bool try_lock()
    {
        std::lock_guard<std::mutex> g(m);
        if (busy) return false;
        busy = true;
        return true;
    }

private:
    std::mutex m;
    std::condition_variable cv;
    bool busy = false;
};

class icy::Model
{
public:
    Model();
    void Reset();

    void Prepare();        // once, before the first Step
    bool Step();           // one output cycle of UpdateEveryNthStep substeps
    void RequestAbort() {abortRequested = true;}

    void UnlockCycleMutex();

    icy::SimParams prms;
    GPU_Implementation5 gpu;
    float compute_time_per_cycle;

    CycleHandshake processing_current_cycle_data;

private:
    bool abortRequested;
    bool cycleTimed = false;   // a start/stop event pair has been recorded
};

#endif
