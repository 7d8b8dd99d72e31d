// host_model.h -- the object behind the icempm_model handle of the host C
// API (icy::Model with its SnapshotManager), shared by the C entry points of
// libicempm_host.so and libicempm_h5.so.
#ifndef ICEMPM_HOST_MODEL_H
#define ICEMPM_HOST_MODEL_H

#include <exception>
#include <string>

#include "icempm_host.h"
#include "model.h"
#include "snapshotmanager.h"

struct icempm_model {
    icy::Model model;
    icy::SnapshotManager snapshot;
    std::string lastError;
};

// runs f, turning an exception into a status code and the model's last error
template <class F>
int icempm_guarded(icempm_model *m, F &&f) {
    try {
        f();
        return ICEMPM_OK;
    } catch (const std::exception &e) {
        if (m) m->lastError = e.what();
        return ICEMPM_E_HIP;
    }
}

#endif
