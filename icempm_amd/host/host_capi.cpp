// host_capi.cpp -- extern "C" entry points over the C++ drop-in surface
// (declared in include/icempm_host.h).
#include <cstring>
#include <stdexcept>
#include <string>

#include "host_model.h"

namespace {

void copy_str(char *dst, size_t cap, const std::string &s) {
    if (!dst || cap == 0) return;
    const size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
    std::memcpy(dst, s.data(), n);
    dst[n] = 0;
}

}  // namespace

extern "C" {

size_t icempm_host_simparams_size(void) { return sizeof(icy::SimParams); }

void icempm_simparams_reset(void *prms) { static_cast<icy::SimParams *>(prms)->Reset(); }

void icempm_simparams_derive(void *prms) {
    icy::SimParams *p = static_cast<icy::SimParams *>(prms);
    p->ComputeCamClayParams2();
    p->ComputeHelperVariables();
}

int icempm_simparams_parse_file(void *prms, const char *path, char *raw_points, size_t cap,
                                char *err, size_t errcap) {
    try {
        std::string r = static_cast<icy::SimParams *>(prms)->ParseFile(path ? path : "");
        copy_str(raw_points, cap, r);
        copy_str(err, errcap, "");
        return ICEMPM_OK;
    } catch (const std::exception &e) {
        copy_str(err, errcap, e.what());
        return ICEMPM_E_ARG;
    }
}

int icempm_model_create(icempm_model **out) {
    if (!out) return ICEMPM_E_ARG;
    *out = nullptr;
    icempm_model *m = nullptr;
    try {
        m = new icempm_model();
    } catch (...) {   // no exception crosses the C interface
        return ICEMPM_E_HOST;
    }
    m->snapshot.model = &m->model;
    int rc = icempm_guarded(m, [&] { m->model.gpu.initialize(); });
    if (rc != ICEMPM_OK) {
        delete m;
        return ICEMPM_E_NODEV;
    }
    *out = m;
    return ICEMPM_OK;
}

void icempm_model_destroy(icempm_model *m) {
    if (!m) return;
    try {
        m->model.gpu.synchronize();
    } catch (...) {
    }
    delete m;
}

const char *icempm_model_last_error(const icempm_model *m) { return m ? m->lastError.c_str() : ""; }

void *icempm_model_params(icempm_model *m) { return m ? &m->model.prms : nullptr; }

icempm_ctx *icempm_model_context(icempm_model *m) { return m ? m->model.gpu.context() : nullptr; }

int icempm_model_load_raw_points(icempm_model *m, const float *xy, const int16_t *grains, int n,
                                 float volume) {
    if (!m) return ICEMPM_E_ARG;
    return icempm_guarded(m, [&] { m->snapshot.LoadRawPoints(xy, grains, n, volume); });
}

int icempm_model_prepare(icempm_model *m) {
    if (!m) return ICEMPM_E_ARG;
    return icempm_guarded(m, [&] { m->model.Prepare(); });
}

int icempm_model_step(icempm_model *m, int *cont) {
    if (!m) return ICEMPM_E_ARG;
    return icempm_guarded(m, [&] {
        bool r = m->model.Step();
        if (cont) *cont = r ? 1 : 0;
    });
}

int icempm_model_wait_cycle(icempm_model *m) {
    if (!m) return ICEMPM_E_ARG;
    return icempm_guarded(m, [&] {
        m->model.gpu.synchronize();
        // the readback callback releases the handshake when it has run
        m->model.processing_current_cycle_data.lock();
        m->model.processing_current_cycle_data.unlock();
        m->model.gpu.refresh_error_code();
    });
}

float icempm_model_cycle_time(const icempm_model *m) { return m ? m->model.compute_time_per_cycle : 0.f; }

double *icempm_model_host_points(icempm_model *m) { return m ? m->model.gpu.tmp_transfer_buffer : nullptr; }

double *icempm_model_host_force(icempm_model *m) {
    return m ? m->model.gpu.host_side_indenter_force_accumulator : nullptr;
}

int icempm_model_set_mode(icempm_model *m, int mode) {
    if (!m) return ICEMPM_E_ARG;
    return icempm_guarded(m, [&] { m->model.gpu.set_mode(mode); });
}

}  // extern "C"
