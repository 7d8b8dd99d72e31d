// snapshot_h5.cpp -- the HDF5 side of icy::SnapshotManager, on the HDF5 C
// library: raw point input (snapshotmanager.cpp:186-218), snapshot and P/Q
// output (snapshotmanager.cpp:14-145), and a writer of raw point files in
// the reference's input format (so synthetic configurations can be fed to an
// unchanged frontend).  File layouts (dataset names, types, shapes,
// attributes) follow the reference's, so its readers and viewers open them.
#include <hdf5.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_model.h"
#include "icempm_h5.h"
#include "point.h"

namespace {

// closes an HDF5 handle at scope exit
struct H {
    hid_t id;
    herr_t (*close)(hid_t);
    H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H() {
        if (id >= 0) close(id);
    }
    H(const H &) = delete;
    H &operator=(const H &) = delete;
    operator hid_t() const { return id; }
};

void check(herr_t e, const char *what) {
    if (e < 0) throw std::runtime_error(std::string("HDF5: ") + what);
}
hid_t check_id(hid_t id, const char *what) {
    if (id < 0) throw std::runtime_error(std::string("HDF5: ") + what);
    return id;
}

void ensure_dir(const std::string &d) {
    std::filesystem::path p(d);
    if (!std::filesystem::is_directory(p)) std::filesystem::create_directory(p);
}

template <class T>
void scalar_attr(hid_t obj, const char *name, hid_t type, const T &v) {
    const hsize_t one = 1;
    H space(check_id(H5Screate_simple(1, &one, nullptr), "attribute space"), H5Sclose);
    H att(check_id(H5Acreate2(obj, name, type, space, H5P_DEFAULT, H5P_DEFAULT), name), H5Aclose);
    check(H5Awrite(att, type, &v), name);
}

// a contiguous 1D or 2D dataset written in one call
hid_t write_dataset(hid_t file, const char *name, hid_t type, int rank, const hsize_t *dims,
                    const void *data) {
    H space(check_id(H5Screate_simple(rank, dims, nullptr), "dataspace"), H5Sclose);
    hid_t ds = check_id(H5Dcreate2(file, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), name);
    if (H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) {
        H5Dclose(ds);
        throw std::runtime_error(std::string("HDF5: writing ") + name);
    }
    return ds;
}

// the 14 attributes of SaveParametersAsAttributes (snapshotmanager.cpp:107-145)
void parameter_attributes(hid_t ds, const icy::SimParams &P) {
    const int64_t pitch = (int64_t)P.nPtsPitch;
    scalar_attr(ds, "indenter_x", H5T_NATIVE_DOUBLE, P.indenter_x);
    scalar_attr(ds, "indenter_y", H5T_NATIVE_DOUBLE, P.indenter_y);
    scalar_attr(ds, "SimulationTime", H5T_NATIVE_DOUBLE, P.SimulationTime);
    scalar_attr(ds, "GridX", H5T_NATIVE_INT, P.GridX);
    scalar_attr(ds, "GridY", H5T_NATIVE_INT, P.GridY);
    scalar_attr(ds, "nPts", H5T_NATIVE_INT, P.nPts);
    scalar_attr(ds, "nPtsPitch", H5T_NATIVE_INT64, pitch);
    scalar_attr(ds, "UpdateEveryNthStep", H5T_NATIVE_INT, P.UpdateEveryNthStep);
    scalar_attr(ds, "n_indenter_subdivisions", H5T_NATIVE_INT, P.n_indenter_subdivisions);
    scalar_attr(ds, "cellsize", H5T_NATIVE_DOUBLE, P.cellsize);
    scalar_attr(ds, "IndDiameter", H5T_NATIVE_DOUBLE, P.IndDiameter);
    scalar_attr(ds, "InitialTimeStep", H5T_NATIVE_DOUBLE, P.InitialTimeStep);
    scalar_attr(ds, "SetupType", H5T_NATIVE_INT, P.SetupType);
    scalar_attr(ds, "Volume", H5T_NATIVE_DOUBLE, P.Volume);
}

// snapshotmanager.cpp:60-105: <dir>/snapshots/dNNNNN.h5 (NNNNN the
// animation frame of P) with the force bins (Indenter_2D), the raw parameter
// struct (Params, with attributes) and the pitched point SoA (Points)
void write_snapshot_file(const std::string &outputDirectory, const std::string &snapshotsDir,
                         const icy::SimParams &P, const double *indenter, const double *points) {
    ensure_dir(outputDirectory);
    ensure_dir(outputDirectory + "/" + snapshotsDir);
    char fileName[20];
    std::snprintf(fileName, sizeof(fileName), "d%05d.h5", P.SimulationStep / P.UpdateEveryNthStep);   // AnimationFrameNumber
    const std::string path = outputDirectory + "/" + snapshotsDir + "/" + fileName;

    H file(check_id(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), path.c_str()), H5Fclose);
    const hsize_t nInd = (hsize_t)P.n_indenter_subdivisions * 2;
    H ind(write_dataset(file, "Indenter_2D", H5T_NATIVE_DOUBLE, 1, &nInd, indenter), H5Dclose);
    const hsize_t nPrm = sizeof(icy::SimParams);
    H prm(write_dataset(file, "Params", H5T_NATIVE_B8, 1, &nPrm, &P), H5Dclose);
    parameter_attributes(prm.id, P);
    const hsize_t nPts = (hsize_t)P.nPtsPitch * icy::SimParams::nPtsArrays;
    H pts(write_dataset(file, "Points", H5T_NATIVE_DOUBLE, 1, &nPts, points), H5Dclose);
}

}  // namespace

void icy::SnapshotManager::SaveParametersAsAttributes(long long dataset)
{
    parameter_attributes((hid_t)dataset, model->prms);
}

// snapshotmanager.cpp:60-105, from the pinned readback buffers
void icy::SnapshotManager::SaveSnapshot(std::string outputDirectory)
{
    if (!model) throw std::runtime_error("SnapshotManager: no model");
    write_snapshot_file(outputDirectory, directory_snapshots, model->prms,
                        model->gpu.host_side_indenter_force_accumulator, model->gpu.tmp_transfer_buffer);
}

// snapshotmanager.cpp:14-56: <dir>/pq/pqNNNNN.h5 with (p, q) of the intact
// points (PQ_intact, with the parameter attributes) and of the crushed ones
// (PQ_crushed), in point order.
void icy::SnapshotManager::SavePQ(std::string outputDirectory)
{
    if (!model) throw std::runtime_error("SnapshotManager: no model");
    ensure_dir(outputDirectory);
    ensure_dir(outputDirectory + "/" + directory_pq);
    char fileName[20];
    std::snprintf(fileName, sizeof(fileName), "pq%05d.h5", model->prms.AnimationFrameNumber());
    const std::string path = outputDirectory + "/" + directory_pq + "/" + fileName;

    const int n = model->prms.nPts;
    const size_t pitch = model->prms.nPtsPitch;
    const double *buf = model->gpu.tmp_transfer_buffer;
    buffer1.clear();
    buffer2.clear();
    buffer1.reserve(2 * (size_t)n);
    buffer2.reserve(2 * (size_t)n);
    const unsigned char *u8 = reinterpret_cast<const unsigned char *>(buf);
    for (int i = 0; i < n; i++) {
        const double p = buf[pitch * icy::SimParams::idx_P + i];
        const double q = buf[pitch * icy::SimParams::idx_Q + i];
        std::vector<double> &dst = u8[i] ? buffer1 : buffer2;
        dst.push_back(p);
        dst.push_back(q);
    }
    H file(check_id(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), path.c_str()), H5Fclose);
    const hsize_t di[2] = {buffer2.size() / 2, 2};
    H intact(write_dataset(file, "PQ_intact", H5T_NATIVE_DOUBLE, 2, di, buffer2.data()), H5Dclose);
    SaveParametersAsAttributes(intact.id);
    const hsize_t dc[2] = {buffer1.size() / 2, 2};
    H crushed(write_dataset(file, "PQ_crushed", H5T_NATIVE_DOUBLE, 2, dc, buffer1.data()), H5Dclose);
}

// snapshotmanager.cpp:186-218: the raw point cloud from HDF5, then the
// array path (sort, placement, upload).
void icy::SnapshotManager::LoadRawPoints(std::string fileName)
{
    if (!model) throw std::runtime_error("SnapshotManager: no model");
    if (!std::filesystem::exists(fileName)) throw std::runtime_error("error reading raw points file - no file");
    H file(check_id(H5Fopen(fileName.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), fileName.c_str()), H5Fclose);
    H ds(check_id(H5Dopen2(file, "Points_Raw_2D", H5P_DEFAULT), "Points_Raw_2D"), H5Dclose);
    H space(check_id(H5Dget_space(ds), "Points_Raw_2D space"), H5Sclose);
    hsize_t dims[2] = {0, 0};
    if (H5Sget_simple_extent_ndims(space) != 2) throw std::runtime_error("error reading raw points file - dimensions mismatch");
    H5Sget_simple_extent_dims(space, dims, nullptr);
    if (dims[1] != (hsize_t)icy::SimParams::dim) throw std::runtime_error("error reading raw points file - dimensions mismatch");
    const int n = (int)dims[0];
    std::vector<float> xy(2 * (size_t)n);
    check(H5Dread(ds, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, xy.data()), "reading Points_Raw_2D");
    H dg(check_id(H5Dopen2(file, "GrainIDs", H5P_DEFAULT), "GrainIDs"), H5Dclose);
    std::vector<short> grains((size_t)n);
    check(H5Dread(dg, H5T_NATIVE_INT16, H5S_ALL, H5S_ALL, H5P_DEFAULT, grains.data()), "reading GrainIDs");
    H att(check_id(H5Aopen(dg, "volume", H5P_DEFAULT), "volume attribute"), H5Aclose);
    float volume = 0;
    check(H5Aread(att, H5T_NATIVE_FLOAT, &volume), "reading volume");
    LoadRawPoints(xy.data(), grains.data(), n, volume);
}

// Resume from a snapshot (the reference's ReadSnapshot, snapshotmanager.cpp:147-184,
// is commented out there): the parameter struct from Params (ParticleViewSize
// kept), arrays allocated for its grid and point count, the point SoA from
// Points (re-pitched if the file's pitch differs), uploaded, constants updated.
void icy::SnapshotManager::ReadSnapshot(std::string fileName)
{
    if (!model) throw std::runtime_error("SnapshotManager: no model");
    if (!std::filesystem::exists(fileName)) throw std::runtime_error("ReadSnapshot: no file");
    H file(check_id(H5Fopen(fileName.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), fileName.c_str()), H5Fclose);
    H dp(check_id(H5Dopen2(file, "Params", H5P_DEFAULT), "Params"), H5Dclose);
    {
        H sp(check_id(H5Dget_space(dp), "Params space"), H5Sclose);
        hsize_t n = 0;
        if (H5Sget_simple_extent_ndims(sp) != 1) throw std::runtime_error("ReadSnapshot: SimParams size mismatch");
        H5Sget_simple_extent_dims(sp, &n, nullptr);
        if (n != sizeof(icy::SimParams)) throw std::runtime_error("ReadSnapshot: SimParams size mismatch");
    }
    icy::SimParams tmp;
    check(H5Dread(dp, H5T_NATIVE_B8, H5S_ALL, H5S_ALL, H5P_DEFAULT, &tmp), "reading Params");
    H dpts(check_id(H5Dopen2(file, "Points", H5P_DEFAULT), "Points"), H5Dclose);
    const size_t filePitch = (size_t)tmp.nPtsPitch;
    {
        H sp(check_id(H5Dget_space(dpts), "Points space"), H5Sclose);
        hsize_t n = 0;
        H5Sget_simple_extent_dims(sp, &n, nullptr);
        if (n != filePitch * icy::SimParams::nPtsArrays || tmp.nPts < 0 || (size_t)tmp.nPts > filePitch)
            throw std::runtime_error("ReadSnapshot: Points size mismatch");
    }
    std::vector<double> pts(filePitch * icy::SimParams::nPtsArrays);
    check(H5Dread(dpts, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, pts.data()), "reading Points");

    const double viewSize = model->prms.ParticleViewSize;
    model->prms = tmp;
    model->prms.ParticleViewSize = viewSize;
    model->gpu.cuda_allocate_arrays(model->prms.GridTotal, model->prms.nPts);
    const size_t pitch = model->prms.nPtsPitch;
    double *dst = model->gpu.tmp_transfer_buffer;
    const int n = model->prms.nPts;
    // row 0 holds bytes (crushed at i, grain at pitch + 2i, viewer byte at
    // 3 pitch + i): re-pitch it byte-wise, the other rows as doubles
    std::fill(dst, dst + pitch * icy::SimParams::nPtsArrays, 0.0);
    const unsigned char *s8 = reinterpret_cast<const unsigned char *>(pts.data());
    unsigned char *d8 = reinterpret_cast<unsigned char *>(dst);
    for (int i = 0; i < n; ++i) {
        d8[i] = s8[i];
        d8[pitch + 2 * (size_t)i] = s8[filePitch + 2 * (size_t)i];
        d8[pitch + 2 * (size_t)i + 1] = s8[filePitch + 2 * (size_t)i + 1];
        d8[3 * pitch + i] = s8[3 * filePitch + i];
    }
    for (size_t r = 1; r < icy::SimParams::nPtsArrays; ++r)
        std::copy(pts.begin() + r * filePitch, pts.begin() + r * filePitch + n, dst + r * pitch);
    model->gpu.transfer_ponts_to_device();
    model->Prepare();
}

extern "C" {

int icempm_model_read_snapshot(icempm_model *m, const char *path) {
    if (!m || !path) return ICEMPM_E_ARG;
    return icempm_guarded(m, [&] { m->snapshot.ReadSnapshot(std::string(path)); });
}

int icempm_model_load_raw_points_file(icempm_model *m, const char *path) {
    if (!m || !path) return ICEMPM_E_ARG;
    return icempm_guarded(m, [&] { m->snapshot.LoadRawPoints(std::string(path)); });
}

int icempm_model_save_snapshot(icempm_model *m, const char *dir) {
    if (!m || !dir) return ICEMPM_E_ARG;
    return icempm_guarded(m, [&] { m->snapshot.SaveSnapshot(std::string(dir)); });
}

int icempm_model_save_pq(icempm_model *m, const char *dir) {
    if (!m || !dir) return ICEMPM_E_ARG;
    return icempm_guarded(m, [&] { m->snapshot.SavePQ(std::string(dir)); });
}

int icempm_h5_write_snapshot(const char *dir, const void *simparams, size_t nbytes, const double *indenter_force,
                             const double *points) {
    if (!dir || !simparams || nbytes != sizeof(icy::SimParams) || !indenter_force || !points)
        return ICEMPM_E_ARG;
    icy::SimParams P;
    std::memcpy(&P, simparams, sizeof(P));
    if (P.n_indenter_subdivisions <= 0 || P.nPtsPitch <= 0 || P.UpdateEveryNthStep <= 0) return ICEMPM_E_ARG;
    try {
        write_snapshot_file(dir, "snapshots", P, indenter_force, points);
        return ICEMPM_OK;
    } catch (const std::exception &) {
        return ICEMPM_E_IO;
    }
}

int icempm_h5_write_raw_points(const char *path, const float *xy, const int16_t *grains, int n, float volume,
                               int deflate) {
    if (!path || !xy || !grains || n <= 0 || deflate < 0 || deflate > 9) return ICEMPM_E_ARG;
    try {
        H file(check_id(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), path), H5Fclose);
        const hsize_t d2[2] = {(hsize_t)n, 2}, d1 = (hsize_t)n;
        H plist2(check_id(H5Pcreate(H5P_DATASET_CREATE), "plist"), H5Pclose);
        H plist1(check_id(H5Pcreate(H5P_DATASET_CREATE), "plist"), H5Pclose);
        if (deflate > 0) {
            const hsize_t c2[2] = {(hsize_t)std::min(n, 1 << 16), 2}, c1 = (hsize_t)std::min(n, 1 << 16);
            check(H5Pset_chunk(plist2, 2, c2), "chunk");
            check(H5Pset_deflate(plist2, (unsigned)deflate), "deflate");
            check(H5Pset_chunk(plist1, 1, &c1), "chunk");
            check(H5Pset_deflate(plist1, (unsigned)deflate), "deflate");
        }
        H s2(check_id(H5Screate_simple(2, d2, nullptr), "space"), H5Sclose);
        H ds(check_id(H5Dcreate2(file, "Points_Raw_2D", H5T_IEEE_F32LE, s2, H5P_DEFAULT, plist2, H5P_DEFAULT),
                      "Points_Raw_2D"), H5Dclose);
        check(H5Dwrite(ds, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, xy), "writing Points_Raw_2D");
        H s1(check_id(H5Screate_simple(1, &d1, nullptr), "space"), H5Sclose);
        H dg(check_id(H5Dcreate2(file, "GrainIDs", H5T_STD_I16LE, s1, H5P_DEFAULT, plist1, H5P_DEFAULT),
                      "GrainIDs"), H5Dclose);
        check(H5Dwrite(dg, H5T_NATIVE_INT16, H5S_ALL, H5S_ALL, H5P_DEFAULT, grains), "writing GrainIDs");
        scalar_attr(dg, "volume", H5T_IEEE_F32LE, volume);
        return ICEMPM_OK;
    } catch (const std::exception &) {
        return ICEMPM_E_IO;
    }
}

}  // extern "C"
