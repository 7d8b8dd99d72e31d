// snapshotmanager.h -- icy::SnapshotManager (reference:
// simulation/snapshotmanager.h:13-28, snapshotmanager.cpp:14-304).
// LoadRawPoints on arrays and SortPoints (the stable cell-key sort on the
// device) are in libicempm_host.so; the HDF5 members -- LoadRawPoints from a
// file (Points_Raw_2D f32[N][2], GrainIDs i16[N] with a f32 volume
// attribute), SaveSnapshot, SavePQ -- are in libicempm_h5.so
// (snapshot_h5.cpp), which links the HDF5 C library.
#ifndef SNAPSHOTMANAGER_H
#define SNAPSHOTMANAGER_H

#include <string>
#include <tuple>
#include <vector>

namespace icy { class SnapshotManager; class Model; }

class icy::SnapshotManager
{
public:
    icy::Model *model = nullptr;

    void LoadRawPoints(const float *xy, const short *grains, int nPoints, float volume);
    void SortPoints(std::vector<std::tuple<float, float, short>> &buffer);

    // libicempm_h5.so
    void LoadRawPoints(std::string fileName);
    void SaveSnapshot(std::string directory);
    void ReadSnapshot(std::string fileName);   // resume from a snapshot
    void SavePQ(std::string directory);
    // the attributes SaveSnapshot and SavePQ attach to a dataset (an HDF5
    // hid_t; the reference takes an H5::DataSet)
    void SaveParametersAsAttributes(long long dataset);

    const std::string directory_snapshots = "snapshots";
    const std::string directory_pq = "pq";
    std::vector<double> buffer1, buffer2;
};

#endif
