// snapshotmanager.h -- the ingest half of icy::SnapshotManager
// (reference: simulation/snapshotmanager.h:13-28, snapshotmanager.cpp:186-304).
// LoadRawPoints takes the raw point cloud as arrays (the reference reads the
// same data -- Points_Raw_2D f32[N][2], GrainIDs i16[N], a volume attribute --
// from HDF5); SortPoints runs the stable cell-key sort on the device.
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
};

#endif
