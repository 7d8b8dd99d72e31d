// snapshotmanager.cpp -- raw point ingest: sort, placement of the block and
// the indenter, particle mass, upload (snapshotmanager.cpp:186-304).
#include "snapshotmanager.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>

#include "model.h"
#include "point.h"

void icy::SnapshotManager::LoadRawPoints(const float *xy, const short *grains, int nPoints,
                                         float volume)
{
    if (!model) throw std::runtime_error("SnapshotManager: no model");
    if (nPoints <= 0 || !xy || !grains) throw std::runtime_error("error reading raw points file - no points");
    icy::SimParams &P = model->prms;
    P.nPts = nPoints;

    std::vector<std::array<float, 2>> buffer((size_t)nPoints);
    std::vector<short> grainIDs(grains, grains + nPoints);
    for (int i = 0; i < nPoints; i++) buffer[i] = {xy[2 * (size_t)i], xy[2 * (size_t)i + 1]};

    if (P.SortPointsByGridCell)
    {
        std::vector<std::tuple<float, float, short>> tmp((size_t)nPoints);
        for (int i = 0; i < nPoints; i++) tmp[i] = std::make_tuple(buffer[i][0], buffer[i][1], grainIDs[i]);
        SortPoints(tmp);
        for (int i = 0; i < nPoints; i++)
        {
            buffer[i][0] = std::get<0>(tmp[i]);
            buffer[i][1] = std::get<1>(tmp[i]);
            grainIDs[i] = std::get<2>(tmp[i]);
        }
    }

    P.Volume = (double)volume;

    auto xr = std::minmax_element(buffer.begin(), buffer.end(),
        [](const std::array<float, 2> &a, const std::array<float, 2> &b) { return a[0] < b[0]; });
    P.xmin = (*xr.first)[0];
    P.xmax = (*xr.second)[0];
    const float length = P.xmax - P.xmin;

    auto yr = std::minmax_element(buffer.begin(), buffer.end(),
        [](const std::array<float, 2> &a, const std::array<float, 2> &b) { return a[1] < b[1]; });
    P.ymin = (*yr.first)[1];
    P.ymax = (*yr.second)[1];

    P.ComputeIntegerBlockCoords();

    const double h = P.cellsize;
    const double box_x = P.GridX * h;
    const double x_offset = (box_x - length) / 2;
    const double y_offset = 2 * h;
    const double block_left = x_offset;
    const double block_top = P.ymax + y_offset;

    const double r = P.IndDiameter / 2;
    const double ht = r - P.IndDepth;
    const double x_ind_offset = std::sqrt(r * r - ht * ht);

    P.indenter_x = std::floor((block_left - x_ind_offset) / h) * h;
    if (P.SetupType == 0)
        P.indenter_y = block_top + ht;
    else if (P.SetupType == 1)
        P.indenter_y = std::ceil(block_top / h) * h;
    P.indenter_x_initial = P.indenter_x;
    P.indenter_y_initial = P.indenter_y;

    P.ParticleVolume = P.Volume / nPoints;
    P.ParticleMass = P.ParticleVolume * P.Density;

    model->gpu.cuda_allocate_arrays(P.GridTotal, nPoints);
    for (int k = 0; k < nPoints; k++)
    {
        Point p;
        p.Reset();
        buffer[k][0] += x_offset;   // float += double: the sum is rounded to float
        buffer[k][1] += y_offset;
        for (int i = 0; i < icy::SimParams::dim; i++) p.pos[i] = buffer[k][i];
        p.grain = grainIDs[k];
        p.TransferToBuffer(model->gpu.tmp_transfer_buffer, (int)P.nPtsPitch, k);
    }

    model->gpu.transfer_ponts_to_device();
    model->Reset();
    model->Prepare();
}

// Stable sort by the cell key (int)(x*hinv-0.5) + (int)(y*hinv-0.5)*GridX of
// the raw float coordinates, on the device.
void icy::SnapshotManager::SortPoints(std::vector<std::tuple<float, float, short>> &buffer)
{
    const int n = (int)buffer.size();
    if (n <= 1) return;
    model->gpu.initialize();
    std::vector<float> xy(2 * (size_t)n);
    std::vector<int16_t> g((size_t)n);
    for (int i = 0; i < n; i++)
    {
        xy[2 * (size_t)i] = std::get<0>(buffer[i]);
        xy[2 * (size_t)i + 1] = std::get<1>(buffer[i]);
        g[i] = std::get<2>(buffer[i]);
    }
    const int rc = icempm_sort_raw_points(model->gpu.context(), xy.data(), g.data(), n,
                                          model->prms.cellsize_inv, model->prms.GridX);
    if (rc != ICEMPM_OK) throw std::runtime_error(std::string("sort points: ") + icempm_last_error(model->gpu.context()));
    for (int i = 0; i < n; i++) buffer[i] = std::make_tuple(xy[2 * (size_t)i], xy[2 * (size_t)i + 1], g[i]);
}
