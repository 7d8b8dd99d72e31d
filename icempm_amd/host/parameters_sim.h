// parameters_sim.h -- icy::SimParams, the simulation parameter POD.
//
// Same members, types and order as the reference's struct
// (simulation/parameters_sim.h:19-90): the reference uploads it verbatim to
// the device (gpu_implementation5.cu:476) and dumps its raw bytes into
// every snapshot (snapshotmanager.cpp:80-83), so the byte layout is part of
// the interface (432 bytes on x86-64).  The device pointers and pitches are
// filled in by GPU_Implementation5::cuda_allocate_arrays.  No Eigen or
// rapidjson: ParseFile reads the flat JSON configs with a small parser of
// its own.
#ifndef P_SIM_H
#define P_SIM_H

#include <cstddef>
#include <string>

namespace icy { struct SimParams; }

struct icy::SimParams
{
public:
    constexpr static double pi = 3.14159265358979323846;
    constexpr static int dim = 2;
    constexpr static int nGridArrays = 3;

    // rows of the point SoA buffer
    constexpr static size_t idx_utility_data = 0;
    constexpr static size_t idx_P = idx_utility_data + 1;
    constexpr static size_t idx_Q = idx_P + 1;
    constexpr static size_t idx_Jp_inv = idx_Q + 1;
    constexpr static size_t posx = idx_Jp_inv + 1;
    constexpr static size_t velx = posx + 2;
    constexpr static size_t Fe00 = velx + 2;
    constexpr static size_t Bp00 = Fe00 + 4;
    constexpr static size_t nPtsArrays = Bp00 + 4;

    double *grid_array;                  // device: 3 rows x nGridPitch
    double *indenter_force_accumulator;  // device: 2 x n_indenter_subdivisions
    double *pts_array;                   // device: 16 rows x nPtsPitch

    size_t nPtsPitch, nGridPitch;        // row strides, in doubles
    int n_indenter_subdivisions;
    int tpb_P2G, tpb_Upd, tpb_G2P;  // threads per block (kept for layout; the HIP kernels pick their own)

    int nPts;
    int GridX, GridY, GridTotal;
    double GridXDimension;

    double InitialTimeStep, SimulationEndTime;
    double Gravity, Density, PoissonsRatio, YoungsModulus;
    double lambda, mu, kappa;            // derived by ComputeLame

    double IceCompressiveStrength, IceTensileStrength, IceShearStrength;
    double NACC_beta, NACC_M, NACC_Msq;  // derived by ComputeCamClayParams2

    double DP_tan_phi, DP_threshold_p;

    double cellsize, cellsize_inv, Dp_inv;

    int UpdateEveryNthStep;              // substeps per output cycle

    double IndDiameter, IndRSq, IndVelocity, IndDepth;
    double xmin, xmax, ymin, ymax;       // extent of the raw point cloud
    int nxmin, nxmax, nymin, nymax;      // that extent in cells

    double ParticleVolume, ParticleMass, ParticleViewSize;

    int SimulationStep;
    double SimulationTime;

    double indenter_x, indenter_x_initial, indenter_y, indenter_y_initial;
    double Volume;                       // area of the ice body (2D)
    int SetupType;                       // 0: side indentation, 1: indenter placed on top
    double GrainVariability;

    int UseWarpLevelReduction; // parsed, unused (as in the reference)
    int SortPointsByGridCell;

    void Reset();
    std::string ParseFile(std::string fileName);

    void ComputeLame();
    void ComputeCamClayParams2();
    void ComputeHelperVariables();
    void ComputeIntegerBlockCoords();
    double PointsPerCell() {return nPts/(Volume/(cellsize*cellsize));}
    int AnimationFrameNumber() { return SimulationStep / UpdateEveryNthStep;}
    size_t IndenterArraySize() { return sizeof(double)*n_indenter_subdivisions*2; }
};

#endif
