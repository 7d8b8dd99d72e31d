// parameters_sim.cpp -- defaults, JSON parsing and derived constants of
// icy::SimParams (reference: simulation/parameters_sim.cpp:5-155).
#include "parameters_sim.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>

void icy::SimParams::Reset()
{
    // zero every byte first so that the dumped POD is deterministic
    std::memset(static_cast<void*>(this), 0, sizeof(*this));
    grid_array = nullptr;
    pts_array = nullptr;
    indenter_force_accumulator = nullptr;

    nPts = 0;
    n_indenter_subdivisions = 100;

    InitialTimeStep = 3.e-5;
    YoungsModulus = 5.e8;
    GridX = 128;
    GridY = 55;
    ParticleViewSize = 2.5f;
    GridXDimension = 3.33;
    SimulationEndTime = 12;
    PoissonsRatio = 0.3;
    Gravity = 9.81;
    Density = 980;
    IndDiameter = 0.324;
    IndVelocity = 0.2;
    IndDepth = 0.25;
    SimulationStep = 0;
    SimulationTime = 0;
    IceCompressiveStrength = 100e6;
    IceTensileStrength = 1e6;
    IceShearStrength = 0.5e6;
    DP_tan_phi = std::tan(30*pi/180.);
    DP_threshold_p = 1e3;
    tpb_P2G = 256;
    tpb_Upd = 512;
    tpb_G2P = 128;
    indenter_x = indenter_x_initial = indenter_y = indenter_y_initial = 0;
    SetupType = 0;
    GrainVariability = 0.50;
    UseWarpLevelReduction = 0;
    SortPointsByGridCell = 1;

    ComputeLame();
    ComputeCamClayParams2();
    ComputeHelperVariables();
}

namespace {

// Minimal reader for the flat JSON objects the configs use: string keys
// mapped to numbers, strings, true/false/null.  Nested values are skipped.
struct FlatJson {
    std::map<std::string, std::string> strings;
    std::map<std::string, double> numbers;
};

void skip_ws(const std::string &s, size_t &i) {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
}

bool parse_string(const std::string &s, size_t &i, std::string &out) {
    if (i >= s.size() || s[i] != '"') return false;
    ++i;
    out.clear();
    while (i < s.size() && s[i] != '"') {
        if (s[i] == '\\' && i + 1 < s.size()) {
            char e = s[i + 1];
            out.push_back(e == 'n' ? '\n' : e == 't' ? '\t' : e);
            i += 2;
        } else {
            out.push_back(s[i++]);
        }
    }
    if (i >= s.size()) return false;
    ++i;
    return true;
}

bool skip_value(const std::string &s, size_t &i) {
    // skip a nested array/object, tracking strings
    int depth = 0;
    do {
        if (i >= s.size()) return false;
        char ch = s[i];
        if (ch == '"') {
            std::string tmp;
            if (!parse_string(s, i, tmp)) return false;
            continue;
        }
        if (ch == '{' || ch == '[') ++depth;
        if (ch == '}' || ch == ']') --depth;
        ++i;
    } while (depth > 0);
    return true;
}

bool parse_flat(const std::string &s, FlatJson &out) {
    size_t i = 0;
    skip_ws(s, i);
    if (i >= s.size() || s[i] != '{') return false;
    ++i;
    skip_ws(s, i);
    if (i < s.size() && s[i] == '}') return true;
    while (i < s.size()) {
        skip_ws(s, i);
        std::string key;
        if (!parse_string(s, i, key)) return false;
        skip_ws(s, i);
        if (i >= s.size() || s[i] != ':') return false;
        ++i;
        skip_ws(s, i);
        if (i >= s.size()) return false;
        if (s[i] == '"') {
            std::string v;
            if (!parse_string(s, i, v)) return false;
            out.strings[key] = v;
        } else if (s[i] == '{' || s[i] == '[') {
            if (!skip_value(s, i)) return false;
        } else if (s.compare(i, 4, "true") == 0) {
            out.numbers[key] = 1;
            i += 4;
        } else if (s.compare(i, 5, "false") == 0) {
            out.numbers[key] = 0;
            i += 5;
        } else if (s.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            const char *b = s.c_str() + i;
            char *e = nullptr;
            double v = std::strtod(b, &e);
            if (e == b) return false;
            out.numbers[key] = v;
            i += (size_t)(e - b);
        }
        skip_ws(s, i);
        if (i < s.size() && s[i] == ',') {
            ++i;
            continue;
        }
        if (i < s.size() && s[i] == '}') return true;
        return false;
    }
    return false;
}

}  // namespace

std::string icy::SimParams::ParseFile(std::string fileName)
{
    if (!std::filesystem::exists(fileName)) throw std::runtime_error("configuration file is not found");
    std::ifstream fileStream(fileName);
    std::stringstream ss;
    ss << fileStream.rdbuf();
    FlatJson doc;
    if (!parse_flat(ss.str(), doc)) throw std::runtime_error("configuration file is not JSON");

    auto num = [&](const char *k, double &dst) {
        auto it = doc.numbers.find(k);
        if (it != doc.numbers.end()) dst = it->second;
    };
    auto inum = [&](const char *k, int &dst) {
        auto it = doc.numbers.find(k);
        if (it != doc.numbers.end()) dst = (int)it->second;
    };
    num("InitialTimeStep", InitialTimeStep);
    num("YoungsModulus", YoungsModulus);
    inum("GridX", GridX);
    inum("GridY", GridY);
    num("GridXDimension", GridXDimension);
    num("ParticleViewSize", ParticleViewSize);
    num("SimulationEndTime", SimulationEndTime);
    num("PoissonsRatio", PoissonsRatio);
    num("Gravity", Gravity);
    num("Density", Density);
    num("IndDiameter", IndDiameter);
    num("IndVelocity", IndVelocity);
    num("IndDepth", IndDepth);
    num("IceCompressiveStrength", IceCompressiveStrength);
    num("IceTensileStrength", IceTensileStrength);
    num("IceShearStrength", IceShearStrength);
    {
        auto it = doc.numbers.find("DP_phi");
        if (it != doc.numbers.end()) DP_tan_phi = std::tan(it->second * pi / 180);
    }
    num("DP_threshold_p", DP_threshold_p);
    num("GrainVariability", GrainVariability);
    inum("UseWarpLevelReduction", UseWarpLevelReduction);
    inum("tpb_P2G", tpb_P2G);
    inum("tpb_Upd", tpb_Upd);
    inum("tpb_G2P", tpb_G2P);
    inum("SortPointsByGridCell", SortPointsByGridCell);

    ComputeCamClayParams2();
    ComputeHelperVariables();

    auto it = doc.strings.find("InputRawPoints");
    if (it == doc.strings.end()) throw std::runtime_error("config parameter missing");
    return it->second;
}

void icy::SimParams::ComputeLame()
{
    lambda = YoungsModulus*PoissonsRatio/((1+PoissonsRatio)*(1-2*PoissonsRatio));
    mu = YoungsModulus/(2*(1+PoissonsRatio));
    kappa = mu*2./3. + lambda;
}

void icy::SimParams::ComputeHelperVariables()
{
    // the float literal is promoted: the quotient is evaluated in double
    UpdateEveryNthStep = (int)(1.f/(200*InitialTimeStep));
    cellsize = GridXDimension/GridX;
    cellsize_inv = 1./cellsize;
    Dp_inv = 4./(cellsize*cellsize);
    IndRSq = IndDiameter*IndDiameter/4.;
    GridTotal = GridX*GridY;
}

void icy::SimParams::ComputeCamClayParams2()
{
    ComputeLame();
    NACC_beta = IceTensileStrength/IceCompressiveStrength;
    const double beta = NACC_beta;
    const double q = IceShearStrength;
    const double p0 = IceCompressiveStrength;
    NACC_M = (2*q*std::sqrt(1+2*beta))/(p0*(1+beta));
    NACC_Msq = NACC_M*NACC_M;
}

void icy::SimParams::ComputeIntegerBlockCoords()
{
    nxmin = (int)std::floor(xmin/cellsize);
    nxmax = (int)std::ceil(xmax/cellsize);
    nymin = (int)std::floor(ymin/cellsize);
    nymax = (int)std::ceil(ymax/cellsize);
}
