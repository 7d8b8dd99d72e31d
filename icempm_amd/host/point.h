// point.h -- icy::Point, one particle's record and the accessors of the
// point SoA buffer (reference: simulation/point.h:14-38, point.cpp:8-70).
// Plain arrays instead of Eigen types; the SoA byte layout is unchanged:
// row 0 holds crushed (u8 at byte i) and grain (i16 at byte pitch + 2i),
// rows 1..15 are P, Q, Jp_inv, pos(2), vel(2), Fe(4, row-major), Bp(4).
#ifndef PARTICLE_H
#define PARTICLE_H

#include <array>
#include <cstdint>
#include <cstring>
#include <utility>

#include "parameters_sim.h"

namespace icy { struct Point; }

struct icy::Point
{
    double pos[2] = {0, 0}, velocity[2] = {0, 0};
    double Bp[4] = {0, 0, 0, 0}, Fe[4] = {1, 0, 0, 1};   // row-major 2x2
    double Jp_inv = 1;
    short grain = 0;
    uint8_t crushed = 0;

    void Reset()
    {
        Fe[0] = 1; Fe[1] = 0; Fe[2] = 0; Fe[3] = 1;
        velocity[0] = velocity[1] = 0;
        Bp[0] = Bp[1] = Bp[2] = Bp[3] = 0;
        crushed = 0;
        Jp_inv = 1;
    }

    void TransferToBuffer(double *buffer, const int pitch, const int point_index) const
    {
        unsigned char *util = reinterpret_cast<unsigned char *>(buffer + (size_t)pitch * SimParams::idx_utility_data);
        util[point_index] = crushed;
        std::memcpy(util + pitch + 2 * (size_t)point_index, &grain, sizeof(short));
        buffer[point_index + (size_t)pitch * SimParams::idx_Jp_inv] = Jp_inv;
        for (int i = 0; i < SimParams::dim; i++)
        {
            buffer[point_index + (size_t)pitch * (SimParams::posx + i)] = pos[i];
            buffer[point_index + (size_t)pitch * (SimParams::velx + i)] = velocity[i];
            for (int j = 0; j < SimParams::dim; j++)
            {
                buffer[point_index + (size_t)pitch * (SimParams::Fe00 + i * SimParams::dim + j)] = Fe[i * 2 + j];
                buffer[point_index + (size_t)pitch * (SimParams::Bp00 + i * SimParams::dim + j)] = Bp[i * 2 + j];
            }
        }
    }

    static std::array<double, 2> getPos(const double *buffer, const int pitch, const int point_index)
    {
        return {buffer[point_index + (size_t)pitch * SimParams::posx],
                buffer[point_index + (size_t)pitch * (SimParams::posx + 1)]};
    }
    static uint8_t getCrushedStatus(const double *buffer, const int /*pitch*/, const int point_index)
    {
        return reinterpret_cast<const uint8_t *>(buffer)[point_index];
    }
    static double getJp_inv(const double *buffer, const int pitch, const int point_index)
    {
        return buffer[point_index + (size_t)pitch * SimParams::idx_Jp_inv];
    }
    static short getGrain(const double *buffer, const int pitch, const int point_index)
    {
        short g;
        std::memcpy(&g, reinterpret_cast<const unsigned char *>(buffer) + pitch + 2 * (size_t)point_index, sizeof(short));
        return g;
    }
    static std::pair<double, double> getPQ(const double *buffer, const int pitch, const int point_index)
    {
        return {buffer[point_index + (size_t)pitch * SimParams::idx_P],
                buffer[point_index + (size_t)pitch * SimParams::idx_Q]};
    }
};

#endif
