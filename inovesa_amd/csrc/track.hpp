// Passive particle tracking: SourceMap::applyTo of every map
// (src/SM/KickMap.cpp:280-304, src/SM/RotationMap.cpp:178-188,
// src/SM/FokkerPlanckMap.cpp:214-268), on the device.
#pragma once

#include "common.hpp"

namespace vfps_impl {

enum TrackKind : int {
    TRACK_KICK_X = 0,   // KickMap, Axis::x (displacement along x per y)
    TRACK_KICK_Y = 1,   // KickMap, Axis::y
    TRACK_ROTATION = 2,
    TRACK_FP = 3,
};

struct TrackOp {
    int kind;
    const float* off;       // kick offsets
    uint32_t pd, kd;        // _meshsize_pd, _meshsize_kd of the kick
    float cs, sn, zb0, zb1; // rotation: _cos_dt, _sin_dt, zerobins
    int fptrack;            // FPTracking 0..3
    uint32_t ip, ysize, nx; // FP taps per y, mesh sizes
    const float* fp_w;      // [ip][ysize]
    const uint32_t* fp_i;   // [ip][ysize]
    const float* data_in;   // the FP's input mesh (approximation2)
    float e1;               // _dampdecr
    float sigma;            // stochastic: sqrt(2 e1) / delta_p
};

constexpr int TRACK_MAX_OPS = 4;

// the applyTo calls of one step (or of one map)
struct TrackProgram {
    TrackOp op[TRACK_MAX_OPS];
    int nops;
    uint64_t seed;
};

// the tracking ops of one map (an Identity gives none); false if the map
// kind has no applyTo here
bool track_op(const vfps_map* m, const float* kick_off, TrackOp* op);

// one launch: every particle through the program; the stochastic noise of
// particle i is drawn from (seed, i, op, *counter + call)
int launch_track(const TrackProgram& prog, float2* pos, uint32_t n,
                 const uint32_t* counter, uint32_t call, hipStream_t st);

}  // namespace vfps_impl
