// gelios_internal.h - the context behind the opaque gel_ctx handle, shared by the
// sweep (gelios_sweep.hip) and the surrogate (surrogate.hip) translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "../../include/gelios.h"

struct gel_ctx {
    int ni = 0, nj = 0;
    size_t nc = 0, nn = 0;
    int nv = 5;                   // primitives per cell
    size_t ts = 8;                // bytes per stored primitive (8 FP64, 4 FP32)
    gel_params prm{};
    double* d_x = nullptr;        // node positions (base positions under prescribed motion)
    double* d_y = nullptr;
    double* d_x2 = nullptr;       // next node positions (explicit moving steps)
    double* d_y2 = nullptr;
    void* d_state[2] = { nullptr, nullptr };
    int cur = 0;
    int8_t* d_ctype = nullptr;
    int8_t* d_otype = nullptr;
    bool has_types = false;
    double* d_src = nullptr;       // [GEL_MK1_NFIELDS][nc] hydrogen moments and k factors
    double* d_qtab = nullptr;      // [5][nc] per-cell source table (GEL_SRC_TABLE)
    unsigned long long* d_dt = nullptr;   // [3]
    unsigned long long* d_cnt = nullptr;  // [3]
    long long steps = 0;
    bool uploaded = false;
    int rows = 8;
    // prescribed motion: nodes at base*s[k], k = step mod period
    std::vector<double> motion;
    long long motion_step = 0;
    hipStream_t stream = nullptr;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};

