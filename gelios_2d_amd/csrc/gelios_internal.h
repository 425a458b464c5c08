// gelios_internal.h - the context behind the opaque gel_ctx handle, shared by the
// sweep (gelios_sweep.hip) and the surrogate (surrogate.hip) translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/gelios.h"

namespace gel {

// The rotating time-step buffer: three slots (the step's dt, the minimum the step
// builds for the next one, the slot cleared for the step after), each of kDtLanes
// entries kDtStride apart, bit patterns of positive doubles.  A wavefront takes its
// atomic minimum into the entry of its index mod kDtLanes: atomics on one address
// serialise (about 11 ns each on MI355X, 23 us for 2048 wavefronts), on addresses
// 256 B apart they proceed side by side.  A slot's value is the minimum of its
// entries; the minimum is exact in any order, so the value does not depend on it.
constexpr int kDtLanes = 16;
constexpr int kDtStride = 32;
constexpr int kDtSlot = kDtLanes * kDtStride;   // unsigned long long per slot
constexpr int kDtWords = 3 * kDtSlot;

// Arguments of one sweep step of one case (also what the fused source refresh reads:
// the grid, the state after the step in `next`, the hydrogen moments and the table).
struct SweepArgs {
    int ni, nj;
    int rows;                     // rows per chunk (the tile of one work-group)
    int dt_rd, dt_wr, dt_clr;     // slots of the rotating time-step buffer
    double s_old, s_new;          // node scale factors before / after the step (moving)
    const double* __restrict__ x; // node positions before the step are x*s_old, y*s_old
    const double* __restrict__ y;
    const double* __restrict__ x2;  // after the step: x2*s_new, y2*s_new
    const double* __restrict__ y2;
    const void* __restrict__ now; // [NV][ncells] of the storage type
    void* __restrict__ next;
    const int8_t* __restrict__ ctype;
    const int8_t* __restrict__ otype;
    const void* __restrict__ src;     // [19][ncells] hydrogen moments and k factors, storage type
    const void* __restrict__ qtab;    // [5][ncells] source table (mass, mom x, mom y, energy, tracer), storage type
    const int* __restrict__ qflag;    // [ncells] 1: the cell takes the analytic source (GEL_SRC_HYBRID)
    const double* __restrict__ geo;   // static grids: the geometry table (geometry_kernel)
    int srow[3];                  // node rows of the fitted surfaces, -1 where none
    int fluid;                    // plasma + hydrogen sweep: 0 plasma, 1..4 hydrogen H1..H4
    int voff[5];                  // its five variables' arrays in the 21-field state
    const double* __restrict__ mfq;   // plasma + hydrogen: [5 fluids][4][ncells] sources of the step
    unsigned long long* dtbuf;    // [3][kDtSlot] time-step slots
    unsigned long long* counters; // [3] rho clamps, p clamps, numeric failures
    gel_params prm;
};

// Device view of the fitted-surface tracking (tracking.hip).
struct TrackDev {
    int rows[3];                  // node rows of the contact, inner shock, outer shock
    int outer_faces, average_lines, n_inner, m4;
    double koef, k_line[3];       // speed factors of the three lines
    double tang_inner, tang_outer;
    double r_inner, r_mid, r_outer;
    const double* cs;             // [nj+1] cos, sin of the rail angles
    const double* sn;
    const double* f_in;           // [n_inner] (j/n_inner)^p_inner
    const double* f_mid;          // [row_inner_shock - n_inner] ((j-n_inner)/(M1-1-n_inner))^p_mid
    const double* f_out;          // [nj+1][M4] ((j+1)/M4)^(p_outer (1 - (s/dd)^2))
    double* face;                 // [3][nj][6] node contributions of every surface face
};

}  // namespace gel

struct gel_mlp;

// Fused source refresh (surrogate.hip): the source table args[k].qtab of each case
// from its state args[k].next, hydrogen moments args[k].src and the network m, with
// the analytic judge; fallback cells are added to *d_nfall.  One work-group per
// tile of 256 cells; d_tiles[g] = (case, tile of the case) for work-group g, or null
// for a single case (args[0]).  Asynchronous on stream s; no allocation.
hipError_t gel_refresh_launch(hipStream_t s, const gel::SweepArgs* d_args, const int2* d_tiles, long long tiles,
                              gel_mlp* m, bool f32, double tol, int judge, unsigned long long* d_nfall);
// tiles the refresh of an ni x nj case takes
inline long long gel_refresh_tiles(int ni, int nj) { return ((long long)ni * nj + 255) / 256; }
// Sticky refresh (judge 2) of a batch over d_list[0 .. d_count[0]), the (case, cell)
// pairs still on the network, then the list rebuilt from the flags (the flagged
// cells added to *d_nfall).  Asynchronous on stream s; no allocation.
hipError_t gel_refresh_list_launch(hipStream_t s, const gel::SweepArgs* d_args, const int2* d_tiles, long long tiles,
                                   long long ncells, int2* d_list, unsigned* d_count, gel_mlp* m, bool f32,
                                   double tol, unsigned long long* d_nfall);
// Build d_list / d_count[0] from the flags (d_count[1] must be 0; it is left 0);
// ncells: the cells of all cases.
hipError_t gel_live_list_build(hipStream_t s, const gel::SweepArgs* d_args, const int2* d_tiles, long long tiles,
                               long long ncells, int2* d_list, unsigned* d_count);

// Record the detail gel_last_error() reports and return st (gelios_sweep.hip).
int gel_set_error(int st, const std::string& msg);

struct gel_ctx;
struct MeshBuf;                   // device face tables of a mesh context (mesh.hip)
// Build / free the face tables of a general mesh (mesh.hip); *err names a rejected input.
int gel_mesh_build(gel_ctx* c, const gel_mesh* g, std::string& err);
int gel_mesh_set_tracking(gel_ctx* c, const gel_tracking* t, const gel_mesh_rails* r, std::string& err);
void gel_mesh_free(gel_ctx* c);
// The mesh tables the Monte Carlo walks (device pointers), each face slot's way on
// (code: the neighbour cell, -1 out of the grid, -2 across the inner sphere, -3 an Axis
// slot, skipped) and its boundary face lists as (A, B, cell) triples: outer (entry)
// faces and inner-sphere faces.
struct MeshMcView {
    const int *coff, *corners, *soff, *sA, *sB;
    const int8_t* ctype;
};
void gel_mesh_mc_view(const gel_ctx* c, MeshMcView* v, std::vector<int>& code, std::vector<int>& outer,
                      std::vector<int>& inner);
namespace gel {
// One step of the general-mesh sweep (mesh.hip); moving: next node positions in a.x2, a.y2.
hipError_t gel_mesh_step(gel_ctx* c, SweepArgs& a, bool moving);
// One tracked step (tracking.hip): surface node speeds, rail redistribution into
// ctx->d_x2/d_y2, and the sweep with surface faces; asynchronous on ctx->stream.
hipError_t launch_tracked_step(gel_ctx* c, SweepArgs& a);
struct BatchItem;
// The same for every case of a batch (args[k].x2 / y2 receive case k's next nodes);
// stiles / rtiles: (case, block of 256) of the surface faces / of the nodes.
hipError_t launch_tracked_batch(const SweepArgs* args, const TrackDev* tdev, const int2* stiles, int nstiles,
                                const int2* rtiles, int nrtiles, const BatchItem* items, int nitems, int workgroups,
                                hipStream_t s, bool hllc, bool types, int src);
int resident_tracked_batch(bool types);
// One step of the plasma + hydrogen sweep (multifluid.hip): five launches, one per fluid.
hipError_t launch_mf_step(gel_ctx* c, SweepArgs& a, bool moving);
// One step of the inner-region solver (inner.hip), static grids.
hipError_t launch_inner_step(gel_ctx* c, SweepArgs& a);
}

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
    void* d_src = nullptr;         // [GEL_MK1_NFIELDS][nc] hydrogen moments and k factors, ts bytes each
    void* d_qtab = nullptr;        // [5][nc] per-cell source table (GEL_SRC_TABLE, GEL_SRC_HYBRID), ts bytes each
    int* d_qflag = nullptr;        // [nc] analytic-fallback flags (GEL_SRC_HYBRID)
    double* d_geo = nullptr;       // geometry table of the static grid (built on first use)
    double* d_mfq = nullptr;       // plasma + hydrogen: charge-exchange sources of the step
    bool geo_valid = false;        // false after the nodes move
    bool tracking = false;         // fitted-surface tracking on (gel_set_tracking)
    gel_tracking trk{};
    gel::TrackDev tdev{};
    double* d_trk = nullptr;       // tables and face contributions of tdev
    unsigned long long* d_dt = nullptr;   // [3][kDtSlot] time-step slots
    unsigned long long* d_cnt = nullptr;  // [3]
    long long steps = 0;
    bool uploaded = false;
    int rows = 8;
    // prescribed motion: nodes at base*s[k], k = step mod period
    std::vector<double> motion;
    long long motion_step = 0;
    hipStream_t stream = nullptr;
    MeshBuf* mesh = nullptr;      // a general mesh (gel_ctx_create_mesh) instead of a polar grid
    int mesh_faces = 0;
    int timing = 0;       // 1: events around every launch; 2: around every gel_step call
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<long long> event_steps;   // steps each event pair spans
};

