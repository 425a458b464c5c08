// tracking.hip - fitted-surface tracking on the polar grid: the node speeds of
// Setka::Move_surface (Setka.cpp:3378-3844), the A-rail redistribution of
// Setka::Move_Setka_Calculate_2 (Setka.cpp:4324-4381), and the sweep instantiations
// with surface faces (gas, FP64, FP contraction off as every FP64 kernel).
//
// One step is three launches on the context's stream: surface_kernel (one wavefront
// per face of the three surface lines: Riemann problem, scaled wave speed,
// contributions to the face's two nodes; the exact solver's bisection runs six of its
// iterations per evaluation over the wavefront's lanes, godunov_bisect_wave),
// rail_kernel (one thread per node: the speeds of its rail's key points, their new
// positions, its own new position by the rail spacing), and the moving sweep from the
// node positions before to those after.  Node speeds are
// combined per node in the reference's loop order, so the result does not depend on
// scheduling; no atomics, no memset.
#include "sweep_kernel.h"

namespace gel {

// Both one-sided face values of surface face j of node row k (master cell (k-1, j)
// below, Gran_copy side (k, j) above); the inner shock's master side is the cell value
// rescaled to the face (Setka.cpp:3560-3575).
__device__ void surface_sides(const SweepArgs& a, int line, int k, int j, double fx, double fy, double gam,
                              double* p1, double* p2)
{
    const int nj = a.nj;
    const size_t nc = (size_t)a.ni * nj;
    const double* st = (const double*)a.now;
    CellT<GasPhys, double> C[4];             // cells (k-2, j), (k-1, j), (k, j), (k+1, j)
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = k - 2 + q;
        const size_t n0 = (size_t)i * (nj + 1) + j, n1 = n0 + (nj + 1);
        const double px[4] = { a.x[n0], a.x[n1], a.x[n1 + 1], a.x[n0 + 1] };
        const double py[4] = { a.y[n0], a.y[n1], a.y[n1 + 1], a.y[n0 + 1] };
        quad_center(px, py, C[q].cx, C[q].cy);
        const size_t c = (size_t)i * nj + j;
#pragma unroll
        for (int v = 0; v < 5; v++) C[q].s[v] = st[v * nc + c];
        C[q].type = GEL_C_2;
        C[q].ok = 1;
    }
    if (line == 1) {
#pragma unroll
        for (int v = 0; v < 5; v++) p1[v] = C[1].s[v];
        const double radius = sqrt(sq(C[1].cx) + sq(C[1].cy));
        const double dis = sqrt(sq(fx) + sq(fy));
        p1[0] = p1[0] * sq(radius) / sq(dis);
        p1[1] = p1[1] * pow(radius / dis, 2.0 * gam);
        polar_perenos(C[1].cx, C[1].cy, fx, fy, p1[2], p1[3]);
    } else {
        tvd_surface<GasPhys>(C[1], C[0], fx, fy, p1);
    }
    tvd_surface<GasPhys>(C[2], C[3], fx, fy, p2);
}

// face[line][j][6] = contribution to node A (x, y, count), to node B (x, y, count);
// line 0 contact, 1 inner shock, 2 outer shock (the order of Move_surface's loops).
// One face per wavefront: every lane computes the same values (the exact solver's
// bisection spreads its midpoints over the lanes, godunov_exact<true>), lane 0 stores.
__device__ __forceinline__ void surface_face(const SweepArgs& a, const TrackDev& t, int g)
{
    const int nj = a.nj;
    if (g >= 3 * nj) return;
    const bool lead = (threadIdx.x & (kWave - 1)) == 0;
    const int line = g / nj, j = g - line * nj;
    double r[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
    if (!(line == 2 && j >= t.outer_faces)) {             // Setka.cpp:3746
        const int k = t.rows[line];
        const size_t ia = (size_t)k * (nj + 1) + j, ib = ia + 1;
        const double fx = (a.x[ia] + a.x[ib]) / 2.0, fy = (a.y[ia] + a.y[ib]) / 2.0;
        double t1 = a.x[ib] - a.x[ia], t2 = a.y[ib] - a.y[ia];   // Gran::Get_normal
        const double nn = sqrt(sq(t1) + sq(t2));
        t1 = t1 / nn; t2 = t2 / nn;
        const double n1 = t2, n2 = -t1;
        const double gam = a.prm.gamma;
        double p1[5], p2[5];
        surface_sides(a, line, k, j, fx, fy, gam, p1, p2);
        const double q1[5] = { p1[0], p1[1], p1[2], p1[3], 0.0 };
        const double q2[5] = { p2[0], p2[1], p2[2], p2[3], 0.0 };
        double q[5], Vl = 0.0, Vp = 0.0, VV = 0.0;
        int st = 0;
        godunov_exact<true>(q1, q2, n1, n2, 0.0, 0.0, 0.0, 0.1, gam, q, Vl, Vp, VV, st);
        if (st && lead) atomicAdd(&a.counters[2], 1ull);
        const double tt1 = -n2, tt2 = n1;
        bool toA = false, toB = false;
        double V = 0.0;
        if (line == 0) {                                   // Setka.cpp:3434-3470
            V = VV * t.koef * t.k_line[0];
            toB = true;
        } else if (line == 1) {                            // :3595-3616
            V = Vl * t.koef * t.k_line[1];
            toB = true;
            toA = fabs(p2[2] * tt1 + p2[3] * tt2) < t.tang_inner;
        } else {                                           // :3782-3825
            V = Vp * t.koef * t.k_line[2];
            const double ut = p1[2] * tt1 + p1[3] * tt2;
            if ((sq(p1[2]) + sq(p1[3]) + sq(0.0)) / (gam * p1[1] / p1[0]) > 1.0) toB = true;
            else if (fabs(ut) < t.tang_outer) { toA = true; toB = true; }
            else if (ut > 0) toB = true;
            else toA = true;
        }
        if (toA) { r[0] = V * n1; r[1] = V * n2; r[2] = 1.0; }
        if (toB) { r[3] = V * n1; r[4] = V * n2; r[5] = 1.0; }
    }
    if (!lead) return;
    double* o = t.face + (size_t)g * 6;
#pragma unroll
    for (int q = 0; q < 6; q++) o[q] = r[q];
}
__global__ void __launch_bounds__(256) surface_kernel(SweepArgs a, TrackDev t)
{
    surface_face(a, t, blockIdx.x * 4 + (int)(threadIdx.x / kWave));
}

// Sum of the contributions to node c of a line, in the reference's loop order (face
// c-1 as its node B, then face c as its node A), and their count.
__device__ __forceinline__ void raw_speed(const TrackDev& t, int nj, int line, int c, double& vx, double& vy, int& n)
{
    vx = 0.0; vy = 0.0; n = 0;
    const double* F = t.face + (size_t)line * nj * 6;
    if (c >= 1) {
        const double* f = F + (size_t)(c - 1) * 6;
        if (f[5] > 0.0) { vx = vx + f[3]; vy = vy + f[4]; n++; }
    }
    if (c < nj) {
        const double* f = F + (size_t)c * 6;
        if (f[2] > 0.0) { vx = vx + f[0]; vy = vy + f[1]; n++; }
    }
}

// After the line's own averaging (Setka.cpp:3522-3530 for the contact; the shocks'
// loops average the contact line again instead of their own, :3718-3731, :3829-3837).
__device__ __forceinline__ void first_speed(const TrackDev& t, int nj, int line, int c, double& vx, double& vy,
                                            int& n)
{
    raw_speed(t, nj, line, c, vx, vy, n);
    if ((line == 0 || t.average_lines) && c < nj && n > 0) {
        vx /= (1.0 * n);
        vy /= (1.0 * n);
    }
}

// Final speed of node c of a line: end-node corrections (Setka.cpp:3532-3535,
// :3733-3736, :3839-3842), then the contact's second and third averaging.
__device__ void line_speed(const SweepArgs& a, const TrackDev& t, int line, int c, double dt, double& vx,
                           double& vy)
{
    const int nj = a.nj;
    const size_t base = (size_t)t.rows[line] * (nj + 1);
    int n;
    first_speed(t, nj, line, c, vx, vy, n);
    if (c == 0) {
        double ux, uy;
        int m;
        first_speed(t, nj, line, 1, ux, uy, m);
        vx = ux + (a.x[base + 1] - a.x[base]) / dt;
    }
    if (c == nj) {
        double ux, uy;
        int m;
        first_speed(t, nj, line, nj - 1, ux, uy, m);
        if (line == 1) vx = ux + (a.x[base + nj - 1] - a.x[base + nj]) / dt;
        else vy = uy + (a.y[base + nj - 1] - a.y[base + nj]) / dt;
    }
    if (line == 0 && !t.average_lines && c < nj && n > 0) {
        vx /= (1.0 * n); vy /= (1.0 * n);
        vx /= (1.0 * n); vy /= (1.0 * n);
    }
}

// Node (r, c) after the step: the rail's key points move along the rail by dt times
// their speed's rail component, the other nodes follow the rail spacing
// (Setka.cpp:4328-4381).  Power-law factors come precomputed from the host.
__device__ __forceinline__ void rail_node(const SweepArgs& a, const TrackDev& t, double* __restrict__ X2,
                                          double* __restrict__ Y2, long long g)
{
    const int ni = a.ni, nj = a.nj;
    if (g >= (long long)(ni + 1) * (nj + 1)) return;
    const int r = (int)(g / (nj + 1)), c = (int)(g % (nj + 1));
    const double dt = dt_read(a);
    const double cs = t.cs[c], sn = t.sn[c];
    const int key_row[3] = { t.rows[1], t.rows[0], t.rows[2] };   // inner shock, contact, outer shock
    const int key_line[3] = { 1, 0, 2 };
    double kx[3], ky[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        double vx, vy;
        line_speed(a, t, key_line[q], c, dt, vx, vy);
        const double V = vx * cs + vy * sn;
        const size_t id = (size_t)key_row[q] * (nj + 1) + c;
        kx[q] = a.x[id] + dt * V * cs;
        ky[q] = a.y[id] + dt * V * sn;
    }
    double x2, y2;
    if (r == key_row[0]) { x2 = kx[0]; y2 = ky[0]; }
    else if (r == key_row[1]) { x2 = kx[1]; y2 = ky[1]; }
    else if (r == key_row[2]) { x2 = kx[2]; y2 = ky[2]; }
    else {
        const double R2 = sqrt(sq(kx[0]) + sq(ky[0]));
        const double R3 = sqrt(sq(kx[1]) + sq(ky[1]));
        const double R4 = sqrt(sq(kx[2]) + sq(ky[2]));
        const int M1 = key_row[0] + 1, M2 = key_row[1] - key_row[0], M3 = key_row[2] - key_row[1];
        double rr;
        if (r < t.n_inner) rr = t.r_inner + t.f_in[r] * (t.r_mid - t.r_inner);
        else if (r < M1 - 1) rr = t.r_mid + t.f_mid[r - t.n_inner] * (R2 - t.r_mid);
        else if (r < M1 + M2 - 1) rr = R2 + (R3 - R2) * (r - M1 + 1) / (M2);
        else if (r < M1 + M2 + M3 - 1) rr = R3 + (R4 - R3) * (r - M1 - M2 + 1) / (M3);
        else rr = R4 + t.f_out[(size_t)c * t.m4 + (r - M1 - M2 - M3)] * (t.r_outer - R4);
        x2 = rr * cs;
        y2 = rr * sn;
    }
    X2[g] = x2;
    Y2[g] = y2;
}
__global__ void __launch_bounds__(256) rail_kernel(SweepArgs a, TrackDev t, double* __restrict__ X2,
                                                   double* __restrict__ Y2)
{
    rail_node(a, t, X2, Y2, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// The same launches for every case of a batch: work-group g takes block tiles[g].y of
// case tiles[g].x (four surface faces, or 256 nodes); the rail kernel writes each
// case's next node positions into its argument set's x2, y2.
__global__ void __launch_bounds__(256) surface_batch_kernel(const SweepArgs* __restrict__ args,
                                                            const TrackDev* __restrict__ tdev,
                                                            const int2* __restrict__ tiles)
{
    const int2 b = tiles[blockIdx.x];
    surface_face(args[b.x], tdev[b.x], b.y * 4 + (int)(threadIdx.x / kWave));
}
__global__ void __launch_bounds__(256) rail_batch_kernel(const SweepArgs* __restrict__ args,
                                                         const TrackDev* __restrict__ tdev,
                                                         const int2* __restrict__ tiles)
{
    const int2 b = tiles[blockIdx.x];
    const SweepArgs& a = args[b.x];
    rail_node(a, tdev[b.x], (double*)a.x2, (double*)a.y2, (long long)b.y * 256 + threadIdx.x);
}

template <bool TYPES, int FX>
static hipError_t tracked_sweep(gel_ctx* c, SweepArgs& a, hipStream_t s)
{
    if (c->prm.source_mode == GEL_SRC_MK1)
        return launch_t<GasPhys, double, double, true, TYPES, GEL_SRC_MK1, FX | kTrackFlux>(c, a, s);
    return launch_t<GasPhys, double, double, true, TYPES, GEL_SRC_NONE, FX | kTrackFlux>(c, a, s);
}

template <bool TYPES, int FX>
static hipError_t tracked_batch_sweep(const SweepArgs* args, const BatchItem* items, int nitems, int workgroups,
                                      hipStream_t s, int src)
{
    if (src == GEL_SRC_MK1)
        return batch_launch_t<GasPhys, double, double, GEL_SRC_MK1, FX | kTrackFlux, true, TYPES>(args, items, nitems,
                                                                                               workgroups, s);
    return batch_launch_t<GasPhys, double, double, GEL_SRC_NONE, FX | kTrackFlux, true, TYPES>(args, items, nitems,
                                                                                            workgroups, s);
}

hipError_t launch_tracked_batch(const SweepArgs* args, const TrackDev* tdev, const int2* stiles, int nstiles,
                                const int2* rtiles, int nrtiles, const BatchItem* items, int nitems, int workgroups,
                                hipStream_t s, bool hllc, bool types, int src)
{
    hipLaunchKernelGGL(surface_batch_kernel, dim3((unsigned)nstiles), dim3(256), 0, s, args, tdev, stiles);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(rail_batch_kernel, dim3((unsigned)nrtiles), dim3(256), 0, s, args, tdev, rtiles);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (types)
        return hllc ? tracked_batch_sweep<true, GEL_FLUX_HLLC>(args, items, nitems, workgroups, s, src)
                    : tracked_batch_sweep<true, GEL_FLUX_HLL>(args, items, nitems, workgroups, s, src);
    return hllc ? tracked_batch_sweep<false, GEL_FLUX_HLLC>(args, items, nitems, workgroups, s, src)
                : tracked_batch_sweep<false, GEL_FLUX_HLL>(args, items, nitems, workgroups, s, src);
}

int resident_tracked_batch(bool types)
{
    const void* kern =
        types ? (const void*)sweep_batch_kernel<GasPhys, double, double, GEL_SRC_NONE, GEL_FLUX_HLLC | kTrackFlux, kNW, true, true>
              : (const void*)sweep_batch_kernel<GasPhys, double, double, GEL_SRC_NONE, GEL_FLUX_HLLC | kTrackFlux, kNW, true, false>;
    return resident_groups(kern, kNW * kWave);
}

hipError_t launch_tracked_step(gel_ctx* c, SweepArgs& a)
{
    hipStream_t s = c->stream;
    const int nj = a.nj;
    const long long nn = (long long)(a.ni + 1) * (nj + 1);
    const dim3 gs((unsigned)((3 * nj + 3) / 4));          // four surface faces per work-group
    hipLaunchKernelGGL(surface_kernel, gs, dim3(256), 0, s, a, c->tdev);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(rail_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, a, c->tdev, c->d_x2, c->d_y2);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    a.x2 = c->d_x2;
    a.y2 = c->d_y2;
    const bool hllc = c->prm.flux == GEL_FLUX_HLLC;
    if (c->has_types) return hllc ? tracked_sweep<true, GEL_FLUX_HLLC>(c, a, s) : tracked_sweep<true, GEL_FLUX_HLL>(c, a, s);
    return hllc ? tracked_sweep<false, GEL_FLUX_HLLC>(c, a, s) : tracked_sweep<false, GEL_FLUX_HLL>(c, a, s);
}

}  // namespace gel
