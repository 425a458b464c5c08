// mesh.hip - the sweep of Setka::Go_5_komponent__MK2 (Setka.cpp:10121-10482) over a
// general mesh: the reference's own block-structured heliosphere grid (rails of four
// kinds, fitted-surface rows, C_centr wedges, Setka.cpp:19-1076) with the face tables
// Setka::TVD_prepare leaves (Setka.cpp:1294-1434).  Gas, FP64, HLL / HLLC.
//
// A step is three launches on the context's stream:
//   mesh_geo_kernel   one thread per cell: centroid (Cell::Get_Center), planar areas
//                     before and after the step (Cell::Get_Volume, Get_Volume_posle:
//                     quads by their diagonals, other polygons by Heron triangles), and
//                     the r-weighted volumes (Get_Volume_rotate / 2 pi); on a static
//                     grid once, on a moving step every step;
//   mesh_face_kernel  one thread per face (a face with two cells once, seen from its
//                     cell of lower index; one wavefront per fitted-surface face, whose
//                     exact solver's bisection spreads over the lanes, in the same
//                     launch): both sides' values (Gran::Get_par,
//                     Get_par_TVD, Get_par_TVD_radial, the C_1 rescaling, fitted-surface
//                     one-sided lines), the face flux (HLL / HLLC, the exact solver on
//                     faces between two nodes of one fitted surface, nul_potok beside
//                     them), flux x area and the largest wave speed;
//   mesh_cell_kernel  one thread per cell: the faces' products summed in Cell::Grans
//                     order (the other cell of a face takes the exact negation), the
//                     CFL time, the ALE update with the MK1 source, the clamps, and the
//                     step's time-step minimum (wave min + one integer atomicMin into one of the
//                     slot's entries).
// The per-face and per-cell arithmetic is that of the structured sweep (sweep_kernel.h,
// sweep_core.h), so the two agree where a polar grid is given in both forms.
#include "sweep_kernel.h"

#include <algorithm>
#include <map>
#include <string>
#include <tuple>

namespace gel {

struct MeshDev {
    int nc, nf;
    const int* coff;      // [nc+1] corner offsets
    const int* corners;   // Cell::contour
    const int8_t* ctype;  // [nc] Cell::type
    const int8_t* caxis;  // [nc] 1: the cell has an Axis face
    const int* soff;      // [nc+1] face-slot offsets (Cell::Grans)
    const int* sface;     // [slots] face index, or ~face where the cell is not the face's owner
    const int* sA;        // [slots] Gran::A, B as the cell sees them (the Heron fan)
    const int* sB;
    const int* fA;        // [nf] nodes as the owner sees them
    const int* fB;
    const int* ftype;     // Gran_type
    const int* fown;      // owner cell
    const int* fnbr;      // the other cell (-1 on boundaries)
    const int* fdown;     // Sosed_down of the owner's slot (the owner's stencil cell)
    const int* fup;       // Sosed_up of the owner's slot (the other cell's stencil cell)
    const int8_t* fsurf;  // bit 0: both nodes on one fitted surface; bit 1: exactly one
    int ng;               // faces with fsurf bit 0 (the exact solver's faces)
    const int* gface;     // [ng] their indices
    double* cgeo;         // [6][nc] cx, cy, vs1, vs2, Volume, Volume2
    double* fout;         // [8][nf] flux x area (5), max wave speed, face centre x, y
};

// Cell::Get_Volume_rotate(360) / (2 pi) over a contour of n corners.
__device__ __forceinline__ double poly_volume_r(const double* X, const double* Y, const int* q, int n)
{
    double A = 0.0, G = 0.0;
    for (int k = 0; k < n; k++) {
        const int a = q[k], b = q[k + 1 == n ? 0 : k + 1];
        const double cr = X[a] * Y[b] - Y[a] * X[b];
        A = A + cr;
        G = G + (Y[a] + Y[b]) * cr;
    }
    A = A * 0.5;
    return fabs(A) * (1.0 / (6.0 * A)) * G * kPi * 360.0 / 180.0 / (2.0 * kPi);
}

__device__ __forceinline__ void poly_center(const double* X, const double* Y, const int* q, int n, double& cx,
                                            double& cy)
{
    double A = 0.0, Gx = 0.0, Gy = 0.0;
    for (int k = 0; k < n; k++) {
        const int a = q[k], b = q[k + 1 == n ? 0 : k + 1];
        const double cr = X[a] * Y[b] - Y[a] * X[b];
        A = A + cr;
        Gx = Gx + (X[a] + X[b]) * cr;
        Gy = Gy + (Y[a] + Y[b]) * cr;
    }
    A = A * 0.5;
    cx = Gx / (6.0 * A);
    cy = Gy / (6.0 * A);
}

// Heron triangles (face, centre (x, y)) summed over the cell's faces (Cell.cpp:167-186).
__device__ __forceinline__ double poly_heron(const MeshDev& m, const double* X, const double* Y, int c, double x,
                                             double y)
{
    double S = 0.0;
    for (int k = m.soff[c]; k < m.soff[c + 1]; k++) {
        const int A = m.sA[k], B = m.sB[k];
        const double a1 = sqrt(sq(X[A] - x) + sq(Y[A] - y));
        const double a2 = sqrt(sq(X[B] - x) + sq(Y[B] - y));
        const double a3 = sqrt(sq(X[A] - X[B]) + sq(Y[A] - Y[B]));
        const double p = (a1 + a2 + a3) / 2.0;
        S = S + sqrt(p * (p - a1) * (p - a2) * (p - a3));
    }
    return S;
}

__global__ void __launch_bounds__(256) mesh_geo_kernel(MeshDev m, const double* __restrict__ X,
                                                       const double* __restrict__ Y, const double* __restrict__ X2,
                                                       const double* __restrict__ Y2)
{
    const int c = xcd_item(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= m.nc) return;
    const int* q = m.corners + m.coff[c];
    const int n = m.coff[c + 1] - m.coff[c];
    double cx, cy, vs1, vs2;
    if (n == 4) {
        const double px[4] = { X[q[0]], X[q[1]], X[q[2]], X[q[3]] };
        const double py[4] = { Y[q[0]], Y[q[1]], Y[q[2]], Y[q[3]] };
        const double px2[4] = { X2[q[0]], X2[q[1]], X2[q[2]], X2[q[3]] };
        const double py2[4] = { Y2[q[0]], Y2[q[1]], Y2[q[2]], Y2[q[3]] };
        quad_center(px, py, cx, cy);
        vs1 = quad_area(px, py);
        vs2 = quad_area(px2, py2);
    } else {
        poly_center(X, Y, q, n, cx, cy);
        vs1 = poly_heron(m, X, Y, c, cx, cy);
        double ax = 0.0, ay = 0.0;                          // Get_Volume_posle: the corners' mean
        for (int k = 0; k < n; k++) { ax += X2[q[k]]; ay += Y2[q[k]]; }
        ax = ax / (1.0 * n);
        ay = ay / (1.0 * n);
        vs2 = poly_heron(m, X2, Y2, c, ax, ay);
    }
    const size_t nc = m.nc;
    m.cgeo[c] = cx;
    m.cgeo[nc + c] = cy;
    m.cgeo[2 * nc + c] = vs1;
    m.cgeo[3 * nc + c] = vs2;
    m.cgeo[4 * nc + c] = poly_volume_r(X, Y, q, n);
    m.cgeo[5 * nc + c] = poly_volume_r(X2, Y2, q, n);
}

// A stencil cell as the face functions of sweep_kernel.h see it; rp (the centroid
// radius to the power 2 gamma) only where a C_1 side's radial reconstruction reads it.
// TS: the state's storage type (FP32-bulk contexts store it in FP32; the cell's values
// are widened, the reconstruction runs in FP64).
template <class TS = double>
__device__ __forceinline__ CellT<GasPhys, double> mesh_cell(const MeshDev& m, const SweepArgs& a, int c, bool rp)
{
    CellT<GasPhys, double> k;
    if (c < 0) {
        k.ok = 0;
        k.type = GEL_C_2;
#pragma unroll
        for (int v = 0; v < 5; v++) k.s[v] = 0.0;
        k.s[0] = 1.0; k.s[1] = 1.0;
        k.cx = 1.0; k.cy = 1.0; k.rp = 0.0; k.ca = 1.0; k.sa = 0.0;
        return k;
    }
    const size_t nc = m.nc;
    const TS* now = (const TS*)a.now;
#pragma unroll
    for (int v = 0; v < 5; v++) k.s[v] = (double)now[v * nc + c];
    k.cx = m.cgeo[c];
    k.cy = m.cgeo[nc + c];
    k.type = m.ctype[c];
    k.ok = 1;
    k.rp = rp ? pow(sqrt(sq(k.cx) + sq(k.cy)), 2.0 * a.prm.gamma) : 0.0;
    k.ca = 1.0; k.sa = 0.0;
    if (k.type == GEL_C_1) {
        const double a1 = polar_angle(k.cx, k.cy);
        k.ca = cos(a1);
        k.sa = sin(a1);
    }
    return k;
}

// Value of cell M on its side of an interior face (Setka.cpp:10213-10251): C_1 cells
// rescaled and rotated with the radial reconstruction; a face between two nodes of one
// fitted surface the one-sided line through D and M (Gran.cpp:345-426); other faces the
// limited line of Gran::Get_par_TVD.
__device__ __forceinline__ void mesh_side(const CellT<GasPhys, double>& M, const CellT<GasPhys, double>& N,
                                          const CellT<GasPhys, double>& D, bool m_axis, bool god, const Face& f,
                                          double gam, double* r, int& bad)
{
    if (M.type == GEL_C_1) side_value<GasPhys, double, double, true>(M, N, D, m_axis, f.fx, f.fy, f.dis, gam, r, bad);
    else if (god && D.ok) tvd_surface<GasPhys>(M, D, f.fx, f.fy, r);
    else tvd_usualy<GasPhys, double, double>(M, N, D, m_axis, f.fx, f.fy, r, bad);
}

// Face f: both sides' values, the flux, flux x area, the wave speed and the face
// centre into fout; returns the numeric-failure count.  WAVE: the whole wavefront
// evaluates this one face (a fitted-surface face, whose exact solver's bisection then
// spreads its midpoints over the lanes), lane 0 stores.  TS = float (FP32 bulk): the
// state read from FP32 storage and, on faces other than the fitted-surface and
// nul_potok ones, the Riemann solver in FP32 on the FP64 area-weighted normal and
// swept volume of face_geom_a (supersonic faces the FP64 upwind flux, as on polar grids).
template <int FX, bool WAVE, class TS = double>
__device__ __forceinline__ int mesh_face(const MeshDev& m, const SweepArgs& a, int f)
{
    int b = 0;
    const double gam = a.prm.gamma;
    const double dt = dt_read(a);
    const int M = m.fown[f], N = m.fnbr[f], t = m.ftype[f];
    const int A = m.fA[f], B = m.fB[f];
    const Face g = face_geom<true, double>(a.x[A], a.y[A], a.x[B], a.y[B], a.x2[A], a.y2[A], a.x2[B], a.y2[B], dt);
    const bool c1 = m.ctype[M] == GEL_C_1 || (N >= 0 && m.ctype[N] == GEL_C_1);
    const CellT<GasPhys, double> Mc = mesh_cell<TS>(m, a, M, c1), Dc = mesh_cell<TS>(m, a, m.fdown[f], c1);
    const bool god = (m.fsurf[f] & 1) != 0, np = (m.fsurf[f] & 2) != 0;
    double L[5], R[5], P[5];
    if (t == GEL_F_USUALY) {
        const CellT<GasPhys, double> Nc = mesh_cell<TS>(m, a, N, c1), Uc = mesh_cell<TS>(m, a, m.fup[f], c1);
        mesh_side(Mc, Nc, Dc, m.caxis[M] != 0, god, g, gam, L, b);
        mesh_side(Nc, Mc, Uc, m.caxis[N] != 0, god, g, gam, R, b);
    } else {
        boundary_states<GasPhys, double, double, true>(a, t, Mc, Dc, g, L, R);
    }
    double uu, S = g.S;
    if (god) {
        uu = surface_flux<FX == GEL_FLUX_HLLC, WAVE>(L, R, g, gam, P, b);
    } else if (FX == GEL_FLUX_HLLC && np) {
        uu = nul_potok_flux(L, R, g, gam, P);
    } else if constexpr (sizeof(TS) == 4) {
        const Face ga = face_geom_a<true>(a.x[A], a.y[A], a.x[B], a.y[B], a.x2[A], a.y2[A], a.x2[B], a.y2[B],
                                          1.0 / dt);
        uu = face_flux<GasPhys, float, FX>(L, R, ga, gam, P);   // already times the area
        S = 1.0;
    } else {
        uu = face_flux<GasPhys, double, FX>(L, R, g, gam, P);
    }
    if (WAVE && (threadIdx.x & (kWave - 1)) != 0) return 0;
    const size_t nf = m.nf;
#pragma unroll
    for (int k = 0; k < 5; k++) m.fout[k * nf + f] = P[k] * S;
    m.fout[5 * nf + f] = uu;
    m.fout[6 * nf + f] = g.fx;
    m.fout[7 * nf + f] = g.fy;
    return b;
}

// Work-groups [0, nfb) take 256 faces each, one per lane, skipping the fitted-surface
// faces; the work-groups after them take four fitted-surface faces each, one per
// wavefront.  nfb is a multiple of eight (xcd_item).
template <int FX, class TS>
__global__ void __launch_bounds__(256) mesh_face_kernel(MeshDev m, SweepArgs a, int nfb)
{
    int b = 0;
    if ((int)blockIdx.x < nfb) {
        const int f = xcd_item(blockIdx.x, nfb) * blockDim.x + threadIdx.x;
        if (f < m.nf && !(m.fsurf[f] & 1)) b = mesh_face<FX, false, TS>(m, a, f);
    } else {
        const int k = ((int)blockIdx.x - nfb) * 4 + (int)(threadIdx.x / kWave);
        if (k < m.ng) b = mesh_face<FX, true, TS>(m, a, m.gface[k]);
    }
    const int s_bad = wave_sum(b);
    if ((threadIdx.x & (kWave - 1)) == 0 && s_bad)
        __hip_atomic_fetch_add(gp(a.counters) + 2, (unsigned long long)s_bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int SRC, class TS>
__global__ void __launch_bounds__(256) mesh_cell_kernel(MeshDev m, SweepArgs a)
{
    const int c = xcd_item(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0) dt_clear(a);
    MarchAcc acc{ __builtin_huge_val(), 0, 0, 0 };
    const size_t nc = m.nc, nf = m.nf;
    if (c < m.nc) {
        const double cx = m.cgeo[c], cy = m.cgeo[nc + c];
        const double radius = sqrt(sq(cx) + sq(cy));
        if (!(radius < a.prm.r_skip)) {                     // Setka.cpp:10169-10172
            double Pot[5];
#pragma unroll
            for (int k = 0; k < 5; k++) Pot[k] = 0.0;
            double time = 10000000;
            for (int k = m.soff[c]; k < m.soff[c + 1]; k++) {
                const int e = m.sface[k];
                const int f = e >= 0 ? e : ~e;
#pragma unroll
                for (int v = 0; v < 5; v++) {
                    const double ps = m.fout[v * nf + f];
                    Pot[v] = Pot[v] + (e >= 0 ? ps : -ps);
                }
                Face g;
                g.fx = m.fout[6 * nf + f];
                g.fy = m.fout[7 * nf + f];
                time = mn(time, cfl_time<double>(a.prm.cfl, cx, cy, g, m.fout[5 * nf + f]));
            }
            acc.tmin = fmin(acc.tmin, time);
            const TS* now = (const TS*)a.now;
            TS* nxt = (TS*)a.next;
            double s[5];
#pragma unroll
            for (int v = 0; v < 5; v++) s[v] = (double)now[v * nc + c];
            if (m.ctype[c] == GEL_C_CENTR) {                // Setka.cpp:10461-10468
#pragma unroll
                for (int v = 0; v < 5; v++) nxt[v * nc + c] = now[v * nc + c];
            } else {
                double qm = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0, qQ = 0.0;
                if constexpr (SRC == GEL_SRC_MK1) {
                    const TS* sp = (const TS*)a.src + c;
                    auto fk = [&](int k) { return (double)sp[(size_t)k * nc]; };
                    source_mk1<decltype(fk), double>(a.prm, fk, s[2], s[3], s[0], s[1], q1, q2, q3);
                } else if constexpr (SRC == GEL_SRC_TABLE) {   // mass, momentum x, y, energy, tracer
                    const TS* qt = (const TS*)a.qtab + c;
                    qm = (double)qt[0];
                    q1 = (double)qt[nc];
                    q2 = (double)qt[2 * nc];
                    q3 = (double)qt[3 * nc];
                    qQ = (double)qt[4 * nc];
                }
                const double dt = dt_read(a);
                double o[5];
                cell_update<GasPhys, false>(s, Pot, m.cgeo[4 * nc + c], m.cgeo[5 * nc + c], m.cgeo[2 * nc + c],
                                            m.cgeo[3 * nc + c], dt, a.prm.gamma, qm, q1, q2, q3, qQ,
                                            SRC == GEL_SRC_TABLE, o, acc.n_rho, acc.n_p);
#pragma unroll
                for (int v = 0; v < 5; v++) nxt[v * nc + c] = (TS)o[v];
            }
        }
    }
    march_flush(a, acc);
}

// ---- fitted-surface tracking on the general mesh --------------------------------
// Setka::Move_surface (Setka.cpp:3378-3844) over the three lines given as face lists
// and Setka::Move_Setka_Calculate_2 (Setka.cpp:4324-4759) over the four rail
// families.  mtrk_face_kernel solves the lines' faces, one wavefront per face (the
// Riemann problem and the face's contributions to its two nodes); mtrk_move_kernel, one
// wavefront, replays the reference's node-speed sums, averages and end corrections in
// its statement order (the line nodes' speeds in LDS), then the rails in the
// reference's order A, B, C, D, a rail's key points uniformly and each of its spacing
// loops across the lanes, so that every node's last assignment is the reference's.
struct MeshTrk {
    int nlf;                  // line faces: Line_Contact, Line_Inner, Line_Outer concatenated
    int nline[3];
    int nln;                  // nodes of the lines (compact slots)
    int nrail, nfree;
    int n_inner, d_chain, outer_faces, average_lines;
    const int* sum_off;       // [4] per line, into sums
    const int* sums;          // [][3] line slot, its contributions (2 face + role, -1)
    const int* avg_off;       // [4] per line, into avgs
    const int* avgs;          // the line slots each averaging loop divides, in order
    int nphase;               // runs of rails that run side by side
    const int* phase;         // [nphase + 1] first rail of each run
    int d_chain_rail;         // the D rail whose outer-shock position the chain copies, -1
    double koef, k_line[3], tang_inner, tang_outer, r_inner, r_mid, r_outer, left;
    const int* lf;            // [nlf][6] A, B (nodes), M, N (Master, Sosed), D, U (their Sosed_down)
    const int* lslot;         // [nlf][2] line slots of A, B
    const int* node_slot;     // [nn] line slot of a node, -1
    const int* rail;          // [nrail][kRailInts]
    const double* rcs;        // [nrail][2] cos s, sin s (host libm, as the reference calls them)
    const double* fin;        // [n_inner] (j / n_inner)^p_inner
    const double* fpow;       // per rail: the mid block's and (A rails) the outer block's factors
    const int* allp;          // Rail::All_point, concatenated
    const int* freep;         // [nfree] nodes on no rail (their next position is their position)
    double* contrib;          // [nlf][6] to A (vx, vy, count), to B (vx, vy, count)
};
// rail row: kind, M1, M2, M3, M4, offset and count in allp, key points k0..k3, offsets
// of its mid and outer factors in fpow, 1 for a D rail that copies the outer-shock
// position of the rail before, the key points the rail stores (bit k: key point k is not
// overwritten by the rail's own spacing loops, which then place it)
constexpr int kRailInts = 15;
constexpr int kMaxLineNodes = 1024;
constexpr int kMaxLineFaces = 1024;
constexpr int kMoveThreads = 512;

// Cell::Get_Center of cell c at node positions X, Y (as mesh_geo_kernel computes it).
__device__ __forceinline__ void mesh_center(const MeshDev& m, const double* X, const double* Y, int c, double& cx,
                                            double& cy)
{
    const int* q = m.corners + m.coff[c];
    const int n = m.coff[c + 1] - m.coff[c];
    if (n == 4) {
        const double px[4] = { X[q[0]], X[q[1]], X[q[2]], X[q[3]] };
        const double py[4] = { Y[q[0]], Y[q[1]], Y[q[2]], Y[q[3]] };
        quad_center(px, py, cx, cy);
    } else {
        poly_center(X, Y, q, n, cx, cy);
    }
}

// Line face g: Gran::Get_par_TVD of both sides (Gran.cpp:345-426: the one-sided line
// through Sosed_down and Master; the inner shock's Master side is the cell value
// rescaled to the face, Setka.cpp:3553-3575), the exact solver (Solvers.cpp:693-1172),
// and the line's rules for its node speeds (Setka.cpp:3394-3470, :3576-3616,
// :3746-3825).  One wavefront per face, every lane with the same values; lane 0 stores.
__global__ void __launch_bounds__(256) mtrk_face_kernel(MeshDev m, MeshTrk t, SweepArgs a)
{
    const int g = blockIdx.x * 4 + (int)(threadIdx.x / kWave);
    if (g >= t.nlf) return;
    const int line = g < t.nline[0] ? 0 : (g < t.nline[0] + t.nline[1] ? 1 : 2);
    const int j = g - (line == 0 ? 0 : (line == 1 ? t.nline[0] : t.nline[0] + t.nline[1]));
    double r[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
    if (!(line == 2 && j >= t.outer_faces)) {              // Setka.cpp:3746
        const int* f = t.lf + (size_t)g * 6;
        const int A = f[0], B = f[1];
        const double* X = a.x;
        const double* Y = a.y;
        const double gam = a.prm.gamma;
        const size_t nc = m.nc;
        const double* st = (const double*)a.now;
        CellT<GasPhys, double> C[4];                       // Master, Sosed_down; Sosed, its Sosed_down
        const int cid[4] = { f[2], f[4], f[3], f[5] };
#pragma unroll
        for (int k = 0; k < 4; k++) {
            mesh_center(m, X, Y, cid[k], C[k].cx, C[k].cy);
#pragma unroll
            for (int v = 0; v < 5; v++) C[k].s[v] = st[v * nc + cid[k]];
            C[k].type = GEL_C_2;
            C[k].ok = 1;
        }
        const double fx = (X[A] + X[B]) / 2.0, fy = (Y[A] + Y[B]) / 2.0;
        double t1 = X[B] - X[A], t2 = Y[B] - Y[A];          // Gran::Get_normal
        const double nn = sqrt(sq(t1) + sq(t2));
        t1 = t1 / nn; t2 = t2 / nn;
        const double n1 = t2, n2 = -t1;
        double p1[5], p2[5];
        if (line == 1) {
#pragma unroll
            for (int v = 0; v < 5; v++) p1[v] = C[0].s[v];
            const double radius = sqrt(sq(C[0].cx) + sq(C[0].cy));
            const double dis = sqrt(sq(fx) + sq(fy));
            p1[0] = p1[0] * sq(radius) / sq(dis);
            p1[1] = p1[1] * pow(radius / dis, 2.0 * gam);
            polar_perenos(C[0].cx, C[0].cy, fx, fy, p1[2], p1[3]);
        } else {
            tvd_surface<GasPhys>(C[0], C[1], fx, fy, p1);
        }
        tvd_surface<GasPhys>(C[2], C[3], fx, fy, p2);
        const double q1[5] = { p1[0], p1[1], p1[2], p1[3], 0.0 };
        const double q2[5] = { p2[0], p2[1], p2[2], p2[3], 0.0 };
        double q[5], Vl = 0.0, Vp = 0.0, VV = 0.0;
        int stt = 0;
        godunov_exact<true>(q1, q2, n1, n2, 0.0, 0.0, 0.0, 0.1, gam, q, Vl, Vp, VV, stt);
        if (stt && (threadIdx.x & (kWave - 1)) == 0) atomicAdd(&a.counters[2], 1ull);
        const double tt1 = -n2, tt2 = n1;
        bool toA = false, toB = false;
        double V = 0.0;
        if (line == 0) {
            V = VV * t.koef * t.k_line[0];
            toB = true;
        } else if (line == 1) {
            V = Vl * t.koef * t.k_line[1];
            toB = true;
            toA = fabs(p2[2] * tt1 + p2[3] * tt2) < t.tang_inner;
        } else {
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
    if ((threadIdx.x & (kWave - 1)) != 0) return;
    double* o = t.contrib + (size_t)g * 6;
#pragma unroll
    for (int k = 0; k < 6; k++) o[k] = r[k];
}

// One work-group: the node speeds (Move_surface's sums, averages and end corrections
// in its statement order, one thread, from LDS), then the rails (Move_Setka_Calculate_2)
// phase by phase: a phase is a run of consecutive rails of which none writes a point
// another writes or reads, so its rails run side by side, one wavefront each, a rail's
// key points uniformly and its spacing loops across the lanes; the phases follow the
// reference's order, so every node's last assignment is the reference's.
__global__ void __launch_bounds__(kMoveThreads) mtrk_move_kernel(MeshTrk t, SweepArgs a, double* __restrict__ X2,
                                                                 double* __restrict__ Y2)
{
    __shared__ double vx[kMaxLineNodes], vy[kMaxLineNodes];
    __shared__ int cnt[kMaxLineNodes];
    __shared__ double sc[kMaxLineFaces * 6];
    __shared__ int sl[kMaxLineFaces * 2];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1), wave = tid / kWave;
    const double* X = a.x;
    const double* Y = a.y;
    const double dt = dt_read(a);
    for (int k = tid; k < t.nln; k += kMoveThreads) { vx[k] = 0.0; vy[k] = 0.0; cnt[k] = 0; }
    for (int k = tid; k < t.nlf * 6; k += kMoveThreads) sc[k] = t.contrib[k];
    for (int k = tid; k < t.nlf * 2; k += kMoveThreads) sl[k] = t.lslot[k];
    __syncthreads();
    // Move_surface section by section (contact, inner shock, outer shock): each line
    // node's contributions in face order (at most two, those of the faces it ends),
    // the averaging loop's nodes (distinct), the two end corrections
    for (int line = 0; line < 3; line++) {
        for (int k = t.sum_off[line] + tid; k < t.sum_off[line + 1]; k += kMoveThreads) {
            const int* e = t.sums + 3 * k;
            for (int c = 1; c <= 2; c++)
                if (e[c] >= 0) {
                    const double* r = sc + (e[c] >> 1) * 6 + (e[c] & 1) * 3;
                    if (r[2] > 0.0) { vx[e[0]] += r[0]; vy[e[0]] += r[1]; cnt[e[0]]++; }
                }
        }
        __syncthreads();
        for (int k = t.avg_off[line] + tid; k < t.avg_off[line + 1]; k += kMoveThreads) {
            const int sa = t.avgs[k];
            if (cnt[sa] > 0) {
                vx[sa] /= (1.0 * cnt[sa]);
                vy[sa] /= (1.0 * cnt[sa]);
            }
        }
        __syncthreads();
        if (tid == 0) {                                     // Setka.cpp:3532-3535, :3733-3736, :3839-3842
            const int base = line == 0 ? 0 : (line == 1 ? t.nline[0] : t.nline[0] + t.nline[1]);
            const int n = t.nline[line];
            const int* f0 = t.lf + (size_t)base * 6;
            const int* fl = t.lf + (size_t)(base + n - 1) * 6;
            const int a0 = sl[base * 2], b0 = sl[base * 2 + 1];
            const int al = sl[(base + n - 1) * 2], bl = sl[(base + n - 1) * 2 + 1];
            vx[a0] = vx[b0] + (X[f0[1]] - X[f0[0]]) / dt;
            if (line == 1) vx[bl] = vx[al] + (X[fl[0]] - X[fl[1]]) / dt;
            else vy[bl] = vy[al] + (Y[fl[0]] - Y[fl[1]]) / dt;
        }
        __syncthreads();
    }
    auto speed = [&](int p, double& ux, double& uy) {
        const int q = t.node_slot[p];
        ux = q >= 0 ? vx[q] : 0.0;
        uy = q >= 0 ? vy[q] : 0.0;
    };
    for (int k = tid; k < t.nfree; k += kMoveThreads) {
        const int p = t.freep[k];
        X2[p] = X[p];
        Y2[p] = Y[p];
    }
    __builtin_amdgcn_s_waitcnt(0);                           // a D rail may read a node no rail places
    __threadfence();
    __syncthreads();
    const double R1 = t.r_inner, R111 = t.r_mid, R5 = t.r_outer;
    // the outer-shock position D rails from d_chain on copy from the rail before, which
    // is the one D rail d_chain - 1 computes (Setka.cpp:4700-4706)
    double dchain = 0.0;
    if (t.d_chain_rail >= 0) {
        const int K2 = t.rail[(size_t)t.d_chain_rail * kRailInts + 9];
        double ux, uy;
        speed(K2, ux, uy);
        const double V = uy;
        dchain = Y[K2] + dt * V;
    }
    for (int ph = 0; ph < t.nphase; ph++) {
        for (int q = t.phase[ph] + wave; q < t.phase[ph + 1]; q += kMoveThreads / kWave) {
            const int* rr = t.rail + (size_t)q * kRailInts;
            const int kind = rr[0], M1 = rr[1], M2 = rr[2], M3 = rr[3], M4 = rr[4];
            const int* P = t.allp + rr[5];
            const int nall = rr[6];
            const int K0 = rr[7], K1 = rr[8], K2 = rr[9], K3 = rr[10];
            const double* fmid = t.fpow + rr[11];
            const double* fout = t.fpow + rr[12];
            const double cs = t.rcs[2 * q], sn = t.rcs[2 * q + 1];
            auto put = [&](int p, double x2, double y2) { X2[p] = x2; Y2[p] = y2; };
            // a key point the rail's loops place later is not stored (its value would be
            // overwritten), so that no two stores of one rail meet at one address
            auto key = [&](int k, int p, double x2, double y2) { if ((rr[14] >> k) & 1) put(p, x2, y2); };
            double ux, uy, V;
            // inner and mid blocks of A, B, C rails (Setka.cpp:4391-4405 and the same
            // loops of the B and C rails) from R2, the radius of key point 0 after its move
            auto inner_blocks = [&](double R2) {
                for (int j = lane; j < M1 - 1; j += kWave) {
                    double r;
                    if (j < t.n_inner) r = R1 + t.fin[j] * (R111 - R1);
                    else r = R111 + fmid[j - t.n_inner] * (R2 - R111);
                    put(P[j], r * cs, r * sn);
                }
            };
            if (kind == 0) {                                 // A (Setka.cpp:4331-4457)
                double k0x, k0y, k1x, k1y, k2x, k2y;
                speed(K0, ux, uy); V = ux * cs + uy * sn; k0x = X[K0] + dt * V * cs; k0y = Y[K0] + dt * V * sn;
                speed(K1, ux, uy); V = ux * cs + uy * sn; k1x = X[K1] + dt * V * cs; k1y = Y[K1] + dt * V * sn;
                speed(K2, ux, uy); V = ux * cs + uy * sn; k2x = X[K2] + dt * V * cs; k2y = Y[K2] + dt * V * sn;
                if (lane == 0) { key(0, K0, k0x, k0y); key(1, K1, k1x, k1y); key(2, K2, k2x, k2y); key(3, K3, X[K3], Y[K3]); }
                const double R2 = sqrt(sq(k0x) + sq(k0y));
                inner_blocks(R2);
                const double R3 = sqrt(sq(k1x) + sq(k1y));
                for (int j = lane; j < M2 - 1; j += kWave) {
                    const double r = R2 + (R3 - R2) * (j + 1) / (M2);
                    put(P[M1 + j], r * cs, r * sn);
                }
                const double R4 = sqrt(sq(k2x) + sq(k2y));
                for (int j = lane; j < M3 - 1; j += kWave) {
                    const double r = R3 + (R4 - R3) * (j + 1) / (M3);
                    put(P[M1 + M2 + j], r * cs, r * sn);
                }
                for (int j = lane; j < M4; j += kWave) {
                    const double r = R4 + fout[j] * (R5 - R4);
                    put(P[M1 + M2 + M3 + j], r * cs, r * sn);
                }
            } else if (kind == 1) {                          // B (Setka.cpp:4468-4593)
                speed(K0, ux, uy); V = ux * cs + uy * sn;
                const double k0x = X[K0] + dt * V * cs, k0y = Y[K0] + dt * V * sn;
                speed(K1, ux, uy); V = ux * 0.0 + uy * 1.0;
                const double k1y = Y[K1] + dt * V;
                speed(K2, ux, uy); V = ux * 0.0 + uy * 1.0;
                const double k2y = Y[K2] + dt * V;
                if (lane == 0) { key(0, K0, k0x, k0y); key(1, K1, k0x, k1y); key(2, K2, k0x, k2y); key(3, K3, k0x, Y[K3]); }
                inner_blocks(sqrt(sq(k0x) + sq(k0y)));
                for (int j = lane; j < M2 - 1; j += kWave) put(P[M1 + j], k0x, k0y + (k1y - k0y) * (j + 1) / (M2));
                for (int j = lane; j < M3 - 1; j += kWave) put(P[M1 + M2 + j], k0x, k1y + (k2y - k1y) * (j + 1) / (M3));
                for (int j = lane; j < M4; j += kWave) put(P[M1 + M2 + M3 + j], k0x, k2y + (R5 - k2y) * (j + 1) / (M4));
            } else if (kind == 2) {                          // C (Setka.cpp:4596-4670)
                speed(K0, ux, uy); V = ux * cs + uy * sn;
                const double k0x = X[K0] + dt * V * cs, k0y = Y[K0] + dt * V * sn;
                if (lane == 0) { key(0, K0, k0x, k0y); key(1, K1, t.left, Y[K1]); }
                inner_blocks(sqrt(sq(k0x) + sq(k0y)));
                const double R2 = k0x, R3 = t.left;
                const double xg = pow(R3 / (R2), 1.0 / (M2));
                for (int j = M1 + lane; j < nall; j += kWave) put(P[j], R2 * pow(xg, (double)(j + 1 - M1)), k0y);
            } else {                                         // D (Setka.cpp:4672-4759)
                const double k0x = X2[K0], k0y = Y2[K0];     // placed by a rail of an earlier phase
                speed(K1, ux, uy); V = uy;
                const double k1y = Y[K1] + dt * V;
                speed(K2, ux, uy); V = uy;
                const double k2y = rr[13] ? dchain : Y[K2] + dt * V;
                if (lane == 0) { key(1, K1, k0x, k1y); key(2, K2, k0x, k2y); key(3, K3, k0x, Y[K3]); }
                for (int j = lane; j < M2 - 1; j += kWave) put(P[j + 1], k0x, k0y + (k1y - k0y) * (j + 1) / (M2));
                for (int j = lane; j < M3 - 1; j += kWave) put(P[M2 + j + 1], k0x, k1y + (k2y - k1y) * (j + 1) / (M3));
                for (int j = lane; j < M4; j += kWave) put(P[M2 + M3 + j + 1], k0x, k2y + (R5 - k2y) * (j + 1) / (M4));
            }
        }
        __builtin_amdgcn_s_waitcnt(0);                       // the phase's stores land, and are seen,
        __threadfence();                                     // before the next phase reads or overwrites
        __syncthreads();
    }
}

}  // namespace gel

using namespace gel;

// Device arrays of a mesh context: one allocation, laid out as MeshDev reads it.
struct MeshBuf {
    MeshDev dev{};
    void* base = nullptr;
    // host copies of the face tables (fitted-surface lines are matched against them)
    std::vector<int> hA, hB, hown, hnbr, hdown, hup;
    // the face slots as given (the Monte Carlo walks them): offsets per cell, Gran::A, B,
    // type and Sosed
    std::vector<int> hsoff, hsA, hsB, hstype, hsnbr;
    // fitted-surface tracking (gel_set_mesh_tracking)
    MeshTrk trk{};
    void* trk_base = nullptr;
};

void gel_mesh_free(gel_ctx* c)
{
    if (!c || !c->mesh) return;
    (void)hipFree(c->mesh->trk_base);
    (void)hipFree(c->mesh->base);
    delete c->mesh;
    c->mesh = nullptr;
}

// Face tables of the mesh: every face with two cells once, owned by the cell of lower
// index, whose slot gives its nodes and stencil; the other cell's slot must be the copy
// face TVD_prepare pairs with it (nodes reversed, Sosed_down and Sosed_up exchanged).
int gel_mesh_build(gel_ctx* c, const gel_mesh* g, std::string& err)
{
    const int nc = g->n_cells, nn = g->n_nodes;
    const int nslot = g->face_offset[nc];
    std::vector<int> sface(nslot), sA(nslot), sB(nslot);
    std::vector<int> fA, fB, ft, fown, fnbr, fdown, fup;
    std::vector<int8_t> fsurf, caxis(nc, 0), ctype(nc);
    std::map<std::tuple<int, int, int, int>, int> pair;   // (low cell, high cell, low node, high node) -> face
    auto pt = [&](int k) { return g->point_type ? (int)g->point_type[k] : 0; };
    for (int cc = 0; cc < nc; cc++) {
        ctype[cc] = g->cell_type[cc];
        if (ctype[cc] < GEL_C_1 || ctype[cc] > GEL_C_NO) { err = "cell_type out of range"; return GEL_E_ARG; }
        const int k0 = g->face_offset[cc], k1 = g->face_offset[cc + 1];
        if (k1 < k0 || k1 - k0 < 3) { err = "a cell needs at least three faces"; return GEL_E_ARG; }
        if (g->corner_offset[cc + 1] - g->corner_offset[cc] < 3) { err = "a cell needs at least three corners"; return GEL_E_ARG; }
        for (int k = k0; k < k1; k++) {
            const int A = g->face_a[k], B = g->face_b[k], t = g->face_type[k], N = g->face_sosed[k];
            const int D = g->face_down[k], U = g->face_up[k];
            if (A < 0 || A >= nn || B < 0 || B >= nn || t < GEL_F_USUALY || t > GEL_F_UPPER_WALL ||
                N < -1 || N >= nc || D < -1 || D >= nc || U < -1 || U >= nc) {
                err = "face slot out of range";
                return GEL_E_ARG;
            }
            sA[k] = A; sB[k] = B;
            if (t == GEL_F_AXIS) caxis[cc] = 1;
            const int pa = pt(A), pb = pt(B);
            const int8_t surf = (int8_t)((pa != 0 && pa == pb ? 1 : 0) | (pa != pb ? 2 : 0));
            if (t != GEL_F_USUALY) {
                sface[k] = (int)fA.size();
                fA.push_back(A); fB.push_back(B); ft.push_back(t); fown.push_back(cc); fnbr.push_back(-1);
                fdown.push_back(D); fup.push_back(-1); fsurf.push_back(surf);
                continue;
            }
            if (N < 0 || N == cc) { err = "an Usualy face needs another cell"; return GEL_E_ARG; }
            const auto key = std::make_tuple(std::min(cc, N), std::max(cc, N), std::min(A, B), std::max(A, B));
            auto it = pair.find(key);
            if (cc < N) {
                if (it != pair.end()) { err = "a face appears twice in a cell"; return GEL_E_ARG; }
                const int f = (int)fA.size();
                pair[key] = f;
                sface[k] = f;
                fA.push_back(A); fB.push_back(B); ft.push_back(t); fown.push_back(cc); fnbr.push_back(N);
                fdown.push_back(D); fup.push_back(U); fsurf.push_back(surf);
            } else {
                // the owner (lower index) comes first in cell order
                if (it == pair.end()) { err = "an Usualy face is missing from its other cell"; return GEL_E_ARG; }
                const int f = it->second;
                if (fA[f] != B || fB[f] != A || fdown[f] != U || fup[f] != D) {
                    err = "the two slots of a face are not a face and its copy (nodes reversed, stencils exchanged)";
                    return GEL_E_ARG;
                }
                sface[k] = ~f;
                pair.erase(it);
            }
        }
    }
    if (!pair.empty()) { err = "an Usualy face is missing from its other cell"; return GEL_E_ARG; }
    const int nf = (int)fA.size();
    std::vector<int> gface;
    for (int f = 0; f < nf; f++)
        if (fsurf[f] & 1) gface.push_back(f);
    const int ng = (int)gface.size();
    // layout: ints, then int8, then doubles (aligned)
    const size_t ints = (size_t)(nc + 1) + g->corner_offset[nc] + (size_t)(nc + 1) + 3 * (size_t)nslot + 7 * (size_t)nf +
                        (size_t)ng;
    const size_t bytes8 = 2 * (size_t)nc + nf;
    const size_t off8 = ints * 4, offd = ((off8 + bytes8 + 255) / 256) * 256;
    const size_t total = offd + (6 * (size_t)nc + 8 * (size_t)nf) * 8;
    MeshBuf* mb = new (std::nothrow) MeshBuf();
    if (!mb) { err = "host allocation"; return GEL_E_NOMEM; }
    if (hipMalloc(&mb->base, total) != hipSuccess) { delete mb; err = "mesh tables"; return GEL_E_NOMEM; }
    std::vector<char> h(offd);
    int* ip = (int*)h.data();
    size_t at = 0;
    auto put = [&](const int* src, size_t n) { const int* p = (const int*)mb->base + at; if (n) std::memcpy(ip + at, src, n * 4); at += n; return p; };
    MeshDev& d = mb->dev;
    d.nc = nc; d.nf = nf;
    d.coff = put(g->corner_offset, nc + 1);
    d.corners = put(g->corners, g->corner_offset[nc]);
    d.soff = put(g->face_offset, nc + 1);
    d.sface = put(sface.data(), nslot);
    d.sA = put(sA.data(), nslot);
    d.sB = put(sB.data(), nslot);
    d.fA = put(fA.data(), nf);
    d.fB = put(fB.data(), nf);
    d.ftype = put(ft.data(), nf);
    d.fown = put(fown.data(), nf);
    d.fnbr = put(fnbr.data(), nf);
    d.fdown = put(fdown.data(), nf);
    d.fup = put(fup.data(), nf);
    d.ng = ng;
    d.gface = put(gface.data(), ng);
    char* b8 = h.data() + off8;
    std::memcpy(b8, ctype.data(), nc);
    std::memcpy(b8 + nc, caxis.data(), nc);
    std::memcpy(b8 + 2 * nc, fsurf.data(), nf);
    d.ctype = (const int8_t*)((char*)mb->base + off8);
    d.caxis = d.ctype + nc;
    d.fsurf = d.ctype + 2 * nc;
    d.cgeo = (double*)((char*)mb->base + offd);
    d.fout = d.cgeo + 6 * (size_t)nc;
    for (int k = 0; k < nc; k++) {
        const int n = g->corner_offset[k + 1] - g->corner_offset[k];
        for (int j = 0; j < n; j++) {
            const int p = g->corners[g->corner_offset[k] + j];
            if (p < 0 || p >= nn) { (void)hipFree(mb->base); delete mb; err = "corner out of range"; return GEL_E_ARG; }
        }
    }
    if (hipMemcpy(mb->base, h.data(), offd, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(mb->base);
        delete mb;
        err = "copy of the mesh tables";
        return GEL_E_HIP;
    }
    mb->hA = fA; mb->hB = fB; mb->hown = fown; mb->hnbr = fnbr; mb->hdown = fdown; mb->hup = fup;
    mb->hsoff.assign(g->face_offset, g->face_offset + nc + 1);
    mb->hsA = sA;
    mb->hsB = sB;
    mb->hstype.assign(g->face_type, g->face_type + nslot);
    mb->hsnbr.assign(g->face_sosed, g->face_sosed + nslot);
    c->mesh = mb;
    c->mesh_faces = nf;
    return GEL_OK;
}

// The mesh tables the Monte Carlo walks: device pointers to the corners and face slots,
// and on the host each slot's way on (code) and the two boundary lists as (A, B, cell)
// triples.  A slot with a neighbour (Sosed, whatever its type: the inner sphere of the
// reference's grid has the central cells behind it) leads to it; an Axis slot is
// skipped (-3); an inner-sphere slot with no neighbour is crossed in a straight line to
// another (-2, the inner list); any other boundary slot leaves the grid (-1, the outer
// list, where primaries enter).
void gel_mesh_mc_view(const gel_ctx* c, MeshMcView* v, std::vector<int>& code, std::vector<int>& outer,
                      std::vector<int>& inner)
{
    const MeshBuf* mb = c->mesh;
    const MeshDev& d = mb->dev;
    v->coff = d.coff; v->corners = d.corners; v->soff = d.soff; v->sA = d.sA; v->sB = d.sB;
    v->ctype = d.ctype;
    const std::vector<int>& soff = mb->hsoff;
    const std::vector<int>& sA = mb->hsA;
    const std::vector<int>& sB = mb->hsB;
    code.assign(mb->hstype.size(), -1);
    outer.clear();
    inner.clear();
    for (int cc = 0; cc + 1 < (int)soff.size(); cc++)
        for (int k = soff[cc]; k < soff[cc + 1]; k++) {
            const int t = mb->hstype[k], n = mb->hsnbr[k];
            if (t == GEL_F_AXIS) { code[k] = -3; continue; }
            if (n >= 0) { code[k] = n; continue; }
            code[k] = t == GEL_F_INNER_SPHERE ? -2 : -1;
            std::vector<int>& L = t == GEL_F_INNER_SPHERE ? inner : outer;
            L.push_back(sA[k]);
            L.push_back(sB[k]);
            L.push_back(cc);
        }
}

// Tables of fitted-surface tracking on a mesh context (gel_set_mesh_tracking).  The
// checks make the device replay the reference's assignments exactly: each line face
// matched to a face of the mesh (its stencil cells from the face tables), the shocks'
// averaging loops inside the contact line's face count, every rail's key points left
// alone by its inner spacing blocks (n_inner <= M1 - 1), a D rail's outer-shock key
// point outside its own loops (the next D rail reads it), and a D rail's key point 0
// placed by an earlier rail or by none.
int gel_mesh_set_tracking(gel_ctx* c, const gel_tracking* tr, const gel_mesh_rails* r, std::string& err)
{
    MeshBuf* mb = c->mesh;
    (void)hipFree(mb->trk_base);
    mb->trk_base = nullptr;
    mb->trk = MeshTrk{};
    if (!tr) return GEL_OK;
    const int nn = (int)c->nn, nc = (int)c->nc;
    if (!r || r->n_rails < 1 || !r->kind || !r->m || !r->s || !r->all_offset || !r->all || !r->key || !r->line) {
        err = "rails and lines needed";
        return GEL_E_ARG;
    }
    const int nl0 = r->n_line[0], nl1 = r->n_line[1], nl2 = r->n_line[2];
    if (nl0 < 1 || nl1 < 1 || nl2 < 1 || nl1 > nl0 || nl2 > nl0) {
        err = "lines: at least one face each, the shocks no longer than the contact (its averaging loops)";
        return GEL_E_ARG;
    }
    const int n_inner = tr->inner_rail_nodes;
    if (n_inner < 1 || tr->outer_faces < 0 || r->d_chain < 1) { err = "tracking constants"; return GEL_E_ARG; }
    // line faces against the mesh's faces
    std::map<std::pair<int, int>, std::vector<int>> by_nodes;
    for (size_t f = 0; f < mb->hA.size(); f++)
        by_nodes[{ std::min(mb->hA[f], mb->hB[f]), std::max(mb->hA[f], mb->hB[f]) }].push_back((int)f);
    const int nlf = nl0 + nl1 + nl2;
    std::vector<int> lf((size_t)nlf * 6), lslot((size_t)nlf * 2), node_slot(nn, -1), lnode;
    for (int g = 0; g < nlf; g++) {
        const int* L = r->line + (size_t)g * 4;
        const int A = L[0], B = L[1], M = L[2], N = L[3];
        if (A < 0 || A >= nn || B < 0 || B >= nn || M < 0 || M >= nc || N < 0 || N >= nc) {
            err = "line face out of range";
            return GEL_E_ARG;
        }
        int D = -1, U = -1;
        auto it = by_nodes.find({ std::min(A, B), std::max(A, B) });
        if (it != by_nodes.end())
            for (int f : it->second) {
                if (mb->hown[f] == M && mb->hnbr[f] == N) { D = mb->hdown[f]; U = mb->hup[f]; }
                else if (mb->hown[f] == N && mb->hnbr[f] == M) { D = mb->hup[f]; U = mb->hdown[f]; }
            }
        if (D < 0 || U < 0) { err = "a line face is not an interior face of the mesh with stencils on both sides"; return GEL_E_ARG; }
        int* o = &lf[(size_t)g * 6];
        o[0] = A; o[1] = B; o[2] = M; o[3] = N; o[4] = D; o[5] = U;
        for (int e = 0; e < 2; e++) {
            const int p = e ? B : A;
            if (node_slot[p] < 0) { node_slot[p] = (int)lnode.size(); lnode.push_back(p); }
            lslot[(size_t)g * 2 + e] = node_slot[p];
        }
    }
    // each line node's contributions in its line's face order, and the averaging loops'
    // nodes: the contact line's A nodes (its own loop, then each shock's loop over the
    // shock's face count), or each line its own (average_lines)
    std::vector<int> sum_off{ 0 }, sums, avg_off{ 0 }, avgs;
    for (int line = 0, base = 0; line < 3; line++) {
        const int n = r->n_line[line];
        std::map<int, std::vector<int>> per;             // line slot -> 2 face + role
        std::vector<int> order;
        for (int j = 0; j < n; j++)
            for (int e = 0; e < 2; e++) {
                const int sl = lslot[(size_t)(base + j) * 2 + e];
                if (!per.count(sl)) order.push_back(sl);
                per[sl].push_back(2 * (base + j) + e);
            }
        for (int sl : order) {
            if (per[sl].size() > 2) { err = "a line node ends more than two of its line's faces"; return GEL_E_ARG; }
            sums.push_back(sl);
            sums.push_back(per[sl][0]);
            sums.push_back(per[sl].size() > 1 ? per[sl][1] : -1);
        }
        sum_off.push_back((int)sums.size() / 3);
        const int avg = (line == 0 || tr->average_lines) ? base : 0;
        std::vector<char> seen(lnode.size() + 1, 0);
        for (int j = 0; j < n; j++) {
            const int sa = lslot[(size_t)(avg + j) * 2];
            if (seen[sa]) { err = "a line node starts two faces of one averaging loop"; return GEL_E_ARG; }
            seen[sa] = 1;
            avgs.push_back(sa);
        }
        avg_off.push_back((int)avgs.size());
        base += n;
    }
    if ((int)lnode.size() > kMaxLineNodes || nlf > kMaxLineFaces) {
        err = "more line faces or nodes than the tracking kernel holds";
        return GEL_E_ARG;
    }
    // rails
    const int nr = r->n_rails;
    const double dd = tr->s_outer / std::sqrt(1.0 - 1.0 / tr->p_outer);
    std::vector<int> rail((size_t)nr * kRailInts), allp;
    std::vector<double> rcs((size_t)nr * 2), fin(n_inner), fpow;
    for (int j = 0; j < n_inner; j++) fin[j] = std::pow(j * 1.0 / (n_inner), tr->p_inner);
    std::vector<std::vector<int>> wrotes;               // every rail's assignments
    int last_kind = 0;
    for (int q = 0; q < nr; q++) {
        const int kind = r->kind[q];
        const int M1 = r->m[4 * q], M2 = r->m[4 * q + 1], M3 = r->m[4 * q + 2], M4 = r->m[4 * q + 3];
        const int o0 = r->all_offset[q], o1 = r->all_offset[q + 1];
        const int nall = o1 - o0;
        const int* K = r->key + 4 * q;
        if (kind < 0 || kind > 3 || kind < last_kind) { err = "rails: kinds 0..3 in the order A, B, C, D"; return GEL_E_ARG; }
        last_kind = kind;
        if (o0 < 0 || nall < 1) { err = "rail offsets"; return GEL_E_ARG; }
        const int nkey = kind == 2 ? 2 : 4;
        for (int k = 0; k < nkey; k++)
            if (K[k] < 0 || K[k] >= nn) { err = "rail key point out of range"; return GEL_E_ARG; }
        for (int k = o0; k < o1; k++)
            if (r->all[k] < 0 || r->all[k] >= nn) { err = "rail point out of range"; return GEL_E_ARG; }
        const int* P = r->all + o0;
        {
            std::vector<int> u(P, P + nall);
            std::sort(u.begin(), u.end());
            if (std::adjacent_find(u.begin(), u.end()) != u.end()) { err = "a rail lists a point twice"; return GEL_E_ARG; }
        }
        std::vector<int> wrote;                          // the rail's assignments, for the checks
        size_t nloop = 0;                                // of them, its loops'
        if (kind <= 2) {
            if (M1 - 1 < n_inner || M1 < 1) { err = "rails: M1 - 1 must be at least n_inner"; return GEL_E_ARG; }
            for (int j = 0; j < M1 - 1; j++) wrote.push_back(P[j]);
        }
        if (kind <= 1) {
            if (M2 < 1 || M3 < 1 || M4 < 1 || M1 + M2 + M3 + M4 - 1 > nall) { err = "rails: All_point shorter than M1..M4"; return GEL_E_ARG; }
            for (int j = 0; j < M2 - 1; j++) wrote.push_back(P[M1 + j]);
            for (int j = 0; j < M3 - 1; j++) wrote.push_back(P[M1 + M2 + j]);
            for (int j = 0; j < M4; j++) wrote.push_back(P[M1 + M2 + M3 + j]);
            nloop = wrote.size();
            for (int k = 0; k < 4; k++) wrote.push_back(K[k]);
        } else if (kind == 2) {
            if (M2 < 1 || nall < M1) { err = "C rails: M2 >= 1, All_point at least M1"; return GEL_E_ARG; }
            for (int j = M1; j < nall; j++) wrote.push_back(P[j]);
            nloop = wrote.size();
            wrote.push_back(K[0]); wrote.push_back(K[1]);
        } else {
            if (M2 < 1 || M3 < 1 || M4 < 1 || M2 + M3 + M4 >= nall || K[2] != P[M2 + M3]) {
                err = "D rails: All_point past M2 + M3 + M4, key point 2 at M2 + M3";
                return GEL_E_ARG;
            }
            for (int j = 0; j < M2 - 1; j++) wrote.push_back(P[j + 1]);
            for (int j = 0; j < M3 - 1; j++) wrote.push_back(P[M2 + j + 1]);
            for (int j = 0; j < M4; j++) wrote.push_back(P[M2 + M3 + j + 1]);
            nloop = wrote.size();
            wrote.push_back(K[1]); wrote.push_back(K[2]); wrote.push_back(K[3]);
        }
        wrotes.push_back(wrote);
        int* rr = &rail[(size_t)q * kRailInts];
        rr[0] = kind; rr[1] = M1; rr[2] = M2; rr[3] = M3; rr[4] = M4;
        rr[5] = (int)allp.size(); rr[6] = nall;
        for (int k = 0; k < 4; k++) rr[7 + k] = k < nkey ? K[k] : K[0];
        int kmask = 0;
        for (int k = 0; k < nkey; k++)
            if (std::find(wrote.begin(), wrote.begin() + nloop, K[k]) == wrote.begin() + nloop) kmask |= 1 << k;
        rr[14] = kmask;
        allp.insert(allp.end(), P, P + nall);
        // host libm, the calls the reference makes at run time (Setka.cpp:4335-4450)
        rcs[2 * q] = std::cos(r->s[q]);
        rcs[2 * q + 1] = std::sin(r->s[q]);
        rr[11] = (int)fpow.size();
        if (kind <= 2)
            for (int j = n_inner; j < M1 - 1; j++)
                fpow.push_back(std::pow((j - n_inner) * 1.0 / (M1 - 1 - n_inner), tr->p_mid));
        rr[12] = (int)fpow.size();
        if (kind == 0)
            for (int j = 0; j < M4; j++) {
                const double x = (j + 1) * 1.0 / (M4);
                fpow.push_back(std::pow(x, tr->p_outer * (1.0 - (r->s[q] / dd) * (r->s[q] / dd))));
            }
    }
    // a D rail reads its key point 0 where an earlier rail placed it, or where it is
    std::vector<char> any(nn, 0), placed(nn, 0);
    for (const auto& w : wrotes)
        for (int p : w) any[p] = 1;
    for (int q = 0; q < nr; q++) {
        if (r->kind[q] == 3 && !placed[r->key[4 * q]] && any[r->key[4 * q]]) {
            err = "D rails: key point 0 placed by a later rail";
            return GEL_E_ARG;
        }
        for (int p : wrotes[q]) placed[p] = 1;
    }
    std::vector<int> freep;
    for (int p = 0; p < nn; p++)
        if (!any[p]) freep.push_back(p);
    // D rails from d_chain on copy the outer-shock position of D rail d_chain - 1
    int d_chain_rail = -1;
    for (int q = 0, dk = 0; q < nr; q++) {
        if (r->kind[q] != 3) continue;
        if (dk == r->d_chain - 1) d_chain_rail = q;
        rail[(size_t)q * kRailInts + 13] = dk >= r->d_chain ? 1 : 0;
        dk++;
    }
    // phases: runs of consecutive rails none of which writes a point another of the run
    // writes or reads (a D rail reads its key point 0)
    std::vector<int> phase{ 0 }, wstamp(nn, -1), rstamp(nn, -1);
    for (int q = 0; q < nr; q++) {
        const int ph = (int)phase.size() - 1;
        bool clash = false;
        for (int p : wrotes[q]) clash = clash || wstamp[p] == ph || rstamp[p] == ph;
        if (r->kind[q] == 3) clash = clash || wstamp[r->key[4 * q]] == ph;
        if (clash) phase.push_back(q);
        const int cur = (int)phase.size() - 1;
        for (int p : wrotes[q]) wstamp[p] = cur;
        if (r->kind[q] == 3) rstamp[r->key[4 * q]] = cur;
    }
    phase.push_back(nr);
    if (fpow.empty()) fpow.push_back(0.0);
    // one allocation: ints, then doubles, then the face contributions
    const size_t ni_ = lf.size() + lslot.size() + node_slot.size() + rail.size() + allp.size() + freep.size() +
                       phase.size() + sum_off.size() + sums.size() + avg_off.size() + avgs.size();
    const size_t offd = ((ni_ * 4 + 255) / 256) * 256;
    const size_t nd_ = rcs.size() + fin.size() + fpow.size();
    const size_t total = offd + (nd_ + (size_t)nlf * 6) * 8;
    void* base = nullptr;
    if (hipMalloc(&base, total) != hipSuccess) { err = "tracking tables"; return GEL_E_NOMEM; }
    std::vector<char> h(offd + nd_ * 8);
    size_t at = 0;
    auto puti = [&](const std::vector<int>& v) {
        const int* p = (const int*)((char*)base + at);
        if (!v.empty()) std::memcpy(h.data() + at, v.data(), v.size() * 4);
        at += v.size() * 4;
        return p;
    };
    MeshTrk t{};
    t.lf = puti(lf); t.lslot = puti(lslot); t.node_slot = puti(node_slot); t.rail = puti(rail);
    t.allp = puti(allp); t.freep = puti(freep); t.phase = puti(phase);
    t.sum_off = puti(sum_off); t.sums = puti(sums); t.avg_off = puti(avg_off); t.avgs = puti(avgs);
    t.nphase = (int)phase.size() - 1;
    t.d_chain_rail = d_chain_rail;
    at = offd;
    auto putd = [&](const std::vector<double>& v) {
        const double* p = (const double*)((char*)base + at);
        std::memcpy(h.data() + at, v.data(), v.size() * 8);
        at += v.size() * 8;
        return p;
    };
    t.rcs = putd(rcs); t.fin = putd(fin); t.fpow = putd(fpow);
    t.contrib = (double*)((char*)base + at);
    if (hipMemcpy(base, h.data(), h.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(t.contrib, 0, (size_t)nlf * 6 * 8) != hipSuccess) {
        (void)hipFree(base);
        err = "copy of the tracking tables";
        return GEL_E_HIP;
    }
    t.nlf = nlf; t.nline[0] = nl0; t.nline[1] = nl1; t.nline[2] = nl2;
    t.nln = (int)lnode.size(); t.nrail = nr; t.nfree = (int)freep.size();
    t.n_inner = n_inner; t.d_chain = r->d_chain; t.outer_faces = tr->outer_faces; t.average_lines = tr->average_lines;
    t.koef = tr->koef; t.k_line[0] = tr->k_contact; t.k_line[1] = tr->k_inner; t.k_line[2] = tr->k_outer;
    t.tang_inner = tr->tang_inner; t.tang_outer = tr->tang_outer;
    t.r_inner = tr->r_inner; t.r_mid = tr->r_mid; t.r_outer = tr->r_outer; t.left = r->left;
    mb->trk = t;
    mb->trk_base = base;
    return GEL_OK;
}

// One step on a mesh context: geometry (when the nodes move, or once on a static grid),
// the faces, the cells.  moving: the step's next node positions are a.x2, a.y2.
hipError_t gel::gel_mesh_step(gel_ctx* c, SweepArgs& a, bool moving)
{
    MeshDev& d = c->mesh->dev;
    hipStream_t s = c->stream;
    if (c->tracking && c->mesh->trk_base) {            // Move_surface, Move_Setka_Calculate_2
        const MeshTrk& t = c->mesh->trk;
        a.x2 = c->d_x2;
        a.y2 = c->d_y2;
        hipLaunchKernelGGL(mtrk_face_kernel, dim3((unsigned)((t.nlf + 3) / 4)), dim3(256), 0, s, d, t, a);
        hipLaunchKernelGGL(mtrk_move_kernel, dim3(1), dim3(kMoveThreads), 0, s, t, a, c->d_x2, c->d_y2);
        moving = true;
    }
    // grids of a multiple of eight work-groups, so that each XCD takes a contiguous
    // run of cells or faces (xcd_item) and neighbours share its L2
    auto groups = [](int n) { return (unsigned)((((n + 255) / 256) + 7) / 8 * 8); };
    if (moving || !c->geo_valid) {
        hipLaunchKernelGGL(mesh_geo_kernel, dim3(groups(d.nc)), dim3(256), 0, s, d, a.x, a.y, a.x2, a.y2);
        c->geo_valid = !moving;
    }
    const int nfb = (int)groups(d.nf);
    const dim3 gf((unsigned)(nfb + (d.ng + 3) / 4)), gc(groups(d.nc)), blk(256);
    const bool hllc = c->prm.flux == GEL_FLUX_HLLC;
    const int src = c->prm.source_mode;
    using FaceK = void (*)(MeshDev, SweepArgs, int);
    using CellK = void (*)(MeshDev, SweepArgs);
    FaceK fk;
    CellK ck;
    if (c->ts == 4) {   // FP32 bulk
        fk = hllc ? mesh_face_kernel<GEL_FLUX_HLLC, float> : mesh_face_kernel<GEL_FLUX_HLL, float>;
        ck = src == GEL_SRC_MK1     ? mesh_cell_kernel<GEL_SRC_MK1, float>
             : src == GEL_SRC_TABLE ? mesh_cell_kernel<GEL_SRC_TABLE, float>
                                    : mesh_cell_kernel<GEL_SRC_NONE, float>;
    } else {
        fk = hllc ? mesh_face_kernel<GEL_FLUX_HLLC, double> : mesh_face_kernel<GEL_FLUX_HLL, double>;
        ck = src == GEL_SRC_MK1     ? mesh_cell_kernel<GEL_SRC_MK1, double>
             : src == GEL_SRC_TABLE ? mesh_cell_kernel<GEL_SRC_TABLE, double>
                                    : mesh_cell_kernel<GEL_SRC_NONE, double>;
    }
    hipLaunchKernelGGL(fk, gf, blk, 0, s, d, a, nfb);
    hipLaunchKernelGGL(ck, gc, blk, 0, s, d, a);
    return hipGetLastError();
}
