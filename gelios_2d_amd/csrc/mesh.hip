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
    const double* now = (const double*)a.now;
#pragma unroll
    for (int v = 0; v < 5; v++) k.s[v] = now[v * nc + c];
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
// spreads its midpoints over the lanes), lane 0 stores.
template <int FX, bool WAVE>
__device__ __forceinline__ int mesh_face(const MeshDev& m, const SweepArgs& a, int f)
{
    int b = 0;
    const double gam = a.prm.gamma;
    const double dt = dt_read(a);
    const int M = m.fown[f], N = m.fnbr[f], t = m.ftype[f];
    const int A = m.fA[f], B = m.fB[f];
    const Face g = face_geom<true, double>(a.x[A], a.y[A], a.x[B], a.y[B], a.x2[A], a.y2[A], a.x2[B], a.y2[B], dt);
    const bool c1 = m.ctype[M] == GEL_C_1 || (N >= 0 && m.ctype[N] == GEL_C_1);
    const CellT<GasPhys, double> Mc = mesh_cell(m, a, M, c1), Dc = mesh_cell(m, a, m.fdown[f], c1);
    const bool god = (m.fsurf[f] & 1) != 0, np = (m.fsurf[f] & 2) != 0;
    double L[5], R[5], P[5];
    if (t == GEL_F_USUALY) {
        const CellT<GasPhys, double> Nc = mesh_cell(m, a, N, c1), Uc = mesh_cell(m, a, m.fup[f], c1);
        mesh_side(Mc, Nc, Dc, m.caxis[M] != 0, god, g, gam, L, b);
        mesh_side(Nc, Mc, Uc, m.caxis[N] != 0, god, g, gam, R, b);
    } else {
        boundary_states<GasPhys, double, double, true>(a, t, Mc, Dc, g, L, R);
    }
    double uu;
    if (god) uu = surface_flux<FX == GEL_FLUX_HLLC, WAVE>(L, R, g, gam, P, b);
    else if (FX == GEL_FLUX_HLLC && np) uu = nul_potok_flux(L, R, g, gam, P);
    else uu = face_flux<GasPhys, double, FX>(L, R, g, gam, P);
    if (WAVE && (threadIdx.x & (kWave - 1)) != 0) return 0;
    const size_t nf = m.nf;
#pragma unroll
    for (int k = 0; k < 5; k++) m.fout[k * nf + f] = P[k] * g.S;
    m.fout[5 * nf + f] = uu;
    m.fout[6 * nf + f] = g.fx;
    m.fout[7 * nf + f] = g.fy;
    return b;
}

// Work-groups [0, nfb) take 256 faces each, one per lane, skipping the fitted-surface
// faces; the work-groups after them take four fitted-surface faces each, one per
// wavefront.  nfb is a multiple of eight (xcd_item).
template <int FX>
__global__ void __launch_bounds__(256) mesh_face_kernel(MeshDev m, SweepArgs a, int nfb)
{
    int b = 0;
    if ((int)blockIdx.x < nfb) {
        const int f = xcd_item(blockIdx.x, nfb) * blockDim.x + threadIdx.x;
        if (f < m.nf && !(m.fsurf[f] & 1)) b = mesh_face<FX, false>(m, a, f);
    } else {
        const int k = ((int)blockIdx.x - nfb) * 4 + (int)(threadIdx.x / kWave);
        if (k < m.ng) b = mesh_face<FX, true>(m, a, m.gface[k]);
    }
    const int s_bad = wave_sum(b);
    if ((threadIdx.x & (kWave - 1)) == 0 && s_bad)
        __hip_atomic_fetch_add(gp(a.counters) + 2, (unsigned long long)s_bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int SRC>
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
            const double* now = (const double*)a.now;
            double* nxt = (double*)a.next;
            double s[5];
#pragma unroll
            for (int v = 0; v < 5; v++) s[v] = now[v * nc + c];
            if (m.ctype[c] == GEL_C_CENTR) {                // Setka.cpp:10461-10468
#pragma unroll
                for (int v = 0; v < 5; v++) nxt[v * nc + c] = s[v];
            } else {
                double q1 = 0.0, q2 = 0.0, q3 = 0.0;
                if constexpr (SRC == GEL_SRC_MK1) {
                    const double* sp = (const double*)a.src + c;
                    auto fk = [&](int k) { return sp[(size_t)k * nc]; };
                    source_mk1<decltype(fk), double>(a.prm, fk, s[2], s[3], s[0], s[1], q1, q2, q3);
                }
                const double dt = dt_read(a);
                double o[5];
                cell_update<GasPhys, false>(s, Pot, m.cgeo[4 * nc + c], m.cgeo[5 * nc + c], m.cgeo[2 * nc + c],
                                            m.cgeo[3 * nc + c], dt, a.prm.gamma, 0.0, q1, q2, q3, 0.0, false, o,
                                            acc.n_rho, acc.n_p);
#pragma unroll
                for (int v = 0; v < 5; v++) nxt[v * nc + c] = o[v];
            }
        }
    }
    march_flush(a, acc);
}

}  // namespace gel

using namespace gel;

// Device arrays of a mesh context: one allocation, laid out as MeshDev reads it.
struct MeshBuf {
    MeshDev dev{};
    void* base = nullptr;
};

void gel_mesh_free(gel_ctx* c)
{
    if (!c || !c->mesh) return;
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
    c->mesh = mb;
    c->mesh_faces = nf;
    return GEL_OK;
}

// One step on a mesh context: geometry (when the nodes move, or once on a static grid),
// the faces, the cells.  moving: the step's next node positions are a.x2, a.y2.
hipError_t gel::gel_mesh_step(gel_ctx* c, SweepArgs& a, bool moving)
{
    MeshDev& d = c->mesh->dev;
    hipStream_t s = c->stream;
    // grids of a multiple of eight work-groups, so that each XCD takes a contiguous
    // run of cells or faces (xcd_item) and neighbours share its L2
    auto groups = [](int n) { return (unsigned)((((n + 255) / 256) + 7) / 8 * 8); };
    if (moving || !c->geo_valid) {
        hipLaunchKernelGGL(mesh_geo_kernel, dim3(groups(d.nc)), dim3(256), 0, s, d, a.x, a.y, a.x2, a.y2);
        c->geo_valid = !moving;
    }
    const int nfb = (int)groups(d.nf);
    const dim3 gf((unsigned)(nfb + (d.ng + 3) / 4)), gc(groups(d.nc)), blk(256);
    if (c->prm.flux == GEL_FLUX_HLLC) hipLaunchKernelGGL(mesh_face_kernel<GEL_FLUX_HLLC>, gf, blk, 0, s, d, a, nfb);
    else hipLaunchKernelGGL(mesh_face_kernel<GEL_FLUX_HLL>, gf, blk, 0, s, d, a, nfb);
    if (c->prm.source_mode == GEL_SRC_MK1) hipLaunchKernelGGL(mesh_cell_kernel<GEL_SRC_MK1>, gc, blk, 0, s, d, a);
    else hipLaunchKernelGGL(mesh_cell_kernel<GEL_SRC_NONE>, gc, blk, 0, s, d, a);
    return hipGetLastError();
}
