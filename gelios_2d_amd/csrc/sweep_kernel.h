// sweep_kernel.h - the CDNA4 (gfx950) Godunov flux-and-update sweep kernels and their
// launch templates, included by gelios_sweep.hip (FP64 kernels, FP contraction off:
// the reference's arithmetic) and by sweep_fp32.hip (FP32-bulk kernels, compiled
// with FP contraction on).
//
// Kernel layout.  A work-group of four wavefronts owns a strip of 60 neighbouring
// rays (columns j) plus 2 halo columns on each side and marches up a chunk of rows
// four rows at a time (march_tile below); every face is solved once, the four faces
// of a cell are summed in the reference's face order (Setka.cpp:10372-10468), and
// the global CFL step is a wave min-reduction plus one integer atomicMin on the bit
// pattern of a positive double per wavefront, into one of the time-step slot's
// sixteen entries (deterministic: the slot's value is the minimum of its entries).
//
// Parity with the reference: the per-face arithmetic is that of sweep_core.h
// (reference operation order); solving a face once instead of once from each side,
// as the reference's cell gather does (Setka.cpp:10195), changes the flux on the
// second side only by a sign, exactly.
#pragma once

#include "sweep_core.h"
#include "gelios_internal.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <type_traits>
#include <string>
#include <vector>

namespace gel {

constexpr int kWave = 64;
constexpr int kHalo = 2;
constexpr int kOut = kWave - 2 * kHalo;   // output columns per wavefront
constexpr int kNW = 4;                    // wavefronts per work-group, marching together

// State vectors.  GAS: (ro, p, u, v, Q), the fields of Parametr (Cell.h:21-25).
// MHD: (ro, p, u, v, w, bx, by, bz, Q).  odd(k): the primitive changes sign across the
// symmetry axis (the reference mirrors v, Gran.cpp:225-233; with swirl and field the
// azimuthal velocity and the radial and azimuthal field change sign as well).
struct GasPhys {
    static constexpr int NV = 5;
    __host__ __device__ static constexpr bool odd(int k) { return k == 3; }
};
struct MhdPhys {
    static constexpr int NV = 9;
    __host__ __device__ static constexpr bool odd(int k) { return k == 3 || k == 4 || k == 6 || k == 7; }
};

template <class Ph, class TS>
struct CellT {          // one cell of the stencil as seen by a lane
    TS s[Ph::NV];
    double cx, cy;
    int type;
    int ok;             // exists (inside the grid)
    double rp;          // grids with cell types: (centroid radius)^(2 gamma), once per row load
    double ca, sa;      // cells the sweep rotates (C_1; every cell of the inner region): cos,
                        // sin of the centroid's polar angle, once per row load
};

// Help.cpp:3-14 (polar_perenos) for a stencil cell: its angle's cosine and sine come
// with the cell (cos(polar_angle(cx, cy)), sin(...), the same calls as polar_perenos).
template <class C>
__device__ __forceinline__ void rotate_cell(const C& M, double x2, double y2, double& u, double& v)
{
    const double a2 = polar_angle(x2, y2);
    const double fr = u * M.ca + v * M.sa;
    const double ff = -u * M.sa + v * M.ca;
    u = fr * cos(a2) - ff * sin(a2);
    v = fr * sin(a2) + ff * cos(a2);
}

struct Nodes {          // nodes (r, j) and (r, j+1) of a lane
    double xa, ya, xb, yb;
};

template <class T> __device__ __forceinline__ T shfl_dn(T v, int d) { return __shfl_down(v, d, kWave); }
template <class T> __device__ __forceinline__ T shfl_up(T v, int d) { return __shfl_up(v, d, kWave); }

template <bool DOWN, class Ph, class TS>
__device__ __forceinline__ CellT<Ph, TS> cell_shfl(const CellT<Ph, TS>& c, int d)
{
    CellT<Ph, TS> r;
#pragma unroll
    for (int k = 0; k < Ph::NV; k++) r.s[k] = DOWN ? shfl_dn(c.s[k], d) : shfl_up(c.s[k], d);
    r.cx = DOWN ? shfl_dn(c.cx, d) : shfl_up(c.cx, d);
    r.cy = DOWN ? shfl_dn(c.cy, d) : shfl_up(c.cy, d);
    r.type = DOWN ? shfl_dn(c.type, d) : shfl_up(c.type, d);
    r.ok = DOWN ? shfl_dn(c.ok, d) : shfl_up(c.ok, d);
    return r;
}

template <bool MOVING>
__device__ __forceinline__ Nodes load_nodes(const SweepArgs& a, const double* __restrict__ X,
                                            const double* __restrict__ Y, double sc, int r, int j, bool col)
{
    Nodes n{1.0, 1.0, 1.0, 1.0};
    if (col && r >= 0 && r <= a.ni) {
        const size_t k = (size_t)r * (a.nj + 1) + j;
        n.xa = X[k]; n.ya = Y[k]; n.xb = X[k + 1]; n.yb = Y[k + 1];
        if (MOVING) { n.xa *= sc; n.ya *= sc; n.xb *= sc; n.yb *= sc; }
    }
    return n;
}

// MF: the plasma + hydrogen sweep, whose five variables are the arrays a.voff[0..4]
// of the 21-field state (one fluid per launch).
template <class Ph, class TS, bool TYPES, bool CENTROID = true, bool MF = false>
__device__ __forceinline__ CellT<Ph, TS> load_cell(const SweepArgs& a, int r, int j, bool col,
                                                   const Nodes& lo, const Nodes& hi)
{
    CellT<Ph, TS> c;
    c.ok = (col && r >= 0 && r < a.ni) ? 1 : 0;
    c.type = GEL_C_2;
    if (c.ok) {
        const size_t nc = (size_t)a.ni * a.nj;
        const size_t k = (size_t)r * a.nj + j;
        const GPtr<const TS> now = gp((const TS*)a.now);
#pragma unroll
        for (int v = 0; v < Ph::NV; v++) c.s[v] = now[(size_t)(MF ? a.voff[v] : v) * nc + k];
        if (TYPES) c.type = gp(a.ctype)[k];
        if constexpr (CENTROID) {
            const double px[4] = { lo.xa, hi.xa, hi.xb, lo.xb };
            const double py[4] = { lo.ya, hi.ya, hi.yb, lo.yb };
            quad_center(px, py, c.cx, c.cy);
        }
    } else {
#pragma unroll
        for (int v = 0; v < Ph::NV; v++) c.s[v] = TS(0);
        c.s[0] = TS(1); c.s[1] = TS(1);
        c.cx = 1.0; c.cy = 1.0;
    }
    return c;
}

__device__ __forceinline__ double dist2d(double ax, double ay, double bx, double by)
{
    return sqn(sq(ax - bx) + sq(ay - by));
}

// Limited linear value at the face (Help.cpp:96-113 with y = 0, x1 = -d1, x2 = -d2,
// x3 = d3): t2 + minmod((t1-t2)/(x1-x2), (t2-t3)/(x2-x3)) * (0 - x2), where d1, d2,
// d3 are the distances of the three cell centres from the face centre.  FP64 keeps
// the reference's divisions.  FP32 bulk computes the distances in FP32, multiplies by
// the reciprocals of the two denominators, which all primitives of a face side
// share, and assembles the face value in FP64 from the exact cell value t2, so that
// only the small slope correction carries FP32 rounding.
struct Lin3 {
    Den r12, r23;              // |x1 - x2|, |x2 - x3| and their reciprocals
    bool n12, n23;             // x1 - x2 < 0, x2 - x3 < 0
    double d2;                 // 0 - x2
};
struct Lin3f {
    float i12, i23, d2;        // 1 / (x1 - x2), 1 / (x2 - x3), 0 - x2
};
template <class TF> using Lin3T = typename std::conditional<sizeof(TF) == 8, Lin3, Lin3f>::type;

__device__ __forceinline__ float dist2f(double ax, double ay, double bx, double by)
{
    const float dx = float(ax - bx), dy = float(ay - by);
    return sr(dx * dx + dy * dy);
}

// stencil weights at face centre (fx, fy) for centres 1 (far, behind), 2 (owner), 3 (across)
template <class TF>
__device__ __forceinline__ Lin3T<TF> lin3_at(double fx, double fy, double x1, double y1, double x2, double y2,
                                             double x3, double y3)
{
    if constexpr (sizeof(TF) == 8) {
        const double d2 = dist2d(fx, fy, x2, y2);
        const double d3 = dist2d(fx, fy, x3, y3);
        const double d1 = dist2d(fx, fy, x1, y1);
        const double den12 = -d1 - (-d2), den23 = -d2 - d3;
        return Lin3{ den(fabs(den12)), den(fabs(den23)), den12 < 0.0, den23 < 0.0, 0.0 - (-d2) };
    } else {
        const float d2 = dist2f(fx, fy, x2, y2);
        const float d3 = dist2f(fx, fy, x3, y3);
        const float d1 = dist2f(fx, fy, x1, y1);
        return Lin3f{ __builtin_amdgcn_rcpf(d2 - d1), __builtin_amdgcn_rcpf(-d2 - d3), d2 };
    }
}
template <class TF>
__device__ __forceinline__ double lin3v(double t1, double t2, double t3, const Lin3T<TF>& g)
{
    if constexpr (sizeof(TF) == 8) {
        const double d = minmod_ratio_d(t1 - t2, g.r12, g.n12, t2 - t3, g.r23, g.n23);
        return d * g.d2 + t2;
    } else {
        const float d = minmod<float>((float(t1) - float(t2)) * g.i12, (float(t2) - float(t3)) * g.i23);
        return double(d * g.d2) + t2;
    }
}

// Gran::Get_par_TVD, Usualy face (Gran.cpp:294-457): master M's value at face centre
// (fx,fy) from the line (D, M, N).  D.ok == 0: mirrored ghost across the axis when M
// touches it (:297-339), else M itself (:340-343).  Distances are FP64 geometry.
template <class Ph, class TS, class TF>
__device__ __forceinline__ void tvd_usualy(const CellT<Ph, TS>& M, const CellT<Ph, TS>& N, const CellT<Ph, TS>& D,
                                           bool m_axis, double fx, double fy, double* r, int& bad)
{
    if (!D.ok) {
#pragma unroll
        for (int k = 0; k < Ph::NV; k++) r[k] = double(M.s[k]);
        if (m_axis) {
            const Lin3T<TF> g = lin3_at<TF>(fx, fy, M.cx, -M.cy, M.cx, M.cy, N.cx, N.cy);
#pragma unroll
            for (int k = 0; k < Ph::NV; k++)
                if (Ph::odd(k)) r[k] = lin3v<TF>(-double(M.s[k]), double(M.s[k]), double(N.s[k]), g);
        }
        return;
    }
    const Lin3T<TF> g = lin3_at<TF>(fx, fy, D.cx, D.cy, M.cx, M.cy, N.cx, N.cy);
#pragma unroll
    for (int k = 0; k < Ph::NV; k++)
        r[k] = lin3v<TF>(double(D.s[k]), double(M.s[k]), double(N.s[k]), g);
    if (r[0] <= 0.0 || r[1] <= 0.0) bad = 1;
}

// Gran::Get_par_TVD_radial, density and pressure (Gran.cpp:562-675); FP64 gas only.
template <class Ph>
__device__ __forceinline__ void tvd_radial(const CellT<Ph, double>& M, const CellT<Ph, double>& N,
                                           const CellT<Ph, double>& D, bool m_axis, double fx, double fy,
                                           double gam, double& ro, double& pr)
{
    double xd, yd, dro, dp;
    if (!D.ok) {
        if (!m_axis) { ro = M.s[0]; pr = M.s[1]; return; }
        xd = M.cx; yd = -M.cy; dro = M.s[0]; dp = M.s[1];
    } else {
        xd = D.cx; yd = D.cy; dro = D.s[0]; dp = D.s[1];
    }
    const double g2 = 2.0 * gam;
    const double r = sqn(sq(fx) + sq(fy));
    const double r2 = sqn(sq(M.cx) + sq(M.cy));
    const double d2 = dist2d(fx, fy, M.cx, M.cy);
    const double r3 = sqn(sq(N.cx) + sq(N.cy));
    const double d3 = dist2d(fx, fy, N.cx, N.cy);
    const double r1 = sqn(sq(xd) + sq(yd));
    const double d1 = dist2d(fx, fy, xd, yd);
    ro = dv(linear3(-d1, dro * sq(r1), -d2, M.s[0] * sq(r2), d3, N.s[0] * sq(r3), 0.0), sq(r));
    // the cells' r^(2 gamma) come with the cells (computed once per row load as
    // pow(sqrt(cx^2 + cy^2), 2 gamma), the same call with the same argument; the mirror
    // image across the axis has the master's radius)
    const double p1 = D.ok ? D.rp : M.rp;
    pr = dv(linear3(-d1, dp * p1, -d2, M.s[1] * M.rp, d3, N.s[1] * N.rp, 0.0), pow(r, g2));
}

// Radial rescaling of a cell value to the face radius (Setka.cpp:10221-10226):
// ro and the tracer by (rc/dis)^2, p by (rc/dis)^(2 gamma); C_1 cells also rotate u,v.
template <int NV, class C>
__device__ __forceinline__ void rescale_c1(double* g, double rc, double dis, double gam, const C& M, double fx,
                                           double fy, bool rotate)
{
    g[0] = dv(g[0] * sq(rc), sq(dis));
    g[NV - 1] = dv(g[NV - 1] * sq(rc), sq(dis));
    g[1] = g[1] * pow(dv(rc, dis), 2 * gam);
    if (rotate) rotate_cell(M, fx, fy, g[2], g[3]);
}

// Value of cell M on its side of an interior face (Setka.cpp:10213-10251).
template <class Ph, class TS, class TF, bool TYPES>
__device__ __forceinline__ void side_value(const CellT<Ph, TS>& M, const CellT<Ph, TS>& N, const CellT<Ph, TS>& D,
                                           bool m_axis, double fx, double fy, double dis, double gam, double* r, int& bad)
{
    if constexpr (TYPES) {
        if (M.type == GEL_C_1) {
            double g[Ph::NV];
#pragma unroll
            for (int k = 0; k < Ph::NV; k++) g[k] = M.s[k];
            const double rc = sqn(sq(M.cx) + sq(M.cy));
            rescale_c1<Ph::NV>(g, rc, dis, gam, M, fx, fy, true);
            double ro, pr;
            tvd_radial<Ph>(M, N, D, m_axis, fx, fy, gam, ro, pr);
            g[0] = ro; g[1] = pr;
#pragma unroll
            for (int k = 0; k < Ph::NV; k++) r[k] = g[k];
            return;
        }
    }
    tvd_usualy<Ph, TS, TF>(M, N, D, m_axis, fx, fy, r, bad);
}

struct Face {          // geometry of a face as seen from its owning cell
    double n1, n2, S, fx, fy, W, dis;
    double A1, A2, WS;  // FP32 bulk: area-weighted normal S n and swept-volume rate W S
};

// CFL time of a cell through a face (Setka.cpp:10388-10393): cfl * |centre - face
// centre| / max|S|.  FP32 bulk forms it in FP32.
template <class TF>
__device__ __forceinline__ double cfl_time(double cfl, double cx, double cy, const Face& f, TF uu)
{
    if constexpr (sizeof(TF) == 8) return dv(cfl * sqn(sq(cx - f.fx) + sq(cy - f.fy)), uu);
    else return double(float(cfl) * dist2f(cx, cy, f.fx, f.fy) * __builtin_amdgcn_rcpf(uu));
}

template <bool MOVING, class TF>
__device__ __forceinline__ Face face_geom(double ax, double ay, double bx, double by,
                                          double ax2, double ay2, double bx2, double by2, double dt)
{
    Face f;
    f.S = sqn(sq(bx - ax) + sq(by - ay)) * (ay + by) * 0.5;            // Setka.cpp:10198
    f.fx = (ax + bx) / 2.0;
    f.fy = (ay + by) / 2.0;
    double t1 = bx - ax, t2 = by - ay;                                    // Gran.cpp:171-181
    const double nn = sqn(sq(t1) + sq(t2));
    if constexpr (sizeof(TF) == 8) {
        t1 = dv(t1, nn); t2 = dv(t2, nn);
    } else {                            // FP32 bulk: one FP64 reciprocal
        const double inn = den(nn).r;
        t1 = t1 * inn; t2 = t2 * inn;
    }
    f.n1 = t2; f.n2 = -t1;
    if (MOVING) {
        const double fx2 = (ax2 + bx2) / 2.0, fy2 = (ay2 + by2) / 2.0;
        f.W = dv((fx2 - f.fx) * f.n1 + (fy2 - f.fy) * f.n2, dt);          // Setka.cpp:10209
    } else {
        f.W = 0.0;
    }
    f.dis = sqn(sq(f.fx) + sq(f.fy));
    return f;
}

// FP32 bulk: a face from its two nodes with no FP64 square root or division.  What
// enters the cell sums is FP64: the area-weighted normal A = S n = (by - ay, -(bx - ax))
// (ay + by) / 2 (the reference's S = |b - a| (ay + by) / 2 times its n, Setka.cpp:10198,
// Gran.cpp:171-181) and the swept-volume rate W S (Setka.cpp:10209 times S).  The unit
// normal, S and W themselves, which only the FP32 Riemann solver and the normal-field
// term use, come from the FP32 reciprocal square root.  idt = 1 / dt.  dis is left for
// the caller (only the inner sphere's faces use it).
template <bool MOVING>
__device__ __forceinline__ Face face_geom_a(double ax, double ay, double bx, double by,
                                            double ax2, double ay2, double bx2, double by2, double idt)
{
    Face f;
    const double dx = bx - ax, dy = by - ay, h = (ay + by) * 0.5;
    f.A1 = dy * h;
    f.A2 = -dx * h;
    f.fx = (ax + bx) * 0.5;
    f.fy = h;
    const float l2 = float(dx * dx + dy * dy);
    const float il = __builtin_amdgcn_rsqf(l2);
    f.n1 = double(float(dy) * il);
    f.n2 = double(-float(dx) * il);
    f.S = double(l2 * il * float(h));
    if constexpr (MOVING) {
        const double ex = (ax2 + bx2) * 0.5 - f.fx, ey = (ay2 + by2) * 0.5 - f.fy;
        f.WS = (ex * f.A1 + ey * f.A2) * idt;
        f.W = (ex * f.n1 + ey * f.n2) * idt;
    } else {
        f.WS = 0.0;
        f.W = 0.0;
    }
    f.dis = 1.0;
    return f;
}

// Face flux from left/right face primitives (FP64); P in the order of the cell
// accumulators.  TF = double: the reference arithmetic.  TF = float (FP32 bulk): the
// wave speeds are estimated in FP32; a supersonic face (the whole solar wind) takes
// the exact FP64 upwind flux, a subsonic one the FP32 Riemann solver.  Rounding the
// flux of a steady supersonic stream to FP32 would bias every step the same way.
// FP32 bulk returns P already times the face area (face_geom_a's A, WS and S); FP64
// returns the flux per unit area, which the caller multiplies by S as the reference does.
template <class Ph, class TF, int FLUX>
__device__ __forceinline__ TF face_flux(const double* L, const double* R, const Face& f, double gam, double* P)
{
    if constexpr (Ph::NV == 5) {
        const Gas<double> gl{ L[0], L[1], L[2], L[3], L[4] };
        const Gas<double> gr{ R[0], R[1], R[2], R[3], R[4] };
        if constexpr (sizeof(TF) == 8) {
            double PQ;
            const double UU = hllc2d<double, FLUX == GEL_FLUX_HLLC>(gl, gr, f.W, f.n1, f.n2, gam, P, PQ);
            P[4] = PQ;
            return UU;
        } else {
            const float n1 = float(f.n1), n2 = float(f.n2), W = float(f.W), g = float(gam);
            const float u1 = float(L[2]) * n1 + float(L[3]) * n2, u2 = float(R[2]) * n1 + float(R[3]) * n2;
            const float cL = sr(dv(g * float(L[1]), float(L[0]))), cR = sr(dv(g * float(R[1]), float(R[0])));
            const float SL = mn(u1, u2) - mx(cL, cR), SR = mx(u1, u2) + mx(cL, cR);
            if (SL >= W || SR <= W) {
                gas_upwind_area<double>(SL >= W ? gl : gr, f.A1, f.A2, f.WS, den(gam - 1.0).r, P);
                return mx(fabsf(SL), fabsf(SR));
            }
            const Gas<float> fl{ float(L[0]), float(L[1]), float(L[2]), float(L[3]), float(L[4]) };
            const Gas<float> fr{ float(R[0]), float(R[1]), float(R[2]), float(R[3]), float(R[4]) };
            float Pf[4], PQ;
            const float UU = hllc2d<float, FLUX == GEL_FLUX_HLLC>(fl, fr, W, n1, n2, g, Pf, PQ);
#pragma unroll
            for (int k = 0; k < 4; k++) P[k] = double(Pf[k]) * f.S;
            P[4] = double(PQ) * f.S;
            return UU;
        }
    } else {
        if constexpr (sizeof(TF) == 8) {
            return hlld<double>(L, R, f.W, f.n1, f.n2, gam, P);
        } else {
            const float n1 = float(f.n1), n2 = float(f.n2), W = float(f.W), g = float(gam);
            const float ispi4 = 0.28209479177387814f, i4pi = 0.07957747154594767f;   // 1/sqrt(4 pi), 1/(4 pi)
            const float bnL = (float(L[5]) * n1 + float(L[6]) * n2) * ispi4;
            const float bnR = (float(R[5]) * n1 + float(R[6]) * n2) * ispi4;
            const float bbL = (sq(float(L[5])) + sq(float(L[6])) + sq(float(L[7]))) * i4pi;
            const float bbR = (sq(float(R[5])) + sq(float(R[6])) + sq(float(R[7]))) * i4pi;
            const float gpL = g * float(L[1]), gpR = g * float(R[1]);
            const float cfL = sr(dv(gpL + bbL + sr(fmaxf(sq(gpL + bbL) - 4.0f * gpL * sq(bnL), 0.0f)), 2.0f * float(L[0])));
            const float cfR = sr(dv(gpR + bbR + sr(fmaxf(sq(gpR + bbR) - 4.0f * gpR * sq(bnR), 0.0f)), 2.0f * float(R[0])));
            const float u1 = float(L[2]) * n1 + float(L[3]) * n2, u2 = float(R[2]) * n1 + float(R[3]) * n2;
            const float SL = mn(u1, u2) - mx(cfL, cfR), SR = mx(u1, u2) + mx(cfL, cfR);
            const float SN = mx(fabsf(SL), fabsf(SR));
            if (SL >= W || SR <= W) {
                const bool left = SL >= W;
                mhd_upwind_area<double>(left ? L : R, left ? R : L, f.A1, f.A2, f.WS, f.n1, f.n2, den(gam - 1.0).r,
                                        double(SN), left, P);
                return SN;
            }
            float fl[9], fr[9], Pf[9];
#pragma unroll
            for (int k = 0; k < 9; k++) { fl[k] = float(L[k]); fr[k] = float(R[k]); }
            const float UU = hlld<float>(fl, fr, W, n1, n2, g, Pf);
#pragma unroll
            for (int k = 0; k < 9; k++) P[k] = double(Pf[k]) * f.S;
            return UU;
        }
    }
}

// Boundary face of cell M: Gran::Get_par ghost (Gran.cpp:205-289) and the sweep's
// treatment of the cell side (Setka.cpp:10252-10274).  D is the cell across M from
// the face (Sosed_down), used on the axis.  L is M's side, R the ghost.
template <class Ph, class TS, class TF, bool TYPES>
__device__ __forceinline__ void boundary_states(const SweepArgs& a, int ftype, const CellT<Ph, TS>& M,
                                                const CellT<Ph, TS>& D, const Face& f, double* L, double* R)
{
    constexpr int NV = Ph::NV;
    const gel_params& p = a.prm;
    const double gam = p.gamma;
    const double radius = sqn(sq(M.cx) + sq(M.cy));
    if (ftype == GEL_F_INNER_SPHERE) {
        double g[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) g[k] = M.s[k];
        rescale_c1<NV>(g, radius, f.dis, gam, M, f.fx, f.fy, false);
        const double d = sqn(f.fx * f.fx + f.fy * f.fy);
        const double ro = dv(1.0, p.in_rho_div * d * d);
        double w[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) w[k] = 0.0;
        w[0] = ro;
        w[1] = dv(ro * p.in_p_a * p.in_p_a, p.in_p_den);
        w[2] = dv(p.in_u * f.fx, d);
        w[3] = dv(p.in_u * f.fy, d);
        w[NV - 1] = ro * p.in_q;
#pragma unroll
        for (int k = 0; k < NV; k++) { L[k] = g[k]; R[k] = w[k]; }
    } else if (ftype == GEL_F_AXIS) {
        bool done = false;
        if constexpr (TYPES) {
            if (M.type == GEL_C_1) {
                double g[NV];
#pragma unroll
                for (int k = 0; k < NV; k++) g[k] = M.s[k];
                rescale_c1<NV>(g, radius, f.dis, gam, M, f.fx, f.fy, true);
#pragma unroll
                for (int k = 0; k < NV; k++) L[k] = g[k];
                done = true;
            }
        }
        if (!done) {                                          // Gran.cpp:486-519
            // the mirror image of M lies as far across the face as M itself
            const Lin3T<TF> g = lin3_at<TF>(f.fx, f.fy, D.cx, D.cy, M.cx, M.cy, M.cx, M.cy);
#pragma unroll
            for (int k = 0; k < NV; k++) {
                L[k] = double(M.s[k]);
                if (Ph::odd(k)) L[k] = lin3v<TF>(double(D.s[k]), double(M.s[k]), -double(M.s[k]), g);
            }
        }
#pragma unroll
        for (int k = 0; k < NV; k++) R[k] = Ph::odd(k) ? -L[k] : L[k];
    } else {
#pragma unroll
        for (int k = 0; k < NV; k++) { L[k] = double(M.s[k]); R[k] = L[k]; }
        if (ftype == GEL_F_INPUT) {
            R[0] = p.inf_rho; R[1] = p.inf_p; R[2] = p.inf_u; R[3] = p.inf_v;
            R[NV - 1] = p.inf_q;
            if constexpr (NV == 9) {
                R[4] = p.inf_w; R[5] = p.inf_bx; R[6] = p.inf_by; R[7] = p.inf_bz;
            }
        } else if (ftype == GEL_F_EXTERN) {
            if (R[2] >= 0.0) R[2] = p.extern_u;
        }
    }
}

// Interior face between M (owner, L side) and N (R side); D lies beyond M and U beyond
// N on the same line (Setka.cpp:10213-10251 from each side).
template <class Ph, class TS, class TF, bool TYPES>
__device__ __forceinline__ void interior_states(const CellT<Ph, TS>& M, const CellT<Ph, TS>& N,
                                                const CellT<Ph, TS>& D, const CellT<Ph, TS>& U, bool m_axis,
                                                bool n_axis, const Face& f, double gam, double* L, double* R,
                                                int& bad)
{
    side_value<Ph, TS, TF, TYPES>(M, N, D, m_axis, f.fx, f.fy, f.dis, gam, L, bad);
    side_value<Ph, TS, TF, TYPES>(N, M, U, n_axis, f.fx, f.fy, f.dis, gam, R, bad);
}

// ------------------------------------------------------------ fitted surfaces ---
// Faces between two nodes of one fitted surface (the radial faces of a surface row)
// and faces with one node on a surface (Setka.cpp:10287-10370).  Gas, FP64.

// Gran::Get_par_TVD on a surface face (Gran.cpp:345-426): the straight line through
// D (behind the master) and M at the face centre; a non-positive density or pressure
// takes M's value.
template <class Ph>
__device__ __forceinline__ void tvd_surface(const CellT<Ph, double>& M, const CellT<Ph, double>& D, double fx,
                                            double fy, double* r)
{
    const double d2 = dist2d(fx, fy, M.cx, M.cy);
    const double d1 = dist2d(fx, fy, D.cx, D.cy);
    const double x1 = -d1, x2 = -d2;
#pragma unroll
    for (int k = 0; k < Ph::NV; k++) {
        const double d = (D.s[k] - M.s[k]) / (x1 - x2);            // Help.cpp:115-120, y = 0
        r[k] = d * (0.0 - x2) + M.s[k];
    }
    if (r[0] <= 0.0) r[0] = M.s[0];
    if (r[1] <= 0.0) r[1] = M.s[1];
}

// Both sides of a surface face: C_1 sides as on any face (rescaled, radial
// reconstruction), the others one-sided.
template <class Ph, bool TYPES>
__device__ __forceinline__ void surface_states(const CellT<Ph, double>& M, const CellT<Ph, double>& N,
                                               const CellT<Ph, double>& D, const CellT<Ph, double>& U,
                                               bool axis, const Face& f, double gam, double* L, double* R, int& bad)
{
    if (TYPES && M.type == GEL_C_1) side_value<Ph, double, double, TYPES>(M, N, D, axis, f.fx, f.fy, f.dis, gam, L, bad);
    else tvd_surface<Ph>(M, D, f.fx, f.fy, L);
    if (TYPES && N.type == GEL_C_1) side_value<Ph, double, double, TYPES>(N, M, U, axis, f.fx, f.fy, f.dis, gam, R, bad);
    else tvd_surface<Ph>(N, U, f.fx, f.fy, R);
}

// Flux through a surface face (Setka.cpp:10335-10370): mass, momentum and energy from
// the exact solver at the face speed, the tracer from the HLL/HLLC flux; returns the
// larger of the two wave-speed maxima, from which cfl * dist / max equals the smaller
// of the two time steps the reference takes.
// The nul_potok HLLC of the faces beside a surface.
__device__ __forceinline__ double nul_potok_flux(const double* L, const double* R, const Face& f, double gam, double* P)
{
    const Gas<double> gl{ L[0], L[1], L[2], L[3], L[4] };
    const Gas<double> gr{ R[0], R[1], R[2], R[3], R[4] };
    double PQ;
    const double uu = hllc2d_nul<double>(gl, gr, f.W, f.n1, f.n2, gam, P, PQ);
    P[4] = PQ;
    return uu;
}

// WAVE: the whole wavefront solves this one face (godunov_exact<true>).
template <bool HLLC, bool WAVE = false>
__device__ __forceinline__ double surface_flux(const double* L, const double* R, const Face& f, double gam,
                                               double* P, int& bad)
{
    const double q1[5] = { L[0], L[1], L[2], L[3], 0.0 };
    const double q2[5] = { R[0], R[1], R[2], R[3], 0.0 };
    double q[5], dsl, dsp, dsc, uug = 0.0;
    int st = 0;
    godunov_exact<WAVE>(q1, q2, f.n1, f.n2, 0.0, f.W, 1.0, 1.0, gam, q, dsl, dsp, dsc, st, &uug);
    if (st) bad += 1;
    const Gas<double> gl{ L[0], L[1], L[2], L[3], L[4] };
    const Gas<double> gr{ R[0], R[1], R[2], R[3], R[4] };
    double Ph_[4], PQ;
    const double uuh = hllc2d<double, HLLC>(gl, gr, f.W, f.n1, f.n2, gam, Ph_, PQ);
    P[0] = q[0]; P[1] = q[1]; P[2] = q[2]; P[3] = q[4]; P[4] = PQ;
    return mx(uug, uuh);
}

// Conservative ALE update of one cell and primitive recovery (Setka.cpp:10423-10460),
// r-weighted: Volume = int y dA before, Volume2 after the step; Ar = planar area.
// MHD adds the axisymmetric geometric sources (hoop stress with field and swirl,
// angular momentum, azimuthal induction) per unit planar area.
// FAST (FP32 bulk): the same update with the divisions by Volume2 and by the new
// density taken as multiplications by reciprocals (FP64 throughout).
template <class Ph, bool FAST>
__device__ __forceinline__ void cell_update(const double* s, const double* Pot, double Volume, double Volume2,
                                            double vs1, double vs2, double T1, double gam, double qm, double q1,
                                            double q2, double q3, double qQ, bool table, double* o, int& n_rho,
                                            int& n_p)
{
    if constexpr (FAST) {
        constexpr int NV = Ph::NV;
        const double iV2 = den(Volume2).r, rV = Volume * iV2, ig1 = den(gam - 1.0).r;
        const double ro = s[0], p = s[1], u = s[2], v = s[3];
        const double Ar = 0.5 * (vs1 + vs2);
        double w = 0.0, bx = 0.0, by = 0.0, bz = 0.0, BB = 0.0, Sv = p, Sw = 0.0, Sb = 0.0;
        if constexpr (NV == 9) {
            constexpr double i4pi = 0.07957747154594767, i8pi = 0.039788735772973836;
            w = s[4]; bx = s[5]; by = s[6]; bz = s[7];
            BB = bx * bx + by * by + bz * bz;
            Sv = ro * w * w + p + BB * i8pi - bz * bz * i4pi;
            Sw = -(ro * v * w - by * bz * i4pi);
            Sb = v * bz - w * by;
        }
        const double Q = s[NV - 1];
        const double E = p * ig1 + ro * (u * u + v * v + w * w) * 0.5 + BB * 0.039788735772973836;
        double ro3 = ro * rV - T1 * (Pot[0] * iV2 - qm);
        if (ro3 <= 0.0) { ro3 = 0.00001; n_rho++; }
        const double iro3 = den(ro3).r;
        const double pv = (NV == 9) ? Sv : p;   // gas: the hoop term is p (Setka.cpp:10440)
        const double u3 = (ro * u * rV - T1 * (Pot[1] * iV2 - q1)) * iro3;
        const double v3 = (ro * v * rV - T1 * ((Pot[2] - pv * Ar) * iV2 - q2)) * iro3;
        double w3 = 0.0, bx3 = 0.0, by3 = 0.0, bz3 = 0.0;
        const int ie = (NV == 9) ? 4 : 3;
        if constexpr (NV == 9) {
            w3 = (ro * w * rV - T1 * ((Pot[3] - Sw * Ar) * iV2)) * iro3;
            bx3 = bx * rV - T1 * (Pot[5] * iV2);
            by3 = by * rV - T1 * (Pot[6] * iV2);
            bz3 = bz * rV - T1 * ((Pot[7] - Sb * Ar) * iV2);
        }
        const double E3 = E * rV - T1 * (Pot[ie] * iV2 - q3);
        double Q3 = Q * rV - T1 * (Pot[NV - 1] * iV2);
        if (table) Q3 = Q3 + T1 * qQ;
        double p3 = (E3 - 0.5 * ro3 * (u3 * u3 + v3 * v3 + w3 * w3) - (bx3 * bx3 + by3 * by3 + bz3 * bz3) * 0.039788735772973836) * (gam - 1);
        if (p3 <= 0) { p3 = 0.000001; n_p++; }
        o[0] = ro3; o[1] = p3; o[2] = u3; o[3] = v3;
        if constexpr (NV == 9) { o[4] = w3; o[5] = bx3; o[6] = by3; o[7] = bz3; }
        o[NV - 1] = Q3;
    } else if constexpr (Ph::NV == 5) {
        const double ro = s[0], p = s[1], u = s[2], v = s[3], Q = s[4];
        double ro3 = dv(ro * Volume, Volume2) - T1 * (dv(Pot[0], Volume2) - qm);
        double Q33 = dv(Q * Volume, Volume2) - dv(T1, Volume2) * Pot[4];
        if (table) Q33 = Q33 + T1 * qQ;
        if (ro3 <= 0.0) { ro3 = 0.00001; n_rho++; }
        const double u3 = dv(dv(ro * u * Volume, Volume2) - T1 * (dv(Pot[1], Volume2) - q1), ro3);
        const double v3 = dv(dv(ro * v * Volume, Volume2) - T1 * (dv(Pot[2] - p * 0.5 * (vs1 + vs2), Volume2) - q2), ro3);
        double p3 = ((dv((dv(p, gam - 1) + ro * (u * u + v * v) * 0.5) * Volume, Volume2) - T1 * (dv(Pot[3], Volume2) - q3)) -
                     0.5 * ro3 * (u3 * u3 + v3 * v3)) * (gam - 1);
        if (p3 <= 0) { p3 = 0.000001; n_p++; }
        o[0] = ro3; o[1] = p3; o[2] = u3; o[3] = v3; o[4] = Q33;
    } else {
        const double pi4 = 4.0 * kPi, pi8 = 8.0 * kPi;
        const double ro = s[0], p = s[1], u = s[2], v = s[3], w = s[4];
        const double bx = s[5], by = s[6], bz = s[7], Q = s[8];
        const double Ar = 0.5 * (vs1 + vs2);
        const double BB = bx * bx + by * by + bz * bz;
        const double Sv = ro * w * w + p + dv(BB, pi8) - dv(bz * bz, pi4);
        const double Sw = -(ro * v * w - dv(by * bz, pi4));
        const double Sb = v * bz - w * by;
        const double E = dv(p, gam - 1) + ro * (u * u + v * v + w * w) * 0.5 + dv(BB, pi8);
        double ro3 = dv(ro * Volume, Volume2) - T1 * (dv(Pot[0], Volume2) - qm);
        if (ro3 <= 0.0) { ro3 = 0.00001; n_rho++; }
        const double u3 = dv(dv(ro * u * Volume, Volume2) - T1 * (dv(Pot[1], Volume2) - q1), ro3);
        const double v3 = dv(dv(ro * v * Volume, Volume2) - T1 * (dv(Pot[2] - Sv * Ar, Volume2) - q2), ro3);
        const double w3 = dv(dv(ro * w * Volume, Volume2) - T1 * dv(Pot[3] - Sw * Ar, Volume2), ro3);
        const double E3 = dv(E * Volume, Volume2) - T1 * (dv(Pot[4], Volume2) - q3);
        const double bx3 = dv(bx * Volume, Volume2) - T1 * dv(Pot[5], Volume2);
        const double by3 = dv(by * Volume, Volume2) - T1 * dv(Pot[6], Volume2);
        const double bz3 = dv(bz * Volume, Volume2) - T1 * dv(Pot[7] - Sb * Ar, Volume2);
        const double Q3 = dv(Q * Volume, Volume2) - T1 * (dv(Pot[8], Volume2) - qQ);
        double p3 = (E3 - 0.5 * ro3 * (u3 * u3 + v3 * v3 + w3 * w3) - dv(bx3 * bx3 + by3 * by3 + bz3 * bz3, pi8)) * (gam - 1);
        if (p3 <= 0) { p3 = 0.000001; n_p++; }
        o[0] = ro3; o[1] = p3; o[2] = u3; o[3] = v3; o[4] = w3;
        o[5] = bx3; o[6] = by3; o[7] = bz3; o[8] = Q3;
    }
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmin(v, __shfl_xor(v, d, kWave));
    return v;
}
__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// ------------------------------------------------------------ geometry table ---
// Static grids read their geometry from a table built once (geometry_kernel) with the
// same functions and the same node order the kernels use on moving grids, so the
// values are identical.  Layout, FP64, structure of arrays:
//   cell  [4][ni*nj]        centroid x, y (Cell::Get_Center), planar area vs1
//                           (Get_Volume), r-weighted volume (Get_Volume_rotate / 2 pi);
//   rad   [6][(ni+1)*nj]    the radial face at node row r, column j, as edge 1 of
//                           cell (r-1, j) (for r = 0 as edge 3 of cell (0, j), the
//                           inner sphere): n1, n2, S, fx, fy, dis;
//   pol   [6][ni*(nj+1)]    the polar face of row i at node column c, as edge 2 of
//                           cell (i, c-1) (for c = 0 as edge 0 of cell (i, 0), the
//                           axis): n1, n2, S, fx, fy, dis.
struct GeoTable {
    GPtr<const double> base;
    size_t nc, nr, np;
    __device__ GeoTable(const double* b, int ni, int nj)
        : base(gp(b)), nc((size_t)ni * nj), nr((size_t)(ni + 1) * nj), np((size_t)ni * (nj + 1)) {}
    __device__ GPtr<const double> cell(int k) const { return base + (size_t)k * nc; }
    __device__ GPtr<const double> rad(int k) const { return base + 4 * nc + (size_t)k * nr; }
    __device__ GPtr<const double> pol(int k) const { return base + 4 * nc + 6 * nr + (size_t)k * np; }
    __host__ __device__ static size_t doubles(int ni, int nj)
    {
        return 4 * (size_t)ni * nj + 6 * (size_t)(ni + 1) * nj + 6 * (size_t)ni * (nj + 1);
    }
};

// One thread per node position (r, c), r in [0, ni], c in [0, nj]: the cell (r, c),
// the radial face at (r, c) and the polar face at (r, c) where they exist.
template <class TF>
__global__ void geometry_kernel(int ni, int nj, const double* __restrict__ X, const double* __restrict__ Y,
                                double* __restrict__ geo)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)(ni + 1) * (nj + 1)) return;
    const int r = (int)(t / (nj + 1)), c = (int)(t % (nj + 1));
    const GeoTable G(geo, ni, nj);
    double* cell[4];
    double* rad[6];
    double* pol[6];
    for (int k = 0; k < 4; k++) cell[k] = (double*)G.cell(k);
    for (int k = 0; k < 6; k++) { rad[k] = (double*)G.rad(k); pol[k] = (double*)G.pol(k); }
    auto nd = [&](int q, int cc, double& x, double& y) {
        const size_t k = (size_t)q * (nj + 1) + cc;
        x = X[k]; y = Y[k];
    };
    auto store = [&](double* const* F, size_t k, const Face& f) {
        F[0][k] = f.n1; F[1][k] = f.n2; F[2][k] = f.S; F[3][k] = f.fx; F[4][k] = f.fy; F[5][k] = f.dis;
    };
    if (r < ni && c < nj) {
        double x0, y0, x1, y1, x2, y2, x3, y3;    // (r, c), (r+1, c), (r+1, c+1), (r, c+1)
        nd(r, c, x0, y0); nd(r + 1, c, x1, y1); nd(r + 1, c + 1, x2, y2); nd(r, c + 1, x3, y3);
        const double px[4] = { x0, x1, x2, x3 };
        const double py[4] = { y0, y1, y2, y3 };
        const size_t k = (size_t)r * nj + c;
        double cx, cy;
        quad_center(px, py, cx, cy);
        cell[0][k] = cx;
        cell[1][k] = cy;
        cell[2][k] = quad_area(px, py);
        cell[3][k] = quad_volume_r(px, py);
    }
    if (c < nj) {                                  // radial face at node row r
        double ax, ay, bx, by;
        nd(r, c, ax, ay); nd(r, c + 1, bx, by);
        if (r == 0) { double t0 = ax; ax = bx; bx = t0; t0 = ay; ay = by; by = t0; }
        store(rad, (size_t)r * nj + c, face_geom<false, TF>(ax, ay, bx, by, ax, ay, bx, by, 1.0));
    }
    if (r < ni) {                                  // polar face of row r at node column c
        double ax, ay, bx, by;
        nd(r + 1, c, ax, ay); nd(r, c, bx, by);
        if (c == 0) { double t0 = ax; ax = bx; bx = t0; t0 = ay; ay = by; by = t0; }
        store(pol, (size_t)r * (nj + 1) + c, face_geom<false, TF>(ax, ay, bx, by, ax, ay, bx, by, 1.0));
    }
}

// ------------------------------------------------- plasma + four hydrogen fluids ---
// Setka::Go_5_komponent_2 (Setka.cpp:8951-9777) as five launches of the gas sweep, one
// per fluid (SweepArgs::fluid): 0 the plasma (ro, p, u, v, Q); k = 1..4 hydrogen Hk
// (ro_Hk, p_Hk, u_Hk, v_Hk, and the plasma tracer in the fifth slot: the reference
// passes it to the hydrogen fluxes and drops their tracer flux).  FP64.

// Gran::Get_par_TVD_radial's H1 fields (Gran.cpp:562-675): ro_H1 on the line through
// r^H_pow-weighted values, p_H1 on the plain line.
__device__ __forceinline__ void mf_tvd_radial_h1(const CellT<GasPhys, double>& M, const CellT<GasPhys, double>& N,
                                                 const CellT<GasPhys, double>& D, bool m_axis, double fx, double fy,
                                                 double hp, double& ro, double& ph)
{
    double xd, yd, dro, dp;
    if (!D.ok) {
        if (!m_axis) { ro = M.s[0]; ph = M.s[1]; return; }
        xd = M.cx; yd = -M.cy; dro = M.s[0]; dp = M.s[1];
    } else {
        xd = D.cx; yd = D.cy; dro = D.s[0]; dp = D.s[1];
    }
    const double r = sqrt(sq(fx) + sq(fy));
    const double r2 = sqrt(sq(M.cx) + sq(M.cy));
    const double d2 = dist2d(fx, fy, M.cx, M.cy);
    const double r3 = sqrt(sq(N.cx) + sq(N.cy));
    const double d3 = dist2d(fx, fy, N.cx, N.cy);
    const double r1 = sqrt(sq(xd) + sq(yd));
    const double d1 = dist2d(fx, fy, xd, yd);
    ph = linear3(-d1, dp, -d2, M.s[1], d3, N.s[1], 0.0);
    ro = linear3(-d1, dro * pow(r1, hp), -d2, M.s[0] * pow(r2, hp), d3, N.s[0] * pow(r3, hp), 0.0) / pow(r, hp);
}

// Value of M on its side of an interior face for the fluid of the launch
// (Setka.cpp:9014-9070): C_1 cells give the plasma its rescaled radial reconstruction,
// H1 its rotated velocity with the radial ro_H1, p_H1, H2..H4 their cell values; other
// cells the limited line of every field.  Non-positive reconstructions count only for
// the plasma (the reference exits there, Gran.cpp:444-455).
template <bool TYPES>
__device__ __forceinline__ void mf_side_value(const SweepArgs& a, const CellT<GasPhys, double>& M,
                                              const CellT<GasPhys, double>& N, const CellT<GasPhys, double>& D,
                                              bool m_axis, double fx, double fy, double dis, double* r, int& bad)
{
    if (TYPES && M.type == GEL_C_1) {
        if (a.fluid == 0) {
            side_value<GasPhys, double, double, TYPES>(M, N, D, m_axis, fx, fy, dis, a.prm.gamma, r, bad);
            return;
        }
#pragma unroll
        for (int k = 0; k < 5; k++) r[k] = M.s[k];
        if (a.fluid == 1) {
            rotate_cell(M, fx, fy, r[2], r[3]);
            mf_tvd_radial_h1(M, N, D, m_axis, fx, fy, a.prm.h_pow, r[0], r[1]);
        }
        return;
    }
    int b = 0;
    tvd_usualy<GasPhys, double, double>(M, N, D, m_axis, fx, fy, r, b);
    if (a.fluid == 0) bad += b;
}

template <bool TYPES>
__device__ __forceinline__ void mf_interior_states(const SweepArgs& a, const CellT<GasPhys, double>& M,
                                                   const CellT<GasPhys, double>& N, const CellT<GasPhys, double>& D,
                                                   const CellT<GasPhys, double>& U, bool m_axis, bool n_axis,
                                                   const Face& f, double* L, double* R, int& bad)
{
    mf_side_value<TYPES>(a, M, N, D, m_axis, f.fx, f.fy, f.dis, L, bad);
    mf_side_value<TYPES>(a, N, M, U, n_axis, f.fx, f.fy, f.dis, R, bad);
}

// Boundary faces for the fluid of the launch (Gran.cpp:205-289, Setka.cpp:9071-9104):
// the plasma as the single-fluid sweep; H1 rescales ro_H1 on the inner sphere and on the
// axis of C_1 cells and meets the hydrogen inflow of the inner sphere; H4 meets the
// interstellar hydrogen on Input faces; the other ghosts copy the cell.
template <bool TYPES>
__device__ __forceinline__ void mf_boundary_states(const SweepArgs& a, int ftype, const CellT<GasPhys, double>& M,
                                                   const CellT<GasPhys, double>& D, const Face& f, double* L,
                                                   double* R)
{
    if (a.fluid == 0) {
        boundary_states<GasPhys, double, double, TYPES>(a, ftype, M, D, f, L, R);
        return;
    }
    const gel_params& p = a.prm;
#pragma unroll
    for (int k = 0; k < 5; k++) L[k] = M.s[k];
    const double radius = sqn(sq(M.cx) + sq(M.cy));
    if (ftype == GEL_F_INNER_SPHERE) {
        if (a.fluid == 1) L[0] = L[0] * pow(radius / f.dis, p.h_pow);
#pragma unroll
        for (int k = 0; k < 5; k++) R[k] = M.s[k];
        if (a.fluid == 1) {
            const double d = sqrt(f.fx * f.fx + f.fy * f.fy);
            R[0] = p.h1_in_rho; R[1] = p.h1_in_p; R[2] = p.in_u * f.fx / d; R[3] = p.in_u * f.fy / d;
        }
    } else if (ftype == GEL_F_AXIS) {
        if (TYPES && M.type == GEL_C_1 && a.fluid == 1) {
            L[0] = L[0] * pow(radius / f.dis, p.h_pow);
            rotate_cell(M, f.fx, f.fy, L[2], L[3]);
        }
#pragma unroll
        for (int k = 0; k < 5; k++) R[k] = (k == 3) ? -L[k] : L[k];
    } else {
#pragma unroll
        for (int k = 0; k < 5; k++) R[k] = L[k];
        if (ftype == GEL_F_INPUT && a.fluid == 4) {
            R[0] = p.h4_inf_rho; R[1] = p.h4_inf_p; R[2] = p.h4_inf_u; R[3] = p.h4_inf_v;
        }
    }
}

// Charge exchange from the cell's 21 values (Setka.cpp:9419-9470, :9554-9617), for the
// five fluids at once (the rates nu are the plasma's with each population): q[0] for
// the plasma = (0, momentum x, momentum y, energy); q[1 + k] for hydrogen fluid H(k+1)
// its mass, momentum and energy sources, the population of the cell's region (C_1 /
// C_centr H1, C_2 / C_3 H2, C_4 H3, else H4) receiving what the plasma loses.  The
// arithmetic of each value is the per-fluid formula's, so that computing them together
// changes no bit.
__device__ __forceinline__ void mf_sources_all(const gel_params& prm, const double* __restrict__ st, size_t nc,
                                               int ct, double q[5][4])
{
    const double ro = st[0], p = st[nc], u = st[2 * nc], v = st[3 * nc];
    double roH[4], pH[4], uH[4], vH[4], UM[4], UH[4], nu[4];
#pragma unroll
    for (int h = 0; h < 4; h++) {
        roH[h] = st[(size_t)(5 + 4 * h) * nc]; pH[h] = st[(size_t)(6 + 4 * h) * nc];
        uH[h] = st[(size_t)(7 + 4 * h) * nc]; vH[h] = st[(size_t)(8 + 4 * h) * nc];
    }
    const double pi = 3.14159265358979323846;         // Solvers.h:9
#pragma unroll
    for (int h = 0; h < 4; h++) {
        UM[h] = sqrt(sq(u - uH[h]) + sq(v - vH[h]) + (64.0 / (9.0 * pi)) * (p / ro + 2.0 * pH[h] / roH[h]));
        UH[h] = sqrt(sq(u - uH[h]) + sq(v - vH[h]) + (4.0 / pi) * (p / ro + 2.0 * pH[h] / roH[h]));
        nu[h] = ro * roH[h] * UM[h] * sq(prm.sig_a - prm.sig_b * log(UM[h]));
    }
    {
        const double cp = prm.src_coef;
        q[0][0] = 0.0;
        q[0][1] = cp * (nu[0] * (uH[0] - u) + nu[1] * (uH[1] - u) + nu[2] * (uH[2] - u) + nu[3] * (uH[3] - u));
        q[0][2] = cp * (nu[0] * (vH[0] - v) + nu[1] * (vH[1] - v) + nu[2] * (vH[2] - v) + nu[3] * (vH[3] - v));
        double e[4];
#pragma unroll
        for (int h = 0; h < 4; h++)
            e[h] = nu[h] * ((sq(uH[h]) + sq(vH[h]) - sq(u) - sq(v)) / 2.0 + (UH[h] / UM[h]) * (2.0 * pH[h] / roH[h] - p / ro));
        q[0][3] = cp * (e[0] + e[1] + e[2] + e[3]);
    }
    int own;                                          // the population of the cell's region
    if (ct == GEL_C_CENTR || ct == GEL_C_1 || ct == GEL_C_2 || ct == GEL_C_3)
        own = (ct == GEL_C_1 || ct == GEL_C_CENTR) ? 0 : 1;
    else
        own = (ct == GEL_C_4) ? 2 : 3;
    const double S1 = nu[0] + nu[1] + nu[2] + nu[3];
    const double S2 = nu[0] * ((sq(u) + sq(v)) / 2.0 + (UH[0] / UM[0]) * (p / ro))
                    + nu[1] * ((sq(u) + sq(v)) / 2.0 + (UH[1] / UM[1]) * (p / ro))
                    + nu[2] * ((sq(u) + sq(v)) / 2.0 + (UH[2] / UM[2]) * (p / ro))
                    + nu[3] * ((sq(u) + sq(v)) / 2.0 + (UH[3] / UM[3]) * (p / ro));
    const double ch = prm.src_coef_h;
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const int kk = h == own ? 1 : 0;
        q[1 + h][0] = ch * (kk * S1 - nu[h]);
        q[1 + h][1] = ch * (kk * S1 * u - nu[h] * uH[h]);
        q[1 + h][2] = ch * (kk * S1 * v - nu[h] * vH[h]);
        q[1 + h][3] = ch * (kk * S2 - nu[h] * ((sq(uH[h]) + sq(vH[h])) / 2.0 + (UH[h] / UM[h]) * 2.0 * (pH[h] / roH[h])));
    }
}

// One thread per cell: the five fluids' sources of the step into q[fluid][4][nc], from
// the state the five launches of the step read (the reference takes them from
// par[now1] for every fluid).
template <bool TYPES>
__device__ __forceinline__ void mf_source_cell(const SweepArgs& a, size_t kc, double* __restrict__ out)
{
    const size_t nc = (size_t)a.ni * a.nj;
    if (kc >= nc) return;
    const int ct = TYPES ? (int)a.ctype[kc] : (int)GEL_C_2;
    double q[5][4];
    mf_sources_all(a.prm, (const double*)a.now + kc, nc, ct, q);
#pragma unroll
    for (int f = 0; f < 5; f++)
#pragma unroll
        for (int k = 0; k < 4; k++) out[((size_t)f * 4 + k) * nc + kc] = q[f][k];
}
template <bool TYPES>
__global__ void __launch_bounds__(256) mf_source_kernel(SweepArgs a, double* __restrict__ out)
{
    mf_source_cell<TYPES>(a, (size_t)blockIdx.x * blockDim.x + threadIdx.x, out);
}
// The same for every case of a batch: work-group g takes the 256 cells of tile
// tiles[g].y of case tiles[g].x; each case's sources go to its own args.mfq.
template <bool TYPES>
__global__ void __launch_bounds__(256) mf_source_batch_kernel(const SweepArgs* __restrict__ args,
                                                              const int2* __restrict__ tiles)
{
    const int2 t = tiles[blockIdx.x];
    const SweepArgs& a = args[t.x];
    mf_source_cell<TYPES>(a, (size_t)t.y * 256 + threadIdx.x, (double*)a.mfq);
}

// Update of a hydrogen fluid (Setka.cpp:9619-9753): the gas update with a mass source
// and the fluid's own floors.
__device__ __forceinline__ void h_update(const double* s, const double* Pot, double Volume, double Volume2, double vs1,
                                         double vs2, double T1, double gam, const double* q, double ro_floor,
                                         double p_floor, double* o)
{
    const double ro = s[0], p = s[1], u = s[2], v = s[3];
    double ro3 = ro * Volume / Volume2 - T1 * (Pot[0] / Volume2 - q[0]);
    if (ro3 <= 0.0) ro3 = ro_floor;
    const double u3 = (ro * u * Volume / Volume2 - T1 * (Pot[1] / Volume2 - q[1])) / ro3;
    const double v3 = (ro * v * Volume / Volume2 - T1 * ((Pot[2] - p * 0.5 * (vs1 + vs2)) / Volume2 - q[2])) / ro3;
    double p3 = (((p / (gam - 1) + ro * (u * u + v * v) * 0.5) * Volume / Volume2 - T1 * (Pot[3] / Volume2 - q[3])) -
                 0.5 * ro3 * (u3 * u3 + v3 * v3)) * (gam - 1);
    if (p3 <= 0) p3 = p_floor;
    o[0] = ro3; o[1] = p3; o[2] = u3; o[3] = v3;
}

// Flag or-ed into the FLUX template argument: the plasma + hydrogen sweep.
constexpr int kMfFlux = 32;

// -------------------------------------------------------------- inner region ---
// Setka::Go_stationary_5_komponent_inner_MK2 (Setka.cpp:6999-7212): the static solver of
// the inner region, the gas sweep with its own face values and update.
constexpr int kInnerFlux = 64;

// Both sides of an interior face: the cells' values rescaled to the face radius and
// rotated to its polar frame, density and pressure from the radial reconstruction
// where both cells are C_1 (Setka.cpp:7064-7091).
template <bool TYPES>
__device__ __forceinline__ void inner_interior_states(const SweepArgs& a, const CellT<GasPhys, double>& M,
                                                      const CellT<GasPhys, double>& N, const CellT<GasPhys, double>& D,
                                                      const CellT<GasPhys, double>& U, bool m_axis, bool n_axis,
                                                      const Face& f, double* L, double* R)
{
    const double gam = a.prm.gamma;
#pragma unroll
    for (int k = 0; k < 5; k++) { L[k] = M.s[k]; R[k] = N.s[k]; }
    rescale_c1<5>(L, sqrt(sq(M.cx) + sq(M.cy)), f.dis, gam, M, f.fx, f.fy, true);
    rescale_c1<5>(R, sqrt(sq(N.cx) + sq(N.cy)), f.dis, gam, N, f.fx, f.fy, true);
    if (TYPES && M.type == GEL_C_1 && N.type == GEL_C_1) {
        tvd_radial<GasPhys>(M, N, D, m_axis, f.fx, f.fy, gam, L[0], L[1]);
        tvd_radial<GasPhys>(N, M, U, n_axis, f.fx, f.fy, gam, R[0], R[1]);
    }
}

// Boundary faces (Setka.cpp:7092-7114): the inner sphere's ghost on both sides; on the
// axis the rescaled, rotated cell value against its mirror image; other faces the
// cell's value against the ghost of Gran::Get_par.
template <bool TYPES>
__device__ __forceinline__ void inner_boundary_states(const SweepArgs& a, int ftype, const CellT<GasPhys, double>& M,
                                                      const CellT<GasPhys, double>& D, const Face& f, double* L,
                                                      double* R)
{
    if (ftype == GEL_F_AXIS) {
#pragma unroll
        for (int k = 0; k < 5; k++) L[k] = M.s[k];
        rescale_c1<5>(L, sqrt(sq(M.cx) + sq(M.cy)), f.dis, a.prm.gamma, M, f.fx, f.fy, true);
#pragma unroll
        for (int k = 0; k < 5; k++) R[k] = (k == 3) ? -L[k] : L[k];
        return;
    }
    boundary_states<GasPhys, double, double, TYPES>(a, ftype, M, D, f, L, R);
    if (ftype == GEL_F_INNER_SPHERE) {
#pragma unroll
        for (int k = 0; k < 5; k++) L[k] = R[k];
    }
}

// Update (Setka.cpp:7168-7200): static, per r-weighted volume, the tracer losing a bare
// dt (Setka.cpp:7173).
__device__ __forceinline__ void inner_update(const double* s, const double* Pot, double Volume, double vs1, double vs2,
                                             double T1, double gam, double q1, double q2, double q3, double* o,
                                             int& n_rho, int& n_p)
{
    const double ro = s[0], p = s[1], u = s[2], v = s[3], Q = s[4];
    double ro3 = ro - T1 * (Pot[0] / Volume);
    const double Q33 = Q - (T1 / Volume) * Pot[4] - T1;
    if (ro3 <= 0) { ro3 = 0.00001; n_rho++; }
    const double u3 = (ro * u - T1 * (Pot[1] / Volume - q1)) / ro3;
    const double v3 = (ro * v - T1 * ((Pot[2] - p * 0.5 * (vs1 + vs2)) / Volume - q2)) / ro3;
    double p3 = (((p / (gam - 1) + ro * (u * u + v * v) * 0.5) - T1 * (Pot[3] / Volume - q3)) -
                 0.5 * ro3 * (u3 * u3 + v3 * v3)) * (gam - 1);
    if (p3 <= 0) { p3 = 0.000001; n_p++; }
    o[0] = ro3; o[1] = p3; o[2] = u3; o[3] = v3; o[4] = Q33;
}

// Flag or-ed into the FLUX template argument of the sweep on grids with fitted surfaces.
constexpr int kTrackFlux = 16;

// Node row r is one of the fitted surfaces (SweepArgs::srow, -1 where unused).
__device__ __forceinline__ bool surface_row(const SweepArgs& a, int r)
{
    return r == a.srow[0] || r == a.srow[1] || r == a.srow[2];
}

// ---------------------------------------------------------------- marching tile ---
//
// A workgroup of NW wavefronts owns a strip of 64 columns (60 output columns and
// 2 halo columns each side) and a chunk of rows [i0, i1).  It marches up the
// chunk NW rows at a time: in a step, wavefront w owns row i = base + w, with
// base = i0 - 1, i0 - 1 + NW, ...  (row i0 - 1 only contributes the radial face
// below the chunk).  A step has three phases separated by barriers:
//   load    each wavefront brings cell row i + 2 (state, centroid, type) and node
//           row i + 3 from HBM into LDS rings of NW + 3 rows;
//   faces   each wavefront solves the radial face above its row and the polar
//           face (j, j+1) of its row, from the LDS rings; the radial face goes to
//           a ring of NW + 1 LDS slots as flux x area products;
//   update  each wavefront sums the four faces of its cells (the polar face left
//           of a column comes from the lane below by a cross-lane shuffle, the
//           radial faces from the slots of its row and of the row below) and writes
//           the updated primitives.
// The update phase reads nothing from the rings, so the next step's loads may
// overwrite them as soon as a wavefront gets there.  Every face is solved once, at
// one call site per direction; boundary faces only choose their two states.
// Static FP32-bulk single-fluid sweeps copy the next step's cell rows by LDS DMA
// (march_tile); their radial-face ring then lives in an LDS object of its own
// (MarchFlux), so that the compiler sees that reading it does not touch the rows being
// filled.  The other sweeps keep it in MarchLds (a separate object costs the FP64
// kernels registers).
template <class TS, bool MOVING, bool TYPES, int FLUX>
constexpr bool kStateDmaT = sizeof(TS) == 4 && !MOVING && !TYPES && (FLUX & (kTrackFlux | kMfFlux | kInnerFlux)) == 0;

template <class Ph, class TS, bool MOVING, bool TYPES, int NW, bool FAPART = false>
struct MarchLds {
    static constexpr int NR = NW + 3;                 // rows held by the rings
    TS st[NR][Ph::NV][kWave];                         // primitives of cell rows
    double cx[NR][kWave], cy[NR][kWave];              // cell centroids
    int8_t ty[TYPES ? NR : 1][kWave];                 // cell types
    double rp[TYPES ? NR : 1][kWave];                 // (centroid radius)^(2 gamma)
    double ca[TYPES ? NR : 1][kWave];                 // cos, sin of the centroid's polar angle
    double sa[TYPES ? NR : 1][kWave];
    double nx[MOVING ? NR : 1][kWave + 1];            // node rows before the step (static
    double ny[MOVING ? NR : 1][kWave + 1];            // grids read the geometry table)
    double nx2[MOVING ? NR : 1][kWave + 1];           // node rows after the step
    double ny2[MOVING ? NR : 1][kWave + 1];
    double fps[FAPART ? 0 : NW + 1][Ph::NV][kWave];   // radial flux x area above a row (none when apart)
    double ftm[FAPART ? 0 : NW + 1][kWave];           // CFL time of the cell above that face
};

// The radial faces' flux x area products and CFL times, apart from the rings.
template <class Ph, int NW>
struct MarchFlux {
    double fps[NW + 1][Ph::NV][kWave];                // radial flux x area above a row
    double ftm[NW + 1][kWave];                        // CFL time of the cell above that face
};

struct MarchAcc {       // per-lane results folded into the case's dt buffer and counters
    double tmin;
    int n_rho, n_p, bad;
};

// The step's dt: the minimum of the read slot's entries (uniform loads).
__device__ __forceinline__ double dt_read(const SweepArgs& a)
{
    const GPtr<const unsigned long long> p = gp((const unsigned long long*)a.dtbuf) + (size_t)a.dt_rd * kDtSlot;
    unsigned long long m = p[0];
#pragma unroll
    for (int k = 1; k < kDtLanes; k++) {
        const unsigned long long v = p[k * kDtStride];
        m = v < m ? v : m;
    }
    return __longlong_as_double((long long)m);
}

// The slot the next step fills, reset to +inf by the first kDtLanes lanes of one
// work-group of the step.
__device__ __forceinline__ void dt_clear(const SweepArgs& a)
{
    if (threadIdx.x < kDtLanes)
        gp(a.dtbuf)[(size_t)a.dt_clr * kDtSlot + threadIdx.x * kDtStride] =
            (unsigned long long)__double_as_longlong(__builtin_huge_val());
}

// The slot entry a wavefront's minimum goes to: its index in the launch mod kDtLanes.
__device__ __forceinline__ int dt_entry()
{
    return (int)((blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) % kDtLanes);
}

// Static FP32-bulk sweeps with a source copy the per-cell source inputs of the rows
// they update from HBM into LDS while they solve the faces (gfx950 LDS DMA,
// global_load_lds_dword: no registers held, the latency hidden behind the polar
// solve).  Fields of a row, 64 lanes each: the hydrogen moments (analytic closure),
// the source table, the fallback flag (hybrid), as the source mode needs them.
template <class TS, bool MOVING, int SRC, int FLUX>
constexpr bool kSrcDma = sizeof(TS) == 4 && !MOVING && SRC != GEL_SRC_NONE &&
                         (FLUX & (kTrackFlux | kMfFlux | kInnerFlux)) == 0;
template <int SRC> constexpr int kSrcMom = (SRC == GEL_SRC_MK1 || SRC == GEL_SRC_HYBRID) ? GEL_MK1_NFIELDS : 0;
template <int SRC> constexpr int kSrcTab = (SRC == GEL_SRC_TABLE || SRC == GEL_SRC_HYBRID) ? 5 : 0;
template <int SRC> constexpr int kSrcFields = kSrcMom<SRC> + kSrcTab<SRC> + (SRC == GEL_SRC_HYBRID ? 1 : 0);
template <class TS, bool MOVING, int SRC, int FLUX, int NW>
constexpr int kSrcLds = kSrcDma<TS, MOVING, SRC, FLUX> ? NW * kSrcFields<SRC> * kWave : 1;

template <class Ph, class TS, class TF, bool MOVING, bool TYPES, int SRC, int FLUX, int NW, class LS, class FL>
__device__ __forceinline__ void march_tile(const SweepArgs& a, LS& S, FL& F, TS* srcl, int bx, int i0, int i1,
                                           MarchAcc& acc)
{
    // FLUX carries kTrackFlux on grids with fitted surfaces (gas, FP64, moving)
    constexpr bool TRACK = (FLUX & kTrackFlux) != 0;
    constexpr bool MF = (FLUX & kMfFlux) != 0;        // plasma + hydrogen, one fluid per launch
    constexpr bool INNER = (FLUX & kInnerFlux) != 0;  // the inner-region solver
    constexpr int FX = FLUX & 15;
    static_assert(!INNER || (Ph::NV == 5 && sizeof(TS) == 8 && sizeof(TF) == 8 && !MOVING && !TRACK && !MF),
                  "inner region: gas, FP64, static");
    static_assert(!TRACK || (Ph::NV == 5 && sizeof(TS) == 8 && sizeof(TF) == 8 && MOVING),
                  "fitted surfaces: gas, FP64, moving grids");
    static_assert(!MF || (Ph::NV == 5 && sizeof(TS) == 8 && sizeof(TF) == 8 && !TRACK),
                  "plasma + hydrogen: gas, FP64");
    constexpr int NV = Ph::NV;
    constexpr int NR = NW + 3;
    using C = CellT<Ph, TS>;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int ni = a.ni, nj = a.nj;
    const int j = bx * kOut + lane - kHalo;
    const bool col = (j >= 0 && j < nj);
    const bool out_lane = col && lane >= kHalo && lane < kWave - kHalo;
    const bool used_polar = col && lane >= 1 && lane < kWave - kHalo;
    const double gam = a.prm.gamma;
    const double cfl = a.prm.cfl;
    const double dt = dt_read(a);
    const bool axis_lo = (j == 0), axis_hi = (j == nj - 1);
    const bool m_axis = axis_lo || axis_hi;
    const bool n_axis = (j + 1 == nj - 1);
    const size_t nc = (size_t)ni * nj;

    if (bx == 0 && i0 == 0) dt_clear(a);

    auto slot = [&](int r) { return (r - i0 + 2) % NR; };     // rows and node rows >= i0 - 2
    auto fslot = [&](int r) { return (r - i0 + 1) % (NW + 1); };  // radial face above row r >= i0 - 1

    const GeoTable G(a.geo, ni, nj);

    // cell row r and node row r + 1 from HBM into the rings
    auto load_row = [&](int r) {
        C c;
        if constexpr (MOVING) {
            const Nodes lo = load_nodes<MOVING>(a, a.x, a.y, a.s_old, r, j, col);
            const Nodes hi = load_nodes<MOVING>(a, a.x, a.y, a.s_old, r + 1, j, col);
            c = load_cell<Ph, TS, TYPES, true, MF>(a, r, j, col, lo, hi);
        } else {
            const Nodes none{ 1.0, 1.0, 1.0, 1.0 };
            c = load_cell<Ph, TS, TYPES, false, MF>(a, r, j, col, none, none);
            if (c.ok) {
                const size_t k = (size_t)r * nj + j;
                c.cx = G.cell(0)[k];
                c.cy = G.cell(1)[k];
            }
        }
        const int s = slot(r);
#pragma unroll
        for (int v = 0; v < NV; v++) S.st[s][v][lane] = c.s[v];
        S.cx[s][lane] = c.cx;
        S.cy[s][lane] = c.cy;
        if constexpr (TYPES) {
            S.ty[s][lane] = (int8_t)c.type;
            S.rp[s][lane] = c.ok ? pow(sqrt(sq(c.cx) + sq(c.cy)), 2.0 * gam) : 0.0;
            double ca = 1.0, sa = 0.0;
            if (c.ok && (INNER || c.type == GEL_C_1)) {
                const double a1 = polar_angle(c.cx, c.cy);
                ca = cos(a1);
                sa = sin(a1);
            }
            S.ca[s][lane] = ca;
            S.sa[s][lane] = sa;
        }
        if constexpr (!MOVING) return;
        const int q = r + 1, sq1 = slot(q);
        auto put = [&](int idx, int jj) {
            const bool v = (q >= 0 && q <= ni && jj >= 0 && jj <= nj);
            const size_t k = (size_t)q * (nj + 1) + jj;
            double x = 1.0, y = 1.0;
            if (v) {
                x = a.x[k]; y = a.y[k];
                if (MOVING) { x *= a.s_old; y *= a.s_old; }
            }
            S.nx[sq1][idx] = x;
            S.ny[sq1][idx] = y;
            if constexpr (MOVING) {
                double x2 = 1.0, y2 = 1.0;
                if (v) { x2 = a.x2[k] * a.s_new; y2 = a.y2[k] * a.s_new; }
                S.nx2[sq1][idx] = x2;
                S.ny2[sq1][idx] = y2;
            }
        };
        put(lane, j);
        if (lane == kWave - 1) put(kWave, j + 1);
    };

    // Static FP32-bulk sweeps: a wavefront copies its next step's cell row from HBM
    // straight into the row's ring slot by LDS DMA at the start of the update phase
    // (which reads no ring; each wavefront waits for its copies before the next load
    // phase's barrier), so that the load phase waits on no HBM load.  Every lane copies (lanes off the grid
    // a cell of the row, unused); rows past the grid take load_row's defaults.
    constexpr bool kStateDma = kStateDmaT<TS, MOVING, TYPES, FLUX>;
    auto dma_row = [&](int r) {
        if (r >= ni) { load_row(r); return; }
        const int s = slot(r);
        const GPtr<const TS> st0 = gp((const TS*)a.now) + (size_t)r * nj + min(max(j, 0), nj - 1);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0x0F70);
#pragma unroll
        for (int v = 0; v < NV; v++)
            __builtin_amdgcn_global_load_lds((const void*)(st0 + (size_t)v * nc), (void*)&S.st[s][v][0], 4, 0, 0);
        // centroids (FP64): 64 x 8 B in two copies of 256 B, lane l taking half l % 2 of
        // cell 32 h + l / 2
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int jj = min(max(bx * kOut + 32 * h + (lane >> 1) - kHalo, 0), nj - 1);
            const size_t k = (size_t)r * nj + jj;
            const GPtr<const char> ox = (GPtr<const char>)(G.cell(0) + k) + 4 * (lane & 1);
            const GPtr<const char> oy = (GPtr<const char>)(G.cell(1) + k) + 4 * (lane & 1);
            __builtin_amdgcn_global_load_lds((const void*)ox, (void*)((char*)&S.cx[s][0] + 256 * h), 4, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)oy, (void*)((char*)&S.cy[s][0] + 256 * h), 4, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    };

    auto cell = [&](int r, int ln) -> C {
        ln = min(max(ln, 0), kWave - 1);
        const int s = slot(r);
        C c;
#pragma unroll
        for (int v = 0; v < NV; v++) c.s[v] = S.st[s][v][ln];
        c.cx = S.cx[s][ln];
        c.cy = S.cy[s][ln];
        c.rp = TYPES ? S.rp[s][ln] : 0.0;
        c.ca = TYPES ? S.ca[s][ln] : 1.0;
        c.sa = TYPES ? S.sa[s][ln] : 0.0;
        c.type = TYPES ? (int)S.ty[s][ln] : (int)GEL_C_2;
        const int jj = bx * kOut + ln - kHalo;
        c.ok = (jj >= 0 && jj < nj && r >= 0 && r < ni) ? 1 : 0;
        return c;
    };

    struct NodeP { double x, y, x2, y2; };
    auto node = [&](int q, int idx) -> NodeP {
        const int s = slot(q);
        NodeP n;
        n.x = S.nx[s][idx];
        n.y = S.ny[s][idx];
        if constexpr (MOVING) { n.x2 = S.nx2[s][idx]; n.y2 = S.ny2[s][idx]; }
        else { n.x2 = n.x; n.y2 = n.y; }
        return n;
    };
    // face geometry of a static grid from the table (defaults where the lane has none)
    auto table_face = [&](const GPtr<const double>* F, bool valid, size_t k) -> Face {
        Face f{ 1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 1.0 };
        if (valid) { f.n1 = F[0][k]; f.n2 = F[1][k]; f.S = F[2][k]; f.fx = F[3][k]; f.fy = F[4][k]; f.dis = F[5][k]; }
        return f;
    };

    // FP32 bulk: faces from their nodes (face_geom_a); static grids read the nodes from
    // HBM (the rows of two neighbouring wavefronts share one node row, an L2 hit for
    // the second), moving grids from the node rings
    constexpr bool F32 = sizeof(TF) == 4;
    static_assert(!F32 || !TYPES, "FP32 bulk: grids without cell types");
    const Face kNoFace{ 1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0 };
    const double idt = (F32 && MOVING) ? 1.0 / dt : 0.0;
    auto gnode = [&](int q, int c) -> NodeP {
        NodeP n{ 1.0, 1.0, 1.0, 1.0 };
        if (q >= 0 && q <= ni && c >= 0 && c <= nj) {
            const size_t k = (size_t)q * (nj + 1) + c;
            n.x = gp(a.x)[k];
            n.y = gp(a.y)[k];
        }
        n.x2 = n.x; n.y2 = n.y;
        return n;
    };

    // Static FP32-bulk sweeps with a source: a row's face nodes (r + 1, j), (r + 1, j + 1),
    // (r, j + 1) and its cells' area and volume are loaded a step ahead, during the
    // update phase before the step that needs them, so that their latency hides behind
    // the update instead of opening the faces phase.
    constexpr bool kGeoAhead = F32 && !MOVING && kSrcDma<TS, MOVING, SRC, FLUX>;
    NodeP gA{ 1.0, 1.0, 1.0, 1.0 }, gB{ 1.0, 1.0, 1.0, 1.0 }, gC{ 1.0, 1.0, 1.0, 1.0 };
    double gv1 = 1.0, gvol = 1.0;
    auto geo_ahead = [&](int r) {
        gA = gnode(r + 1, j);
        gB = gnode(r + 1, j + 1);
        gC = gnode(r, j + 1);
        gv1 = 1.0; gvol = 1.0;
        if (col && r >= 0 && r < ni) {
            const size_t k = (size_t)r * nj + j;
            gv1 = G.cell(2)[k];
            gvol = G.cell(3)[k];
        }
    };

    // initial rows i0 - 2 .. i0 (and node rows i0 - 1 .. i0 + 1)
    for (int r = i0 - 2 + w; r <= i0; r += NW) load_row(r);
    if constexpr (kGeoAhead) geo_ahead(i0 - 1 + w);

    for (int base = i0 - 1; base < i1; base += NW) {
        const int i = base + w;
        const bool act = (i < i1);
        const bool upd = act && (i >= i0);

        // ---- load (after the first step of a tile, a DMA path copied the rows already)
        if (act && (!kStateDma || base == i0 - 1)) load_row(i + 2);
        // a barrier does not wait for a wavefront's LDS DMA: each waits for its own
        // copies first (s_waitcnt vmcnt(0)), so that the rows are whole for every reader
        if constexpr (kStateDma) __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();

        // ---- faces
        C M;
        double psp[NV], ps0[NV];
        double t_dn = 0.0, t_pol = 0.0, t0 = 0.0;
        double vs1 = 1.0, vol = 1.0, vs2 = 1.0, vol2 = 1.0;
        if (act) {
            // static FP32 bulk: node (i + 1, j + 1), shared by the two faces
            const NodeP nB = kGeoAhead ? gB : (F32 && !MOVING) ? gnode(i + 1, j + 1) : NodeP{ 1.0, 1.0, 1.0, 1.0 };
            M = cell(i, lane);
            const C D = cell(i - 1, lane), N = cell(i + 1, lane), U = cell(i + 2, lane);
            // radial face above row i, nodes (i+1, j) -> (i+1, j+1); below row 0 it is the
            // inner sphere, oriented from row 0 inward: (0, j+1) -> (0, j)
            {
                const bool inner = (i < 0);
                Face f;
                if constexpr (F32) {
                    NodeP na = MOVING ? node(i + 1, lane) : kGeoAhead ? gA : gnode(i + 1, j);
                    NodeP nb = MOVING ? node(i + 1, lane + 1) : nB;
                    if (inner) { const NodeP t = na; na = nb; nb = t; }
                    f = face_geom_a<MOVING>(na.x, na.y, nb.x, nb.y, na.x2, na.y2, nb.x2, nb.y2, idt);
                    if (!(col && i + 1 <= ni)) f = kNoFace;
                    if (inner) f.dis = sqrt(sq(f.fx) + sq(f.fy));
                } else if constexpr (MOVING) {
                    NodeP na = node(i + 1, lane), nb = node(i + 1, lane + 1);
                    if (inner) { const NodeP t = na; na = nb; nb = t; }
                    f = face_geom<MOVING, TF>(na.x, na.y, nb.x, nb.y, na.x2, na.y2, nb.x2, nb.y2, dt);
                } else {
                    const GPtr<const double> F[6] = { G.rad(0), G.rad(1), G.rad(2), G.rad(3), G.rad(4), G.rad(5) };
                    f = table_face(F, col && i + 1 <= ni, (size_t)(i + 1) * nj + j);
                }
                double L[NV], R[NV], P[NV];
                int b = 0;
                // a face of a fitted surface: both nodes on surface row i + 1
                const bool god = TRACK && !inner && i + 1 < ni && surface_row(a, i + 1);
                if (inner) {
                    if constexpr (MF) mf_boundary_states<TYPES>(a, GEL_F_INNER_SPHERE, N, U, f, L, R);
                    else if constexpr (INNER) inner_boundary_states<TYPES>(a, GEL_F_INNER_SPHERE, N, U, f, L, R);
                    else boundary_states<Ph, TS, TF, TYPES>(a, GEL_F_INNER_SPHERE, N, U, f, L, R);
                } else if (i + 1 >= ni) {
                    const int ot = col ? (int)gp(a.otype)[j] : (int)GEL_F_EXTERN;
                    if constexpr (MF) mf_boundary_states<TYPES>(a, ot, M, D, f, L, R);
                    else if constexpr (INNER) inner_boundary_states<TYPES>(a, ot, M, D, f, L, R);
                    else boundary_states<Ph, TS, TF, TYPES>(a, ot, M, D, f, L, R);
                } else if (god) {
                    if constexpr (TRACK) surface_states<Ph, TYPES>(M, N, D, U, m_axis, f, gam, L, R, b);
                } else {
                    if constexpr (MF) mf_interior_states<TYPES>(a, M, N, D, U, m_axis, m_axis, f, L, R, b);
                    else if constexpr (INNER) inner_interior_states<TYPES>(a, M, N, D, U, m_axis, m_axis, f, L, R);
                    else interior_states<Ph, TS, TF, TYPES>(M, N, D, U, m_axis, m_axis, f, gam, L, R, b);
                }
                TF uu;
                if constexpr (TRACK) {
                    uu = god ? surface_flux<FX == GEL_FLUX_HLLC>(L, R, f, gam, P, b) : face_flux<Ph, TF, FX>(L, R, f, gam, P);
                } else {
                    uu = face_flux<Ph, TF, FX>(L, R, f, gam, P);
                }
                const int fs = fslot(i);
#pragma unroll
                for (int k = 0; k < NV; k++) F.fps[fs][k][lane] = (inner ? -P[k] : P[k]) * (F32 ? 1.0 : f.S);
                F.ftm[fs][lane] = cfl_time<TF>(cfl, N.cx, N.cy, f, uu);
                t_dn = cfl_time<TF>(cfl, M.cx, M.cy, f, uu);
                if (out_lane) acc.bad += b;
            }
            if (upd) {
                // polar face (j, j+1), nodes (i+1, j+1) -> (i, j+1); the lane left of
                // column 0 solves the axis face of column 0, oriented (i, 0) -> (i+1, 0)
                const bool axl = (j == -1);
                Face f;
                if constexpr (F32) {
                    NodeP na = MOVING ? node(i + 1, lane + 1) : nB;
                    NodeP nb = MOVING ? node(i, lane + 1) : kGeoAhead ? gC : gnode(i, j + 1);
                    if (axl) { const NodeP t = na; na = nb; nb = t; }
                    f = face_geom_a<MOVING>(na.x, na.y, nb.x, nb.y, na.x2, na.y2, nb.x2, nb.y2, idt);
                    if (!(j >= -1 && j < nj)) f = kNoFace;
                } else if constexpr (MOVING) {
                    NodeP na = node(i + 1, lane + 1), nb = node(i, lane + 1);
                    if (axl) { const NodeP t = na; na = nb; nb = t; }
                    f = face_geom<MOVING, TF>(na.x, na.y, nb.x, nb.y, na.x2, na.y2, nb.x2, nb.y2, dt);
                } else {
                    const GPtr<const double> F[6] = { G.pol(0), G.pol(1), G.pol(2), G.pol(3), G.pol(4), G.pol(5) };
                    f = table_face(F, j >= -1 && j < nj, (size_t)i * (nj + 1) + (j + 1));
                }
                if constexpr (!MOVING) {
                    if constexpr (kGeoAhead) {
                        vs1 = gv1; vol = gvol;
                        vs2 = vs1; vol2 = vol;
                    } else if constexpr (kSrcDma<TS, MOVING, SRC, FLUX>) {
                        // the cell's geometry, loaded before the source inputs' LDS DMA
                        // below waits for every load issued so far
                        if (col) {
                            const size_t k = (size_t)i * nj + j;
                            vs1 = G.cell(2)[k];
                            vol = G.cell(3)[k];
                        }
                        vs2 = vs1; vol2 = vol;
                    }
                }
                const C Mr = cell(i, lane + 1), Ml = cell(i, lane - 1), Mr2 = cell(i, lane + 2);
                double L[NV], R[NV], P[NV];
                int b = 0;
                if constexpr (MF) {
                    if (axl) mf_boundary_states<TYPES>(a, GEL_F_AXIS, Mr, Mr2, f, L, R);
                    else if (axis_hi) mf_boundary_states<TYPES>(a, GEL_F_AXIS, M, Ml, f, L, R);
                    else mf_interior_states<TYPES>(a, M, Mr, Ml, Mr2, m_axis, n_axis, f, L, R, b);
                } else if constexpr (INNER) {
                    if (axl) inner_boundary_states<TYPES>(a, GEL_F_AXIS, Mr, Mr2, f, L, R);
                    else if (axis_hi) inner_boundary_states<TYPES>(a, GEL_F_AXIS, M, Ml, f, L, R);
                    else inner_interior_states<TYPES>(a, M, Mr, Ml, Mr2, m_axis, n_axis, f, L, R);
                } else {
                    if (axl) boundary_states<Ph, TS, TF, TYPES>(a, GEL_F_AXIS, Mr, Mr2, f, L, R);
                    else if (axis_hi) boundary_states<Ph, TS, TF, TYPES>(a, GEL_F_AXIS, M, Ml, f, L, R);
                    else interior_states<Ph, TS, TF, TYPES>(M, Mr, Ml, Mr2, m_axis, n_axis, f, gam, L, R, b);
                }
                if constexpr (kSrcDma<TS, MOVING, SRC, FLUX>) {
                    // the source inputs of row i into srcl[w] (field f, lane l -> column
                    // l), after the last global load this phase waits for (memory
                    // counters retire in order), so that they arrive during the polar
                    // solve; an explicit wait before the barrier ending the phase
                    // completes them.  Every lane
                    // copies (lanes off the grid a cell of the row, unused), so that no
                    // branch leaves the memory counters path-dependent.  The compiler
                    // takes LDS DMA and loads into registers as retiring out of order and
                    // would wait for the copies at the next use of a loaded register: the
                    // loads issued so far are waited for first (s_waitcnt vmcnt(0)).
                    __builtin_amdgcn_sched_barrier(0);
                    __builtin_amdgcn_s_waitcnt(0x0F70);
                    const size_t kc0 = (size_t)i * nj + min(max(j, 0), nj - 1);
                    TS* dst = srcl + w * kSrcFields<SRC> * kWave;
                    auto copy = [&](const void* p) {
                        __builtin_amdgcn_global_load_lds(p, (void*)dst, 4, 0, 0);
                        dst += kWave;
                    };
                    if constexpr (kSrcMom<SRC> > 0) {
                        const GPtr<const TS> m = gp((const TS*)a.src) + kc0;
#pragma unroll
                        for (int k = 0; k < kSrcMom<SRC>; k++) copy((const void*)(m + (size_t)k * nc));
                    }
                    if constexpr (kSrcTab<SRC> > 0) {
                        const GPtr<const TS> t = gp((const TS*)a.qtab) + kc0;
#pragma unroll
                        for (int k = 0; k < kSrcTab<SRC>; k++) copy((const void*)(t + (size_t)k * nc));
                    }
                    if constexpr (SRC == GEL_SRC_HYBRID) copy((const void*)(gp(a.qflag) + kc0));
                    __builtin_amdgcn_sched_barrier(0);
                }
                TF uu;
                if constexpr (TRACK && FX == GEL_FLUX_HLLC) {
                    // one node of the face on a surface row: nul_potok (Setka.cpp:10299-10327)
                    if (surface_row(a, i) || surface_row(a, i + 1)) {
                        uu = nul_potok_flux(L, R, f, gam, P);
                    } else {
                        uu = face_flux<Ph, TF, FX>(L, R, f, gam, P);
                    }
                } else if constexpr (INNER) {
                    // the axis takes the HLL flux (met = 0, Setka.cpp:7108)
                    uu = (axl || axis_hi) ? face_flux<Ph, TF, GEL_FLUX_HLL>(L, R, f, gam, P) : face_flux<Ph, TF, FX>(L, R, f, gam, P);
                } else {
                    uu = face_flux<Ph, TF, FX>(L, R, f, gam, P);
                }
#pragma unroll
                for (int k = 0; k < NV; k++) psp[k] = (axl ? -P[k] : P[k]) * (F32 ? 1.0 : f.S);
                t_pol = cfl_time<TF>(cfl, M.cx, M.cy, f, uu);
                const double t_right = cfl_time<TF>(cfl, Mr.cx, Mr.cy, f, uu);
                if (used_polar) acc.bad += b;
                // the face left of each column is the lane below's face, seen from the other side
#pragma unroll
                for (int k = 0; k < NV; k++) ps0[k] = -shfl_up(psp[k], 1);
                t0 = shfl_up(t_right, 1);
                // cell geometry before and after the step (static grids with the source
                // DMA: read above)
                if constexpr (!MOVING && !kSrcDma<TS, MOVING, SRC, FLUX>) {
                    if (col) {
                        const size_t k = (size_t)i * nj + j;
                        vs1 = G.cell(2)[k];
                        vol = G.cell(3)[k];
                    }
                    vs2 = vs1; vol2 = vol;
                } else if constexpr (MOVING) {
                const NodeP c00 = node(i, lane), c10 = node(i + 1, lane), c11 = node(i + 1, lane + 1),
                            c01 = node(i, lane + 1);
                const double px[4] = { c00.x, c10.x, c11.x, c01.x };
                const double py[4] = { c00.y, c10.y, c11.y, c01.y };
                vs1 = quad_area(px, py);
                vol = quad_volume_r(px, py);
                vs2 = vs1; vol2 = vol;
                    const double px2[4] = { c00.x2, c10.x2, c11.x2, c01.x2 };
                    const double py2[4] = { c00.y2, c10.y2, c11.y2, c01.y2 };
                    vs2 = quad_area(px2, py2);
                    vol2 = quad_volume_r(px2, py2);
                }
            }
        }
        // the source inputs' LDS DMA of this phase: waited for before the barrier
        if constexpr (kSrcDma<TS, MOVING, SRC, FLUX>) __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();

        // ---- update (Setka.cpp:10372-10468)
        if constexpr (kStateDma)
            if (i + NW < i1) dma_row(i + NW + 2);
        if constexpr (kGeoAhead) geo_ahead(i + NW);
        if (upd) {
            const int fa = fslot(i), fb = fslot(i - 1);
            const double x = M.cx, y = M.cy;
            const double radius = sqn(sq(x) + sq(y));
            if (out_lane && (INNER || !(radius < a.prm.r_skip))) {
                // faces in the reference's order: edge 0 (polar, left), 1 (radial, above),
                // 2 (polar, right), 3 (radial, below)
                double Pot[NV];
#pragma unroll
                for (int k = 0; k < NV; k++) {
                    double s = 0.0;
                    s = s + ps0[k];
                    s = s + F.fps[fa][k][lane];
                    s = s + psp[k];
                    s = s + (-F.fps[fb][k][lane]);
                    Pot[k] = s;
                }
                double time = 10000000;
                time = mn(time, t0);
                time = mn(time, t_dn);
                time = mn(time, t_pol);
                time = mn(time, F.ftm[fb][lane]);
                const size_t kc = (size_t)i * nj + j;
                const GPtr<TS> nxt = gp((TS*)a.next);
                // the cell's type for the update rules, read here from the grid's type
                // array: the update depends on no value carried over from the faces phase
                // but the flux sums and the cell's own state
                const int ctu = TYPES ? (int)gp(a.ctype)[kc] : (int)GEL_C_2;
                if constexpr (INNER) {
                    // C_centr cells are skipped: no time step, no update (Setka.cpp:7025-7028)
                    const int ct = ctu;
                    double o[5];
                    if (ct == GEL_C_CENTR) {
#pragma unroll
                        for (int k = 0; k < 5; k++) o[k] = M.s[k];
                    } else {
                        acc.tmin = fmin(acc.tmin, time);
                        double q1s = 0.0, q2s = 0.0, q3s = 0.0;
                        if (SRC == GEL_SRC_MK1) {
                            const double* __restrict__ sp = (const double*)a.src + kc;
                            auto fk = [&](int k) { return sp[(size_t)k * nc]; };
                            source_mk1<decltype(fk), double>(a.prm, fk, M.s[2], M.s[3], M.s[0], M.s[1], q1s, q2s, q3s);
                        }
                        double sd[5];
#pragma unroll
                        for (int k = 0; k < 5; k++) sd[k] = M.s[k];
                        inner_update(sd, Pot, vol, vs1, vs2, dt, gam, q1s, q2s, q3s, o, acc.n_rho, acc.n_p);
                    }
#pragma unroll
                    for (int k = 0; k < 5; k++) nxt[(size_t)k * nc + kc] = TS(o[k]);
                } else if constexpr (MF) {
                    // C_centr cells solve no H1 faces (Setka.cpp:9172) and keep the plasma
                    // and H1 (:9525-9531, :9645-9651); H2..H4 update everywhere
                    const int ct = ctu;
                    if (!(a.fluid == 1 && ct == GEL_C_CENTR)) acc.tmin = fmin(acc.tmin, time);
                    const int nw = a.fluid == 0 ? 5 : 4;
                    double o[5];
                    if (ct == GEL_C_CENTR && a.fluid <= 1) {
#pragma unroll
                        for (int k = 0; k < 5; k++) o[k] = M.s[k];
                    } else {
                        // this fluid's sources of the step (mf_source_kernel)
                        double q[4];
                        const GPtr<const double> qs = gp(a.mfq) + (size_t)a.fluid * 4 * nc + kc;
#pragma unroll
                        for (int k = 0; k < 4; k++) q[k] = qs[(size_t)k * nc];
                        double sd[5];
#pragma unroll
                        for (int k = 0; k < 5; k++) sd[k] = M.s[k];
                        if (a.fluid == 0) {
                            cell_update<Ph, false>(sd, Pot, vol, vol2, vs1, vs2, dt, gam, 0.0, q[1], q[2], q[3], 0.0, false,
                                                   o, acc.n_rho, acc.n_p);
                        } else {
                            // floors of Setka.cpp:9627, :9642, :9672, :9687, :9701, :9716, :9733, :9748
                            const double rf = a.fluid == 1 ? 0.0000001 : 0.000001;
                            const double pf = a.fluid == 4 ? 0.0000001 : 0.000001;
                            h_update(sd, Pot, vol, vol2, vs1, vs2, dt, gam, q, rf, pf, o);
                        }
                    }
                    for (int k = 0; k < nw; k++) nxt[(size_t)a.voff[k] * nc + kc] = TS(o[k]);
                } else {
                acc.tmin = fmin(acc.tmin, time);
                if (TYPES && ctu == GEL_C_CENTR) {
#pragma unroll
                    for (int k = 0; k < NV; k++) nxt[(size_t)k * nc + kc] = M.s[k];
                } else {
                    double sd[NV];
#pragma unroll
                    for (int k = 0; k < NV; k++) sd[k] = double(M.s[k]);
                    double qms = 0.0, q1s = 0.0, q2s = 0.0, q3s = 0.0, qQs = 0.0;
                    constexpr bool DMA = kSrcDma<TS, MOVING, SRC, FLUX>;
                    const TS* sl = srcl + w * kSrcFields<SRC> * kWave + lane;   // this cell's copied inputs
                    bool flag = false;
                    if constexpr (SRC == GEL_SRC_HYBRID) {
                        if constexpr (DMA) flag = __builtin_bit_cast(int, sl[(kSrcFields<SRC> - 1) * kWave]) != 0;
                        else flag = gp(a.qflag)[kc] != 0;
                    }
                    const bool analytic = (SRC == GEL_SRC_MK1) || (SRC == GEL_SRC_HYBRID && flag);
                    if (analytic) {
                        // moments stored in the state's type, the closure in the face type
                        const GPtr<const TS> sp = gp((const TS*)a.src) + kc;
                        auto fk = [&](int k) {
                            if constexpr (DMA) return sl[k * kWave];
                            else return sp[(size_t)k * nc];
                        };
                        TF r1, r2, r3;
                        source_mk1<decltype(fk), TF>(a.prm, fk, TF(M.s[2]), TF(M.s[3]), TF(M.s[0]), TF(M.s[1]), r1, r2, r3);
                        q1s = r1; q2s = r2; q3s = r3;
                    } else if (SRC == GEL_SRC_TABLE || SRC == GEL_SRC_HYBRID) {
                        const GPtr<const TS> qt = gp((const TS*)a.qtab);
                        auto tk = [&](int k) {
                            if constexpr (DMA) return sl[(kSrcMom<SRC> + k) * kWave];
                            else return qt[(size_t)k * nc + kc];
                        };
                        qms = tk(0);
                        q1s = tk(1);
                        q2s = tk(2);
                        q3s = tk(3);
                        qQs = tk(4);
                    }
                    double o[NV];
                    cell_update<Ph, sizeof(TF) == 4>(sd, Pot, vol, vol2, vs1, vs2, dt, gam, qms, q1s, q2s, q3s, qQs,
                                    (SRC == GEL_SRC_TABLE || SRC == GEL_SRC_HYBRID) && !analytic, o, acc.n_rho,
                                    acc.n_p);
#pragma unroll
                    for (int k = 0; k < NV; k++) nxt[(size_t)k * nc + kc] = TS(o[k]);
                }
                }
            }
        }
    }
}

// Fold a wavefront's results into the case's dt buffer and counters.
__device__ __forceinline__ void march_flush(const SweepArgs& a, MarchAcc& acc)
{
    const double tmin = wave_min(acc.tmin);
    const int s_rho = wave_sum(acc.n_rho), s_p = wave_sum(acc.n_p), s_bad = wave_sum(acc.bad);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (tmin < __builtin_huge_val())
            __hip_atomic_fetch_min(gp(a.dtbuf) + (size_t)a.dt_wr * kDtSlot + dt_entry() * kDtStride,
                                   (unsigned long long)__double_as_longlong(tmin), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        const GPtr<unsigned long long> cnt = gp(a.counters);
        if (s_rho) __hip_atomic_fetch_add(cnt, (unsigned long long)s_rho, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s_p) __hip_atomic_fetch_add(cnt + 1, (unsigned long long)s_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s_bad) __hip_atomic_fetch_add(cnt + 2, (unsigned long long)s_bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    acc = MarchAcc{ __builtin_huge_val(), 0, 0, 0 };
}

// Work-group index -> work item, so that the work-groups one XCD runs at a time
// (dispatch deals work-groups to the 8 XCDs round-robin) take neighbouring tiles.
__device__ __forceinline__ int xcd_item(int g, int ng)
{
    constexpr int kXcd = 8;
    if (ng % kXcd != 0) return g;
    return (g % kXcd) * (ng / kXcd) + g / kXcd;
}

// One step of one case: persistent work-groups over the (strip, chunk) tiles.
// Two work-groups of four wavefronts per CU (the LDS rings of the FP64 MHD kernel
// take 78 KB of the 160 KB) need at most 256 registers per lane.
template <class Ph, class TS, class TF, bool MOVING, bool TYPES, int SRC, int FLUX, int NW>
__global__ void __launch_bounds__(NW * kWave) __attribute__((amdgpu_waves_per_eu(2)))
sweep_kernel(SweepArgs a, int strips, int ntiles)
{
    constexpr bool apart = kStateDmaT<TS, MOVING, TYPES, FLUX>;
    __shared__ MarchLds<Ph, TS, MOVING, TYPES, NW, apart> S;
    __shared__ MarchFlux<Ph, apart ? NW : 0> F;
    __shared__ TS srcl[kSrcLds<TS, MOVING, SRC, FLUX, NW>];
    MarchAcc acc{ __builtin_huge_val(), 0, 0, 0 };
    for (int t = xcd_item(blockIdx.x, gridDim.x); t < ntiles; t += gridDim.x) {
        const int bx = t % strips, by = t / strips;
        const int i0 = by * a.rows, i1 = min(i0 + a.rows, a.ni);
        // no barrier between tiles: the update phase of the previous tile reads no ring,
        // and its faces phase ended at a barrier every wavefront has passed
        if constexpr (apart) march_tile<Ph, TS, TF, MOVING, TYPES, SRC, FLUX, NW>(a, S, F, srcl, bx, i0, i1, acc);
        else march_tile<Ph, TS, TF, MOVING, TYPES, SRC, FLUX, NW>(a, S, S, srcl, bx, i0, i1, acc);
    }
    march_flush(a, acc);
}

// One step of many independent cases in one persistent launch: the work-groups walk
// the (case, strip, chunk) list round-robin; args[case] carries each case's grid,
// buffers, time-step slots and parameters.
struct BatchItem {
    int cs, bx, by;
};

// MOVING: every case of the batch moves its nodes by its prescribed breathing (the
// step's scale factors in its argument set).  TYPES: every case has C_1 / C_centr cells.
template <class Ph, class TS, class TF, int SRC, int FLUX, int NW, bool MOVING = false, bool TYPES = false>
__global__ void __launch_bounds__(NW * kWave) __attribute__((amdgpu_waves_per_eu(2)))
sweep_batch_kernel(const SweepArgs* __restrict__ args, const BatchItem* __restrict__ items, int nitems)
{
    constexpr bool apart = kStateDmaT<TS, MOVING, TYPES, FLUX>;
    __shared__ MarchLds<Ph, TS, MOVING, TYPES, NW, apart> S;
    __shared__ MarchFlux<Ph, apart ? NW : 0> F;
    __shared__ TS srcl[kSrcLds<TS, MOVING, SRC, FLUX, NW>];
    MarchAcc acc{ __builtin_huge_val(), 0, 0, 0 };
    for (int it = xcd_item(blockIdx.x, gridDim.x); it < nitems; it += gridDim.x) {
        const BatchItem b = items[it];
        const SweepArgs& a = args[b.cs];
        const int i0 = b.by * a.rows, i1 = min(i0 + a.rows, a.ni);
        if constexpr (apart) march_tile<Ph, TS, TF, MOVING, TYPES, SRC, FLUX, NW>(a, S, F, srcl, b.bx, i0, i1, acc);
        else march_tile<Ph, TS, TF, MOVING, TYPES, SRC, FLUX, NW>(a, S, S, srcl, b.bx, i0, i1, acc);
        march_flush(a, acc);
    }
}


// ------------------------------------------------------------------ launches ---

// Resident work-groups of a kernel on the current device (occupancy x CUs).
inline int resident_groups(const void* kern, int block)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, block, 0) != hipSuccess || per < 1 || cus < 1)
        return 256;
    return per * cus;
}

// Rows per chunk for one case run by `groups` resident work-groups: a work-group
// takes ceil((rows + 1) / kNW) marching steps per tile (the extra row is the face
// below the chunk), and the tiles run in ceil(tiles / groups) rounds; pick the rows
// with the fewest steps on the critical path, the longest chunks among equals.
inline int choose_rows(int ni, int nj, int groups)
{
    const int strips = (nj + kOut - 1) / kOut;
    long best = -1;
    int br = kNW - 1;
    for (int k = 1; k <= 64; k++) {
        const int rows = k * kNW - 1;
        const long tiles = (long)strips * ((ni + rows - 1) / rows);
        const long rounds = (tiles + groups - 1) / groups;
        const long cost = rounds * k;
        if (best < 0 || cost <= best) { best = cost; br = rows; }
        if (rows >= ni) break;
    }
    return br;
}

template <class K>
hipError_t launch_k(K kern, gel_ctx* c, SweepArgs& a, hipStream_t s)
{
    const int groups = resident_groups((const void*)kern, kNW * kWave);
    if (c->rows <= 0) c->rows = choose_rows(c->ni, c->nj, groups);
    a.rows = c->rows;
    const int strips = (a.nj + kOut - 1) / kOut;
    const int ntiles = strips * ((a.ni + a.rows - 1) / a.rows);
    const int grid = ntiles < groups ? ntiles : groups;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kNW * kWave), 0, s, a, strips, ntiles);
    return hipGetLastError();
}

template <class Ph, class TS, class TF, bool MOVING, bool TYPES, int SRC, int FLUX>
hipError_t launch_t(gel_ctx* c, SweepArgs& a, hipStream_t s)
{
    return launch_k(sweep_kernel<Ph, TS, TF, MOVING, TYPES, SRC, FLUX, kNW>, c, a, s);
}

template <class Ph, class TS, class TF, bool TYPES, int FLUX>
hipError_t launch_ms(gel_ctx* c, SweepArgs& a, hipStream_t s, bool moving, int src)
{
#define GEL_L(MV, SR) launch_t<Ph, TS, TF, MV, TYPES, SR, FLUX>(c, a, s)
    if (moving) {
        switch (src) {
        case GEL_SRC_MK1: return GEL_L(true, GEL_SRC_MK1);
        case GEL_SRC_TABLE: return GEL_L(true, GEL_SRC_TABLE);
        case GEL_SRC_HYBRID: return GEL_L(true, GEL_SRC_HYBRID);
        default: return GEL_L(true, GEL_SRC_NONE);
        }
    }
    switch (src) {
    case GEL_SRC_MK1: return GEL_L(false, GEL_SRC_MK1);
    case GEL_SRC_TABLE: return GEL_L(false, GEL_SRC_TABLE);
    case GEL_SRC_HYBRID: return GEL_L(false, GEL_SRC_HYBRID);
    default: return GEL_L(false, GEL_SRC_NONE);
    }
#undef GEL_L
}

// One step of a batch: the persistent launch over the (case, strip, chunk) items;
// workgroups = 0: the kernel's resident work-groups.
template <class Ph, class TS, class TF, int SRC, int FLUX, bool MOVING = false, bool TYPES = false>
hipError_t batch_launch_t(const SweepArgs* d_args, const BatchItem* d_items, int nitems, int workgroups,
                          hipStream_t s)
{
    const auto kern = sweep_batch_kernel<Ph, TS, TF, SRC, FLUX, kNW, MOVING, TYPES>;
    int grid = workgroups > 0 ? workgroups : resident_groups((const void*)kern, kNW * kWave);
    if (grid > nitems) grid = nitems;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kNW * kWave), 0, s, d_args, d_items, nitems);
    return hipGetLastError();
}

// moving batches (prescribed breathing) take no source or the analytic closure
template <class Ph, class TS, class TF, int FLUX>
hipError_t batch_launch_s(const SweepArgs* d_args, const BatchItem* d_items, int nitems, int workgroups,
                          hipStream_t s, int src, bool moving)
{
    if (moving) {
        if (src == GEL_SRC_MK1)
            return batch_launch_t<Ph, TS, TF, GEL_SRC_MK1, FLUX, true>(d_args, d_items, nitems, workgroups, s);
        if (src == GEL_SRC_NONE)
            return batch_launch_t<Ph, TS, TF, GEL_SRC_NONE, FLUX, true>(d_args, d_items, nitems, workgroups, s);
        return hipErrorInvalidValue;
    }
    if (src == GEL_SRC_MK1) return batch_launch_t<Ph, TS, TF, GEL_SRC_MK1, FLUX>(d_args, d_items, nitems, workgroups, s);
    if (src == GEL_SRC_TABLE) return batch_launch_t<Ph, TS, TF, GEL_SRC_TABLE, FLUX>(d_args, d_items, nitems, workgroups, s);
    if (src == GEL_SRC_HYBRID) return batch_launch_t<Ph, TS, TF, GEL_SRC_HYBRID, FLUX>(d_args, d_items, nitems, workgroups, s);
    return batch_launch_t<Ph, TS, TF, GEL_SRC_NONE, FLUX>(d_args, d_items, nitems, workgroups, s);
}

// Batches of grids with C_1 / C_centr cells (gas, FP64): no source or the analytic
// closure, static or moving.
template <int FLUX>
hipError_t batch_launch_typed(const SweepArgs* d_args, const BatchItem* d_items, int nitems, int workgroups,
                              hipStream_t s, int src, bool moving)
{
    using G = GasPhys;
    if (src != GEL_SRC_NONE && src != GEL_SRC_MK1) return hipErrorInvalidValue;
    if (moving)
        return src == GEL_SRC_MK1
                   ? batch_launch_t<G, double, double, GEL_SRC_MK1, FLUX, true, true>(d_args, d_items, nitems, workgroups, s)
                   : batch_launch_t<G, double, double, GEL_SRC_NONE, FLUX, true, true>(d_args, d_items, nitems, workgroups, s);
    return src == GEL_SRC_MK1
               ? batch_launch_t<G, double, double, GEL_SRC_MK1, FLUX, false, true>(d_args, d_items, nitems, workgroups, s)
               : batch_launch_t<G, double, double, GEL_SRC_NONE, FLUX, false, true>(d_args, d_items, nitems, workgroups, s);
}

// The FP32-bulk launches live in sweep_fp32.hip (FP contraction on there).
hipError_t launch_sweep_fp32(gel_ctx* c, SweepArgs& a, bool moving);
hipError_t launch_batch_fp32(const gel_params& p, const SweepArgs* d_args, const BatchItem* d_items, int nitems,
                             int workgroups, hipStream_t s, bool moving);
int resident_batch_fp32(const gel_params& p, bool moving);

}  // namespace gel

