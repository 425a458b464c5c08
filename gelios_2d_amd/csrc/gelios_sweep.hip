// gelios_sweep.hip - CDNA4 (gfx950) Godunov flux-and-update sweep on polar grids,
// and the C-ABI of include/gelios.h.
//
// Kernel layout.  One wavefront (64 lanes) owns a strip of 60 neighbouring rays
// (columns j) plus 2 halo columns on each side, and marches along the rays over a
// chunk of rows i.  Per row it
//   - solves the radial face (i, i+1) once, from registers: the four-cell TVD
//     stencil (i-1, i, i+1, i+2) of a column lives in the same lane;
//   - solves the polar face (j, j+1) once, taking the stencil (j-1 .. j+2) from the
//     neighbouring lanes with cross-lane shuffles;
//   - adds the four face fluxes of each cell in the reference's face order and
//     writes the updated primitives (Setka.cpp:10372-10468).
// Every state value is read from HBM once per row chunk plus a 2-row halo; the
// global CFL step is a wave min-reduction and one integer atomicMin on the bit
// pattern of a positive double per wave (deterministic).
//
// Parity with the reference: the per-face arithmetic is that of sweep_core.h
// (reference operation order, FP contraction off); solving a face once instead of
// once from each side, as the reference's cell gather does (Setka.cpp:10195),
// changes the flux on the second side only by rounding.
#include "sweep_core.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace gel {

constexpr int kWave = 64;
constexpr int kHalo = 2;
constexpr int kOut = kWave - 2 * kHalo;   // output columns per wavefront

struct SweepArgs {
    int ni, nj;
    int rows;                     // rows per wavefront chunk
    int dt_rd, dt_wr, dt_clr;     // slots of the rotating time-step buffer
    const double* __restrict__ x;
    const double* __restrict__ y;
    const double* __restrict__ x2;
    const double* __restrict__ y2;
    const double* __restrict__ now;
    double* __restrict__ next;
    const int8_t* __restrict__ ctype;
    const int8_t* __restrict__ otype;
    const double* __restrict__ src;
    unsigned long long* dtbuf;    // [3] bit patterns of positive doubles
    unsigned long long* counters; // [3] rho clamps, p clamps, numeric failures
    gel_params prm;
};

struct Cell {          // one cell of the stencil as seen by a lane
    Gas<double> s;
    double cx, cy;
    int type;
    int ok;            // exists (inside the grid)
};

struct Nodes {         // nodes (r, j) and (r, j+1) of a lane
    double xa, ya, xb, yb;
};

__device__ __forceinline__ double shfl_dn(double v, int d) { return __shfl_down(v, d, kWave); }
__device__ __forceinline__ double shfl_up(double v, int d) { return __shfl_up(v, d, kWave); }
__device__ __forceinline__ int shfl_dn(int v, int d) { return __shfl_down(v, d, kWave); }
__device__ __forceinline__ int shfl_up(int v, int d) { return __shfl_up(v, d, kWave); }

__device__ __forceinline__ Cell cell_dn(const Cell& c, int d)
{
    Cell r;
    r.s.ro = shfl_dn(c.s.ro, d); r.s.p = shfl_dn(c.s.p, d); r.s.u = shfl_dn(c.s.u, d);
    r.s.v = shfl_dn(c.s.v, d); r.s.Q = shfl_dn(c.s.Q, d);
    r.cx = shfl_dn(c.cx, d); r.cy = shfl_dn(c.cy, d);
    r.type = shfl_dn(c.type, d); r.ok = shfl_dn(c.ok, d);
    return r;
}
__device__ __forceinline__ Cell cell_up(const Cell& c, int d)
{
    Cell r;
    r.s.ro = shfl_up(c.s.ro, d); r.s.p = shfl_up(c.s.p, d); r.s.u = shfl_up(c.s.u, d);
    r.s.v = shfl_up(c.s.v, d); r.s.Q = shfl_up(c.s.Q, d);
    r.cx = shfl_up(c.cx, d); r.cy = shfl_up(c.cy, d);
    r.type = shfl_up(c.type, d); r.ok = shfl_up(c.ok, d);
    return r;
}

__device__ __forceinline__ Nodes load_nodes(const SweepArgs& a, const double* __restrict__ X,
                                            const double* __restrict__ Y, int r, int j, bool col)
{
    Nodes n{1.0, 1.0, 1.0, 1.0};
    if (col && r >= 0 && r <= a.ni) {
        const size_t k = (size_t)r * (a.nj + 1) + j;
        n.xa = X[k]; n.ya = Y[k]; n.xb = X[k + 1]; n.yb = Y[k + 1];
    }
    return n;
}

template <bool TYPES>
__device__ __forceinline__ Cell load_cell(const SweepArgs& a, int r, int j, bool col,
                                          const Nodes& lo, const Nodes& hi)
{
    Cell c;
    c.ok = (col && r >= 0 && r < a.ni) ? 1 : 0;
    c.type = GEL_C_2;
    if (c.ok) {
        const size_t nc = (size_t)a.ni * a.nj;
        const size_t k = (size_t)r * a.nj + j;
        c.s.ro = a.now[k]; c.s.p = a.now[nc + k]; c.s.u = a.now[2 * nc + k];
        c.s.v = a.now[3 * nc + k]; c.s.Q = a.now[4 * nc + k];
        if (TYPES) c.type = a.ctype[k];
        const double px[4] = { lo.xa, hi.xa, hi.xb, lo.xb };
        const double py[4] = { lo.ya, hi.ya, hi.yb, lo.yb };
        quad_center(px, py, c.cx, c.cy);
    } else {
        c.s.ro = 1.0; c.s.p = 1.0; c.s.u = 0.0; c.s.v = 0.0; c.s.Q = 0.0;
        c.cx = 1.0; c.cy = 1.0;
    }
    return c;
}

__device__ __forceinline__ double dist2d(double ax, double ay, double bx, double by)
{
    return sqrt(sq(ax - bx) + sq(ay - by));
}

// Gran::Get_par_TVD, Usualy face (Gran.cpp:294-457): master M's value at face centre
// (fx,fy) from the line (D, M, N).  D.ok == 0: mirrored ghost across the axis when M
// touches it (:297-339), else M itself (:340-343).
__device__ __forceinline__ Gas<double> tvd_usualy(const Cell& M, const Cell& N, const Cell& D, bool m_axis,
                                                  double fx, double fy, int& bad)
{
    if (!D.ok) {
        Gas<double> r = M.s;
        if (m_axis) {
            const double d2 = dist2d(fx, fy, M.cx, M.cy);
            const double d3 = dist2d(fx, fy, N.cx, N.cy);
            const double d1 = dist2d(fx, fy, M.cx, -M.cy);
            r.v = linear3(-d1, -M.s.v, -d2, M.s.v, d3, N.s.v, 0.0);
        }
        return r;
    }
    const double d2 = dist2d(fx, fy, M.cx, M.cy);
    const double d3 = dist2d(fx, fy, N.cx, N.cy);
    const double d1 = dist2d(fx, fy, D.cx, D.cy);
    Gas<double> r;
    r.ro = linear3(-d1, D.s.ro, -d2, M.s.ro, d3, N.s.ro, 0.0);
    r.p = linear3(-d1, D.s.p, -d2, M.s.p, d3, N.s.p, 0.0);
    r.u = linear3(-d1, D.s.u, -d2, M.s.u, d3, N.s.u, 0.0);
    r.v = linear3(-d1, D.s.v, -d2, M.s.v, d3, N.s.v, 0.0);
    r.Q = linear3(-d1, D.s.Q, -d2, M.s.Q, d3, N.s.Q, 0.0);
    if (r.ro <= 0.0 || r.p <= 0.0) bad = 1;
    return r;
}

// Gran::Get_par_TVD_radial, density and pressure (Gran.cpp:562-675).
__device__ __forceinline__ void tvd_radial(const Cell& M, const Cell& N, const Cell& D, bool m_axis,
                                           double fx, double fy, double gam, double& ro, double& pr)
{
    double xd, yd;
    Gas<double> sd;
    if (!D.ok) {
        if (!m_axis) { ro = M.s.ro; pr = M.s.p; return; }
        xd = M.cx; yd = -M.cy; sd = M.s;
    } else {
        xd = D.cx; yd = D.cy; sd = D.s;
    }
    const double g2 = 2.0 * gam;
    const double r = sqrt(sq(fx) + sq(fy));
    const double r2 = sqrt(sq(M.cx) + sq(M.cy));
    const double d2 = dist2d(fx, fy, M.cx, M.cy);
    const double r3 = sqrt(sq(N.cx) + sq(N.cy));
    const double d3 = dist2d(fx, fy, N.cx, N.cy);
    const double r1 = sqrt(sq(xd) + sq(yd));
    const double d1 = dist2d(fx, fy, xd, yd);
    ro = linear3(-d1, sd.ro * sq(r1), -d2, M.s.ro * sq(r2), d3, N.s.ro * sq(r3), 0.0) / sq(r);
    pr = linear3(-d1, sd.p * pow(r1, g2), -d2, M.s.p * pow(r2, g2), d3, N.s.p * pow(r3, g2), 0.0) / pow(r, g2);
}

// Radial rescaling of a C_1 value to the face radius (Setka.cpp:10221-10226).
__device__ __forceinline__ void rescale_c1(Gas<double>& g, double rc, double dis, double gam,
                                           double cx, double cy, double fx, double fy, bool rotate)
{
    g.ro = g.ro * sq(rc) / sq(dis);
    g.Q = g.Q * sq(rc) / sq(dis);
    g.p = g.p * pow(rc / dis, 2 * gam);
    if (rotate) polar_perenos(cx, cy, fx, fy, g.u, g.v);
}

// Value of cell M on its side of an interior face (both cell kinds of the sweep,
// Setka.cpp:10213-10251).
template <bool TYPES>
__device__ __forceinline__ Gas<double> side_value(const Cell& M, const Cell& N, const Cell& D, bool m_axis,
                                                  double fx, double fy, double dis, double gam, int& bad)
{
    if (TYPES && M.type == GEL_C_1) {
        Gas<double> g = M.s;
        const double rc = sqrt(sq(M.cx) + sq(M.cy));
        rescale_c1(g, rc, dis, gam, M.cx, M.cy, fx, fy, true);
        double ro, pr;
        tvd_radial(M, N, D, m_axis, fx, fy, gam, ro, pr);
        g.ro = ro; g.p = pr;
        return g;
    }
    return tvd_usualy(M, N, D, m_axis, fx, fy, bad);
}

struct Face {          // geometry of a face as seen from its owning cell
    double n1, n2, S, fx, fy, W, dis;
};

template <bool MOVING>
__device__ __forceinline__ Face face_geom(double ax, double ay, double bx, double by,
                                          double ax2, double ay2, double bx2, double by2, double dt)
{
    Face f;
    f.S = sqrt(sq(bx - ax) + sq(by - ay)) * (ay + by) * 0.5;           // Setka.cpp:10198
    f.fx = (ax + bx) / 2.0;
    f.fy = (ay + by) / 2.0;
    double t1 = bx - ax, t2 = by - ay;                                    // Gran.cpp:171-181
    const double nn = sqrt(sq(t1) + sq(t2));
    t1 = t1 / nn; t2 = t2 / nn;
    f.n1 = t2; f.n2 = -t1;
    if (MOVING) {
        const double fx2 = (ax2 + bx2) / 2.0, fy2 = (ay2 + by2) / 2.0;
        f.W = ((fx2 - f.fx) * f.n1 + (fy2 - f.fy) * f.n2) / dt;          // Setka.cpp:10209
    } else {
        f.W = 0.0;
    }
    f.dis = sqrt(sq(f.fx) + sq(f.fy));
    return f;
}

template <bool HLLC>
__device__ __forceinline__ double flux(const Gas<double>& L, const Gas<double>& R, const Face& f,
                                       double gam, double P[5])
{
    double PQ;
    const double UU = hllc2d<double, HLLC>(L, R, f.W, f.n1, f.n2, gam, P, PQ);
    P[4] = PQ;
    return UU;
}

// Boundary face of cell M (not Usualy): Gran::Get_par ghost (Gran.cpp:205-289) and the
// sweep's treatment of the cell side (Setka.cpp:10252-10274).  D is the cell across
// M from the face (Sosed_down); used on the axis.
template <bool TYPES, bool HLLC>
__device__ __forceinline__ double boundary_flux(const SweepArgs& a, int ftype, const Cell& M, const Cell& D,
                                                const Face& f, double P[5])
{
    const gel_params& p = a.prm;
    const double gam = p.gamma;
    Gas<double> L = M.s, R = M.s;
    const double radius = sqrt(sq(M.cx) + sq(M.cy));
    if (ftype == GEL_F_INNER_SPHERE) {
        rescale_c1(L, radius, f.dis, gam, M.cx, M.cy, f.fx, f.fy, false);
        const double d = sqrt(f.fx * f.fx + f.fy * f.fy);
        const double ro = 1.0 / (p.in_rho_div * d * d);
        R.ro = ro;
        R.p = ro * p.in_p_a * p.in_p_a / p.in_p_den;
        R.u = p.in_u * f.fx / d;
        R.v = p.in_u * f.fy / d;
        R.Q = ro * p.in_q;
    } else if (ftype == GEL_F_AXIS) {
        if (TYPES && M.type == GEL_C_1) {
            rescale_c1(L, radius, f.dis, gam, M.cx, M.cy, f.fx, f.fy, true);
        } else {                                             // Gran.cpp:486-519
            const double d2 = dist2d(f.fx, f.fy, M.cx, M.cy);
            const double d1 = dist2d(f.fx, f.fy, D.cx, D.cy);
            L.v = linear3(-d1, D.s.v, -d2, M.s.v, d2, -M.s.v, 0.0);
        }
        R = L;
        R.v = -R.v;
    } else if (ftype == GEL_F_INPUT) {
        R.ro = p.inf_rho; R.p = p.inf_p; R.u = p.inf_u; R.v = p.inf_v; R.Q = p.inf_q;
    } else if (ftype == GEL_F_EXTERN) {
        if (R.u >= 0.0) R.u = p.extern_u;
    }
    return flux<HLLC>(L, R, f, gam, P);
}


template <bool TYPES, bool HLLC>
__device__ __forceinline__ double interior_flux(const Cell& M, const Cell& N, const Cell& D, const Cell& U,
                                                bool m_axis, bool n_axis, const Face& f, double gam,
                                                double P[5], int& bad)
{
    const Gas<double> L = side_value<TYPES>(M, N, D, m_axis, f.fx, f.fy, f.dis, gam, bad);
    const Gas<double> R = side_value<TYPES>(N, M, U, n_axis, f.fx, f.fy, f.dis, gam, bad);
    return flux<HLLC>(L, R, f, gam, P);
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

// One step of the sweep over a chunk of rows for a strip of columns.
template <bool MOVING, bool TYPES, bool SRC, bool HLLC>
__global__ void __launch_bounds__(kWave)
sweep_kernel(SweepArgs a)
{
    const int lane = threadIdx.x;
    const int j = blockIdx.x * kOut + lane - kHalo;
    const bool col = (j >= 0 && j < a.nj);
    const bool out_lane = col && lane >= kHalo && lane < kWave - kHalo;
    const int i0 = blockIdx.y * a.rows;
    const int i1 = min(i0 + a.rows, a.ni);
    const double gam = a.prm.gamma;
    const double cfl = a.prm.cfl;
    const double dt = __longlong_as_double((long long)a.dtbuf[a.dt_rd]);
    const bool axis_lo = (j == 0), axis_hi = (j == a.nj - 1);
    const bool m_axis = axis_lo || axis_hi;
    const bool n_axis = (j + 1 == a.nj - 1);
    const size_t nc = (size_t)a.ni * a.nj;

    if (blockIdx.x == 0 && blockIdx.y == 0 && lane == 0)
        a.dtbuf[a.dt_clr] = (unsigned long long)__double_as_longlong(__builtin_huge_val());

    // window: cells of rows i-1, i, i+1, i+2 and node rows i .. i+3
    Nodes nr[4];
    Cell w[4];
    const Nodes nm2 = load_nodes(a, a.x, a.y, i0 - 2, j, col);
    const Nodes nm1 = load_nodes(a, a.x, a.y, i0 - 1, j, col);
#pragma unroll
    for (int k = 0; k < 4; k++) nr[k] = load_nodes(a, a.x, a.y, i0 + k, j, col);
    const Cell cm2 = load_cell<TYPES>(a, i0 - 2, j, col, nm2, nm1);
    w[0] = load_cell<TYPES>(a, i0 - 1, j, col, nm1, nr[0]);
#pragma unroll
    for (int k = 1; k < 4; k++) w[k] = load_cell<TYPES>(a, i0 + k - 1, j, col, nr[k - 1], nr[k]);

    int bad = 0;
    // flux through the radial face below row i0, oriented along +i
    double flo[5], uulo;
    {
        Nodes q = nr[0], qn = nr[0];
        if (MOVING) qn = load_nodes(a, a.x2, a.y2, i0, j, col);
        double P[5];
        if (i0 == 0) {
            // inner sphere, edge 3 of cell (0, j): (0, j+1) -> (0, j)
            const Face f = face_geom<MOVING>(q.xb, q.yb, q.xa, q.ya, qn.xb, qn.yb, qn.xa, qn.ya, dt);
            uulo = boundary_flux<TYPES, HLLC>(a, GEL_F_INNER_SPHERE, w[1], w[2], f, P);
#pragma unroll
            for (int k = 0; k < 5; k++) flo[k] = -P[k];
        } else {
            // edge 1 of cell (i0-1, j): (i0, j) -> (i0, j+1)
            const Face f = face_geom<MOVING>(q.xa, q.ya, q.xb, q.yb, qn.xa, qn.ya, qn.xb, qn.yb, dt);
            uulo = interior_flux<TYPES, HLLC>(w[0], w[1], cm2, w[2], m_axis, m_axis, f, gam, flo, bad);
        }
    }

    double tmin = __builtin_huge_val();
    int n_rho = 0, n_p = 0;
    for (int i = i0; i < i1; i++) {
        const Nodes q0 = nr[0], q1 = nr[1];
        Nodes q0n = q0, q1n = q1;
        if (MOVING) {
            q0n = load_nodes(a, a.x2, a.y2, i, j, col);
            q1n = load_nodes(a, a.x2, a.y2, i + 1, j, col);
        }
        // edge 1: radial face (i, i+1), (i+1, j) -> (i+1, j+1)
        const Face f1 = face_geom<MOVING>(q1.xa, q1.ya, q1.xb, q1.yb, q1n.xa, q1n.ya, q1n.xb, q1n.yb, dt);
        double P1[5], uu1;
        if (i + 1 < a.ni) {
            uu1 = interior_flux<TYPES, HLLC>(w[1], w[2], w[0], w[3], m_axis, m_axis, f1, gam, P1, bad);
        } else {
            const int ot = col ? (int)a.otype[j] : (int)GEL_F_EXTERN;
            uu1 = boundary_flux<TYPES, HLLC>(a, ot, w[1], w[0], f1, P1);
        }
        // edge 2: polar face (j, j+1), (i+1, j+1) -> (i, j+1)
        const Face f2 = face_geom<MOVING>(q1.xb, q1.yb, q0.xb, q0.yb, q1n.xb, q1n.yb, q0n.xb, q0n.yb, dt);
        const Cell cN = cell_dn(w[1], 1), cD = cell_up(w[1], 1), cU = cell_dn(w[1], 2);
        double P2[5], uu2;
        if (axis_hi) uu2 = boundary_flux<TYPES, HLLC>(a, GEL_F_AXIS, w[1], cD, f2, P2);
        else uu2 = interior_flux<TYPES, HLLC>(w[1], cN, cD, cU, m_axis, n_axis, f2, gam, P2, bad);
        // edge 0: polar face (j-1, j), (i, j) -> (i+1, j); solved by the lane below
        const Face f0 = face_geom<MOVING>(q0.xa, q0.ya, q1.xa, q1.ya, q0n.xa, q0n.ya, q1n.xa, q1n.ya, dt);
        double P0[5], uu0;
        {
            double Pl[5];
#pragma unroll
            for (int k = 0; k < 5; k++) Pl[k] = shfl_up(P2[k], 1);
            const double uul = shfl_up(uu2, 1);
            if (axis_lo) {
                uu0 = boundary_flux<TYPES, HLLC>(a, GEL_F_AXIS, w[1], cN, f0, P0);
            } else {
#pragma unroll
                for (int k = 0; k < 5; k++) P0[k] = -Pl[k];
                uu0 = uul;
            }
        }
        // edge 3: radial face (i-1, i), (i, j+1) -> (i, j)
        const Face f3 = face_geom<MOVING>(q0.xb, q0.yb, q0.xa, q0.ya, q0n.xb, q0n.yb, q0n.xa, q0n.ya, dt);

        // cell update (Setka.cpp:10372-10468)
        const Cell& K = w[1];
        const double x = K.cx, y = K.cy;
        const double radius = sqrt(sq(x) + sq(y));
        if (out_lane && !(radius < a.prm.r_skip)) {
            double Pot[5];
#pragma unroll
            for (int k = 0; k < 5; k++) {
                double s = 0.0;
                s = s + P0[k] * f0.S;
                s = s + P1[k] * f1.S;
                s = s + P2[k] * f2.S;
                s = s + (-flo[k]) * f3.S;
                Pot[k] = s;
            }
            double time = 10000000;
            time = mn(time, cfl * sqrt(sq(x - f0.fx) + sq(y - f0.fy)) / uu0);
            time = mn(time, cfl * sqrt(sq(x - f1.fx) + sq(y - f1.fy)) / uu1);
            time = mn(time, cfl * sqrt(sq(x - f2.fx) + sq(y - f2.fy)) / uu2);
            time = mn(time, cfl * sqrt(sq(x - f3.fx) + sq(y - f3.fy)) / uulo);
            tmin = fmin(tmin, time);

            const size_t kc = (size_t)i * a.nj + j;
            if (TYPES && K.type == GEL_C_CENTR) {
                a.next[kc] = K.s.ro; a.next[nc + kc] = K.s.p; a.next[2 * nc + kc] = K.s.u;
                a.next[3 * nc + kc] = K.s.v; a.next[4 * nc + kc] = K.s.Q;
            } else {
                const double px[4] = { q0.xa, q1.xa, q1.xb, q0.xb };
                const double py[4] = { q0.ya, q1.ya, q1.yb, q0.yb };
                const double vs1 = quad_area(px, py);
                const double Volume = quad_volume_r(px, py);
                double vs2 = vs1, Volume2 = Volume;
                if (MOVING) {
                    const double px2[4] = { q0n.xa, q1n.xa, q1n.xb, q0n.xb };
                    const double py2[4] = { q0n.ya, q1n.ya, q1n.yb, q0n.yb };
                    vs2 = quad_area(px2, py2);
                    Volume2 = quad_volume_r(px2, py2);
                }
                const double u = K.s.u, v = K.s.v, ro = K.s.ro, p = K.s.p, Q = K.s.Q;
                double q1s = 0.0, q2s = 0.0, q3s = 0.0;
                if (SRC) {
                    const double* __restrict__ sp = a.src + kc;
                    source_mk1(a.prm, [&](int k) { return sp[(size_t)k * nc]; }, u, v, ro, p, q1s, q2s, q3s);
                }
                const double T1 = dt;
                double ro3 = ro * Volume / Volume2 - T1 * (Pot[0] / Volume2);
                const double Q33 = Q * Volume / Volume2 - (T1 / Volume2) * Pot[4];
                if (ro3 <= 0.0) { ro3 = 0.00001; n_rho++; }
                const double u3 = (ro * u * Volume / Volume2 - T1 * (Pot[1] / Volume2 - q1s)) / ro3;
                const double v3 = (ro * v * Volume / Volume2 - T1 * ((Pot[2] - p * 0.5 * (vs1 + vs2)) / Volume2 - q2s)) / ro3;
                double p3 = (((p / (gam - 1) + ro * (u * u + v * v) * 0.5) * Volume / Volume2 - T1 * (Pot[3] / Volume2 - q3s)) -
                             0.5 * ro3 * (u3 * u3 + v3 * v3)) * (gam - 1);
                if (p3 <= 0) { p3 = 0.000001; n_p++; }
                a.next[kc] = ro3;
                a.next[nc + kc] = p3;
                a.next[2 * nc + kc] = u3;
                a.next[3 * nc + kc] = v3;
                a.next[4 * nc + kc] = Q33;
            }
        }

        // slide the window one row up
#pragma unroll
        for (int k = 0; k < 5; k++) flo[k] = P1[k];
        uulo = uu1;
        w[0] = w[1]; w[1] = w[2]; w[2] = w[3];
        nr[0] = nr[1]; nr[1] = nr[2]; nr[2] = nr[3];
        nr[3] = load_nodes(a, a.x, a.y, i + 4, j, col);
        w[3] = load_cell<TYPES>(a, i + 3, j, col, nr[2], nr[3]);
    }

    // the numeric flag counts on lanes whose faces are used (edge 0 of lane 2 is lane 1's)
    if (!col || lane < 1 || lane >= kWave - kHalo) bad = 0;
    tmin = wave_min(tmin);
    const int s_rho = wave_sum(n_rho), s_p = wave_sum(n_p), s_bad = wave_sum(bad);
    if (lane == 0) {
        if (tmin < __builtin_huge_val())
            atomicMin(&a.dtbuf[a.dt_wr], (unsigned long long)__double_as_longlong(tmin));
        if (s_rho) atomicAdd(&a.counters[0], (unsigned long long)s_rho);
        if (s_p) atomicAdd(&a.counters[1], (unsigned long long)s_p);
        if (s_bad) atomicAdd(&a.counters[2], (unsigned long long)s_bad);
    }
}

// Batched face flux for tests (gel_flux_hllc2d_batch).
__global__ void flux_batch_kernel(int n, const double* __restrict__ L, const double* __restrict__ R,
                                  const double* __restrict__ W, const double* __restrict__ n1,
                                  const double* __restrict__ n2, const double* __restrict__ rad, double cfl,
                                  double gam, int metod, int nul, double* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Gas<double> gl{ L[5 * k], L[5 * k + 2], L[5 * k + 3], L[5 * k + 4], L[5 * k + 1] };
    const Gas<double> gr{ R[5 * k], R[5 * k + 2], R[5 * k + 3], R[5 * k + 4], R[5 * k + 1] };
    double P[4], PQ, Vl, Vc, Vp;
    const double t = hllc2d_full<double>(gl, gr, W[k], n1[k], n2[k], rad[k], cfl, gam, metod, nul, P, PQ, Vl, Vc, Vp);
    double* o = out + 9 * (size_t)k;
    o[0] = P[0]; o[1] = P[1]; o[2] = P[2]; o[3] = P[3]; o[4] = PQ; o[5] = Vl; o[6] = Vc; o[7] = Vp; o[8] = t;
}

}  // namespace gel

// ------------------------------------------------------------------ C-ABI ------

using namespace gel;

static thread_local std::string g_last_error;

static int fail(int st, const std::string& msg)
{
    g_last_error = msg;
    return st;
}

#define HIP_OK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(GEL_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct gel_ctx {
    int ni = 0, nj = 0;
    size_t nc = 0, nn = 0;
    gel_params prm{};
    double* d_x = nullptr;
    double* d_y = nullptr;
    double* d_x2 = nullptr;       // next node positions (moving steps)
    double* d_y2 = nullptr;
    double* d_state[2] = { nullptr, nullptr };
    int cur = 0;
    int8_t* d_ctype = nullptr;
    int8_t* d_otype = nullptr;
    bool has_types = false;
    double* d_src = nullptr;
    unsigned long long* d_dt = nullptr;   // [3]
    unsigned long long* d_cnt = nullptr;  // [3]
    long long steps = 0;
    bool uploaded = false;
    int rows = 8;
    hipStream_t stream = nullptr;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};

extern "C" {

const char* gel_strerror(int st)
{
    switch (st) {
    case GEL_OK: return "ok";
    case GEL_E_ARG: return "invalid argument";
    case GEL_E_HIP: return "HIP runtime error";
    case GEL_E_NOMEM: return "device allocation failed";
    case GEL_E_UNSUPPORTED: return "unsupported option combination";
    case GEL_E_STATE: return "call out of order";
    case GEL_E_NUMERIC: return "non-positive reconstructed density or pressure";
    default: return "unknown status";
    }
}

const char* gel_last_error(void) { return g_last_error.c_str(); }

void gel_params_reference(gel_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    const double ga = 5.0 / 3.0, chi = 36.1275;
    p->gamma = ga;
    p->cfl = 0.1;
    p->dt0 = 0.00000001;
    p->r_skip = 38.0 / 390.222;
    p->flux = GEL_FLUX_HLL;
    p->physics = GEL_PHYS_GAS;
    p->source_mode = GEL_SRC_NONE;
    p->in_rho_div = chi * chi;
    p->in_p_a = chi;
    p->in_p_den = ga * 10.0 * 10.0;
    p->in_u = chi;
    p->in_q = 1.0;
    p->inf_rho = 1.0; p->inf_p = 1.0; p->inf_u = -2.54338; p->inf_v = 0.0; p->inf_q = 100.0;
    p->extern_u = -0.1;
    p->src_coef = 3.0 / 6.0;
    p->sig_a = 1.0;
    p->sig_b = 0.1307345665;
}

static void free_ctx(gel_ctx* c)
{
    if (!c) return;
    hipFree(c->d_x); hipFree(c->d_y); hipFree(c->d_x2); hipFree(c->d_y2);
    hipFree(c->d_state[0]); hipFree(c->d_state[1]);
    hipFree(c->d_ctype); hipFree(c->d_otype); hipFree(c->d_src);
    hipFree(c->d_dt); hipFree(c->d_cnt);
    for (auto& e : c->events) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

void gel_ctx_destroy(gel_ctx* ctx) { free_ctx(ctx); }

static int choose_rows(int ni, int nj)
{
    // enough wavefronts to give every SIMD of the 256 CUs several of them
    const int strips = (nj + kOut - 1) / kOut;
    const long target = 4096;
    int rows = (int)((long)ni * strips / target);
    if (rows < 2) rows = 2;
    if (rows > 64) rows = 64;
    return rows;
}

int gel_ctx_create(gel_ctx** out, const gel_grid* g, const gel_params* prm)
{
    if (!out || !g || !prm || !g->x || !g->y) return fail(GEL_E_ARG, "null argument");
    *out = nullptr;
    if (g->ni < 3 || g->nj < 3) return fail(GEL_E_ARG, "grid needs at least 3x3 cells");
    if (prm->physics != GEL_PHYS_GAS) return fail(GEL_E_UNSUPPORTED, "physics");
    if (prm->flux != GEL_FLUX_HLL && prm->flux != GEL_FLUX_HLLC) return fail(GEL_E_UNSUPPORTED, "flux");
    if (prm->source_mode != GEL_SRC_NONE && prm->source_mode != GEL_SRC_MK1) return fail(GEL_E_UNSUPPORTED, "source_mode");
    gel_ctx* c = new (std::nothrow) gel_ctx();
    if (!c) return fail(GEL_E_NOMEM, "host allocation");
    c->ni = g->ni; c->nj = g->nj;
    c->nc = (size_t)g->ni * g->nj;
    c->nn = (size_t)(g->ni + 1) * (g->nj + 1);
    c->prm = *prm;
    c->rows = choose_rows(g->ni, g->nj);
    auto bad = [&](hipError_t e, const char* what) {
        free_ctx(c);
        return fail(e == hipErrorOutOfMemory ? GEL_E_NOMEM : GEL_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e, "stream");
    const size_t nb = c->nn * sizeof(double);
    if ((e = hipMalloc(&c->d_x, nb)) != hipSuccess) return bad(e, "nodes");
    if ((e = hipMalloc(&c->d_y, nb)) != hipSuccess) return bad(e, "nodes");
    if ((e = hipMalloc(&c->d_x2, nb)) != hipSuccess) return bad(e, "nodes");
    if ((e = hipMalloc(&c->d_y2, nb)) != hipSuccess) return bad(e, "nodes");
    for (int b = 0; b < 2; b++)
        if ((e = hipMalloc(&c->d_state[b], 5 * c->nc * sizeof(double))) != hipSuccess) return bad(e, "state");
    if ((e = hipMalloc(&c->d_otype, (size_t)g->nj)) != hipSuccess) return bad(e, "outer types");
    if ((e = hipMalloc(&c->d_dt, 3 * sizeof(unsigned long long))) != hipSuccess) return bad(e, "dt");
    if ((e = hipMalloc(&c->d_cnt, 3 * sizeof(unsigned long long))) != hipSuccess) return bad(e, "counters");
    if ((e = hipMemcpy(c->d_x, g->x, nb, hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "copy");
    if ((e = hipMemcpy(c->d_y, g->y, nb, hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "copy");
    std::vector<int8_t> ot((size_t)g->nj);
    for (int jj = 0; jj < g->nj; jj++) {
        if (g->outer_type) {
            ot[jj] = g->outer_type[jj];
        } else {
            const size_t a = (size_t)g->ni * (g->nj + 1) + jj;
            ot[jj] = (0.5 * (g->x[a] + g->x[a + 1]) > 0.0) ? GEL_F_INPUT : GEL_F_EXTERN;
        }
        if (ot[jj] != GEL_F_INPUT && ot[jj] != GEL_F_EXTERN && ot[jj] != GEL_F_UPPER_WALL) {
            free_ctx(c);
            return fail(GEL_E_ARG, "outer_type must be Input, Extern or Upper_wall");
        }
    }
    if ((e = hipMemcpy(c->d_otype, ot.data(), ot.size(), hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "copy");
    if (g->cell_type) {
        bool any = false;
        for (size_t k = 0; k < c->nc; k++) {
            const int t = g->cell_type[k];
            if (t < GEL_C_1 || t > GEL_C_NO) { free_ctx(c); return fail(GEL_E_ARG, "cell_type out of range"); }
            if (t == GEL_C_1 || t == GEL_C_CENTR) any = true;
        }
        if (any) {
            if ((e = hipMalloc(&c->d_ctype, c->nc)) != hipSuccess) return bad(e, "cell types");
            if ((e = hipMemcpy(c->d_ctype, g->cell_type, c->nc, hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "copy");
            c->has_types = true;
        }
    }
    if ((e = hipMemset(c->d_cnt, 0, 3 * sizeof(unsigned long long))) != hipSuccess) return bad(e, "memset");
    *out = c;
    return GEL_OK;
}

static unsigned long long dbits(double v)
{
    unsigned long long b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}
static double bitsd(unsigned long long b)
{
    double v;
    std::memcpy(&v, &b, sizeof v);
    return v;
}

int gel_upload_state(gel_ctx* c, const double* prim)
{
    if (!c || !prim) return fail(GEL_E_ARG, "null argument");
    HIP_OK(hipMemcpyAsync(c->d_state[0], prim, 5 * c->nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(c->d_state[1], prim, 5 * c->nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const unsigned long long dt[3] = { dbits(c->prm.dt0), dbits(__builtin_huge_val()), dbits(__builtin_huge_val()) };
    HIP_OK(hipMemcpyAsync(c->d_dt, dt, sizeof dt, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    c->cur = 0;
    c->steps = 0;
    c->uploaded = true;
    return GEL_OK;
}

int gel_download_state(gel_ctx* c, double* prim)
{
    if (!c || !prim) return fail(GEL_E_ARG, "null argument");
    if (!c->uploaded) return fail(GEL_E_STATE, "no state uploaded");
    HIP_OK(hipMemcpyAsync(prim, c->d_state[c->cur], 5 * c->nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return GEL_OK;
}

int gel_set_sources(gel_ctx* c, const double* f)
{
    if (!c || !f) return fail(GEL_E_ARG, "null argument");
    const size_t nb = (size_t)GEL_MK1_NFIELDS * c->nc * sizeof(double);
    if (!c->d_src) {
        hipError_t e = hipMalloc(&c->d_src, nb);
        if (e != hipSuccess) return fail(GEL_E_NOMEM, "sources");
    }
    HIP_OK(hipMemcpy(c->d_src, f, nb, hipMemcpyHostToDevice));
    return GEL_OK;
}

int gel_set_rows_per_wave(gel_ctx* c, int rows)
{
    if (!c || rows < 1 || rows > 4096) return fail(GEL_E_ARG, "rows");
    c->rows = rows;
    return GEL_OK;
}

}  // extern "C"

template <bool MOVING, bool TYPES, bool SRC, bool HLLC>
static hipError_t launch_t(const SweepArgs& a, dim3 grid, hipStream_t s)
{
    hipLaunchKernelGGL((sweep_kernel<MOVING, TYPES, SRC, HLLC>), grid, dim3(kWave), 0, s, a);
    return hipGetLastError();
}

static hipError_t launch(const SweepArgs& a, dim3 grid, hipStream_t s, bool moving, bool types, bool src, bool hllc)
{
    const int key = (moving ? 8 : 0) | (types ? 4 : 0) | (src ? 2 : 0) | (hllc ? 1 : 0);
    switch (key) {
#define K(m, t, sr, h) case ((m) * 8 + (t) * 4 + (sr) * 2 + (h)): return launch_t<m, t, sr, h>(a, grid, s);
    K(0, 0, 0, 0) K(0, 0, 0, 1) K(0, 0, 1, 0) K(0, 0, 1, 1) K(0, 1, 0, 0) K(0, 1, 0, 1) K(0, 1, 1, 0) K(0, 1, 1, 1)
    K(1, 0, 0, 0) K(1, 0, 0, 1) K(1, 0, 1, 0) K(1, 0, 1, 1) K(1, 1, 0, 0) K(1, 1, 0, 1) K(1, 1, 1, 0) K(1, 1, 1, 1)
#undef K
    }
    return hipErrorInvalidValue;
}

static int one_step(gel_ctx* c, bool moving)
{
    SweepArgs a;
    a.ni = c->ni; a.nj = c->nj; a.rows = c->rows;
    a.dt_rd = (int)(c->steps % 3);
    a.dt_wr = (int)((c->steps + 1) % 3);
    a.dt_clr = (int)((c->steps + 2) % 3);
    a.x = c->d_x; a.y = c->d_y;
    a.x2 = moving ? c->d_x2 : c->d_x;
    a.y2 = moving ? c->d_y2 : c->d_y;
    a.now = c->d_state[c->cur];
    a.next = c->d_state[c->cur ^ 1];
    a.ctype = c->d_ctype; a.otype = c->d_otype;
    a.src = c->d_src;
    a.dtbuf = c->d_dt;
    a.counters = c->d_cnt;
    a.prm = c->prm;
    const bool src = (c->prm.source_mode == GEL_SRC_MK1);
    if (src && !c->d_src) return fail(GEL_E_STATE, "sources not set");
    dim3 grid((c->nj + kOut - 1) / kOut, (c->ni + c->rows - 1) / c->rows);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        HIP_OK(hipEventRecord(e0, c->stream));
    }
    hipError_t e = launch(a, grid, c->stream, moving, c->has_types, src, c->prm.flux == GEL_FLUX_HLLC);
    if (e != hipSuccess) return fail(GEL_E_HIP, std::string("sweep launch: ") + hipGetErrorString(e));
    if (c->timing) {
        HIP_OK(hipEventRecord(e1, c->stream));
        c->events.emplace_back(e0, e1);
    }
    c->cur ^= 1;
    c->steps++;
    return GEL_OK;
}

static int fill_stats(gel_ctx* c, gel_stats* out)
{
    if (!out) return GEL_OK;
    unsigned long long dt[3], cnt[3];
    HIP_OK(hipMemcpyAsync(dt, c->d_dt, sizeof dt, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(cnt, c->d_cnt, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    out->steps = c->steps;
    out->dt_next = bitsd(dt[c->steps % 3]);
    out->dt_last = c->steps > 0 ? bitsd(dt[(c->steps - 1) % 3]) : c->prm.dt0;
    out->n_rho_clamp = (long long)cnt[0];
    out->n_p_clamp = (long long)cnt[1];
    out->n_numeric = (long long)cnt[2];
    return GEL_OK;
}

extern "C" {

int gel_step(gel_ctx* c, int nsteps, gel_stats* out)
{
    if (!c || nsteps < 0) return fail(GEL_E_ARG, "bad argument");
    if (!c->uploaded) return fail(GEL_E_STATE, "no state uploaded");
    for (int s = 0; s < nsteps; s++) {
        const int st = one_step(c, false);
        if (st != GEL_OK) return st;
    }
    return fill_stats(c, out);
}

int gel_step_moving(gel_ctx* c, const double* x2, const double* y2, gel_stats* out)
{
    if (!c || !x2 || !y2) return fail(GEL_E_ARG, "null argument");
    if (!c->uploaded) return fail(GEL_E_STATE, "no state uploaded");
    const size_t nb = c->nn * sizeof(double);
    HIP_OK(hipMemcpyAsync(c->d_x2, x2, nb, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(c->d_y2, y2, nb, hipMemcpyHostToDevice, c->stream));
    const int st = one_step(c, true);
    if (st != GEL_OK) return st;
    std::swap(c->d_x, c->d_x2);       // Point::x = x2 after the step (Setka.cpp:10473-10480)
    std::swap(c->d_y, c->d_y2);
    return fill_stats(c, out);
}

int gel_sync(gel_ctx* c, gel_stats* out)
{
    if (!c) return fail(GEL_E_ARG, "null argument");
    HIP_OK(hipStreamSynchronize(c->stream));
    return fill_stats(c, out);
}

int gel_set_dt(gel_ctx* c, double dt)
{
    if (!c || !(dt > 0)) return fail(GEL_E_ARG, "dt must be positive");
    const unsigned long long b = dbits(dt);
    HIP_OK(hipMemcpyAsync(c->d_dt + (c->steps % 3), &b, sizeof b, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return GEL_OK;
}

int gel_enable_timing(gel_ctx* c, int on)
{
    if (!c) return fail(GEL_E_ARG, "null argument");
    c->timing = on != 0;
    return GEL_OK;
}

int gel_kernel_timing(gel_ctx* c, double* avg_ms, long long* n)
{
    if (!c || !avg_ms || !n) return fail(GEL_E_ARG, "null argument");
    HIP_OK(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (auto& e : c->events) {
        float ms = 0.0f;
        HIP_OK(hipEventElapsedTime(&ms, e.first, e.second));
        tot += ms;
        hipEventDestroy(e.first);
        hipEventDestroy(e.second);
    }
    *n = (long long)c->events.size();
    *avg_ms = c->events.empty() ? 0.0 : tot / (double)c->events.size();
    c->events.clear();
    return GEL_OK;
}

int gel_flux_hllc2d_batch(int n, const double* L, const double* R, const double* W, const double* n1,
                          const double* n2, const double* rad, double cfl, double gamma, int metod, int nul_potok,
                          double* P, double* PQ, double* Vl, double* Vc, double* Vp, double* dt)
{
    if (n < 0 || (n > 0 && (!L || !R || !W || !n1 || !n2 || !rad || !P || !PQ || !Vl || !Vc || !Vp || !dt)))
        return fail(GEL_E_ARG, "null argument");
    if (n == 0) return GEL_OK;
    double *dL = nullptr, *dR = nullptr, *dW = nullptr, *d1 = nullptr, *d2 = nullptr, *dr = nullptr, *dout = nullptr;
    auto cleanup = [&]() { hipFree(dL); hipFree(dR); hipFree(dW); hipFree(d1); hipFree(d2); hipFree(dr); hipFree(dout); };
    const size_t s = (size_t)n * sizeof(double);
    if (hipMalloc(&dL, 5 * s) != hipSuccess || hipMalloc(&dR, 5 * s) != hipSuccess || hipMalloc(&dW, s) != hipSuccess ||
        hipMalloc(&d1, s) != hipSuccess || hipMalloc(&d2, s) != hipSuccess || hipMalloc(&dr, s) != hipSuccess ||
        hipMalloc(&dout, 9 * s) != hipSuccess) {
        cleanup();
        return fail(GEL_E_NOMEM, "flux batch buffers");
    }
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMemcpy(dL, L, 5 * s, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dR, R, 5 * s, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dW, W, s, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d1, n1, s, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d2, n2, s, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dr, rad, s, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(flux_batch_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, n, dL, dR, dW, d1, d2, dr,
                           cfl, gamma, metod, nul_potok, dout);
        e = hipGetLastError();
    }
    std::vector<double> o((size_t)9 * n);
    if (e == hipSuccess) e = hipMemcpy(o.data(), dout, 9 * s, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return fail(GEL_E_HIP, std::string("flux batch: ") + hipGetErrorString(e));
    for (int k = 0; k < n; k++) {
        const double* q = &o[(size_t)9 * k];
        P[4 * k] = q[0]; P[4 * k + 1] = q[1]; P[4 * k + 2] = q[2]; P[4 * k + 3] = q[3];
        PQ[k] = q[4]; Vl[k] = q[5]; Vc[k] = q[6]; Vp[k] = q[7]; dt[k] = q[8];
    }
    return GEL_OK;
}

}  // extern "C"
