/*
 * gelios_oracle.c - CPU restatement of the reference sweep.  TEST INFRASTRUCTURE:
 * loaded only by tests, the smoke check and the CPU-baseline leg of the benchmark,
 * never by the product path.
 *
 * Restates Setka::Go_5_komponent__MK2 (Setka.cpp:10121-10482), the face flux
 * Setka::HLLC_2d_Korolkov_b_s (Setka.cpp:10484-10700), the boundary states
 * Gran::Get_par (Gran.cpp:205-289), the reconstructions Gran::Get_par_TVD
 * (Gran.cpp:292-559) and Gran::Get_par_TVD_radial (Gran.cpp:562-675), the cell and
 * face geometry (Cell.cpp:66-275, Gran.cpp:159-198), the limiter helpers
 * (Help.cpp:3-120) and the charge-exchange source Cell::Get_Sourse_MK1
 * (Cell.cpp:498-665).  Expressions keep the reference's operation order.
 *
 * Fitted surfaces: the sweep's exact-Godunov faces between nodes of one surface and
 * nul_potok faces beside them (Setka.cpp:10287-10370), the one-sided reconstruction on
 * surface faces (Gran.cpp:345-426), and the surface tracking of Setka::Move_surface
 * (Setka.cpp:3378-3844) with the A-rail redistribution of Move_Setka_Calculate_2
 * (Setka.cpp:4324-4381) on the polar grid (ora_track_nodes).
 *
 * Not restated here: the radial-chain re-interpolation of the MC moments in
 * Get_Sourse_MK1 (Cell.cpp:528-589), which on a static grid returns the cell's own
 * moments, and the B/C/D rails of the reference's own block grid.
 */
#include "gelios_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_ 3.14159265358979323846   /* Help.h:50 */

/* std::min / std::max semantics (return the first argument on ties). */
static inline double mn(double a, double b) { return (b < a) ? b : a; }
static inline double mx(double a, double b) { return (a < b) ? b : a; }
static inline double sq(double a) { return a * a; }

/* ------------------------------------------------------------------ helpers --- */

static double sgn(double x)   /* Help.cpp:80-94 */
{
    if (x > 0) return 1.0;
    if (x < 0) return -1.0;
    return 0.0;
}

double ora_minmod(double x, double y)   /* Help.cpp:67-78 */
{
    const double s = sgn(x) + sgn(y);
    if (s == 0) return 0.0;
    return (s / 2.0) * mn(fabs(x), fabs(y));
}

double ora_linear3(double x1, double t1, double x2, double t2, double x3, double t3, double y)
{   /* Help.cpp:96-113, the minmod branch */
    const double d = ora_minmod((t1 - t2) / (x1 - x2), (t2 - t3) / (x2 - x3));
    return d * (y - x2) + t2;
}

double ora_linear2(double x1, double t1, double x2, double t2, double y)  /* Help.cpp:115-120 */
{
    const double d = (t1 - t2) / (x1 - x2);
    return d * (y - x2) + t2;
}

double ora_polar_angle(double x, double y)   /* Help.cpp:37-65 */
{
    if (fabs(x) + fabs(y) < 0.000001 / 390.222) return 0.0;   /* RR_, Help.h:79 */
    if (x < 0) return atan(y / x) + 1.0 * PI_;
    if (x > 0 && y >= 0) return atan(y / x);
    if (x > 0 && y < 0) return atan(y / x) + 2.0 * PI_;
    if (y > 0) return PI_ / 2.0;
    if (y < 0) return 3.0 * PI_ / 2.0;
    return 0.0;
}

void ora_polar_perenos(double x1, double y1, double x2, double y2, double* u, double* v)
{   /* Help.cpp:3-14: rotate (u,v) from the polar frame at (x1,y1) to the one at (x2,y2) */
    const double a1 = ora_polar_angle(x1, y1);
    const double a2 = ora_polar_angle(x2, y2);
    const double fr = *u * cos(a1) + *v * sin(a1);
    const double ff = -*u * sin(a1) + *v * cos(a1);
    *u = fr * cos(a2) - ff * sin(a2);
    *v = fr * sin(a2) + ff * cos(a2);
}

/* ----------------------------------------------------------------- geometry --- */

void ora_cell_center(const double* px, const double* py, const int* q, double* cx, double* cy)
{   /* Cell::Get_Center, polygon barycentre (Cell.cpp:66-88) */
    double A = 0.0, Gx = 0.0, Gy = 0.0;
    for (int k = 0; k < 4; k++) {
        const int a = q[k], b = q[(k + 1) & 3];
        const double cr = px[a] * py[b] - py[a] * px[b];
        A = A + cr;
        Gx = Gx + (px[a] + px[b]) * cr;
        Gy = Gy + (py[a] + py[b]) * cr;
    }
    A = A * 0.5;
    *cx = Gx / (6.0 * A);
    *cy = Gy / (6.0 * A);
}

double ora_cell_area(const double* px, const double* py, const int* q)
{   /* Cell::Get_Volume for quads: half product of the diagonals times sin (Cell.cpp:146-166) */
    double a1 = px[q[0]] - px[q[2]], a2 = py[q[0]] - py[q[2]];
    const double a = sqrt(sq(a1) + sq(a2));
    double b1 = px[q[1]] - px[q[3]], b2 = py[q[1]] - py[q[3]];
    const double b = sqrt(sq(b1) + sq(b2));
    a1 = a1 / a; a2 = a2 / a;
    b1 = b1 / b; b2 = b2 / b;
    const double c = a1 * b1 + a2 * b2;
    const double s = sqrt(1 - sq(c));
    return 0.5 * a * b * s;
}

double ora_cell_volume_rotate(const double* px, const double* py, const int* q, double angle)
{   /* Cell::Get_Volume_rotate (Cell.cpp:188-206): Pappus volume of the swept polygon */
    double A = 0.0, G = 0.0;
    for (int k = 0; k < 4; k++) {
        const int a = q[k], b = q[(k + 1) & 3];
        const double cr = px[a] * py[b] - py[a] * px[b];
        A = A + cr;
        G = G + (py[a] + py[b]) * cr;
    }
    A = A * 0.5;
    return fabs(A) * (1.0 / (6.0 * A)) * G * PI_ * angle / 180.0;
}

/* Cell geometry of cell c of a mesh (quads or polygons), on node positions px, py. */
static inline int mesh_ncorner(const ora_mesh* m, int c) { return m->coff ? m->coff[c + 1] - m->coff[c] : 4; }
static inline const int* mesh_corners(const ora_mesh* m, int c) { return m->contour + (m->coff ? m->coff[c] : 4 * c); }

static void mesh_center(const ora_mesh* m, const double* px, const double* py, int c, double* cx, double* cy)
{   /* Cell::Get_Center (Cell.cpp:66-88) over the cell's contour */
    const int n = mesh_ncorner(m, c);
    const int* q = mesh_corners(m, c);
    double A = 0.0, Gx = 0.0, Gy = 0.0;
    for (int k = 0; k < n; k++) {
        const int a = q[k], b = q[k + 1 == n ? 0 : k + 1];
        const double cr = px[a] * py[b] - py[a] * px[b];
        A = A + cr;
        Gx = Gx + (px[a] + px[b]) * cr;
        Gy = Gy + (py[a] + py[b]) * cr;
    }
    A = A * 0.5;
    *cx = Gx / (6.0 * A);
    *cy = Gy / (6.0 * A);
}

/* Cell::Get_Volume (posle = 0, Cell.cpp:146-186) and Get_Volume_posle (posle = 1,
 * :208-254): quads by their diagonals; other polygons as the sum of Heron areas of the
 * triangles (face, centre), the centre being the barycentre before the step and the
 * mean of the corners after it. */
static double mesh_area(const ora_mesh* m, const double* px, const double* py, int c, int posle)
{
    const int n = mesh_ncorner(m, c);
    const int* q = mesh_corners(m, c);
    if (n == 4) return ora_cell_area(px, py, q);
    double x = 0.0, y = 0.0;
    if (posle) {
        for (int k = 0; k < n; k++) { x += px[q[k]]; y += py[q[k]]; }
        x = x / (1.0 * n);
        y = y / (1.0 * n);
    } else {
        mesh_center(m, px, py, c, &x, &y);
    }
    double S = 0.0;
    for (int k = 0; k < m->nface[c]; k++) {
        const int s = c * ORA_MAXF + k, A = m->fa[s], B = m->fb[s];
        const double a1 = sqrt(sq(px[A] - x) + sq(py[A] - y));
        const double a2 = sqrt(sq(px[B] - x) + sq(py[B] - y));
        const double a3 = sqrt(sq(px[A] - px[B]) + sq(py[A] - py[B]));
        const double p = (a1 + a2 + a3) / 2.0;
        S = S + sqrt(p * (p - a1) * (p - a2) * (p - a3));
    }
    return S;
}

static double mesh_volume_rotate(const ora_mesh* m, const double* px, const double* py, int c, double angle)
{   /* Cell::Get_Volume_rotate (Cell.cpp:188-206) over the cell's contour */
    const int n = mesh_ncorner(m, c);
    const int* q = mesh_corners(m, c);
    double A = 0.0, G = 0.0;
    for (int k = 0; k < n; k++) {
        const int a = q[k], b = q[k + 1 == n ? 0 : k + 1];
        const double cr = px[a] * py[b] - py[a] * px[b];
        A = A + cr;
        G = G + (py[a] + py[b]) * cr;
    }
    A = A * 0.5;
    return fabs(A) * (1.0 / (6.0 * A)) * G * PI_ * angle / 180.0;
}

/* ---------------------------------------------------------------- face flux --- */

double ora_hllc2d(const ora_gas* L, const ora_gas* R, double W, double n1, double n2,
                  double rad, double cfl, double gam, int metod, int nul_potok,
                  double P[4], double* PQ, double* Vl, double* Vc, double* Vp)
{
    /* Setka.cpp:10496-10503: normal/tangential components, tangent t = (n2, -n1) */
    const double t1 = n2, t2 = -n1;
    const double u1 = L->u * n1 + L->v * n2, v1 = L->u * t1 + L->v * t2;
    const double u2 = R->u * n1 + R->v * n2, v2 = R->u * t1 + R->v * t2;
    const double g1 = gam - 1.0;

    const double cL = sqrt(gam * L->p / L->ro);   /* :10507-10508 */
    const double cR = sqrt(gam * R->p / R->ro);
    const double kL = (sq(L->u) + sq(L->v)) / 2.0;
    const double kR = (sq(R->u) + sq(R->v)) / 2.0;

    const double SL = mn(u1, u2) - mx(cL, cR);    /* Davis speeds, :10514-10515 */
    const double SR = mx(u1, u2) + mx(cL, cR);
    *Vl = SL;
    *Vp = SR;
    const double suR = SR - u2, suL = SL - u1;
    const double SM = (suR * R->ro * u2 - suL * L->ro * u1 - R->p + L->p)
                      / (suR * R->ro - suL * L->ro);          /* :10526-10527 */
    *Vc = SM;
    const double time = cfl * rad / mx(fabs(SL), fabs(SR)); /* :10532-10536 */

    const double eL = L->p / g1 + L->ro * kL;
    const double eR = R->p / g1 + R->ro * kR;
    const double FL[5] = { L->ro * u1, L->ro * u1 * u1 + L->p, L->ro * u1 * v1, (eL + L->p) * u1, L->Q * u1 };
    const double FR[5] = { R->ro * u2, R->ro * u2 * u2 + R->p, R->ro * u2 * v2, (eR + R->p) * u2, R->Q * u2 };
    const double UL[5] = { L->ro, L->ro * u1, L->ro * v1, eL, L->Q };
    const double UR[5] = { R->ro, R->ro * u2, R->ro * v2, eR, R->Q };

    /* supersonic sides, flux minus W*U (ALE), :10569-10589 */
    if (SL >= W || SR <= W) {
        const double* F = (SL >= W) ? FL : FR;
        const double* U = (SL >= W) ? UL : UR;
        P[1] = n1 * (F[1] - W * U[1]) + t1 * (F[2] - W * U[2]);
        P[2] = n2 * (F[1] - W * U[1]) + t2 * (F[2] - W * U[2]);
        P[0] = F[0] - W * U[0];
        P[3] = F[3] - W * U[3];
        *PQ = F[4] - W * U[4];
        return time;
    }

    const double roLL = L->ro * (SL - u1) / (SL - SM);
    const double roRR = R->ro * (SR - u2) / (SR - SM);
    const double QLL = L->Q * (SL - u1) / (SL - SM);
    const double QRR = R->Q * (SR - u2) / (SR - SM);
    double UZ[5];                                    /* HLL average state, :10601-10605 */
    for (int k = 0; k < 5; k++) UZ[k] = (SR * UR[k] - SL * UL[k] + FL[k] - FR[k]) / (SR - SL);

    double PO[5];
    if (metod == 0) {                                 /* HLL, :10606-10618 */
        for (int k = 0; k < 5; k++)
            PO[k] = (SR * FL[k] - SL * FR[k] + SR * SL * (UR[k] - UL[k])) / (SR - SL);
        P[1] = n1 * (PO[1] - W * UZ[1]) + t1 * (PO[2] - W * UZ[2]);
        P[2] = n2 * (PO[1] - W * UZ[1]) + t2 * (PO[2] - W * UZ[2]);
        P[0] = PO[0] - W * UZ[0];
        P[3] = PO[3] - W * UZ[3];
        *PQ = PO[4] - W * UZ[4];
        return time;
    }

    /* HLLC star states, :10621-10672 */
    const double suRm = suR / (SR - SM), suLm = suL / (SL - SM);
    const double ptzR = R->p + R->ro * suR * (SM - u2);
    const double ptzL = L->p + L->ro * suL * (SM - u1);
    const double ptz = (ptzR + ptzL) / 2.0;
    double vLL, vRR;
    if (nul_potok) { vRR = UZ[2] / UZ[0]; vLL = vRR; }   /* :10644-10653 */
    else { vLL = v1; vRR = v2; }
    const double ee2 = eR * suRm + (ptz * SM - R->p * u2) / (SR - SM);
    const double ee1 = eL * suLm + (ptz * SM - L->p * u1) / (SL - SM);
    const double ULL[5] = { roLL, roLL * SM, roLL * vLL, ee1, QLL };
    const double URR[5] = { roRR, roRR * SM, roRR * vRR, ee2, QRR };
    if (SL < W && SM >= W) {
        for (int k = 0; k < 5; k++) PO[k] = FL[k] + SL * ULL[k] - SL * UL[k] - W * ULL[k];
    } else if (SR > W && SM < W) {
        for (int k = 0; k < 5; k++) PO[k] = FR[k] + SR * URR[k] - SR * UR[k] - W * URR[k];
    } else {
        for (int k = 0; k < 5; k++) PO[k] = NAN;    /* unreachable unless SM is NaN */
    }
    P[1] = n1 * PO[1] + t1 * PO[2];                  /* back rotation, :10693-10697 */
    P[2] = n2 * PO[1] + t2 * PO[2];
    P[0] = PO[0];
    P[3] = PO[3];
    *PQ = PO[4];
    return time;
}

/* ------------------------------------------------------------------ sources --- */

void ora_source_mk1(const gel_params* prm, const double* f, double u, double v,
                    double ro, double p, double* q1, double* q2, double* q3)
{   /* Cell::Get_Sourse_MK1, own-cell branch (Cell.cpp:590-595, :598-664);
       f is strided by the caller: f[k] is field k of this cell. */
    double Mu = 0.0, Mv = 0.0, MT = 0.0;
    double nu[4], uH[4], vH[4], roH[4], pH[4], UM[4], UU[4];
    for (int k = 0; k < 4; k++) {
        uH[k] = f[4 + k];
        vH[k] = f[8 + k];
        roH[k] = f[k];
        pH[k] = 0.5 * f[12 + k] * f[k];
        if (roH[k] <= 0.0) { roH[k] = 0.0000001; pH[k] = 0.0; }
    }
    for (int k = 0; k < 4; k++) {
        UM[k] = sqrt(sq(u - uH[k]) + sq(v - vH[k]) + (64.0 / (9.0 * PI_)) * (p / ro + 2.0 * pH[k] / roH[k]));
        UU[k] = sqrt(sq(u - uH[k]) + sq(v - vH[k]) + (4.0 / PI_) * (p / ro + 2.0 * pH[k] / roH[k]));
        const double sig = sq(prm->sig_a - prm->sig_b * log(UM[k]));
        nu[k] = ro * roH[k] * UM[k] * sig;
    }
    /* sums in the reference's left-to-right order, :645-659 */
    Mu = nu[0] * (uH[0] - u) + nu[1] * (uH[1] - u) + nu[2] * (uH[2] - u) + nu[3] * (uH[3] - u);
    Mv = nu[0] * (vH[0] - v) + nu[1] * (vH[1] - v) + nu[2] * (vH[2] - v) + nu[3] * (vH[3] - v);
    {
        double t[4];
        for (int k = 0; k < 4; k++)
            t[k] = nu[k] * ((sq(uH[k]) + sq(vH[k]) - sq(u) - sq(v)) / 2.0 +
                            (UU[k] / UM[k]) * (2.0 * pH[k] / roH[k] - p / ro));
        MT = t[0] + t[1] + t[2] + t[3];
    }
    Mu = prm->src_coef * Mu;
    Mv = prm->src_coef * Mv;
    MT = prm->src_coef * MT;
    *q1 = Mu * f[16];
    *q2 = Mv * f[17];
    *q3 = MT * f[18];
}

/* ------------------------------------------------------------------- sweep --- */

typedef struct sweep_ctx {
    const ora_mesh* m;
    const gel_params* prm;
    const double* now;
    int nc;
    const double* cx;   /* cell centres at the old node positions */
    const double* cy;
} sweep_ctx;

static inline ora_gas load(const sweep_ctx* S, int c)
{
    const int nc = S->nc;
    ora_gas g = { S->now[c], S->now[nc + c], S->now[2 * nc + c], S->now[3 * nc + c], S->now[4 * nc + c] };
    return g;
}

static int cell_has_axis(const ora_mesh* m, int c)
{
    for (int k = 0; k < m->nface[c]; k++)
        if (m->ftype[c * ORA_MAXF + k] == GEL_F_AXIS) return 1;
    return 0;
}

static inline double dist2d(double ax, double ay, double bx, double by)
{
    return sqrt(sq(ax - bx) + sq(ay - by));
}

/* Gran::Get_par (Gran.cpp:205-289): state on the far side of the face. */
static ora_gas get_par(const sweep_ctx* S, int M, int slot, double fx, double fy)
{
    const gel_params* p = S->prm;
    const int t = S->m->ftype[slot];
    ora_gas g = load(S, M);
    switch (t) {
    case GEL_F_USUALY:
        return load(S, S->m->fnbr[slot]);
    case GEL_F_EXTERN:
        if (g.u >= 0.0) g.u = p->extern_u;
        return g;
    case GEL_F_AXIS:
        g.v = -g.v;
        return g;
    case GEL_F_INNER_SPHERE: {
        const double d = sqrt(fx * fx + fy * fy);
        const double ro = 1.0 / (p->in_rho_div * d * d);
        ora_gas w = { ro, ro * p->in_p_a * p->in_p_a / p->in_p_den, p->in_u * fx / d,
                      p->in_u * fy / d, ro * p->in_q };
        return w;
    }
    case GEL_F_INPUT: {
        ora_gas w = { p->inf_rho, p->inf_p, p->inf_u, p->inf_v, p->inf_q };
        return w;
    }
    default: /* Upper_wall */
        return g;
    }
}

/* Gran::Get_par_TVD on a face whose two nodes lie on one fitted surface
 * (Gran.cpp:345-426): the straight line through down D and master M at the face centre,
 * non-positive density or pressure replaced by the master's. */
static ora_gas tvd_surface(const double* now, int nc, int M, int D, double xm, double ym, double xd, double yd,
                           double fx, double fy)
{
    const ora_gas pm = { now[M], now[nc + M], now[2 * nc + M], now[3 * nc + M], now[4 * nc + M] };
    const ora_gas pd = { now[D], now[nc + D], now[2 * nc + D], now[3 * nc + D], now[4 * nc + D] };
    const double d2 = sqrt(sq(fx - xm) + sq(fy - ym));
    const double d1 = sqrt(sq(fx - xd) + sq(fy - yd));
    ora_gas r;
    r.ro = ora_linear2(-d1, pd.ro, -d2, pm.ro, 0.0);
    r.p = ora_linear2(-d1, pd.p, -d2, pm.p, 0.0);
    r.u = ora_linear2(-d1, pd.u, -d2, pm.u, 0.0);
    r.v = ora_linear2(-d1, pd.v, -d2, pm.v, 0.0);
    r.Q = ora_linear2(-d1, pd.Q, -d2, pm.Q, 0.0);
    if (r.ro <= 0.0) r.ro = pm.ro;
    if (r.p <= 0.0) r.p = pm.p;
    return r;
}

/* Gran::Get_par_TVD for an Usualy face (Gran.cpp:294-457): value of master M at the
 * face centre (fx,fy) from the line (down D, M, sosed N); surf: the face's nodes lie on
 * one fitted surface.  *bad set where the reference exits. */
static ora_gas tvd_usualy(const sweep_ctx* S, int M, int N, int D, double fx, double fy, int surf, int* bad)
{
    ora_gas pm = load(S, M);
    if (D < 0) {
        if (cell_has_axis(S->m, M)) {          /* mirrored ghost across the axis, :297-339 */
            ora_gas pn = load(S, N);
            const double d2 = dist2d(fx, fy, S->cx[M], S->cy[M]);
            const double d3 = dist2d(fx, fy, S->cx[N], S->cy[N]);
            const double d1 = dist2d(fx, fy, S->cx[M], -S->cy[M]);
            ora_gas r = pm;
            r.v = ora_linear3(-d1, -pm.v, -d2, pm.v, d3, pn.v, 0.0);
            return r;
        }
        return pm;                             /* :340-343 */
    }
    if (surf) return tvd_surface(S->now, S->nc, M, D, S->cx[M], S->cy[M], S->cx[D], S->cy[D], fx, fy);
    ora_gas pd = load(S, D), pn = load(S, N);
    const double d2 = dist2d(fx, fy, S->cx[M], S->cy[M]);
    const double d3 = dist2d(fx, fy, S->cx[N], S->cy[N]);
    const double d1 = dist2d(fx, fy, S->cx[D], S->cy[D]);
    ora_gas r;                                  /* :427-456 */
    r.ro = ora_linear3(-d1, pd.ro, -d2, pm.ro, d3, pn.ro, 0.0);
    r.p = ora_linear3(-d1, pd.p, -d2, pm.p, d3, pn.p, 0.0);
    r.u = ora_linear3(-d1, pd.u, -d2, pm.u, d3, pn.u, 0.0);
    r.v = ora_linear3(-d1, pd.v, -d2, pm.v, d3, pn.v, 0.0);
    r.Q = ora_linear3(-d1, pd.Q, -d2, pm.Q, d3, pn.Q, 0.0);
    if (r.ro <= 0.0 || r.p <= 0.0) *bad = 1;
    return r;
}

/* Gran::Get_par_TVD_radial, density and pressure only (Gran.cpp:562-675): the line
 * fit is done on ro*r^2 and p*r^(2*gamma). */
static void tvd_radial(const sweep_ctx* S, int M, int N, int D, double fx, double fy,
                       double* ro, double* pr)
{
    const double g2 = 2.0 * S->prm->gamma;
    ora_gas pm = load(S, M), pn = load(S, N);
    double xd, yd;
    ora_gas pd;
    if (D < 0) {
        if (!cell_has_axis(S->m, M)) { *ro = pm.ro; *pr = pm.p; return; }
        xd = S->cx[M]; yd = -S->cy[M]; pd = pm;      /* mirrored master, :575-604 */
    } else {
        xd = S->cx[D]; yd = S->cy[D]; pd = load(S, D);
    }
    const double r = sqrt(sq(fx) + sq(fy));
    const double r2 = sqrt(sq(S->cx[M]) + sq(S->cy[M]));
    const double d2 = dist2d(fx, fy, S->cx[M], S->cy[M]);
    const double r3 = sqrt(sq(S->cx[N]) + sq(S->cy[N]));
    const double d3 = dist2d(fx, fy, S->cx[N], S->cy[N]);
    const double r1 = sqrt(sq(xd) + sq(yd));
    const double d1 = dist2d(fx, fy, xd, yd);
    *ro = ora_linear3(-d1, pd.ro * sq(r1), -d2, pm.ro * sq(r2), d3, pn.ro * sq(r3), 0.0) / sq(r);
    *pr = ora_linear3(-d1, pd.p * pow(r1, g2), -d2, pm.p * pow(r2, g2), d3, pn.p * pow(r3, g2), 0.0) / pow(r, g2);
}

/* Gran::Get_par_TVD for an Axis face (Gran.cpp:486-519). */
static ora_gas tvd_axis(const sweep_ctx* S, int M, int D, double fx, double fy)
{
    ora_gas pm = load(S, M), pd = load(S, D);
    const double d2 = dist2d(fx, fy, S->cx[M], S->cy[M]);
    const double d1 = dist2d(fx, fy, S->cx[D], S->cy[D]);
    ora_gas r = pm;
    r.v = ora_linear3(-d1, pd.v, -d2, pm.v, d2, -pm.v, 0.0);
    return r;
}

/* radial rescaling of a C_1 cell value from its centre radius rc to the face
 * radius dis, and rotation of the velocity to the face's polar frame
 * (Setka.cpp:10221-10226). */
static void rescale_c1(ora_gas* g, double rc, double dis, double gam, double cx, double cy,
                       double fx, double fy, int rotate)
{
    g->ro = g->ro * sq(rc) / sq(dis);
    g->Q = g->Q * sq(rc) / sq(dis);
    g->p = g->p * pow(rc / dis, 2 * gam);
    if (rotate) ora_polar_perenos(cx, cy, fx, fy, &g->u, &g->v);
}

int ora_step_mk2(const ora_mesh* m, const gel_params* prm, const double* now,
                 double* next, const double* src, double dt_prev, double* dt_new,
                 long long* n_rho_clamp, long long* n_p_clamp, int nthreads)
{
    const int nc = m->n_cells;
    const double gam = prm->gamma;
    double* cx = (double*)malloc(sizeof(double) * (size_t)nc);
    double* cy = (double*)malloc(sizeof(double) * (size_t)nc);
    if (!cx || !cy) { free(cx); free(cy); return GEL_E_NOMEM; }
    for (int c = 0; c < nc; c++) mesh_center(m, m->x, m->y, c, &cx[c], &cy[c]);
    sweep_ctx S = { m, prm, now, nc, cx, cy };

    double Tmin = 100000000;          /* T[now2], Setka.cpp:10142 */
    long long nrc = 0, npc = 0;
    int bad_any = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(min:Tmin) reduction(+:nrc,npc) reduction(|:bad_any)
#endif
    for (int c = 0; c < nc; c++) {
        const double x = cx[c], y = cy[c];
        const double vs1 = mesh_area(m, m->x, m->y, c, 0);
        const double vs2 = mesh_area(m, m->x2, m->y2, c, 1);
        const double radius = sqrt(sq(x) + sq(y));
        if (radius < prm->r_skip) continue;
        const double Volume = mesh_volume_rotate(m, m->x, m->y, c, 360.0) / (2.0 * PI_);
        const double Volume2 = mesh_volume_rotate(m, m->x2, m->y2, c, 360.0) / (2.0 * PI_);
        const ora_gas par1 = load(&S, c);
        const int K1 = (m->cell_type[c] == GEL_C_1);
        double Pot[5] = { 0, 0, 0, 0, 0 };
        double time = 10000000;
        int bad = 0;

        for (int k = 0; k < m->nface[c]; k++) {
            const int s = c * ORA_MAXF + k;
            const int A = m->fa[s], B = m->fb[s], t = m->ftype[s];
            /* Gran geometry at the old and new node positions (Gran.cpp:159-198) */
            const double S_ = sqrt(sq(m->x[B] - m->x[A]) + sq(m->y[B] - m->y[A])) * (m->y[A] + m->y[B]) * 0.5;
            const double fx = (m->x[A] + m->x[B]) / 2.0, fy = (m->y[A] + m->y[B]) / 2.0;
            const double fx2 = (m->x2[A] + m->x2[B]) / 2.0, fy2 = (m->y2[A] + m->y2[B]) / 2.0;
            const double dist = sqrt(sq(x - fx) + sq(y - fy));
            double n1, n2;
            {
                double t1 = m->x[B] - m->x[A], t2 = m->y[B] - m->y[A];
                const double nn = sqrt(sq(t1) + sq(t2));
                t1 = t1 / nn; t2 = t2 / nn;
                n1 = t2; n2 = -t1;
            }
            const double W = ((fx2 - fx) * n1 + (fy2 - fy) * n2) / dt_prev;   /* :10209 */
            const double dis = sqrt(sq(fx) + sq(fy));
            ora_gas p11 = par1;
            ora_gas p2 = get_par(&S, c, s, fx, fy);
            /* fitted surfaces (Setka.cpp:10287-10327): both nodes on one surface -> exact
               solver; exactly one node on a surface (or nodes on two) -> nul_potok */
            const int pa = m->ptype ? m->ptype[A] : 0, pb = m->ptype ? m->ptype[B] : 0;
            const int god = (pa != 0 && pa == pb);
            const int np = (pa != pb);

            if (t == GEL_F_USUALY) {                    /* :10213-10251 */
                const int N = m->fnbr[s];
                const int N1 = (m->cell_type[N] == GEL_C_1);
                if (K1) {
                    rescale_c1(&p11, radius, dis, gam, x, y, fx, fy, 1);
                    double ro, pr;
                    tvd_radial(&S, c, N, m->fdown[s], fx, fy, &ro, &pr);
                    p11.ro = ro; p11.p = pr;
                } else {
                    p11 = tvd_usualy(&S, c, N, m->fdown[s], fx, fy, god, &bad);
                }
                if (N1) {
                    const double dd = sqrt(sq(cx[N]) + sq(cy[N]));
                    rescale_c1(&p2, dd, dis, gam, cx[N], cy[N], fx, fy, 1);
                    double ro, pr;
                    tvd_radial(&S, N, c, m->fup[s], fx, fy, &ro, &pr);
                    p2.ro = ro; p2.p = pr;
                } else {
                    p2 = tvd_usualy(&S, N, c, m->fup[s], fx, fy, god, &bad);
                }
            } else if (t == GEL_F_INNER_SPHERE) {      /* :10252-10257 */
                rescale_c1(&p11, radius, dis, gam, x, y, fx, fy, 0);
            } else if (K1 && t == GEL_F_AXIS) {        /* :10258-10268 */
                rescale_c1(&p11, radius, dis, gam, x, y, fx, fy, 1);
                p2 = p11;
                p2.v = -p2.v;
            } else if (t == GEL_F_AXIS) {              /* :10269-10274 */
                p11 = tvd_axis(&S, c, m->fdown[s], fx, fy);
                p2 = p11;
                p2.v = -p2.v;
            }

            double P[4], PQ, Vl, Vc, Vp;
            if (god) {                                  /* :10335-10370 */
                const double q1[5] = { p11.ro, p11.p, p11.u, p11.v, 0.0 };
                const double q2[5] = { p2.ro, p2.p, p2.u, p2.v, 0.0 };
                const double nv[3] = { n1, n2, 0.0 };
                double qqq[5], dl, dp, dc;
                int gst = 0;
                time = mn(time, ora_godunov(q1, q2, nv, W, dist, prm->cfl, gam, qqq, &dl, &dp, &dc, &gst));
                if (gst) bad = 1;
                Pot[0] = Pot[0] + qqq[0] * S_;
                Pot[1] = Pot[1] + qqq[1] * S_;
                Pot[2] = Pot[2] + qqq[2] * S_;
                Pot[3] = Pot[3] + qqq[4] * S_;
                time = mn(time, ora_hllc2d(&p11, &p2, W, n1, n2, dist, prm->cfl, gam, prm->flux, np,
                                           P, &PQ, &Vl, &Vc, &Vp));
                Pot[4] = Pot[4] + PQ * S_;
                continue;
            }
            const double tf = ora_hllc2d(&p11, &p2, W, n1, n2, dist, prm->cfl, gam, prm->flux, np,
                                         P, &PQ, &Vl, &Vc, &Vp);
            time = mn(time, tf);
            for (int v = 0; v < 4; v++) Pot[v] = Pot[v] + P[v] * S_;   /* :10372-10376 */
            Pot[4] = Pot[4] + PQ * S_;
        }
        if (bad) bad_any = 1;
        if (time < Tmin) Tmin = time;

        const double u = par1.u, v = par1.v, ro = par1.ro, p = par1.p, Q = par1.Q;
        double q1 = 0.0, q2 = 0.0, q3 = 0.0;
        if (prm->source_mode == GEL_SRC_MK1 && src) {
            double f[GEL_MK1_NFIELDS];
            for (int k = 0; k < GEL_MK1_NFIELDS; k++) f[k] = src[(size_t)k * nc + c];
            ora_source_mk1(prm, f, u, v, ro, p, &q1, &q2, &q3);
        }
        const double T1 = dt_prev;
        if (m->cell_type[c] != GEL_C_CENTR) {           /* :10423-10460 */
            double ro3 = ro * Volume / Volume2 - T1 * (Pot[0] / Volume2);
            const double Q33 = Q * Volume / Volume2 - (T1 / Volume2) * Pot[4];
            if (ro3 <= 0.0) { ro3 = 0.00001; nrc++; }
            const double u3 = (ro * u * Volume / Volume2 - T1 * (Pot[1] / Volume2 - q1)) / ro3;
            const double v3 = (ro * v * Volume / Volume2 - T1 * ((Pot[2] - p * 0.5 * (vs1 + vs2)) / Volume2 - q2)) / ro3;
            double p3 = (((p / (gam - 1) + ro * (u * u + v * v) * 0.5) * Volume / Volume2 - T1 * (Pot[3] / Volume2 - q3)) -
                         0.5 * ro3 * (u3 * u3 + v3 * v3)) * (gam - 1);
            if (p3 <= 0) { p3 = 0.000001; npc++; }
            next[c] = ro3;
            next[nc + c] = p3;
            next[2 * nc + c] = u3;
            next[3 * nc + c] = v3;
            next[4 * nc + c] = Q33;
        } else {
            for (int k = 0; k < 5; k++) next[(size_t)k * nc + c] = now[(size_t)k * nc + c];
        }
    }
    free(cx);
    free(cy);
    *dt_new = Tmin;
    if (n_rho_clamp) *n_rho_clamp += nrc;
    if (n_p_clamp) *n_p_clamp += npc;
    return bad_any ? GEL_E_NUMERIC : GEL_OK;
}

/* ------------------------------------------------------------------ MHD HLLD --- */

static double sgn_ref(double x) { return sgn(x); }

/* HLLD flux of Solvers::HLLDQ_Korolkov, metod 3 (Solvers.cpp:1642-1740 set-up,
 * :1980-2185 HLLD), extended to any normal n = (n1, n2, 0) in the meridional plane and
 * to a face moving with speed W along n: the fan is sampled at xi = W and every
 * region's flux is taken minus W*U, as the gas flux does (Setka.cpp:10569-10618).
 * Basis: t = (-n2, n1, 0), m = (0, 0, 1) (the reference's choice for n = (1,0,0),
 * Solvers.cpp:1664-1668).  B is Gaussian; b = B/sqrt(4 pi) inside.
 * P: [0] mass, [1..3] momentum x,y,phi, [4] energy, [5..7] B x,y,phi, [8] tracer.
 * Returns max(|SL|,|SR|). */
double ora_hlld(const ora_mhd* L, const ora_mhd* R, double W, double n1, double n2, double gam, double P[9])
{
    const double spi4 = sqrt(4 * 3.14159265358979323846);   /* Solvers.h:9-12 */
    const double g1 = gam - 1.0;
    const double n3 = 0.0;
    const double t1 = -n2, t2 = n1, t3 = 0.0;
    const double m1 = 0.0, m2 = 0.0, m3 = 1.0;
    const double bx_L = L->bx / spi4, by_L = L->by / spi4, bz_L = L->bz / spi4;
    const double bx_R = R->bx / spi4, by_R = R->by / spi4, bz_R = R->bz / spi4;

    const double u1 = L->u * n1 + L->v * n2 + L->w * n3;
    const double v1 = L->u * t1 + L->v * t2 + L->w * t3;
    const double w1 = L->u * m1 + L->v * m2 + L->w * m3;
    const double u2 = R->u * n1 + R->v * n2 + R->w * n3;
    const double v2 = R->u * t1 + R->v * t2 + R->w * t3;
    const double w2 = R->u * m1 + R->v * m2 + R->w * m3;
    const double bn1 = bx_L * n1 + by_L * n2 + bz_L * n3;
    const double bt1 = bx_L * t1 + by_L * t2 + bz_L * t3;
    const double bm1 = bx_L * m1 + by_L * m2 + bz_L * m3;
    const double bn2 = bx_R * n1 + by_R * n2 + bz_R * n3;
    const double bt2 = bx_R * t1 + by_R * t2 + bz_R * t3;
    const double bm2 = bx_R * m1 + by_R * m2 + bz_R * m3;

    const double ro_L = L->ro, ro_R = R->ro, p_L = L->p, p_R = R->p, Q_L = L->Q, Q_R = R->Q;
    const double bb_L = sq(bx_L) + sq(by_L) + sq(bz_L);
    const double bb_R = sq(bx_R) + sq(by_R) + sq(bz_R);
    const double uu_L = (sq(L->u) + sq(L->v) + sq(L->w)) / 2.0;
    const double uu_R = (sq(R->u) + sq(R->v) + sq(R->w)) / 2.0;
    /* fast magnetosonic speeds, :1734-1737 */
    const double cfL = sqrt((gam * p_L + bb_L + sqrt(sq(gam * p_L + bb_L) - 4.0 * gam * p_L * sq(bn1))) / (2.0 * ro_L));
    const double cfR = sqrt((gam * p_R + bb_R + sqrt(sq(gam * p_R + bb_R) - 4.0 * gam * p_R * sq(bn2))) / (2.0 * ro_R));
    const double SL = mn(u1, u2) - mx(cfL, cfR);
    const double SR = mx(u1, u2) + mx(cfL, cfR);
    const double pTL = p_L + bb_L / 2.0;
    const double pTR = p_R + bb_R / 2.0;
    const double suR = SR - u2, suL = SL - u1;
    const double SM = (suR * ro_R * u2 - suL * ro_L * u1 - pTR + pTL) / (suR * ro_R - suL * ro_L);
    const double PTT = (suR * ro_R * pTL - suL * ro_L * pTR + ro_L * ro_R * suR * suL * (u2 - u1)) /
                       (suR * ro_R - suL * ro_L);
    const double UU = mx(fabs(SL), fabs(SR));

    const double e1 = p_L / g1 + ro_L * uu_L + bb_L / 2.0;
    const double e2 = p_R / g1 + ro_R * uu_R + bb_R / 2.0;
    double FL[9], FR[9], UL[9], UR[9];
    FL[0] = ro_L * u1;
    FL[1] = ro_L * u1 * u1 + pTL - sq(bn1);
    FL[2] = ro_L * u1 * v1 - bn1 * bt1;
    FL[3] = ro_L * u1 * w1 - bn1 * bm1;
    FL[4] = (e1 + pTL) * u1 - bn1 * (u1 * bn1 + v1 * bt1 + w1 * bm1);
    FL[5] = 0.0;
    FL[6] = u1 * bt1 - v1 * bn1;
    FL[7] = u1 * bm1 - w1 * bn1;
    FL[8] = Q_L * u1;
    FR[0] = ro_R * u2;
    FR[1] = ro_R * u2 * u2 + pTR - sq(bn2);
    FR[2] = ro_R * u2 * v2 - bn2 * bt2;
    FR[3] = ro_R * u2 * w2 - bn2 * bm2;
    FR[4] = (e2 + pTR) * u2 - bn2 * (u2 * bn2 + v2 * bt2 + w2 * bm2);
    FR[5] = 0.0;
    FR[6] = u2 * bt2 - v2 * bn2;
    FR[7] = u2 * bm2 - w2 * bn2;
    FR[8] = Q_R * u2;
    UL[0] = ro_L; UL[1] = ro_L * u1; UL[2] = ro_L * v1; UL[3] = ro_L * w1; UL[4] = e1;
    UL[5] = bn1; UL[6] = bt1; UL[7] = bm1; UL[8] = Q_L;
    UR[0] = ro_R; UR[1] = ro_R * u2; UR[2] = ro_R * v2; UR[3] = ro_R * w2; UR[4] = e2;
    UR[5] = bn2; UR[6] = bt2; UR[7] = bm2; UR[8] = Q_R;

    const double bn = (SR * UR[5] - SL * UL[5] + FL[5] - FR[5]) / (SR - SL);
    const double bbn = bn * bn;
    const double ro_LL = ro_L * (SL - u1) / (SL - SM);
    const double ro_RR = ro_R * (SR - u2) / (SR - SM);
    const double Q_LL = Q_L * (SL - u1) / (SL - SM);
    const double Q_RR = Q_R * (SR - u2) / (SR - SM);

    /* :1983-2017 */
    const double ttL = ro_L * suL * (SL - SM) - bbn;
    const double ttR = ro_R * suR * (SR - SM) - bbn;
    double vLL, wLL, vRR, wRR, btt1, bmm1, btt2, bmm2;
    if (fabs(ttL) >= 0.000001) {
        vLL = v1 - bn * bt1 * (SM - u1) / ttL;
        wLL = w1 - bn * bm1 * (SM - u1) / ttL;
        btt1 = bt1 * (ro_L * suL * suL - bbn) / ttL;
        bmm1 = bm1 * (ro_L * suL * suL - bbn) / ttL;
    } else {
        vLL = v1; wLL = w1; btt1 = 0.0; bmm1 = 0.0;
    }
    if (fabs(ttR) >= 0.000001) {
        vRR = v2 - bn * bt2 * (SM - u2) / ttR;
        wRR = w2 - bn * bm2 * (SM - u2) / ttR;
        btt2 = bt2 * (ro_R * suR * suR - bbn) / ttR;
        bmm2 = bm2 * (ro_R * suR * suR - bbn) / ttR;
    } else {
        vRR = v2; wRR = w2; btt2 = 0.0; bmm2 = 0.0;
    }
    /* :2019-2024 */
    const double eLL = (e1 * suL + PTT * SM - pTL * u1 + bn * ((u1 * bn1 + v1 * bt1 + w1 * bm1) - (SM * bn + vLL * btt1 + wLL * bmm1))) / (SL - SM);
    const double eRR = (e2 * suR + PTT * SM - pTR * u2 + bn * ((u2 * bn2 + v2 * bt2 + w2 * bm2) - (SM * bn + vRR * btt2 + wRR * bmm2))) / (SR - SM);
    const double sqrtroLL = sqrt(ro_LL), sqrtroRR = sqrt(ro_RR);
    double SLL = SM - fabs(bn) / sqrtroLL;
    double SRR = SM + fabs(bn) / sqrtroRR;
    double idbn;
    if (fabs(bn) > 0.001) {
        idbn = 1.0 * sgn_ref(bn);
    } else {
        idbn = 0.0; SLL = SM; SRR = SM;
    }
    const double vLLL = (sqrtroLL * vLL + sqrtroRR * vRR + idbn * (btt2 - btt1)) / (sqrtroLL + sqrtroRR);
    const double wLLL = (sqrtroLL * wLL + sqrtroRR * wRR + idbn * (bmm2 - bmm1)) / (sqrtroLL + sqrtroRR);
    const double bttt = (sqrtroLL * btt2 + sqrtroRR * btt1 + idbn * sqrtroLL * sqrtroRR * (vRR - vLL)) / (sqrtroLL + sqrtroRR);
    const double bmmm = (sqrtroLL * bmm2 + sqrtroRR * bmm1 + idbn * sqrtroLL * sqrtroRR * (wRR - wLL)) / (sqrtroLL + sqrtroRR);
    const double eLLL = eLL - idbn * sqrtroLL * ((SM * bn + vLL * btt1 + wLL * bmm1) - (SM * bn + vLLL * bttt + wLLL * bmmm));
    const double eRRR = eRR + idbn * sqrtroRR * ((SM * bn + vRR * btt2 + wRR * bmm2) - (SM * bn + vLLL * bttt + wLLL * bmmm));
    const double ULL[9] = { ro_LL, ro_LL * SM, ro_LL * vLL, ro_LL * wLL, eLL, bn, btt1, bmm1, Q_LL };
    const double URR[9] = { ro_RR, ro_RR * SM, ro_RR * vRR, ro_RR * wRR, eRR, bn, btt2, bmm2, Q_RR };
    const double ULLL[9] = { ro_LL, ro_LL * SM, ro_LL * vLLL, ro_LL * wLLL, eLLL, bn, bttt, bmmm, Q_LL };
    const double URRR[9] = { ro_RR, ro_RR * SM, ro_RR * vLLL, ro_RR * wLLL, eRRR, bn, bttt, bmmm, Q_RR };

    double PO[9];
    double bnW = 0.0;   /* normal field of the sampled region, for the -W*U term */
    /* regions of the fan at xi = W (:2106-2156).  The reference tests L, L*, L**, R**,
     * R*, R in that order.  Where the fan is not monotone (the Alfven speeds use the
     * HLL normal field bn, which the fast speeds of the two sides need not bound, so
     * that SLL < SL or SRR > SR), that order takes a star region on one side of the
     * face and the upwind state on the other: the mirrored problem F(R, L, -n, -W),
     * which a cell loop solving every face from both sides evaluates, then differs from
     * -F(L, R, n, W) at O(1).  Both upwind regions are tested first here, then the side
     * of the contact, so that every branch is mirror-symmetric; where the fan is
     * monotone the region is the reference's. */
    int reg;
    if (SL >= W) reg = 0;
    else if (SR <= W) reg = 5;
    else if (SM >= W) reg = (SLL >= W) ? 1 : 2;
    else reg = (SRR <= W) ? 4 : 3;
    if (reg == 0) {
        for (int i = 0; i < 9; i++) PO[i] = FL[i] - W * UL[i];
        bnW = UL[5];
    } else if (reg == 1) {
        for (int i = 0; i < 9; i++) PO[i] = FL[i] + SL * ULL[i] - SL * UL[i] - W * ULL[i];
        bnW = bn;
    } else if (reg == 2) {
        for (int i = 0; i < 9; i++) PO[i] = FL[i] + SLL * ULLL[i] - (SLL - SL) * ULL[i] - SL * UL[i] - W * ULLL[i];
        bnW = bn;
    } else if (reg == 3) {
        for (int i = 0; i < 9; i++) PO[i] = FR[i] + SRR * URRR[i] - (SRR - SR) * URR[i] - SR * UR[i] - W * URRR[i];
        bnW = bn;
    } else if (reg == 4) {
        for (int i = 0; i < 9; i++) PO[i] = FR[i] + SR * URR[i] - SR * UR[i] - W * URR[i];
        bnW = bn;
    } else {
        for (int i = 0; i < 9; i++) PO[i] = FR[i] - W * UR[i];
        bnW = UR[5];
    }
    /* normal-field flux replaced by the jump times the fastest speed, :2161-2163 */
    const double SN = mx(fabs(SL), fabs(SR));
    PO[5] = -SN * (bn2 - bn1) - W * bnW;

    P[0] = PO[0];
    P[1] = n1 * PO[1] + t1 * PO[2] + m1 * PO[3];
    P[2] = n2 * PO[1] + t2 * PO[2] + m2 * PO[3];
    P[3] = n3 * PO[1] + t3 * PO[2] + m3 * PO[3];
    P[4] = PO[4];
    P[5] = spi4 * (n1 * PO[5] + t1 * PO[6] + m1 * PO[7]);
    P[6] = spi4 * (n2 * PO[5] + t2 * PO[6] + m2 * PO[7]);
    P[7] = spi4 * (n3 * PO[5] + t3 * PO[6] + m3 * PO[7]);
    P[8] = PO[8];
    return UU;
}

/* ------------------------------------------------------------------ MHD sweep --- */
/* The MHD extension of the cell loop of Go_5_komponent__MK2 (Setka.cpp:10148-10470):
 * same gather over Cell::Grans, same boundary states and minmod reconstruction
 * (extended to the 9 primitives), HLLD faces, and the r-weighted ALE update with the
 * geometric sources of axisymmetric MHD with swirl, per unit planar area:
 *   r-momentum   + (ro w^2 + p + B^2/8pi - Bphi^2/4pi)   (hoop stress; gas: + p)
 *   phi-momentum - (ro v w - By Bphi/4pi)
 *   Bphi         + (v Bphi - w By)
 * Variables: 0 ro, 1 p, 2 u, 3 v, 4 w, 5 bx, 6 by, 7 bz (= Bphi), 8 Q. */

#define MV 9
static const int mirror9[MV] = { 0, 0, 0, 1, 1, 0, 1, 1, 0 };   /* odd across the axis */

typedef struct msw { const ora_mesh* m; const gel_params* prm; const double* now; int nc;
                     const double* cx; const double* cy; } msw;

static void mload(const msw* S, int c, double* q) { for (int k = 0; k < MV; k++) q[k] = S->now[(size_t)k * S->nc + c]; }

static void mget_par(const msw* S, int M, int slot, double fx, double fy, double* g)
{
    const gel_params* p = S->prm;
    const int t = S->m->ftype[slot];
    mload(S, M, g);
    if (t == GEL_F_USUALY) { mload(S, S->m->fnbr[slot], g); return; }
    if (t == GEL_F_EXTERN) { if (g[2] >= 0.0) g[2] = p->extern_u; return; }
    if (t == GEL_F_AXIS) { for (int k = 0; k < MV; k++) if (mirror9[k]) g[k] = -g[k]; return; }
    if (t == GEL_F_INNER_SPHERE) {
        const double d = sqrt(fx * fx + fy * fy);
        const double ro = 1.0 / (p->in_rho_div * d * d);
        g[0] = ro; g[1] = ro * p->in_p_a * p->in_p_a / p->in_p_den;
        g[2] = p->in_u * fx / d; g[3] = p->in_u * fy / d; g[4] = 0.0;
        g[5] = 0.0; g[6] = 0.0; g[7] = 0.0; g[8] = ro * p->in_q;
        return;
    }
    if (t == GEL_F_INPUT) {
        g[0] = p->inf_rho; g[1] = p->inf_p; g[2] = p->inf_u; g[3] = p->inf_v; g[4] = p->inf_w;
        g[5] = p->inf_bx; g[6] = p->inf_by; g[7] = p->inf_bz; g[8] = p->inf_q;
    }
}

static void mtvd_usualy(const msw* S, int M, int N, int D, double fx, double fy, double* r, int* bad)
{
    double pm[MV], pn[MV];
    mload(S, M, pm);
    mload(S, N, pn);
    if (D < 0) {
        for (int k = 0; k < MV; k++) r[k] = pm[k];
        if (cell_has_axis(S->m, M)) {
            const double d2 = dist2d(fx, fy, S->cx[M], S->cy[M]);
            const double d3 = dist2d(fx, fy, S->cx[N], S->cy[N]);
            const double d1 = dist2d(fx, fy, S->cx[M], -S->cy[M]);
            for (int k = 0; k < MV; k++)
                if (mirror9[k]) r[k] = ora_linear3(-d1, -pm[k], -d2, pm[k], d3, pn[k], 0.0);
        }
        return;
    }
    double pd[MV];
    mload(S, D, pd);
    const double d2 = dist2d(fx, fy, S->cx[M], S->cy[M]);
    const double d3 = dist2d(fx, fy, S->cx[N], S->cy[N]);
    const double d1 = dist2d(fx, fy, S->cx[D], S->cy[D]);
    for (int k = 0; k < MV; k++) r[k] = ora_linear3(-d1, pd[k], -d2, pm[k], d3, pn[k], 0.0);
    if (r[0] <= 0.0 || r[1] <= 0.0) *bad = 1;
}

static void mtvd_axis(const msw* S, int M, int D, double fx, double fy, double* r)
{
    double pm[MV], pd[MV];
    mload(S, M, pm);
    mload(S, D, pd);
    const double d2 = dist2d(fx, fy, S->cx[M], S->cy[M]);
    const double d1 = dist2d(fx, fy, S->cx[D], S->cy[D]);
    for (int k = 0; k < MV; k++) r[k] = pm[k];
    for (int k = 0; k < MV; k++)
        if (mirror9[k]) r[k] = ora_linear3(-d1, pd[k], -d2, pm[k], d2, -pm[k], 0.0);
}

static ora_mhd as_mhd(const double* q)
{
    ora_mhd s = { q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8] };
    return s;
}

/* Face-once summation (ora_set_face_once): each interior face is solved once, by the
 * cell of lower index, and the other cell takes the negated flux x area product, as
 * the device sweep sums them.  The reference solves every face from both sides
 * (Setka.cpp:10195); this mode exists to measure what that difference does. */
static int g_face_once = 0;
void ora_set_face_once(int on) { g_face_once = on != 0; }

/* One face slot of cell c as the cell loop of the MHD sweep evaluates it: the flux x
 * area products P[9]*S and the max wave speed. */
static void mhd_face(const msw* S, int c, int s, double dt_prev, double* PS, double* UUo, int* bad, double* LRW)
{
    const ora_mesh* m = S->m;
    const double gam = S->prm->gamma;
    const double x = S->cx[c], y = S->cy[c];
    const double radius = sqrt(sq(x) + sq(y));
    const int A = m->fa[s], B = m->fb[s], t = m->ftype[s];
    const double S_ = sqrt(sq(m->x[B] - m->x[A]) + sq(m->y[B] - m->y[A])) * (m->y[A] + m->y[B]) * 0.5;
    const double fx = (m->x[A] + m->x[B]) / 2.0, fy = (m->y[A] + m->y[B]) / 2.0;
    const double fx2 = (m->x2[A] + m->x2[B]) / 2.0, fy2 = (m->y2[A] + m->y2[B]) / 2.0;
    double n1, n2;
    {
        double t1 = m->x[B] - m->x[A], t2 = m->y[B] - m->y[A];
        const double nn = sqrt(sq(t1) + sq(t2));
        t1 = t1 / nn; t2 = t2 / nn;
        n1 = t2; n2 = -t1;
    }
    const double W = ((fx2 - fx) * n1 + (fy2 - fy) * n2) / dt_prev;
    const double dis = sqrt(sq(fx) + sq(fy));
    double p11[MV], p2[MV];
    mload(S, c, p11);
    mget_par(S, c, s, fx, fy, p2);
    if (t == GEL_F_USUALY) {
        const int N = m->fnbr[s];
        mtvd_usualy(S, c, N, m->fdown[s], fx, fy, p11, bad);
        mtvd_usualy(S, N, c, m->fup[s], fx, fy, p2, bad);
    } else if (t == GEL_F_INNER_SPHERE) {
        p11[0] = p11[0] * sq(radius) / sq(dis);
        p11[8] = p11[8] * sq(radius) / sq(dis);
        p11[1] = p11[1] * pow(radius / dis, 2 * gam);
    } else if (t == GEL_F_AXIS) {
        mtvd_axis(S, c, m->fdown[s], fx, fy, p11);
        for (int v = 0; v < MV; v++) p2[v] = mirror9[v] ? -p11[v] : p11[v];
    }
    const ora_mhd Lm = as_mhd(p11), Rm = as_mhd(p2);
    if (LRW) {
        for (int v = 0; v < MV; v++) { LRW[v] = p11[v]; LRW[MV + v] = p2[v]; }
        LRW[2 * MV] = W; LRW[2 * MV + 1] = n1; LRW[2 * MV + 2] = n2; LRW[2 * MV + 3] = S_;
    }
    double P[MV];
    *UUo = ora_hlld(&Lm, &Rm, W, n1, n2, gam, P);
    for (int v = 0; v < MV; v++) PS[v] = P[v] * S_;
}

int ora_mhd_face(const ora_mesh* m, const gel_params* prm, const double* now, int c, int k, double dt_prev,
                 double PS[9], double* UU, double* LRW)
{
    const int nc = m->n_cells;
    if (c < 0 || c >= nc || k < 0 || k >= m->nface[c]) return GEL_E_ARG;
    double* cx = (double*)malloc(sizeof(double) * (size_t)nc);
    double* cy = (double*)malloc(sizeof(double) * (size_t)nc);
    if (!cx || !cy) { free(cx); free(cy); return GEL_E_NOMEM; }
    for (int i = 0; i < nc; i++) mesh_center(m, m->x, m->y, i, &cx[i], &cy[i]);
    msw S = { m, prm, now, nc, cx, cy };
    int bad = 0;
    mhd_face(&S, c, c * ORA_MAXF + k, dt_prev, PS, UU, &bad, LRW);
    free(cx);
    free(cy);
    return bad ? GEL_E_NUMERIC : GEL_OK;
}

int ora_step_mhd(const ora_mesh* m, const gel_params* prm, const double* now, double* next,
                 const double* src, double dt_prev, double* dt_new, long long* n_rho_clamp,
                 long long* n_p_clamp, int nthreads)
{
    const int nc = m->n_cells;
    const double gam = prm->gamma;
    const double pi4 = 4.0 * PI_, pi8 = 8.0 * PI_;
    const int once = g_face_once;
    for (int c = 0; c < nc; c++)
        if (m->cell_type[c] == GEL_C_1) return GEL_E_UNSUPPORTED;
    double* cx = (double*)malloc(sizeof(double) * (size_t)nc);
    double* cy = (double*)malloc(sizeof(double) * (size_t)nc);
    if (!cx || !cy) { free(cx); free(cy); return GEL_E_NOMEM; }
    for (int c = 0; c < nc; c++) mesh_center(m, m->x, m->y, c, &cx[c], &cy[c]);
    msw S = { m, prm, now, nc, cx, cy };
    double Tmin = 100000000;
    long long nrc = 0, npc = 0;
    int bad_any = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(min:Tmin) reduction(+:nrc,npc) reduction(|:bad_any)
#endif
    for (int c = 0; c < nc; c++) {
        const double x = cx[c], y = cy[c];
        const double vs1 = mesh_area(m, m->x, m->y, c, 0);
        const double vs2 = mesh_area(m, m->x2, m->y2, c, 1);
        const double radius = sqrt(sq(x) + sq(y));
        if (radius < prm->r_skip) continue;
        const double Volume = mesh_volume_rotate(m, m->x, m->y, c, 360.0) / (2.0 * PI_);
        const double Volume2 = mesh_volume_rotate(m, m->x2, m->y2, c, 360.0) / (2.0 * PI_);
        double par1[MV];
        mload(&S, c, par1);
        double Pot[MV] = { 0 };
        double time = 10000000;
        int bad = 0;
        for (int k = 0; k < m->nface[c]; k++) {
            const int s = c * ORA_MAXF + k;
            const int A = m->fa[s], B = m->fb[s];
            const double fx = (m->x[A] + m->x[B]) / 2.0, fy = (m->y[A] + m->y[B]) / 2.0;
            const double dist = sqrt(sq(x - fx) + sq(y - fy));
            double PS[MV], UU;
            const int N = m->fnbr[s];
            int s2 = -1;
            if (once && m->ftype[s] == GEL_F_USUALY && N < c)
                for (int k2 = 0; k2 < m->nface[N]; k2++)
                    if (m->fnbr[N * ORA_MAXF + k2] == c) { s2 = N * ORA_MAXF + k2; break; }
            if (s2 >= 0) {
                int b2 = 0;
                mhd_face(&S, N, s2, dt_prev, PS, &UU, &b2, NULL);
                for (int v = 0; v < MV; v++) PS[v] = -PS[v];
            } else {
                mhd_face(&S, c, s, dt_prev, PS, &UU, &bad, NULL);
            }
            time = mn(time, prm->cfl * dist / UU);
            for (int v = 0; v < MV; v++) Pot[v] = Pot[v] + PS[v];
        }
        if (bad) bad_any = 1;
        if (time < Tmin) Tmin = time;

        const double ro = par1[0], p = par1[1], u = par1[2], v = par1[3], w = par1[4];
        const double bx = par1[5], by = par1[6], bz = par1[7], Q = par1[8];
        double q1 = 0.0, q2 = 0.0, q3 = 0.0;
        if (prm->source_mode == GEL_SRC_MK1 && src) {
            double f[GEL_MK1_NFIELDS];
            for (int k = 0; k < GEL_MK1_NFIELDS; k++) f[k] = src[(size_t)k * nc + c];
            ora_source_mk1(prm, f, u, v, ro, p, &q1, &q2, &q3);
        }
        const double T1 = dt_prev;
        const double Ar = 0.5 * (vs1 + vs2);
        const double BB = bx * bx + by * by + bz * bz;
        const double Sv = ro * w * w + p + BB / pi8 - bz * bz / pi4;
        const double Sw = -(ro * v * w - by * bz / pi4);
        const double Sb = v * bz - w * by;
        const double E = p / (gam - 1) + ro * (u * u + v * v + w * w) * 0.5 + BB / pi8;
        if (m->cell_type[c] != GEL_C_CENTR) {
            double ro3 = ro * Volume / Volume2 - T1 * (Pot[0] / Volume2);
            if (ro3 <= 0.0) { ro3 = 0.00001; nrc++; }
            const double u3 = (ro * u * Volume / Volume2 - T1 * (Pot[1] / Volume2 - q1)) / ro3;
            const double v3 = (ro * v * Volume / Volume2 - T1 * ((Pot[2] - Sv * Ar) / Volume2 - q2)) / ro3;
            const double w3 = (ro * w * Volume / Volume2 - T1 * ((Pot[3] - Sw * Ar) / Volume2)) / ro3;
            const double E3 = E * Volume / Volume2 - T1 * (Pot[4] / Volume2 - q3);
            const double bx3 = bx * Volume / Volume2 - T1 * (Pot[5] / Volume2);
            const double by3 = by * Volume / Volume2 - T1 * (Pot[6] / Volume2);
            const double bz3 = bz * Volume / Volume2 - T1 * ((Pot[7] - Sb * Ar) / Volume2);
            const double Q3 = Q * Volume / Volume2 - T1 * (Pot[8] / Volume2);
            double p3 = (E3 - 0.5 * ro3 * (u3 * u3 + v3 * v3 + w3 * w3) - (bx3 * bx3 + by3 * by3 + bz3 * bz3) / pi8) * (gam - 1);
            if (p3 <= 0) { p3 = 0.000001; npc++; }
            const double o[MV] = { ro3, p3, u3, v3, w3, bx3, by3, bz3, Q3 };
            for (int k = 0; k < MV; k++) next[(size_t)k * nc + c] = o[k];
        } else {
            for (int k = 0; k < MV; k++) next[(size_t)k * nc + c] = now[(size_t)k * nc + c];
        }
    }
    free(cx);
    free(cy);
    *dt_new = Tmin;
    if (n_rho_clamp) *n_rho_clamp += nrc;
    if (n_p_clamp) *n_p_clamp += npc;
    return bad_any ? GEL_E_NUMERIC : GEL_OK;
}

/* ---- exact Riemann solver (Solvers::Godunov_Solver_Alexashov, Solvers.cpp:693-1172,
 * with newton :1174-1229, devtwo :1231-1305, lev :1307-1321) -------------------------
 * Constants as the reference's macros (Help.h:39-41, Solvers.cpp:7-12) for gamma = gam;
 * eps = 10e-10, eps8 = 10e-8 (Solvers.h:7-8).  Returns the local time step; *status
 * = 1 where the reference would exit. */
typedef struct { double ga, g1, g2, gp, gm; } ora_gk;

static double ora_gd_lev(const ora_gk* k, double enI, double pI, double rI, double enII, double pII, double rII,
                         double uuu)
{   /* Solvers.cpp:1307-1321 */
    double cI = sqrt(k->ga * pI / rI);
    double cII = sqrt(k->ga * pII / rII);
    double fI = (uuu - pI) / (rI * cI * sqrt(k->gp * uuu / pI + k->gm));
    double fII = 2.0 / k->g1 * cII * (pow((uuu / pII), k->gm) - 1.0);
    double f1 = fI + fII;
    double f2 = enI - enII;
    return f1 - f2;
}

static int ora_gd_devtwo(const ora_gk* k, double enI, double pI, double rI, double enII, double pII, double rII,
                         double* p)
{   /* Solvers.cpp:1231-1305: bisection between pI and pII */
    const double epsil = 10e-10;
    double kl = pI, kp = pII, ksi, ksir, ksit, kc;
    int it;
    ksi = ora_gd_lev(k, enI, pI, rI, enII, pII, rII, kl);
    ksir = ora_gd_lev(k, enI, pI, rI, enII, pII, rII, kp);
    if (fabs(ksi) <= epsil) { *p = kl; return 0; }
    if (fabs(ksir) <= epsil) { *p = kp; return 0; }
    for (it = 1; it < 1100; it++) {
        kc = (kl + kp) / 2.0;
        ksit = ora_gd_lev(k, enI, pI, rI, enII, pII, rII, kc);
        if (fabs(ksit) <= epsil) { *p = kc; return 0; }
        if ((ksi * ksit) <= 0.0) { kp = kc; ksir = ksit; }
        else { kl = kc; ksi = ksit; }
    }
    (void)ksir;
    return 1;
}

static int ora_gd_newton(const ora_gk* k, double enI, double pI, double rI, double enII, double pII, double rII,
                         double* pout)
{   /* Solvers.cpp:1174-1229 */
    double cI = sqrt(k->ga * pI / rI);
    double cII = sqrt(k->ga * pII / rII);
    double pn = pI * rII * cII + pII * rI * cI + (enI - enII) * rI * cI * rII * cII;
    double pee, p, fI, fIs, fII, fIIs;
    int kiter = 0;
    pn = pn / (rI * cI + rII * cII);
    pee = pn;
    for (;;) {
        p = pn;
        if (p <= 0.0) return 1;
        kiter = kiter + 1;
        fI = (p - pI) / (rI * cI * sqrt(k->gp * p / pI + k->gm));
        fIs = (k->ga + 1.0) * p / pI + (3.0 * k->ga - 1.0);
        fIs = fIs / (4.0 * k->ga * rI * cI * pow((k->gp * p / pI + k->gm), (3.0 / 2.0)));
        fII = (p - pII) / (rII * cII * sqrt(k->gp * p / pII + k->gm));
        fIIs = (k->ga + 1.0) * p / pII + (3.0 * k->ga - 1.0);
        fIIs = fIIs / (4.0 * k->ga * rII * cII * pow((k->gp * p / pII + k->gm), (3.0 / 2.0)));
        if (kiter == 1100) return 1;
        pn = p - (fI + fII - (enI - enII)) / (fIs + fIIs);
        if (!(fabs(pn / pee - p / pee) >= 10e-10)) break;
    }
    *pout = pn;
    return 0;
}

static double ora_gd_time(double cfl, double dist, double UU)
{
    if (UU > 10e-8) return cfl * dist / UU;
    return cfl * dist / 10e-8;
}

double ora_godunov(const double q1[5], const double q2[5], const double n[3], double w, double dist, double cfl,
                   double gam, double qqq[5], double* dsl, double* dsp, double* dsc, int* status)
{
    ora_gk K;
    double al = n[0], be = n[1], ge = n[2];
    double al2 = be, be2 = -al, ge2 = 0.0, al3 = 0.0, be3 = 0.0, ge3 = 1.0;
    double enI, teI2, teI3, enII, teII2, teII3, pI, pII, rI, rII, cI, cII, a, Uud, Urz, Uvk, Udf;
    double time = 0.0, p = 0.0, r = 0.0, te2 = 0.0, te3 = 0.0, en = 0.0, uo, vo, wo, eo;
    int ipiz = 0, il = -2, ip = -2;
    K.ga = gam; K.g1 = gam - 1.0; K.g2 = gam + 1.0;
    K.gp = (K.g2 / gam) / 2.0; K.gm = (K.g1 / gam) / 2.0;
    *status = 0;
    *dsl = *dsp = *dsc = 0.0;
    enI = al * q1[2] + be * q1[3] + ge * q1[4];
    teI2 = al2 * q1[2] + be2 * q1[3] + ge2 * q1[4];
    teI3 = al3 * q1[2] + be3 * q1[3] + ge3 * q1[4];
    enII = al * q2[2] + be * q2[3] + ge * q2[4];
    teII2 = al2 * q2[2] + be2 * q2[3] + ge2 * q2[4];
    teII3 = al3 * q2[2] + be3 * q2[3] + ge3 * q2[4];
    pI = q1[1]; pII = q2[1]; rI = q1[0]; rII = q2[0];
    if (pI > pII) {                                    /* Solvers.cpp:750-777 */
        double eno2 = enII, teo22 = teII2, teo23 = teII3, p2 = pII, r2 = rII;
        double eno1 = enI, teo12 = teI2, teo13 = teI3, p1 = pI, r1 = rI;
        enI = -eno2; teI2 = teo22; teI3 = teo23; pI = p2; rI = r2;
        enII = -eno1; teII2 = teo12; teII3 = teo13; pII = p1; rII = r1;
        w = -w;
        ipiz = 1;
    }
    cI = (pI != 0.0) ? sqrt(K.ga * pI / rI) : 0.0;
    cII = (pII != 0.0) ? sqrt(K.ga * pII / rII) : 0.0;
    a = sqrt(rI * (K.g2 * pII + K.g1 * pI) / 2.0);
    Uud = (pII - pI) / a;
    Urz = -2.0 * cII / K.g1 * (1.0 - pow((pI / pII), K.gm));
    Uvk = -2.0 * (cII + cI) / K.g1;
    Udf = enI - enII;
    if (Udf < Uvk) { il = -1; ip = -1; }               /* Solvers.cpp:793-824 */
    else if ((Udf >= Uvk) && (Udf <= Urz)) { p = pI * pow(((Udf - Uvk) / (Urz - Uvk)), (1.0 / K.gm)); il = 0; ip = 0; }
    else if ((Udf > Urz) && (Udf <= Uud)) {
        if (ora_gd_devtwo(&K, enI, pI, rI, enII, pII, rII, &p)) { *status = 1; memset(qqq, 0, 5 * sizeof(double)); return 0.0; }
        il = 1; ip = 0;
    } else if (Udf > Uud) {
        if (ora_gd_newton(&K, enI, pI, rI, enII, pII, rII, &p)) { *status = 1; memset(qqq, 0, 5 * sizeof(double)); return 0.0; }
        il = 1; ip = 1;
    }
    if (il == 1 && ip == 1) {                          /* two shocks, :828-883 */
        double aI = sqrt(rI * (K.g2 / 2.0 * p + K.g1 / 2.0 * pI));
        double aII = sqrt(rII * (K.g2 / 2.0 * p + K.g1 / 2.0 * pII));
        double u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        double dI = enI - aI / rI, dII = enII + aII / rII;
        *dsl = dI; *dsp = dII; *dsc = u;
        time = ora_gd_time(cfl, dist, mx(fabs(*dsl), fabs(*dsp)));
        if (w <= dI) { en = enI; p = pI; r = rI; te2 = teI2; te3 = teI3; }
        else if ((w > dI) && (w <= u)) { en = u; r = rI * aI / (aI - rI * (enI - u)); te2 = teI2; te3 = teI3; }
        else if ((w > u) && (w < dII)) { en = u; r = rII * aII / (aII + rII * (enII - u)); te2 = teII2; te3 = teII3; }
        else if (w >= dII) { en = enII; p = pII; r = rII; te2 = teII2; te3 = teII3; }
    }
    if (il == 1 && ip == 0) {                          /* shock + fan, :886-966 */
        double aI = sqrt(rI * (K.g2 / 2.0 * p + K.g1 / 2.0 * pI)), aII, u, dI, dII, ddII, UU;
        if (fabs(p - pII) < 10e-10) aII = rII * cII;
        else aII = K.gm * rII * cII * (1.0 - p / pII) / (1.0 - pow((p / pII), K.gm));
        u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        dI = enI - aI / rI; dII = enII + cII; ddII = u + cII - K.g1 * (enII - u) / 2.0;
        *dsl = dI; *dsp = dII; *dsc = u;
        UU = mx(fabs(*dsl), fabs(*dsp));
        UU = mx(UU, fabs(ddII));
        time = ora_gd_time(cfl, dist, UU);
        if (w <= dI) { en = enI; p = pI; r = rI; te2 = teI2; te3 = teI3; }
        if ((w > dI) && (w <= u)) { en = u; r = rI * aI / (aI - rI * (enI - u)); te2 = teI2; te3 = teI3; }
        if ((w > u) && (w <= ddII)) {
            double ce = cII - K.g1 / 2.0 * (enII - u);
            en = u; r = K.ga * p / ce / ce; te2 = teII2; te3 = teII3;
        }
        if ((w > ddII) && (w < dII)) {
            double ce = -K.g1 / K.g2 * (enII - w) + 2.0 / K.g2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / K.gm)); r = K.ga * p / ce / ce; te2 = teII2; te3 = teII3;
        }
        if (w >= dII) { en = enII; p = pII; r = rII; te2 = teII2; te3 = teII3; }
    }
    if (il == 0 && ip == 0) {                          /* two fans, :968-1065 */
        double aI, aII, u, dI, ddI, dII, ddII, UU;
        if (fabs(p - pI) < 10e-10) aI = rI * cI;
        else aI = K.gm * rI * cI * (1.0 - p / pI) / (1.0 - pow((p / pI), K.gm));
        if (fabs(p - pII) < 10e-10) aII = rII * cII;
        else aII = K.gm * rII * cII * (1.0 - p / pII) / (1.0 - pow((p / pII), K.gm));
        u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        dI = enI - cI; ddI = u - cI - K.g1 * (enI - u) / 2.0;
        dII = enII + cII; ddII = u + cII - K.g1 * (enII - u) / 2.0;
        *dsl = dI; *dsp = dII; *dsc = u;
        UU = mx(fabs(*dsl), fabs(*dsp));
        UU = mx(UU, fabs(ddII));
        UU = mx(UU, fabs(ddI));
        time = ora_gd_time(cfl, dist, UU);
        if (w <= dI) { en = enI; p = pI; r = rI; te2 = teI2; te3 = teI3; }
        if ((w > dI) && (w < ddI)) {
            double ce = K.g1 / K.g2 * (enI - w) + 2.0 / K.g2 * cI;
            en = w + ce; p = pI * pow((ce / cI), (1.0 / K.gm)); r = K.ga * p / ce / ce; te2 = teI2; te3 = teI3;
        }
        if ((w >= ddI) && (w <= u)) {
            double ce = cI + K.g1 / 2.0 * (enI - u);
            en = u; r = K.ga * p / ce / ce; te2 = teI2; te3 = teI3;
        }
        if ((w > u) && (w <= ddII)) {
            double ce = cII - K.g1 / 2.0 * (enII - u);
            en = u; r = K.ga * p / ce / ce; te2 = teII2; te3 = teII3;
        }
        if ((w > ddII) && (w < dII)) {
            double ce = -K.g1 / K.g2 * (enII - w) + 2.0 / K.g2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / K.gm)); r = K.ga * p / ce / ce; te2 = teII2; te3 = teII3;
        }
        if (w >= dII) { en = enII; p = pII; r = rII; te2 = teII2; te3 = teII3; }
    }
    if (il == -1 && ip == -1) {                        /* vacuum, :1068-1140 */
        double gg1 = K.ga - 1.0, gg2 = K.ga + 1.0;
        double dI = enI - cI, ddI = enI + 2.0 / gg1 * cI, dII = enII + cII, ddII = enII - 2.0 / gg1 * cII, UU;
        *dsl = dI; *dsp = dII; *dsc = (dI + dII) / 2.0;
        UU = mx(fabs(*dsl), fabs(*dsp));
        UU = mx(UU, fabs(ddII));
        UU = mx(UU, fabs(ddI));
        time = ora_gd_time(cfl, dist, UU);
        if (w <= dI) { en = enI; p = pI; r = rI; te2 = teI2; te3 = teI3; }
        if ((w > dI) && (w < ddI)) {
            double ce = gg1 / gg2 * (enI - w) + 2.0 / gg2 * cI;
            en = w + ce; p = pI * pow((ce / cI), (1.0 / K.gm)); r = K.ga * p / ce / ce; te2 = teI2; te3 = teI3;
        }
        if ((w >= ddI) && (w <= ddII)) { en = w; p = 0.0; r = 0.0; te2 = 0.0; te3 = 0.0; }
        if ((w > ddII) && (w < dII)) {
            double ce = -gg1 / gg2 * (enII - w) + 2.0 / gg2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / K.gm)); r = K.ga * p / ce / ce; te2 = teII2; te3 = teII3;
        }
        if (w >= dII) { en = enII; p = pII; r = rII; te2 = teII2; te3 = teII3; }
    }
    if (ipiz == 1) {                                   /* :1144-1153 */
        double dsl1 = *dsl, dsp1 = *dsp;
        en = -en;
        *dsl = -dsp1; *dsp = -dsl1; *dsc = -*dsc;
        w = -w;
    }
    uo = al * en + al2 * te2 + al3 * te3;              /* :1155-1168 */
    vo = be * en + be2 * te2 + be3 * te3;
    wo = ge * en + ge2 * te2 + ge3 * te3;
    eo = p / K.g1 + 0.5 * r * (uo * uo + vo * vo + wo * wo);
    en = al * uo + be * vo + ge * wo;
    qqq[0] = (r * (en - w));
    qqq[1] = (r * (en - w) * uo + al * p);
    qqq[2] = (r * (en - w) * vo + be * p);
    qqq[3] = (r * (en - w) * wo + ge * p);
    qqq[4] = ((en - w) * eo + en * p);
    return time;
}

/* ------------------------------------------------------ fitted-surface tracking --- */

/* Node speeds of the fitted surfaces (Setka::Move_surface, Setka.cpp:3378-3844) and the
 * A-rail redistribution of Setka::Move_Setka_Calculate_2 (Setka.cpp:4324-4381) on the
 * polar grid: each surface is the node row t->row_*, its line the radial faces of that
 * row in the order of the polar angle (face j joins nodes (k, j) -> (k, j+1), master
 * cell (k-1, j), Setka.cpp:1007-1045), each node column c is a rail at angle rail_s[c]. */

typedef struct trk_acc { double* vx; double* vy; int* cnt; } trk_acc;

/* Get_par_TVD of both sides of the surface face j of row k, with the inner shock's
 * master side taken from the cell centre rescaled to the face (Setka.cpp:3560-3575). */
static void trk_face(const gel_tracking* t, int line, int ni, int nj, int k, int j, const double* x, const double* y,
                     const double* now, double gam, double* n1o, double* n2o, ora_gas* p1, ora_gas* p2)
{
    (void)t; (void)ni;
    const int nc = ni * nj;
    const int a = k * (nj + 1) + j, b = a + 1;
    const double fx = (x[a] + x[b]) / 2.0, fy = (y[a] + y[b]) / 2.0;
    double t1 = x[b] - x[a], t2 = y[b] - y[a];
    const double nn = sqrt(sq(t1) + sq(t2));
    t1 = t1 / nn; t2 = t2 / nn;
    *n1o = t2; *n2o = -t1;
    int q[4];
    double cx[4], cy[4];               /* cells (k-2, j), (k-1, j), (k, j), (k+1, j) */
    for (int r = 0; r < 4; r++) {
        const int i = k - 2 + r;
        q[0] = i * (nj + 1) + j; q[1] = (i + 1) * (nj + 1) + j; q[2] = q[1] + 1; q[3] = q[0] + 1;
        ora_cell_center(x, y, q, &cx[r], &cy[r]);
    }
    const int D = (k - 2) * nj + j, M = (k - 1) * nj + j, N = k * nj + j, U = (k + 1) * nj + j;
    if (line == 1) {                   /* inner shock: Setka.cpp:3560-3575 */
        ora_gas g = { now[M], now[nc + M], now[2 * nc + M], now[3 * nc + M], now[4 * nc + M] };
        const double radius = sqrt(sq(cx[1]) + sq(cy[1]));
        const double dis = sqrt(sq(fx) + sq(fy));
        g.ro = g.ro * sq(radius) / sq(dis);
        g.p = g.p * pow(radius / dis, 2.0 * gam);
        ora_polar_perenos(cx[1], cy[1], fx, fy, &g.u, &g.v);
        *p1 = g;
    } else {
        *p1 = tvd_surface(now, nc, M, D, cx[1], cy[1], cx[0], cy[0], fx, fy);
    }
    *p2 = tvd_surface(now, nc, N, U, cx[2], cy[2], cx[3], cy[3], fx, fy);
}

static void trk_add(trk_acc* acc, int node, double v, double n1, double n2)
{
    acc->vx[node] += v * n1;
    acc->vy[node] += v * n2;
    acc->cnt[node]++;
}

static void trk_norm(trk_acc* acc, int node)
{
    if (acc->cnt[node] > 0) {
        acc->vx[node] /= (1.0 * acc->cnt[node]);
        acc->vy[node] /= (1.0 * acc->cnt[node]);
    }
}

int ora_track_nodes(const gel_tracking* t, int ni, int nj, const double* x, const double* y, const double* now,
                    const double* rail_s, double gam, double dt, double* x2, double* y2)
{
    const int nn = (ni + 1) * (nj + 1);
    trk_acc acc;
    acc.vx = (double*)calloc((size_t)nn, sizeof(double));
    acc.vy = (double*)calloc((size_t)nn, sizeof(double));
    acc.cnt = (int*)calloc((size_t)nn, sizeof(int));
    if (!acc.vx || !acc.vy || !acc.cnt) { free(acc.vx); free(acc.vy); free(acc.cnt); return GEL_E_NOMEM; }
    int bad = 0;
    const int rows[3] = { t->row_contact, t->row_inner_shock, t->row_outer_shock };
    for (int line = 0; line < 3; line++) {          /* contact, inner shock, outer shock */
        const int k = rows[line];
        const int base = k * (nj + 1);
        for (int j = 0; j < nj; j++) {
            if (line == 2 && j >= t->outer_faces) continue;     /* Setka.cpp:3746 */
            double n1, n2;
            ora_gas p1, p2;
            trk_face(t, line, ni, nj, k, j, x, y, now, gam, &n1, &n2, &p1, &p2);
            const double q1[5] = { p1.ro, p1.p, p1.u, p1.v, 0.0 }, q2[5] = { p2.ro, p2.p, p2.u, p2.v, 0.0 };
            const double nv[3] = { n1, n2, 0.0 };
            double qqq[5], Vl, Vp, VV;
            int st = 0;
            ora_godunov(q1, q2, nv, 0.0, 0.0, 0.1, gam, qqq, &Vl, &Vp, &VV, &st);
            if (st) bad = 1;
            const int A = base + j, B = A + 1;
            const double t1 = -n2, t2 = n1;
            if (line == 0) {                            /* Setka.cpp:3434-3470 */
                VV = VV * t->koef * t->k_contact;
                trk_add(&acc, B, VV, n1, n2);
            } else if (line == 1) {                     /* :3595-3616 */
                Vl = Vl * t->koef * t->k_inner;
                if (fabs(p2.u * t1 + p2.v * t2) < t->tang_inner) {
                    trk_add(&acc, A, Vl, n1, n2);
                    trk_add(&acc, B, Vl, n1, n2);
                } else {
                    trk_add(&acc, B, Vl, n1, n2);
                }
            } else {                                    /* :3782-3825 */
                Vp = Vp * t->koef * t->k_outer;
                const double ut = p1.u * t1 + p1.v * t2;
                if ((sq(p1.u) + sq(p1.v) + sq(0.0)) / (gam * p1.p / p1.ro) > 1.0) {
                    trk_add(&acc, B, Vp, n1, n2);
                } else if (fabs(ut) < t->tang_outer) {
                    trk_add(&acc, A, Vp, n1, n2);
                    trk_add(&acc, B, Vp, n1, n2);
                } else if (ut > 0) {
                    trk_add(&acc, B, Vp, n1, n2);
                } else {
                    trk_add(&acc, A, Vp, n1, n2);
                }
            }
        }
        /* averaging: the reference divides the A nodes of the contact line after each
           of the three loops (Setka.cpp:3522-3530, :3718-3731, :3829-3837) */
        const int avg_row = t->average_lines ? k : t->row_contact;
        if (line == 0 || !t->average_lines) {
            for (int j = 0; j < nj; j++) trk_norm(&acc, avg_row * (nj + 1) + j);
        } else {
            for (int j = 0; j < nj; j++) trk_norm(&acc, base + j);
        }
        /* end nodes (Setka.cpp:3532-3535, :3733-3736, :3839-3842) */
        const int n0 = base, n1_ = base + 1, nl = base + nj, nl1 = base + nj - 1;
        acc.vx[n0] = acc.vx[n1_] + (x[n1_] - x[n0]) / dt;
        if (line == 1) acc.vx[nl] = acc.vx[nl1] + (x[nl1] - x[nl]) / dt;
        else acc.vy[nl] = acc.vy[nl1] + (y[nl1] - y[nl]) / dt;
    }

    /* A rails (Setka.cpp:4328-4381): keys = inner shock, contact, outer shock, outer end */
    const int M1 = t->row_inner_shock + 1;
    const int M2 = t->row_contact - t->row_inner_shock;
    const int M3 = t->row_outer_shock - t->row_contact;
    const int M4 = ni - t->row_outer_shock;
    const double dd = t->s_outer / sqrt(1.0 - 1.0 / t->p_outer);
    for (int c = 0; c <= nj; c++) {
        const double s = rail_s[c];
        const int key[3] = { t->row_inner_shock, t->row_contact, t->row_outer_shock };
        for (int q = 0; q < 3; q++) {
            const int id = key[q] * (nj + 1) + c;
            const double V = acc.vx[id] * cos(s) + acc.vy[id] * sin(s);
            x2[id] = x[id] + dt * V * cos(s);
            y2[id] = y[id] + dt * V * sin(s);
        }
        {
            const int id = ni * (nj + 1) + c;
            x2[id] = x[id];
            y2[id] = y[id];
        }
        const int k0 = key[0] * (nj + 1) + c, k1 = key[1] * (nj + 1) + c, k2 = key[2] * (nj + 1) + c;
        const double R2 = sqrt(sq(x2[k0]) + sq(y2[k0]));
        double xx, r;
        for (int j = 0; j < t->inner_rail_nodes; j++) {
            xx = j * 1.0 / (t->inner_rail_nodes);
            r = t->r_inner + pow(xx, t->p_inner) * (t->r_mid - t->r_inner);
            x2[j * (nj + 1) + c] = r * cos(s);
            y2[j * (nj + 1) + c] = r * sin(s);
        }
        for (int j = t->inner_rail_nodes; j < M1 - 1; j++) {
            xx = (j - t->inner_rail_nodes) * 1.0 / (M1 - 1 - t->inner_rail_nodes);
            r = t->r_mid + pow(xx, t->p_mid) * (R2 - t->r_mid);
            x2[j * (nj + 1) + c] = r * cos(s);
            y2[j * (nj + 1) + c] = r * sin(s);
        }
        const double R3 = sqrt(sq(x2[k1]) + sq(y2[k1]));
        for (int j = 0; j < M2 - 1; j++) {
            r = R2 + (R3 - R2) * (j + 1) / (M2);
            x2[(M1 + j) * (nj + 1) + c] = r * cos(s);
            y2[(M1 + j) * (nj + 1) + c] = r * sin(s);
        }
        const double R4 = sqrt(sq(x2[k2]) + sq(y2[k2]));
        for (int j = 0; j < M3 - 1; j++) {
            r = R3 + (R4 - R3) * (j + 1) / (M3);
            x2[(M1 + M2 + j) * (nj + 1) + c] = r * cos(s);
            y2[(M1 + M2 + j) * (nj + 1) + c] = r * sin(s);
        }
        for (int j = 0; j < M4; j++) {
            xx = (j + 1) * 1.0 / (M4);
            r = R4 + pow(xx, t->p_outer * (1.0 - sq(s / dd))) * (t->r_outer - R4);
            x2[(M1 + M2 + M3 + j) * (nj + 1) + c] = r * cos(s);
            y2[(M1 + M2 + M3 + j) * (nj + 1) + c] = r * sin(s);
        }
    }
    free(acc.vx);
    free(acc.vy);
    free(acc.cnt);
    return bad ? GEL_E_NUMERIC : GEL_OK;
}

/* ------------------------------------------------ plasma + four hydrogen fluids --- */

/* Setka::Go_5_komponent_2 (Setka.cpp:8951-9777): the plasma and the hydrogen fluids
 * H1..H4 on one grid, 21 primitives per cell in Parametr order (ro, p, u, v, Q, then
 * ro_Hk, p_Hk, u_Hk, v_Hk).  Every fluid takes the HLL/HLLC face flux; the fluids
 * exchange mass, momentum and energy by charge exchange. */

#define MFV 21
static const int mf_odd[MFV] = { 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1 };

typedef struct mfs { const ora_mesh* m; const gel_params* prm; const double* now; int nc;
                     const double* cx; const double* cy; } mfs;

static void mf_load(const mfs* S, int c, double* g)
{
    for (int k = 0; k < MFV; k++) g[k] = S->now[(size_t)k * S->nc + c];
}

/* Gran::Get_par_TVD for all 21 fields (Gran.cpp:294-474): no down cell -> master (with
 * the mirrored ghost across the axis for every v), else the limited line through
 * (down, master, sosed).  *bad where the plasma density or pressure is non-positive
 * (the reference exits, :444-455). */
static void mf_tvd(const mfs* S, int M, int N, int D, double fx, double fy, double* r, int* bad)
{
    double pm[MFV];
    mf_load(S, M, pm);
    if (D < 0) {
        for (int k = 0; k < MFV; k++) r[k] = pm[k];
        if (cell_has_axis(S->m, M)) {
            double pn[MFV];
            mf_load(S, N, pn);
            const double d2 = dist2d(fx, fy, S->cx[M], S->cy[M]);
            const double d3 = dist2d(fx, fy, S->cx[N], S->cy[N]);
            const double d1 = dist2d(fx, fy, S->cx[M], -S->cy[M]);
            for (int k = 0; k < MFV; k++)
                if (mf_odd[k]) r[k] = ora_linear3(-d1, -pm[k], -d2, pm[k], d3, pn[k], 0.0);
        }
        return;
    }
    double pd[MFV], pn[MFV];
    mf_load(S, D, pd);
    mf_load(S, N, pn);
    const double d2 = dist2d(fx, fy, S->cx[M], S->cy[M]);
    const double d3 = dist2d(fx, fy, S->cx[N], S->cy[N]);
    const double d1 = dist2d(fx, fy, S->cx[D], S->cy[D]);
    for (int k = 0; k < MFV; k++) r[k] = ora_linear3(-d1, pd[k], -d2, pm[k], d3, pn[k], 0.0);
    if (r[0] <= 0.0 || r[1] <= 0.0) *bad = 1;
}

/* Gran::Get_par_TVD_radial, the fields the sweep takes from it (Gran.cpp:562-675,
 * Setka.cpp:9034-9040): plasma ro (r^2-weighted), p (r^(2 gamma)), p_H1 (plain),
 * ro_H1 (r^H_pow-weighted). */
static void mf_tvd_radial(const mfs* S, int M, int N, int D, double fx, double fy, double* ro, double* pr,
                          double* ph1, double* roh1)
{
    const double g2 = 2.0 * S->prm->gamma, hp = S->prm->h_pow;
    double pm[MFV], pn[MFV], pd[MFV];
    mf_load(S, M, pm);
    mf_load(S, N, pn);
    double xd, yd;
    if (D < 0) {
        if (!cell_has_axis(S->m, M)) { *ro = pm[0]; *pr = pm[1]; *ph1 = pm[6]; *roh1 = pm[5]; return; }
        xd = S->cx[M]; yd = -S->cy[M];
        for (int k = 0; k < MFV; k++) pd[k] = pm[k];
    } else {
        xd = S->cx[D]; yd = S->cy[D];
        mf_load(S, D, pd);
    }
    const double r = sqrt(sq(fx) + sq(fy));
    const double r2 = sqrt(sq(S->cx[M]) + sq(S->cy[M]));
    const double d2 = dist2d(fx, fy, S->cx[M], S->cy[M]);
    const double r3 = sqrt(sq(S->cx[N]) + sq(S->cy[N]));
    const double d3 = dist2d(fx, fy, S->cx[N], S->cy[N]);
    const double r1 = sqrt(sq(xd) + sq(yd));
    const double d1 = dist2d(fx, fy, xd, yd);
    *ro = ora_linear3(-d1, pd[0] * sq(r1), -d2, pm[0] * sq(r2), d3, pn[0] * sq(r3), 0.0) / sq(r);
    *pr = ora_linear3(-d1, pd[1] * pow(r1, g2), -d2, pm[1] * pow(r2, g2), d3, pn[1] * pow(r3, g2), 0.0) / pow(r, g2);
    *ph1 = ora_linear3(-d1, pd[6], -d2, pm[6], d3, pn[6], 0.0);
    *roh1 = ora_linear3(-d1, pd[5] * pow(r1, hp), -d2, pm[5] * pow(r2, hp), d3, pn[5] * pow(r3, hp), 0.0) / pow(r, hp);
}

/* Gran::Get_par of the 21 fields (Gran.cpp:205-289). */
static void mf_ghost(const mfs* S, int M, int slot, double fx, double fy, double* g)
{
    const gel_params* p = S->prm;
    const int t = S->m->ftype[slot];
    mf_load(S, M, g);
    switch (t) {
    case GEL_F_USUALY:
        mf_load(S, S->m->fnbr[slot], g);
        return;
    case GEL_F_EXTERN:
        if (g[2] >= 0.0) g[2] = p->extern_u;
        return;
    case GEL_F_AXIS:
        for (int k = 0; k < MFV; k++) if (mf_odd[k]) g[k] = -g[k];
        return;
    case GEL_F_INNER_SPHERE: {
        const double d = sqrt(fx * fx + fy * fy);
        const double ro = 1.0 / (p->in_rho_div * d * d);
        g[0] = ro; g[1] = ro * p->in_p_a * p->in_p_a / p->in_p_den;
        g[2] = p->in_u * fx / d; g[3] = p->in_u * fy / d; g[4] = ro * p->in_q;
        g[5] = p->h1_in_rho; g[6] = p->h1_in_p; g[7] = p->in_u * fx / d; g[8] = p->in_u * fy / d;
        return;
    }
    case GEL_F_INPUT:
        g[0] = p->inf_rho; g[1] = p->inf_p; g[2] = p->inf_u; g[3] = p->inf_v; g[4] = p->inf_q;
        g[17] = p->h4_inf_rho; g[18] = p->h4_inf_p; g[19] = p->h4_inf_u; g[20] = p->h4_inf_v;
        return;
    default:
        return;
    }
}

/* rescaling of a C_1 cell value to the face (Setka.cpp:9021-9028): plasma ro, Q by
 * (rc/dis)^2, p by (rc/dis)^(2 gamma), ro_H1 by (rc/dis)^H_pow; rotate: (u, v) and
 * (u_H1, v_H1) to the face's polar frame. */
static void mf_rescale(const gel_params* p, double* g, double rc, double dis, double cx, double cy, double fx,
                       double fy, int rotate)
{
    g[0] = g[0] * sq(rc) / sq(dis);
    g[5] = g[5] * pow(rc / dis, p->h_pow);
    g[4] = g[4] * sq(rc) / sq(dis);
    g[1] = g[1] * pow(rc / dis, 2 * p->gamma);
    if (rotate) {
        ora_polar_perenos(cx, cy, fx, fy, &g[2], &g[3]);
        ora_polar_perenos(cx, cy, fx, fy, &g[7], &g[8]);
    }
}

static double mf_flux(const double* L, const double* R, int h, double W, double n1, double n2, double dist,
                      const gel_params* prm, int metod, double P[4], double* PQ)
{
    /* fluid h: 0 plasma (its own tracer), 1..4 hydrogen with the plasma tracer
       (Setka.cpp:9176-9199; the tracer flux of the hydrogen calls is not used) */
    double Vl, Vc, Vp;
    ora_gas a, b;
    if (h == 0) {
        a.ro = L[0]; a.p = L[1]; a.u = L[2]; a.v = L[3]; a.Q = L[4];
        b.ro = R[0]; b.p = R[1]; b.u = R[2]; b.v = R[3]; b.Q = R[4];
    } else {
        const int o = 5 + 4 * (h - 1);
        a.ro = L[o]; a.p = L[o + 1]; a.u = L[o + 2]; a.v = L[o + 3]; a.Q = L[4];
        b.ro = R[o]; b.p = R[o + 1]; b.u = R[o + 2]; b.v = R[o + 3]; b.Q = R[4];
    }
    return ora_hllc2d(&a, &b, W, n1, n2, dist, prm->cfl, prm->gamma, metod, 0, P, PQ, &Vl, &Vc, &Vp);
}

int ora_step_mf(const ora_mesh* m, const gel_params* prm, const double* now, double* next, double dt_prev,
                double* dt_new, long long* n_rho_clamp, long long* n_p_clamp, int nthreads)
{
    const int nc = m->n_cells;
    const double gam = prm->gamma;
    double* cx = (double*)malloc(sizeof(double) * (size_t)nc);
    double* cy = (double*)malloc(sizeof(double) * (size_t)nc);
    if (!cx || !cy) { free(cx); free(cy); return GEL_E_NOMEM; }
    for (int c = 0; c < nc; c++) mesh_center(m, m->x, m->y, c, &cx[c], &cy[c]);
    mfs S = { m, prm, now, nc, cx, cy };
    double Tmin = 100000000;
    long long nrc = 0, npc = 0;
    int bad_any = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(min:Tmin) reduction(+:nrc,npc) reduction(|:bad_any)
#endif
    for (int c = 0; c < nc; c++) {
        const double x = cx[c], y = cy[c];
        const double vs1 = mesh_area(m, m->x, m->y, c, 0);
        const double vs2 = mesh_area(m, m->x2, m->y2, c, 1);
        const double radius = sqrt(sq(x) + sq(y));
        if (radius < prm->r_skip) continue;                       /* Setka.cpp:8997-9000 */
        const double Volume = mesh_volume_rotate(m, m->x, m->y, c, 360.0) / (2.0 * PI_);
        const double Volume2 = mesh_volume_rotate(m, m->x2, m->y2, c, 360.0) / (2.0 * PI_);
        double par1[MFV];
        mf_load(&S, c, par1);
        const int ct = m->cell_type[c];
        const int K1 = (ct == GEL_C_1), KC = (ct == GEL_C_CENTR);
        double Pot[5] = { 0, 0, 0, 0, 0 }, PH[4][4];
        memset(PH, 0, sizeof PH);
        double time = 10000000;
        int bad = 0;
        for (int k = 0; k < m->nface[c]; k++) {
            const int s = c * ORA_MAXF + k;
            const int A = m->fa[s], B = m->fb[s], t = m->ftype[s];
            const double S_ = sqrt(sq(m->x[B] - m->x[A]) + sq(m->y[B] - m->y[A])) * (m->y[A] + m->y[B]) * 0.5;
            const double fx = (m->x[A] + m->x[B]) / 2.0, fy = (m->y[A] + m->y[B]) / 2.0;
            const double fx2 = (m->x2[A] + m->x2[B]) / 2.0, fy2 = (m->y2[A] + m->y2[B]) / 2.0;
            const double dist = sqrt(sq(x - fx) + sq(y - fy));
            double n1, n2;
            {
                double t1 = m->x[B] - m->x[A], t2 = m->y[B] - m->y[A];
                const double nn = sqrt(sq(t1) + sq(t2));
                t1 = t1 / nn; t2 = t2 / nn;
                n1 = t2; n2 = -t1;
            }
            const double W = ((fx2 - fx) * n1 + (fy2 - fy) * n2) / dt_prev;
            const double dis = sqrt(sq(fx) + sq(fy));
            double p11[MFV], p2[MFV];
            for (int v = 0; v < MFV; v++) p11[v] = par1[v];
            mf_ghost(&S, c, s, fx, fy, p2);
            if (t == GEL_F_USUALY) {                              /* :9014-9070 */
                const int N = m->fnbr[s];
                if (K1) {
                    mf_rescale(prm, p11, radius, dis, x, y, fx, fy, 1);
                    double ro, pr, ph1, roh1;
                    mf_tvd_radial(&S, c, N, m->fdown[s], fx, fy, &ro, &pr, &ph1, &roh1);
                    p11[0] = ro; p11[1] = pr; p11[6] = ph1; p11[5] = roh1;
                } else {
                    mf_tvd(&S, c, N, m->fdown[s], fx, fy, p11, &bad);
                }
                if (m->cell_type[N] == GEL_C_1) {
                    const double dd = sqrt(sq(cx[N]) + sq(cy[N]));
                    mf_rescale(prm, p2, dd, dis, cx[N], cy[N], fx, fy, 1);
                    double ro, pr, ph1, roh1;
                    mf_tvd_radial(&S, N, c, m->fup[s], fx, fy, &ro, &pr, &ph1, &roh1);
                    p2[0] = ro; p2[1] = pr; p2[6] = ph1; p2[5] = roh1;
                } else {
                    mf_tvd(&S, N, c, m->fup[s], fx, fy, p2, &bad);
                }
            } else if (t == GEL_F_INNER_SPHERE) {                 /* :9071-9077 */
                mf_rescale(prm, p11, radius, dis, x, y, fx, fy, 0);
            } else if (K1 && t == GEL_F_AXIS) {                   /* :9078-9093 */
                mf_rescale(prm, p11, radius, dis, x, y, fx, fy, 1);
                for (int v = 0; v < MFV; v++) p2[v] = mf_odd[v] ? -p11[v] : p11[v];
            } else if (t == GEL_F_AXIS) {                         /* :9094-9104 */
                /* Get_par_TVD on an axis face (Gran.cpp:486-519): the master's fields with
                   the plasma v on the line through down, master and the master's mirror */
                const int D = m->fdown[s];
                const double d2 = dist2d(fx, fy, cx[c], cy[c]);
                const double d1 = dist2d(fx, fy, cx[D], cy[D]);
                p11[3] = ora_linear3(-d1, now[(size_t)3 * nc + D], -d2, par1[3], d2, -par1[3], 0.0);
                for (int v = 0; v < MFV; v++) p2[v] = mf_odd[v] ? -p11[v] : p11[v];
            }
            double P[4], PQ;
            if (!KC) {                                            /* :9172-9181 */
                time = mn(time, mf_flux(p11, p2, 1, W, n1, n2, dist, prm, prm->flux, P, &PQ));
                for (int v = 0; v < 4; v++) PH[0][v] = PH[0][v] + P[v] * S_;
            }
            for (int h = 2; h <= 4; h++) {                        /* :9182-9199 */
                time = mn(time, mf_flux(p11, p2, h, W, n1, n2, dist, prm, prm->flux, P, &PQ));
                for (int v = 0; v < 4; v++) PH[h - 1][v] = PH[h - 1][v] + P[v] * S_;
            }
            time = mn(time, mf_flux(p11, p2, 0, W, n1, n2, dist, prm, prm->flux, P, &PQ));   /* :9226-9234 */
            for (int v = 0; v < 4; v++) Pot[v] = Pot[v] + P[v] * S_;
            Pot[4] = Pot[4] + PQ * S_;
        }
        if (bad) bad_any = 1;
        if (time < Tmin) Tmin = time;

        /* charge exchange (Setka.cpp:9419-9470) from the cell's values */
        const double u = par1[0 + 2], v = par1[3], ro = par1[0], p = par1[1], Q = par1[4];
        double uH[4], vH[4], roH[4], pH[4], UM[4], UH[4], nu[4];
        for (int h = 0; h < 4; h++) {
            roH[h] = par1[5 + 4 * h]; pH[h] = par1[6 + 4 * h]; uH[h] = par1[7 + 4 * h]; vH[h] = par1[8 + 4 * h];
        }
        const double pi = 3.14159265358979323846;    /* Solvers.h:9 */
        for (int h = 0; h < 4; h++) {
            UM[h] = sqrt(sq(u - uH[h]) + sq(v - vH[h]) + (64.0 / (9.0 * pi)) * (p / ro + 2.0 * pH[h] / roH[h]));
            UH[h] = sqrt(sq(u - uH[h]) + sq(v - vH[h]) + (4.0 / pi) * (p / ro + 2.0 * pH[h] / roH[h]));
        }
        for (int h = 0; h < 4; h++) {
            const double sg = sq(prm->sig_a - prm->sig_b * log(UM[h]));
            nu[h] = ro * roH[h] * UM[h] * sg;
        }
        const double cp = prm->src_coef, ch = prm->src_coef_h;
        const double q2_1 = cp * (nu[0] * (uH[0] - u) + nu[1] * (uH[1] - u) + nu[2] * (uH[2] - u) + nu[3] * (uH[3] - u));
        const double q2_2 = cp * (nu[0] * (vH[0] - v) + nu[1] * (vH[1] - v) + nu[2] * (vH[2] - v) + nu[3] * (vH[3] - v));
        double e3[4];
        for (int h = 0; h < 4; h++)
            e3[h] = nu[h] * ((sq(uH[h]) + sq(vH[h]) - sq(u) - sq(v)) / 2.0 + (UH[h] / UM[h]) * (2.0 * pH[h] / roH[h] - p / ro));
        const double q3 = cp * (e3[0] + e3[1] + e3[2] + e3[3]);
        const double T1 = dt_prev;
        if (!KC) {                                                /* :9491-9535 */
            double ro3 = ro * Volume / Volume2 - T1 * (Pot[0] / Volume2);
            const double Q33 = Q * Volume / Volume2 - (T1 / Volume2) * Pot[4];
            if (ro3 <= 0.0) { ro3 = 0.00001; nrc++; }
            const double u3 = (ro * u * Volume / Volume2 - T1 * (Pot[1] / Volume2 - q2_1)) / ro3;
            const double v3 = (ro * v * Volume / Volume2 - T1 * ((Pot[2] - p * 0.5 * (vs1 + vs2)) / Volume2 - q2_2)) / ro3;
            double p3 = (((p / (gam - 1) + ro * (u * u + v * v) * 0.5) * Volume / Volume2 - T1 * (Pot[3] / Volume2 - q3)) -
                         0.5 * ro3 * (u3 * u3 + v3 * v3)) * (gam - 1);
            if (p3 <= 0) { p3 = 0.000001; npc++; }
            next[c] = ro3; next[(size_t)nc + c] = p3; next[(size_t)2 * nc + c] = u3; next[(size_t)3 * nc + c] = v3;
            next[(size_t)4 * nc + c] = Q33;
        } else {
            for (int k = 0; k < 5; k++) next[(size_t)k * nc + c] = now[(size_t)k * nc + c];
        }
        /* populations of the source regions (Setka.cpp:9554-9586) */
        int kk[4] = { 0, 0, 0, 0 };
        if (ct == GEL_C_CENTR || ct == GEL_C_1 || ct == GEL_C_2 || ct == GEL_C_3) {
            if (ct == GEL_C_1 || ct == GEL_C_CENTR) kk[0] = 1; else kk[1] = 1;
        } else {
            if (ct == GEL_C_4) kk[2] = 1; else kk[3] = 1;
        }
        const double S1 = nu[0] + nu[1] + nu[2] + nu[3];
        const double S2 = nu[0] * ((sq(u) + sq(v)) / 2.0 + (UH[0] / UM[0]) * (p / ro))
                        + nu[1] * ((sq(u) + sq(v)) / 2.0 + (UH[1] / UM[1]) * (p / ro))
                        + nu[2] * ((sq(u) + sq(v)) / 2.0 + (UH[2] / UM[2]) * (p / ro))
                        + nu[3] * ((sq(u) + sq(v)) / 2.0 + (UH[3] / UM[3]) * (p / ro));
        static const double ro_floor[4] = { 0.0000001, 0.000001, 0.000001, 0.000001 };   /* :9627, :9672, :9701, :9733 */
        static const double p_floor[4] = { 0.000001, 0.000001, 0.000001, 0.0000001 };     /* :9642, :9687, :9716, :9748 */
        for (int h = 0; h < 4; h++) {                             /* :9593-9753 */
            const int o = 5 + 4 * h;
            if (h == 0 && KC) {
                for (int k = 0; k < 4; k++) next[(size_t)(o + k) * nc + c] = now[(size_t)(o + k) * nc + c];
                continue;
            }
            const double q1h = ch * (kk[h] * S1 - nu[h]);
            const double q21h = ch * (kk[h] * S1 * u - nu[h] * uH[h]);
            const double q22h = ch * (kk[h] * S1 * v - nu[h] * vH[h]);
            const double q3h = ch * (kk[h] * S2 - nu[h] * ((sq(uH[h]) + sq(vH[h])) / 2.0 + (UH[h] / UM[h]) * 2.0 * (pH[h] / roH[h])));
            double ro3 = roH[h] * Volume / Volume2 - T1 * (PH[h][0] / Volume2 - q1h);
            if (ro3 <= 0.0) ro3 = ro_floor[h];
            const double u3 = (roH[h] * uH[h] * Volume / Volume2 - T1 * (PH[h][1] / Volume2 - q21h)) / ro3;
            const double v3 = (roH[h] * vH[h] * Volume / Volume2 - T1 * ((PH[h][2] - pH[h] * 0.5 * (vs1 + vs2)) / Volume2 - q22h)) / ro3;
            double p3 = (((pH[h] / (gam - 1) + roH[h] * (uH[h] * uH[h] + vH[h] * vH[h]) * 0.5) * Volume / Volume2 -
                          T1 * (PH[h][3] / Volume2 - q3h)) - 0.5 * ro3 * (u3 * u3 + v3 * v3)) * (gam - 1);
            if (p3 <= 0) p3 = p_floor[h];
            next[(size_t)o * nc + c] = ro3; next[(size_t)(o + 1) * nc + c] = p3;
            next[(size_t)(o + 2) * nc + c] = u3; next[(size_t)(o + 3) * nc + c] = v3;
        }
    }
    free(cx);
    free(cy);
    *dt_new = Tmin;
    if (n_rho_clamp) *n_rho_clamp += nrc;
    if (n_p_clamp) *n_p_clamp += npc;
    return bad_any ? GEL_E_NUMERIC : GEL_OK;
}

/* ------------------------------------------------------------ inner region --- */

/* Setka::Go_stationary_5_komponent_inner_MK2 (Setka.cpp:6999-7212): the static solver
 * of the inner region.  Every interior face takes both cells' values rescaled to the
 * face radius (density and tracer by r^2, pressure by r^(2 gamma), velocity rotated to
 * the face's polar frame), replaced by the radial reconstruction of density and
 * pressure where both cells are C_1; inner-sphere faces take the ghost on both sides;
 * axis faces the rescaled, mirrored cell value with the HLL flux; C_centr cells are
 * skipped; the update is the static one with the tracer losing a bare dt per step
 * (Setka.cpp:7173). */
int ora_step_inner(const ora_mesh* m, const gel_params* prm, const double* now, double* next, const double* src,
                   double dt_prev, double* dt_new, long long* n_rho_clamp, long long* n_p_clamp, int nthreads)
{
    const int nc = m->n_cells;
    const double gam = prm->gamma;
    double* cx = (double*)malloc(sizeof(double) * (size_t)nc);
    double* cy = (double*)malloc(sizeof(double) * (size_t)nc);
    if (!cx || !cy) { free(cx); free(cy); return GEL_E_NOMEM; }
    for (int c = 0; c < nc; c++) mesh_center(m, m->x, m->y, c, &cx[c], &cy[c]);
    sweep_ctx S = { m, prm, now, nc, cx, cy };
    double Tmin = 100000000;
    long long nrc = 0, npc = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(min:Tmin) reduction(+:nrc,npc)
#endif
    for (int c = 0; c < nc; c++) {
        if (m->cell_type[c] == GEL_C_CENTR) continue;             /* Setka.cpp:7025-7028 */
        const double x = cx[c], y = cy[c];
        const double vs1 = mesh_area(m, m->x, m->y, c, 0);
        const double vs2 = mesh_area(m, m->x2, m->y2, c, 1);
        const double radius = sqrt(sq(x) + sq(y));
        const double Volume = mesh_volume_rotate(m, m->x, m->y, c, 360.0) / (2.0 * PI_);
        const ora_gas par1 = load(&S, c);
        const int K1 = (m->cell_type[c] == GEL_C_1);
        double Pot[5] = { 0, 0, 0, 0, 0 };
        double time = 10000000;
        for (int k = 0; k < m->nface[c]; k++) {
            const int s = c * ORA_MAXF + k;
            const int A = m->fa[s], B = m->fb[s], t = m->ftype[s];
            const double S_ = sqrt(sq(m->x[B] - m->x[A]) + sq(m->y[B] - m->y[A])) * (m->y[A] + m->y[B]) * 0.5;
            const double fx = (m->x[A] + m->x[B]) / 2.0, fy = (m->y[A] + m->y[B]) / 2.0;
            const double dist = sqrt(sq(x - fx) + sq(y - fy));
            double n1, n2;
            {
                double t1 = m->x[B] - m->x[A], t2 = m->y[B] - m->y[A];
                const double nn = sqrt(sq(t1) + sq(t2));
                t1 = t1 / nn; t2 = t2 / nn;
                n1 = t2; n2 = -t1;
            }
            const double dis = sqrt(sq(fx) + sq(fy));
            ora_gas p11 = par1;
            ora_gas p2 = get_par(&S, c, s, fx, fy);
            int met = prm->flux;
            if (t == GEL_F_USUALY) {                               /* :7064-7091 */
                const int N = m->fnbr[s];
                const double dd = sqrt(sq(cx[N]) + sq(cy[N]));
                p11.ro = p11.ro * sq(radius) / sq(dis);
                p2.ro = p2.ro * sq(dd) / sq(dis);
                p11.Q = p11.Q * sq(radius) / sq(dis);
                p2.Q = p2.Q * sq(dd) / sq(dis);
                p11.p = p11.p * pow(radius / dis, 2.0 * gam);
                p2.p = p2.p * pow(dd / dis, 2.0 * gam);
                ora_polar_perenos(x, y, fx, fy, &p11.u, &p11.v);
                ora_polar_perenos(cx[N], cy[N], fx, fy, &p2.u, &p2.v);
                if (K1 && m->cell_type[N] == GEL_C_1) {
                    double ro1, pr1, ro2, pr2;
                    tvd_radial(&S, c, N, m->fdown[s], fx, fy, &ro1, &pr1);
                    tvd_radial(&S, N, c, m->fup[s], fx, fy, &ro2, &pr2);
                    p11.ro = ro1; p2.ro = ro2;
                    p11.p = pr1; p2.p = pr2;
                }
            } else if (t == GEL_F_INNER_SPHERE) {                 /* :7092-7099 */
                p11 = p2;
            } else if (t == GEL_F_AXIS) {                         /* :7100-7114 */
                p11.ro = p11.ro * sq(radius) / sq(dis);
                p11.Q = p11.Q * sq(radius) / sq(dis);
                p11.p = p11.p * pow(radius / dis, 2.0 * gam);
                ora_polar_perenos(x, y, fx, fy, &p11.u, &p11.v);
                p2 = p11;
                p2.v = -p2.v;
                met = 0;
            }
            double P[4], PQ, Vl, Vc, Vp;
            time = mn(time, ora_hllc2d(&p11, &p2, 0.0, n1, n2, dist, prm->cfl, gam, met, 0, P, &PQ, &Vl, &Vc, &Vp));
            for (int v = 0; v < 4; v++) Pot[v] = Pot[v] + P[v] * S_;
            Pot[4] = Pot[4] + PQ * S_;
        }
        if (time < Tmin) Tmin = time;
        const double u = par1.u, v = par1.v, ro = par1.ro, p = par1.p, Q = par1.Q;
        double q1 = 0.0, q2 = 0.0, q3 = 0.0;
        if (prm->source_mode == GEL_SRC_MK1 && src) {
            double f[GEL_MK1_NFIELDS];
            for (int k = 0; k < GEL_MK1_NFIELDS; k++) f[k] = src[(size_t)k * nc + c];
            ora_source_mk1(prm, f, u, v, ro, p, &q1, &q2, &q3);
        }
        const double T1 = dt_prev;                                 /* :7168-7200 */
        double ro3 = ro - T1 * (Pot[0] / Volume);
        const double Q33 = Q - (T1 / Volume) * Pot[4] - T1;
        if (ro3 <= 0) { ro3 = 0.00001; nrc++; }
        const double u3 = (ro * u - T1 * (Pot[1] / Volume - q1)) / ro3;
        const double v3 = (ro * v - T1 * ((Pot[2] - p * 0.5 * (vs1 + vs2)) / Volume - q2)) / ro3;
        double p3 = (((p / (gam - 1) + ro * (u * u + v * v) * 0.5) - T1 * (Pot[3] / Volume - q3)) -
                     0.5 * ro3 * (u3 * u3 + v3 * v3)) * (gam - 1);
        if (p3 <= 0) { p3 = 0.000001; npc++; }
        next[c] = ro3; next[(size_t)nc + c] = p3; next[(size_t)2 * nc + c] = u3; next[(size_t)3 * nc + c] = v3;
        next[(size_t)4 * nc + c] = Q33;
    }
    free(cx);
    free(cy);
    *dt_new = Tmin;
    if (n_rho_clamp) *n_rho_clamp += nrc;
    if (n_p_clamp) *n_p_clamp += npc;
    return GEL_OK;
}
