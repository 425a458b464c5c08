/* mc_oracle.c - CPU restatement of the reference's Monte Carlo of the interstellar
 * hydrogen, in the form the device tracer implements (gel_mc_run).
 *
 * TEST INFRASTRUCTURE ONLY: tests/ call it through oracle/oracle.py to check the HIP
 * library; nothing in the product links or calls it.
 *
 * What follows the reference step for step (and is checked against it, bit for bit,
 * through oracle/ref_harness.cpp, mode_mcunit):
 *   - the random sensor of sensor.cpp:11-31;
 *   - Setka::Velosity_1..3 (Setka.cpp:15830-15866);
 *   - the exchange-velocity sampler of MKmethod::Change_Velosity4 with no splitting
 *     zones (MKmethod.cpp:2270-2531, I = 0: the call at Setka.cpp:13465).
 * What follows the reference's definition with other arithmetic:
 *   - MKmethod::int_1..3 (MKmethod.cpp:2741-2858) are polynomial fits of three
 *     Maxwellian integrals of |u - v| sigma(|u - v|) times 1, (u - v) and |u - v|^2; here
 *     the integrals are taken by quadrature (mc_tab_build) and agree with the fits to
 *     about 1e-5.
 * What is a restatement of the algorithm of Setka::Fly_exchenge_Imit_Korol
 * (Setka.cpp:12814-13981) on a polar grid:
 *   - a cell's way out from the faces' surfaces of revolution (the quadratics of
 *     :12913-12954), the step cut into pieces of at most one mean free time (:13076-13081),
 *     the optical depth at the piece's midpoint (:13085-13101), the exchange point from
 *     the truncated exponential (:13121-13128), the tallies before and after it
 *     (:13399-13443, :13700-13721) and the source integrals at it (:13408-13432), the
 *     secondary's population from the cell type (:13384-13398), its velocity from the
 *     sampler, its weight roulette;
 *   - the post-processing of MK_start_new (:12241-12449) into the sweep's 19 inputs.
 * Where the device design departs from the reference (documented with gel_mc_run):
 *   - primaries enter through the sphere around the grid from the flux of the drifting
 *     Maxwellian (velocity from the speed-weighted Maxwellian, the point uniform on the
 *     disc across it), not by the zone-split sampler MKmethod::Init_Parametrs;
 *   - secondaries take one velocity from the unsplit kernel (no Change_Velosity4 zones)
 *     and the roulette weight w_min x primary (Setka::Fly_exchenge's weight_) instead of
 *     the zone table pred();
 *   - every atom carries its own sensor (a primary's seeded from the run's seed and its
 *     number, a secondary's from three draws of its parent), so that atoms can be traced
 *     in any order;
 *   - the plasma velocity at a point is (u, v y/r, v z/r), where the reference takes
 *     cos and sin of polar_angle(y, z).
 */
#include "gelios_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MC_PI 3.14159265358979323846      /* pi_ (Help.h:50), pi (Solvers.h:9) */
#define MC_SQRTPI 1.77245385              /* sqrtpi_ (Help.h:51), as the reference rounds it */

/* ------------------------------------------------------------------ sensor ---- */
double ora_sensor_draw(int s[3])
{   /* Sensor::MakeRandom (sensor.cpp:11-31) */
    const int ic15 = 32768, ic10 = 1024;
    const int mz = 710, my = 17784, mx = 11973;
    const double xi = 9.0949470177292824E-13, c = 1.073741824E9;
    int i13 = mz * s[0] + my * s[1] + mx * s[2];
    int i12 = my * s[0] + mx * s[1];
    const int i11 = mx * s[0];
    int ii = i11 / ic15;
    i12 = i12 + ii;
    s[0] = i11 - ic15 * ii;
    ii = i12 / ic15;
    i13 = i13 + ii;
    s[1] = i12 - ic15 * ii;
    s[2] = i13 % ic10;
    return xi * (c * s[2] + ic15 * s[1] + s[0]);
}

/* Seeds: a primary's sensor from the run's seed and its number (splitmix64), a
 * secondary's from three draws of its parent's.  The low word is odd: the sensor is a
 * multiplicative congruential generator modulo 2^40. */
static uint64_t mc_splitmix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
void ora_mc_seed_primary(unsigned long long seed, long long k, int s[3])
{
    const uint64_t z = mc_splitmix(mc_splitmix((uint64_t)seed) + (uint64_t)k);
    s[0] = (int)(z & 0x7fff) | 1;
    s[1] = (int)((z >> 15) & 0x7fff);
    s[2] = (int)((z >> 30) & 0x3ff);
}
static void mc_seed_child(int parent[3], int s[3])
{
    const double d1 = ora_sensor_draw(parent), d2 = ora_sensor_draw(parent), d3 = ora_sensor_draw(parent);
    s[0] = ((int)(d1 * 32768.0) & 0x7fff) | 1;
    s[1] = (int)(d2 * 32768.0) & 0x7fff;
    s[2] = (int)(d3 * 1024.0) & 0x3ff;
}

/* ------------------------------------------------------------------- rates ---- */
void ora_mc_velosity(double u, double cp, double o[3])
{   /* Setka::Velosity_1..3 (Setka.cpp:15830-15866): mean relative speed, momentum and
       energy weights of a Maxwellian of thermal speed cp seen at relative speed u */
    const double pi = MC_PI;
    if (u < 0.00001) {
        o[0] = 2.0 * cp / MC_SQRTPI + 2.0 * u * u / (3.0 * cp * MC_SQRTPI) - u * u * u * u / (15.0 * cp * cp * cp * MC_SQRTPI);
        o[1] = (8.0 / 3.0) * (cp * cp) * (cp * cp) * pi * u + (8.0 / 15.0) * (cp * cp) * pi * u * u * u -
               (4.0 / 105.0) * pi * (u * u) * (u * u) * u;
        o[2] = 8.0 * cp / (3.0 * MC_SQRTPI) + 8.0 * u * u / (9.0 * cp * MC_SQRTPI) -
               44.0 * u * u * u * u / (135.0 * cp * cp * cp * MC_SQRTPI);
    } else {
        o[0] = exp(-u * u / (cp * cp)) * cp / MC_SQRTPI + (u + (cp * cp) / (2.0 * u)) * erf(u / cp);
        o[1] = cp * cp * cp * pi * (exp(-u * u / (cp * cp)) * cp * u * 2.0 * ((cp * cp) + 2.0 * (u * u)) +
                                    MC_SQRTPI * (4.0 * (u * u) * (u * u) + 4.0 * cp * cp * (u * u) - (cp * cp) * (cp * cp)) *
                                        erf(u / cp)) /
               (4.0 * u * u);
        o[2] = exp(-u * u / (cp * cp)) * cp * (5.0 * (cp * cp) + 2.0 * (u * u)) / (MC_SQRTPI * (3.0 * (cp * cp) + 2.0 * (u * u))) +
               (4.0 * (u * u) * (u * u) + 12.0 * cp * cp * (u * u) + 3.0 * (cp * cp) * (cp * cp)) * erf(u / cp) /
                   (2.0 * u * (3.0 * (cp * cp) + 2.0 * (u * u)));
    }
}

/* The nine base integrals at x = u / cp, for psi_k(s) = (ln s)^k, k = 0, 1, 2, over the
 * Maxwellian exp(-|y|^2) seen at relative speed x (s = |x - y|):
 *   f_k = int |x - y| psi_k d^3y                = (pi / x) int s^2 psi_k [E- - E+] ds
 *   g_k = int |x - y| (x - y).x^ psi_k d^3y     = 2 pi int s^4 psi_k [a (E- + E+) - (E- - E+)] / a^2 ds
 *   h_k = int |x - y|^3 psi_k d^3y              = (pi / x) int s^4 psi_k [E- - E+] ds
 * with E-+ = exp(-(s -+ x)^2), a = 2 x s; at x = 0 the limits 4 pi int s^3 psi e^{-s^2},
 * 0, 4 pi int s^5 psi e^{-s^2}.  With sigma(w cp) = (b - a ln w)^2, b = 1 - a ln cp,
 * int_1 = cp / sqrtpi^3 (b^2 f_0 - 2 a b f_1 + a^2 f_2), int_2 = -cp^2 / sqrtpi^3 (...g),
 * int_3 = cp^3 / sqrtpi^3 (...h) (MKmethod.cpp:2741-2858).  Composite Simpson over
 * s in [0, x + 9]. */
#define MC_TAB_PER 128
#define MC_TAB_N (7 * MC_TAB_PER + 4)
static double mc_tab[9][MC_TAB_N];
static int mc_tab_ready;

static void mc_base_quad(double x, double o[9])
{
    const int N = 3000;
    const double smax = x + 9.0, hs = smax / N;
    double acc[9] = { 0 };
    for (int i = 1; i <= N; i++) {
        const double s = i * hs;
        const double w = (i == N) ? 1.0 : (i & 1) ? 4.0 : 2.0;
        const double ls = log(s);
        const double psi[3] = { 1.0, ls, ls * ls };
        const double s2 = s * s, s4 = s2 * s2;
        double kf, kg, kh;
        if (x == 0.0) {
            const double e = exp(-s2);
            kf = 4.0 * MC_PI * s2 * s * e;
            kg = 0.0;
            kh = 4.0 * MC_PI * s4 * s * e;
        } else {
            const double a = 2.0 * x * s;
            double dm, sm;   /* E- - E+, E- + E+ */
            if (a < 1.0) {
                const double e0 = exp(-(x * x + s2));
                dm = e0 * (expm1(a) - expm1(-a));
                sm = e0 * (exp(a) + exp(-a));
            } else {
                const double em = exp(-(s - x) * (s - x)), ep = exp(-(s + x) * (s + x));
                dm = em - ep;
                sm = em + ep;
            }
            kf = MC_PI / x * s2 * dm;
            kh = MC_PI / x * s4 * dm;
            if (a < 0.1) {
                const double e0 = exp(-(x * x + s2));
                kg = 2.0 * MC_PI * s4 * e0 * 2.0 * (a / 3.0 + a * a * a / 30.0 + a * a * a * a * a / 840.0);
            } else {
                kg = 2.0 * MC_PI * s4 * (a * sm - dm) / (a * a);
            }
        }
        for (int k = 0; k < 3; k++) {
            acc[k] += w * kf * psi[k];
            acc[3 + k] += w * kg * psi[k];
            acc[6 + k] += w * kh * psi[k];
        }
    }
    for (int m = 0; m < 9; m++) o[m] = acc[m] * hs / 3.0;
}

static void mc_tab_build(void)
{
    if (mc_tab_ready) return;
    for (int k = 0; k < MC_TAB_N; k++) {
        double o[9];
        mc_base_quad((double)k / MC_TAB_PER, o);
        for (int m = 0; m < 9; m++) mc_tab[m][k] = o[m];
    }
    mc_tab_ready = 1;
}

/* Cubic (four-point Lagrange) interpolation of the base integrals, x in [0, 7]. */
static void mc_base(double x, double o[9])
{
    const double t = x * MC_TAB_PER;
    int k = (int)t;
    if (k < 1) k = 1;
    if (k > MC_TAB_N - 3) k = MC_TAB_N - 3;
    const double r = t - k;    /* nodes k-1, k, k+1, k+2 at -1, 0, 1, 2 */
    const double w0 = -r * (r - 1.0) * (r - 2.0) / 6.0, w1 = (r + 1.0) * (r - 1.0) * (r - 2.0) / 2.0;
    const double w2 = -(r + 1.0) * r * (r - 2.0) / 2.0, w3 = (r + 1.0) * r * (r - 1.0) / 6.0;
    for (int m = 0; m < 9; m++)
        o[m] = w0 * mc_tab[m][k - 1] + w1 * mc_tab[m][k] + w2 * mc_tab[m][k + 1] + w3 * mc_tab[m][k + 2];
}

void ora_mc_int(double u, double cp, double sig_a, double o[3])
{
    mc_tab_build();
    double B[9];
    mc_base(u / cp, B);
    const double b = 1.0 - sig_a * log(cp);
    const double c0 = b * b, c1 = -2.0 * sig_a * b, c2 = sig_a * sig_a;
    const double n3 = MC_SQRTPI * MC_SQRTPI * MC_SQRTPI;
    o[0] = (cp / n3) * (c0 * B[0] + c1 * B[1] + c2 * B[2]);
    o[1] = -(cp * cp / n3) * (c0 * B[3] + c1 * B[4] + c2 * B[5]);
    o[2] = (cp * cp * cp / n3) * (c0 * B[6] + c1 * B[7] + c2 * B[8]);
}

/* ----------------------------------------------------------------- sampler ---- */
/* MKmethod::Change_Velosity4, I = 0 (MKmethod.cpp:2270-2531): the plasma proton a
 * hydrogen atom of velocity V exchanges with, from the Maxwellian around U weighted by
 * the relative speed times the cross section.  Velocities in units of cp.  sigma_on = 0
 * drops the cross section (the speed-weighted Maxwellian of the inflow). */
static int mc_sample(int s[3], const double U[3], const double V[3], double cp, double sig_a, int sigma_on, double W[3])
{
    const double pi_ = MC_PI;
    const double X = sqrt((V[0] - U[0]) * (V[0] - U[0]) + (V[1] - U[1]) * (V[1] - U[1]) + (V[2] - U[2]) * (V[2] - U[2]));
    const double p4 = 0.5 * MC_SQRTPI * X / (1.0 + 0.5 * MC_SQRTPI * X);
    double v1, v2, v3, h, ksi6;
    int tries = 0;
    do {
        const double ksi1 = ora_sensor_draw(s), ksi2 = ora_sensor_draw(s), ksi3 = ora_sensor_draw(s);
        const double ksi4 = ora_sensor_draw(s), ksi5 = ora_sensor_draw(s);
        ksi6 = ora_sensor_draw(s);
        double y1, y2, y3;
        if (p4 < ksi1) {
            const double om1 = 1.0 - 2.0 * ksi4;
            const double om2 = sqrt(1.0 - om1 * om1) * cos(2.0 * pi_ * ksi5);
            const double om3 = sqrt(1.0 - om1 * om1) * sin(2.0 * pi_ * ksi5);
            const double lo = sqrt(-log(ksi2 * ksi3));
            y1 = lo * om1;
            y2 = lo * om2;
            y3 = lo * om3;
        } else {
            y1 = sqrt(-log(ksi2)) * cos(pi_ * ksi3);
            y2 = sqrt(-log(ksi4)) * cos(2.0 * pi_ * ksi5);
            y3 = sqrt(-log(ksi4)) * sin(2.0 * pi_ * ksi5);
        }
        v1 = y1 + U[0];
        v2 = y2 + U[1];
        v3 = y3 + U[2];
        const double u1 = V[0] - v1, u2 = V[1] - v2, u3 = V[2] - v3;
        const double uuu = sqrt(u1 * u1 + u2 * u2 + u3 * u3);
        const double yy = sqrt(y1 * y1 + y2 * y2 + y3 * y3);
        if (sigma_on) {
            const double su = 1.0 - sig_a * log(uuu * cp), sx = 1.0 - sig_a * log(X * cp);
            h = (uuu * (su * su)) / ((sx * sx) * (X + yy));
        } else {
            h = uuu / (X + yy);
        }
    } while (h < ksi6 && ++tries < 100000);   /* the device's bound on the rejections */
    W[0] = v1;
    W[1] = v2;
    W[2] = v3;
    return (v1 >= 0.0 || v2 * v2 + v3 * v3 > 0.0 * (v1 * v1)) ? 1 : 0;
}

int ora_mc_exchange(int s[3], const double in[7], double sig_a, double out[4])
{
    double W[3];
    const int mu = mc_sample(s, in, in + 3, in[6], sig_a, 1, W);
    out[0] = W[0];
    out[1] = W[1];
    out[2] = W[2];
    out[3] = mu;
    return mu;
}

/* ---------------------------------------------------------------- geometry ---- */
/* Smallest t > tmin at which p + t v meets the surface swept by the segment
 * (xa, ra) - (xb, rb) of the (x, r) half plane turning about the x axis; -1 if none.
 * The quadratic (dx r)^2 = (dx ra + dr (x - xa))^2 of Setka.cpp:12913-12930 in the
 * segment's own terms; roots on the mirror cone (r < 0) or off the segment are dropped. */
static double mc_hit(const double p[3], const double v[3], double xa, double ra, double xb, double rb, double tmin)
{
    const double dx = xb - xa, dr = rb - ra;
    const double len2 = dx * dx + dr * dr, len = sqrt(len2);
    double t[2];
    int nt = 0;
    if (fabs(dx) <= 1e-13 * len) {
        if (v[0] != 0.0) t[nt++] = (xa - p[0]) / v[0];
    } else {
        const double c0 = dx * ra + dr * (p[0] - xa), c1 = dr * v[0], dx2 = dx * dx;
        const double A = dx2 * (v[1] * v[1] + v[2] * v[2]) - c1 * c1;
        const double B = 2.0 * (dx2 * (p[1] * v[1] + p[2] * v[2]) - c0 * c1);
        const double C = dx2 * (p[1] * p[1] + p[2] * p[2]) - c0 * c0;
        if (fabs(A) <= 1e-13 * dx2 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])) {
            if (B != 0.0) t[nt++] = -C / B;
        } else {
            const double D = B * B - 4.0 * A * C;
            if (D >= 0.0) {
                const double q = -0.5 * (B + (B >= 0.0 ? sqrt(D) : -sqrt(D)));
                t[nt++] = q / A;
                if (q != 0.0) t[nt++] = C / q;
            }
        }
    }
    double best = -1.0;
    for (int k = 0; k < nt; k++) {
        const double tk = t[k];
        if (!(tk > tmin)) continue;
        const double xh = p[0] + tk * v[0];
        const double yh = p[1] + tk * v[1], zh = p[2] + tk * v[2];
        const double rh = sqrt(yh * yh + zh * zh);
        const double s = ((xh - xa) * dx + (rh - ra) * dr) / len2;
        if (s < -1e-9 || s > 1.0 + 1e-9) continue;
        const double off = (dr * (xh - xa) - dx * (rh - ra)) / len;
        if (fabs(off) > 1e-7 * len) continue;
        if (best < 0.0 || tk < best) best = tk;
    }
    return best;
}

/* ------------------------------------------------------------------ tracer ---- */
typedef struct {
    double p[3], v[3], mu;
    int i, j, area, s[3];
} mc_atom;

typedef struct {
    int ni, nj;
    const double *x, *y, *prim;
    const int8_t* pop;
    size_t nc;
    const gel_mc_params* prm;
    double* tal;          /* [22][nc] */
    mc_atom* stack;
    long long nstack, cap;
    gel_mc_stats* st;
    double wmin;
    /* a general mesh: corners and face slots per cell, each slot's way on (the next
     * cell, -1 out of the grid, -2 across the inner sphere, -3 an Axis slot, skipped),
     * the boundary lists as (A, B, cell) triples (outer: entry; inner: inner sphere) */
    int mesh;
    const int *coff, *corners, *soff, *sA, *sB, *scode;
    int *bout, *bin, nbout, nbin;
} mc_run;

static inline int mc_nd(const mc_run* R, int q, int c) { return q * (R->nj + 1) + c; }

/* exit of cell (i, j): face 0 below (row i), 1 above, 2 left (column j), 3 right */
static double mc_exit(const mc_run* R, int i, int j, const double p[3], const double v[3], double tmin, int* face)
{
    const int a = mc_nd(R, i, j), b = mc_nd(R, i, j + 1), c = mc_nd(R, i + 1, j), d = mc_nd(R, i + 1, j + 1);
    const double* X = R->x;
    const double* Y = R->y;
    double best = -1.0;
    *face = -1;
    const double t0 = mc_hit(p, v, X[a], Y[a], X[b], Y[b], tmin);
    if (t0 > 0.0) { best = t0; *face = 0; }
    const double t1 = mc_hit(p, v, X[c], Y[c], X[d], Y[d], tmin);
    if (t1 > 0.0 && (best < 0.0 || t1 < best)) { best = t1; *face = 1; }
    if (j > 0) {
        const double t2 = mc_hit(p, v, X[a], Y[a], X[c], Y[c], tmin);
        if (t2 > 0.0 && (best < 0.0 || t2 < best)) { best = t2; *face = 2; }
    }
    if (j < R->nj - 1) {
        const double t3 = mc_hit(p, v, X[b], Y[b], X[d], Y[d], tmin);
        if (t3 > 0.0 && (best < 0.0 || t3 < best)) { best = t3; *face = 3; }
    }
    return best;
}

/* first crossing of the ray with the faces of node row q (q = ni: the grid's outside,
 * q = 0: the inner sphere); the column through *jj */
static double mc_row_hit(const mc_run* R, int q, const double p[3], const double v[3], double tmin, int* jj)
{
    double best = -1.0;
    *jj = -1;
    for (int c = 0; c < R->nj; c++) {
        const int a = mc_nd(R, q, c), b = mc_nd(R, q, c + 1);
        const double t = mc_hit(p, v, R->x[a], R->y[a], R->x[b], R->y[b], tmin);
        if (t > 0.0 && (best < 0.0 || t < best)) { best = t; *jj = c; }
    }
    return best;
}

/* mesh cell c: the nearest crossing over its face slots in Cell::Grans order, Axis
 * slots skipped (Setka.cpp:12893-12960); *slot = the slot crossed */
static double mc_mesh_exit(const mc_run* R, int c, const double p[3], const double v[3], double tmin, int* slot)
{
    double best = -1.0;
    *slot = -1;
    for (int s = R->soff[c]; s < R->soff[c + 1]; s++) {
        if (R->scode[s] == -3) continue;
        const int A = R->sA[s], B = R->sB[s];
        const double t = mc_hit(p, v, R->x[A], R->y[A], R->x[B], R->y[B], tmin);
        if (t > 0.0 && (best < 0.0 || t < best)) { best = t; *slot = s; }
    }
    return best;
}

/* first crossing with n (A, B, cell) faces; *jj = the entry, -1 if none */
static double mc_list_hit(const mc_run* R, const int* L, int n, const double p[3], const double v[3], double tmin,
                          int* jj)
{
    double best = -1.0;
    *jj = -1;
    for (int c = 0; c < n; c++) {
        const double t = mc_hit(p, v, R->x[L[3 * c]], R->y[L[3 * c]], R->x[L[3 * c + 1]], R->y[L[3 * c + 1]], tmin);
        if (t > 0.0 && (best < 0.0 || t < best)) { best = t; *jj = c; }
    }
    return best;
}

static void mc_push(mc_run* R, const mc_atom* a)
{
    if (R->nstack == R->cap) {
        R->cap = R->cap ? 2 * R->cap : 1024;
        R->stack = (mc_atom*)realloc(R->stack, (size_t)R->cap * sizeof(mc_atom));
    }
    R->stack[R->nstack++] = *a;
}

static inline double mc_sig(double sig_a, double u)
{
    const double t = 1.0 - sig_a * log(u);
    return t * t;
}

/* plasma velocity of cell k at the point (y, z) off the axis */
static inline void mc_plasma(const mc_run* R, size_t k, double y, double z, double U[3], double* ca, double* sa)
{
    const double r = sqrt(y * y + z * z);
    *ca = r > 0.0 ? y / r : 1.0;
    *sa = r > 0.0 ? z / r : 0.0;
    const double u = R->prim[2 * R->nc + k], v = R->prim[3 * R->nc + k];
    U[0] = u;
    U[1] = v * *ca;
    U[2] = v * *sa;
}

static void mc_trace(mc_run* R, mc_atom a)
{
    const gel_mc_params* P = R->prm;
    const size_t nc = R->nc;
    double* T = R->tal;
    R->st->atoms++;
    for (int seg = 0; seg < P->max_segments; seg++) {
        const size_t k = R->mesh ? (size_t)a.i : (size_t)a.i * R->nj + a.j;
        const double vn = sqrt(a.v[0] * a.v[0] + a.v[1] * a.v[1] + a.v[2] * a.v[2]);
        const double* X = R->x;
        const double* Y = R->y;
        /* a diagonal of the cell: corners 0 and 2 */
        const int n00 = R->mesh ? R->corners[R->coff[k]] : mc_nd(R, a.i, a.j);
        const int n11 = R->mesh ? R->corners[R->coff[k] + 2] : mc_nd(R, a.i + 1, a.j + 1);
        const double L = sqrt((X[n11] - X[n00]) * (X[n11] - X[n00]) + (Y[n11] - Y[n00]) * (Y[n11] - Y[n00]));
        const double tmin = 1e-9 * L / vn;
        int face;
        const double t_exit = R->mesh ? mc_mesh_exit(R, a.i, a.p, a.v, tmin, &face)
                                      : mc_exit(R, a.i, a.j, a.p, a.v, tmin, &face);
        if (face < 0) { R->st->lost++; return; }
        R->st->segments++;
        const double ro = R->prim[k], pr = R->prim[nc + k];
        const double cp = sqrt(pr / ro);
        double time = t_exit;
        int ch_now = 0;
        {   /* pieces of at most one mean free time (Setka.cpp:13076-13081) */
            double o[3];
            ora_mc_velosity(0.0, cp, o);
            const double nu0 = (ro * o[0] * mc_sig(P->sig_a, o[0])) / P->kn;
            if (time > 1.0 / nu0) {
                time = time / ((int)(time / (1.0 / nu0)) + 1);
                ch_now = 1;
            }
        }
        double U[3], ca, sa;
        mc_plasma(R, k, a.p[1] + 0.5 * time * a.v[1], a.p[2] + 0.5 * time * a.v[2], U, &ca, &sa);
        double w1 = a.v[0] - U[0], w2 = a.v[1] - U[1], w3 = a.v[2] - U[2];
        double u = sqrt(w1 * w1 + w2 * w2 + w3 * w3);
        double nu;
        if (u / cp > 7.0) {
            double o[3];
            ora_mc_velosity(u, cp, o);
            nu = (ro * o[0] * mc_sig(P->sig_a, o[0])) / P->kn;
        } else {
            double o[3];
            ora_mc_int(u, cp, P->sig_a, o);
            nu = (ro * o[0]) / P->kn;
        }
        const double kappa = nu * time;
        const double xi = ora_sensor_draw(a.s);
        const double t_ex = -(time / kappa) * log(1.0 - xi * (1.0 - exp(-kappa)));
        const double mu_ex = a.mu * (1.0 - exp(-kappa));
        const double mu2 = a.mu - mu_ex;
        const double pe[3] = { a.p[0] + t_ex * a.v[0], a.p[1] + t_ex * a.v[1], a.p[2] + t_ex * a.v[2] };
        const double V2 = a.v[0] * a.v[0] + a.v[1] * a.v[1] + a.v[2] * a.v[2];
        const int h = a.area;
        /* the weight before the exchange point, the rest after it (:13399-13443, :13700-13721) */
        {
            double U1[3], c1, s1, U2[3], c2, s2;
            mc_plasma(R, k, a.p[1] + 0.5 * t_ex * a.v[1], a.p[2] + 0.5 * t_ex * a.v[2], U1, &c1, &s1);
            const double t2 = time - t_ex;
            mc_plasma(R, k, pe[1] + 0.5 * t2 * a.v[1], pe[2] + 0.5 * t2 * a.v[2], U2, &c2, &s2);
            const double wn = t_ex * a.mu + t2 * mu2;
            T[(0 + h) * nc + k] += wn;
            T[(4 + h) * nc + k] += a.v[0] * wn;
            T[(8 + h) * nc + k] += t_ex * (a.v[1] * c1 + a.v[2] * s1) * a.mu + t2 * (a.v[1] * c2 + a.v[2] * s2) * mu2;
            T[(12 + h) * nc + k] += V2 * wn;
        }
        /* the source integrals at the exchange point (:13408-13432) */
        mc_plasma(R, k, pe[1], pe[2], U, &ca, &sa);
        w1 = a.v[0] - U[0];
        w2 = a.v[1] - U[1];
        w3 = a.v[2] - U[2];
        u = sqrt(w1 * w1 + w2 * w2 + w3 * w3);
        const double u1 = -w1, u2 = -w2, u3 = -w3;
        const double skalar = a.v[0] * u1 + a.v[1] * u2 + a.v[2] * u3;
        if (u > 0.0) {
            if (u / cp > 7.0) {
                double o[3];
                ora_mc_velosity(u, cp, o);
                const double uz = o[0];
                const double uz_M = o[1] / (uz * (cp * cp) * cp * MC_PI * MC_SQRTPI);
                const double uz_E = o[2];
                T[16 * nc + k] += -mu_ex * uz_M * u1 / u;
                T[17 * nc + k] += -mu_ex * uz_M * (u2 * ca + u3 * sa) / u;
                T[18 * nc + k] += mu_ex * (-0.25 * (3.0 * (cp * cp) + 2.0 * (u * u)) * (uz_E / uz) - uz_M * skalar / u);
            } else {
                double o[3];
                ora_mc_int(u, cp, P->sig_a, o);
                T[16 * nc + k] += mu_ex * (o[1] / o[0]) * u1 / u;
                T[17 * nc + k] += mu_ex * (o[1] / o[0]) * (u2 * ca + u3 * sa) / u;
                T[18 * nc + k] += mu_ex * (-0.5 * o[2] / o[0] + o[1] / o[0] * skalar / u);
            }
        }
        /* the secondary (:13445-13481): population of the cell, kernel velocity,
           roulette at w_min primary weights */
        {
            double Un[3] = { U[0] / cp, U[1] / cp, U[2] / cp }, Vn[3] = { a.v[0] / cp, a.v[1] / cp, a.v[2] / cp };
            double W[3];
            mc_sample(a.s, Un, Vn, cp, P->sig_a, 1, W);
            W[0] *= cp;
            W[1] *= cp;
            W[2] *= cp;
            double mu3 = mu_ex;
            int keep = 1;
            if (!(mu3 >= R->wmin)) {
                if (mu3 >= ora_sensor_draw(a.s) * R->wmin) mu3 = R->wmin;
                else keep = 0;
            }
            if (keep) {
                T[19 * nc + k] += mu3 * (a.v[0] - W[0]);
                T[20 * nc + k] += mu3 * ((a.v[1] - W[1]) * ca + (a.v[2] - W[2]) * sa);
                T[21 * nc + k] += mu3 * 0.5 * (V2 - (W[0] * W[0] + W[1] * W[1] + W[2] * W[2]));
                mc_atom c;
                c.p[0] = pe[0]; c.p[1] = pe[1]; c.p[2] = pe[2];
                c.v[0] = W[0]; c.v[1] = W[1]; c.v[2] = W[2];
                c.mu = mu3;
                c.i = a.i;
                c.j = a.j;
                c.area = R->pop[k];
                mc_seed_child(a.s, c.s);
                mc_push(R, &c);
            }
        }
        /* on: the rest of the cell, or the next cell */
        a.mu = mu2;
        if (ch_now) {
            for (int d = 0; d < 3; d++) a.p[d] += time * a.v[d];
            continue;
        }
        for (int d = 0; d < 3; d++) a.p[d] += t_exit * a.v[d];
        if (R->mesh) {
            const int to = R->scode[face];
            if (to >= 0) {
                a.i = to;
            } else if (to == -2) {                     /* through the inner sphere */
                int jj;
                const double th = mc_list_hit(R, R->bin, R->nbin, a.p, a.v, tmin, &jj);
                if (jj < 0) { R->st->lost++; return; }
                for (int d = 0; d < 3; d++) a.p[d] += th * a.v[d];
                a.i = R->bin[3 * jj + 2];
            } else {
                return;                                /* out of the grid */
            }
        } else if (face == 1) {
            if (a.i == R->ni - 1) return;              /* out of the grid */
            a.i++;
        } else if (face == 0) {
            if (a.i == 0) {                            /* through the inner sphere */
                int jj;
                const double th = mc_row_hit(R, 0, a.p, a.v, tmin, &jj);
                if (jj < 0) { R->st->lost++; return; }
                for (int d = 0; d < 3; d++) a.p[d] += th * a.v[d];
                a.j = jj;
            } else {
                a.i--;
            }
        } else if (face == 2) {
            a.j--;
        } else {
            a.j++;
        }
    }
    R->st->lost++;
}

/* primary k: velocity from the speed-weighted Maxwellian, entry point uniform on the
 * disc across it of the sphere of radius Rs around the grid, then straight to the grid */
static int mc_primary(mc_run* R, long long kk, double Rs, mc_atom* a)
{
    const gel_mc_params* P = R->prm;
    ora_mc_seed_primary(P->seed, kk, a->s);
    const double U[3] = { P->v_inf, 0.0, 0.0 }, Z[3] = { 0.0, 0.0, 0.0 };
    double v[3];
    mc_sample(a->s, U, Z, 1.0, P->sig_a, 0, v);
    const double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double e[3] = { v[0] / vn, v[1] / vn, v[2] / vn };
    /* a unit vector across e from the axis least aligned with it */
    double q[3] = { 0.0, 0.0, 0.0 };
    const double ax = fabs(e[0]), ay = fabs(e[1]), az = fabs(e[2]);
    if (ax <= ay && ax <= az) q[0] = 1.0;
    else if (ay <= az) q[1] = 1.0;
    else q[2] = 1.0;
    double e1[3] = { e[1] * q[2] - e[2] * q[1], e[2] * q[0] - e[0] * q[2], e[0] * q[1] - e[1] * q[0] };
    const double n1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    e1[0] /= n1; e1[1] /= n1; e1[2] /= n1;
    const double e2[3] = { e[1] * e1[2] - e[2] * e1[1], e[2] * e1[0] - e[0] * e1[2], e[0] * e1[1] - e[1] * e1[0] };
    const double rho = Rs * sqrt(ora_sensor_draw(a->s));
    const double phi = 2.0 * MC_PI * ora_sensor_draw(a->s);
    const double c = rho * cos(phi), s = rho * sin(phi);
    const double back = sqrt(Rs * Rs - rho * rho);
    double p[3];
    for (int d = 0; d < 3; d++) p[d] = c * e1[d] + s * e2[d] - back * e[d];
    int jj;
    const double t = R->mesh ? mc_list_hit(R, R->bout, R->nbout, p, v, 0.0, &jj) : mc_row_hit(R, R->ni, p, v, 0.0, &jj);
    if (jj < 0) return 0;
    for (int d = 0; d < 3; d++) { a->p[d] = p[d] + t * v[d]; a->v[d] = v[d]; }
    a->mu = 1.0;
    a->i = R->mesh ? R->bout[3 * jj + 2] : R->ni - 1;
    a->j = R->mesh ? 0 : jj;
    a->area = 3;
    return 1;
}

double ora_mc_inflow(double v_inf)
{   /* mean speed of the drifting Maxwellian (thermal speed 1): Velosity_1(|v_inf|, 1) */
    double o[3];
    ora_mc_velosity(fabs(v_inf), 1.0, o);
    return o[0];
}

static void mc_run_all(mc_run* RR, int nn, double* out);

int ora_mc_run(int ni, int nj, const double* x, const double* y, const double* prim, const int8_t* pop,
               const gel_mc_params* prm, double* out, gel_mc_stats* st)
{
    mc_run R;
    memset(&R, 0, sizeof(R));
    memset(st, 0, sizeof(*st));
    R.ni = ni;
    R.nj = nj;
    R.x = x;
    R.y = y;
    R.prim = prim;
    R.pop = pop;
    R.nc = (size_t)ni * nj;
    R.prm = prm;
    R.st = st;
    R.wmin = prm->w_min;
    mc_run_all(&R, (ni + 1) * (nj + 1), out);
    return 0;
}

/* The tracer on a general mesh: cell c's corners are corners[coff[c] .. coff[c + 1]),
 * its face slots soff[c] .. soff[c + 1]) in Cell::Grans order with nodes sA, sB as the
 * cell sees them and their way on scode (the next cell, -1 out of the grid, -2 across
 * the inner sphere, -3 an Axis slot).  Primaries enter by the -1 slots. */
int ora_mc_run_mesh(int nc, int nn, const double* x, const double* y, const int* coff, const int* corners,
                    const int* soff, const int* sA, const int* sB, const int* scode, const double* prim,
                    const int8_t* pop, const gel_mc_params* prm, double* out, gel_mc_stats* st)
{
    mc_run R;
    memset(&R, 0, sizeof(R));
    memset(st, 0, sizeof(*st));
    R.x = x;
    R.y = y;
    R.prim = prim;
    R.pop = pop;
    R.nc = (size_t)nc;
    R.prm = prm;
    R.st = st;
    R.wmin = prm->w_min;
    R.mesh = 1;
    R.coff = coff; R.corners = corners; R.soff = soff; R.sA = sA; R.sB = sB; R.scode = scode;
    const int ns = soff[nc];
    R.bout = (int*)malloc((3 * (size_t)ns + 1) * sizeof(int));
    R.bin = (int*)malloc((3 * (size_t)ns + 1) * sizeof(int));
    for (int c = 0; c < nc; c++)
        for (int s = soff[c]; s < soff[c + 1]; s++) {
            if (scode[s] != -1 && scode[s] != -2) continue;
            int* L = scode[s] == -1 ? R.bout + 3 * R.nbout++ : R.bin + 3 * R.nbin++;
            L[0] = sA[s]; L[1] = sB[s]; L[2] = c;
        }
    mc_run_all(&R, nn, out);
    free(R.bout);
    free(R.bin);
    return 0;
}

static void mc_run_all(mc_run* RR, int nn, double* out)
{
    mc_run R = *RR;
    const gel_mc_params* prm = R.prm;
    gel_mc_stats* st = R.st;
    const double *x = R.x, *y = R.y, *prim = R.prim;
    const int nj = R.nj;
    const size_t nc = R.nc;
    R.tal = (double*)calloc(22 * nc, sizeof(double));
    double Rs = 0.0;
    for (int k = 0; k < nn; k++) Rs = fmax(Rs, sqrt(x[k] * x[k] + y[k] * y[k]));
    Rs *= 1.0 + 1e-12;
    for (long long kk = 0; kk < prm->n_primary; kk++) {
        mc_atom a;
        if (!mc_primary(&R, kk, Rs, &a)) continue;
        st->primaries_in++;
        mc_trace(&R, a);
        while (R.nstack > 0) {
            const mc_atom c = R.stack[--R.nstack];
            mc_trace(&R, c);
        }
    }
    /* post-processing (Setka.cpp:12241-12449) */
    const double Phi = MC_PI * Rs * Rs * ora_mc_inflow(prm->v_inf);
    double* T = R.tal;
    for (size_t k = 0; k < nc; k++) {
        double vol;
        if (R.mesh) {   /* Cell::Get_Volume_rotate(360) (Cell.cpp:188-206) over the contour */
            const int* q = R.corners + R.coff[k];
            const int n = R.coff[k + 1] - R.coff[k];
            double A = 0.0, G = 0.0;
            for (int m = 0; m < n; m++) {
                const int a = q[m], b = q[m + 1 == n ? 0 : m + 1];
                const double cr = x[a] * y[b] - y[a] * x[b];
                A = A + cr;
                G = G + (y[a] + y[b]) * cr;
            }
            A = A * 0.5;
            vol = fabs(A) * (1.0 / (6.0 * A)) * G * MC_PI * 360.0 / 180.0;
        } else {
            const int i = (int)(k / nj), j = (int)(k % nj);
            const int q[4] = { i * (nj + 1) + j, (i + 1) * (nj + 1) + j, (i + 1) * (nj + 1) + j + 1, i * (nj + 1) + j + 1 };
            vol = ora_cell_volume_rotate(x, y, q, 360.0);
        }
        const double no = (double)prm->n_primary * vol;
        double Hn[4], Hu[4], Hv[4], HT[4];
        double Fn = 0.0, Fu = 0.0, Fv = 0.0, FT = 0.0;
        for (int h = 0; h < 4; h++) {
            Hn[h] = T[(0 + h) * nc + k];
            Hu[h] = T[(4 + h) * nc + k];
            Hv[h] = T[(8 + h) * nc + k];
            HT[h] = T[(12 + h) * nc + k];
            Fn += Hn[h]; Fu += Hu[h]; Fv += Hv[h]; FT += HT[h];
        }
        double I[3], II[3];
        for (int m = 0; m < 3; m++) {
            I[m] = prm->n_p_lism * (Phi * T[(16 + m) * nc + k] / no);
            II[m] = prm->n_p_lism * (Phi * T[(19 + m) * nc + k] / no);
        }
        if (Fn > 0.0) {
            Fu = Fu / Fn;
            Fv = Fv / Fn;
            FT = (2.0 / 3.0) * (FT / Fn - (Fu * Fu + Fv * Fv + 0.0 * 0.0));
            for (int h = 0; h < 4; h++) {
                if (Hn[h] > 0.0) {
                    Hu[h] = Hu[h] / Hn[h];
                    Hv[h] = Hv[h] / Hn[h];
                    HT[h] = (2.0 / 3.0) * (HT[h] / Hn[h] - (Hu[h] * Hu[h] + Hv[h] * Hv[h] + 0.0 * 0.0));
                } else {
                    Hu[h] = 0.0; Hv[h] = 0.0; HT[h] = 0.0;
                }
            }
        } else {
            Fn = 0.0; Fu = 0.0; Fv = 0.0; FT = 0.0;
        }
        Fn = Phi * Fn / no;
        for (int h = 0; h < 4; h++) Hn[h] = Phi * Hn[h] / no;
        /* the closure with these moments (:12298-12393) and k = I / M (:12395-12449) */
        double u = prim[2 * nc + k], v = prim[3 * nc + k], ro = prim[k], p = prim[nc + k];
        if (ro <= 0.0) { ro = 0.0000001; p = 0.0; }
        double Mu = 0.0, Mv = 0.0, MT = 0.0;
        for (int h = 0; h < 4; h++) {
            double roH = Hn[h], pH = 0.5 * HT[h] * Hn[h];
            if (roH <= 0.0) { roH = 0.0000001; pH = 0.0; }
            const double UM = sqrt((u - Hu[h]) * (u - Hu[h]) + (v - Hv[h]) * (v - Hv[h]) +
                                   (64.0 / (9.0 * MC_PI)) * (p / ro + 2.0 * pH / roH));
            const double UU = sqrt((u - Hu[h]) * (u - Hu[h]) + (v - Hv[h]) * (v - Hv[h]) + (4.0 / MC_PI) * (p / ro + 2.0 * pH / roH));
            const double sg = (1.0 - prm->sig_a * log(UM)) * (1.0 - prm->sig_a * log(UM));
            const double nuh = ro * roH * UM * sg;
            Mu += nuh * (Hu[h] - u);
            Mv += nuh * (Hv[h] - v);
            MT += nuh * (((Hu[h] * Hu[h]) + (Hv[h] * Hv[h]) - (u * u) - (v * v)) / 2.0 + (UU / UM) * (2.0 * pH / roH - p / ro));
        }
        const double coef = prm->n_p_lism / prm->kn;
        Mu *= coef; Mv *= coef; MT *= coef;
        double kf[3];
        const double M[3] = { Mu, Mv, MT };
        for (int m = 0; m < 3; m++) {
            kf[m] = 1.0;
            if (fabs(M[m]) > 0.0001) {
                kf[m] = I[m] / M[m];
                if (kf[m] > 3.0 || kf[m] < 0.33) kf[m] = 1.0;
            }
        }
        for (int h = 0; h < 4; h++) {
            out[(0 + h) * nc + k] = Hn[h];
            out[(4 + h) * nc + k] = Hu[h];
            out[(8 + h) * nc + k] = Hv[h];
            out[(12 + h) * nc + k] = HT[h];
        }
        for (int m = 0; m < 3; m++) {
            out[(16 + m) * nc + k] = kf[m];
            out[(19 + m) * nc + k] = I[m];
            out[(22 + m) * nc + k] = II[m];
        }
        out[25 * nc + k] = Fn;
        out[26 * nc + k] = Fu;
        out[27 * nc + k] = Fv;
        out[28 * nc + k] = FT;
    }
    free(R.tal);
    free(R.stack);
}
