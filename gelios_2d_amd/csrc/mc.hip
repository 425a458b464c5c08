// mc.hip - the Monte Carlo of the interstellar hydrogen on the device: the atoms'
// trajectories through the grid and the per-cell moments and charge-exchange sources
// they leave (the reference's Setka::MK_start_new, Setka.cpp:11989-12535, with its tracer
// Setka::Fly_exchenge_Imit_Korol, :12814-13981, and the post-processing, :12241-12449).
//
// Design.  One thread per atom, traced to the end of its way through the grid.  An
// exchange in a cell leaves a secondary atom: it goes to a queue (atomic slot), and the
// next launch traces the queue, generation by generation, until no secondary is left.
// Primaries come in batches so that the queues stay bounded.  Every atom carries its own
// random sensor (the reference's generator, sensor.cpp:11-31), seeded from the run's seed
// and the primary's number, or from three draws of its parent, so the result does not
// depend on the order the device traces the atoms in.  Cell tallies are FP64 atomic sums.
// The charge-exchange rate integrals int_1..3 (MKmethod.cpp:2741-2858) come from a table
// of their defining integrals built on the device by quadrature (table_kernel) and
// interpolated to fourth order.  The tracing kernels are built for three wavefronts per
// SIMD (168 registers, about 270 B of spills per lane): 17 % more segments per second than
// at two (no spills) on the 200 x 100 heliosphere, the atoms' latency-bound walks gaining
// more from the third wavefront than the spills cost.
#include "gelios_internal.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace gel {
namespace mc {

constexpr double kPiM = 3.14159265358979323846;   // pi_ (Help.h:50)
constexpr double kSqrtPiR = 1.77245385;           // sqrtpi_ (Help.h:51), as the reference rounds it
constexpr int kTabPer = 128;                      // table nodes per unit of u / cp
constexpr int kTabN = 7 * kTabPer + 4;            // u / cp in [0, 7 + 3 / kTabPer]
constexpr int kTal = 22;                          // tallies per cell
constexpr int kCnt = 5;                           // primaries in, atoms, segments, lost, dropped
constexpr int kMaxGenerations = 4096;             // bound on the secondary generations of a batch
constexpr int kMaxTries = 100000;                 // bound on the sampler's rejections (a proposal is
                                                  // accepted with probability of order one)

// i, j: the cell on a polar grid; on a general mesh i is the cell and j is 0
struct Atom {
    double p[3], v[3], mu;
    int i, j, area, s[3];
};

struct Args {
    int ni, nj;
    long long nc;
    const double* __restrict__ x;     // nodes
    const double* __restrict__ y;
    const double* __restrict__ pl;    // [5][nc] ro, cp, u, v, p of the plasma
    const int8_t* __restrict__ pop;   // [nc] population of a secondary born in the cell
    const double* __restrict__ tab;   // [9][kTabN] base integrals
    double* tal;                      // [kTal][nc]
    double sig_a, kn, wmin, v_inf, Rs;
    unsigned long long seed;
    int max_segments;
    const Atom* qin;                  // atoms to trace (trace_kernel)
    long long nin;
    Atom* qout;                       // secondaries born
    unsigned long long* nout;
    long long cap;
    unsigned long long* cnt;          // [kCnt]
    // a general mesh (the reference's own grid): the cells' corners and face slots in
    // Cell::Grans order with each slot's way on (scode: the next cell, -1 out of the grid,
    // -2 across the inner sphere, -3 an Axis slot), and the boundary faces the atoms
    // enter by (outer) and cross the inner sphere to (inner), as (A, B, cell) triples
    const int* __restrict__ coff;
    const int* __restrict__ corners;
    const int* __restrict__ soff;
    const int* __restrict__ sA;
    const int* __restrict__ sB;
    const int* __restrict__ scode;
    const int* __restrict__ bout;
    const int* __restrict__ bin;
    int nbout, nbin;
};

// ------------------------------------------------------------------- sensor ----
// Sensor::MakeRandom (sensor.cpp:11-31): a multiplicative congruential generator modulo
// 2^40 kept in three integers.
__device__ __forceinline__ double draw(int* s)
{
    const int ic15 = 32768, ic10 = 1024;
    const int mz = 710, my = 17784, mx = 11973;
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
    return 9.0949470177292824E-13 * (1.073741824E9 * s[2] + ic15 * s[1] + s[0]);
}

__device__ __forceinline__ unsigned long long splitmix(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// a primary's sensor from the seed and its number (odd low word: full period)
__device__ __forceinline__ void seed_primary(unsigned long long seed, long long k, int* s)
{
    const unsigned long long z = splitmix(splitmix(seed) + (unsigned long long)k);
    s[0] = (int)(z & 0x7fff) | 1;
    s[1] = (int)((z >> 15) & 0x7fff);
    s[2] = (int)((z >> 30) & 0x3ff);
}

// a secondary's sensor from three draws of its parent's
__device__ __forceinline__ void seed_child(int* parent, int* s)
{
    const double d1 = draw(parent), d2 = draw(parent), d3 = draw(parent);
    s[0] = ((int)(d1 * 32768.0) & 0x7fff) | 1;
    s[1] = (int)(d2 * 32768.0) & 0x7fff;
    s[2] = (int)(d3 * 1024.0) & 0x3ff;
}

// -------------------------------------------------------------------- rates ----
// Setka::Velosity_1..3 (Setka.cpp:15830-15866): the mean relative speed of a Maxwellian
// of thermal speed cp seen at relative speed u, and its momentum and energy weights.
__device__ __forceinline__ void velosity(double u, double cp, double* o)
{
    const double pi = kPiM;
    const double c2 = cp * cp, u2 = u * u;
    if (u < 0.00001) {
        o[0] = 2.0 * cp / kSqrtPiR + 2.0 * u * u / (3.0 * cp * kSqrtPiR) - u * u * u * u / (15.0 * cp * cp * cp * kSqrtPiR);
        o[1] = (8.0 / 3.0) * c2 * c2 * pi * u + (8.0 / 15.0) * c2 * pi * u * u * u - (4.0 / 105.0) * pi * u2 * u2 * u;
        o[2] = 8.0 * cp / (3.0 * kSqrtPiR) + 8.0 * u * u / (9.0 * cp * kSqrtPiR) -
               44.0 * u * u * u * u / (135.0 * cp * cp * cp * kSqrtPiR);
    } else {
        const double e = exp(-u * u / c2), f = erf(u / cp);
        o[0] = e * cp / kSqrtPiR + (u + c2 / (2.0 * u)) * f;
        o[1] = cp * cp * cp * pi * (e * cp * u * 2.0 * (c2 + 2.0 * u2) + kSqrtPiR * (4.0 * u2 * u2 + 4.0 * cp * cp * u2 - c2 * c2) * f) /
               (4.0 * u * u);
        o[2] = e * cp * (5.0 * c2 + 2.0 * u2) / (kSqrtPiR * (3.0 * c2 + 2.0 * u2)) +
               (4.0 * u2 * u2 + 12.0 * cp * cp * u2 + 3.0 * c2 * c2) * f / (2.0 * u * (3.0 * c2 + 2.0 * u2));
    }
}

// The nine integrals over the Maxwellian exp(-|y|^2) seen at relative speed x, with
// s = |x - y| and psi_k = (ln s)^k, k = 0, 1, 2:
//   f_k = int |x - y| psi_k d^3y,  g_k = int |x - y| (x - y).x^ psi_k d^3y,
//   h_k = int |x - y|^3 psi_k d^3y,
// reduced to one dimension: f_k = (pi/x) int s^2 psi_k (E- - E+) ds,
// g_k = 2 pi int s^4 psi_k (a (E- + E+) - (E- - E+)) / a^2 ds, h_k = (pi/x) int s^4 psi_k
// (E- - E+) ds, with E-+ = exp(-(s -+ x)^2) and a = 2 x s (series for a < 0.1), their
// x = 0 limits at x = 0.  Composite Simpson, s in [0, x + 9], 3000 intervals.
__device__ void base_quad(double x, double* o)
{
    const int N = 3000;
    const double smax = x + 9.0, hs = smax / N;
    double acc[9];
#pragma unroll
    for (int m = 0; m < 9; m++) acc[m] = 0.0;
    for (int i = 1; i <= N; i++) {
        const double s = i * hs;
        const double w = (i == N) ? 1.0 : (i & 1) ? 4.0 : 2.0;
        const double ls = log(s);
        const double psi[3] = { 1.0, ls, ls * ls };
        const double s2 = s * s, s4 = s2 * s2;
        double kf, kg, kh;
        if (x == 0.0) {
            const double e = exp(-s2);
            kf = 4.0 * kPiM * s2 * s * e;
            kg = 0.0;
            kh = 4.0 * kPiM * s4 * s * e;
        } else {
            const double a = 2.0 * x * s;
            double dm, sm;
            if (a < 1.0) {
                const double e0 = exp(-(x * x + s2));
                dm = e0 * (expm1(a) - expm1(-a));
                sm = e0 * (exp(a) + exp(-a));
            } else {
                const double em = exp(-(s - x) * (s - x)), ep = exp(-(s + x) * (s + x));
                dm = em - ep;
                sm = em + ep;
            }
            kf = kPiM / x * s2 * dm;
            kh = kPiM / x * s4 * dm;
            if (a < 0.1) {
                const double e0 = exp(-(x * x + s2));
                kg = 2.0 * kPiM * s4 * e0 * 2.0 * (a / 3.0 + a * a * a / 30.0 + a * a * a * a * a / 840.0);
            } else {
                kg = 2.0 * kPiM * s4 * (a * sm - dm) / (a * a);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            acc[k] += w * kf * psi[k];
            acc[3 + k] += w * kg * psi[k];
            acc[6 + k] += w * kh * psi[k];
        }
    }
#pragma unroll
    for (int m = 0; m < 9; m++) o[m] = acc[m] * hs / 3.0;
}

__global__ void __launch_bounds__(64) table_kernel(double* tab)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kTabN) return;
    double o[9];
    base_quad((double)k / kTabPer, o);
#pragma unroll
    for (int m = 0; m < 9; m++) tab[m * kTabN + k] = o[m];
}

// MKmethod::int_1..3 at (u, cp), u / cp <= 7: the table combined for the cross section
// (1 - a ln(w cp))^2 = (b - a ln w)^2, b = 1 - a ln cp.
__device__ __forceinline__ void rate_int(const double* __restrict__ tab, double u, double cp, double sa, double* o)
{
    const double t = (u / cp) * kTabPer;
    int k = (int)t;
    k = k < 1 ? 1 : (k > kTabN - 3 ? kTabN - 3 : k);
    const double r = t - k;
    const double w0 = -r * (r - 1.0) * (r - 2.0) / 6.0, w1 = (r + 1.0) * (r - 1.0) * (r - 2.0) / 2.0;
    const double w2 = -(r + 1.0) * r * (r - 2.0) / 2.0, w3 = (r + 1.0) * r * (r - 1.0) / 6.0;
    double B[9];
#pragma unroll
    for (int m = 0; m < 9; m++) {
        const double* c = tab + m * kTabN + k;
        B[m] = w0 * c[-1] + w1 * c[0] + w2 * c[1] + w3 * c[2];
    }
    const double b = 1.0 - sa * log(cp);
    const double c0 = b * b, c1 = -2.0 * sa * b, c2 = sa * sa;
    const double n3 = kSqrtPiR * kSqrtPiR * kSqrtPiR;
    o[0] = (cp / n3) * (c0 * B[0] + c1 * B[1] + c2 * B[2]);
    o[1] = -(cp * cp / n3) * (c0 * B[3] + c1 * B[4] + c2 * B[5]);
    o[2] = (cp * cp * cp / n3) * (c0 * B[6] + c1 * B[7] + c2 * B[8]);
}

__device__ __forceinline__ double sig(double sa, double u)
{
    const double t = 1.0 - sa * log(u);
    return t * t;
}

// ------------------------------------------------------------------ sampler ----
// MKmethod::Change_Velosity4 with no splitting zones (MKmethod.cpp:2270-2531, I = 0):
// the proton an atom of velocity V exchanges with, from the Maxwellian around U weighted
// by |V - v| sigma(|V - v|): proposals from the mixture (X + |y|) exp(-y^2) (plain
// Maxwellian with probability p4, else the speed-weighted one), accepted with
// |V - v| sigma / (sigma(X) (X + |y|)).  Velocities in units of cp.  Without sigma: the
// speed-weighted Maxwellian of the inflow.  Returns the reference's mu flag.  The loop
// stops after kMaxTries proposals (the reference's has no bound) with the last one.
template <bool SIGMA>
__device__ __forceinline__ int sample(int* s, const double* U, const double* V, double cp, double sa, double* W)
{
    const double pi_ = kPiM;
    const double X = sqrt((V[0] - U[0]) * (V[0] - U[0]) + (V[1] - U[1]) * (V[1] - U[1]) + (V[2] - U[2]) * (V[2] - U[2]));
    const double p4 = 0.5 * kSqrtPiR * X / (1.0 + 0.5 * kSqrtPiR * X);
    double v1, v2, v3, h, ksi6;
    int tries = 0;
    do {
        const double ksi1 = draw(s), ksi2 = draw(s), ksi3 = draw(s), ksi4 = draw(s), ksi5 = draw(s);
        ksi6 = draw(s);
        double y1, y2, y3;
        if (p4 < ksi1) {
            const double om1 = 1.0 - 2.0 * ksi4;
            const double sq = sqrt(1.0 - om1 * om1);
            const double om2 = sq * cos(2.0 * pi_ * ksi5);
            const double om3 = sq * sin(2.0 * pi_ * ksi5);
            const double lo = sqrt(-log(ksi2 * ksi3));
            y1 = lo * om1;
            y2 = lo * om2;
            y3 = lo * om3;
        } else {
            const double l4 = sqrt(-log(ksi4));
            y1 = sqrt(-log(ksi2)) * cos(pi_ * ksi3);
            y2 = l4 * cos(2.0 * pi_ * ksi5);
            y3 = l4 * sin(2.0 * pi_ * ksi5);
        }
        v1 = y1 + U[0];
        v2 = y2 + U[1];
        v3 = y3 + U[2];
        const double u1 = V[0] - v1, u2 = V[1] - v2, u3 = V[2] - v3;
        const double uuu = sqrt(u1 * u1 + u2 * u2 + u3 * u3);
        const double yy = sqrt(y1 * y1 + y2 * y2 + y3 * y3);
        if constexpr (SIGMA) {
            const double su = 1.0 - sa * log(uuu * cp), sx = 1.0 - sa * log(X * cp);
            h = (uuu * (su * su)) / ((sx * sx) * (X + yy));
        } else {
            h = uuu / (X + yy);
        }
    } while (h < ksi6 && ++tries < kMaxTries);
    W[0] = v1;
    W[1] = v2;
    W[2] = v3;
    return (v1 >= 0.0 || v2 * v2 + v3 * v3 > 0.0 * (v1 * v1)) ? 1 : 0;
}

// ----------------------------------------------------------------- geometry ----
// First t > tmin where p + t v crosses the surface the segment (xa, ra) - (xb, rb) of the
// (x, r) half plane sweeps about the x axis (the quadratic of Setka.cpp:12913-12930 in
// the segment's own terms); -1 if none.  Roots on the mirror cone or off the segment
// are dropped.
__device__ double hit(const double* p, const double* v, double xa, double ra, double xb, double rb, double tmin)
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

// Way out of cell (i, j): face 0 (node row i), 1 (row i + 1), 2 (column j, not on the
// axis), 3 (column j + 1, not on the axis).
__device__ double cell_exit(const Args& a, int i, int j, const double* p, const double* v, double tmin, int& face)
{
    const int w = a.nj + 1;
    const int n0 = i * w + j, n1 = n0 + 1, n2 = n0 + w, n3 = n2 + 1;
    const double x0 = a.x[n0], y0 = a.y[n0], x1 = a.x[n1], y1 = a.y[n1];
    const double x2 = a.x[n2], y2 = a.y[n2], x3 = a.x[n3], y3 = a.y[n3];
    double best = -1.0;
    face = -1;
    double t = hit(p, v, x0, y0, x1, y1, tmin);
    if (t > 0.0) { best = t; face = 0; }
    t = hit(p, v, x2, y2, x3, y3, tmin);
    if (t > 0.0 && (best < 0.0 || t < best)) { best = t; face = 1; }
    if (j > 0) {
        t = hit(p, v, x0, y0, x2, y2, tmin);
        if (t > 0.0 && (best < 0.0 || t < best)) { best = t; face = 2; }
    }
    if (j < a.nj - 1) {
        t = hit(p, v, x1, y1, x3, y3, tmin);
        if (t > 0.0 && (best < 0.0 || t < best)) { best = t; face = 3; }
    }
    return best;
}

// first crossing with the faces of node row q (ni: the grid's outside, 0: the inner sphere)
__device__ double row_hit(const Args& a, int q, const double* p, const double* v, double tmin, int& jj)
{
    double best = -1.0;
    jj = -1;
    const int w = a.nj + 1;
    for (int c = 0; c < a.nj; c++) {
        const int n0 = q * w + c;
        const double t = hit(p, v, a.x[n0], a.y[n0], a.x[n0 + 1], a.y[n0 + 1], tmin);
        if (t > 0.0 && (best < 0.0 || t < best)) { best = t; jj = c; }
    }
    return best;
}

// Way out of mesh cell cc: the nearest crossing over its face slots, Axis faces skipped
// (the reference's loop over now->Grans, Setka.cpp:12893-12960); slot = the slot crossed.
__device__ double mesh_exit(const Args& a, int cc, const double* p, const double* v, double tmin, int& slot)
{
    double best = -1.0;
    slot = -1;
    for (int s = a.soff[cc]; s < a.soff[cc + 1]; s++) {
        if (a.scode[s] == -3) continue;
        const int A = a.sA[s], B = a.sB[s];
        const double t = hit(p, v, a.x[A], a.y[A], a.x[B], a.y[B], tmin);
        if (t > 0.0 && (best < 0.0 || t < best)) { best = t; slot = s; }
    }
    return best;
}

// first crossing with a list of n (A, B, cell) faces; jj = the list entry, -1 if none
__device__ double list_hit(const Args& a, const int* L, int n, const double* p, const double* v, double tmin, int& jj)
{
    double best = -1.0;
    jj = -1;
    for (int c = 0; c < n; c++) {
        const int A = L[3 * c], B = L[3 * c + 1];
        const double t = hit(p, v, a.x[A], a.y[A], a.x[B], a.y[B], tmin);
        if (t > 0.0 && (best < 0.0 || t < best)) { best = t; jj = c; }
    }
    return best;
}

// the plasma velocity of cell k at the point (y, z): (u, v cos, v sin) of its angle about the axis
__device__ __forceinline__ void plasma_at(const Args& a, long long k, double y, double z, double* U, double& ca,
                                          double& sa)
{
    const double r = sqrt(y * y + z * z);
    ca = r > 0.0 ? y / r : 1.0;
    sa = r > 0.0 ? z / r : 0.0;
    const double u = a.pl[2 * a.nc + k], v = a.pl[3 * a.nc + k];
    U[0] = u;
    U[1] = v * ca;
    U[2] = v * sa;
}

__device__ __forceinline__ void tally(double* p, double v) { unsafeAtomicAdd(p, v); }

// ------------------------------------------------------------------- tracer ----
// One atom to the end of its way through the grid (Setka::Fly_exchenge_Imit_Korol,
// Setka.cpp:12814-13981): per cell piece, the optical depth at the piece's midpoint, the
// exchange point from the truncated exponential, the weight before it and the rest after
// it tallied with their midpoint directions, the source integrals at the exchange point,
// the secondary (population of the cell, kernel velocity, roulette at wmin) queued.
template <bool MESH>
__device__ void trace(const Args& a, Atom at)
{
    const long long nc = a.nc;
    double* T = a.tal;
    unsigned long long segs = 0, lost = 0;
    int seg = 0;
    for (; seg < a.max_segments; seg++) {
        const long long k = MESH ? (long long)at.i : (long long)at.i * a.nj + at.j;
        const double vn = sqrt(at.v[0] * at.v[0] + at.v[1] * at.v[1] + at.v[2] * at.v[2]);
        int n00, n11;   // a diagonal of the cell: its corners 0 and 2
        if constexpr (MESH) {
            n00 = a.corners[a.coff[k]];
            n11 = a.corners[a.coff[k] + 2];
        } else {
            n00 = at.i * (a.nj + 1) + at.j;
            n11 = n00 + a.nj + 2;
        }
        const double dX = a.x[n11] - a.x[n00], dY = a.y[n11] - a.y[n00];
        const double tmin = 1e-9 * sqrt(dX * dX + dY * dY) / vn;
        int face;
        const double t_exit = MESH ? mesh_exit(a, at.i, at.p, at.v, tmin, face)
                                   : cell_exit(a, at.i, at.j, at.p, at.v, tmin, face);
        if (face < 0) { lost++; break; }
        segs++;
        const double ro = a.pl[k], cp = a.pl[nc + k];
        double time = t_exit;
        bool ch_now = false;
        double o[3];
        {   // pieces of at most one mean free time (Setka.cpp:13076-13081)
            velosity(0.0, cp, o);
            const double nu0 = (ro * o[0] * sig(a.sig_a, o[0])) / a.kn;
            if (time > 1.0 / nu0) {
                time = time / ((int)(time / (1.0 / nu0)) + 1);
                ch_now = true;
            }
        }
        double U[3], ca, sa;
        plasma_at(a, k, at.p[1] + 0.5 * time * at.v[1], at.p[2] + 0.5 * time * at.v[2], U, ca, sa);
        double w1 = at.v[0] - U[0], w2 = at.v[1] - U[1], w3 = at.v[2] - U[2];
        double u = sqrt(w1 * w1 + w2 * w2 + w3 * w3);
        double nu;
        if (u / cp > 7.0) {
            velosity(u, cp, o);
            nu = (ro * o[0] * sig(a.sig_a, o[0])) / a.kn;
        } else {
            rate_int(a.tab, u, cp, a.sig_a, o);
            nu = (ro * o[0]) / a.kn;
        }
        const double kappa = nu * time;
        const double xi = draw(at.s);
        const double t_ex = -(time / kappa) * log(1.0 - xi * (1.0 - exp(-kappa)));
        const double mu_ex = at.mu * (1.0 - exp(-kappa));
        const double mu2 = at.mu - mu_ex;
        const double pe[3] = { at.p[0] + t_ex * at.v[0], at.p[1] + t_ex * at.v[1], at.p[2] + t_ex * at.v[2] };
        const double V2 = at.v[0] * at.v[0] + at.v[1] * at.v[1] + at.v[2] * at.v[2];
        const int h = at.area;
        {   // the weight before the exchange point and the rest after it (:13399-13443, :13700-13721)
            double U1[3], c1, s1, U2[3], c2, s2;
            plasma_at(a, k, at.p[1] + 0.5 * t_ex * at.v[1], at.p[2] + 0.5 * t_ex * at.v[2], U1, c1, s1);
            const double t2 = time - t_ex;
            plasma_at(a, k, pe[1] + 0.5 * t2 * at.v[1], pe[2] + 0.5 * t2 * at.v[2], U2, c2, s2);
            const double wn = t_ex * at.mu + t2 * mu2;
            tally(T + (0 + h) * nc + k, wn);
            tally(T + (4 + h) * nc + k, at.v[0] * wn);
            tally(T + (8 + h) * nc + k,
                  t_ex * (at.v[1] * c1 + at.v[2] * s1) * at.mu + t2 * (at.v[1] * c2 + at.v[2] * s2) * mu2);
            tally(T + (12 + h) * nc + k, V2 * wn);
        }
        // the source integrals at the exchange point (:13408-13432)
        plasma_at(a, k, pe[1], pe[2], U, ca, sa);
        w1 = at.v[0] - U[0];
        w2 = at.v[1] - U[1];
        w3 = at.v[2] - U[2];
        u = sqrt(w1 * w1 + w2 * w2 + w3 * w3);
        const double u1 = -w1, u2 = -w2, u3 = -w3;
        const double skalar = at.v[0] * u1 + at.v[1] * u2 + at.v[2] * u3;
        if (u > 0.0) {
            if (u / cp > 7.0) {
                velosity(u, cp, o);
                const double uz = o[0];
                const double uz_M = o[1] / (uz * (cp * cp) * cp * kPiM * kSqrtPiR);
                const double uz_E = o[2];
                tally(T + 16 * nc + k, -mu_ex * uz_M * u1 / u);
                tally(T + 17 * nc + k, -mu_ex * uz_M * (u2 * ca + u3 * sa) / u);
                tally(T + 18 * nc + k, mu_ex * (-0.25 * (3.0 * (cp * cp) + 2.0 * (u * u)) * (uz_E / uz) - uz_M * skalar / u));
            } else {
                rate_int(a.tab, u, cp, a.sig_a, o);
                tally(T + 16 * nc + k, mu_ex * (o[1] / o[0]) * u1 / u);
                tally(T + 17 * nc + k, mu_ex * (o[1] / o[0]) * (u2 * ca + u3 * sa) / u);
                tally(T + 18 * nc + k, mu_ex * (-0.5 * o[2] / o[0] + o[1] / o[0] * skalar / u));
            }
        }
        {   // the secondary (:13445-13481)
            const double Un[3] = { U[0] / cp, U[1] / cp, U[2] / cp };
            const double Vn[3] = { at.v[0] / cp, at.v[1] / cp, at.v[2] / cp };
            double W[3];
            sample<true>(at.s, Un, Vn, cp, a.sig_a, W);
            W[0] *= cp;
            W[1] *= cp;
            W[2] *= cp;
            double mu3 = mu_ex;
            bool keep = true;
            if (!(mu3 >= a.wmin)) {
                if (mu3 >= draw(at.s) * a.wmin) mu3 = a.wmin;
                else keep = false;
            }
            if (keep) {
                tally(T + 19 * nc + k, mu3 * (at.v[0] - W[0]));
                tally(T + 20 * nc + k, mu3 * ((at.v[1] - W[1]) * ca + (at.v[2] - W[2]) * sa));
                tally(T + 21 * nc + k, mu3 * 0.5 * (V2 - (W[0] * W[0] + W[1] * W[1] + W[2] * W[2])));
                Atom c;
                c.p[0] = pe[0]; c.p[1] = pe[1]; c.p[2] = pe[2];
                c.v[0] = W[0]; c.v[1] = W[1]; c.v[2] = W[2];
                c.mu = mu3;
                c.i = at.i;
                c.j = at.j;
                c.area = a.pop[k];
                seed_child(at.s, c.s);
                const unsigned long long slot = atomicAdd(a.nout, 1ull);
                if ((long long)slot < a.cap) a.qout[slot] = c;
                else atomicAdd(a.cnt + 4, 1ull);
            }
        }
        at.mu = mu2;
        if (ch_now) {
#pragma unroll
            for (int d = 0; d < 3; d++) at.p[d] += time * at.v[d];
            continue;
        }
#pragma unroll
        for (int d = 0; d < 3; d++) at.p[d] += t_exit * at.v[d];
        if constexpr (MESH) {
            const int to = a.scode[face];
            if (to >= 0) {
                at.i = to;
            } else if (to == -2) {                 // across the inner sphere
                int jj;
                const double th = list_hit(a, a.bin, a.nbin, at.p, at.v, tmin, jj);
                if (jj < 0) { lost++; break; }
#pragma unroll
                for (int d = 0; d < 3; d++) at.p[d] += th * at.v[d];
                at.i = a.bin[3 * jj + 2];
            } else {
                break;                             // out of the grid
            }
        } else if (face == 1) {
            if (at.i == a.ni - 1) break;           // out of the grid
            at.i++;
        } else if (face == 0) {
            if (at.i == 0) {                       // across the inner sphere
                int jj;
                const double th = row_hit(a, 0, at.p, at.v, tmin, jj);
                if (jj < 0) { lost++; break; }
#pragma unroll
                for (int d = 0; d < 3; d++) at.p[d] += th * at.v[d];
                at.j = jj;
            } else {
                at.i--;
            }
        } else if (face == 2) {
            at.j--;
        } else {
            at.j++;
        }
    }
    if (seg == a.max_segments) lost++;
    atomicAdd(a.cnt + 1, 1ull);
    atomicAdd(a.cnt + 2, segs);
    if (lost) atomicAdd(a.cnt + 3, lost);
}

// Primary number kk: velocity from the speed-weighted drifting Maxwellian, entry point
// uniform on the disc across it of the sphere of radius Rs around the grid, then in a
// straight line to the grid's outer surface.  False if it misses the grid.
template <bool MESH>
__device__ bool primary(const Args& a, long long kk, Atom& at)
{
    seed_primary(a.seed, kk, at.s);
    const double U[3] = { a.v_inf, 0.0, 0.0 }, Z[3] = { 0.0, 0.0, 0.0 };
    double v[3];
    sample<false>(at.s, U, Z, 1.0, a.sig_a, v);
    const double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double e[3] = { v[0] / vn, v[1] / vn, v[2] / vn };
    double q[3] = { 0.0, 0.0, 0.0 };
    const double ax = fabs(e[0]), ay = fabs(e[1]), az = fabs(e[2]);
    if (ax <= ay && ax <= az) q[0] = 1.0;
    else if (ay <= az) q[1] = 1.0;
    else q[2] = 1.0;
    double e1[3] = { e[1] * q[2] - e[2] * q[1], e[2] * q[0] - e[0] * q[2], e[0] * q[1] - e[1] * q[0] };
    const double n1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    e1[0] /= n1;
    e1[1] /= n1;
    e1[2] /= n1;
    const double e2[3] = { e[1] * e1[2] - e[2] * e1[1], e[2] * e1[0] - e[0] * e1[2], e[0] * e1[1] - e[1] * e1[0] };
    const double rho = a.Rs * sqrt(draw(at.s));
    const double phi = 2.0 * kPiM * draw(at.s);
    const double c = rho * cos(phi), s = rho * sin(phi);
    const double back = sqrt(a.Rs * a.Rs - rho * rho);
    double p[3];
#pragma unroll
    for (int d = 0; d < 3; d++) p[d] = c * e1[d] + s * e2[d] - back * e[d];
    int jj;
    const double t = MESH ? list_hit(a, a.bout, a.nbout, p, v, 0.0, jj) : row_hit(a, a.ni, p, v, 0.0, jj);
    if (jj < 0) return false;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        at.p[d] = p[d] + t * v[d];
        at.v[d] = v[d];
    }
    at.mu = 1.0;
    at.i = MESH ? a.bout[3 * jj + 2] : a.ni - 1;
    at.j = MESH ? 0 : jj;
    at.area = 3;
    return true;
}

template <bool MESH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) primary_kernel(Args a, long long k0, long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    Atom at;
    if (!primary<MESH>(a, k0 + t, at)) return;
    atomicAdd(a.cnt, 1ull);
    trace<MESH>(a, at);
}

template <bool MESH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) trace_kernel(Args a)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nin) return;
    trace<MESH>(a, a.qin[t]);
}

// rho, cp = sqrt(p / rho), u, v, p of the plasma; bad counts cells with rho <= 0 or p <= 0
template <class TS>
__global__ void __launch_bounds__(256) plasma_kernel(const TS* __restrict__ st, long long nc, double* pl,
                                                     unsigned long long* bad)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    const double ro = (double)st[k], p = (double)st[nc + k];
    if (!(ro > 0.0) || !(p > 0.0)) atomicAdd(bad, 1ull);
    pl[k] = ro;
    pl[nc + k] = sqrt(p / ro);
    pl[2 * nc + k] = (double)st[2 * nc + k];
    pl[3 * nc + k] = (double)st[3 * nc + k];
    pl[4 * nc + k] = p;
}

// The post-processing of MK_start_new (Setka.cpp:12241-12449): moments per population and
// of all, sources per unit volume, the closure with the new moments and k = I / M.
template <bool MESH>
__global__ void __launch_bounds__(256) finalize_kernel(Args a, double Phi, double n_p, long long nprim, double* out)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nc = a.nc;
    if (k >= nc) return;
    double A = 0.0, G = 0.0;
    if constexpr (MESH) {   // Cell::Get_Volume_rotate(360) (Cell.cpp:188-206) over the contour
        const int* q = a.corners + a.coff[k];
        const int n = a.coff[k + 1] - a.coff[k];
        for (int m = 0; m < n; m++) {
            const int p0 = q[m], p1 = q[m + 1 == n ? 0 : m + 1];
            const double cr = a.x[p0] * a.y[p1] - a.y[p0] * a.x[p1];
            A = A + cr;
            G = G + (a.y[p0] + a.y[p1]) * cr;
        }
    } else {
        const int i = (int)(k / a.nj), j = (int)(k % a.nj), w = a.nj + 1;
        const int q[4] = { i * w + j, (i + 1) * w + j, (i + 1) * w + j + 1, i * w + j + 1 };
        for (int m = 0; m < 4; m++) {   // Cell::Get_Volume_rotate(360) (Cell.cpp:188-206)
            const int p0 = q[m], p1 = q[(m + 1) & 3];
            const double cr = a.x[p0] * a.y[p1] - a.y[p0] * a.x[p1];
            A = A + cr;
            G = G + (a.y[p0] + a.y[p1]) * cr;
        }
    }
    A = A * 0.5;
    const double vol = fabs(A) * (1.0 / (6.0 * A)) * G * kPiM * 360.0 / 180.0;
    const double no = (double)nprim * vol;
    const double* T = a.tal;
    double Hn[4], Hu[4], Hv[4], HT[4];
    double Fn = 0.0, Fu = 0.0, Fv = 0.0, FT = 0.0;
    for (int hh = 0; hh < 4; hh++) {
        Hn[hh] = T[(0 + hh) * nc + k];
        Hu[hh] = T[(4 + hh) * nc + k];
        Hv[hh] = T[(8 + hh) * nc + k];
        HT[hh] = T[(12 + hh) * nc + k];
        Fn += Hn[hh];
        Fu += Hu[hh];
        Fv += Hv[hh];
        FT += HT[hh];
    }
    double I[3], II[3];
    for (int m = 0; m < 3; m++) {
        I[m] = n_p * (Phi * T[(16 + m) * nc + k] / no);
        II[m] = n_p * (Phi * T[(19 + m) * nc + k] / no);
    }
    if (Fn > 0.0) {
        Fu = Fu / Fn;
        Fv = Fv / Fn;
        FT = (2.0 / 3.0) * (FT / Fn - (Fu * Fu + Fv * Fv + 0.0 * 0.0));
        for (int hh = 0; hh < 4; hh++) {
            if (Hn[hh] > 0.0) {
                Hu[hh] = Hu[hh] / Hn[hh];
                Hv[hh] = Hv[hh] / Hn[hh];
                HT[hh] = (2.0 / 3.0) * (HT[hh] / Hn[hh] - (Hu[hh] * Hu[hh] + Hv[hh] * Hv[hh] + 0.0 * 0.0));
            } else {
                Hu[hh] = 0.0;
                Hv[hh] = 0.0;
                HT[hh] = 0.0;
            }
        }
    } else {
        Fn = 0.0; Fu = 0.0; Fv = 0.0; FT = 0.0;
    }
    Fn = Phi * Fn / no;
    for (int hh = 0; hh < 4; hh++) Hn[hh] = Phi * Hn[hh] / no;
    const double u = a.pl[2 * nc + k], v = a.pl[3 * nc + k], ro = a.pl[k];
    const double p = a.pl[4 * nc + k];
    double Mu = 0.0, Mv = 0.0, MT = 0.0;
    for (int hh = 0; hh < 4; hh++) {
        double roH = Hn[hh], pH = 0.5 * HT[hh] * Hn[hh];
        if (roH <= 0.0) { roH = 0.0000001; pH = 0.0; }
        const double du = u - Hu[hh], dv = v - Hv[hh];
        const double UM = sqrt(du * du + dv * dv + (64.0 / (9.0 * kPiM)) * (p / ro + 2.0 * pH / roH));
        const double UU = sqrt(du * du + dv * dv + (4.0 / kPiM) * (p / ro + 2.0 * pH / roH));
        const double nuh = ro * roH * UM * sig(a.sig_a, UM);
        Mu += nuh * (Hu[hh] - u);
        Mv += nuh * (Hv[hh] - v);
        MT += nuh * (((Hu[hh] * Hu[hh]) + (Hv[hh] * Hv[hh]) - (u * u) - (v * v)) / 2.0 + (UU / UM) * (2.0 * pH / roH - p / ro));
    }
    const double coef = n_p / a.kn;
    const double M[3] = { Mu * coef, Mv * coef, MT * coef };
    for (int hh = 0; hh < 4; hh++) {
        out[(0 + hh) * nc + k] = Hn[hh];
        out[(4 + hh) * nc + k] = Hu[hh];
        out[(8 + hh) * nc + k] = Hv[hh];
        out[(12 + hh) * nc + k] = HT[hh];
    }
    for (int m = 0; m < 3; m++) {
        double kf = 1.0;
        if (fabs(M[m]) > 0.0001) {
            kf = I[m] / M[m];
            if (kf > 3.0 || kf < 0.33) kf = 1.0;
        }
        out[(16 + m) * nc + k] = kf;
        out[(19 + m) * nc + k] = I[m];
        out[(22 + m) * nc + k] = II[m];
    }
    out[25 * nc + k] = Fn;
    out[26 * nc + k] = Fu;
    out[27 * nc + k] = Fv;
    out[28 * nc + k] = FT;
}

// fields 0..18 of the output into the context's source inputs, in the state's type
template <class TS>
__global__ void __launch_bounds__(256) install_kernel(const double* __restrict__ out, TS* src, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) src[k] = (TS)out[k];
}

__global__ void __launch_bounds__(64) rates_kernel(int n, const double* __restrict__ u, const double* __restrict__ cp,
                                                   const double* __restrict__ tab, double sa, double* out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double o[3];
    velosity(u[k], cp[k], o);
    out[6 * k] = o[0];
    out[6 * k + 1] = o[1];
    out[6 * k + 2] = o[2];
    o[0] = o[1] = o[2] = 0.0;
    if (u[k] / cp[k] <= 7.0) rate_int(tab, u[k], cp[k], sa, o);
    out[6 * k + 3] = o[0];
    out[6 * k + 4] = o[1];
    out[6 * k + 5] = o[2];
}

__global__ void __launch_bounds__(64) exchange_kernel(int n, const double* __restrict__ in, const int* __restrict__ seeds,
                                                      int reps, double sa, double* out, int* state)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int s[3] = { seeds[3 * k], seeds[3 * k + 1], seeds[3 * k + 2] };
    const double* r = in + 7 * k;
    for (int q = 0; q < reps; q++) {
        double W[3];
        const int mu = sample<true>(s, r, r + 3, r[6], sa, W);
        double* o = out + ((size_t)k * reps + q) * 4;
        o[0] = W[0];
        o[1] = W[1];
        o[2] = W[2];
        o[3] = mu;
    }
    state[3 * k] = s[0];
    state[3 * k + 1] = s[1];
    state[3 * k + 2] = s[2];
}

}  // namespace mc
}  // namespace gel

using namespace gel;
using gel::mc::kMaxGenerations;

namespace {

// Device buffers of one call, freed on every exit path.
struct DevBufs {
    std::vector<void*> p;
    ~DevBufs()
    {
        for (void* q : p) (void)hipFree(q);
    }
    template <class T> bool get(T** out, size_t n)
    {
        void* q = nullptr;
        if (hipMalloc(&q, n * sizeof(T)) != hipSuccess) return false;
        p.push_back(q);
        *out = (T*)q;
        return true;
    }
};

constexpr double kSigA = 0.1307345665;   // a_2 (Help.h:63)

#define MC_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return gel_set_error(GEL_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

extern "C" {

void gel_mc_params_reference(gel_mc_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->v_inf = -2.54338;       // Velosity_inf (Help.h:53)
    p->kn = 6.0;               // Kn_ (Help.h:60)
    p->sig_a = kSigA;          // a_2 (Help.h:63)
    p->n_p_lism = 3.0;         // n_p_LISM_ (Help.h:64)
    p->w_min = 0.25;           // weight_ (Help.h:93)
    p->n_primary = 1000000;
    p->seed = 1;
    p->max_segments = 1000000;
}

int gel_mc_rates(int n, const double* u, const double* cp, double* out)
{
    if (n < 0 || (n > 0 && (!u || !cp || !out))) return gel_set_error(GEL_E_ARG, "null argument");
    if (n == 0) return GEL_OK;
    for (int k = 0; k < n; k++)
        if (!(cp[k] > 0.0) || !(u[k] >= 0.0)) return gel_set_error(GEL_E_ARG, "rates need u >= 0 and cp > 0");
    DevBufs b;
    double *du, *dc, *dt, *dout;
    if (!b.get(&du, n) || !b.get(&dc, n) || !b.get(&dt, 9 * (size_t)mc::kTabN) || !b.get(&dout, 6 * (size_t)n))
        return gel_set_error(GEL_E_NOMEM, "rates buffers");
    MC_HIP(hipMemcpy(du, u, n * sizeof(double), hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(dc, cp, n * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mc::table_kernel, dim3((mc::kTabN + 63) / 64), dim3(64), 0, 0, dt);
    MC_HIP(hipGetLastError());
    hipLaunchKernelGGL(mc::rates_kernel, dim3((n + 63) / 64), dim3(64), 0, 0, n, du, dc, dt, kSigA, dout);
    MC_HIP(hipGetLastError());
    MC_HIP(hipMemcpy(out, dout, 6 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    return GEL_OK;
}

int gel_mc_exchange(int n, const double* in, const int* seeds, int reps, double* out, int* state)
{
    if (n < 0 || reps < 0 || (n > 0 && (!in || !seeds || !out || !state))) return gel_set_error(GEL_E_ARG, "null argument");
    if (n == 0) return GEL_OK;
    for (int k = 0; k < n; k++) {
        if (!(in[7 * k + 6] > 0.0)) return gel_set_error(GEL_E_ARG, "cp must be positive");
        if (seeds[3 * k] < 0 || seeds[3 * k] > 32767 || seeds[3 * k + 1] < 0 || seeds[3 * k + 1] > 32767 ||
            seeds[3 * k + 2] < 0 || seeds[3 * k + 2] > 1023)
            return gel_set_error(GEL_E_ARG, "sensor seeds must lie in [0, 32768) x [0, 32768) x [0, 1024)");
    }
    DevBufs b;
    double *din, *dout;
    int *ds, *dst;
    if (!b.get(&din, 7 * (size_t)n) || !b.get(&ds, 3 * (size_t)n) || !b.get(&dout, 4 * (size_t)n * (reps ? reps : 1)) ||
        !b.get(&dst, 3 * (size_t)n))
        return gel_set_error(GEL_E_NOMEM, "sampler buffers");
    MC_HIP(hipMemcpy(din, in, 7 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(ds, seeds, 3 * (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mc::exchange_kernel, dim3((n + 63) / 64), dim3(64), 0, 0, n, din, ds, reps, kSigA, dout, dst);
    MC_HIP(hipGetLastError());
    MC_HIP(hipMemcpy(out, dout, 4 * (size_t)n * reps * sizeof(double), hipMemcpyDeviceToHost));
    MC_HIP(hipMemcpy(state, dst, 3 * (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    return GEL_OK;
}

int gel_mc_run(gel_ctx* c, const gel_mc_params* p, double* out, gel_mc_stats* st)
{
    if (!c || !p) return gel_set_error(GEL_E_ARG, "null argument");
    const bool mesh = c->mesh != nullptr;
    if (!c->motion.empty() || c->tracking)
        return gel_set_error(GEL_E_UNSUPPORTED, "the Monte Carlo runs on a grid at rest (no motion, no tracking)");
    if (!c->uploaded) return gel_set_error(GEL_E_STATE, "no state uploaded");
    if (p->n_primary <= 0 || !(p->kn > 0.0) || !(p->w_min > 0.0) || p->max_segments <= 0 || !(p->v_inf != 0.0))
        return gel_set_error(GEL_E_ARG, "n_primary, kn, w_min, max_segments must be positive, v_inf nonzero");
    const auto t0 = std::chrono::steady_clock::now();
    MC_HIP(hipStreamSynchronize(c->stream));
    hipStream_t s = c->stream;
    const long long nc = (long long)c->nc;
    // nodes on the host for the radius of the sphere around the grid
    std::vector<double> hx(c->nn), hy(c->nn);
    MC_HIP(hipMemcpy(hx.data(), c->d_x, c->nn * sizeof(double), hipMemcpyDeviceToHost));
    MC_HIP(hipMemcpy(hy.data(), c->d_y, c->nn * sizeof(double), hipMemcpyDeviceToHost));
    double Rs = 0.0;
    for (size_t k = 0; k < c->nn; k++) Rs = std::fmax(Rs, std::sqrt(hx[k] * hx[k] + hy[k] * hy[k]));
    Rs *= 1.0 + 1e-12;
    MeshMcView mv{};
    std::vector<int> scode, bout, bin;
    if (mesh) {
        gel_mesh_mc_view(c, &mv, scode, bout, bin);
        if (bout.empty()) return gel_set_error(GEL_E_ARG, "the mesh has no outer boundary faces to enter by");
    }
    std::vector<int8_t> pop(nc, 1);
    if (c->has_types) {   // population of a secondary born in a cell (Setka.cpp:13384-13398)
        std::vector<int8_t> ct(nc);
        MC_HIP(hipMemcpy(ct.data(), mesh ? mv.ctype : c->d_ctype, nc, hipMemcpyDeviceToHost));
        for (long long k = 0; k < nc; k++)
            pop[k] = ct[k] == GEL_C_5 ? 3 : ct[k] == GEL_C_4 ? 2 : (ct[k] == GEL_C_1 || ct[k] == GEL_C_CENTR) ? 0 : 1;
    }
    // queue room: each generation of one batch fits; a batch of primaries is an eighth of it
    const long long cap = std::max(1LL << 20, std::min(8LL << 20, 8 * p->n_primary));
    const long long batch = std::max(1LL, cap / 8);
    DevBufs b;
    double *pl, *tab, *tal, *dout;
    int8_t* dpop;
    mc::Atom *qa, *qb;
    unsigned long long *cnt, *nout;
    int *dcode = nullptr, *dbout = nullptr, *dbin = nullptr;
    if (!b.get(&pl, 5 * (size_t)nc) || !b.get(&tab, 9 * (size_t)mc::kTabN) || !b.get(&tal, mc::kTal * (size_t)nc) ||
        !b.get(&dout, GEL_MC_NFIELDS * (size_t)nc) || !b.get(&dpop, (size_t)nc) || !b.get(&qa, (size_t)cap) ||
        !b.get(&qb, (size_t)cap) || !b.get(&cnt, mc::kCnt + 1) || !b.get(&nout, 1) ||
        (mesh && (!b.get(&dcode, scode.size()) || !b.get(&dbout, bout.size()) || !b.get(&dbin, bin.size() + 1))))
        return gel_set_error(GEL_E_NOMEM, "Monte Carlo buffers");
    MC_HIP(hipMemcpyAsync(dpop, pop.data(), nc, hipMemcpyHostToDevice, s));
    if (mesh) {
        MC_HIP(hipMemcpyAsync(dcode, scode.data(), scode.size() * sizeof(int), hipMemcpyHostToDevice, s));
        MC_HIP(hipMemcpyAsync(dbout, bout.data(), bout.size() * sizeof(int), hipMemcpyHostToDevice, s));
        if (!bin.empty()) MC_HIP(hipMemcpyAsync(dbin, bin.data(), bin.size() * sizeof(int), hipMemcpyHostToDevice, s));
    }
    MC_HIP(hipMemsetAsync(cnt, 0, (mc::kCnt + 1) * sizeof(unsigned long long), s));
    MC_HIP(hipMemsetAsync(tal, 0, mc::kTal * (size_t)nc * sizeof(double), s));
    const dim3 gc((unsigned)((nc + 255) / 256));
    if (c->ts == 8)
        hipLaunchKernelGGL(mc::plasma_kernel<double>, gc, dim3(256), 0, s, (const double*)c->d_state[c->cur], nc, pl,
                           cnt + mc::kCnt);
    else
        hipLaunchKernelGGL(mc::plasma_kernel<float>, gc, dim3(256), 0, s, (const float*)c->d_state[c->cur], nc, pl,
                           cnt + mc::kCnt);
    hipLaunchKernelGGL(mc::table_kernel, dim3((mc::kTabN + 63) / 64), dim3(64), 0, s, tab);
    MC_HIP(hipGetLastError());
    unsigned long long bad = 0;
    MC_HIP(hipMemcpyAsync(&bad, cnt + mc::kCnt, sizeof(bad), hipMemcpyDeviceToHost, s));
    MC_HIP(hipStreamSynchronize(s));
    if (bad) return gel_set_error(GEL_E_ARG, "the plasma needs positive density and pressure in every cell");

    mc::Args a{};
    a.ni = c->ni;
    a.nj = c->nj;
    a.nc = nc;
    a.x = c->d_x;
    a.y = c->d_y;
    a.pl = pl;
    a.pop = dpop;
    a.tab = tab;
    a.tal = tal;
    a.sig_a = p->sig_a;
    a.kn = p->kn;
    a.wmin = p->w_min;
    a.v_inf = p->v_inf;
    a.Rs = Rs;
    a.seed = p->seed;
    a.max_segments = p->max_segments;
    a.cap = cap;
    a.cnt = cnt;
    a.nout = nout;
    if (mesh) {
        a.ni = 0;
        a.nj = 0;
        a.coff = mv.coff; a.corners = mv.corners; a.soff = mv.soff; a.sA = mv.sA; a.sB = mv.sB;
        a.scode = dcode;
        a.bout = dbout;
        a.nbout = (int)(bout.size() / 3);
        a.bin = dbin;
        a.nbin = (int)(bin.size() / 3);
    }
    const auto primaries = mesh ? mc::primary_kernel<true> : mc::primary_kernel<false>;
    const auto tracer = mesh ? mc::trace_kernel<true> : mc::trace_kernel<false>;
    hipEvent_t e0, e1;
    MC_HIP(hipEventCreate(&e0));
    MC_HIP(hipEventCreate(&e1));
    MC_HIP(hipEventRecord(e0, s));
    int gens = 0;
    int rc = GEL_OK;
    for (long long k0 = 0; k0 < p->n_primary && rc == GEL_OK; k0 += batch) {
        const long long nb = std::min(batch, p->n_primary - k0);
        mc::Atom *qin = qa, *qo = qb;
        a.qout = qo;
        if (hipMemsetAsync(nout, 0, sizeof(unsigned long long), s) != hipSuccess) { rc = GEL_E_HIP; break; }
        hipLaunchKernelGGL(primaries, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, a, k0, nb);
        int g = 0;
        for (;;) {
            unsigned long long n = 0;
            if (hipMemcpyAsync(&n, nout, sizeof(n), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) { rc = GEL_E_HIP; break; }
            if (n == 0) break;
            if (++g > kMaxGenerations) { rc = GEL_E_NUMERIC; break; }
            std::swap(qin, qo);      // the secondaries just born are traced next
            a.qin = qin;
            a.nin = std::min((long long)n, cap);
            a.qout = qo;
            if (hipMemsetAsync(nout, 0, sizeof(unsigned long long), s) != hipSuccess) { rc = GEL_E_HIP; break; }
            hipLaunchKernelGGL(tracer, dim3((unsigned)((a.nin + 255) / 256)), dim3(256), 0, s, a);
        }
        gens = std::max(gens, g);
    }
    (void)hipEventRecord(e1, s);
    if (rc != GEL_OK) {
        (void)hipStreamSynchronize(s);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return gel_set_error(rc, rc == GEL_E_NUMERIC ? "secondaries kept multiplying (lower the exchange rate or raise w_min)"
                                                     : "Monte Carlo launches");
    }
    const double Phi = mc::kPiM * Rs * Rs;
    double mean_speed[3];
    {   // inflow per unit density through the sphere: pi Rs^2 x the mean speed, Velosity_1(|v_inf|, 1)
        const double u = std::fabs(p->v_inf), cp = 1.0;
        const double e = std::exp(-u * u / (cp * cp)), f = std::erf(u / cp);
        mean_speed[0] = e * cp / mc::kSqrtPiR + (u + cp * cp / (2.0 * u)) * f;
    }
    hipLaunchKernelGGL(mesh ? mc::finalize_kernel<true> : mc::finalize_kernel<false>, gc, dim3(256), 0, s, a,
                       Phi * mean_speed[0], p->n_p_lism, p->n_primary, dout);
    MC_HIP(hipGetLastError());
    if (c->prm.source_mode == GEL_SRC_MK1 || c->prm.source_mode == GEL_SRC_HYBRID) {
        const size_t n = (size_t)GEL_MK1_NFIELDS * c->nc;
        if (!c->d_src && hipMalloc(&c->d_src, n * c->ts) != hipSuccess)
            return gel_set_error(GEL_E_NOMEM, "sources");
        const dim3 gi((unsigned)((n + 255) / 256));
        if (c->ts == 8) hipLaunchKernelGGL(mc::install_kernel<double>, gi, dim3(256), 0, s, dout, (double*)c->d_src, (long long)n);
        else hipLaunchKernelGGL(mc::install_kernel<float>, gi, dim3(256), 0, s, dout, (float*)c->d_src, (long long)n);
        MC_HIP(hipGetLastError());
    }
    unsigned long long h[mc::kCnt];
    MC_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    if (out) MC_HIP(hipMemcpyAsync(out, dout, GEL_MC_NFIELDS * (size_t)nc * sizeof(double), hipMemcpyDeviceToHost, s));
    MC_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) {
        st->primaries_in = (long long)h[0];
        st->atoms = (long long)h[1];
        st->segments = (long long)h[2];
        st->lost = (long long)h[3];
        st->dropped = (long long)h[4];
        st->generations = gens;
        st->ms = ms;
    }
    (void)t0;
    return GEL_OK;
}

}  // extern "C"
