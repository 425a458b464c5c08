// sweep_core.h - per-face and per-cell arithmetic of the Godunov sweep, shared by the
// CDNA4 kernels.  Floating-point expressions keep the operation order of the
// reference (cited per function) so that FP64 runs agree with it to rounding.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/gelios.h"

#define GEL_HD __host__ __device__ __forceinline__

namespace gel {

constexpr double kPi = 3.14159265358979323846;   // pi_ (Help.h:50)

template <class T> GEL_HD T mn(T a, T b) { return (b < a) ? b : a; }   // std::min
template <class T> GEL_HD T mx(T a, T b) { return (a < b) ? b : a; }   // std::max
template <class T> GEL_HD T sq(T a) { return a * a; }

// FP64 division and square root, bit for bit what the compiler's IEEE sequences give
// for operands in the normal range, in fewer instructions.  The compiler divides by
// v_div_scale (operand scaling for extreme exponents), v_rcp_f64 and two Newton steps
// on the reciprocal, the quotient a * r corrected once by its remainder, v_div_fmas
// and v_div_fixup (special operands): 11 instructions.  Where no scaling applies
// (|a| above 2^-969, exponents less than 768 apart, no denormal quotient) the scaling
// steps pass the operands through and the fix-up passes the quotient through, so the
// same reciprocal and correction alone (den, qt) give the same bits in 8, and a
// denominator shared by several quotients pays for its reciprocal once (3 more per
// quotient).  A zero numerator gives a zero (its sign may differ).  The callers'
// denominators are positive distances, areas, volumes, densities, wave speeds and
// wave-speed differences and gamma - 1, their numerators physical quantities.
struct Den {
    double b, r;        // the denominator and its refined reciprocal
};
__device__ __forceinline__ Den den(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    const double e0 = fma(-b, r, 1.0);
    r = fma(r, e0, r);
    const double e1 = fma(-b, r, 1.0);
    r = fma(r, e1, r);
    return Den{ b, r };
}
__device__ __forceinline__ double qt(double a, const Den& d)
{
    const double q = a * d.r;
    const double rem = fma(-d.b, q, a);
    return fma(rem, d.r, q);
}
// The compiler's square root is v_rsq_f64 with two Goldschmidt steps and a final
// correction, wrapped in a scaling of arguments below 2^-767 and a pass-through of
// zero and infinity: here the same steps without the scaling (the callers' arguments
// are zero or far above 2^-767: squared distances, pressures over densities).
__device__ __forceinline__ double sqn(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    const double d0 = fma(-g, g, x);
    h = fma(h, r, h);
    g = fma(d0, h, g);
    const double d1 = fma(-g, g, x);
    g = fma(d1, h, g);
    return x == 0.0 ? x : g;
}

// Division and square root inside the Riemann solvers: IEEE in FP64 (the
// reference's arithmetic, by den / qt and sqn); in FP32 (the FP32-bulk solvers) the
// hardware reciprocal and square root, 1 ulp each, instead of the correctly rounded
// sequences.
__device__ __forceinline__ double dv(double a, double b) { return qt(a, den(b)); }
__device__ __forceinline__ double sr(double x) { return sqn(x); }
__device__ __forceinline__ float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float sr(float x) { return __builtin_amdgcn_sqrtf(x); }

// Every buffer the kernels touch is in global memory, but kernels that read the
// buffers' addresses from a table in memory (the batched sweep, the fused source
// refresh) leave the compiler unable to tell their address space: it would emit flat
// instructions, which also count on the LDS counter, so that every wait for an LDS
// read would wait for the HBM traffic in flight as well.  The hot accesses go through
// gp(), the address typed as global.
template <class T> using GPtr = __attribute__((address_space(1))) T*;
template <class T> __device__ __forceinline__ GPtr<T> gp(T* p) { return (GPtr<T>)p; }

// Primitive gas state in the reference's order (Parametr, Cell.h:21-25).
template <class T> struct Gas { T ro, p, u, v, Q; };

// Help.cpp:80-94 and :67-78.
template <class T> GEL_HD T sgn(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }
template <class T> GEL_HD T minmod(T x, T y)
{
    const T s = sgn(x) + sgn(y);
    if (s == T(0)) return T(0);
    return (s / T(2)) * mn(fabs(x), fabs(y));
}
// FP32 (the FP32-bulk limiter): the same value, the smaller magnitude with x's sign
// where the sign bits agree and 0 where they differ (a zero operand gives 0 either way),
// in five instructions instead of the two sign functions.
template <> GEL_HD float minmod<float>(float x, float y)
{
    const float m = fminf(fabsf(x), fabsf(y));
    const bool same = (__builtin_bit_cast(int, x) ^ __builtin_bit_cast(int, y)) >= 0;
    return same ? copysignf(m, x) : 0.0f;
}
// minmod(a / p, b / q) with one division: each ratio's sign from the signs of its
// operands, the smaller magnitude chosen by the cross products |a| |q| and |b| |p|,
// and only that ratio divided out (|a| / |p| has the bits of |a / p|).  It returns what
// minmod does except where the two magnitudes agree to about an ulp, where either may
// come back.  p, q nonzero (the limiter's centre-distance differences).
template <class T> GEL_HD T minmod_ratio(T a, T p, T b, T q)
{
    if (a == T(0) || b == T(0)) return T(0);
    const bool nx = (a < T(0)) != (p < T(0)), ny = (b < T(0)) != (q < T(0));
    if (nx != ny) return T(0);
    const T fa = fabs(a), fp = fabs(p), fb = fabs(b), fq = fabs(q);
    const bool first = fa * fq < fb * fp;
    const T m = dv(first ? fa : fb, first ? fp : fq);      // the one division
    return nx ? -m : m;
}
// minmod_ratio(a, p, b, q) with the reciprocals of |p| and |q| given (FP64: den of
// fabs(p), fabs(q)), for the limiter, whose nine variables share a face side's two
// centre-distance differences.  The same value as minmod_ratio.
__device__ __forceinline__ double minmod_ratio_d(double a, const Den& p, bool p_neg, double b, const Den& q,
                                                 bool q_neg)
{
    if (a == 0.0 || b == 0.0) return 0.0;
    const bool nx = (a < 0.0) != p_neg, ny = (b < 0.0) != q_neg;
    if (nx != ny) return 0.0;
    const double fa = fabs(a), fb = fabs(b);
    const bool first = fa * q.b < fb * p.b;
    const double m = first ? qt(fa, p) : qt(fb, q);
    return nx ? -m : m;
}
// Help.cpp:96-113: minmod slope through (x1,t1),(x2,t2),(x3,t3), evaluated at y.
template <class T> GEL_HD T linear3(T x1, T t1, T x2, T t2, T x3, T t3, T y)
{
    const T d = minmod_ratio(t1 - t2, x1 - x2, t2 - t3, x2 - x3);
    return d * (y - x2) + t2;
}
// Help.cpp:37-65.
GEL_HD double polar_angle(double x, double y)
{
    if (fabs(x) + fabs(y) < 0.000001 / 390.222) return 0.0;
    if (x < 0) return atan(y / x) + 1.0 * kPi;
    if (x > 0 && y >= 0) return atan(y / x);
    if (x > 0 && y < 0) return atan(y / x) + 2.0 * kPi;
    if (y > 0) return kPi / 2.0;
    if (y < 0) return 3.0 * kPi / 2.0;
    return 0.0;
}
// Help.cpp:3-14.
GEL_HD void polar_perenos(double x1, double y1, double x2, double y2, double& u, double& v)
{
    const double a1 = polar_angle(x1, y1), a2 = polar_angle(x2, y2);
    const double fr = u * cos(a1) + v * sin(a1);
    const double ff = -u * sin(a1) + v * cos(a1);
    u = fr * cos(a2) - ff * sin(a2);
    v = fr * sin(a2) + ff * cos(a2);
}

// Cell::Get_Center on a quad (Cell.cpp:66-88).
GEL_HD void quad_center(const double* px, const double* py, double& cx, double& cy)
{
    double A = 0.0, Gx = 0.0, Gy = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int b = (k + 1) & 3;
        const double cr = px[k] * py[b] - py[k] * px[b];
        A = A + cr;
        Gx = Gx + (px[k] + px[b]) * cr;
        Gy = Gy + (py[k] + py[b]) * cr;
    }
    A = A * 0.5;
    const Den d6A = den(6.0 * A);
    cx = qt(Gx, d6A);
    cy = qt(Gy, d6A);
}
// Cell::Get_Volume for quads (Cell.cpp:146-166).
GEL_HD double quad_area(const double* px, const double* py)
{
    double a1 = px[0] - px[2], a2 = py[0] - py[2];
    const double a = sqn(sq(a1) + sq(a2));
    double b1 = px[1] - px[3], b2 = py[1] - py[3];
    const double b = sqn(sq(b1) + sq(b2));
    const Den da = den(a), db = den(b);
    a1 = qt(a1, da); a2 = qt(a2, da);
    b1 = qt(b1, db); b2 = qt(b2, db);
    const double c = a1 * b1 + a2 * b2;
    const double s = sqn(1 - sq(c));
    return 0.5 * a * b * s;
}
// Cell::Get_Volume_rotate(360) / (2 pi) (Cell.cpp:188-206, Setka.cpp:10187-10189).
GEL_HD double quad_volume_r(const double* px, const double* py)
{
    double A = 0.0, G = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int b = (k + 1) & 3;
        const double cr = px[k] * py[b] - py[k] * px[b];
        A = A + cr;
        G = G + (py[k] + py[b]) * cr;
    }
    A = A * 0.5;
    return qt(qt(fabs(A) * qt(1.0, den(6.0 * A)) * G * kPi * 360.0, den(180.0)), den(2.0 * kPi));
}

// Face flux of Setka::HLLC_2d_Korolkov_b_s (Setka.cpp:10484-10700), W = face speed
// along n.  Writes P[0..3] (mass, x-momentum, y-momentum, energy) and PQ (tracer);
// returns max(|SL|,|SR|), from which the caller forms cfl*dist/max.
template <class T, bool HLLC>
GEL_HD T hllc2d(const Gas<T>& L, const Gas<T>& R, T W, T n1, T n2, T gam, T P[4], T& PQ)
{
    const T t1 = n2, t2 = -n1;
    const T u1 = L.u * n1 + L.v * n2, v1 = L.u * t1 + L.v * t2;
    const T u2 = R.u * n1 + R.v * n2, v2 = R.u * t1 + R.v * t2;
    const T g1 = gam - T(1);
    const T cL = sr(dv(gam * L.p, L.ro));
    const T cR = sr(dv(gam * R.p, R.ro));
    const T kL = (sq(L.u) + sq(L.v)) / T(2);
    const T kR = (sq(R.u) + sq(R.v)) / T(2);
    const T SL = mn(u1, u2) - mx(cL, cR);
    const T SR = mx(u1, u2) + mx(cL, cR);
    const T UU = mx(fabs(SL), fabs(SR));
    const T eL = dv(L.p, g1) + L.ro * kL;
    const T eR = dv(R.p, g1) + R.ro * kR;
    const T FL[5] = { L.ro * u1, L.ro * u1 * u1 + L.p, L.ro * u1 * v1, (eL + L.p) * u1, L.Q * u1 };
    const T FR[5] = { R.ro * u2, R.ro * u2 * u2 + R.p, R.ro * u2 * v2, (eR + R.p) * u2, R.Q * u2 };
    const T UL[5] = { L.ro, L.ro * u1, L.ro * v1, eL, L.Q };
    const T UR[5] = { R.ro, R.ro * u2, R.ro * v2, eR, R.Q };
    if (SL >= W || SR <= W) {           // supersonic, Setka.cpp:10569-10589
        const bool left = SL >= W;
        T F[5], U[5];
#pragma unroll
        for (int k = 0; k < 5; k++) { F[k] = left ? FL[k] : FR[k]; U[k] = left ? UL[k] : UR[k]; }
        P[1] = n1 * (F[1] - W * U[1]) + t1 * (F[2] - W * U[2]);
        P[2] = n2 * (F[1] - W * U[1]) + t2 * (F[2] - W * U[2]);
        P[0] = F[0] - W * U[0];
        P[3] = F[3] - W * U[3];
        PQ = F[4] - W * U[4];
        return UU;
    }
    const T suR = SR - u2, suL = SL - u1;
    const T SM = dv(suR * R.ro * u2 - suL * L.ro * u1 - R.p + L.p, suR * R.ro - suL * L.ro);
    T UZ[5];
#pragma unroll
    for (int k = 0; k < 5; k++) UZ[k] = dv(SR * UR[k] - SL * UL[k] + FL[k] - FR[k], SR - SL);
    T PO[5];
    if (!HLLC) {                        // HLL, Setka.cpp:10606-10618
#pragma unroll
        for (int k = 0; k < 5; k++)
            PO[k] = dv(SR * FL[k] - SL * FR[k] + SR * SL * (UR[k] - UL[k]), SR - SL);
        P[1] = n1 * (PO[1] - W * UZ[1]) + t1 * (PO[2] - W * UZ[2]);
        P[2] = n2 * (PO[1] - W * UZ[1]) + t2 * (PO[2] - W * UZ[2]);
        P[0] = PO[0] - W * UZ[0];
        P[3] = PO[3] - W * UZ[3];
        PQ = PO[4] - W * UZ[4];
        return UU;
    }
    // HLLC, Setka.cpp:10621-10697 (nul_potok = false on the sweep's faces)
    const T roLL = dv(L.ro * (SL - u1), SL - SM);
    const T roRR = dv(R.ro * (SR - u2), SR - SM);
    const T QLL = dv(L.Q * (SL - u1), SL - SM);
    const T QRR = dv(R.Q * (SR - u2), SR - SM);
    const T suRm = dv(suR, SR - SM), suLm = dv(suL, SL - SM);
    const T ptzR = R.p + R.ro * suR * (SM - u2);
    const T ptzL = L.p + L.ro * suL * (SM - u1);
    const T ptz = (ptzR + ptzL) / T(2);
    const T ee2 = eR * suRm + dv(ptz * SM - R.p * u2, SR - SM);
    const T ee1 = eL * suLm + dv(ptz * SM - L.p * u1, SL - SM);
    if (SM >= W) {
        const T ULL[5] = { roLL, roLL * SM, roLL * v1, ee1, QLL };
#pragma unroll
        for (int k = 0; k < 5; k++) PO[k] = FL[k] + SL * ULL[k] - SL * UL[k] - W * ULL[k];
    } else {
        const T URR[5] = { roRR, roRR * SM, roRR * v2, ee2, QRR };
#pragma unroll
        for (int k = 0; k < 5; k++) PO[k] = FR[k] + SR * URR[k] - SR * UR[k] - W * URR[k];
    }
    P[1] = n1 * PO[1] + t1 * PO[2];
    P[2] = n2 * PO[1] + t2 * PO[2];
    P[0] = PO[0];
    P[3] = PO[3];
    PQ = PO[4];
    return UU;
}

// HLLC with the reference's nul_potok option (Setka.cpp:10644-10653): the star states
// of both sides take the tangential velocity UZ2/UZ0 of the HLL average state.  The
// sweep sets it on faces with exactly one node on a fitted surface
// (Setka.cpp:10300-10327).  Supersonic faces take the upwind flux as without it.
template <class T>
GEL_HD T hllc2d_nul(const Gas<T>& L, const Gas<T>& R, T W, T n1, T n2, T gam, T P[4], T& PQ)
{
    const T t1 = n2, t2 = -n1;
    const T u1 = L.u * n1 + L.v * n2, v1 = L.u * t1 + L.v * t2;
    const T u2 = R.u * n1 + R.v * n2, v2 = R.u * t1 + R.v * t2;
    const T cL = sqrt(gam * L.p / L.ro);
    const T cR = sqrt(gam * R.p / R.ro);
    const T SL = mn(u1, u2) - mx(cL, cR);
    const T SR = mx(u1, u2) + mx(cL, cR);
    if (SL >= W || SR <= W) return hllc2d<T, true>(L, R, W, n1, n2, gam, P, PQ);
    const T suR = SR - u2, suL = SL - u1;
    const T SM = (suR * R.ro * u2 - suL * L.ro * u1 - R.p + L.p) / (suR * R.ro - suL * L.ro);
    const T g1 = gam - T(1);
    const T kL = (sq(L.u) + sq(L.v)) / T(2), kR = (sq(R.u) + sq(R.v)) / T(2);
    const T eL = L.p / g1 + L.ro * kL, eR = R.p / g1 + R.ro * kR;
    const T FL[5] = { L.ro * u1, L.ro * u1 * u1 + L.p, L.ro * u1 * v1, (eL + L.p) * u1, L.Q * u1 };
    const T FR[5] = { R.ro * u2, R.ro * u2 * u2 + R.p, R.ro * u2 * v2, (eR + R.p) * u2, R.Q * u2 };
    const T UL[5] = { L.ro, L.ro * u1, L.ro * v1, eL, L.Q };
    const T UR[5] = { R.ro, R.ro * u2, R.ro * v2, eR, R.Q };
    const T UZ0 = (SR * UR[0] - SL * UL[0] + FL[0] - FR[0]) / (SR - SL);
    const T UZ2 = (SR * UR[2] - SL * UL[2] + FL[2] - FR[2]) / (SR - SL);
    const T vv = UZ2 / UZ0;
    const T suRm = suR / (SR - SM), suLm = suL / (SL - SM);
    const T ptz = ((R.p + R.ro * suR * (SM - u2)) + (L.p + L.ro * suL * (SM - u1))) / T(2);
    T PO[5];
    if (SM >= W) {
        const T r = L.ro * (SL - u1) / (SL - SM);
        const T ULL[5] = { r, r * SM, r * vv, eL * suLm + (ptz * SM - L.p * u1) / (SL - SM), L.Q * (SL - u1) / (SL - SM) };
        for (int k = 0; k < 5; k++) PO[k] = FL[k] + SL * ULL[k] - SL * UL[k] - W * ULL[k];
    } else {
        const T r = R.ro * (SR - u2) / (SR - SM);
        const T URR[5] = { r, r * SM, r * vv, eR * suRm + (ptz * SM - R.p * u2) / (SR - SM), R.Q * (SR - u2) / (SR - SM) };
        for (int k = 0; k < 5; k++) PO[k] = FR[k] + SR * URR[k] - SR * UR[k] - W * URR[k];
    }
    P[1] = n1 * PO[1] + t1 * PO[2];
    P[2] = n2 * PO[1] + t2 * PO[2];
    P[0] = PO[0];
    P[3] = PO[3];
    PQ = PO[4];

    return mx(fabs(SL), fabs(SR));
}

// Flux into the Gas-independent form used by the batch test entry point.
template <class T>
GEL_HD T hllc2d_full(const Gas<T>& L, const Gas<T>& R, T W, T n1, T n2, T rad, T cfl, T gam,
                     int metod, int nul_potok, T P[4], T& PQ, T& Vl, T& Vc, T& Vp)
{
    const T t1 = n2, t2 = -n1;
    const T u1 = L.u * n1 + L.v * n2, v1 = L.u * t1 + L.v * t2;
    const T u2 = R.u * n1 + R.v * n2, v2 = R.u * t1 + R.v * t2;
    const T cL = sqrt(gam * L.p / L.ro);
    const T cR = sqrt(gam * R.p / R.ro);
    const T SL = mn(u1, u2) - mx(cL, cR);
    const T SR = mx(u1, u2) + mx(cL, cR);
    Vl = SL;
    Vp = SR;
    const T suR = SR - u2, suL = SL - u1;
    Vc = (suR * R.ro * u2 - suL * L.ro * u1 - R.p + L.p) / (suR * R.ro - suL * L.ro);
    T UU;
    if (metod == 0) {
        UU = hllc2d<T, false>(L, R, W, n1, n2, gam, P, PQ);
    } else if (!nul_potok || SL >= W || SR <= W) {
        UU = hllc2d<T, true>(L, R, W, n1, n2, gam, P, PQ);
    } else {
        UU = hllc2d_nul<T>(L, R, W, n1, n2, gam, P, PQ);
    }
    return cfl * rad / UU;
}

// Cell::Get_Sourse_MK1, own-cell moments (Cell.cpp:590-664).  f(k) returns field k.
// R is the arithmetic type: double for the FP64 sweep (the reference's arithmetic);
// float for the FP32-bulk sweep, whose moments are stored in FP32, with the
// hardware's reciprocal, square root and logarithm (a source term's relative error
// of a few 1e-7 enters the step as dt * q, below the rounding of the FP32 state).
template <class F, class R = double>
GEL_HD void source_mk1(const gel_params& prm, F f, R u, R v, R ro, R p, R& q1, R& q2, R& q3)
{
    R nu[4], uH[4], vH[4], roH[4], pH[4], UM[4], UU[4];
    const R half = R(0.5), two = R(2.0);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uH[k] = R(f(4 + k));
        vH[k] = R(f(8 + k));
        roH[k] = R(f(k));
        pH[k] = half * R(f(12 + k)) * R(f(k));
        if (roH[k] <= R(0)) { roH[k] = R(0.0000001); pH[k] = R(0); }
    }
    const R c_m = R(64.0 / (9.0 * kPi)), c_u = R(4.0 / kPi);
    const R sig_a = R(prm.sig_a), sig_b = R(prm.sig_b);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        UM[k] = sr(sq(u - uH[k]) + sq(v - vH[k]) + c_m * (dv(p, ro) + dv(two * pH[k], roH[k])));
        UU[k] = sr(sq(u - uH[k]) + sq(v - vH[k]) + c_u * (dv(p, ro) + dv(two * pH[k], roH[k])));
        R lg;
        if constexpr (sizeof(R) == 4) lg = __logf(UM[k]);
        else lg = log(UM[k]);
        const R sig = sq(sig_a - sig_b * lg);
        nu[k] = ro * roH[k] * UM[k] * sig;
    }
    R Mu = nu[0] * (uH[0] - u) + nu[1] * (uH[1] - u) + nu[2] * (uH[2] - u) + nu[3] * (uH[3] - u);
    R Mv = nu[0] * (vH[0] - v) + nu[1] * (vH[1] - v) + nu[2] * (vH[2] - v) + nu[3] * (vH[3] - v);
    R t[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        t[k] = nu[k] * ((sq(uH[k]) + sq(vH[k]) - sq(u) - sq(v)) / two + dv(UU[k], UM[k]) * (dv(two * pH[k], roH[k]) - dv(p, ro)));
    R MT = t[0] + t[1] + t[2] + t[3];
    const R coef = R(prm.src_coef);
    Mu = coef * Mu;
    Mv = coef * Mv;
    MT = coef * MT;
    q1 = Mu * R(f(16));
    q2 = Mv * R(f(17));
    q3 = MT * R(f(18));
}

}  // namespace gel

namespace gel {

// HLLD of Solvers::HLLDQ_Korolkov metod 3 (Solvers.cpp:1642-2185) for an in-plane
// normal n = (n1, n2, 0) and a face moving with speed W along n: the fan is sampled
// at xi = W and each region's flux is taken minus W*U, as the gas flux does
// (Setka.cpp:10569-10618).  Basis t = (-n2, n1, 0), m = (0, 0, 1).  Primitive input
// q = (ro, p, u, v, w, bx, by, bz, Q), B Gaussian.  P: mass, momentum x/y/phi,
// energy, B x/y/phi, tracer.  Returns max(|SL|,|SR|).
template <class T>
GEL_HD T hlld(const T* L, const T* R, T W, T n1, T n2, T gam, T P[9])
{
    const T spi4 = T(3.5449077018110318);    // sqrt(4 pi), Solvers.h:9-12
    const T g1 = gam - T(1);
    const T t1 = -n2, t2 = n1;
    const T ro_L = L[0], p_L = L[1], Q_L = L[8], ro_R = R[0], p_R = R[1], Q_R = R[8];
    const T bx_L = dv(L[5], spi4), by_L = dv(L[6], spi4), bz_L = dv(L[7], spi4);
    const T bx_R = dv(R[5], spi4), by_R = dv(R[6], spi4), bz_R = dv(R[7], spi4);
    // three-term dot products with the zero components kept, as the reference forms them
    const T z = T(0), o = T(1);
    const T u1 = L[2] * n1 + L[3] * n2 + L[4] * z;
    const T v1 = L[2] * t1 + L[3] * t2 + L[4] * z;
    const T w1 = L[2] * z + L[3] * z + L[4] * o;
    const T u2 = R[2] * n1 + R[3] * n2 + R[4] * z;
    const T v2 = R[2] * t1 + R[3] * t2 + R[4] * z;
    const T w2 = R[2] * z + R[3] * z + R[4] * o;
    const T bn1 = bx_L * n1 + by_L * n2 + bz_L * z;
    const T bt1 = bx_L * t1 + by_L * t2 + bz_L * z;
    const T bm1 = bx_L * z + by_L * z + bz_L * o;
    const T bn2 = bx_R * n1 + by_R * n2 + bz_R * z;
    const T bt2 = bx_R * t1 + by_R * t2 + bz_R * z;
    const T bm2 = bx_R * z + by_R * z + bz_R * o;
    const T bb_L = sq(bx_L) + sq(by_L) + sq(bz_L);
    const T bb_R = sq(bx_R) + sq(by_R) + sq(bz_R);
    const T uu_L = (sq(L[2]) + sq(L[3]) + sq(L[4])) / T(2);
    const T uu_R = (sq(R[2]) + sq(R[3]) + sq(R[4])) / T(2);
    const T cfL = sr(dv(gam * p_L + bb_L + sr(sq(gam * p_L + bb_L) - T(4) * gam * p_L * sq(bn1)), T(2) * ro_L));
    const T cfR = sr(dv(gam * p_R + bb_R + sr(sq(gam * p_R + bb_R) - T(4) * gam * p_R * sq(bn2)), T(2) * ro_R));
    const T SL = mn(u1, u2) - mx(cfL, cfR);
    const T SR = mx(u1, u2) + mx(cfL, cfR);
    const T pTL = p_L + bb_L / T(2);
    const T pTR = p_R + bb_R / T(2);
    const T UU = mx(fabs(SL), fabs(SR));
    const T e1 = dv(p_L, g1) + ro_L * uu_L + bb_L / T(2);
    const T e2 = dv(p_R, g1) + ro_R * uu_R + bb_R / T(2);
    const T FL[9] = { ro_L * u1, ro_L * u1 * u1 + pTL - sq(bn1), ro_L * u1 * v1 - bn1 * bt1, ro_L * u1 * w1 - bn1 * bm1,
                      (e1 + pTL) * u1 - bn1 * (u1 * bn1 + v1 * bt1 + w1 * bm1), T(0), u1 * bt1 - v1 * bn1,
                      u1 * bm1 - w1 * bn1, Q_L * u1 };
    const T FR[9] = { ro_R * u2, ro_R * u2 * u2 + pTR - sq(bn2), ro_R * u2 * v2 - bn2 * bt2, ro_R * u2 * w2 - bn2 * bm2,
                      (e2 + pTR) * u2 - bn2 * (u2 * bn2 + v2 * bt2 + w2 * bm2), T(0), u2 * bt2 - v2 * bn2,
                      u2 * bm2 - w2 * bn2, Q_R * u2 };
    const T UL[9] = { ro_L, ro_L * u1, ro_L * v1, ro_L * w1, e1, bn1, bt1, bm1, Q_L };
    const T UR[9] = { ro_R, ro_R * u2, ro_R * v2, ro_R * w2, e2, bn2, bt2, bm2, Q_R };
    // Regions of the fan at xi = W.  Both upwind regions are tested before the star
    // regions, then the side of the contact: the Alfven speeds come from the HLL normal
    // field bn, which the two sides' fast speeds need not bound (SLL < SL or SRR > SR),
    // and the reference's order (L, L*, L**, R**, R*, R) would then sample a star region
    // for (L, R, n, W) and the upwind state for the mirrored (R, L, -n, -W).  Here every
    // branch is mirror-symmetric, so a face solved once gives what a cell loop solving
    // it from both sides gives, to rounding; on monotone fans the region is the
    // reference's (Solvers.cpp:2106-2156).
    T PO[9];
    T bnW;
    if (SL >= W) {
#pragma unroll
        for (int i = 0; i < 9; i++) PO[i] = FL[i] - W * UL[i];
        bnW = UL[5];
    } else if (SR <= W) {
#pragma unroll
        for (int i = 0; i < 9; i++) PO[i] = FR[i] - W * UR[i];
        bnW = UR[5];
    } else {
        // subsonic faces only: the star states (a wavefront of supersonic faces, the
        // radial faces of a row in the wind, skips all of it)
        const T suR = SR - u2, suL = SL - u1;
        const T SM = dv(suR * ro_R * u2 - suL * ro_L * u1 - pTR + pTL, suR * ro_R - suL * ro_L);
        const T PTT = dv(suR * ro_R * pTL - suL * ro_L * pTR + ro_L * ro_R * suR * suL * (u2 - u1), suR * ro_R - suL * ro_L);
        const T bn = dv(SR * UR[5] - SL * UL[5] + FL[5] - FR[5], SR - SL);
        const T bbn = bn * bn;
        const T ro_LL = dv(ro_L * (SL - u1), SL - SM);
        const T ro_RR = dv(ro_R * (SR - u2), SR - SM);
        const T Q_LL = dv(Q_L * (SL - u1), SL - SM);
        const T Q_RR = dv(Q_R * (SR - u2), SR - SM);
        const T ttL = ro_L * suL * (SL - SM) - bbn;
        const T ttR = ro_R * suR * (SR - SM) - bbn;
        T vLL = v1, wLL = w1, btt1 = T(0), bmm1 = T(0);
        T vRR = v2, wRR = w2, btt2 = T(0), bmm2 = T(0);
        if (fabs(ttL) >= T(0.000001)) {
            vLL = v1 - dv(bn * bt1 * (SM - u1), ttL);
            wLL = w1 - dv(bn * bm1 * (SM - u1), ttL);
            btt1 = dv(bt1 * (ro_L * suL * suL - bbn), ttL);
            bmm1 = dv(bm1 * (ro_L * suL * suL - bbn), ttL);
        }
        if (fabs(ttR) >= T(0.000001)) {
            vRR = v2 - dv(bn * bt2 * (SM - u2), ttR);
            wRR = w2 - dv(bn * bm2 * (SM - u2), ttR);
            btt2 = dv(bt2 * (ro_R * suR * suR - bbn), ttR);
            bmm2 = dv(bm2 * (ro_R * suR * suR - bbn), ttR);
        }
        const T eLL = dv(e1 * suL + PTT * SM - pTL * u1 + bn * ((u1 * bn1 + v1 * bt1 + w1 * bm1) - (SM * bn + vLL * btt1 + wLL * bmm1)), SL - SM);
        const T eRR = dv(e2 * suR + PTT * SM - pTR * u2 + bn * ((u2 * bn2 + v2 * bt2 + w2 * bm2) - (SM * bn + vRR * btt2 + wRR * bmm2)), SR - SM);
        const T sLL = sr(ro_LL), sRR = sr(ro_RR);
        T SLL = SM - dv(fabs(bn), sLL);
        T SRR = SM + dv(fabs(bn), sRR);
        T idbn;
        if (fabs(bn) > T(0.001)) {
            idbn = T(1) * sgn(bn);
        } else {
            idbn = T(0); SLL = SM; SRR = SM;
        }
        const T vLLL = dv(sLL * vLL + sRR * vRR + idbn * (btt2 - btt1), sLL + sRR);
        const T wLLL = dv(sLL * wLL + sRR * wRR + idbn * (bmm2 - bmm1), sLL + sRR);
        const T bttt = dv(sLL * btt2 + sRR * btt1 + idbn * sLL * sRR * (vRR - vLL), sLL + sRR);
        const T bmmm = dv(sLL * bmm2 + sRR * bmm1 + idbn * sLL * sRR * (wRR - wLL), sLL + sRR);
        const T eLLL = eLL - idbn * sLL * ((SM * bn + vLL * btt1 + wLL * bmm1) - (SM * bn + vLLL * bttt + wLLL * bmmm));
        const T eRRR = eRR + idbn * sRR * ((SM * bn + vRR * btt2 + wRR * bmm2) - (SM * bn + vLLL * bttt + wLLL * bmmm));
        if (SM >= W && SLL >= W) {
            const T ULL[9] = { ro_LL, ro_LL * SM, ro_LL * vLL, ro_LL * wLL, eLL, bn, btt1, bmm1, Q_LL };
#pragma unroll
            for (int i = 0; i < 9; i++) PO[i] = FL[i] + SL * ULL[i] - SL * UL[i] - W * ULL[i];
            bnW = bn;
        } else if (SM >= W) {
            const T ULL[9] = { ro_LL, ro_LL * SM, ro_LL * vLL, ro_LL * wLL, eLL, bn, btt1, bmm1, Q_LL };
            const T ULLL[9] = { ro_LL, ro_LL * SM, ro_LL * vLLL, ro_LL * wLLL, eLLL, bn, bttt, bmmm, Q_LL };
#pragma unroll
            for (int i = 0; i < 9; i++) PO[i] = FL[i] + SLL * ULLL[i] - (SLL - SL) * ULL[i] - SL * UL[i] - W * ULLL[i];
            bnW = bn;
        } else if (SRR <= W) {
            const T URR[9] = { ro_RR, ro_RR * SM, ro_RR * vRR, ro_RR * wRR, eRR, bn, btt2, bmm2, Q_RR };
#pragma unroll
            for (int i = 0; i < 9; i++) PO[i] = FR[i] + SR * URR[i] - SR * UR[i] - W * URR[i];
            bnW = bn;
        } else {
            const T URR[9] = { ro_RR, ro_RR * SM, ro_RR * vRR, ro_RR * wRR, eRR, bn, btt2, bmm2, Q_RR };
            const T URRR[9] = { ro_RR, ro_RR * SM, ro_RR * vLLL, ro_RR * wLLL, eRRR, bn, bttt, bmmm, Q_RR };
#pragma unroll
            for (int i = 0; i < 9; i++) PO[i] = FR[i] + SRR * URRR[i] - (SRR - SR) * URR[i] - SR * UR[i] - W * URRR[i];
            bnW = bn;
        }
    }
    const T SN = mx(fabs(SL), fabs(SR));
    PO[5] = -SN * (bn2 - bn1) - W * bnW;
    P[0] = PO[0];
    P[1] = n1 * PO[1] + t1 * PO[2] + z * PO[3];
    P[2] = n2 * PO[1] + t2 * PO[2] + z * PO[3];
    P[3] = z * PO[1] + z * PO[2] + o * PO[3];
    P[4] = PO[4];
    P[5] = spi4 * (n1 * PO[5] + t1 * PO[6] + z * PO[7]);
    P[6] = spi4 * (n2 * PO[5] + t2 * PO[6] + z * PO[7]);
    P[7] = spi4 * (z * PO[5] + z * PO[6] + o * PO[7]);
    P[8] = PO[8];
    return UU;
}

}  // namespace gel

namespace gel {

// Upwind (supersonic) fluxes of the FP32-bulk faces, the branches SL >= W / SR <= W
// of HLLC_2d_Korolkov_b_s (Setka.cpp:10569-10589) and of the HLLD below
// (Solvers.cpp:2106-2113, :2150-2156), times the face area.  The flux of a state
// through a face is linear in the face's area-weighted normal A = S n, so it is
// formed in Cartesian components from A and the swept-volume rate WS = W S, both
// FP64, with no unit normal: the products entering the cell sums carry no rounding
// of a normalisation, and a uniform stream through a closed cell stays balanced to
// FP64 rounding.  P: the cell accumulators' order; ig1 = 1 / (gamma - 1).
template <class T>
GEL_HD void gas_upwind_area(const Gas<T>& S, T A1, T A2, T WS, T ig1, T P[5])
{
    const T uS = S.u * A1 + S.v * A2;                   // normal velocity x S
    const T e = S.p * ig1 + S.ro * ((sq(S.u) + sq(S.v)) * T(0.5));
    const T m = S.ro * uS;
    P[0] = m - WS * S.ro;
    P[1] = m * S.u + S.p * A1 - WS * (S.ro * S.u);
    P[2] = m * S.v + S.p * A2 - WS * (S.ro * S.v);
    P[3] = (e + S.p) * uS - WS * e;
    P[4] = S.Q * uS - WS * S.Q;
}

// The MHD upwind flux of state q (the side the whole fan leaves behind; o is the other
// side) times the area, Gaussian B: mass, momentum x/y/phi, energy, B x/y/phi, tracer.
// The normal-field term of the HLLD, -SN (Bn_R - Bn_L) along n (Solvers.cpp:2161-2163),
// is the one part that needs the unit normal (n1, n2).
template <class T>
GEL_HD void mhd_upwind_area(const T* q, const T* o, T A1, T A2, T WS, T n1, T n2, T ig1, T SN, bool q_is_left,
                            T P[9])
{
    const T i4pi = T(0.07957747154594767), i8pi = T(0.039788735772973836);   // 1 / (4 pi), 1 / (8 pi)
    const T ro = q[0], p = q[1], u = q[2], v = q[3], w = q[4], bx = q[5], by = q[6], bz = q[7], Q = q[8];
    const T uS = u * A1 + v * A2;                       // normal velocity x S
    const T BA = bx * A1 + by * A2;                     // normal field x S
    const T BB = sq(bx) + sq(by) + sq(bz);
    const T pT = p + BB * i8pi;
    const T e = p * ig1 + ro * ((sq(u) + sq(v) + sq(w)) * T(0.5)) + BB * i8pi;
    const T VB = u * bx + v * by + w * bz;
    const T BAi = BA * i4pi;
    const T dB = q_is_left ? ((o[5] - bx) * n1 + (o[6] - by) * n2) : ((bx - o[5]) * n1 + (by - o[6]) * n2);
    const T m = ro * uS;
    P[0] = m - WS * ro;
    P[1] = m * u + pT * A1 - bx * BAi - WS * (ro * u);
    P[2] = m * v + pT * A2 - by * BAi - WS * (ro * v);
    P[3] = m * w - bz * BAi - WS * (ro * w);
    P[4] = (e + pT) * uS - VB * BAi - WS * e;
    P[5] = uS * bx - BA * u - WS * bx - SN * dB * A1;
    P[6] = uS * by - BA * v - WS * by - SN * dB * A2;
    P[7] = uS * bz - BA * w - WS * bz;
    P[8] = Q * uS - WS * Q;
}

}  // namespace gel

namespace gel {

// ---- exact Riemann solver for a gamma-gas with a moving face ---------------------
// Solvers::Godunov_Solver_Alexashov (Solvers.cpp:693-1172) with its pressure solvers
// newton (:1174-1229), devtwo (:1231-1305) and lev (:1307-1321).  States are
// (ro, p, u, v, w) in Cartesian components; n = (n1, n2, n3) with the tangents
// t1 = (n2, -n1, 0), t2 = (0, 0, 1) of the reference.  The pattern (vacuum, two
// rarefactions, shock + rarefaction, two shocks) follows from the velocity jump
// against the thresholds Uvk < Urz < Uud after ordering the sides so that pI <= pII;
// the star pressure is closed-form, bisection or Newton accordingly; the solution is
// sampled at xi = w and the flux is taken relative to the face.  Returns
// cfl * dist / max|wave speed| (the reference's local time step); status 1 when an
// iteration does not converge or meets a non-positive pressure (the reference exits).
struct GodunovK {
    double ga, g1, g2, gp, gm;
};
GEL_HD GodunovK godunov_k(double ga)
{
    GodunovK k;
    k.ga = ga;
    k.g1 = ga - 1.0;
    k.g2 = ga + 1.0;
    k.gp = (k.g2 / ga) / 2.0;
    k.gm = (k.g1 / ga) / 2.0;
    return k;
}

// lev: f(P) = fI(P) + fII(P) - (enI - enII), shock branch on side I, rarefaction on II
GEL_HD double godunov_lev(const GodunovK& K, double enI, double pI, double rI, double enII, double pII, double rII,
                          double P)
{
    const double cI = sqrt(K.ga * pI / rI);
    const double cII = sqrt(K.ga * pII / rII);
    const double fI = (P - pI) / (rI * cI * sqrt(K.gp * P / pI + K.gm));
    const double fII = 2.0 / K.g1 * cII * (pow((P / pII), K.gm) - 1.0);
    return (fI + fII) - (enI - enII);
}

// Where devtwo's bisection of lev on [pI, pII] may stop.  lev is increasing and
// concave on [pI, pII] (its shock term has a negative second derivative, its
// rarefaction term is a power below one); as a function of ln P it is convex (its
// terms grow as powers of P), so Newton's iteration in ln P from the chord's zero
// (right of the root) descends to the root without overshooting, in a handful of
// steps for pressure ratios up to 1e6.  That iteration (kept inside the bracket it
// narrows, its step's exponential in FP32: it only locates the root) ends within
// tol / 8 of the root in lev; around the point it
// reaches, the interval (a, b) is widened until lev, evaluated as the bisection
// evaluates it, is below -(tol + 2E) at a and above tol + 2E at b, where E bounds the
// rounding error of one evaluation anywhere on [pI, pII] (fifty times the worst case
// of its terms: the shock term below |lev(pII)| + |Udf|, the rarefaction term's factor
// 2 cII / g1 times the power's one-ulp error, the velocity jump).  Then at any P <= a
// the evaluated lev is below -tol and at any P >= b above tol, so the bisection's
// outcome at such a P is known without evaluating it.  false (a = -inf, b = +inf):
// nothing is known and every midpoint is evaluated.
GEL_HD bool godunov_bracket(const GodunovK& K, double enI, double pI, double rI, double enII, double pII, double rII,
                            double flo, double fhi, double tol, double& a, double& b)
{
    a = -__builtin_huge_val();
    b = __builtin_huge_val();
    const double cI = sqrt(K.ga * pI / rI), cII = sqrt(K.ga * pII / rII);
    const double rcI = rI * cI, kII = 2.0 / K.g1 * cII, Udf = enI - enII;
    const double E = 1e-13 * (fabs(flo) + fabs(fhi) + 4.0 * (fabs(enI) + fabs(enII)) + kII);
    const double D = tol + 2.0 * E;
    double L = pI, R = pII;
    double x = pI - flo * (pII - pI) / (fhi - flo);
    if (!(x > L && x < R)) x = 0.5 * (L + R);
    double d = 0.0;
    for (int it = 0; it < 16; it++) {
        const double A = K.gp * x / pI + K.gm, sA = sqrt(A);
        const double xg = pow(x / pII, K.gm);
        const double f = (x - pI) / (rcI * sA) + kII * (xg - 1.0) - Udf;
        d = (K.gp * x / (2.0 * pI) + K.gm + K.gp / 2.0) / (rcI * A * sA) + kII * K.gm * xg / x;
        if (!(d > 0.0) || !(f == f)) return false;
        if (f < 0.0) L = x;
        else R = x;
        if (fabs(f) <= 0.125 * tol) break;
        const double del = f / (x * d);                 // the Newton step in ln P
        double xn = fabs(del) < 1e-3 ? x * (1.0 - del + 0.5 * del * del) : x * (double)expf((float)-del);
        if (!(xn > L && xn < R)) xn = 0.5 * (L + R);
        if (xn == x) break;
        x = xn;
    }
    double h = 2.0 * D / d;
    for (int t = 0; t < 3; t++, h *= 4.0) {
        const double lo = x - h, hi = x + h;
        const double fl = lo > pI ? godunov_lev(K, enI, pI, rI, enII, pII, rII, lo) : flo;
        const double fh = hi < pII ? godunov_lev(K, enI, pI, rI, enII, pII, rII, hi) : fhi;
        if (fl < -D && fh > D) {
            a = lo > pI ? lo : pI;
            b = hi < pII ? hi : pII;
            return true;
        }
    }
    return false;
}

// devtwo: bisection of lev on [pI, pII].  The midpoints, the stopping test and the
// returned point are the reference's; a midpoint outside the bracket of
// godunov_bracket is not evaluated, its lev taking the sign it certainly has (only
// the sign enters the update: |flo|, |fm| > tol, so flo * fm cannot underflow to 0).
GEL_HD int godunov_bisect(const GodunovK& K, double enI, double pI, double rI, double enII, double pII, double rII,
                          double& p)
{
    const double tol = 10e-10;
    double lo = pI, hi = pII;
    double flo = godunov_lev(K, enI, pI, rI, enII, pII, rII, lo);
    const double fhi = godunov_lev(K, enI, pI, rI, enII, pII, rII, hi);
    if (fabs(flo) <= tol) { p = lo; return 0; }
    if (fabs(fhi) <= tol) { p = hi; return 0; }
    double a, b;
    godunov_bracket(K, enI, pI, rI, enII, pII, rII, flo, fhi, tol, a, b);
    for (int it = 1; it < 1100; it++) {
        const double mid = (lo + hi) / 2.0;
        const double fm = mid <= a ? -1.0 : mid >= b ? 1.0 : godunov_lev(K, enI, pI, rI, enII, pII, rII, mid);
        if (fabs(fm) <= tol) { p = mid; return 0; }
        if ((flo * fm) <= 0.0) {
            hi = mid;
        } else {
            lo = mid;
            flo = fm;
        }
    }
    p = 0.0;
    return 1;
}

// Lane l's value of v, l the same in every lane (two v_readlane_b32).
__device__ __forceinline__ double lane_value(double v, int l)
{
    l = __builtin_amdgcn_readfirstlane(l);
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// devtwo by one wavefront whose 64 lanes hold the same problem: the bisection's next
// six midpoints along whichever path it takes are the 63 nodes of a binary tree
// (node n's children 2n, the lower half, and 2n + 1); lane n - 1 reaches node n by the
// halvings its path takes ((lo + hi) / 2 at every level, as the bisection computes
// them) and evaluates lev there, all lanes at once; then the wavefront walks the path
// with those values, six of the bisection's iterations for one evaluation's latency.
// The midpoints, values, stopping test and iteration count are the bisection's.
// Every lane of the wavefront must call it with identical arguments.
__device__ __forceinline__ int godunov_bisect_wave(const GodunovK& K, double enI, double pI, double rI, double enII,
                                                   double pII, double rII, double& p)
{
    const double tol = 10e-10;
    double lo = pI, hi = pII;
    double flo = godunov_lev(K, enI, pI, rI, enII, pII, rII, lo);
    const double fhi = godunov_lev(K, enI, pI, rI, enII, pII, rII, hi);
    if (fabs(flo) <= tol) { p = lo; return 0; }
    if (fabs(fhi) <= tol) { p = hi; return 0; }
    const int n = min((int)(threadIdx.x & 63) + 1, 63);      // this lane's tree node
    const int depth = 31 - __builtin_clz(n);
    int it = 1;
    while (it < 1100) {
        double l = lo, h = hi;
        for (int d = depth - 1; d >= 0; d--) {
            const double m = (l + h) / 2.0;
            if ((n >> d) & 1) l = m;
            else h = m;
        }
        const double mine = (l + h) / 2.0;
        const double fmine = godunov_lev(K, enI, pI, rI, enII, pII, rII, mine);
        int node = 1;
        for (int s = 0; s < 6 && it < 1100; s++, it++) {
            const double mid = lane_value(mine, node - 1);
            const double fm = lane_value(fmine, node - 1);
            if (fabs(fm) <= tol) { p = mid; return 0; }
            if ((flo * fm) <= 0.0) {
                hi = mid;
                node = 2 * node;
            } else {
                lo = mid;
                flo = fm;
                node = 2 * node + 1;
            }
        }
    }
    p = 0.0;
    return 1;
}

// newton: two shocks, from the acoustic estimate
GEL_HD int godunov_newton(const GodunovK& K, double enI, double pI, double rI, double enII, double pII, double rII,
                          double& p)
{
    const double cI = sqrt(K.ga * pI / rI);
    const double cII = sqrt(K.ga * pII / rII);
    double pn = pI * rII * cII + pII * rI * cI + (enI - enII) * rI * cI * rII * cII;
    pn = pn / (rI * cI + rII * cII);
    const double p0 = pn;
    for (int it = 1; it < 1100; it++) {
        p = pn;
        if (p <= 0.0) return 1;
        double fI = (p - pI) / (rI * cI * sqrt(K.gp * p / pI + K.gm));
        double dfI = (K.ga + 1.0) * p / pI + (3.0 * K.ga - 1.0);
        dfI = dfI / (4.0 * K.ga * rI * cI * pow((K.gp * p / pI + K.gm), (3.0 / 2.0)));
        double fII = (p - pII) / (rII * cII * sqrt(K.gp * p / pII + K.gm));
        double dfII = (K.ga + 1.0) * p / pII + (3.0 * K.ga - 1.0);
        dfII = dfII / (4.0 * K.ga * rII * cII * pow((K.gp * p / pII + K.gm), (3.0 / 2.0)));
        pn = p - (fI + fII - (enI - enII)) / (dfI + dfII);
        if (!(fabs(pn / p0 - p / p0) >= 10e-10)) {
            p = pn;
            return 0;
        }
    }
    return 1;
}

// q1, q2: (ro, p, u, v, w); q out: mass, momentum x/y/z, energy fluxes through the face
// uu_out (optional): the wave speed the returned time step is formed from, floored at
// 1e-7 as there, so that cfl * dist / *uu_out equals the returned time.
// WAVE: one problem per wavefront, every lane with the same arguments (the bisection
// by godunov_bisect_wave); otherwise one problem per lane.
template <bool WAVE = false>
GEL_HD double godunov_exact(const double* q1, const double* q2, double n1, double n2, double n3, double w,
                            double dist, double cfl, double gam, double* q, double& dsl, double& dsp, double& dsc,
                            int& status, double* uu_out = nullptr)
{
    const GodunovK K = godunov_k(gam);
    const double ga = K.ga, g1 = K.g1, g2 = K.g2, gm = K.gm;
    const double al = n1, be = n2, ge = n3;
    const double al2 = be, be2 = -al, ge2 = 0.0;
    const double al3 = 0.0, be3 = 0.0, ge3 = 1.0;
    double enI = al * q1[2] + be * q1[3] + ge * q1[4];
    double teI2 = al2 * q1[2] + be2 * q1[3] + ge2 * q1[4];
    double teI3 = al3 * q1[2] + be3 * q1[3] + ge3 * q1[4];
    double enII = al * q2[2] + be * q2[3] + ge * q2[4];
    double teII2 = al2 * q2[2] + be2 * q2[3] + ge2 * q2[4];
    double teII3 = al3 * q2[2] + be3 * q2[3] + ge3 * q2[4];
    double pI = q1[1], pII = q2[1], rI = q1[0], rII = q2[0];
    status = 0;
    double time = 0.0;
    const bool swapped = pI > pII;      // side I becomes the low-pressure side, mirrored
    if (swapped) {
        const double e1 = enI, t12 = teI2, t13 = teI3, p1 = pI, r1 = rI;
        enI = -enII; teI2 = teII2; teI3 = teII3; pI = pII; rI = rII;
        enII = -e1; teII2 = t12; teII3 = t13; pII = p1; rII = r1;
        w = -w;
    }
    const double cI = (pI != 0.0) ? sqrt(ga * pI / rI) : 0.0;
    const double cII = (pII != 0.0) ? sqrt(ga * pII / rII) : 0.0;
    const double a = sqrt(rI * (g2 * pII + g1 * pI) / 2.0);
    const double Uud = (pII - pI) / a;
    const double Urz = -2.0 * cII / g1 * (1.0 - pow((pI / pII), gm));
    const double Uvk = -2.0 * (cII + cI) / g1;
    const double Udf = enI - enII;

    double p = 0.0, r = 0.0, te2 = 0.0, te3 = 0.0, en = 0.0;
    // star-region density behind a shock (side I, left; side II, right)
    auto shock_rho_I = [&](double aI, double u) { return rI * aI / (aI - rI * (enI - u)); };
    auto shock_rho_II = [&](double aII, double u) { return rII * aII / (aII + rII * (enII - u)); };
    // mass flux across a rarefaction with star pressure p (|p - p_side| < eps: acoustic limit)
    auto fan_a = [&](double rs, double cs, double ps, double P) {
        return (fabs(P - ps) < 10e-10) ? rs * cs : gm * rs * cs * (1.0 - P / ps) / (1.0 - pow((P / ps), gm));
    };
    double uu_eff = 10e-8;
    auto time_of = [&](double UU) {
        uu_eff = (UU > 10e-8) ? UU : 10e-8;
        return (UU > 10e-8) ? cfl * dist / UU : cfl * dist / 10e-8;
    };
    auto state_I = [&]() { en = enI; p = pI; r = rI; te2 = teI2; te3 = teI3; };
    auto state_II = [&]() { en = enII; p = pII; r = rII; te2 = teII2; te3 = teII3; };

    if (Udf < Uvk) {                                        // vacuum between two fans
        const double gg1 = ga - 1.0, gg2 = ga + 1.0;
        const double dI = enI - cI, ddI = enI + 2.0 / gg1 * cI;
        const double dII = enII + cII, ddII = enII - 2.0 / gg1 * cII;
        dsl = dI; dsp = dII; dsc = (dI + dII) / 2.0;
        double UU = mx(fabs(dsl), fabs(dsp));
        UU = mx(UU, fabs(ddII));
        UU = mx(UU, fabs(ddI));
        time = time_of(UU);
        if (w <= dI) state_I();
        if ((w > dI) && (w < ddI)) {
            const double ce = gg1 / gg2 * (enI - w) + 2.0 / gg2 * cI;
            en = w + ce; p = pI * pow((ce / cI), (1.0 / gm)); r = ga * p / ce / ce; te2 = teI2; te3 = teI3;
        }
        if ((w >= ddI) && (w <= ddII)) { en = w; p = 0.0; r = 0.0; te2 = 0.0; te3 = 0.0; }
        if ((w > ddII) && (w < dII)) {
            const double ce = -gg1 / gg2 * (enII - w) + 2.0 / gg2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / gm)); r = ga * p / ce / ce; te2 = teII2; te3 = teII3;
        }
        if (w >= dII) state_II();
    } else if ((Udf >= Uvk) && (Udf <= Urz)) {              // two rarefactions
        p = pI * pow(((Udf - Uvk) / (Urz - Uvk)), (1.0 / gm));
        const double aI = fan_a(rI, cI, pI, p), aII = fan_a(rII, cII, pII, p);
        const double u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        const double dI = enI - cI, ddI = u - cI - g1 * (enI - u) / 2.0;
        const double dII = enII + cII, ddII = u + cII - g1 * (enII - u) / 2.0;
        dsl = dI; dsp = dII; dsc = u;
        double UU = mx(fabs(dsl), fabs(dsp));
        UU = mx(UU, fabs(ddII));
        UU = mx(UU, fabs(ddI));
        time = time_of(UU);
        if (w <= dI) state_I();
        if ((w > dI) && (w < ddI)) {
            const double ce = g1 / g2 * (enI - w) + 2.0 / g2 * cI;
            en = w + ce; p = pI * pow((ce / cI), (1.0 / gm)); r = ga * p / ce / ce; te2 = teI2; te3 = teI3;
        }
        if ((w >= ddI) && (w <= u)) {
            const double ce = cI + g1 / 2.0 * (enI - u);
            en = u; r = ga * p / ce / ce; te2 = teI2; te3 = teI3;
        }
        if ((w > u) && (w <= ddII)) {
            const double ce = cII - g1 / 2.0 * (enII - u);
            en = u; r = ga * p / ce / ce; te2 = teII2; te3 = teII3;
        }
        if ((w > ddII) && (w < dII)) {
            const double ce = -g1 / g2 * (enII - w) + 2.0 / g2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / gm)); r = ga * p / ce / ce; te2 = teII2; te3 = teII3;
        }
        if (w >= dII) state_II();
    } else if ((Udf > Urz) && (Udf <= Uud)) {               // shock into side I, fan into side II
        if constexpr (WAVE) status = godunov_bisect_wave(K, enI, pI, rI, enII, pII, rII, p);
        else status = godunov_bisect(K, enI, pI, rI, enII, pII, rII, p);
        if (status) return 0.0;
        const double aI = sqrt(rI * (g2 / 2.0 * p + g1 / 2.0 * pI));
        const double aII = (fabs(p - pII) < 10e-10) ? rII * cII
                                                     : gm * rII * cII * (1.0 - p / pII) / (1.0 - pow((p / pII), gm));
        const double u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        const double dI = enI - aI / rI, dII = enII + cII;
        const double ddII = u + cII - g1 * (enII - u) / 2.0;
        dsl = dI; dsp = dII; dsc = u;
        double UU = mx(fabs(dsl), fabs(dsp));
        UU = mx(UU, fabs(ddII));
        time = time_of(UU);
        if (w <= dI) state_I();
        if ((w > dI) && (w <= u)) { en = u; r = shock_rho_I(aI, u); te2 = teI2; te3 = teI3; }
        if ((w > u) && (w <= ddII)) {
            const double ce = cII - g1 / 2.0 * (enII - u);
            en = u; r = ga * p / ce / ce; te2 = teII2; te3 = teII3;
        }
        if ((w > ddII) && (w < dII)) {
            const double ce = -g1 / g2 * (enII - w) + 2.0 / g2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / gm)); r = ga * p / ce / ce; te2 = teII2; te3 = teII3;
        }
        if (w >= dII) state_II();
    } else if (Udf > Uud) {                                 // two shocks
        status = godunov_newton(K, enI, pI, rI, enII, pII, rII, p);
        if (status) return 0.0;
        const double aI = sqrt(rI * (g2 / 2.0 * p + g1 / 2.0 * pI));
        const double aII = sqrt(rII * (g2 / 2.0 * p + g1 / 2.0 * pII));
        const double u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        const double dI = enI - aI / rI, dII = enII + aII / rII;
        dsl = dI; dsp = dII; dsc = u;
        time = time_of(mx(fabs(dsl), fabs(dsp)));
        if (w <= dI) state_I();
        else if ((w > dI) && (w <= u)) { en = u; r = shock_rho_I(aI, u); te2 = teI2; te3 = teI3; }
        else if ((w > u) && (w < dII)) { en = u; r = shock_rho_II(aII, u); te2 = teII2; te3 = teII3; }
        else if (w >= dII) state_II();
    }
    if (swapped) {
        en = -en;
        const double l = dsl, rr = dsp;
        dsl = -rr; dsp = -l; dsc = -dsc;
        w = -w;
    }
    const double uo = al * en + al2 * te2 + al3 * te3;
    const double vo = be * en + be2 * te2 + be3 * te3;
    const double wo = ge * en + ge2 * te2 + ge3 * te3;
    const double eo = p / g1 + 0.5 * r * (uo * uo + vo * vo + wo * wo);
    en = al * uo + be * vo + ge * wo;
    q[0] = (r * (en - w));
    q[1] = (r * (en - w) * uo + al * p);
    q[2] = (r * (en - w) * vo + be * p);
    q[3] = (r * (en - w) * wo + ge * p);
    q[4] = ((en - w) * eo + en * p);
    if (uu_out) *uu_out = uu_eff;
    return time;
}

}  // namespace gel
