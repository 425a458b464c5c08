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
// Help.cpp:96-113: minmod slope through (x1,t1),(x2,t2),(x3,t3), evaluated at y.
template <class T> GEL_HD T linear3(T x1, T t1, T x2, T t2, T x3, T t3, T y)
{
    const T d = minmod((t1 - t2) / (x1 - x2), (t2 - t3) / (x2 - x3));
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
    cx = Gx / (6.0 * A);
    cy = Gy / (6.0 * A);
}
// Cell::Get_Volume for quads (Cell.cpp:146-166).
GEL_HD double quad_area(const double* px, const double* py)
{
    double a1 = px[0] - px[2], a2 = py[0] - py[2];
    const double a = sqrt(sq(a1) + sq(a2));
    double b1 = px[1] - px[3], b2 = py[1] - py[3];
    const double b = sqrt(sq(b1) + sq(b2));
    a1 = a1 / a; a2 = a2 / a;
    b1 = b1 / b; b2 = b2 / b;
    const double c = a1 * b1 + a2 * b2;
    const double s = sqrt(1 - sq(c));
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
    return fabs(A) * (1.0 / (6.0 * A)) * G * kPi * 360.0 / 180.0 / (2.0 * kPi);
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
    const T cL = sqrt(gam * L.p / L.ro);
    const T cR = sqrt(gam * R.p / R.ro);
    const T kL = (sq(L.u) + sq(L.v)) / T(2);
    const T kR = (sq(R.u) + sq(R.v)) / T(2);
    const T SL = mn(u1, u2) - mx(cL, cR);
    const T SR = mx(u1, u2) + mx(cL, cR);
    const T UU = mx(fabs(SL), fabs(SR));
    const T eL = L.p / g1 + L.ro * kL;
    const T eR = R.p / g1 + R.ro * kR;
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
    const T SM = (suR * R.ro * u2 - suL * L.ro * u1 - R.p + L.p) / (suR * R.ro - suL * L.ro);
    T UZ[5];
#pragma unroll
    for (int k = 0; k < 5; k++) UZ[k] = (SR * UR[k] - SL * UL[k] + FL[k] - FR[k]) / (SR - SL);
    T PO[5];
    if (!HLLC) {                        // HLL, Setka.cpp:10606-10618
#pragma unroll
        for (int k = 0; k < 5; k++)
            PO[k] = (SR * FL[k] - SL * FR[k] + SR * SL * (UR[k] - UL[k])) / (SR - SL);
        P[1] = n1 * (PO[1] - W * UZ[1]) + t1 * (PO[2] - W * UZ[2]);
        P[2] = n2 * (PO[1] - W * UZ[1]) + t2 * (PO[2] - W * UZ[2]);
        P[0] = PO[0] - W * UZ[0];
        P[3] = PO[3] - W * UZ[3];
        PQ = PO[4] - W * UZ[4];
        return UU;
    }
    // HLLC, Setka.cpp:10621-10697 (nul_potok = false on the sweep's faces)
    const T roLL = L.ro * (SL - u1) / (SL - SM);
    const T roRR = R.ro * (SR - u2) / (SR - SM);
    const T QLL = L.Q * (SL - u1) / (SL - SM);
    const T QRR = R.Q * (SR - u2) / (SR - SM);
    const T suRm = suR / (SR - SM), suLm = suL / (SL - SM);
    const T ptzR = R.p + R.ro * suR * (SM - u2);
    const T ptzL = L.p + L.ro * suL * (SM - u1);
    const T ptz = (ptzR + ptzL) / T(2);
    const T ee2 = eR * suRm + (ptz * SM - R.p * u2) / (SR - SM);
    const T ee1 = eL * suLm + (ptz * SM - L.p * u1) / (SL - SM);
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
        // nul_potok: tangential star velocity UZ2/UZ0 on both sides (Setka.cpp:10644-10653)
        const T g1 = gam - T(1);
        const T kL = (sq(L.u) + sq(L.v)) / T(2), kR = (sq(R.u) + sq(R.v)) / T(2);
        const T eL = L.p / g1 + L.ro * kL, eR = R.p / g1 + R.ro * kR;
        const T FL[5] = { L.ro * u1, L.ro * u1 * u1 + L.p, L.ro * u1 * v1, (eL + L.p) * u1, L.Q * u1 };
        const T FR[5] = { R.ro * u2, R.ro * u2 * u2 + R.p, R.ro * u2 * v2, (eR + R.p) * u2, R.Q * u2 };
        const T UL[5] = { L.ro, L.ro * u1, L.ro * v1, eL, L.Q };
        const T UR[5] = { R.ro, R.ro * u2, R.ro * v2, eR, R.Q };
        const T SM = Vc;
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
        UU = mx(fabs(SL), fabs(SR));
    }
    return cfl * rad / UU;
}

// Cell::Get_Sourse_MK1, own-cell moments (Cell.cpp:590-664).  f(k) returns field k.
template <class F>
GEL_HD void source_mk1(const gel_params& prm, F f, double u, double v, double ro, double p,
                       double& q1, double& q2, double& q3)
{
    double nu[4], uH[4], vH[4], roH[4], pH[4], UM[4], UU[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uH[k] = f(4 + k);
        vH[k] = f(8 + k);
        roH[k] = f(k);
        pH[k] = 0.5 * f(12 + k) * f(k);
        if (roH[k] <= 0.0) { roH[k] = 0.0000001; pH[k] = 0.0; }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        UM[k] = sqrt(sq(u - uH[k]) + sq(v - vH[k]) + (64.0 / (9.0 * kPi)) * (p / ro + 2.0 * pH[k] / roH[k]));
        UU[k] = sqrt(sq(u - uH[k]) + sq(v - vH[k]) + (4.0 / kPi) * (p / ro + 2.0 * pH[k] / roH[k]));
        const double sig = sq(prm.sig_a - prm.sig_b * log(UM[k]));
        nu[k] = ro * roH[k] * UM[k] * sig;
    }
    double Mu = nu[0] * (uH[0] - u) + nu[1] * (uH[1] - u) + nu[2] * (uH[2] - u) + nu[3] * (uH[3] - u);
    double Mv = nu[0] * (vH[0] - v) + nu[1] * (vH[1] - v) + nu[2] * (vH[2] - v) + nu[3] * (vH[3] - v);
    double t[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        t[k] = nu[k] * ((sq(uH[k]) + sq(vH[k]) - sq(u) - sq(v)) / 2.0 + (UU[k] / UM[k]) * (2.0 * pH[k] / roH[k] - p / ro));
    double MT = t[0] + t[1] + t[2] + t[3];
    Mu = prm.src_coef * Mu;
    Mv = prm.src_coef * Mv;
    MT = prm.src_coef * MT;
    q1 = Mu * f(16);
    q2 = Mv * f(17);
    q3 = MT * f(18);
}

}  // namespace gel
