// surrogate.hip - learned stand-in for the charge-exchange source terms.
//
// A multilayer perceptron 14 -> 128 -> 128 -> 128 -> 128 -> 5 (SiLU after each hidden
// layer) evaluated and trained on CDNA4 matrix cores: every product is a BF16 MFMA
// (v_mfma_f32_16x16x16_bf16) with FP32 accumulation; biases, pre-activations, the loss
// and the Adam moments are FP32; the FP32 master weights are re-quantised to BF16 after
// each optimiser step.
//
// The reference computes these sources by Monte Carlo (Setka::MK_start_new,
// Setka.cpp:11989-12535) and refreshes them between sweeps; its analytic closure
// Cell::Get_Sourse_MK1 (Cell.cpp:498-665) is what the sweep then applies.  Here that
// analytic path generates the training targets, judges the network per cell, and is
// the fallback where the network misses its tolerance.
//
// All matrix products are one kernel shape, C[m][n] = sum_k A[m][k] * Bt[n][k] with
// both operands contiguous in k (row-major A, row-major B transposed).  Forward:
// Z = H W^T (A = activations [cells][in], Bt = W [out][in]).  Backprop of activations:
// dH = dZ W (Bt = W^T kept as a second BF16 copy).  Weight gradients: dW = dZ^T H
// (A = dZ^T [out][cells], Bt = H^T [in][cells]); the epilogues store the transposed
// copies these need.  The sum over the batch happens inside one GEMM (K = cells), so
// gradients are deterministic without atomics.
#include "sweep_core.h"
#include "gelios_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace gel {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
using bf16 = unsigned short;

__device__ inline bf16 f2bf(float f)
{   // round to nearest even (v_cvt_pk_bf16_f32)
    return __builtin_bit_cast(bf16, (__bf16)f);
}
__host__ __device__ inline float bf2f(bf16 b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ inline float sigmoidf(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

// network shape (inputs and outputs padded to the 16-wide MFMA tile)
constexpr int kIn = 14, kInP = 16, kHid = 128, kOut = 5, kOutP = 16, kLayers = 5;
constexpr int kIn_[kLayers] = { kInP, kHid, kHid, kHid, kHid };
constexpr int kOut_[kLayers] = { kHid, kHid, kHid, kHid, kOutP };
constexpr int kInReal[kLayers] = { kIn, kHid, kHid, kHid, kHid };
constexpr int kOutReal[kLayers] = { kHid, kHid, kHid, kHid, kOut };

enum Epi {
    EPI_F32 = 0,        // C fp32
    EPI_BIAS_F32 = 1,   // C + bias, fp32 (output layer)
    EPI_BIAS_SILU = 2,  // Z = C + bias (fp32), H = silu(Z) as bf16 [m][n] and [n][m]
    EPI_DSILU = 3,      // dZ = C * silu'(Z) as bf16 [m][n] and [n][m]
};

struct GemmOut {
    float* cf;          // [M][ldc] fp32 (Z or C)
    int ldc;
    const float* bias;  // [N]
    bf16* cb;           // [M][ldc] bf16
    bf16* cbt;          // [N][ldt] bf16 transposed
    int ldt;
    const float* z;     // [M][ldc] saved pre-activation (EPI_DSILU)
};

// 256 threads = 4 wavefronts, each a 32x32 output tile of 2x2 MFMA 16x16x16 blocks.
// M, N, K are multiples of 16; tiles past M or N are skipped whole (wave-uniform).
template <int EPI>
__global__ void __launch_bounds__(256)
gemm_nt(int M, int N, int K, const bf16* __restrict__ A, int lda, const bf16* __restrict__ Bt, int ldb, GemmOut o)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = blockIdx.x * 64 + (wave >> 1) * 32;
    const int n0 = blockIdx.y * 64 + (wave & 1) * 32;
    const int r = lane & 15, q = lane >> 4;
    f4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f4{ 0.f, 0.f, 0.f, 0.f };
    const bool mv[2] = { m0 < M, m0 + 16 < M };
    const bool nv[2] = { n0 < N, n0 + 16 < N };
    for (int k = 0; k < K; k += 16) {
        s4 af[2], bfr[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            if (mv[t]) af[t] = *(const s4*)(A + (size_t)(m0 + t * 16 + r) * lda + k + 4 * q);
            if (nv[t]) bfr[t] = *(const s4*)(Bt + (size_t)(n0 + t * 16 + r) * ldb + k + 4 * q);
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
                if (mv[a] && nv[b]) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bfr[b], acc[a][b], 0, 0, 0);
    }
    // D layout: lane holds rows 4q..4q+3 of column r of each 16x16 block
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            if (!(mv[a] && nv[b])) continue;
            const int n = n0 + b * 16 + r;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int m = m0 + a * 16 + 4 * q + j;
                float v = acc[a][b][j];
                const size_t ij = (size_t)m * o.ldc + n;
                if (EPI == EPI_F32) {
                    o.cf[ij] = v;
                } else if (EPI == EPI_BIAS_F32) {
                    o.cf[ij] = v + o.bias[n];
                } else if (EPI == EPI_BIAS_SILU) {
                    const float z = v + o.bias[n];
                    o.cf[ij] = z;
                    const bf16 h = f2bf(z * sigmoidf(z));
                    o.cb[ij] = h;
                    o.cbt[(size_t)n * o.ldt + m] = h;
                } else {
                    const float z = o.z[ij];
                    const float s = sigmoidf(z);
                    const bf16 d = f2bf(v * (s * (1.0f + z * (1.0f - s))));
                    o.cb[ij] = d;
                    o.cbt[(size_t)n * o.ldt + m] = d;
                }
            }
        }
}

// dY = 2 (Y - T) / (B * kOut) on the real outputs, 0 on the padding; loss partials.
__global__ void loss_grad_kernel(int B, const float* __restrict__ Y, const float* __restrict__ T,
                                 bf16* __restrict__ dy, bf16* __restrict__ dyt, float* __restrict__ partial)
{
    __shared__ float red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < B * kOutP) {
        const int b = i / kOutP, c = i % kOutP;
        float g = 0.f;
        if (c < kOut) {
            const float e = Y[i] - T[(size_t)b * kOut + c];
            l = e * e;
            g = 2.0f * e / (float)(B * kOut);
        }
        const bf16 gb = f2bf(g);
        dy[i] = gb;
        dyt[(size_t)c * B + b] = gb;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// db[n] = sum over the batch of dZ[b][n] (dZ as bf16 [B][ld])
__global__ void bias_grad_kernel(int B, int N, const bf16* __restrict__ dz, int ld, float* __restrict__ db)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int b = 0; b < B; b++) s += bf2f(dz[(size_t)b * ld + n]);
    db[n] = s;
}

// Adam on the FP32 master copy, then BF16 copies W [out][in] and W^T [in][out].
__global__ void adam_kernel(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void quantize_kernel(int rows, int cols, const float* __restrict__ w, bf16* __restrict__ wb,
                                bf16* __restrict__ wt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const int r = i / cols, c = i % cols;
    const bf16 b = f2bf(w[i]);
    wb[i] = b;
    wt[(size_t)c * rows + r] = b;
}

// X fp32 [n][14] -> bf16 [B][16] and [16][B] (zero padding)
__global__ void pack_input_kernel(int n, int B, const float* __restrict__ x, bf16* __restrict__ xb,
                                  bf16* __restrict__ xt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * kInP) return;
    const int b = i / kInP, c = i % kInP;
    const float v = (b < n && c < kIn) ? x[(size_t)b * kIn + c] : 0.f;
    const bf16 h = f2bf(v);
    xb[i] = h;
    xt[(size_t)c * B + b] = h;
}


// ---------------------------------------------------------- source features ---
// Features of the source model per cell (raw, before normalisation):
//   log ro, u, v, log(p/ro)  -- plasma (first four primitives of either state vector)
//   log nA, uA, vA, log TA   -- populations 1+2 (density-weighted velocity, temperature)
//   log nB, uB, vB, log TB   -- populations 3+4
//   y, x of the cell centre  -- position (cylindrical radius, axis coordinate)
// The hydrogen moments are the GEL_MK1_NFIELDS fields of gel_set_sources.
template <class TS>
__global__ void features_kernel(int ni, int nj, const double* __restrict__ X, const double* __restrict__ Y,
                                const TS* __restrict__ st, const TS* __restrict__ src,
                                const float* __restrict__ mu, const float* __restrict__ sd, float* __restrict__ out)
{
    const size_t nc = (size_t)ni * nj;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int i = (int)(c / nj), j = (int)(c % nj);
    const size_t a = (size_t)i * (nj + 1) + j, b = a + (nj + 1);
    const double px[4] = { X[a], X[b], X[b + 1], X[a + 1] };
    const double py[4] = { Y[a], Y[b], Y[b + 1], Y[a + 1] };
    double cx, cy;
    quad_center(px, py, cx, cy);
    const double ro = st[c], p = st[nc + c], u = st[2 * nc + c], v = st[3 * nc + c];
    double f[kIn];
    f[0] = log(ro); f[1] = u; f[2] = v; f[3] = log(p / ro);
    for (int g = 0; g < 2; g++) {
        const int k0 = 2 * g, k1 = 2 * g + 1;
        const double n0 = src[k0 * nc + c], n1 = src[k1 * nc + c];
        const double n = n0 + n1;
        const double w0 = n > 0 ? n0 / n : 0.5, w1 = n > 0 ? n1 / n : 0.5;
        f[4 + 4 * g] = log(fmax(n, 1e-30));
        f[5 + 4 * g] = w0 * src[(4 + k0) * nc + c] + w1 * src[(4 + k1) * nc + c];
        f[6 + 4 * g] = w0 * src[(8 + k0) * nc + c] + w1 * src[(8 + k1) * nc + c];
        f[7 + 4 * g] = log(fmax(w0 * src[(12 + k0) * nc + c] + w1 * src[(12 + k1) * nc + c], 1e-30));
    }
    f[12] = cy; f[13] = cx;
#pragma unroll
    for (int k = 0; k < kIn; k++) out[c * kIn + k] = mu ? (float)((f[k] - mu[k]) / sd[k]) : (float)f[k];
}

// Analytic charge-exchange source per cell (the judge and fallback): the device
// closure of Cell::Get_Sourse_MK1 (Cell.cpp:590-664) on the cell's own moments, in
// the state's type, as the sweep evaluates it.
template <class TS, class P>
__device__ inline void analytic_q(const gel_params& prm, size_t nc, size_t c, P st, P src, double q[5])
{
    TS r1, r2, r3;
    auto fk = [&](int k) { return src[(size_t)k * nc + c]; };
    source_mk1<decltype(fk), TS>(prm, fk, st[2 * nc + c], st[3 * nc + c], st[c], st[nc + c], r1, r2, r3);
    q[0] = 0.0; q[1] = r1; q[2] = r2; q[3] = r3; q[4] = 0.0;
}

// The closure without the Monte Carlo's k factors (k_u = k_v = k_T = 1).
template <class TS, class P>
__device__ inline void analytic_q_k1(const gel_params& prm, size_t nc, size_t c, P st, P src, double q[5])
{
    TS r1, r2, r3;
    auto fk = [&](int k) { return k >= 16 ? TS(1) : src[(size_t)k * nc + c]; };
    source_mk1<decltype(fk), TS>(prm, fk, st[2 * nc + c], st[3 * nc + c], st[c], st[nc + c], r1, r2, r3);
    q[0] = 0.0; q[1] = r1; q[2] = r2; q[3] = r3; q[4] = 0.0;
}

template <class TS>
__global__ void analytic_table_kernel(size_t nc, gel_params prm, const TS* __restrict__ st,
                                      const TS* __restrict__ src, double* __restrict__ qtab)
{
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double q[5];
    analytic_q<TS>(prm, nc, c, st, src, q);
#pragma unroll
    for (int k = 0; k < 5; k++) qtab[k * nc + c] = q[k];
}

// ------------------------------------------------------------ fused refresh ---
// One work-group of 4 wavefronts per tile of 256 cells of one case: features (a
// cell per thread) -> 5 MFMA layers over four passes of 64 cells with the
// activations in LDS -> de-normalisation, analytic judge and fallback (a cell per
// thread) -> source table.  Bias, SiLU and BF16 rounding are the forward's
// (gemm_nt); the 128-wide products run on the K = 32 MFMA, whose FP32 sums group
// the terms differently from the forward's K = 16 steps.
constexpr int kTile = 256;         // cells per work-group (one per thread)
constexpr int kSub = 64;           // rows per MFMA pass
constexpr int kLd = kHid + 8;      // LDS row stride (BF16) of the activations

struct NetDev {
    const bf16* w[kLayers];        // BF16 W_l [out_l][in_l] (padded)
    const float* b[kLayers];
    const float* mu;
    const float* sd;
    const float* ymu;
    const float* ysd;
    int mode;                      // GEL_MLP_OUT_SOURCE or GEL_MLP_OUT_FACTOR
};

// normalised features of cell c (features_kernel), BF16-rounded input row
template <class TS>
__device__ inline void refresh_features(const SweepArgs& a, size_t c, const NetDev& net, bf16* row)
{
    const size_t nc = (size_t)a.ni * a.nj;
    const int nj = a.nj;
    const int i = (int)(c / nj), j = (int)(c % nj);
    const size_t ka = (size_t)i * (nj + 1) + j, kb = ka + (nj + 1);
    const GPtr<const double> X = gp(a.x), Y = gp(a.y);
    const double px[4] = { X[ka], X[kb], X[kb + 1], X[ka + 1] };
    const double py[4] = { Y[ka], Y[kb], Y[kb + 1], Y[ka + 1] };
    double cx, cy;
    quad_center(px, py, cx, cy);
    const GPtr<const TS> st = gp((const TS*)a.next);
    const GPtr<const TS> src = gp((const TS*)a.src);
    const double ro = st[c], p = st[nc + c], u = st[2 * nc + c], v = st[3 * nc + c];
    double f[kIn];
    f[0] = log(ro); f[1] = u; f[2] = v; f[3] = log(p / ro);
    for (int g = 0; g < 2; g++) {
        const int k0 = 2 * g, k1 = 2 * g + 1;
        const double n0 = src[k0 * nc + c], n1 = src[k1 * nc + c];
        const double n = n0 + n1;
        const double w0 = n > 0 ? n0 / n : 0.5, w1 = n > 0 ? n1 / n : 0.5;
        f[4 + 4 * g] = log(fmax(n, 1e-30));
        f[5 + 4 * g] = w0 * src[(4 + k0) * nc + c] + w1 * src[(4 + k1) * nc + c];
        f[6 + 4 * g] = w0 * src[(8 + k0) * nc + c] + w1 * src[(8 + k1) * nc + c];
        f[7 + 4 * g] = log(fmax(w0 * src[(12 + k0) * nc + c] + w1 * src[(12 + k1) * nc + c], 1e-30));
    }
    f[12] = cy; f[13] = cx;
#pragma unroll
    for (int k = 0; k < kIn; k++) row[k] = f2bf((float)((f[k] - net.mu[k]) / net.sd[k]));
    row[14] = 0; row[15] = 0;
}

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));

// One hidden layer for the 64 cells of a pass, as D[unit][cell] = sum_k W[unit][k] *
// A[cell][k] (weights as the MFMA's first operand, cells as its columns), so that a
// lane's four results are four consecutive units of one cell: one 8-byte LDS store.
// Wavefront w computes units [32 w, 32 w + 32).  K = 16 (the padded input) takes
// v_mfma_f32_16x16x16_bf16, K = 128 v_mfma_f32_16x16x32_bf16 (lane l: row l % 16,
// k 8 (l / 16) .. + 7).
template <int K>
__device__ inline void refresh_hidden(const bf16* __restrict__ A, int lda, const bf16* __restrict__ W,
                                      const float* __restrict__ bias, bf16* __restrict__ H)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int u0 = wave * 32;
    f4 acc[2][4];
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[b][c] = f4{ 0.f, 0.f, 0.f, 0.f };
    if constexpr (K == 16) {
        s4 wa[2], hb[4];
#pragma unroll
        for (int b = 0; b < 2; b++) wa[b] = *(const s4*)(W + (size_t)(u0 + b * 16 + r) * K + 4 * q);
#pragma unroll
        for (int c = 0; c < 4; c++) hb[c] = *(const s4*)(A + (c * 16 + r) * lda + 4 * q);
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[b][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa[b], hb[c], acc[b][c], 0, 0, 0);
    } else {
#pragma unroll
        for (int k = 0; k < K; k += 32) {
            bf8 wa[2], hb[4];
#pragma unroll
            for (int b = 0; b < 2; b++) wa[b] = *(const bf8*)(W + (size_t)(u0 + b * 16 + r) * K + k + 8 * q);
#pragma unroll
            for (int c = 0; c < 4; c++) hb[c] = *(const bf8*)(A + (c * 16 + r) * lda + k + 8 * q);
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
                for (int c = 0; c < 4; c++) acc[b][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[b], hb[c], acc[b][c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int b = 0; b < 2; b++) {
        const int u = u0 + b * 16 + 4 * q;
        const float b0 = bias[u], b1 = bias[u + 1], b2 = bias[u + 2], b3 = bias[u + 3];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const float z0 = acc[b][c][0] + b0, z1 = acc[b][c][1] + b1, z2 = acc[b][c][2] + b2, z3 = acc[b][c][3] + b3;
            const bf2 lo = { (__bf16)(z0 * sigmoidf(z0)), (__bf16)(z1 * sigmoidf(z1)) };
            const bf2 hi = { (__bf16)(z2 * sigmoidf(z2)), (__bf16)(z3 * sigmoidf(z3)) };
            uint2 v;
            v.x = __builtin_bit_cast(unsigned int, lo);
            v.y = __builtin_bit_cast(unsigned int, hi);
            *(uint2*)(H + (c * 16 + r) * kLd + u) = v;
        }
    }
}

// The network and the judge for up to kTile cells packed into rows [0, nlive) of a
// work-group: thread t brings cell c of case cs (live: evaluate it) into row pos.
// Features (a cell per thread) -> 5 MFMA layers over ceil(nlive / 64) passes of 64
// rows with the activations in LDS -> de-normalisation, analytic judge and fallback
// (a cell per thread) -> source table and flag.  kept: a cell not evaluated that
// counts as fallen back.  Uniform across the work-group: nlive.
struct RefreshLds {
    // xs: the BF16 input rows; a row's 32 bytes take its 5 FP32 outputs once the
    // output layer has run (the row's input is consumed by then)
    __attribute__((aligned(16))) bf16 xs[kTile * kInP];
    __attribute__((aligned(16))) bf16 ha[kSub * kLd];
    __attribute__((aligned(16))) bf16 hb[kSub * kLd];
    int wcount[kTile / 64];
};

template <class TS>
__device__ __forceinline__ void refresh_rows(const SweepArgs& a, size_t c, bool live, bool kept, int pos, int nlive,
                                             const NetDev& net, double tol, int judge,
                                             unsigned long long* __restrict__ nfall, RefreshLds& L)
{
    bf16* xs = L.xs;
    float* ys = (float*)xs;                 // [kTile][8] FP32
    const int tid = threadIdx.x;
    if (live) refresh_features<TS>(a, c, net, xs + pos * kInP);
    if (tid >= nlive)
        for (int k = 0; k < kInP; k++) xs[tid * kInP + k] = 0;
    __syncthreads();
    for (int sub = 0; sub < kTile / kSub; sub++) {
        if (sub * kSub >= nlive) break;                 // uniform
        refresh_hidden<kInP>(xs + sub * kSub * kInP, kInP, net.w[0], net.b[0], L.ha);
        __syncthreads();
        refresh_hidden<kHid>(L.ha, kLd, net.w[1], net.b[1], L.hb);
        __syncthreads();
        refresh_hidden<kHid>(L.hb, kLd, net.w[2], net.b[2], L.ha);
        __syncthreads();
        refresh_hidden<kHid>(L.ha, kLd, net.w[3], net.b[3], L.hb);
        __syncthreads();
        {   // output layer, transposed: wavefront w takes cells [16 w, 16 w + 16); a lane
            // holds outputs 4 q .. 4 q + 3 of cell 16 w + r (outputs 0 .. 4 are real)
            const int lane = tid & 63, wave = tid >> 6;
            const int r = lane & 15, q = lane >> 4;
            f4 acc = f4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for (int k = 0; k < kHid; k += 32) {
                const bf8 wa = *(const bf8*)(net.w[4] + (size_t)r * kHid + k + 8 * q);
                const bf8 hv = *(const bf8*)(L.hb + (wave * 16 + r) * kLd + k + 8 * q);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, hv, acc, 0, 0, 0);
            }
            float* yrow = ys + (sub * kSub + wave * 16 + r) * 8;
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (4 * q + j < kOut) yrow[4 * q + j] = acc[j] + net.b[4][4 * q + j];
        }
        __syncthreads();
    }
    bool fall = kept;
    if (live) {
        const size_t nc = (size_t)a.ni * a.nj;
        const GPtr<const TS> st = gp((const TS*)a.next);
        const GPtr<const TS> src = gp((const TS*)a.src);
        const GPtr<TS> qtab = gp((TS*)const_cast<void*>(a.qtab));
        const double ro = st[c];
        const double nH = (double)src[c] + (double)src[nc + c] + (double)src[2 * nc + c] + (double)src[3 * nc + c];
        const double scale = a.prm.src_coef * ro * nH;
        double qm[5], q1[5];
        if (net.mode == GEL_MLP_OUT_FACTOR) {
            // the closure without k factors times the network's factors 1 + y
            analytic_q_k1<TS>(a.prm, nc, c, st, src, q1);
#pragma unroll
            for (int k = 0; k < 5; k++) qm[k] = q1[k] * (1.0 + ((double)ys[pos * 8 + k] * net.ysd[k] + net.ymu[k]));
        } else {
#pragma unroll
            for (int k = 0; k < 5; k++) qm[k] = ((double)ys[pos * 8 + k] * net.ysd[k] + net.ymu[k]) * scale;
        }
        if (judge) {
            double qa[5];
            analytic_q<TS>(a.prm, nc, c, st, src, qa);
            double err = 0.0;
#pragma unroll
            for (int k = 0; k < 5; k++) {
                // factors: relative to the closure with the Monte Carlo's k factors
                const double den = net.mode == GEL_MLP_OUT_FACTOR ? fabs(qa[k])
                                                                   : fmax(fabs(qa[k]), (double)net.ysd[k] * scale);
                if (den > 0.0) err = fmax(err, fabs(qm[k] - qa[k]) / den);
            }
            if (!(err <= tol)) {
                fall = true;
#pragma unroll
                for (int k = 0; k < 5; k++) qm[k] = qa[k];
            }
        }
#pragma unroll
        for (int k = 0; k < 5; k++) qtab[k * nc + c] = TS(qm[k]);
        if (a.qflag) gp(const_cast<int*>(a.qflag))[c] = fall ? 1 : 0;
    }
    if (nfall) {
        const unsigned long long mask = __ballot(fall);
        if ((tid & 63) == 0 && mask) atomicAdd(nfall, (unsigned long long)__popcll(mask));
    }
}

// One work-group per tile of 256 cells of one case.  judge 2 (hybrid cases): cells
// the last refresh sent to the analytic closure keep it and are not evaluated; the
// others are packed to the front of the tile, so that the MFMA passes run over
// ceil(live / 64) rows of 64 instead of four.
template <class TS>
__global__ void __launch_bounds__(256)
refresh_kernel(const SweepArgs* __restrict__ args, const int2* __restrict__ tiles, NetDev net, double tol,
               int judge, unsigned long long* __restrict__ nfall)
{
    __shared__ RefreshLds L;
    int cs = 0;
    long long t = blockIdx.x;
    if (tiles) {                        // (case, tile of that case)
        const int2 e = tiles[blockIdx.x];
        cs = e.x;
        t = e.y;
    }
    const SweepArgs& a = args[cs];
    const size_t nc = (size_t)a.ni * a.nj;
    const size_t c0 = (size_t)t * kTile;
    const int tid = threadIdx.x;
    const size_t c = c0 + tid;
    const bool sticky = judge == 2 && a.qflag;
    const bool kept = sticky && c < nc && gp(a.qflag)[c] != 0;
    const bool live = c < nc && !kept;
    int pos = tid, nlive = (int)min((size_t)kTile, nc > c0 ? nc - c0 : (size_t)0);
    if (sticky) {
        const unsigned long long m = __ballot(live);
        const int lane = tid & 63, wave = tid >> 6;
        if (lane == 0) L.wcount[wave] = __popcll(m);
        __syncthreads();
        int before = 0;
        nlive = 0;
        for (int k = 0; k < kTile / 64; k++) {
            if (k < wave) before += L.wcount[k];
            nlive += L.wcount[k];
        }
        pos = before + __popcll(m & ((1ull << lane) - 1ull));
    }
    refresh_rows<TS>(a, c, live, kept, pos, nlive, net, tol, judge, nfall, L);
}

// judge 2 over a batch: the (case, cell) list of the cells still on the network, in
// chunks of 256 (a persistent grid walks the chunks), so that the work scales with
// the cells evaluated and not with the grid.  The list is rebuilt after every refresh
// (live_list_kernel); the fallback count comes from there.
template <class TS>
__global__ void __launch_bounds__(256)
refresh_list_kernel(const SweepArgs* __restrict__ args, const int2* __restrict__ list,
                    const unsigned* __restrict__ count, NetDev net, double tol)
{
    __shared__ RefreshLds L;
    const int n = (int)count[0];
    const int tid = threadIdx.x;
    for (int g = blockIdx.x; g * kTile < n; g += gridDim.x) {
        const int nlive = min(kTile, n - g * kTile);
        const bool live = tid < nlive;
        int2 e = make_int2(0, 0);
        if (live) e = list[g * kTile + tid];
        refresh_rows<TS>(args[e.x], (size_t)e.y, live, false, tid, nlive, net, tol, 2, nullptr, L);
        __syncthreads();    // the rows are rewritten by the next chunk
    }
}

// The list of cells still on the network (flag 0) after a refresh, appended to
// list[count[1]...].  A work-group takes kListTiles (case, tile) pairs: it counts its
// cells, reserves their places with one atomic (a single counter takes one atomic
// per work-group, not per wavefront: the reservations serialise), then each
// wavefront writes its cells from its own offset.  The order varies from run to run;
// the cells' results do not depend on it.
constexpr int kListTiles = 16;
__global__ void __launch_bounds__(256)
live_list_kernel(const SweepArgs* __restrict__ args, const int2* __restrict__ tiles, int ntiles,
                 int2* __restrict__ list, unsigned* __restrict__ count)
{
    __shared__ unsigned wsum[kTile / 64];
    __shared__ unsigned gbase;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t0 = blockIdx.x * kListTiles, t1 = min(t0 + kListTiles, ntiles);
    auto is_live = [&](int t, int2& e) -> bool {
        const int2 tl = tiles[t];
        const SweepArgs& a = args[tl.x];
        const size_t nc = (size_t)a.ni * a.nj;
        const size_t c = (size_t)tl.y * kTile + threadIdx.x;
        e = make_int2(tl.x, (int)c);
        return c < nc && gp(a.qflag)[c] == 0;
    };
    unsigned n = 0;
    for (int t = t0; t < t1; t++) {
        int2 e;
        n += (unsigned)__popcll(__ballot(is_live(t, e)));
    }
    if (lane == 0) wsum[wave] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned tot = 0;
        for (int k = 0; k < kTile / 64; k++) tot += wsum[k];
        gbase = tot ? atomicAdd(&count[1], tot) : 0u;
    }
    __syncthreads();
    unsigned off = gbase;
    for (int k = 0; k < wave; k++) off += wsum[k];
    for (int t = t0; t < t1; t++) {
        int2 e;
        const bool live = is_live(t, e);
        const unsigned long long m = __ballot(live);
        if (live) list[off + __popcll(m & ((1ull << lane) - 1ull))] = e;
        off += (unsigned)__popcll(m);
    }
}

// count[0] = the list just built, count[1] cleared for the next build; the cells not
// on it (flagged) added to *nfall (one thread).
__global__ void live_list_commit(unsigned* __restrict__ count, unsigned long long ncells,
                                 unsigned long long* __restrict__ nfall)
{
    if (nfall) *nfall += ncells - count[1];
    count[0] = count[1];
    count[1] = 0;
}

}  // namespace gel

using namespace gel;

static thread_local std::string g_mlp_err;
static int mfail(int st, const std::string& m)
{
    g_mlp_err = m;
    return st;
}

struct gel_mlp {
    // FP32 master parameters, padded layouts: W_l [out_l][in_l], b_l [out_l]
    float* w[kLayers] = {};
    float* b[kLayers] = {};
    float* gw[kLayers] = {};
    float* gb[kLayers] = {};
    float* mw[kLayers] = {};
    float* vw[kLayers] = {};
    float* mb[kLayers] = {};
    float* vb[kLayers] = {};
    bf16* wb[kLayers] = {};   // BF16 W
    bf16* wt[kLayers] = {};   // BF16 W^T
    long long adam_t = 0;
    // activations for a batch of capacity cap
    int cap = 0;
    bf16 *x = nullptr, *xt = nullptr;
    float* z[4] = {};
    bf16* h[4] = {};
    bf16* ht[4] = {};
    float* y = nullptr;
    float* t = nullptr;
    bf16 *dy = nullptr, *dyt = nullptr;
    bf16 *dz = nullptr, *dzt = nullptr, *dz2 = nullptr, *dz2t = nullptr;
    float* partial = nullptr;
    float* xin = nullptr;
    // training set
    int nset = 0;
    float* set_x = nullptr;
    float* set_t = nullptr;
    // normalisation of features and targets (device), set by gel_mlp_set_scaling
    int mode = GEL_MLP_OUT_SOURCE;   // what the outputs are (gel_mlp_set_output_mode)
    float* mu = nullptr;
    float* sd = nullptr;
    float* ymu = nullptr;
    float* ysd = nullptr;
    hipStream_t stream = nullptr;
    // single-case refresh: its argument block and fallback counter
    SweepArgs* d_rargs = nullptr;
    unsigned long long* d_rcount = nullptr;
};

static size_t wsz(int l) { return (size_t)kOut_[l] * kIn_[l]; }

static void mlp_free_batch(gel_mlp* m)
{
    (void)hipFree(m->x); (void)hipFree(m->xt); (void)hipFree(m->y); (void)hipFree(m->t);
    (void)hipFree(m->dy); (void)hipFree(m->dyt); (void)hipFree(m->dz); (void)hipFree(m->dzt);
    (void)hipFree(m->dz2); (void)hipFree(m->dz2t); (void)hipFree(m->partial); (void)hipFree(m->xin);
    for (int l = 0; l < 4; l++) { (void)hipFree(m->z[l]); (void)hipFree(m->h[l]); (void)hipFree(m->ht[l]); }
    m->x = m->xt = nullptr; m->y = m->t = nullptr; m->dy = m->dyt = m->dz = m->dzt = m->dz2 = m->dz2t = nullptr;
    m->partial = m->xin = nullptr;
    for (int l = 0; l < 4; l++) { m->z[l] = nullptr; m->h[l] = nullptr; m->ht[l] = nullptr; }
    m->cap = 0;
}

static int mlp_reserve(gel_mlp* m, int B)
{
    if (B <= m->cap) return GEL_OK;
    mlp_free_batch(m);
    const size_t b = (size_t)B;
    bool ok = true;
    auto al = [&](void** p, size_t bytes) { if (ok && hipMalloc(p, bytes) != hipSuccess) ok = false; };
    al((void**)&m->x, b * kInP * 2); al((void**)&m->xt, b * kInP * 2);
    for (int l = 0; l < 4; l++) { al((void**)&m->z[l], b * kHid * 4); al((void**)&m->h[l], b * kHid * 2); al((void**)&m->ht[l], b * kHid * 2); }
    al((void**)&m->y, b * kOutP * 4); al((void**)&m->t, b * kOut * 4);
    al((void**)&m->dy, b * kOutP * 2); al((void**)&m->dyt, b * kOutP * 2);
    al((void**)&m->dz, b * kHid * 2); al((void**)&m->dzt, b * kHid * 2);
    al((void**)&m->dz2, b * kHid * 2); al((void**)&m->dz2t, b * kHid * 2);
    al((void**)&m->partial, (b * kOutP / 256 + 2) * 4); al((void**)&m->xin, b * kIn * 4);
    if (!ok) { mlp_free_batch(m); return mfail(GEL_E_NOMEM, "mlp batch buffers"); }
    m->cap = B;
    return GEL_OK;
}

template <int EPI>
static void gemm(hipStream_t s, int M, int N, int K, const bf16* A, int lda, const bf16* Bt, int ldb, GemmOut o)
{
    dim3 grid((M + 63) / 64, (N + 63) / 64);
    hipLaunchKernelGGL((gemm_nt<EPI>), grid, dim3(256), 0, s, M, N, K, A, lda, Bt, ldb, o);
}

// forward over a padded batch B (activations stay on the device)
static void mlp_forward_dev(gel_mlp* m, int B)
{
    const bf16* in = m->x;
    int ldin = kInP;
    for (int l = 0; l < 4; l++) {
        GemmOut o{ m->z[l], kHid, m->b[l], m->h[l], m->ht[l], B, nullptr };
        gemm<EPI_BIAS_SILU>(m->stream, B, kHid, kIn_[l], in, ldin, m->wb[l], kIn_[l], o);
        in = m->h[l];
        ldin = kHid;
    }
    GemmOut o{ m->y, kOutP, m->b[4], nullptr, nullptr, 0, nullptr };
    gemm<EPI_BIAS_F32>(m->stream, B, kOutP, kHid, in, kHid, m->wb[4], kHid, o);
}

// backward for the batch in m->x / m->t (forward already run); gradients in gw/gb
static void mlp_backward_dev(gel_mlp* m, int B)
{
    hipLaunchKernelGGL(loss_grad_kernel, dim3((B * kOutP + 255) / 256), dim3(256), 0, m->stream, B, m->y, m->t, m->dy,
                       m->dyt, m->partial);
    // layer 4 (linear): dW4 = dY^T H3, db4, dH3 = dY W4 -> dZ3
    {
        GemmOut o{ m->gw[4], kHid, nullptr, nullptr, nullptr, 0, nullptr };
        gemm<EPI_F32>(m->stream, kOutP, kHid, B, m->dyt, B, m->ht[3], B, o);
        hipLaunchKernelGGL(bias_grad_kernel, dim3(1), dim3(64), 0, m->stream, B, kOutP, m->dy, kOutP, m->gb[4]);
        GemmOut d{ nullptr, kHid, nullptr, m->dz, m->dzt, B, m->z[3] };
        d.cf = nullptr;
        d.z = m->z[3];
        gemm<EPI_DSILU>(m->stream, B, kHid, kOutP, m->dy, kOutP, m->wt[4], kOutP, d);
    }
    bf16 *dz = m->dz, *dzt = m->dzt, *dzn = m->dz2, *dznt = m->dz2t;
    for (int l = 3; l >= 0; l--) {
        // dW_l = dZ_l^T H_{l-1} (H_{-1} = X)
        const bf16* hprevT = (l > 0) ? m->ht[l - 1] : m->xt;
        GemmOut o{ m->gw[l], kIn_[l], nullptr, nullptr, nullptr, 0, nullptr };
        gemm<EPI_F32>(m->stream, kHid, kIn_[l], B, dzt, B, hprevT, B, o);
        hipLaunchKernelGGL(bias_grad_kernel, dim3((kHid + 63) / 64), dim3(64), 0, m->stream, B, kHid, dz, kHid, m->gb[l]);
        if (l > 0) {
            GemmOut d{ nullptr, kHid, nullptr, dzn, dznt, B, m->z[l - 1] };
            gemm<EPI_DSILU>(m->stream, B, kHid, kHid, dz, kHid, m->wt[l], kHid, d);
            std::swap(dz, dzn);
            std::swap(dzt, dznt);
        }
    }
}

static void mlp_quantize(gel_mlp* m)
{
    for (int l = 0; l < kLayers; l++) {
        const int n = (int)wsz(l);
        hipLaunchKernelGGL(quantize_kernel, dim3((n + 255) / 256), dim3(256), 0, m->stream, kOut_[l], kIn_[l], m->w[l],
                           m->wb[l], m->wt[l]);
    }
}

hipError_t gel_refresh_launch(hipStream_t s, const SweepArgs* d_args, const int2* d_tiles, long long tiles,
                              gel_mlp* m, bool f32, double tol, int judge, unsigned long long* d_nfall)
{
    if (tiles <= 0) return hipSuccess;
    NetDev net;
    for (int l = 0; l < kLayers; l++) { net.w[l] = m->wb[l]; net.b[l] = m->b[l]; }
    net.mu = m->mu; net.sd = m->sd; net.ymu = m->ymu; net.ysd = m->ysd; net.mode = m->mode;
    if (f32)
        hipLaunchKernelGGL(refresh_kernel<float>, dim3((unsigned)tiles), dim3(256), 0, s, d_args, d_tiles, net, tol,
                           judge, d_nfall);
    else
        hipLaunchKernelGGL(refresh_kernel<double>, dim3((unsigned)tiles), dim3(256), 0, s, d_args, d_tiles, net, tol,
                           judge, d_nfall);
    return hipGetLastError();
}

static void live_list_launch(hipStream_t s, const SweepArgs* d_args, const int2* d_tiles, long long tiles,
                             long long ncells, int2* d_list, unsigned* d_count, unsigned long long* d_nfall)
{
    const unsigned groups = (unsigned)((tiles + kListTiles - 1) / kListTiles);
    hipLaunchKernelGGL(live_list_kernel, dim3(groups), dim3(256), 0, s, d_args, d_tiles, (int)tiles, d_list, d_count);
    hipLaunchKernelGGL(live_list_commit, dim3(1), dim3(1), 0, s, d_count, (unsigned long long)ncells, d_nfall);
}

hipError_t gel_refresh_list_launch(hipStream_t s, const SweepArgs* d_args, const int2* d_tiles, long long tiles,
                                   long long ncells, int2* d_list, unsigned* d_count, gel_mlp* m, bool f32,
                                   double tol, unsigned long long* d_nfall)
{
    if (tiles <= 0) return hipSuccess;
    NetDev net;
    for (int l = 0; l < kLayers; l++) { net.w[l] = m->wb[l]; net.b[l] = m->b[l]; }
    net.mu = m->mu; net.sd = m->sd; net.ymu = m->ymu; net.ysd = m->ysd; net.mode = m->mode;
    // a persistent grid over the list's chunks: about three work-groups per CU
    const dim3 grid(768);
    if (f32)
        hipLaunchKernelGGL(refresh_list_kernel<float>, grid, dim3(256), 0, s, d_args, d_list, d_count, net, tol);
    else
        hipLaunchKernelGGL(refresh_list_kernel<double>, grid, dim3(256), 0, s, d_args, d_list, d_count, net, tol);
    live_list_launch(s, d_args, d_tiles, tiles, ncells, d_list, d_count, d_nfall);
    return hipGetLastError();
}

hipError_t gel_live_list_build(hipStream_t s, const SweepArgs* d_args, const int2* d_tiles, long long tiles,
                               long long ncells, int2* d_list, unsigned* d_count)
{
    if (tiles <= 0) return hipSuccess;
    live_list_launch(s, d_args, d_tiles, tiles, ncells, d_list, d_count, nullptr);
    return hipGetLastError();
}

static uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

extern "C" {

const char* gel_mlp_last_error(void) { return g_mlp_err.c_str(); }

int gel_mlp_param_count(void)
{
    int n = 0;
    for (int l = 0; l < kLayers; l++) n += kOutReal[l] * kInReal[l] + kOutReal[l];
    return n;
}

void gel_mlp_destroy(gel_mlp* m)
{
    if (!m) return;
    mlp_free_batch(m);
    for (int l = 0; l < kLayers; l++) {
        (void)hipFree(m->w[l]); (void)hipFree(m->b[l]); (void)hipFree(m->gw[l]); (void)hipFree(m->gb[l]);
        (void)hipFree(m->mw[l]); (void)hipFree(m->vw[l]); (void)hipFree(m->mb[l]); (void)hipFree(m->vb[l]);
        (void)hipFree(m->wb[l]); (void)hipFree(m->wt[l]);
    }
    (void)hipFree(m->set_x); (void)hipFree(m->set_t);
    (void)hipFree(m->mu); (void)hipFree(m->sd); (void)hipFree(m->ymu); (void)hipFree(m->ysd);
    (void)hipFree(m->d_rargs); (void)hipFree(m->d_rcount);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int gel_mlp_set_params(gel_mlp* m, const float* p)
{
    if (!m || !p) return mfail(GEL_E_ARG, "null argument");
    size_t o = 0;
    for (int l = 0; l < kLayers; l++) {
        std::vector<float> W(wsz(l), 0.f), b((size_t)kOut_[l], 0.f);
        for (int r = 0; r < kOutReal[l]; r++)
            for (int c = 0; c < kInReal[l]; c++) W[(size_t)r * kIn_[l] + c] = p[o++];
        for (int r = 0; r < kOutReal[l]; r++) b[r] = p[o++];
        if (hipMemcpy(m->w[l], W.data(), W.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->b[l], b.data(), b.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return mfail(GEL_E_HIP, "set params");
    }
    mlp_quantize(m);
    return hipStreamSynchronize(m->stream) == hipSuccess ? GEL_OK : mfail(GEL_E_HIP, "quantize");
}

int gel_mlp_get_params(gel_mlp* m, float* p)
{
    if (!m || !p) return mfail(GEL_E_ARG, "null argument");
    if (hipStreamSynchronize(m->stream) != hipSuccess) return mfail(GEL_E_HIP, "sync");
    size_t o = 0;
    for (int l = 0; l < kLayers; l++) {
        std::vector<float> W(wsz(l)), b((size_t)kOut_[l]);
        if (hipMemcpy(W.data(), m->w[l], W.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(b.data(), m->b[l], b.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return mfail(GEL_E_HIP, "get params");
        for (int r = 0; r < kOutReal[l]; r++)
            for (int c = 0; c < kInReal[l]; c++) p[o++] = W[(size_t)r * kIn_[l] + c];
        for (int r = 0; r < kOutReal[l]; r++) p[o++] = b[r];
    }
    return GEL_OK;
}

int gel_mlp_create(gel_mlp** out, uint64_t seed)
{
    if (!out) return mfail(GEL_E_ARG, "null argument");
    *out = nullptr;
    gel_mlp* m = new (std::nothrow) gel_mlp();
    if (!m) return mfail(GEL_E_NOMEM, "host");
    bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
    for (int l = 0; l < kLayers && ok; l++) {
        const size_t n = wsz(l), nb = (size_t)kOut_[l];
        float** fs[] = { &m->w[l], &m->gw[l], &m->mw[l], &m->vw[l] };
        for (auto f : fs) ok = ok && hipMalloc((void**)f, n * 4) == hipSuccess && hipMemset(*f, 0, n * 4) == hipSuccess;
        float** bs[] = { &m->b[l], &m->gb[l], &m->mb[l], &m->vb[l] };
        for (auto f : bs) ok = ok && hipMalloc((void**)f, nb * 4) == hipSuccess && hipMemset(*f, 0, nb * 4) == hipSuccess;
        ok = ok && hipMalloc((void**)&m->wb[l], n * 2) == hipSuccess && hipMalloc((void**)&m->wt[l], n * 2) == hipSuccess;
    }
    // the clears ran on the null stream, with which the network's stream does not
    // synchronise
    ok = ok && hipDeviceSynchronize() == hipSuccess;
    if (!ok) { gel_mlp_destroy(m); return mfail(GEL_E_NOMEM, "mlp parameters"); }
    // uniform(-a, a) with a = sqrt(6 / fan_in) for SiLU layers, sqrt(3 / fan_in) for the output
    std::vector<float> p((size_t)gel_mlp_param_count());
    uint64_t s = seed;
    size_t o = 0;
    for (int l = 0; l < kLayers; l++) {
        const float a = std::sqrt((l < 4 ? 6.0f : 3.0f) / (float)kInReal[l]);
        for (int i = 0; i < kOutReal[l] * kInReal[l]; i++)
            p[o++] = a * (2.0f * (float)((splitmix(s) >> 11) * (1.0 / 9007199254740992.0)) - 1.0f);
        for (int i = 0; i < kOutReal[l]; i++) p[o++] = 0.0f;
    }
    const int st = gel_mlp_set_params(m, p.data());
    if (st != GEL_OK) { gel_mlp_destroy(m); return st; }
    *out = m;
    return GEL_OK;
}

// Forward on host arrays x [n][14] (normalised features) -> y [n][5].
int gel_mlp_forward(gel_mlp* m, int n, const float* x, float* y)
{
    if (!m || n < 0 || (n > 0 && (!x || !y))) return mfail(GEL_E_ARG, "null argument");
    if (n == 0) return GEL_OK;
    const int B = (n + 15) / 16 * 16;
    int st = mlp_reserve(m, B);
    if (st != GEL_OK) return st;
    if (hipMemcpyAsync(m->xin, x, (size_t)n * kIn * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess)
        return mfail(GEL_E_HIP, "copy");
    hipLaunchKernelGGL(pack_input_kernel, dim3((B * kInP + 255) / 256), dim3(256), 0, m->stream, n, B, m->xin, m->x, m->xt);
    mlp_forward_dev(m, B);
    std::vector<float> Y((size_t)B * kOutP);
    if (hipMemcpyAsync(Y.data(), m->y, Y.size() * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return mfail(GEL_E_HIP, "forward");
    for (int i = 0; i < n; i++)
        for (int c = 0; c < kOut; c++) y[(size_t)i * kOut + c] = Y[(size_t)i * kOutP + c];
    return GEL_OK;
}

static int load_batch(gel_mlp* m, int n, const float* x_dev, const float* t_dev, hipMemcpyKind kind)
{
    const int B = (n + 15) / 16 * 16;
    int st = mlp_reserve(m, B);
    if (st != GEL_OK) return st;
    if (hipMemsetAsync(m->t, 0, (size_t)B * kOut * 4, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->xin, x_dev, (size_t)n * kIn * 4, kind, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->t, t_dev, (size_t)n * kOut * 4, kind, m->stream) != hipSuccess)
        return mfail(GEL_E_HIP, "batch copy");
    hipLaunchKernelGGL(pack_input_kernel, dim3((B * kInP + 255) / 256), dim3(256), 0, m->stream, n, B, m->xin, m->x, m->xt);
    return GEL_OK;
}

// Loss (mean squared error over the n x 5 outputs) and its gradient with respect to
// every parameter, in the packing order of gel_mlp_get_params.  Rows past n in the
// padded batch carry zero targets and zero inputs; the gradient divides by n*5 only
// through the rows that exist (padding rows are excluded by the loss kernel's B).
int gel_mlp_grad(gel_mlp* m, int n, const float* x, const float* t, float* grad, float* loss)
{
    if (!m || n <= 0 || n % 16 || !x || !t || !grad || !loss) return mfail(GEL_E_ARG, "n must be a positive multiple of 16");
    int st = load_batch(m, n, x, t, hipMemcpyHostToDevice);
    if (st != GEL_OK) return st;
    mlp_forward_dev(m, n);
    mlp_backward_dev(m, n);
    const int nb = (n * kOutP + 255) / 256;
    std::vector<float> part((size_t)nb);
    if (hipMemcpyAsync(part.data(), m->partial, (size_t)nb * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return mfail(GEL_E_HIP, "grad");
    double s = 0.0;
    for (float v : part) s += v;
    *loss = (float)(s / (double)(n * kOut));
    size_t o = 0;
    for (int l = 0; l < kLayers; l++) {
        std::vector<float> W(wsz(l)), b((size_t)kOut_[l]);
        if (hipMemcpy(W.data(), m->gw[l], W.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(b.data(), m->gb[l], b.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return mfail(GEL_E_HIP, "grad copy");
        for (int r = 0; r < kOutReal[l]; r++)
            for (int c = 0; c < kInReal[l]; c++) grad[o++] = W[(size_t)r * kIn_[l] + c];
        for (int r = 0; r < kOutReal[l]; r++) grad[o++] = b[r];
    }
    return GEL_OK;
}

// Training set on the device: x [n][14] normalised features, t [n][5] targets.
int gel_mlp_set_dataset(gel_mlp* m, int n, const float* x, const float* t)
{
    if (!m || n <= 0 || !x || !t) return mfail(GEL_E_ARG, "null argument");
    (void)hipFree(m->set_x);
    (void)hipFree(m->set_t);
    m->set_x = m->set_t = nullptr;
    if (hipMalloc((void**)&m->set_x, (size_t)n * kIn * 4) != hipSuccess ||
        hipMalloc((void**)&m->set_t, (size_t)n * kOut * 4) != hipSuccess)
        return mfail(GEL_E_NOMEM, "dataset");
    if (hipMemcpy(m->set_x, x, (size_t)n * kIn * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->set_t, t, (size_t)n * kOut * 4, hipMemcpyHostToDevice) != hipSuccess)
        return mfail(GEL_E_HIP, "dataset copy");
    m->nset = n;
    return GEL_OK;
}

// steps minibatch Adam steps over the stored set (batch rows taken in order,
// wrapping; the caller shuffles the set).  loss_out[steps] receives the batch losses.
int gel_mlp_train(gel_mlp* m, int steps, int batch, float lr, float* loss_out)
{
    if (!m || steps < 0 || batch <= 0 || batch % 16 || !m->set_x) return mfail(GEL_E_ARG, "train arguments");
    if (batch > m->nset) return mfail(GEL_E_ARG, "batch larger than the set");
    int st = mlp_reserve(m, batch);
    if (st != GEL_OK) return st;
    const int nb = (batch * kOutP + 255) / 256;
    std::vector<float> part((size_t)nb);
    long long pos = 0;
    for (int it = 0; it < steps; it++) {
        if (pos + batch > m->nset) pos = 0;
        st = load_batch(m, batch, m->set_x + (size_t)pos * kIn, m->set_t + (size_t)pos * kOut, hipMemcpyDeviceToDevice);
        if (st != GEL_OK) return st;
        pos += batch;
        mlp_forward_dev(m, batch);
        mlp_backward_dev(m, batch);
        m->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)m->adam_t), c2 = 1.0f - std::pow(b2, (float)m->adam_t);
        for (int l = 0; l < kLayers; l++) {
            const int n = (int)wsz(l), nbias = kOut_[l];
            hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, m->stream, n, m->w[l], m->gw[l], m->mw[l],
                               m->vw[l], lr, b1, b2, 1e-8f, c1, c2);
            hipLaunchKernelGGL(adam_kernel, dim3((nbias + 255) / 256), dim3(256), 0, m->stream, nbias, m->b[l], m->gb[l],
                               m->mb[l], m->vb[l], lr, b1, b2, 1e-8f, c1, c2);
        }
        mlp_quantize(m);
        if (loss_out) {
            if (hipMemcpyAsync(part.data(), m->partial, (size_t)nb * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
                hipStreamSynchronize(m->stream) != hipSuccess)
                return mfail(GEL_E_HIP, "train");
            double s = 0.0;
            for (float v : part) s += v;
            loss_out[it] = (float)(s / (double)(batch * kOut));
        }
    }
    if (hipStreamSynchronize(m->stream) != hipSuccess) return mfail(GEL_E_HIP, "train sync");
    if (hipGetLastError() != hipSuccess) return mfail(GEL_E_HIP, "train launch");
    return GEL_OK;
}

}  // extern "C"

extern "C" {

// Feature and target normalisation of the source model: x_n = (x - mu)/sd,
// y = y_n * ysd + ymu (targets in units of src_coef * ro * n_H).
int gel_mlp_set_scaling(gel_mlp* m, const float* mu, const float* sd, const float* ymu, const float* ysd)
{
    if (!m || !mu || !sd || !ymu || !ysd) return mfail(GEL_E_ARG, "null argument");
    float** bufs[] = { &m->mu, &m->sd, &m->ymu, &m->ysd };
    const float* src[] = { mu, sd, ymu, ysd };
    const int n[] = { kIn, kIn, kOut, kOut };
    for (int k = 0; k < 4; k++) {
        if (!*bufs[k] && hipMalloc((void**)bufs[k], 16 * sizeof(float)) != hipSuccess) return mfail(GEL_E_NOMEM, "scaling");
        if (hipMemcpy(*bufs[k], src[k], n[k] * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return mfail(GEL_E_HIP, "scaling copy");
    }
    return GEL_OK;
}

int gel_mlp_set_output_mode(gel_mlp* m, int mode)
{
    if (!m) return mfail(GEL_E_ARG, "null argument");
    if (mode != GEL_MLP_OUT_SOURCE && mode != GEL_MLP_OUT_FACTOR) return mfail(GEL_E_ARG, "unknown output mode");
    m->mode = mode;
    return GEL_OK;
}

static int check_ctx(gel_ctx* c)
{
    if (!c) return mfail(GEL_E_ARG, "null context");
    if (c->mesh) return mfail(GEL_E_UNSUPPORTED, "the learned source runs on polar grids");
    if (!c->uploaded) return mfail(GEL_E_STATE, "no state uploaded");
    if (!c->d_src) return mfail(GEL_E_STATE, "hydrogen moments not set (gel_set_sources)");
    return GEL_OK;
}

// Raw source-model features of every cell of the current state: x [ncells][14].
int gel_surrogate_features(gel_ctx* c, float* x)
{
    int st = check_ctx(c);
    if (st != GEL_OK) return st;
    if (!x) return mfail(GEL_E_ARG, "null argument");
    float* d = nullptr;
    if (hipMalloc((void**)&d, c->nc * kIn * sizeof(float)) != hipSuccess) return mfail(GEL_E_NOMEM, "features");
    const dim3 g((unsigned)((c->nc + 255) / 256));
    if (c->ts == 8)
        hipLaunchKernelGGL(features_kernel<double>, g, dim3(256), 0, c->stream, c->ni, c->nj, c->d_x, c->d_y,
                           (const double*)c->d_state[c->cur], (const double*)c->d_src, nullptr, nullptr, d);
    else
        hipLaunchKernelGGL(features_kernel<float>, g, dim3(256), 0, c->stream, c->ni, c->nj, c->d_x, c->d_y,
                           (const float*)c->d_state[c->cur], (const float*)c->d_src, nullptr, nullptr, d);
    hipError_t e = hipMemcpyAsync(x, d, c->nc * kIn * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    return e == hipSuccess ? GEL_OK : mfail(GEL_E_HIP, "features");
}

// Analytic charge-exchange source table of the current state: q [5][ncells].
int gel_source_analytic(gel_ctx* c, double* q)
{
    int st = check_ctx(c);
    if (st != GEL_OK) return st;
    if (!q) return mfail(GEL_E_ARG, "null argument");
    double* d = nullptr;
    if (hipMalloc((void**)&d, 5 * c->nc * sizeof(double)) != hipSuccess) return mfail(GEL_E_NOMEM, "table");
    const dim3 g((unsigned)((c->nc + 255) / 256));
    if (c->ts == 8)
        hipLaunchKernelGGL(analytic_table_kernel<double>, g, dim3(256), 0, c->stream, c->nc, c->prm,
                           (const double*)c->d_state[c->cur], (const double*)c->d_src, d);
    else
        hipLaunchKernelGGL(analytic_table_kernel<float>, g, dim3(256), 0, c->stream, c->nc, c->prm,
                           (const float*)c->d_state[c->cur], (const float*)c->d_src, d);
    hipError_t e = hipMemcpyAsync(q, d, 5 * c->nc * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    return e == hipSuccess ? GEL_OK : mfail(GEL_E_HIP, "analytic table");
}

// Refresh the context's source table from the network on the current state.  judge:
// compare every cell with the analytic source and fall back where the error exceeds
// tol; *n_fallback receives the number of such cells.  Runs on the context's stream.
int gel_surrogate_refresh(gel_ctx* c, gel_mlp* m, double tol, int judge, long long* n_fallback)
{
    int st = check_ctx(c);
    if (st != GEL_OK) return st;
    if (!m || !m->mu) return mfail(GEL_E_STATE, "network scaling not set");
    if (judge < 0 || judge > 2) return mfail(GEL_E_ARG, "judge is 0, 1 or 2");
    if (judge == 2 && c->prm.source_mode != GEL_SRC_HYBRID) return mfail(GEL_E_UNSUPPORTED, "judge 2 needs GEL_SRC_HYBRID");
    if (!c->d_qtab && hipMalloc(&c->d_qtab, 5 * c->nc * c->ts) != hipSuccess)
        return mfail(GEL_E_NOMEM, "source table");
    if (!c->d_qflag && (hipMalloc((void**)&c->d_qflag, c->nc * sizeof(int)) != hipSuccess ||
                        hipMemsetAsync(c->d_qflag, 0, c->nc * sizeof(int), c->stream) != hipSuccess))
        return mfail(GEL_E_NOMEM, "source flags");
    if (!m->d_rargs && (hipMalloc((void**)&m->d_rargs, sizeof(SweepArgs)) != hipSuccess ||
                        hipMalloc((void**)&m->d_rcount, sizeof(unsigned long long)) != hipSuccess))
        return mfail(GEL_E_NOMEM, "refresh buffers");
    SweepArgs a{};
    a.ni = c->ni; a.nj = c->nj;
    a.x = c->d_x; a.y = c->d_y; a.x2 = c->d_x; a.y2 = c->d_y;
    a.now = c->d_state[c->cur];
    a.next = c->d_state[c->cur];
    a.src = c->d_src;
    a.qtab = c->d_qtab;
    a.qflag = c->d_qflag;
    a.prm = c->prm;
    hipError_t e = hipMemcpyAsync(m->d_rargs, &a, sizeof a, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->d_rcount, 0, sizeof(unsigned long long), c->stream);
    if (e == hipSuccess)
        e = gel_refresh_launch(c->stream, m->d_rargs, nullptr, gel_refresh_tiles(c->ni, c->nj), m, c->ts == 4, tol,
                               judge, m->d_rcount);
    unsigned long long cnt = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&cnt, m->d_rcount, sizeof cnt, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return mfail(GEL_E_HIP, std::string("refresh: ") + hipGetErrorString(e));
    if (n_fallback) *n_fallback = (long long)cnt;
    return GEL_OK;
}

}  // extern "C"
