// gelios_sweep.hip - the C-ABI of include/gelios.h around the sweep kernels of
// sweep_kernel.h: contexts, state transfer, the FP64 launches (FP contraction off,
// the reference's arithmetic), batches, and the face-flux test kernels.
#include "sweep_kernel.h"

#include <mutex>

namespace gel {

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

// Batched HLLD for tests (gel_flux_hlld_batch): primitives [n][9] in, fluxes [n][9]
// and max wave speed [n] out.
__global__ void hlld_batch_kernel(int n, const double* __restrict__ L, const double* __restrict__ R,
                                  const double* __restrict__ W, const double* __restrict__ n1,
                                  const double* __restrict__ n2, double gam, double* __restrict__ P,
                                  double* __restrict__ uu)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uu[k] = hlld<double>(L + 9 * (size_t)k, R + 9 * (size_t)k, W[k], n1[k], n2[k], gam, P + 9 * (size_t)k);
}

// The sweep's FP64 division and square root (gel_fp64_div_sqrt): q = a / b as dv, three
// quotients sharing b's reciprocal as the limiter and the update take them, and the
// square root of a as sqn.
__global__ void fp64_ops_kernel(int n, const double* __restrict__ a, const double* __restrict__ b,
                                double* __restrict__ q, double* __restrict__ q2, double* __restrict__ s)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Den d = den(b[k]);
    q[k] = dv(a[k], b[k]);
    q2[k] = qt(a[(k + 1) % n], d);
    s[k] = sqn(a[k]);
}

// Batched exact Riemann solver for tests (gel_flux_godunov_batch): out [n][10] =
// q0..q4, dsl, dsp, dsc, dt, status.
__global__ void godunov_batch_kernel(int n, const double* __restrict__ L, const double* __restrict__ R,
                                     const double* __restrict__ nrm, const double* __restrict__ W,
                                     const double* __restrict__ dist, double cfl, double gam,
                                     double* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double q[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 }, dsl = 0.0, dsp = 0.0, dsc = 0.0;
    int st = 0;
    const double* nk = nrm + 3 * (size_t)k;
    const double t = godunov_exact(L + 5 * (size_t)k, R + 5 * (size_t)k, nk[0], nk[1], nk[2], W[k], dist[k], cfl, gam,
                                   q, dsl, dsp, dsc, st);
    double* o = out + 10 * (size_t)k;
    for (int i = 0; i < 5; i++) o[i] = q[i];
    o[5] = dsl; o[6] = dsp; o[7] = dsc; o[8] = t; o[9] = st;
}

// The same problems one per wavefront (four per work-group), as the fitted-surface
// kernel solves them (godunov_exact<true>): out as godunov_batch_kernel.
__global__ void __launch_bounds__(256) godunov_wave_kernel(int n, const double* __restrict__ L,
                                                           const double* __restrict__ R,
                                                           const double* __restrict__ nrm,
                                                           const double* __restrict__ W,
                                                           const double* __restrict__ dist, double cfl, double gam,
                                                           double* __restrict__ out)
{
    const int k = blockIdx.x * 4 + (int)(threadIdx.x / kWave);
    if (k >= n) return;
    double q[5] = { 0.0, 0.0, 0.0, 0.0, 0.0 }, dsl = 0.0, dsp = 0.0, dsc = 0.0;
    int st = 0;
    const double* nk = nrm + 3 * (size_t)k;
    const double t = godunov_exact<true>(L + 5 * (size_t)k, R + 5 * (size_t)k, nk[0], nk[1], nk[2], W[k], dist[k],
                                         cfl, gam, q, dsl, dsp, dsc, st);
    if ((threadIdx.x & (kWave - 1)) != 0) return;
    double* o = out + 10 * (size_t)k;
    for (int i = 0; i < 5; i++) o[i] = q[i];
    o[5] = dsl; o[6] = dsp; o[7] = dsc; o[8] = t; o[9] = st;
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

int gel_set_error(int st, const std::string& msg) { return fail(st, msg); }

#define HIP_OK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(GEL_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

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
    p->precision = GEL_PREC_FP64;
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
    p->src_coef_h = 1.0 / 6.0;                              // n_H_LISM_ / Kn_
    p->h_pow = 0.5;
    p->h1_in_rho = 0.0000001;
    p->h1_in_p = (0.0000001 * chi * chi / (ga * 14.1344 * 14.1344));
    p->h4_inf_rho = 1.0; p->h4_inf_p = 0.5; p->h4_inf_u = -2.54338; p->h4_inf_v = 0.0;
}

static void free_ctx(gel_ctx* c)
{
    if (!c) return;
    (void)hipFree(c->d_x); (void)hipFree(c->d_y); (void)hipFree(c->d_x2); (void)hipFree(c->d_y2);
    (void)hipFree(c->d_state[0]); (void)hipFree(c->d_state[1]);
    (void)hipFree(c->d_ctype); (void)hipFree(c->d_otype); (void)hipFree(c->d_src); (void)hipFree(c->d_qtab);
    (void)hipFree(c->d_qflag);
    (void)hipFree(c->d_mfq);
    (void)hipFree(c->d_geo);
    (void)hipFree(c->d_trk);
    (void)hipFree(c->d_dt); (void)hipFree(c->d_cnt);
    gel_mesh_free(c);
    for (auto& e : c->events) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void gel_ctx_destroy(gel_ctx* ctx) { free_ctx(ctx); }

int gel_ctx_create(gel_ctx** out, const gel_grid* g, const gel_params* prm)
{
    if (!out || !g || !prm || !g->x || !g->y) return fail(GEL_E_ARG, "null argument");
    *out = nullptr;
    if (g->ni < 3 || g->nj < 3) return fail(GEL_E_ARG, "grid needs at least 3x3 cells");
    if (prm->physics < GEL_PHYS_GAS || prm->physics > GEL_PHYS_GAS_INNER) return fail(GEL_E_UNSUPPORTED, "physics");
    if (prm->physics == GEL_PHYS_GAS_INNER && (prm->precision != GEL_PREC_FP64 ||
                                               (prm->source_mode != GEL_SRC_NONE && prm->source_mode != GEL_SRC_MK1)))
        return fail(GEL_E_UNSUPPORTED, "inner-region solver: FP64, no source or the MK1 source");
    if (prm->physics != GEL_PHYS_MHD && prm->flux != GEL_FLUX_HLL && prm->flux != GEL_FLUX_HLLC)
        return fail(GEL_E_UNSUPPORTED, "flux");
    if (prm->physics == GEL_PHYS_GAS4H && (prm->precision != GEL_PREC_FP64 || prm->source_mode != GEL_SRC_NONE))
        return fail(GEL_E_UNSUPPORTED, "plasma + hydrogen: FP64, charge exchange built in (source_mode none)");
    if (prm->source_mode < GEL_SRC_NONE || prm->source_mode > GEL_SRC_HYBRID)
        return fail(GEL_E_UNSUPPORTED, "source_mode");
    if (prm->precision != GEL_PREC_FP64 && prm->precision != GEL_PREC_FP32_BULK) return fail(GEL_E_UNSUPPORTED, "precision");
    if (!(prm->dt0 > 0.0) || !(prm->cfl > 0.0) || !(prm->gamma > 1.0)) return fail(GEL_E_ARG, "dt0, cfl, gamma");
    gel_ctx* c = new (std::nothrow) gel_ctx();
    if (!c) return fail(GEL_E_NOMEM, "host allocation");
    c->ni = g->ni; c->nj = g->nj;
    c->nc = (size_t)g->ni * g->nj;
    c->nn = (size_t)(g->ni + 1) * (g->nj + 1);
    c->nv = prm->physics == GEL_PHYS_MHD ? 9 : (prm->physics == GEL_PHYS_GAS4H ? 21 : 5);
    c->ts = prm->precision == GEL_PREC_FP32_BULK ? 4 : 8;
    c->prm = *prm;
    c->rows = 0;          // chosen at the first launch from the kernel's occupancy
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
        if ((e = hipMalloc(&c->d_state[b], (size_t)c->nv * c->nc * c->ts)) != hipSuccess) return bad(e, "state");
    if ((e = hipMalloc(&c->d_otype, (size_t)g->nj)) != hipSuccess) return bad(e, "outer types");
    if ((e = hipMalloc(&c->d_dt, kDtWords * sizeof(unsigned long long))) != hipSuccess) return bad(e, "dt");
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
        // the plasma + hydrogen sweep assigns hydrogen populations by every cell type
        if (any || prm->physics == GEL_PHYS_GAS4H || prm->physics == GEL_PHYS_GAS_INNER) {
            if (prm->physics == GEL_PHYS_MHD || prm->precision != GEL_PREC_FP64) {
                free_ctx(c);
                return fail(GEL_E_UNSUPPORTED, "C_1 / C_centr cells need FP64 gas physics");
            }
            if ((e = hipMalloc(&c->d_ctype, c->nc)) != hipSuccess) return bad(e, "cell types");
            if ((e = hipMemcpy(c->d_ctype, g->cell_type, c->nc, hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "copy");
            c->has_types = true;
        }
    }
    // the context's stream does not synchronise with the null stream: clear on it
    if ((e = hipMemsetAsync(c->d_cnt, 0, 3 * sizeof(unsigned long long), c->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(c->stream)) != hipSuccess)
        return bad(e, "memset");
    *out = c;
    return GEL_OK;
}

int gel_ctx_create_mesh(gel_ctx** out, const gel_mesh* g, const gel_params* prm)
{
    if (!out || !g || !prm || !g->x || !g->y || !g->corner_offset || !g->corners || !g->cell_type ||
        !g->face_offset || !g->face_a || !g->face_b || !g->face_type || !g->face_sosed || !g->face_down || !g->face_up)
        return fail(GEL_E_ARG, "null argument");
    *out = nullptr;
    if (g->n_cells < 1 || g->n_nodes < 3) return fail(GEL_E_ARG, "empty mesh");
    if (g->corner_offset[0] != 0 || g->face_offset[0] != 0) return fail(GEL_E_ARG, "offsets start at 0");
    for (int k = 0; k < g->n_cells; k++)
        if (g->corner_offset[k + 1] < g->corner_offset[k] || g->face_offset[k + 1] < g->face_offset[k])
            return fail(GEL_E_ARG, "offsets decrease");
    if (prm->physics != GEL_PHYS_GAS || (prm->precision != GEL_PREC_FP64 && prm->precision != GEL_PREC_FP32_BULK))
        return fail(GEL_E_UNSUPPORTED, "general meshes: gas physics, FP64 or FP32 bulk");
    if (prm->flux != GEL_FLUX_HLL && prm->flux != GEL_FLUX_HLLC) return fail(GEL_E_UNSUPPORTED, "flux");
    if (prm->source_mode != GEL_SRC_NONE && prm->source_mode != GEL_SRC_MK1 && prm->source_mode != GEL_SRC_TABLE)
        return fail(GEL_E_UNSUPPORTED, "general meshes: no source, the MK1 source or a source table");
    if (!(prm->dt0 > 0.0) || !(prm->cfl > 0.0) || !(prm->gamma > 1.0)) return fail(GEL_E_ARG, "dt0, cfl, gamma");
    gel_ctx* c = new (std::nothrow) gel_ctx();
    if (!c) return fail(GEL_E_NOMEM, "host allocation");
    c->nc = (size_t)g->n_cells;
    c->nn = (size_t)g->n_nodes;
    c->nv = 5;
    c->ts = prm->precision == GEL_PREC_FP32_BULK ? 4 : 8;
    c->prm = *prm;
    c->has_types = true;
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
        if ((e = hipMalloc(&c->d_state[b], (size_t)c->nv * c->nc * c->ts)) != hipSuccess) return bad(e, "state");
    if ((e = hipMalloc(&c->d_dt, kDtWords * sizeof(unsigned long long))) != hipSuccess) return bad(e, "dt");
    if ((e = hipMalloc(&c->d_cnt, 3 * sizeof(unsigned long long))) != hipSuccess) return bad(e, "counters");
    if ((e = hipMemcpy(c->d_x, g->x, nb, hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "copy");
    if ((e = hipMemcpy(c->d_y, g->y, nb, hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "copy");
    std::string err;
    const int st = gel_mesh_build(c, g, err);
    if (st != GEL_OK) { free_ctx(c); return fail(st, err); }
    if ((e = hipMemsetAsync(c->d_cnt, 0, 3 * sizeof(unsigned long long), c->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(c->stream)) != hipSuccess)
        return bad(e, "memset");
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

static int copy_state_in(gel_ctx* c, void* dst, const double* prim)
{
    const size_t n = (size_t)c->nv * c->nc;
    if (c->ts == 8) {
        HIP_OK(hipMemcpyAsync(dst, prim, n * 8, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> f(n);
        for (size_t k = 0; k < n; k++) f[k] = (float)prim[k];
        HIP_OK(hipMemcpyAsync(dst, f.data(), n * 4, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    return GEL_OK;
}

int gel_upload_state(gel_ctx* c, const double* prim)
{
    if (!c || !prim) return fail(GEL_E_ARG, "null argument");
    int st = copy_state_in(c, c->d_state[0], prim);
    if (st != GEL_OK) return st;
    st = copy_state_in(c, c->d_state[1], prim);
    if (st != GEL_OK) return st;
    std::vector<unsigned long long> dt(kDtWords, dbits(__builtin_huge_val()));
    std::fill(dt.begin(), dt.begin() + kDtSlot, dbits(c->prm.dt0));
    HIP_OK(hipMemcpyAsync(c->d_dt, dt.data(), kDtWords * sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    c->cur = 0;
    c->steps = 0;
    c->motion_step = 0;
    c->uploaded = true;
    return GEL_OK;
}

int gel_download_state(gel_ctx* c, double* prim)
{
    if (!c || !prim) return fail(GEL_E_ARG, "null argument");
    if (!c->uploaded) return fail(GEL_E_STATE, "no state uploaded");
    const size_t n = (size_t)c->nv * c->nc;
    if (c->ts == 8) {
        HIP_OK(hipMemcpyAsync(prim, c->d_state[c->cur], n * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> f(n);
        HIP_OK(hipMemcpyAsync(f.data(), c->d_state[c->cur], n * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < n; k++) prim[k] = (double)f[k];
    }
    return GEL_OK;
}

int gel_set_sources(gel_ctx* c, const double* f)
{
    if (!c || !f) return fail(GEL_E_ARG, "null argument");
    if (c->prm.physics == GEL_PHYS_GAS4H) return fail(GEL_E_UNSUPPORTED, "the hydrogen fluids are part of the state");
    // stored in the state's type: FP32 for the FP32-bulk sweep, whose closure runs in FP32;
    // the work queued on the context's stream finishes first (it does not synchronise
    // with the null stream the blocking copy runs on)
    HIP_OK(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)GEL_MK1_NFIELDS * c->nc;
    if (!c->d_src) {
        hipError_t e = hipMalloc(&c->d_src, n * c->ts);
        if (e != hipSuccess) return fail(GEL_E_NOMEM, "sources");
    }
    if (c->ts == 8) {
        HIP_OK(hipMemcpy(c->d_src, f, n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> h(n);
        for (size_t k = 0; k < n; k++) h[k] = (float)f[k];
        HIP_OK(hipMemcpy(c->d_src, h.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    return GEL_OK;
}

int gel_set_source_table(gel_ctx* c, const double* q)
{
    if (!c || !q) return fail(GEL_E_ARG, "null argument");
    if (c->prm.physics == GEL_PHYS_GAS4H) return fail(GEL_E_UNSUPPORTED, "the hydrogen fluids are part of the state");
    // stored in the state's type, like the moments; ordered after the work queued on
    // the context's stream
    HIP_OK(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)5 * c->nc;
    if (!c->d_qflag && hipMalloc((void**)&c->d_qflag, c->nc * sizeof(int)) != hipSuccess)
        return fail(GEL_E_NOMEM, "source flags");
    HIP_OK(hipMemsetAsync(c->d_qflag, 0, c->nc * sizeof(int), c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (!c->d_qtab) {
        hipError_t e = hipMalloc(&c->d_qtab, n * c->ts);
        if (e != hipSuccess) return fail(GEL_E_NOMEM, "source table");
    }
    if (c->ts == 8) {
        HIP_OK(hipMemcpy(c->d_qtab, q, n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> h(n);
        for (size_t k = 0; k < n; k++) h[k] = (float)q[k];
        HIP_OK(hipMemcpy(c->d_qtab, h.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    return GEL_OK;
}

int gel_get_source_table(gel_ctx* c, double* q)
{
    if (!c || !q) return fail(GEL_E_ARG, "null argument");
    if (!c->d_qtab) return fail(GEL_E_STATE, "source table not set");
    HIP_OK(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)5 * c->nc;
    if (c->ts == 8) {
        HIP_OK(hipMemcpy(q, c->d_qtab, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> h(n);
        HIP_OK(hipMemcpy(h.data(), c->d_qtab, n * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; k++) q[k] = h[k];
    }
    return GEL_OK;
}

int gel_get_source_flags(gel_ctx* c, int* flags)
{
    if (!c || !flags) return fail(GEL_E_ARG, "null argument");
    if (!c->d_qflag) return fail(GEL_E_STATE, "source table not set");
    HIP_OK(hipStreamSynchronize(c->stream));
    HIP_OK(hipMemcpy(flags, c->d_qflag, c->nc * sizeof(int), hipMemcpyDeviceToHost));
    return GEL_OK;
}

int gel_set_rows_per_wave(gel_ctx* c, int rows)
{
    if (c && c->mesh) return fail(GEL_E_UNSUPPORTED, "rows per chunk: polar grids");
    if (!c || rows < 1 || rows > 4096) return fail(GEL_E_ARG, "rows");
    c->rows = rows;
    return GEL_OK;
}

int gel_set_motion(gel_ctx* c, double amp, int period, double phase)
{
    if (c && c->mesh) return fail(GEL_E_UNSUPPORTED, "prescribed motion: polar grids (general meshes: gel_step_moving)");
    if (!c || period < 1 || !(amp >= 0.0) || !(amp < 0.5)) return fail(GEL_E_ARG, "motion: amp in [0,0.5), period >= 1");
    if (c->prm.physics == GEL_PHYS_GAS_INNER) return fail(GEL_E_UNSUPPORTED, "the inner-region solver is static");
    if (c->tracking) return fail(GEL_E_STATE, "fitted-surface tracking is on");
    c->motion.assign((size_t)period, 1.0);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < period; k++) c->motion[k] = 1.0 + amp * std::sin(2.0 * pi * k / period + phase);
    c->motion_step = 0;
    return GEL_OK;
}

void gel_tracking_reference(gel_tracking* t)
{
    if (!t) return;
    std::memset(t, 0, sizeof(*t));
    const double RR = 390.222;                         // Help.h:75
    t->row_inner_shock = t->row_contact = t->row_outer_shock = -1;
    t->outer_faces = 1 << 30;
    t->average_lines = 0;
    t->inner_rail_nodes = 20;                                   // Help.h:66
    t->r_inner = 1.0 / RR;                             // R1_, R111_, R5_ (Help.h:76-83)
    t->r_mid = 50.0 / RR;
    t->r_outer = 2200.0 / RR;
    t->p_inner = 1.3; t->p_mid = 1.05; t->p_outer = 1.76; t->s_outer = 1.5708;   // Setka.cpp:4349-4380
    t->koef = 0.1;                                     // Setka.cpp:3380
    t->k_contact = 0.3; t->k_inner = 0.1; t->k_outer = 1.0;
    t->tang_inner = 0.1; t->tang_outer = 0.01;
}

int gel_set_mesh_tracking(gel_ctx* c, const gel_tracking* t, const gel_mesh_rails* r)
{
    if (!c) return fail(GEL_E_ARG, "null argument");
    if (!c->mesh) return fail(GEL_E_UNSUPPORTED, "mesh tracking: general-mesh contexts (polar grids: gel_set_tracking)");
    if (t && (c->prm.physics != GEL_PHYS_GAS || c->prm.precision != GEL_PREC_FP64))
        return fail(GEL_E_UNSUPPORTED, "fitted surfaces: gas physics in FP64");
    HIP_OK(hipStreamSynchronize(c->stream));
    std::string err;
    const int st = gel_mesh_set_tracking(c, t, r, err);
    if (st != GEL_OK) { c->tracking = false; return fail(st, err); }
    c->tracking = t != nullptr;
    if (t) c->trk = *t;
    return GEL_OK;
}

int gel_set_tracking(gel_ctx* c, const gel_tracking* t, const double* rail)
{
    if (c && c->mesh) return fail(GEL_E_UNSUPPORTED, "fitted-surface tracking: polar grids");
    if (!c) return fail(GEL_E_ARG, "null argument");
    if (!t) { c->tracking = false; return GEL_OK; }
    if (!rail) return fail(GEL_E_ARG, "rail angles needed");
    if (c->prm.physics != GEL_PHYS_GAS || c->prm.precision != GEL_PREC_FP64)
        return fail(GEL_E_UNSUPPORTED, "fitted surfaces: gas physics in FP64 (the single-fluid sweep)");
    if (c->prm.source_mode != GEL_SRC_NONE && c->prm.source_mode != GEL_SRC_MK1)
        return fail(GEL_E_UNSUPPORTED, "fitted surfaces: no source or the MK1 source");
    if (!c->motion.empty()) return fail(GEL_E_STATE, "prescribed motion is set");
    const int ni = c->ni, nj = c->nj;
    const int r0 = t->row_inner_shock, r1 = t->row_contact, r2 = t->row_outer_shock;
    if (nj < 2 || t->inner_rail_nodes < 2 || r0 < t->inner_rail_nodes || r0 < 2 || !(r0 < r1 && r1 < r2) || r2 > ni - 2)
        return fail(GEL_E_ARG, "surface rows: inner_rail_nodes <= inner shock < contact < outer shock <= ni - 2");
    if (t->outer_faces < 0) return fail(GEL_E_ARG, "outer_faces must be >= 0");
    const int M1 = r0 + 1, M4 = ni - r2;
    const int nmid = M1 - 1 - t->inner_rail_nodes;
    const size_t n_cs = (size_t)(nj + 1), n_in = (size_t)t->inner_rail_nodes, n_mid = (size_t)(nmid > 0 ? nmid : 0);
    const size_t n_out = (size_t)(nj + 1) * M4, n_face = (size_t)3 * nj * 6;
    std::vector<double> h(2 * n_cs + n_in + n_mid + n_out);
    double* cs = h.data();
    double* sn = cs + n_cs;
    double* fin = sn + n_cs;
    double* fmid = fin + n_in;
    double* fout = fmid + n_mid;
    // host libm, the functions the reference calls at run time (Setka.cpp:4331-4380)
    const double dd = t->s_outer / std::sqrt(1.0 - 1.0 / t->p_outer);
    for (int k = 0; k <= nj; k++) {
        cs[k] = std::cos(rail[k]);
        sn[k] = std::sin(rail[k]);
        const double ex = t->p_outer * (1.0 - (rail[k] / dd) * (rail[k] / dd));
        for (int j = 0; j < M4; j++) fout[(size_t)k * M4 + j] = std::pow((j + 1) * 1.0 / (M4), ex);
    }
    for (int j = 0; j < t->inner_rail_nodes; j++) fin[j] = std::pow(j * 1.0 / (t->inner_rail_nodes), t->p_inner);
    for (int j = t->inner_rail_nodes; j < M1 - 1; j++)
        fmid[j - t->inner_rail_nodes] = std::pow((j - t->inner_rail_nodes) * 1.0 / (M1 - 1 - t->inner_rail_nodes), t->p_mid);
    (void)hipFree(c->d_trk);
    c->d_trk = nullptr;
    c->tracking = false;
    if (hipMalloc(&c->d_trk, (h.size() + n_face) * sizeof(double)) != hipSuccess)
        return fail(GEL_E_NOMEM, "tracking tables");
    HIP_OK(hipMemcpy(c->d_trk, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    TrackDev& d = c->tdev;
    d.rows[0] = r1; d.rows[1] = r0; d.rows[2] = r2;
    d.outer_faces = t->outer_faces;
    d.average_lines = t->average_lines ? 1 : 0;
    d.n_inner = t->inner_rail_nodes;
    d.m4 = M4;
    d.koef = t->koef;
    d.k_line[0] = t->k_contact; d.k_line[1] = t->k_inner; d.k_line[2] = t->k_outer;
    d.tang_inner = t->tang_inner; d.tang_outer = t->tang_outer;
    d.r_inner = t->r_inner; d.r_mid = t->r_mid; d.r_outer = t->r_outer;
    d.cs = c->d_trk;
    d.sn = d.cs + n_cs;
    d.f_in = d.sn + n_cs;
    d.f_mid = d.f_in + n_in;
    d.f_out = d.f_mid + n_mid;
    d.face = c->d_trk + h.size();
    c->trk = *t;
    c->tracking = true;
    return GEL_OK;
}

int gel_get_nodes(gel_ctx* c, double* x, double* y)
{
    if (!c || !x || !y) return fail(GEL_E_ARG, "null argument");
    HIP_OK(hipStreamSynchronize(c->stream));
    HIP_OK(hipMemcpy(x, c->d_x, c->nn * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(y, c->d_y, c->nn * sizeof(double), hipMemcpyDeviceToHost));
    return GEL_OK;
}

}  // extern "C"

static hipError_t launch(gel_ctx* c, SweepArgs& a, bool moving)
{
    const int src = c->prm.source_mode;
    if (c->prm.precision == GEL_PREC_FP32_BULK) return launch_sweep_fp32(c, a, moving);
    hipStream_t s = c->stream;
    if (c->prm.physics == GEL_PHYS_MHD) return launch_ms<MhdPhys, double, double, false, 3>(c, a, s, moving, src);
    if (c->prm.physics == GEL_PHYS_GAS4H) return launch_mf_step(c, a, moving);
    if (c->prm.physics == GEL_PHYS_GAS_INNER) {
        if (moving) return hipErrorInvalidValue;           // rejected by the step entry points first
        return launch_inner_step(c, a);
    }
    const bool hllc = (c->prm.flux == GEL_FLUX_HLLC);
    if (c->has_types) {
        return hllc ? launch_ms<GasPhys, double, double, true, GEL_FLUX_HLLC>(c, a, s, moving, src)
                    : launch_ms<GasPhys, double, double, true, GEL_FLUX_HLL>(c, a, s, moving, src);
    }
    return hllc ? launch_ms<GasPhys, double, double, false, GEL_FLUX_HLLC>(c, a, s, moving, src)
                : launch_ms<GasPhys, double, double, false, GEL_FLUX_HLL>(c, a, s, moving, src);
}

// The geometry table of a static grid (GeoTable in sweep_kernel.h), built on the
// context's stream at the first static step after the nodes last moved.
static int ensure_geometry(gel_ctx* c)
{
    if (c->geo_valid) return GEL_OK;
    if (!c->d_geo && hipMalloc(&c->d_geo, GeoTable::doubles(c->ni, c->nj) * sizeof(double)) != hipSuccess)
        return fail(GEL_E_NOMEM, "geometry table");
    const long long n = (long long)(c->ni + 1) * (c->nj + 1);
    const dim3 g((unsigned)((n + 255) / 256));
    if (c->prm.precision == GEL_PREC_FP32_BULK)
        hipLaunchKernelGGL(geometry_kernel<float>, g, dim3(256), 0, c->stream, c->ni, c->nj, c->d_x, c->d_y, c->d_geo);
    else
        hipLaunchKernelGGL(geometry_kernel<double>, g, dim3(256), 0, c->stream, c->ni, c->nj, c->d_x, c->d_y, c->d_geo);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GEL_E_HIP, std::string("geometry: ") + hipGetErrorString(e));
    c->geo_valid = true;
    return GEL_OK;
}

// moving: 0 static, 1 explicit next positions in d_x2/d_y2, 2 prescribed motion,
// 3 fitted-surface tracking (next positions computed into d_x2/d_y2, then committed)
static int one_step(gel_ctx* c, int moving)
{
    if (moving == 0 && !c->mesh) {
        const int st = ensure_geometry(c);
        if (st != GEL_OK) return st;
    }
    SweepArgs a;
    a.ni = c->ni; a.nj = c->nj; a.rows = c->rows;
    a.dt_rd = (int)(c->steps % 3);
    a.dt_wr = (int)((c->steps + 1) % 3);
    a.dt_clr = (int)((c->steps + 2) % 3);
    a.s_old = 1.0; a.s_new = 1.0;
    a.x = c->d_x; a.y = c->d_y;
    a.x2 = c->d_x; a.y2 = c->d_y;
    if (moving == 1) { a.x2 = c->d_x2; a.y2 = c->d_y2; }
    if (moving == 2) {
        const size_t P = c->motion.size();
        a.s_old = c->motion[(size_t)(c->motion_step % (long long)P)];
        a.s_new = c->motion[(size_t)((c->motion_step + 1) % (long long)P)];
    }
    a.now = c->d_state[c->cur];
    a.next = c->d_state[c->cur ^ 1];
    a.ctype = c->d_ctype; a.otype = c->d_otype;
    a.src = c->d_src;
    a.qtab = c->d_qtab;
    a.qflag = c->d_qflag;
    a.geo = c->d_geo;
    a.srow[0] = a.srow[1] = a.srow[2] = -1;
    a.fluid = 0;
    for (int k = 0; k < 5; k++) a.voff[k] = k;
    if (moving == 3) { a.srow[0] = c->trk.row_inner_shock; a.srow[1] = c->trk.row_contact; a.srow[2] = c->trk.row_outer_shock; }
    a.dtbuf = c->d_dt;
    a.counters = c->d_cnt;
    a.prm = c->prm;
    if ((c->prm.source_mode == GEL_SRC_MK1 || c->prm.source_mode == GEL_SRC_HYBRID) && !c->d_src)
        return fail(GEL_E_STATE, "sources not set");
    if ((c->prm.source_mode == GEL_SRC_TABLE || c->prm.source_mode == GEL_SRC_HYBRID) && !c->d_qtab)
        return fail(GEL_E_STATE, "source table not set");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing == 1) {
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        HIP_OK(hipEventRecord(e0, c->stream));
    }
    hipError_t e = c->mesh ? gel_mesh_step(c, a, moving != 0)
                           : (moving == 3) ? launch_tracked_step(c, a) : launch(c, a, moving != 0);
    if (e != hipSuccess) return fail(GEL_E_HIP, std::string("sweep launch: ") + hipGetErrorString(e));
    if (c->timing == 1) {
        HIP_OK(hipEventRecord(e1, c->stream));
        c->events.emplace_back(e0, e1);
        c->event_steps.push_back(1);
    }
    c->cur ^= 1;
    c->steps++;
    if (moving == 2) c->motion_step++;
    if (moving == 3) {                 // Point::x = x2 (Setka.cpp:10473-10480)
        std::swap(c->d_x, c->d_x2);
        std::swap(c->d_y, c->d_y2);
        c->geo_valid = false;
    }
    return GEL_OK;
}

static int fill_stats(gel_ctx* c, gel_stats* out)
{
    if (!out) return GEL_OK;
    unsigned long long dt[kDtWords], cnt[3];
    HIP_OK(hipMemcpyAsync(dt, c->d_dt, sizeof dt, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(cnt, c->d_cnt, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    auto slot = [&](long long k) {                          // the minimum of a slot's entries
        unsigned long long m = dt[k * kDtSlot];
        for (int e = 1; e < kDtLanes; e++) m = std::min(m, dt[k * kDtSlot + e * kDtStride]);
        return bitsd(m);
    };
    out->steps = c->steps;
    out->dt_next = slot(c->steps % 3);
    out->dt_last = c->steps > 0 ? slot((c->steps - 1) % 3) : c->prm.dt0;
    out->n_rho_clamp = (long long)cnt[0];
    out->n_p_clamp = (long long)cnt[1];
    out->n_numeric = (long long)cnt[2];
    return GEL_OK;
}

template <class F>
static int batch_run(int n, std::initializer_list<std::pair<const double*, size_t>> ins,
                     std::initializer_list<std::pair<double*, size_t>> outs, F&& launch_fn)
{
    std::vector<double*> din, dout;
    auto cleanup = [&]() { for (auto p : din) (void)hipFree(p); for (auto p : dout) (void)hipFree(p); };
    for (auto& in : ins) {
        double* d = nullptr;
        if (hipMalloc(&d, in.second * n * sizeof(double)) != hipSuccess) { cleanup(); return fail(GEL_E_NOMEM, "batch"); }
        din.push_back(d);
        if (hipMemcpy(d, in.first, in.second * n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            cleanup();
            return fail(GEL_E_HIP, "batch copy");
        }
    }
    for (auto& o : outs) {
        double* d = nullptr;
        if (hipMalloc(&d, o.second * n * sizeof(double)) != hipSuccess) { cleanup(); return fail(GEL_E_NOMEM, "batch"); }
        dout.push_back(d);
    }
    launch_fn(din, dout);
    hipError_t e = hipGetLastError();
    size_t k = 0;
    for (auto& o : outs) {
        if (e == hipSuccess) e = hipMemcpy(o.first, dout[k], o.second * n * sizeof(double), hipMemcpyDeviceToHost);
        k++;
    }
    cleanup();
    if (e != hipSuccess) return fail(GEL_E_HIP, std::string("batch: ") + hipGetErrorString(e));
    return GEL_OK;
}

extern "C" {

int gel_step(gel_ctx* c, int nsteps, gel_stats* out)
{
    if (!c || nsteps < 0) return fail(GEL_E_ARG, "bad argument");
    if (!c->uploaded) return fail(GEL_E_STATE, "no state uploaded");
    const int mode = c->tracking ? 3 : (c->motion.empty() ? 0 : 2);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing == 2 && nsteps > 0) {
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        HIP_OK(hipEventRecord(e0, c->stream));
    }
    for (int s = 0; s < nsteps; s++) {
        const int st = one_step(c, mode);
        if (st != GEL_OK) return st;
    }
    if (e0) {
        HIP_OK(hipEventRecord(e1, c->stream));
        c->events.emplace_back(e0, e1);
        c->event_steps.push_back(nsteps);
    }
    return fill_stats(c, out);
}

int gel_step_moving(gel_ctx* c, const double* x2, const double* y2, gel_stats* out)
{
    if (!c || !x2 || !y2) return fail(GEL_E_ARG, "null argument");
    if (!c->uploaded) return fail(GEL_E_STATE, "no state uploaded");
    if (!c->motion.empty()) return fail(GEL_E_STATE, "prescribed motion is set");
    if (c->tracking) return fail(GEL_E_STATE, "fitted-surface tracking moves the nodes");
    if (c->prm.physics == GEL_PHYS_GAS_INNER) return fail(GEL_E_UNSUPPORTED, "the inner-region solver is static");
    const size_t nb = c->nn * sizeof(double);
    HIP_OK(hipMemcpyAsync(c->d_x2, x2, nb, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(c->d_y2, y2, nb, hipMemcpyHostToDevice, c->stream));
    const int st = one_step(c, 1);
    if (st != GEL_OK) return st;
    std::swap(c->d_x, c->d_x2);       // Point::x = x2 after the step (Setka.cpp:10473-10480)
    std::swap(c->d_y, c->d_y2);
    c->geo_valid = false;
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
    const std::vector<unsigned long long> b(kDtSlot, dbits(dt));   // every entry of the step's slot
    HIP_OK(hipMemcpyAsync(c->d_dt + (c->steps % 3) * kDtSlot, b.data(), kDtSlot * sizeof(unsigned long long),
                          hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return GEL_OK;
}

int gel_enable_timing(gel_ctx* c, int on)
{
    if (!c || on < 0 || on > 2) return fail(GEL_E_ARG, "timing mode is 0, 1 or 2");
    c->timing = on;
    return GEL_OK;
}

int gel_kernel_timing(gel_ctx* c, double* avg_ms, long long* n)
{
    if (!c || !avg_ms || !n) return fail(GEL_E_ARG, "null argument");
    HIP_OK(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    long long steps = 0;
    for (size_t k = 0; k < c->events.size(); k++) {
        float ms = 0.0f;
        HIP_OK(hipEventElapsedTime(&ms, c->events[k].first, c->events[k].second));
        tot += ms;
        steps += c->event_steps[k];
        (void)hipEventDestroy(c->events[k].first);
        (void)hipEventDestroy(c->events[k].second);
    }
    *n = steps;
    *avg_ms = steps == 0 ? 0.0 : tot / (double)steps;
    c->events.clear();
    c->event_steps.clear();
    return GEL_OK;
}

int gel_flux_hllc2d_batch(int n, const double* L, const double* R, const double* W, const double* n1,
                          const double* n2, const double* rad, double cfl, double gamma, int metod, int nul_potok,
                          double* P, double* PQ, double* Vl, double* Vc, double* Vp, double* dt)
{
    if (n < 0 || (n > 0 && (!L || !R || !W || !n1 || !n2 || !rad || !P || !PQ || !Vl || !Vc || !Vp || !dt)))
        return fail(GEL_E_ARG, "null argument");
    if (n == 0) return GEL_OK;
    std::vector<double> o((size_t)9 * n);
    const int st = batch_run(n, { { L, 5 }, { R, 5 }, { W, 1 }, { n1, 1 }, { n2, 1 }, { rad, 1 } }, { { o.data(), 9 } },
                             [&](std::vector<double*>& i, std::vector<double*>& d) {
                                 hipLaunchKernelGGL(flux_batch_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, n, i[0],
                                                    i[1], i[2], i[3], i[4], i[5], cfl, gamma, metod, nul_potok, d[0]);
                             });
    if (st != GEL_OK) return st;
    for (int k = 0; k < n; k++) {
        const double* q = &o[(size_t)9 * k];
        P[4 * k] = q[0]; P[4 * k + 1] = q[1]; P[4 * k + 2] = q[2]; P[4 * k + 3] = q[3];
        PQ[k] = q[4]; Vl[k] = q[5]; Vc[k] = q[6]; Vp[k] = q[7]; dt[k] = q[8];
    }
    return GEL_OK;
}

static int godunov_batch(bool wave, int n, const double* L, const double* R, const double* nrm, const double* W,
                         const double* dist, double cfl, double gamma, double* q, double* speeds, double* dt,
                         int* status)
{
    if (n < 0 || (n > 0 && (!L || !R || !nrm || !W || !dist || !q || !speeds || !dt || !status)))
        return fail(GEL_E_ARG, "null argument");
    if (n == 0) return GEL_OK;
    std::vector<double> out((size_t)n * 10);
    const int st = batch_run(n, { { L, 5 }, { R, 5 }, { nrm, 3 }, { W, 1 }, { dist, 1 } }, { { out.data(), 10 } },
                             [&](std::vector<double*>& i, std::vector<double*>& d) {
                                 if (wave)
                                     hipLaunchKernelGGL(godunov_wave_kernel, dim3((n + 3) / 4), dim3(256), 0, 0, n,
                                                        i[0], i[1], i[2], i[3], i[4], cfl, gamma, d[0]);
                                 else
                                     hipLaunchKernelGGL(godunov_batch_kernel, dim3((n + 255) / 256), dim3(256), 0, 0,
                                                        n, i[0], i[1], i[2], i[3], i[4], cfl, gamma, d[0]);
                             });
    if (st != GEL_OK) return st;
    for (int k = 0; k < n; k++) {
        const double* o = &out[(size_t)k * 10];
        for (int i = 0; i < 5; i++) q[(size_t)k * 5 + i] = o[i];
        for (int i = 0; i < 3; i++) speeds[(size_t)k * 3 + i] = o[5 + i];
        dt[k] = o[8];
        status[k] = (int)o[9];
    }
    return GEL_OK;
}

int gel_flux_godunov_batch(int n, const double* L, const double* R, const double* nrm, const double* W,
                           const double* dist, double cfl, double gamma, double* q, double* speeds, double* dt,
                           int* status)
{
    return godunov_batch(false, n, L, R, nrm, W, dist, cfl, gamma, q, speeds, dt, status);
}

int gel_flux_godunov_wave(int n, const double* L, const double* R, const double* nrm, const double* W,
                          const double* dist, double cfl, double gamma, double* q, double* speeds, double* dt,
                          int* status)
{
    return godunov_batch(true, n, L, R, nrm, W, dist, cfl, gamma, q, speeds, dt, status);
}

int gel_fp64_div_sqrt(int n, const double* a, const double* b, double* q, double* q2, double* s)
{
    if (n < 0 || (n > 0 && (!a || !b || !q || !q2 || !s))) return fail(GEL_E_ARG, "null argument");
    if (n == 0) return GEL_OK;
    return batch_run(n, { { a, 1 }, { b, 1 } }, { { q, 1 }, { q2, 1 }, { s, 1 } },
                     [&](std::vector<double*>& i, std::vector<double*>& d) {
                         hipLaunchKernelGGL(fp64_ops_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, n, i[0], i[1],
                                            d[0], d[1], d[2]);
                     });
}

int gel_flux_hlld_batch(int n, const double* L, const double* R, const double* W, const double* n1,
                        const double* n2, double gamma, double* P, double* smax)
{
    if (n < 0 || (n > 0 && (!L || !R || !W || !n1 || !n2 || !P || !smax))) return fail(GEL_E_ARG, "null argument");
    if (n == 0) return GEL_OK;
    return batch_run(n, { { L, 9 }, { R, 9 }, { W, 1 }, { n1, 1 }, { n2, 1 } }, { { P, 9 }, { smax, 1 } },
                     [&](std::vector<double*>& i, std::vector<double*>& d) {
                         hipLaunchKernelGGL(hlld_batch_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, n, i[0], i[1],
                                            i[2], i[3], i[4], gamma, d[0], d[1]);
                     });
}

}  // extern "C"

// ------------------------------------------------- the no-allocation flux entry ---
// gel_flux_hlld_batch_noalloc stages its host arrays in a file-scope device array
// instead of device memory it allocates: [F][9] L, R, P, then [F] W, n1, n2, smax,
// F = GEL_NOALLOC_MAX_FACES (31 doubles a face, 8 MB).  Each device holds its own
// copy from the library's load until the process ends; calls on one device share it.
static constexpr size_t kNoAllocFaces = GEL_NOALLOC_MAX_FACES;
static __device__ __align__(256) double scratch[kNoAllocFaces * 31];

static constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kScratchDevices] = {};

// Device address of `scratch` on the current device, looked up once per device with
// that device current (the array's name in host code is not a device address).
static int scratch_on_device(double** out)
{
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices) return fail(GEL_E_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GEL_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return GEL_OK;
}

extern "C" {

int gel_flux_hlld_batch_noalloc(int n, const double* L, const double* R, const double* W, const double* n1,
                                const double* n2, double gamma, double* P, double* smax)
{
    if (n < 0 || (n > 0 && (!L || !R || !W || !n1 || !n2 || !P || !smax))) return fail(GEL_E_ARG, "null argument");
    if ((size_t)n > kNoAllocFaces)
        return fail(GEL_E_ARG, "n exceeds GEL_NOALLOC_MAX_FACES (" + std::to_string(kNoAllocFaces) + ") faces");
    if (n == 0) return GEL_OK;
    double* s = nullptr;
    const int st = scratch_on_device(&s);
    if (st != GEL_OK) return st;
    const size_t F = kNoAllocFaces, b = sizeof(double);
    double *dL = s, *dR = s + 9 * F, *dP = s + 18 * F, *dW = s + 27 * F, *dn1 = s + 28 * F, *dn2 = s + 29 * F,
           *du = s + 30 * F;
    HIP_OK(hipMemcpy(dL, L, 9 * (size_t)n * b, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dR, R, 9 * (size_t)n * b, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dW, W, (size_t)n * b, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dn1, n1, (size_t)n * b, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dn2, n2, (size_t)n * b, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(hlld_batch_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, n, dL, dR, dW, dn1, dn2, gamma, dP,
                       du);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(P, dP, 9 * (size_t)n * b, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(smax, du, (size_t)n * b, hipMemcpyDeviceToHost));
    return GEL_OK;
}

}  // extern "C"

// ------------------------------------------------------------ batched cases ---

struct gel_batch {
    std::vector<gel_ctx*> cases;
    int rows = 0;                 // rows per chunk for every case; 0 = per-case default
    int workgroups = 0;           // persistent grid; 0 = the resident work-groups of the kernel
    SweepArgs* d_args = nullptr;  // [cycle + 1][ncases]: one argument set per phase of the cycle
    int cycle = 6;                // phases: two state buffers x three time-step slots (x the
                                  // motion period of a moving batch)
    bool moving = false;          // every case moves its nodes by its prescribed breathing
    bool types = false;           // every case has C_1 / C_centr cells (gas, FP64)
    int nfl = 1;                  // sweep launches per step: 5 for the plasma + four hydrogen
                                  // fluids (one argument block per fluid and phase)
    // fitted-surface tracking: every case's tracking tables, and the (case, block) work
    // lists of the batched surface and rail launches
    bool tracked = false;
    TrackDev* d_tdev = nullptr;
    int2* d_stiles = nullptr;
    int2* d_rtiles = nullptr;
    int nstiles = 0, nrtiles = 0;
    BatchItem* d_items = nullptr;
    int nitems = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t exec = nullptr;  // one cycle of batch_cycle() steps
    bool graph_valid = false;
    int graph_off = 0;              // refresh offset (steps since the last refresh) it was captured at
    hipGraphExec_t tail = nullptr;  // a call's steps after its whole cycles
    int tail_len = 0, tail_off = 0;
    long long gstep = 0;            // steps the batch has advanced its cases
    long long src_base = 0;         // gstep when the source model was attached
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // learned source refreshed inside gel_batch_step every `every` steps
    gel_mlp* model = nullptr;
    int every = 0;
    double tol = 1e-3;
    int judge = 1;
    long long refresh_tiles = 0;
    int2* d_tiles = nullptr;                 // refresh work-group -> (case, tile of the case)
    unsigned long long* d_nfall = nullptr;   // [1] fallback cells since the last stats call
    // judge 2: the (case, cell) pairs still on the network, [0] their number, [1] the
    // counter of the next build (0 between builds)
    int2* d_list = nullptr;
    unsigned* d_lcount = nullptr;
    long long refreshes = 0;
};

static constexpr int kCycle = 6;  // two state buffers x three time-step slots
static int lcm_int(int a, int b)
{
    int x = a, y = b;
    while (y) { const int t = x % y; x = y; y = t; }
    return a / x * b;
}

static constexpr int kBatchRows = 4 * kNW - 1;   // rows per chunk until the autotuner picks
static int batch_rows(const gel_batch* b, const gel_ctx*) { return b->rows > 0 ? b->rows : kBatchRows; }

static void batch_drop_graph(gel_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->graph_valid = false;
    if (b->tail) (void)hipGraphExecDestroy(b->tail);
    b->tail = nullptr;
    b->tail_len = 0;
}

static int batch_build_items(gel_batch* b)
{
    std::vector<BatchItem> items;
    for (int cs = 0; cs < (int)b->cases.size(); cs++) {
        const gel_ctx* c = b->cases[cs];
        const int rows = batch_rows(b, c);
        const int strips = (c->nj + kOut - 1) / kOut, chunks = (c->ni + rows - 1) / rows;
        for (int by = 0; by < chunks; by++)
            for (int bx = 0; bx < strips; bx++) items.push_back(BatchItem{ cs, bx, by });
    }
    (void)hipFree(b->d_items);
    b->d_items = nullptr;
    if (hipMalloc(&b->d_items, items.size() * sizeof(BatchItem)) != hipSuccess) return fail(GEL_E_NOMEM, "batch items");
    HIP_OK(hipMemcpy(b->d_items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice));
    b->nitems = (int)items.size();
    batch_drop_graph(b);
    return GEL_OK;
}

// argument sets of the next b->cycle steps, from each case's own buffer/slot phase
// (and motion phase), and (block b->cycle) the cases as they are now, for a refresh
// before any step
static int batch_upload_args(gel_batch* b)
{
    const size_t n = b->cases.size();
    const int cyc = b->cycle, nfl = b->nfl;
    std::vector<SweepArgs> args((size_t)(cyc + 1) * nfl * n);
    for (int pf = 0; pf < (cyc + 1) * nfl; pf++)
        for (size_t cs = 0; cs < n; cs++) {
            const int p = pf / nfl, fl = pf % nfl;
            gel_ctx* c = b->cases[cs];
            SweepArgs& a = args[pf * n + cs];
            const long long st = c->steps + p;
            const int cur = c->cur ^ (p & 1);
            a.ni = c->ni; a.nj = c->nj; a.rows = batch_rows(b, c);
            a.dt_rd = (int)(st % 3); a.dt_wr = (int)((st + 1) % 3); a.dt_clr = (int)((st + 2) % 3);
            a.s_old = 1.0; a.s_new = 1.0;
            if (b->moving) {
                const long long P = (long long)c->motion.size();
                a.s_old = c->motion[(size_t)((c->motion_step + p) % P)];
                a.s_new = c->motion[(size_t)((c->motion_step + p + 1) % P)];
            }
            a.x = c->d_x; a.y = c->d_y; a.x2 = c->d_x; a.y2 = c->d_y;
            if (b->tracked) {
                // the rail launch writes the next nodes into the other buffer; the two
                // buffers change roles every step (Point::x = x2, Setka.cpp:10473-10480)
                const bool odd = (p & 1) != 0;
                a.x = odd ? c->d_x2 : c->d_x;  a.y = odd ? c->d_y2 : c->d_y;
                a.x2 = odd ? c->d_x : c->d_x2; a.y2 = odd ? c->d_y : c->d_y2;
                a.srow[0] = c->trk.row_inner_shock; a.srow[1] = c->trk.row_contact; a.srow[2] = c->trk.row_outer_shock;
            }
            a.now = c->d_state[cur];
            a.next = c->d_state[cur ^ 1];
            a.ctype = b->types ? c->d_ctype : nullptr; a.otype = c->d_otype;
            a.src = c->d_src; a.qtab = c->d_qtab; a.qflag = c->d_qflag; a.geo = c->d_geo;
            if (!b->tracked) a.srow[0] = a.srow[1] = a.srow[2] = -1;
            a.fluid = 0;
            for (int v = 0; v < 5; v++) a.voff[v] = v;
            a.mfq = c->d_mfq;
            if (nfl == 5) {
                // the plasma: ro, p, u, v, Q; fluid Hf: ro_Hf, p_Hf, u_Hf, v_Hf and the plasma Q
                a.fluid = fl;
                const int o = fl == 0 ? 0 : 5 + 4 * (fl - 1);
                for (int k = 0; k < 4; k++) a.voff[k] = o + k;
                a.voff[4] = 4;
            }
            a.dtbuf = c->d_dt; a.counters = c->d_cnt;
            a.prm = c->prm;
            if (p == cyc) a.next = c->d_state[c->cur];
        }
    HIP_OK(hipMemcpyAsync(b->d_args, args.data(), args.size() * sizeof(SweepArgs), hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    return GEL_OK;
}

template <int FX>
static hipError_t batch_launch_mf_sweep(const gel_batch* b, const SweepArgs* args)
{
    if (b->moving)
        return batch_launch_t<GasPhys, double, double, GEL_SRC_NONE, FX | kMfFlux, true, true>(args, b->d_items, b->nitems,
                                                                                            b->workgroups, b->stream);
    return batch_launch_t<GasPhys, double, double, GEL_SRC_NONE, FX | kMfFlux, false, true>(args, b->d_items, b->nitems,
                                                                                         b->workgroups, b->stream);
}

// One step of a batch of plasma + hydrogen cases: the charge-exchange sources of every
// case, then the sweep of each of the five fluids over every case.
static hipError_t batch_launch_mf(const gel_batch* b, int phase)
{
    const size_t n = b->cases.size();
    const SweepArgs* args0 = b->d_args + (size_t)phase * 5 * n;
    hipLaunchKernelGGL(mf_source_batch_kernel<true>, dim3((unsigned)b->refresh_tiles), dim3(256), 0, b->stream, args0,
                       b->d_tiles);
    hipError_t e = hipGetLastError();
    for (int f = 0; f < 5 && e == hipSuccess; f++) {
        const SweepArgs* args = args0 + (size_t)f * n;
        e = b->cases[0]->prm.flux == GEL_FLUX_HLLC ? batch_launch_mf_sweep<GEL_FLUX_HLLC>(b, args)
                                                   : batch_launch_mf_sweep<GEL_FLUX_HLL>(b, args);
    }
    return e;
}

static hipError_t batch_launch(const gel_batch* b, int phase)
{
    const gel_params& p = b->cases[0]->prm;
    if (b->nfl == 5) return batch_launch_mf(b, phase);
    const SweepArgs* args = b->d_args + (size_t)phase * b->cases.size();
    if (b->tracked)
        return launch_tracked_batch(args, b->d_tdev, b->d_stiles, b->nstiles, b->d_rtiles, b->nrtiles, b->d_items,
                                    b->nitems, b->workgroups, b->stream, p.flux == GEL_FLUX_HLLC, b->types,
                                    p.source_mode);
    if (p.precision == GEL_PREC_FP32_BULK)
        return launch_batch_fp32(p, args, b->d_items, b->nitems, b->workgroups, b->stream, b->moving);
    if (b->types)
        return p.flux == GEL_FLUX_HLLC
                   ? batch_launch_typed<GEL_FLUX_HLLC>(args, b->d_items, b->nitems, b->workgroups, b->stream,
                                                       p.source_mode, b->moving)
                   : batch_launch_typed<GEL_FLUX_HLL>(args, b->d_items, b->nitems, b->workgroups, b->stream,
                                                      p.source_mode, b->moving);
    if (p.physics == GEL_PHYS_MHD)
        return batch_launch_s<MhdPhys, double, double, 3>(args, b->d_items, b->nitems, b->workgroups, b->stream,
                                                          p.source_mode, b->moving);
    if (p.flux == GEL_FLUX_HLLC)
        return batch_launch_s<GasPhys, double, double, GEL_FLUX_HLLC>(args, b->d_items, b->nitems, b->workgroups,
                                                                      b->stream, p.source_mode, b->moving);
    return batch_launch_s<GasPhys, double, double, GEL_FLUX_HLL>(args, b->d_items, b->nitems, b->workgroups,
                                                                 b->stream, p.source_mode, b->moving);
}

// steps per captured graph: the time-step / buffer cycle, times what makes the
// refresh schedule repeat
static int batch_cycle(const gel_batch* b)
{
    if (!b->model || b->every <= 0) return b->cycle;
    return lcm_int(b->cycle, b->every);
}

static long long batch_cells(const gel_batch* b)
{
    long long n = 0;
    for (const gel_ctx* c : b->cases) n += (long long)c->nc;
    return n;
}

// Steps since the last refresh of the batch's schedule: a refresh follows every
// `every`-th step the batch advances after the model was attached, whatever the
// lengths of the calls.
static int batch_refresh_offset(const gel_batch* b)
{
    return (b->model && b->every > 0) ? (int)((b->gstep - b->src_base) % b->every) : 0;
}

// step i of a call that starts `off` steps after a refresh: the sweep (argument block
// i mod 6), then the source refresh when one is due
static hipError_t batch_launch_step(const gel_batch* b, int i, int off)
{
    const int phase = i % b->cycle;
    hipError_t e = batch_launch(b, phase);
    if (e == hipSuccess && b->model && b->every > 0 && (off + i + 1) % b->every == 0) {
        const SweepArgs* args = b->d_args + (size_t)phase * b->nfl * b->cases.size();
        const bool f32 = b->cases[0]->prm.precision == GEL_PREC_FP32_BULK;
        if (b->judge == 2)
            e = gel_refresh_list_launch(b->stream, args, b->d_tiles, b->refresh_tiles, batch_cells(b), b->d_list,
                                        b->d_lcount, b->model, f32, b->tol, b->d_nfall);
        else
            e = gel_refresh_launch(b->stream, args, b->d_tiles, b->refresh_tiles, b->model, f32, b->tol, b->judge,
                                   b->d_nfall);
    }
    return e;
}

// Capture steps [0, n) of a call starting `off` steps after a refresh as a graph.
static int batch_capture(gel_batch* b, int n, int off, hipGraphExec_t* out)
{
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeGlobal));
    hipError_t e = hipSuccess;
    for (int p = 0; p < n && e == hipSuccess; p++) e = batch_launch_step(b, p, off);
    const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(GEL_E_HIP, "graph capture");
    }
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(GEL_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    return GEL_OK;
}

// The graphs a call of nsteps steps replays: whole cycles (a cycle starts and ends at
// the same buffer, time-step slot and refresh phase), then one graph of the rest.
static int batch_prepare_graphs(gel_batch* b, int nsteps)
{
    const int cyc = batch_cycle(b), off = batch_refresh_offset(b);
    if (nsteps >= cyc && !(b->graph_valid && b->graph_off == off)) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        b->graph_valid = false;
        const int st = batch_capture(b, cyc, off, &b->exec);
        if (st != GEL_OK) return st;
        b->graph_valid = true;
        b->graph_off = off;
    }
    const int rem = nsteps % cyc;
    if (rem > 0 && !(b->tail && b->tail_len == rem && b->tail_off == off)) {
        if (b->tail) (void)hipGraphExecDestroy(b->tail);
        b->tail = nullptr;
        b->tail_len = 0;
        const int st = batch_capture(b, rem, off, &b->tail);
        if (st != GEL_OK) return st;
        b->tail_len = rem;
        b->tail_off = off;
    }
    return GEL_OK;
}

// judge 2: (re)build the list of cells still on the network from the current flags,
// on the batch stream, before a call steps or refreshes with it
static int batch_build_list(gel_batch* b)
{
    if (!b->d_list) {
        size_t n = 0;
        for (const gel_ctx* c : b->cases) n += c->nc;
        if (hipMalloc((void**)&b->d_list, n * sizeof(int2)) != hipSuccess ||
            hipMalloc((void**)&b->d_lcount, 2 * sizeof(unsigned)) != hipSuccess ||
            hipMemsetAsync(b->d_lcount, 0, 2 * sizeof(unsigned), b->stream) != hipSuccess)
            return fail(GEL_E_NOMEM, "list of the cells on the network");
    }
    const hipError_t e = gel_live_list_build(b->stream, b->d_args, b->d_tiles, b->refresh_tiles, batch_cells(b),
                                             b->d_list, b->d_lcount);
    return e == hipSuccess ? GEL_OK : fail(GEL_E_HIP, std::string("list build: ") + hipGetErrorString(e));
}

static void batch_advance(gel_batch* b, int n)
{
    b->gstep += n;
    for (gel_ctx* c : b->cases) {
        c->steps += n;
        if (n & 1) c->cur ^= 1;
        if (b->moving) c->motion_step += n;
        if (b->tracked && (n & 1)) {
            std::swap(c->d_x, c->d_x2);
            std::swap(c->d_y, c->d_y2);
            c->geo_valid = false;
        }
    }
}

extern "C" {

void gel_batch_destroy(gel_batch* b)
{
    if (!b) return;
    batch_drop_graph(b);
    (void)hipFree(b->d_args);
    (void)hipFree(b->d_items);
    (void)hipFree(b->d_nfall);
    (void)hipFree(b->d_tiles);
    (void)hipFree(b->d_list);
    (void)hipFree(b->d_lcount);
    (void)hipFree(b->d_tdev);
    (void)hipFree(b->d_stiles);
    (void)hipFree(b->d_rtiles);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int gel_batch_create(gel_batch** out, int ncases, gel_ctx* const* cases)
{
    if (!out || ncases <= 0 || !cases) return fail(GEL_E_ARG, "null argument");
    *out = nullptr;
    for (int k = 0; k < ncases; k++) {
        if (!cases[k]) return fail(GEL_E_ARG, "null case");
        if (cases[k]->mesh) return fail(GEL_E_UNSUPPORTED, "batched cases are polar grids");
        // one context twice would have two sweeps of a step write its one next buffer
        for (int l = 0; l < k; l++)
            if (cases[l] == cases[k]) return fail(GEL_E_ARG, "a context appears twice in the batch");
    }
    const gel_params& p0 = cases[0]->prm;
    const bool moving = !cases[0]->motion.empty();
    const size_t period = cases[0]->motion.size();
    const bool types = cases[0]->has_types;
    for (int k = 0; k < ncases; k++) {
        const gel_ctx* c = cases[k];
        if (!c->uploaded) return fail(GEL_E_STATE, "every case needs an uploaded state");
        if (c->prm.physics == GEL_PHYS_GAS_INNER)
            return fail(GEL_E_UNSUPPORTED, "the inner-region solver is not batched");
        if (c->tracking != cases[0]->tracking)
            return fail(GEL_E_UNSUPPORTED, "batched cases all track their fitted surfaces or none does");
        if (c->tracking && c->prm.source_mode != GEL_SRC_NONE && c->prm.source_mode != GEL_SRC_MK1)
            return fail(GEL_E_UNSUPPORTED, "tracked batches take no source or the analytic closure");
        // grids with C_1 / C_centr cells: all of them, gas in FP64
        if (c->has_types != types)
            return fail(GEL_E_UNSUPPORTED, "batched cases all have C_1 / C_centr cells or none has");
        if (types && c->prm.physics != GEL_PHYS_GAS4H &&
            (c->prm.physics != GEL_PHYS_GAS || c->prm.precision != GEL_PREC_FP64 ||
             (c->prm.source_mode != GEL_SRC_NONE && c->prm.source_mode != GEL_SRC_MK1)))
            return fail(GEL_E_UNSUPPORTED, "batched cases with C_1 / C_centr cells: gas, FP64, no source or MK1");
        // a moving batch: every case breathes with the same period (each with its own
        // amplitude and phase), so that one cycle of argument sets repeats
        if (c->motion.empty() == moving || c->motion.size() != period)
            return fail(GEL_E_UNSUPPORTED, "batched cases are all static or all moving with one motion period");
        if (c->prm.physics != p0.physics || c->prm.precision != p0.precision || c->prm.flux != p0.flux ||
            c->prm.source_mode != p0.source_mode)
            return fail(GEL_E_UNSUPPORTED, "batched cases share physics, precision, flux and source mode");
        if ((c->prm.source_mode == GEL_SRC_MK1 || c->prm.source_mode == GEL_SRC_HYBRID) && !c->d_src)
            return fail(GEL_E_STATE, "sources not set");
        if ((c->prm.source_mode == GEL_SRC_TABLE || c->prm.source_mode == GEL_SRC_HYBRID) && !c->d_qtab)
            return fail(GEL_E_STATE, "source table not set");
        if (moving && (c->prm.source_mode == GEL_SRC_TABLE || c->prm.source_mode == GEL_SRC_HYBRID))
            return fail(GEL_E_UNSUPPORTED, "moving batches take no source or the analytic closure");
    }
    gel_batch* b = new (std::nothrow) gel_batch();
    if (!b) return fail(GEL_E_NOMEM, "host");
    b->cases.assign(cases, cases + ncases);
    b->moving = moving;
    b->types = types;
    b->nfl = p0.physics == GEL_PHYS_GAS4H ? 5 : 1;
    b->tracked = cases[0]->tracking;
    if (b->tracked) {
        std::vector<TrackDev> td(ncases);
        std::vector<int2> st, rt;
        for (int k = 0; k < ncases; k++) {
            const gel_ctx* c = cases[k];
            td[k] = c->tdev;
            for (int q = 0; q < (3 * c->nj + 3) / 4; q++) st.push_back(make_int2(k, q));   // four faces each
            const long long nn = (long long)(c->ni + 1) * (c->nj + 1);
            for (long long q = 0; q < (nn + 255) / 256; q++) rt.push_back(make_int2(k, (int)q));
        }
        b->nstiles = (int)st.size();
        b->nrtiles = (int)rt.size();
        if (hipMalloc(&b->d_tdev, td.size() * sizeof(TrackDev)) != hipSuccess ||
            hipMalloc(&b->d_stiles, st.size() * sizeof(int2)) != hipSuccess ||
            hipMalloc(&b->d_rtiles, rt.size() * sizeof(int2)) != hipSuccess ||
            hipMemcpy(b->d_tdev, td.data(), td.size() * sizeof(TrackDev), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->d_stiles, st.data(), st.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->d_rtiles, rt.data(), rt.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess) {
            gel_batch_destroy(b);
            return fail(GEL_E_NOMEM, "batch tracking tables");
        }
    }
    for (int k = 0; k < ncases && b->nfl == 5; k++) {
        gel_ctx* c = cases[k];
        if (!c->d_mfq && hipMalloc((void**)&c->d_mfq, 20 * c->nc * sizeof(double)) != hipSuccess) {
            gel_batch_destroy(b);
            return fail(GEL_E_NOMEM, "plasma + hydrogen sources");
        }
    }
    b->cycle = moving ? lcm_int(kCycle, (int)period) : kCycle;
    for (int k = 0; k < ncases && !moving && !b->tracked; k++) {
        const int st = ensure_geometry(cases[k]);
        if (st != GEL_OK) { gel_batch_destroy(b); return st; }
    }
    std::vector<int2> tiles;
    for (int k = 0; k < ncases; k++)
        for (long long t = 0; t < gel_refresh_tiles(cases[k]->ni, cases[k]->nj); t++) tiles.push_back(make_int2(k, (int)t));
    b->refresh_tiles = (long long)tiles.size();
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
        hipMalloc(&b->d_args, (size_t)(b->cycle + 1) * b->nfl * ncases * sizeof(SweepArgs)) != hipSuccess ||
        hipMalloc(&b->d_nfall, sizeof(unsigned long long)) != hipSuccess ||
        hipMemsetAsync(b->d_nfall, 0, sizeof(unsigned long long), b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        gel_batch_destroy(b);
        return fail(GEL_E_NOMEM, "batch");
    }
    if (hipMalloc(&b->d_tiles, tiles.size() * sizeof(int2)) != hipSuccess ||
        hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess) {
        gel_batch_destroy(b);
        return fail(GEL_E_NOMEM, "batch refresh tiles");
    }
    const int st = batch_build_items(b);
    if (st != GEL_OK) { gel_batch_destroy(b); return st; }
    *out = b;
    return GEL_OK;
}

int gel_batch_set_launch(gel_batch* b, int rows, int workgroups)
{
    if (!b || rows < 0 || rows > 4096 || workgroups < 0) return fail(GEL_E_ARG, "launch shape");
    b->rows = rows;
    b->workgroups = workgroups;
    return batch_build_items(b);
}

// Advance every case nsteps steps.  use_graph: whole cycles of batch_cycle() steps
// replay one captured HIP graph, the remaining steps another (both kept for the next
// call of the same length and refresh phase); otherwise plain launches.  A refresh
// follows every `every`-th step of the batch, counted over all calls.
// *ms (may be NULL) receives the device time of the call (HIP events on the batch stream).
int gel_batch_step(gel_batch* b, int nsteps, int use_graph, double* ms)
{
    if (!b || nsteps < 0) return fail(GEL_E_ARG, "bad argument");
    for (gel_ctx* c : b->cases) HIP_OK(hipStreamSynchronize(c->stream));
    int st = batch_upload_args(b);
    if (st != GEL_OK) return st;
    if (b->model && b->every > 0 && b->judge == 2 && (st = batch_build_list(b)) != GEL_OK) return st;
    const int cyc = batch_cycle(b), off = batch_refresh_offset(b);
    if (use_graph && (st = batch_prepare_graphs(b, nsteps)) != GEL_OK) return st;
    HIP_OK(hipEventRecord(b->ev0, b->stream));
    int done = 0;
    if (use_graph) {
        while (nsteps - done >= cyc) {
            HIP_OK(hipGraphLaunch(b->exec, b->stream));
            done += cyc;
        }
        if (done < nsteps) {
            HIP_OK(hipGraphLaunch(b->tail, b->stream));
            done = nsteps;
        }
    }
    for (; done < nsteps; done++) {
        const hipError_t e = batch_launch_step(b, done, off);
        if (e != hipSuccess) return fail(GEL_E_HIP, std::string("batch launch: ") + hipGetErrorString(e));
    }
    if (b->model && b->every > 0) b->refreshes += (off + nsteps) / b->every;
    HIP_OK(hipEventRecord(b->ev1, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    if (ms) {
        float t = 0.f;
        HIP_OK(hipEventElapsedTime(&t, b->ev0, b->ev1));
        *ms = t;
    }
    batch_advance(b, nsteps);
    return GEL_OK;
}

int gel_batch_prepare(gel_batch* b, int nsteps)
{
    if (!b || nsteps < 0) return fail(GEL_E_ARG, "bad argument");
    for (gel_ctx* c : b->cases) HIP_OK(hipStreamSynchronize(c->stream));
    return batch_prepare_graphs(b, nsteps);
}

int gel_batch_set_source_model(gel_batch* b, gel_mlp* m, int every, double tol, int judge)
{
    if (!b || every < 0 || (every > 0 && !m) || !(tol >= 0.0)) return fail(GEL_E_ARG, "bad argument");
    if (every > 0)
        for (const gel_ctx* c : b->cases)
            if ((c->prm.source_mode != GEL_SRC_TABLE && c->prm.source_mode != GEL_SRC_HYBRID) || !c->d_src)
                return fail(GEL_E_STATE, "a learned source needs GEL_SRC_TABLE or GEL_SRC_HYBRID cases with "
                                         "hydrogen moments set");
    if (judge < 0 || judge > 2) return fail(GEL_E_ARG, "judge is 0, 1 or 2");
    if (judge == 2)
        for (const gel_ctx* c : b->cases)
            if (c->prm.source_mode != GEL_SRC_HYBRID) return fail(GEL_E_UNSUPPORTED, "judge 2 needs GEL_SRC_HYBRID cases");
    b->model = every > 0 ? m : nullptr;
    b->every = every;
    b->src_base = b->gstep;
    b->tol = tol;
    b->judge = judge;
    batch_drop_graph(b);
    return GEL_OK;
}

int gel_batch_refresh(gel_batch* b, gel_mlp* m, double tol, int judge, long long* n_fallback)
{
    if (!b || !m || !(tol >= 0.0) || judge < 0 || judge > 2) return fail(GEL_E_ARG, "bad argument");
    if (b->moving || b->tracked) return fail(GEL_E_UNSUPPORTED, "moving batches take no learned source");
    for (const gel_ctx* c : b->cases) {
        if (judge == 2 && c->prm.source_mode != GEL_SRC_HYBRID)
            return fail(GEL_E_UNSUPPORTED, "judge 2 needs GEL_SRC_HYBRID cases");
        if (!c->d_src) return fail(GEL_E_STATE, "hydrogen moments not set");
        if (!c->d_qtab) return fail(GEL_E_STATE, "source table not allocated (gel_set_source_table)");
    }
    for (gel_ctx* c : b->cases) HIP_OK(hipStreamSynchronize(c->stream));
    int st = batch_upload_args(b);
    if (st != GEL_OK) return st;
    if (judge == 2 && (st = batch_build_list(b)) != GEL_OK) return st;
    unsigned long long* cnt = nullptr;
    if (hipMalloc(&cnt, sizeof *cnt) != hipSuccess) return fail(GEL_E_NOMEM, "counter");
    hipError_t e = hipMemsetAsync(cnt, 0, sizeof *cnt, b->stream);
    const SweepArgs* args = b->d_args + (size_t)b->cycle * b->nfl * b->cases.size();
    const bool f32 = b->cases[0]->prm.precision == GEL_PREC_FP32_BULK;
    if (e == hipSuccess)
        e = judge == 2 ? gel_refresh_list_launch(b->stream, args, b->d_tiles, b->refresh_tiles, batch_cells(b),
                                                 b->d_list, b->d_lcount, m, f32, tol, cnt)
                       : gel_refresh_launch(b->stream, args, b->d_tiles, b->refresh_tiles, m, f32, tol, judge, cnt);
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, cnt, sizeof h, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(cnt);
    if (e != hipSuccess) return fail(GEL_E_HIP, std::string("batch refresh: ") + hipGetErrorString(e));
    if (n_fallback) *n_fallback = (long long)h;
    return GEL_OK;
}

int gel_batch_source_stats(gel_batch* b, long long* refreshes, long long* fallback_cells)
{
    if (!b) return fail(GEL_E_ARG, "null batch");
    unsigned long long h = 0;
    HIP_OK(hipStreamSynchronize(b->stream));
    HIP_OK(hipMemcpy(&h, b->d_nfall, sizeof h, hipMemcpyDeviceToHost));
    HIP_OK(hipMemsetAsync(b->d_nfall, 0, sizeof h, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    if (refreshes) *refreshes = b->refreshes;
    if (fallback_cells) *fallback_cells = (long long)h;
    b->refreshes = 0;
    return GEL_OK;
}

// Project autotuner: time trial_steps steps for each candidate (rows per chunk,
// persistent workgroups) on the batch's own cases, restoring every case's state,
// time-step slots and counters after each trial; keep the fastest.
int gel_batch_autotune(gel_batch* b, int trial_steps, int* best_rows, int* best_wg, double* best_ms)
{
    if (!b || trial_steps < 1) return fail(GEL_E_ARG, "bad argument");
    struct Saved {
        void* st[2]; unsigned long long dt[kDtWords]; int cur; long long steps, motion_step;
        double *x, *y, *x2, *y2;      // tracked cases: the node buffers and (nx, ny) their positions
        void *nx, *ny;
    };
    std::vector<Saved> saved(b->cases.size());
    for (size_t k = 0; k < b->cases.size(); k++) {
        gel_ctx* c = b->cases[k];
        HIP_OK(hipStreamSynchronize(c->stream));
        const size_t bytes = (size_t)c->nv * c->nc * c->ts;
        for (int q = 0; q < 2; q++) {
            if (hipMalloc(&saved[k].st[q], bytes) != hipSuccess) return fail(GEL_E_NOMEM, "autotune snapshot");
            HIP_OK(hipMemcpy(saved[k].st[q], c->d_state[q], bytes, hipMemcpyDeviceToDevice));
        }
        HIP_OK(hipMemcpy(saved[k].dt, c->d_dt, sizeof saved[k].dt, hipMemcpyDeviceToHost));
        saved[k].cur = c->cur;
        saved[k].steps = c->steps;
        saved[k].motion_step = c->motion_step;
        saved[k].x = c->d_x; saved[k].y = c->d_y; saved[k].x2 = c->d_x2; saved[k].y2 = c->d_y2;
        saved[k].nx = saved[k].ny = nullptr;
        if (b->tracked) {
            const size_t nb = c->nn * sizeof(double);
            if (hipMalloc(&saved[k].nx, nb) != hipSuccess || hipMalloc(&saved[k].ny, nb) != hipSuccess)
                return fail(GEL_E_NOMEM, "autotune snapshot");
            HIP_OK(hipMemcpy(saved[k].nx, c->d_x, nb, hipMemcpyDeviceToDevice));
            HIP_OK(hipMemcpy(saved[k].ny, c->d_y, nb, hipMemcpyDeviceToDevice));
        }
    }
    auto restore = [&]() -> int {
        for (size_t k = 0; k < b->cases.size(); k++) {
            gel_ctx* c = b->cases[k];
            const size_t bytes = (size_t)c->nv * c->nc * c->ts;
            for (int q = 0; q < 2; q++) HIP_OK(hipMemcpy(c->d_state[q], saved[k].st[q], bytes, hipMemcpyDeviceToDevice));
            HIP_OK(hipMemcpy(c->d_dt, saved[k].dt, sizeof saved[k].dt, hipMemcpyHostToDevice));
            c->cur = saved[k].cur;
            c->steps = saved[k].steps;
            c->motion_step = saved[k].motion_step;
            if (b->tracked) {
                c->d_x = saved[k].x; c->d_y = saved[k].y; c->d_x2 = saved[k].x2; c->d_y2 = saved[k].y2;
                const size_t nb = c->nn * sizeof(double);
                HIP_OK(hipMemcpy(c->d_x, saved[k].nx, nb, hipMemcpyDeviceToDevice));
                HIP_OK(hipMemcpy(c->d_y, saved[k].ny, nb, hipMemcpyDeviceToDevice));
            }
        }
        return GEL_OK;
    };
    // chunks of k * kNW - 1 rows fill the marching steps; grids of the resident
    // work-groups and of twice as many (the hardware then balances the tail)
    const int rows_c[] = { kNW - 1, 2 * kNW - 1, 3 * kNW - 1, 4 * kNW - 1, 6 * kNW - 1, 8 * kNW - 1, 12 * kNW - 1,
                           16 * kNW - 1 };
    b->workgroups = 0;
    int resident = 0;
    {
        const gel_params& p = b->cases[0]->prm;
        if (p.precision == GEL_PREC_FP32_BULK) {
            resident = resident_batch_fp32(p, b->moving);
        } else if (b->tracked) {
            resident = resident_tracked_batch(b->types);
        } else if (b->nfl == 5) {
            const void* kern =
                b->moving ? (const void*)sweep_batch_kernel<GasPhys, double, double, GEL_SRC_NONE, GEL_FLUX_HLLC | kMfFlux, kNW, true, true>
                          : (const void*)sweep_batch_kernel<GasPhys, double, double, GEL_SRC_NONE, GEL_FLUX_HLLC | kMfFlux, kNW, false, true>;
            resident = resident_groups(kern, kNW * kWave);
        } else if (b->types) {
            const void* kern =
                b->moving ? (const void*)sweep_batch_kernel<GasPhys, double, double, GEL_SRC_NONE, GEL_FLUX_HLLC, kNW, true, true>
                          : (const void*)sweep_batch_kernel<GasPhys, double, double, GEL_SRC_NONE, GEL_FLUX_HLLC, kNW, false, true>;
            resident = resident_groups(kern, kNW * kWave);
        } else if (b->moving) {
            const void* kern =
                p.physics == GEL_PHYS_MHD
                    ? (const void*)sweep_batch_kernel<MhdPhys, double, double, GEL_SRC_NONE, 3, kNW, true>
                    : (const void*)sweep_batch_kernel<GasPhys, double, double, GEL_SRC_NONE, GEL_FLUX_HLLC, kNW, true>;
            resident = resident_groups(kern, kNW * kWave);
        } else {
            const void* kern = p.physics == GEL_PHYS_MHD
                                   ? (const void*)sweep_batch_kernel<MhdPhys, double, double, GEL_SRC_TABLE, 3, kNW>
                                   : (const void*)sweep_batch_kernel<GasPhys, double, double, GEL_SRC_TABLE, GEL_FLUX_HLLC, kNW>;
            resident = resident_groups(kern, kNW * kWave);
        }
    }
    const int wg_c[] = { resident, 2 * resident };
    gel_mlp* model = b->model;      // trials run the sweep alone (the refresh does not depend on the shape)
    b->model = nullptr;
    const long long gstep = b->gstep;
    batch_drop_graph(b);
    double best = 1e300;
    int br = 0, bw = 0;
    int st = GEL_OK;
    for (int r : rows_c) {
        for (int w : wg_c) {
            b->rows = r;
            b->workgroups = w;
            st = batch_build_items(b);
            if (st != GEL_OK) break;
            double ms = 0.0;
            st = gel_batch_step(b, trial_steps, 0, &ms);
            if (st != GEL_OK) break;
            st = restore();
            if (st != GEL_OK) break;
            if (ms < best) { best = ms; br = r; bw = w; }
        }
        if (st != GEL_OK) break;
    }
    for (auto& s : saved) {
        (void)hipFree(s.st[0]); (void)hipFree(s.st[1]);
        (void)hipFree(s.nx); (void)hipFree(s.ny);
    }
    b->model = model;
    b->gstep = gstep;
    batch_drop_graph(b);
    if (st != GEL_OK) return st;
    b->rows = br;
    b->workgroups = bw;
    st = batch_build_items(b);
    if (st != GEL_OK) return st;
    if (best_rows) *best_rows = br;
    if (best_wg) *best_wg = bw;
    if (best_ms) *best_ms = best / trial_steps;
    return GEL_OK;
}

}  // extern "C"
