// ref_harness.cpp - drives the compiled reference (SergeyMSU/Gelios_2D sources under
// /root/reference, built by oracle/Makefile into oracle/_ref/) to write golden vectors.
// TEST INFRASTRUCTURE: runs only in the build container, never on the GPU box.
//
// It builds a polar quad grid out of the reference's own Point/Cell/Gran objects in a
// Setka, lets the reference's Setka::TVD_prepare (Setka.cpp:1294-1492) build the face
// stencils, and runs the reference's sweep Setka::Go_5_komponent__MK2
// (Setka.cpp:10121-10482).  Point-wise modes call the flux, geometry, limiter and
// source functions directly.
//
// Binary I/O: little-endian int32 / float64 arrays, laid out as documented per mode.
#include "Setka.h"
#include "Cell.h"
#include "Gran.h"
#include "Point.h"
#include "Help.h"
#include "Solvers.h"
#include "Rail.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <vector>

static FILE* xopen(const char* p, const char* m)
{
    FILE* f = std::fopen(p, m);
    if (!f) { std::fprintf(stderr, "cannot open %s\n", p); std::exit(2); }
    return f;
}
template <class T> static void rd(FILE* f, T* v, size_t n)
{
    if (std::fread(v, sizeof(T), n, f) != n) { std::fprintf(stderr, "short read\n"); std::exit(2); }
}
template <class T> static void wr(FILE* f, const T* v, size_t n) { std::fwrite(v, sizeof(T), n, f); }

// ---- point-wise: HLLC_2d_Korolkov_b_s -------------------------------------------
// in : int32 n; float64 [n][16] = roL QL pL uL vL roR QR pR uR vR W n1 n2 rad metod nul
// out: float64 [n][9] = P0 P1 P2 P3 PQ Vl Vc Vp time
static int mode_hllc(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t n; rd(f, &n, 1);
    std::vector<double> a((size_t)n * 16); rd(f, a.data(), a.size()); std::fclose(f);
    Setka* S = new Setka();
    std::vector<double> o((size_t)n * 9);
    for (int k = 0; k < n; k++) {
        const double* r = &a[(size_t)k * 16];
        double P[4], PQ, Vl, Vc, Vp;
        double t = S->HLLC_2d_Korolkov_b_s(r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10],
                                           P, PQ, r[11], r[12], r[13], (int)r[14], Vl, Vc, Vp, r[15] != 0.0);
        double* q = &o[(size_t)k * 9];
        q[0] = P[0]; q[1] = P[1]; q[2] = P[2]; q[3] = P[3]; q[4] = PQ; q[5] = Vl; q[6] = Vc; q[7] = Vp; q[8] = t;
    }
    f = xopen(out, "wb"); wr(f, o.data(), o.size()); std::fclose(f);
    return 0;
}

// ---- point-wise: Help.cpp helpers --------------------------------------------------
// in : int32 n; float64 [n][7] = x1 t1 x2 t2 x3 t3 y
// out: float64 [n][8] = minmod(x1,t1) linear3 linear2(x1..t2,y) polar_angle(x1,t1)
//                       perenos(x1,t1 -> x2,t2 applied to (x3,t3)) -> u v, sign(x1)
static int mode_helpers(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t n; rd(f, &n, 1);
    std::vector<double> a((size_t)n * 7); rd(f, a.data(), a.size()); std::fclose(f);
    std::vector<double> o((size_t)n * 8);
    for (int k = 0; k < n; k++) {
        const double* r = &a[(size_t)k * 7];
        double* q = &o[(size_t)k * 8];
        q[0] = minmod(r[0], r[1]);
        q[1] = linear(r[0], r[1], r[2], r[3], r[4], r[5], r[6]);
        q[2] = linear(r[0], r[1], r[2], r[3], r[6]);
        q[3] = polar_angle(r[0], r[1]);
        double u = r[4], v = r[5];
        polar_perenos(r[0], r[1], r[2], r[3], u, v);
        q[4] = u; q[5] = v;
        q[6] = sign(r[0]);
        q[7] = 0.0;
    }
    f = xopen(out, "wb"); wr(f, o.data(), o.size()); std::fclose(f);
    return 0;
}

// ---- point-wise: cell geometry and Get_Sourse_MK1 ----------------------------------
// in : int32 n; float64 [n][8+4+19] = quad x0 y0 .. x3 y3, (u v ro p), 19 MK1 fields
// out: float64 [n][7] = center x y, Get_Volume, Get_Volume_rotate(360), q1 q2 q3
static int mode_cell(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t n; rd(f, &n, 1);
    const int W = 8 + 4 + 19;
    std::vector<double> a((size_t)n * W); rd(f, a.data(), a.size()); std::fclose(f);
    std::vector<double> o((size_t)n * 7);
    for (int k = 0; k < n; k++) {
        const double* r = &a[(size_t)k * W];
        Point* P0 = new Point(r[0], r[1]);
        Point* P1 = new Point(r[2], r[3]);
        Point* P2 = new Point(r[4], r[5]);
        Point* P3 = new Point(r[6], r[7]);
        Cell* C = new Cell(P0, P1, P2, P3);
        const double* fld = r + 12;
        for (int h = 0; h < 4; h++) {
            C->par[0].H_n[h] = fld[h];
            C->par[0].H_u[h] = fld[4 + h];
            C->par[0].H_v[h] = fld[8 + h];
            C->par[0].H_T[h] = fld[12 + h];
        }
        C->par[0].k_u = fld[16]; C->par[0].k_v = fld[17]; C->par[0].k_T = fld[18];
        double* q = &o[(size_t)k * 7];
        C->Get_Center(q[0], q[1]);
        q[2] = C->Get_Volume();
        q[3] = C->Get_Volume_rotate(360.0);
        C->Get_Sourse_MK1(q[4], q[5], q[6], r[8], r[9], r[10], r[11]);
    }
    f = xopen(out, "wb"); wr(f, o.data(), o.size()); std::fclose(f);
    return 0;
}

// ---- point-wise: Solvers::HLLDQ_Korolkov (MHD, W = 0) ------------------------------
// in : int32 n; float64 [n][22] = roL QL pL v1L v2L v3L BxL ByL BzL roR QR pR v1R v2R v3R
//                                 BxR ByR BzR n1 n2 n3 metod
// out: float64 [n][10] = P0..P7 PQ time
static int mode_hlld(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t n; rd(f, &n, 1);
    std::vector<double> a((size_t)n * 22); rd(f, a.data(), a.size()); std::fclose(f);
    Solvers S;
    std::vector<double> o((size_t)n * 10);
    for (int k = 0; k < n; k++) {
        const double* r = &a[(size_t)k * 22];
        double P[8] = {0, 0, 0, 0, 0, 0, 0, 0}, PQ = 0.0;
        double t = S.HLLDQ_Korolkov(r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11],
                                    r[12], r[13], r[14], r[15], r[16], r[17], P, PQ, r[18], r[19], r[20], 1.0,
                                    (int)r[21]);
        double* q = &o[(size_t)k * 10];
        for (int i = 0; i < 8; i++) q[i] = P[i];
        q[8] = PQ; q[9] = t;
    }
    f = xopen(out, "wb"); wr(f, o.data(), o.size()); std::fclose(f);
    return 0;
}

// ---- point-wise: Solvers::Godunov_Solver_Alexashov (exact, moving face) ------------
// in : int32 n; float64 [n][15] = roL pL uL vL wL roR pR uR vR wR n1 n2 n3 W DIST
// out: float64 [n][9] = q0..q4 dsl dsp dsc time
static int mode_godunov(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t n; rd(f, &n, 1);
    std::vector<double> a((size_t)n * 15); rd(f, a.data(), a.size()); std::fclose(f);
    Solvers S;
    std::vector<double> o((size_t)n * 9);
    for (int k = 0; k < n; k++) {
        const double* r = &a[(size_t)k * 15];
        std::vector<double> q1(r, r + 5), q2(r + 5, r + 10), nn(r + 10, r + 13), q(5);
        double dsl = 0.0, dsp = 0.0, dsc = 0.0;
        const double t = S.Godunov_Solver_Alexashov(q1, q2, nn, q, dsl, dsp, dsc, r[13], r[14]);
        double* z = &o[(size_t)k * 9];
        for (int i = 0; i < 5; i++) z[i] = q[i];
        z[5] = dsl; z[6] = dsp; z[7] = dsc; z[8] = t;
    }
    f = xopen(out, "wb"); wr(f, o.data(), o.size()); std::fclose(f);
    return 0;
}

// ---- whole sweep -----------------------------------------------------------------
// The polar grid as the reference's own objects: nodes (i, j) at i*(nj+1)+j, cells
// (i, j) with contour (i,j),(i+1,j),(i+1,j+1),(i,j+1), the state in both buffers.
// own1[c]: the radial face above cell c (main face, master c), own2[c]: the polar face
// towards j+1.
struct PolarRef {
    Setka* S;
    std::vector<Gran*> own1, own2;
    int gnum;
};

static PolarRef build_polar(int ni, int nj, const std::vector<double>& x, const std::vector<double>& y,
                            const std::vector<int32_t>& ctype, const std::vector<int32_t>& otype,
                            const std::vector<double>& prim, const std::vector<double>& src, bool has_src)
{
    const int nn = (ni + 1) * (nj + 1), nc = ni * nj;
    PolarRef P;
    Setka* S = new Setka();
    P.S = S;
    auto nid = [&](int i, int j) { return i * (nj + 1) + j; };
    auto cid = [&](int i, int j) { return i * nj + j; };
    for (int k = 0; k < nn; k++) {
        Point* Pt = new Point(x[k], y[k]);
        Pt->number = k;
        S->All_Points.push_back(Pt);
    }
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++) {
            Cell* C = new Cell(S->All_Points[nid(i, j)], S->All_Points[nid(i + 1, j)],
                               S->All_Points[nid(i + 1, j + 1)], S->All_Points[nid(i, j + 1)]);
            C->number = cid(i, j);
            C->type = (Cell_type)ctype[cid(i, j)];
            S->All_Cells.push_back(C);
        }
    int gnum = 0;
    // Face of edge k of cell (i,j): contour[k] -> contour[k+1] (normal out of the cell).
    // Edges 1 (outer, i+1) and 2 (towards j+1) are owned by the cell; edges 0 and 3 are
    // the copies of the neighbours' faces.  Boundary copies have no master, as the
    // reference builds its axis faces (Setka.cpp:711-741).
    std::vector<Gran*>& own1 = P.own1;
    std::vector<Gran*>& own2 = P.own2;
    own1.assign(nc, nullptr);
    own2.assign(nc, nullptr);
    auto mk_boundary = [&](Cell* K, Point* A, Point* B, Gran_type t) {
        Gran* G = new Gran(A, B, t);
        Gran* G2 = new Gran(B, A, t);
        G->main_gran = true; G2->main_gran = false;
        G->Master = K; G2->Master = nullptr; G2->Sosed = K;
        G->Gran_copy = G2; G2->Gran_copy = G;
        G->number = gnum++; G2->number = gnum++;
        S->All_Gran.push_back(G); S->All_Gran_copy.push_back(G2);
        return G;
    };
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++) {
            Cell* K = S->All_Cells[cid(i, j)];
            Point* c1 = K->contour[1]; Point* c2 = K->contour[2]; Point* c3 = K->contour[3];
            if (i + 1 < ni) {
                Cell* N = S->All_Cells[cid(i + 1, j)];
                Gran* G = new Gran(c1, c2, Usualy);
                Gran* G2 = new Gran(c2, c1, Usualy);
                G->Master = K; G->Sosed = N; G2->Master = N; G2->Sosed = K;
                G->Gran_copy = G2; G2->Gran_copy = G; G2->main_gran = false;
                G->number = gnum++; G2->number = gnum++;
                S->All_Gran.push_back(G); S->All_Gran_copy.push_back(G2);
                own1[cid(i, j)] = G;
            } else {
                own1[cid(i, j)] = mk_boundary(K, c1, c2, (Gran_type)otype[j]);
            }
            if (j + 1 < nj) {
                Cell* N = S->All_Cells[cid(i, j + 1)];
                Gran* G = new Gran(c2, c3, Usualy);
                Gran* G2 = new Gran(c3, c2, Usualy);
                G->Master = K; G->Sosed = N; G2->Master = N; G2->Sosed = K;
                G->Gran_copy = G2; G2->Gran_copy = G; G2->main_gran = false;
                G->number = gnum++; G2->number = gnum++;
                S->All_Gran.push_back(G); S->All_Gran_copy.push_back(G2);
                own2[cid(i, j)] = G;
            } else {
                own2[cid(i, j)] = mk_boundary(K, c2, c3, Axis);
            }
        }
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++) {
            Cell* K = S->All_Cells[cid(i, j)];
            Gran* e0 = (j > 0) ? own2[cid(i, j - 1)]->Gran_copy : mk_boundary(K, K->contour[0], K->contour[1], Axis);
            Gran* e3 = (i > 0) ? own1[cid(i - 1, j)]->Gran_copy : mk_boundary(K, K->contour[3], K->contour[0], Inner_sphere);
            K->Grans.push_back(e0);
            K->Grans.push_back(own1[cid(i, j)]);
            K->Grans.push_back(own2[cid(i, j)]);
            K->Grans.push_back(e3);
        }
    for (int c = 0; c < nc; c++) {
        Cell* K = S->All_Cells[c];
        for (int b = 0; b < 2; b++) {
            Parametr& p = K->par[b];
            p.ro = prim[c]; p.p = prim[nc + c]; p.u = prim[2 * nc + c]; p.v = prim[3 * nc + c]; p.Q = prim[4 * nc + c];
            for (int hh = 0; hh < 4; hh++) {
                p.H_n[hh] = has_src ? src[(size_t)hh * nc + c] : 0.0;
                p.H_u[hh] = has_src ? src[(size_t)(4 + hh) * nc + c] : 0.0;
                p.H_v[hh] = has_src ? src[(size_t)(8 + hh) * nc + c] : 0.0;
                p.H_T[hh] = has_src ? src[(size_t)(12 + hh) * nc + c] : 0.0;
            }
            p.k_u = has_src ? src[(size_t)16 * nc + c] : 0.0;
            p.k_v = has_src ? src[(size_t)17 * nc + c] : 0.0;
            p.k_T = has_src ? src[(size_t)18 * nc + c] : 0.0;
        }
        K->Back = nullptr;
        K->Next = nullptr;
    }
    P.gnum = gnum;
    return P;
}

// Setka::Move_surface (called by every step of Go_5_komponent__MK2 and
// Go_5_komponent_2) indexes the fitted-surface face lists without checking that they
// are non-empty (Setka.cpp:3532, :3734, :3840).  A polar grid without fitted surfaces
// gets one stand-in face per list, between two detached cells far from the grid.  Its
// nodes are not in All_Points and there are no rails, so Move_Setka_Calculate_2 moves
// nothing and the stand-in never touches the grid.
static void add_standin_lines(Setka* S, int& gnum)
{
    {
        S->N1 = S->N2 = S->N3 = S->N4 = S->M1 = S->M2 = S->M3 = S->M4 = 0;
        std::vector<Point*> dp;
        for (int k = 0; k <= 4; k++) { dp.push_back(new Point(1000.0 + k, 1000.0)); dp.push_back(new Point(1000.0 + k, 1001.0)); }
        std::vector<Cell*> dc;
        for (int k = 0; k < 4; k++) {
            Cell* C = new Cell(dp[2 * k], dp[2 * k + 2], dp[2 * k + 3], dp[2 * k + 1]);
            C->type = C_2;
            for (int b = 0; b < 2; b++) { C->par[b].ro = 1.0; C->par[b].p = 1.0; C->par[b].u = 0.0; C->par[b].v = 0.0; C->par[b].Q = 1.0; }
            dc.push_back(C);
        }
        Gran* G = new Gran(dp[4], dp[5], Usualy);
        Gran* G2 = new Gran(dp[5], dp[4], Usualy);
        G->Master = dc[1]; G->Sosed = dc[2]; G->Sosed_down = dc[0]; G->Sosed_up = dc[3];
        G2->Master = dc[2]; G2->Sosed = dc[1]; G2->Sosed_down = dc[3]; G2->Sosed_up = dc[0];
        G->Gran_copy = G2; G2->Gran_copy = G;
        G->number = gnum++; G2->number = gnum++;
        S->Line_Contact.push_back(G);
        S->Line_Inner.push_back(G);
        S->Line_Outer.push_back(G);
    }
}

// in : int32 ni nj steps moving has_src
//      float64 x[nn] y[nn]; if moving: float64 [steps][2][nn] node positions after
//      each step (x2 then y2)
//      int32 cell_type[nc] outer_type[nj]
//      float64 prim[5][nc] (ro p u v Q); if has_src: float64 src[19][nc]
// Static: Go_5_komponent__MK2(steps).  Moving: for each step set Point::x2,y2 and
// call Go_5_komponent__MK2(1), so every step uses the reference's first-step dt.
// out: float64 prim[5][nc]; int32 contour[nc][4]; int32 nface[nc];
//      int32 slots[nc][8][6] = A B type sosed sosed_down sosed_up (-1 when absent)
static int mode_sweep(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t h[5]; rd(f, h, 5);
    const int ni = h[0], nj = h[1], steps = h[2], moving = h[3], has_src = h[4];
    const int nn = (ni + 1) * (nj + 1), nc = ni * nj;
    std::vector<double> x(nn), y(nn);
    rd(f, x.data(), nn); rd(f, y.data(), nn);
    std::vector<double> traj;
    if (moving) { traj.resize((size_t)steps * 2 * nn); rd(f, traj.data(), traj.size()); }
    std::vector<int32_t> ctype(nc), otype(nj);
    rd(f, ctype.data(), nc); rd(f, otype.data(), nj);
    std::vector<double> prim((size_t)5 * nc); rd(f, prim.data(), prim.size());
    std::vector<double> src;
    if (has_src) { src.resize((size_t)19 * nc); rd(f, src.data(), src.size()); }
    std::fclose(f);

    PolarRef PR = build_polar(ni, nj, x, y, ctype, otype, prim, src, has_src != 0);
    Setka* S = PR.S;
    int gnum = PR.gnum;
    S->TVD_prepare();

    add_standin_lines(S, gnum);

    int final_buf;
    if (!moving) {
        // the step loop alone, timed (the CPU baseline's fairness check)
        const auto t0 = std::chrono::steady_clock::now();
        S->Go_5_komponent__MK2(steps);
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::fprintf(stderr, "ref_sweep_seconds %.6f\n", sec);
        final_buf = steps % 2;
    } else {
        for (int s = 0; s < steps; s++) {
            const double* t = &traj[(size_t)s * 2 * nn];
            for (int k = 0; k < nn; k++) { S->All_Points[k]->x2 = t[k]; S->All_Points[k]->y2 = t[nn + k]; }
            S->Go_5_komponent__MK2(1);
            // the single step wrote par[1]; carry it to par[0], the next call's input
            for (auto* K : S->All_Cells) {
                K->par[0].ro = K->par[1].ro; K->par[0].p = K->par[1].p; K->par[0].u = K->par[1].u;
                K->par[0].v = K->par[1].v; K->par[0].Q = K->par[1].Q;
            }
        }
        final_buf = 0;
    }

    std::vector<double> res((size_t)5 * nc);
    for (int c = 0; c < nc; c++) {
        const Parametr& p = S->All_Cells[c]->par[final_buf];
        res[c] = p.ro; res[nc + c] = p.p; res[2 * nc + c] = p.u; res[3 * nc + c] = p.v; res[4 * nc + c] = p.Q;
    }
    std::vector<int32_t> contour((size_t)nc * 4), nface(nc), slots((size_t)nc * 8 * 6, -1);
    for (int c = 0; c < nc; c++) {
        Cell* K = S->All_Cells[c];
        for (int k = 0; k < 4; k++) contour[(size_t)c * 4 + k] = K->contour[k]->number;
        nface[c] = (int32_t)K->Grans.size();
        for (size_t k = 0; k < K->Grans.size(); k++) {
            Gran* G = K->Grans[k];
            int32_t* s = &slots[((size_t)c * 8 + k) * 6];
            s[0] = G->A->number; s[1] = G->B->number; s[2] = (int32_t)G->type;
            s[3] = G->Sosed ? G->Sosed->number : -1;
            s[4] = G->Sosed_down ? G->Sosed_down->number : -1;
            s[5] = G->Sosed_up ? G->Sosed_up->number : -1;
        }
    }
    f = xopen(out, "wb");
    wr(f, res.data(), res.size());
    wr(f, contour.data(), contour.size());
    wr(f, nface.data(), nface.size());
    wr(f, slots.data(), slots.size());
    std::fclose(f);
    return 0;
}

// ---- whole sweep with fitted-surface tracking ----------------------------------------
// in : int32 ni nj steps has_src row_inner_shock row_contact row_outer_shock outer_faces
//      float64 x[nn] y[nn] rail_s[nj+1]
//      int32 cell_type[nc] outer_type[nj]
//      float64 prim[5][nc]; if has_src: float64 src[19][nc]
// The three node rows get the point types P_Inner_shock, P_Contact, P_Outer_shock; the
// radial faces of those rows (masters below) form Line_Inner, Line_Contact, Line_Outer
// in the order of the polar angle (Setka.cpp:1007-1045); every node column is an A rail
// with the four key points (Rail.cpp:19-75).  Setka::N1 is set so that the outer shock's
// face limit n_outer_shock + N1 + N2 + 1 (Setka.cpp:3746) equals outer_faces.  Then
// Go_5_komponent__MK2(steps), which calls Move_surface and Move_Setka_Calculate_2 in
// every step (Setka.cpp:10144-10145).
// out: float64 prim[5][nc]; float64 x[nn] y[nn] (node positions after the last step)
static int mode_track(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t h[8]; rd(f, h, 8);
    const int ni = h[0], nj = h[1], steps = h[2], has_src = h[3];
    const int rin = h[4], rc = h[5], rout = h[6], outer_faces = h[7];
    const int nn = (ni + 1) * (nj + 1), nc = ni * nj;
    std::vector<double> x(nn), y(nn), rail(nj + 1);
    rd(f, x.data(), nn); rd(f, y.data(), nn); rd(f, rail.data(), nj + 1);
    std::vector<int32_t> ctype(nc), otype(nj);
    rd(f, ctype.data(), nc); rd(f, otype.data(), nj);
    std::vector<double> prim((size_t)5 * nc); rd(f, prim.data(), prim.size());
    std::vector<double> src;
    if (has_src) { src.resize((size_t)19 * nc); rd(f, src.data(), src.size()); }
    std::fclose(f);

    PolarRef PR = build_polar(ni, nj, x, y, ctype, otype, prim, src, has_src != 0);
    Setka* S = PR.S;
    auto nid = [&](int i, int j) { return i * (nj + 1) + j; };
    for (int c = 0; c <= nj; c++) {
        S->All_Points[nid(rin, c)]->type = P_Inner_shock;
        S->All_Points[nid(rc, c)]->type = P_Contact;
        S->All_Points[nid(rout, c)]->type = P_Outer_shock;
    }
    S->TVD_prepare();
    // The source interpolation of Cell::Get_Sourse_MK1 (Cell.cpp:528-589) walks the
    // radial Back/Next chains for 10/RR_ < r < 1500/RR_, x >= 0.  TVD_prepare links only
    // neighbours whose centres align with the ray within 1e-3 (Setka.cpp:1438-1489),
    // which the tracked grid's skewed cells can miss; the chains are set to the rays
    // here.  The walk steps up to two cells past a cell whose radius moved, so the end
    // rows link to themselves.  No source is applied in this mode (k_u = k_v = k_T = 0
    // make it zero whatever the interpolated moments).
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++) {
            Cell* K = S->All_Cells[i * nj + j];
            K->Back = S->All_Cells[(i > 0 ? i - 1 : i) * nj + j];
            K->Next = S->All_Cells[(i + 1 < ni ? i + 1 : i) * nj + j];
        }
    for (int j = 0; j < nj; j++) {
        S->Line_Inner.push_back(PR.own1[(rin - 1) * nj + j]);
        S->Line_Contact.push_back(PR.own1[(rc - 1) * nj + j]);
        S->Line_Outer.push_back(PR.own1[(rout - 1) * nj + j]);
    }
    S->N1 = outer_faces - n_outer_shock - 1;
    S->N2 = 0;
    for (int c = 0; c <= nj; c++) {
        Rail* R = new Rail(rail[c]);
        R->type = A;
        for (int i = 0; i <= ni; i++) R->All_point.push_back(S->All_Points[nid(i, c)]);
        R->Key_point.push_back(S->All_Points[nid(rin, c)]);
        R->Key_point.push_back(S->All_Points[nid(rc, c)]);
        R->Key_point.push_back(S->All_Points[nid(rout, c)]);
        R->Key_point.push_back(S->All_Points[nid(ni, c)]);
        R->M1 = rin + 1;
        R->M2 = rc - rin;
        R->M3 = rout - rc;
        R->M4 = ni - rout;
        S->A_Rails.push_back(R);
    }
    S->Go_5_komponent__MK2(steps);
    const int final_buf = steps % 2;

    std::vector<double> res((size_t)5 * nc), xo(nn), yo(nn);
    for (int c = 0; c < nc; c++) {
        const Parametr& p = S->All_Cells[c]->par[final_buf];
        res[c] = p.ro; res[nc + c] = p.p; res[2 * nc + c] = p.u; res[3 * nc + c] = p.v; res[4 * nc + c] = p.Q;
    }
    for (int k = 0; k < nn; k++) { xo[k] = S->All_Points[k]->x; yo[k] = S->All_Points[k]->y; }
    f = xopen(out, "wb");
    wr(f, res.data(), res.size());
    wr(f, xo.data(), xo.size());
    wr(f, yo.data(), yo.size());
    std::fclose(f);
    return 0;
}

// ---- plasma + four hydrogen fluids ---------------------------------------------------
// in : int32 ni nj steps moving
//      float64 x[nn] y[nn]; if moving: float64 [steps][2][nn] node positions after
//      each step (x2 then y2)
//      int32 cell_type[nc] outer_type[nj]
//      float64 prim[21][nc] (ro p u v Q, then ro_Hk p_Hk u_Hk v_Hk for k = 1..4)
// Static: Go_5_komponent_2(steps).  Moving: Point::x2, y2 then Go_5_komponent_2(1) per
// step.  out: float64 prim[21][nc]
static int mode_mf(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t h[4]; rd(f, h, 4);
    const int ni = h[0], nj = h[1], steps = h[2], moving = h[3];
    const int nn = (ni + 1) * (nj + 1), nc = ni * nj;
    std::vector<double> x(nn), y(nn);
    rd(f, x.data(), nn); rd(f, y.data(), nn);
    std::vector<double> traj;
    if (moving) { traj.resize((size_t)steps * 2 * nn); rd(f, traj.data(), traj.size()); }
    std::vector<int32_t> ctype(nc), otype(nj);
    rd(f, ctype.data(), nc); rd(f, otype.data(), nj);
    std::vector<double> prim((size_t)21 * nc); rd(f, prim.data(), prim.size());
    std::fclose(f);

    std::vector<double> none;
    PolarRef PR = build_polar(ni, nj, x, y, ctype, otype, prim, none, false);
    Setka* S = PR.S;
    int gnum = PR.gnum;
    auto set_h = [&](Parametr& p, int c) {
        const double* g = &prim[0];
        p.ro_H1 = g[(size_t)5 * nc + c]; p.p_H1 = g[(size_t)6 * nc + c]; p.u_H1 = g[(size_t)7 * nc + c]; p.v_H1 = g[(size_t)8 * nc + c];
        p.ro_H2 = g[(size_t)9 * nc + c]; p.p_H2 = g[(size_t)10 * nc + c]; p.u_H2 = g[(size_t)11 * nc + c]; p.v_H2 = g[(size_t)12 * nc + c];
        p.ro_H3 = g[(size_t)13 * nc + c]; p.p_H3 = g[(size_t)14 * nc + c]; p.u_H3 = g[(size_t)15 * nc + c]; p.v_H3 = g[(size_t)16 * nc + c];
        p.ro_H4 = g[(size_t)17 * nc + c]; p.p_H4 = g[(size_t)18 * nc + c]; p.u_H4 = g[(size_t)19 * nc + c]; p.v_H4 = g[(size_t)20 * nc + c];
    };
    for (int c = 0; c < nc; c++)
        for (int b = 0; b < 2; b++) set_h(S->All_Cells[c]->par[b], c);
    S->TVD_prepare();
    add_standin_lines(S, gnum);
    int final_buf;
    if (!moving) {
        S->Go_5_komponent_2(steps);
        final_buf = steps % 2;
    } else {
        for (int s = 0; s < steps; s++) {
            const double* t = &traj[(size_t)s * 2 * nn];
            for (int k = 0; k < nn; k++) { S->All_Points[k]->x2 = t[k]; S->All_Points[k]->y2 = t[nn + k]; }
            S->Go_5_komponent_2(1);
            for (auto* K : S->All_Cells) K->par[0] = K->par[1];
        }
        final_buf = 0;
    }
    std::vector<double> res((size_t)21 * nc);
    for (int c = 0; c < nc; c++) {
        const Parametr& p = S->All_Cells[c]->par[final_buf];
        const double v[21] = { p.ro, p.p, p.u, p.v, p.Q, p.ro_H1, p.p_H1, p.u_H1, p.v_H1, p.ro_H2, p.p_H2, p.u_H2,
                               p.v_H2, p.ro_H3, p.p_H3, p.u_H3, p.v_H3, p.ro_H4, p.p_H4, p.u_H4, p.v_H4 };
        for (int k = 0; k < 21; k++) res[(size_t)k * nc + c] = v[k];
    }
    f = xopen(out, "wb");
    wr(f, res.data(), res.size());
    std::fclose(f);
    return 0;
}

// ---- inner-region solver ------------------------------------------------------------
// in : int32 ni nj steps has_src
//      float64 x[nn] y[nn]; int32 cell_type[nc] outer_type[nj]
//      float64 prim[5][nc]; if has_src: float64 src[19][nc]
// Every cell of the polar grid is an inner-region cell (Setka::All_Cells_Inner); then
// Go_stationary_5_komponent_inner_MK2(steps).  Get_Sourse_MK1 walks the radial Back/Next
// chains (Cell.cpp:528-589), set to the rays; on a static grid the walk from a cell's
// own radius returns the cell's own moments.
// out: float64 prim[5][nc]
static int mode_inner(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t h[4]; rd(f, h, 4);
    const int ni = h[0], nj = h[1], steps = h[2], has_src = h[3];
    const int nn = (ni + 1) * (nj + 1), nc = ni * nj;
    std::vector<double> x(nn), y(nn);
    rd(f, x.data(), nn); rd(f, y.data(), nn);
    std::vector<int32_t> ctype(nc), otype(nj);
    rd(f, ctype.data(), nc); rd(f, otype.data(), nj);
    std::vector<double> prim((size_t)5 * nc); rd(f, prim.data(), prim.size());
    std::vector<double> src;
    if (has_src) { src.resize((size_t)19 * nc); rd(f, src.data(), src.size()); }
    std::fclose(f);

    PolarRef PR = build_polar(ni, nj, x, y, ctype, otype, prim, src, has_src != 0);
    Setka* S = PR.S;
    S->TVD_prepare();
    // each ray ends in a detached stand-in cell on both sides, so that the walk from an
    // end row also comes back to that row (Cell.cpp:528-561)
    for (int j = 0; j < nj; j++) {
        Cell* first = S->All_Cells[j];
        Cell* last = S->All_Cells[(ni - 1) * nj + j];
        Cell* lo = new Cell(first->contour[0], first->contour[1], first->contour[2], first->contour[3]);
        Cell* hi = new Cell(last->contour[0], last->contour[1], last->contour[2], last->contour[3]);
        lo->r_istoch = -1.0; lo->Back = nullptr; lo->Next = first;
        hi->r_istoch = 1e300; hi->Back = last; hi->Next = nullptr;
        for (int i = 0; i < ni; i++) {
            Cell* K = S->All_Cells[i * nj + j];
            K->Back = i > 0 ? S->All_Cells[(i - 1) * nj + j] : lo;
            K->Next = i + 1 < ni ? S->All_Cells[(i + 1) * nj + j] : hi;
        }
    }
    for (Cell* K : S->All_Cells) S->All_Cells_Inner.push_back(K);
    S->Go_stationary_5_komponent_inner_MK2(steps);
    const int final_buf = steps % 2;
    std::vector<double> res((size_t)5 * nc);
    for (int c = 0; c < nc; c++) {
        const Parametr& p = S->All_Cells[c]->par[final_buf];
        res[c] = p.ro; res[nc + c] = p.p; res[2 * nc + c] = p.u; res[3 * nc + c] = p.v; res[4 * nc + c] = p.Q;
    }
    f = xopen(out, "wb");
    wr(f, res.data(), res.size());
    std::fclose(f);
    return 0;
}

// ---- checkpoint writer ---------------------------------------------------------------
// in : int32 ni nj has_src; float64 x[nn] y[nn]; int32 cell_type[nc] outer_type[nj];
//      float64 prim[21][nc]; if has_src: float64 src[19][nc]
// Builds the polar grid, records every cell in the lists of its four contour points in
// contour order (as the grid constructor does, Setka.cpp:199-265) and writes the
// reference's checkpoint with Setka::Save_Setka_ALL_ALPHA (Setka.cpp:2355-2574) to out.
static int mode_save(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t h[3]; rd(f, h, 3);
    const int ni = h[0], nj = h[1], has_src = h[2];
    const int nn = (ni + 1) * (nj + 1), nc = ni * nj;
    std::vector<double> x(nn), y(nn);
    rd(f, x.data(), nn); rd(f, y.data(), nn);
    std::vector<int32_t> ctype(nc), otype(nj);
    rd(f, ctype.data(), nc); rd(f, otype.data(), nj);
    std::vector<double> prim((size_t)21 * nc); rd(f, prim.data(), prim.size());
    std::vector<double> src;
    if (has_src) { src.resize((size_t)19 * nc); rd(f, src.data(), src.size()); }
    std::fclose(f);
    PolarRef PR = build_polar(ni, nj, x, y, ctype, otype, prim, src, has_src != 0);
    Setka* S = PR.S;
    for (int c = 0; c < nc; c++) {
        Cell* K = S->All_Cells[c];
        Parametr& p = K->par[0];
        double* hv[16] = { &p.ro_H1, &p.p_H1, &p.u_H1, &p.v_H1, &p.ro_H2, &p.p_H2, &p.u_H2, &p.v_H2,
                           &p.ro_H3, &p.p_H3, &p.u_H3, &p.v_H3, &p.ro_H4, &p.p_H4, &p.u_H4, &p.v_H4 };
        for (int k = 0; k < 16; k++) *hv[k] = prim[(size_t)(5 + k) * nc + c];
        for (int k = 0; k < 4; k++) K->contour[k]->my_cell.push_back(K);
    }
    S->N1 = S->N2 = S->N3 = S->N4 = S->M1 = S->M2 = S->M3 = S->M4 = 0;   // no rails
    S->Save_Setka_ALL_ALPHA(std::string(out));
    return 0;
}

// ---- checkpoint reader ---------------------------------------------------------------
// Setka::Download_Setka_ALL_ALPHA_2_0 (Setka.cpp:2899-3258) of the checkpoint in, then
// Save_Setka_ALL_ALPHA to out: a checkpoint the reference reads back unchanged comes
// out byte for byte.
static int mode_resave(const char* in, const char* out)
{
    Setka* S = new Setka();
    S->Download_Setka_ALL_ALPHA_2_0(std::string(in));
    S->Save_Setka_ALL_ALPHA(std::string(out));
    return 0;
}

// ---- the reference's own block grid -------------------------------------------------
// in : int32 N1 N2 N3 N4 M1 M2 M3 M4 steps
// argv: in, out, checkpoint path (written)
// Setka(N1..M4) builds the reference's heliosphere grid (rails of four kinds, C_centr
// wedges, Setka.cpp:19-1076); Init_conditions fills it (Setka.cpp:5149), the hydrogen
// moments and k factors are zeroed, and Save_Setka_ALL_ALPHA writes the checkpoint.  A
// second Setka reads it back with Download_Setka_ALL_ALPHA_2_0 (so both sides of the
// comparison start from the file's six-digit values), TVD_prepare builds the stencils,
// and Go_5_komponent__MK2(steps) runs.  The fitted-surface lines are replaced by one
// stand-in face each (add_standin_lines), so Move_surface moves no node of the grid;
// Move_Setka_Calculate_2 then places every rail node by its rail's spacing law in the
// first step (moving the nodes from the file's rounded positions) and leaves them there.
// out: int32 nn nc; float64 x[nn] y[nn]; int32 ptype[nn] (0, 1 contact, 2 inner shock,
//      3 outer shock); int32 ncorner[nc]; int32 corners[sum]; int32 cell_type[nc];
//      int32 nface[nc]; int32 slots[nc][8][6]; float64 prim0[5][nc];
//      float64 prim[5][nc] after the steps; float64 x[nn] y[nn] after the steps
static void add_standin_lines_keep(Setka* S)
{
    std::vector<Point*> dp;
    for (int k = 0; k <= 4; k++) { dp.push_back(new Point(1000.0 + k, 1000.0)); dp.push_back(new Point(1000.0 + k, 1001.0)); }
    std::vector<Cell*> dc;
    for (int k = 0; k < 4; k++) {
        Cell* C = new Cell(dp[2 * k], dp[2 * k + 2], dp[2 * k + 3], dp[2 * k + 1]);
        C->type = C_2;
        for (int b = 0; b < 2; b++) { C->par[b].ro = 1.0; C->par[b].p = 1.0; C->par[b].u = 0.0; C->par[b].v = 0.0; C->par[b].Q = 1.0; }
        dc.push_back(C);
    }
    Gran* G = new Gran(dp[4], dp[5], Usualy);
    Gran* G2 = new Gran(dp[5], dp[4], Usualy);
    G->Master = dc[1]; G->Sosed = dc[2]; G->Sosed_down = dc[0]; G->Sosed_up = dc[3];
    G2->Master = dc[2]; G2->Sosed = dc[1]; G2->Sosed_down = dc[3]; G2->Sosed_up = dc[0];
    G->Gran_copy = G2; G2->Gran_copy = G;
    S->Line_Contact.clear(); S->Line_Inner.clear(); S->Line_Outer.clear();
    S->Line_Contact.push_back(G);
    S->Line_Inner.push_back(G);
    S->Line_Outer.push_back(G);
}

static void zero_moments(Parametr& p)
{
    for (int h = 0; h < 4; h++) { p.H_n[h] = 0.0; p.H_u[h] = 0.0; p.H_v[h] = 0.0; p.H_T[h] = 0.0; }
    p.k_u = 0.0; p.k_v = 0.0; p.k_T = 0.0;
    p.I_u = 0.0; p.I_v = 0.0; p.I_T = 0.0;
}

static int mode_block(const char* in, const char* out, const char* ckpt)
{
    FILE* f = xopen(in, "rb");
    int32_t h[9]; rd(f, h, 9);
    std::fclose(f);
    const int steps = h[8];
    Setka* S0 = new Setka(h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7]);
    S0->Init_conditions();
    for (auto* K : S0->All_Cells) { zero_moments(K->par[0]); zero_moments(K->par[1]); }
    S0->Save_Setka_ALL_ALPHA(ckpt);
    Setka* S = new Setka();
    S->Download_Setka_ALL_ALPHA_2_0(ckpt);
    S->TVD_prepare();
    add_standin_lines_keep(S);

    const int nn = (int)S->All_Points.size(), nc = (int)S->All_Cells.size();
    std::vector<double> x(nn), y(nn);
    std::vector<int32_t> ptype(nn);
    for (int k = 0; k < nn; k++) {
        Point* P = S->All_Points[k];
        P->number = k;
        x[k] = P->x; y[k] = P->y;
        ptype[k] = P->type == P_Contact ? 1 : P->type == P_Inner_shock ? 2 : P->type == P_Outer_shock ? 3 : 0;
    }
    std::vector<int32_t> ncorner(nc), corners, ctype(nc), nface(nc), slots((size_t)nc * 8 * 6, -1);
    std::vector<double> prim0((size_t)5 * nc);
    for (int c = 0; c < nc; c++) {
        Cell* K = S->All_Cells[c];
        K->number = c;
    }
    for (int c = 0; c < nc; c++) {
        Cell* K = S->All_Cells[c];
        ncorner[c] = (int32_t)K->contour.size();
        for (auto* P : K->contour) corners.push_back(P->number);
        ctype[c] = (int32_t)K->type;
        nface[c] = (int32_t)K->Grans.size();
        if (K->Grans.size() > 8) { std::fprintf(stderr, "cell %d has %zu faces (> 8)\n", c, K->Grans.size()); return 3; }
        for (size_t k = 0; k < K->Grans.size(); k++) {
            Gran* G = K->Grans[k];
            int32_t* sl = &slots[((size_t)c * 8 + k) * 6];
            sl[0] = G->A->number; sl[1] = G->B->number; sl[2] = (int32_t)G->type;
            sl[3] = G->Sosed ? G->Sosed->number : -1;
            sl[4] = G->Sosed_down ? G->Sosed_down->number : -1;
            sl[5] = G->Sosed_up ? G->Sosed_up->number : -1;
        }
        const Parametr& p = K->par[0];
        prim0[c] = p.ro; prim0[nc + c] = p.p; prim0[2 * nc + c] = p.u; prim0[3 * nc + c] = p.v; prim0[4 * nc + c] = p.Q;
    }
    S->Go_5_komponent__MK2(steps);
    const int fb = steps % 2;
    std::vector<double> res((size_t)5 * nc), x2(nn), y2(nn);
    for (int c = 0; c < nc; c++) {
        const Parametr& p = S->All_Cells[c]->par[fb];
        res[c] = p.ro; res[nc + c] = p.p; res[2 * nc + c] = p.u; res[3 * nc + c] = p.v; res[4 * nc + c] = p.Q;
    }
    for (int k = 0; k < nn; k++) { x2[k] = S->All_Points[k]->x; y2[k] = S->All_Points[k]->y; }
    f = xopen(out, "wb");
    const int32_t hh[2] = { nn, nc };
    wr(f, hh, 2);
    wr(f, x.data(), nn); wr(f, y.data(), nn); wr(f, ptype.data(), nn);
    wr(f, ncorner.data(), nc); wr(f, corners.data(), corners.size()); wr(f, ctype.data(), nc);
    wr(f, nface.data(), nc); wr(f, slots.data(), slots.size());
    wr(f, prim0.data(), prim0.size()); wr(f, res.data(), res.size());
    wr(f, x2.data(), nn); wr(f, y2.data(), nn);
    std::fclose(f);
    return 0;
}

// ---- the reference's own block grid with its fitted surfaces moving --------------------
// in : int32 N1 N2 N3 N4 M1 M2 M3 M4 steps
// Setka(N1..M4) as the constructor builds it (rails of four kinds, the fitted-surface
// lines Line_Contact / Line_Inner / Line_Outer, Setka.cpp:19-1076), Init_conditions,
// hydrogen moments and k factors zeroed, TVD_prepare; then Go_5_komponent__MK2, whose
// every step calls Move_surface and Move_Setka_Calculate_2 (all four rail kinds) before
// the sweep.  Two runs from the same start: 1 step and `steps` steps (each call starts
// its own lagged time step at 1e-8).
// out: int32 nn nc; float64 x[nn] y[nn]; int32 ptype[nn]; int32 ncorner[nc];
//      int32 corners[sum]; int32 cell_type[nc]; int32 nface[nc]; int32 slots[nc][8][6];
//      float64 prim0[5][nc];
//      rails: int32 count[4] (A, B, C, D); per rail int32 M1 M2 M3 M4 nall nkey,
//             float64 s, int32 all[nall], int32 key[nkey];
//      lines: int32 count[3] (Contact, Inner, Outer); per face int32 A B Master Sosed
//             (-1 where none), in the list's order;
//      then for the 1-step and the `steps`-step runs: float64 prim[5][nc], x[nn], y[nn]
static void block_numbering(Setka* S)
{
    for (size_t k = 0; k < S->All_Points.size(); k++) S->All_Points[k]->number = (int)k;
    for (size_t c = 0; c < S->All_Cells.size(); c++) S->All_Cells[c]->number = (int)c;
}

static Setka* block_tracked_setka(const int32_t* h)
{
    Setka* S = new Setka(h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7]);
    S->Init_conditions();
    for (auto* K : S->All_Cells) { zero_moments(K->par[0]); zero_moments(K->par[1]); }
    S->TVD_prepare();
    block_numbering(S);
    return S;
}

static int mode_blocktrack(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t h[9]; rd(f, h, 9);
    std::fclose(f);
    const int steps = h[8];
    Setka* S = block_tracked_setka(h);
    const int nn = (int)S->All_Points.size(), nc = (int)S->All_Cells.size();
    f = xopen(out, "wb");
    const int32_t hh[2] = { nn, nc };
    wr(f, hh, 2);
    std::vector<double> x(nn), y(nn);
    std::vector<int32_t> ptype(nn);
    for (int k = 0; k < nn; k++) {
        Point* P = S->All_Points[k];
        x[k] = P->x; y[k] = P->y;
        ptype[k] = P->type == P_Contact ? 1 : P->type == P_Inner_shock ? 2 : P->type == P_Outer_shock ? 3 : 0;
    }
    wr(f, x.data(), nn); wr(f, y.data(), nn); wr(f, ptype.data(), nn);
    std::vector<int32_t> ncorner(nc), corners, ctype(nc), nface(nc), slots((size_t)nc * 8 * 6, -1);
    std::vector<double> prim0((size_t)5 * nc);
    for (int c = 0; c < nc; c++) {
        Cell* K = S->All_Cells[c];
        ncorner[c] = (int32_t)K->contour.size();
        for (auto* P : K->contour) corners.push_back(P->number);
        ctype[c] = (int32_t)K->type;
        nface[c] = (int32_t)K->Grans.size();
        if (K->Grans.size() > 8) { std::fprintf(stderr, "cell %d has %zu faces (> 8)\n", c, K->Grans.size()); return 3; }
        for (size_t k = 0; k < K->Grans.size(); k++) {
            Gran* G = K->Grans[k];
            int32_t* sl = &slots[((size_t)c * 8 + k) * 6];
            sl[0] = G->A->number; sl[1] = G->B->number; sl[2] = (int32_t)G->type;
            sl[3] = G->Sosed ? G->Sosed->number : -1;
            sl[4] = G->Sosed_down ? G->Sosed_down->number : -1;
            sl[5] = G->Sosed_up ? G->Sosed_up->number : -1;
        }
        const Parametr& p = K->par[0];
        prim0[c] = p.ro; prim0[nc + c] = p.p; prim0[2 * nc + c] = p.u; prim0[3 * nc + c] = p.v; prim0[4 * nc + c] = p.Q;
    }
    wr(f, ncorner.data(), nc); wr(f, corners.data(), corners.size()); wr(f, ctype.data(), nc);
    wr(f, nface.data(), nc); wr(f, slots.data(), slots.size()); wr(f, prim0.data(), prim0.size());
    const std::vector<Rail*>* kinds[4] = { &S->A_Rails, &S->B_Rails, &S->C_Rails, &S->D_Rails };
    int32_t rc[4];
    for (int q = 0; q < 4; q++) rc[q] = (int32_t)kinds[q]->size();
    wr(f, rc, 4);
    for (int q = 0; q < 4; q++)
        for (Rail* R : *kinds[q]) {
            const int32_t rh[6] = { R->M1, R->M2, R->M3, R->M4, (int32_t)R->All_point.size(), (int32_t)R->Key_point.size() };
            wr(f, rh, 6);
            wr(f, &R->s, 1);
            for (Point* P : R->All_point) { const int32_t k = P->number; wr(f, &k, 1); }
            for (Point* P : R->Key_point) { const int32_t k = P->number; wr(f, &k, 1); }
        }
    const std::vector<Gran*>* lines[3] = { &S->Line_Contact, &S->Line_Inner, &S->Line_Outer };
    int32_t lc[3];
    for (int q = 0; q < 3; q++) lc[q] = (int32_t)lines[q]->size();
    wr(f, lc, 3);
    for (int q = 0; q < 3; q++)
        for (Gran* G : *lines[q]) {
            const int32_t g[4] = { G->A->number, G->B->number, G->Master ? G->Master->number : -1,
                                   G->Sosed ? G->Sosed->number : -1 };
            wr(f, g, 4);
        }
    for (int run = 0; run < 2; run++) {
        Setka* T = run == 0 ? S : block_tracked_setka(h);
        const int n = run == 0 ? 1 : steps;
        T->Go_5_komponent__MK2(n);
        const int fb = n % 2;
        std::vector<double> res((size_t)5 * nc), x2(nn), y2(nn);
        for (int c = 0; c < nc; c++) {
            const Parametr& p = T->All_Cells[c]->par[fb];
            res[c] = p.ro; res[nc + c] = p.p; res[2 * nc + c] = p.u; res[3 * nc + c] = p.v; res[4 * nc + c] = p.Q;
        }
        for (int k = 0; k < nn; k++) { x2[k] = T->All_Points[k]->x; y2[k] = T->All_Points[k]->y; }
        wr(f, res.data(), res.size()); wr(f, x2.data(), nn); wr(f, y2.data(), nn);
    }
    std::fclose(f);
    return 0;
}

// ---- the reference's Monte Carlo of the hydrogen atoms on its own block grid -----------
// in : int32 N1 N2 N3 N4 M1 M2 M3 M4 seed
// Writes the seed tables M_K_prepare reads (my_dat_822.txt: 822 seed triples of
// Sensor, Sensors276405.txt: seed quadruples of sensor2) into the working directory from
// the seed, builds Setka(N1..M4) with Init_conditions, and runs M_K_prepare and
// MK_start_new (Setka.cpp:11228-12535).
// out: int32 nc; float64 per cell [nc][24]: cx cy ro p u v, H_n[4] H_u[4] H_v[4] H_T[4],
//      I_u I_v I_T (the Monte Carlo's momentum and energy sources), k_u k_v k_T
static int mode_mc(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t h[9]; rd(f, h, 9);
    std::fclose(f);
    {
        std::srand((unsigned)h[8]);
        FILE* g = xopen("my_dat_822.txt", "w");
        for (int i = 0; i < 822; i++)
            std::fprintf(g, "%d %d %d\n", 2 * (std::rand() % 16384) + 1, std::rand() % 32768, 2 * (std::rand() % 512) + 1);
        std::fclose(g);
        g = xopen("Sensors276405.txt", "w");
        for (int i = 0; i < 276405; i++)
            std::fprintf(g, "%u %u %u %u\n", 128u + (unsigned)std::rand(), 128u + (unsigned)std::rand(),
                         128u + (unsigned)std::rand(), 128u + (unsigned)std::rand());
        std::fclose(g);
    }
    Setka* S = new Setka(h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7]);
    S->Init_conditions();
    for (auto* K : S->All_Cells) { zero_moments(K->par[0]); zero_moments(K->par[1]); }
    S->TVD_prepare();
    const auto t0 = std::chrono::steady_clock::now();
    S->M_K_prepare();
    S->MK_start_new();
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "ref_mc_seconds %.3f\n", sec);
    const int nc = (int)S->All_Cells.size();
    std::vector<double> o((size_t)nc * 24);
    for (int c = 0; c < nc; c++) {
        Cell* K = S->All_Cells[c];
        const Parametr& p = K->par[0];
        double x, y;
        K->Get_Center(x, y);
        double* r = &o[(size_t)c * 24];
        r[0] = x; r[1] = y; r[2] = p.ro; r[3] = p.p; r[4] = p.u; r[5] = p.v;
        for (int k = 0; k < 4; k++) { r[6 + k] = p.H_n[k]; r[10 + k] = p.H_u[k]; r[14 + k] = p.H_v[k]; r[18 + k] = p.H_T[k]; }
        r[22] = p.I_u; r[23] = p.I_v;
    }
    f = xopen(out, "wb");
    const int32_t n32 = nc;
    wr(f, &n32, 1);
    wr(f, o.data(), o.size());
    std::fclose(f);
    return 0;
}

// Point-wise pieces of the reference's Monte Carlo (MKmethod.cpp, sensor.cpp,
// Setka.cpp:15830-15866): the random sensor, the charge-exchange rate integrals and the
// exchange-velocity sampler of Change_Velosity4 with no splitting zones (I = 0, the call
// of Fly_exchenge_Imit_Korol at Setka.cpp:13465).
// in : int32 nr; float64 [nr][2] = u cp
//      int32 ns, nd; int32 [ns][3] sensor seeds
//      int32 nx, reps; float64 [nx][7] = Ur Uthe Uphi Vr Vthe Vphi cp; int32 [nx][3] seeds
// out: float64 [nr][6] = Velosity_1 Velosity_2 Velosity_3 int_1 int_2 int_3 (int_k only
//      for u / cp <= 7, else 0)
//      float64 [ns][nd] draws
//      float64 [nx][reps][4] = Wr Wthe Wphi mu; int32 [nx][3] sensor state afterwards
static int mode_mcunit(const char* in, const char* out)
{
    FILE* f = xopen(in, "rb");
    int32_t nr; rd(f, &nr, 1);
    std::vector<double> ra((size_t)nr * 2); rd(f, ra.data(), ra.size());
    int32_t nsd[2]; rd(f, nsd, 2);
    std::vector<int32_t> seeds((size_t)nsd[0] * 3); rd(f, seeds.data(), seeds.size());
    int32_t nxr[2]; rd(f, nxr, 2);
    std::vector<double> xa((size_t)nxr[0] * 7); rd(f, xa.data(), xa.size());
    std::vector<int32_t> xs((size_t)nxr[0] * 3); rd(f, xs.data(), xs.size());
    std::fclose(f);
    Setka* S = new Setka();
    MKmethod MK;
    f = xopen(out, "wb");
    for (int k = 0; k < nr; k++) {
        const double u = ra[2 * k], cp = ra[2 * k + 1];
        double o[6] = { S->Velosity_1(u, cp), S->Velosity_2(u, cp), S->Velosity_3(u, cp), 0.0, 0.0, 0.0 };
        if (u / cp <= 7.0) { o[3] = MK.int_1(u, cp); o[4] = MK.int_2(u, cp); o[5] = MK.int_3(u, cp); }
        wr(f, o, 6);
    }
    for (int s = 0; s < nsd[0]; s++) {
        Sensor sen(seeds[3 * s], seeds[3 * s + 1], seeds[3 * s + 2]);
        for (int d = 0; d < nsd[1]; d++) { const double v = sen.MakeRandom(); wr(f, &v, 1); }
    }
    std::vector<int32_t> st((size_t)nxr[0] * 3);
    for (int x = 0; x < nxr[0]; x++) {
        const double* a = &xa[(size_t)x * 7];
        Sensor sen(xs[3 * x], xs[3 * x + 1], xs[3 * x + 2]);
        for (int r = 0; r < nxr[1]; r++) {
            std::vector<double> Wr(1), Wt(1), Wp(1), mu(1);
            MK.Change_Velosity4(&sen, a[0], a[1], a[2], a[3], a[4], a[5], Wr, Wt, Wp, mu, a[6], 1.0, 0);
            const double o[4] = { Wr[0], Wt[0], Wp[0], mu[0] };
            wr(f, o, 4);
        }
        st[3 * x] = sen.a1_; st[3 * x + 1] = sen.a2_; st[3 * x + 2] = sen.a3_;
    }
    wr(f, st.data(), st.size());
    std::fclose(f);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc < 4) { std::fprintf(stderr, "usage: ref_harness hllc|helpers|cell|sweep|hlld|godunov|track|mf|inner|save|resave|blocktrack in out | block in out checkpoint\n"); return 2; }
    if (!std::strcmp(argv[1], "hllc")) return mode_hllc(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "helpers")) return mode_helpers(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "cell")) return mode_cell(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "sweep")) return mode_sweep(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "hlld")) return mode_hlld(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "godunov")) return mode_godunov(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "track")) return mode_track(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "mf")) return mode_mf(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "inner")) return mode_inner(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "save")) return mode_save(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "resave")) return mode_resave(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "block") && argc >= 5) return mode_block(argv[2], argv[3], argv[4]);
    if (!std::strcmp(argv[1], "blocktrack")) return mode_blocktrack(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "mc")) return mode_mc(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "mcunit")) return mode_mcunit(argv[2], argv[3]);
    std::fprintf(stderr, "unknown mode %s\n", argv[1]);
    return 2;
}
