/*
 * gelios_oracle.h - CPU restatement of the reference sweep.  TEST INFRASTRUCTURE.
 *
 * This library is the checker for the HIP path: only tests, the smoke check and the
 * CPU-baseline leg of the benchmark load it.  The product path (gelios_2d_amd) never
 * links or calls it.
 *
 * It restates, in plain C and FP64, the cell loop of Setka::Go_5_komponent__MK2
 * (Setka.cpp:10121-10482) on a general mesh given as per-cell face slots, the way the
 * reference walks Cell::Grans.  Each function cites the reference lines it follows.
 * Floating-point expressions keep the reference's operation order so that the
 * restatement agrees with the compiled reference to the last bits where the
 * compilers agree.
 *
 * Pinning (see tests/test_oracle_*.py): the face flux, geometry, limiter and the
 * whole sweep are compared with vectors written by the compiled reference
 * (oracle/_ref, tests/golden/); the flux is also checked against closed forms
 * (consistency F(U,U), mirror antisymmetry, supersonic upwinding).
 */
#ifndef GELIOS_ORACLE_H
#define GELIOS_ORACLE_H

#include "../include/gelios.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_MAXF 8   /* face slots per cell */

/* General mesh: per-cell face slots in Cell::Grans order.  Slot s = c*ORA_MAXF + k.
 * fa/fb: nodes A,B of the face as seen from the cell (normal to the right of A->B,
 * pointing out of the cell, Gran.h:22).  fnbr: Gran::Sosed (-1 if none),
 * fdown: Gran::Sosed_down, fup: Gran::Sosed_up (-1 if none), ftype: Gran_type. */
typedef struct ora_mesh {
    int n_nodes, n_cells;
    const double *x, *y;      /* node positions now (Point::x, y) */
    const double *x2, *y2;    /* node positions after the step (Point::x2, y2) */
    const int *contour;       /* 4 node indices per cell, counter-clockwise */
    const int *cell_type;     /* enum gel_cell_type */
    const int *nface;         /* faces per cell */
    const int *fa, *fb, *ftype, *fnbr, *fdown, *fup;
    /* node types of the fitted surfaces (0 none, 1 P_Contact, 2 P_Inner_shock,
       3 P_Outer_shock, Point.h:15-17), or NULL: no surfaces */
    const int *ptype;
    /* polygons: corners of cell c are contour[coff[c] .. coff[c+1]) (Cell::contour);
       NULL: every cell is a quad with contour[4c .. 4c+4) */
    const int *coff;
} ora_mesh;

typedef struct ora_gas { double ro, p, u, v, Q; } ora_gas;
/* MHD primitive state: velocity (u, v, w) and field (bx, by, bz) in (x, y, phi). */
typedef struct ora_mhd { double ro, p, u, v, w, bx, by, bz, Q; } ora_mhd;

/* HLLD of Solvers::HLLDQ_Korolkov metod 3 (Solvers.cpp:1642-2185) for any in-plane
 * normal and face speed W.  P: mass, momentum x/y/phi, energy, B x/y/phi, tracer.
 * Returns max(|SL|,|SR|). */
double ora_hlld(const ora_mhd* L, const ora_mhd* R, double W, double n1, double n2, double gam, double P[9]);

/* Exact Riemann solver for a gamma-gas with face speed w (Solvers::Godunov_Solver_
 * Alexashov, Solvers.cpp:693-1321).  q1, q2 = (ro, p, u, v, w); n = unit normal.
 * qqq = mass, momentum x/y/z, energy fluxes relative to the face; returns the local
 * time step cfl * dist / max|speed|; *status = 1 where the reference would exit. */
double ora_godunov(const double q1[5], const double q2[5], const double n[3], double w, double dist, double cfl,
                   double gam, double qqq[5], double* dsl, double* dsp, double* dsc, int* status);

/* Setka::HLLC_2d_Korolkov_b_s (Setka.cpp:10484-10700). Returns cfl*rad/max|S|. */
double ora_hllc2d(const ora_gas* L, const ora_gas* R, double W, double n1, double n2,
                  double rad, double cfl, double gam, int metod, int nul_potok,
                  double P[4], double* PQ, double* Vl, double* Vc, double* Vp);

/* Help.cpp helpers. */
double ora_minmod(double x, double y);                                  /* :67-78 */
double ora_linear3(double x1, double t1, double x2, double t2,
                   double x3, double t3, double y);                    /* :96-113 */
double ora_linear2(double x1, double t1, double x2, double t2, double y); /* :115-120 */
double ora_polar_angle(double x, double y);                             /* :37-65 */
void ora_polar_perenos(double x1, double y1, double x2, double y2,
                       double* u, double* v);                           /* :3-14 */

/* Cell geometry on 4 nodes (Cell.cpp:66-88, :146-186, :188-206). */
void ora_cell_center(const double* px, const double* py, const int* q, double* cx, double* cy);
double ora_cell_area(const double* px, const double* py, const int* q);
double ora_cell_volume_rotate(const double* px, const double* py, const int* q, double angle);

/* Cell::Get_Sourse_MK1 (Cell.cpp:498-665) with the cell's own moments.
 * f points at the 19 fields of one cell (H_n[4],H_u[4],H_v[4],H_T[4],k_u,k_v,k_T). */
void ora_source_mk1(const gel_params* prm, const double* f, double u, double v,
                    double ro, double p, double* q1, double* q2, double* q3);

/* One step of the cell loop of Setka::Go_5_komponent__MK2 (Setka.cpp:10148-10470).
 * now/next: 5 x n_cells primitive SoA (ro,p,u,v,Q).  src: 19 x n_cells or NULL.
 * dt_prev is T[now1]; *dt_new receives T[now2].  Cells skipped by r_skip keep
 * whatever next held (Setka.cpp:10169-10172).  Returns GEL_OK or GEL_E_NUMERIC
 * where the reference would exit(-1) (Gran.cpp:444-455). */
int ora_step_mk2(const ora_mesh* m, const gel_params* prm, const double* now,
                 double* next, const double* src, double dt_prev, double* dt_new,
                 long long* n_rho_clamp, long long* n_p_clamp, int nthreads);

/* MHD form of the same cell loop: 9 x n_cells primitives (ro,p,u,v,w,bx,by,bz,Q),
 * HLLD faces, axisymmetric MHD geometric sources.  C_1 cells are not supported. */
int ora_step_mhd(const ora_mesh* m, const gel_params* prm, const double* now, double* next,
                 const double* src, double dt_prev, double* dt_new, long long* n_rho_clamp,
                 long long* n_p_clamp, int nthreads);

/* Face-once summation for ora_step_mhd (on != 0): every interior face is solved once,
 * from the cell of lower index, and its neighbour takes the negated flux x area
 * product (the device sweep's summation); 0 (default): every face from both sides, as
 * the reference's cell loop does (Setka.cpp:10195). */
void ora_set_face_once(int on);

/* Face slot k of cell c as ora_step_mhd evaluates it from that cell: the flux x area
 * products PS[9] and the largest wave speed; LRW (22 values, or NULL) receives the two
 * face states (9 + 9 primitives), the face speed W, the normal (n1, n2) and the area S. */
int ora_mhd_face(const ora_mesh* m, const gel_params* prm, const double* now, int c, int k, double dt_prev,
                 double PS[9], double* UU, double* LRW);

/* The static inner-region solver Setka::Go_stationary_5_komponent_inner_MK2
 * (Setka.cpp:6999-7212): 5 x n_cells gas primitives; src as for ora_step_mk2. */
int ora_step_inner(const ora_mesh* m, const gel_params* prm, const double* now, double* next, const double* src,
                   double dt_prev, double* dt_new, long long* n_rho_clamp, long long* n_p_clamp, int nthreads);

/* The plasma and four hydrogen fluids of Setka::Go_5_komponent_2 (Setka.cpp:8951-9777):
 * 21 x n_cells primitives in Parametr order (ro, p, u, v, Q, then ro_Hk, p_Hk, u_Hk,
 * v_Hk for k = 1..4), every fluid with the HLL/HLLC flux (prm->flux), charge exchange
 * between the plasma and each fluid, hydrogen sources by cell-type region. */
int ora_step_mf(const ora_mesh* m, const gel_params* prm, const double* now, double* next, double dt_prev,
                double* dt_new, long long* n_rho_clamp, long long* n_p_clamp, int nthreads);

/* Fitted-surface tracking on the polar grid (Setka::Move_surface, Setka.cpp:3378-3844,
 * and the A rails of Setka::Move_Setka_Calculate_2, Setka.cpp:4324-4381): node
 * positions x2, y2 after a step of length dt from the state now (5 x ni*nj) and the
 * node positions x, y; rail_s[c] is the angle of node column c.  Returns GEL_OK, or
 * GEL_E_NUMERIC where the exact solver would have exited. */
int ora_track_nodes(const gel_tracking* t, int ni, int nj, const double* x, const double* y, const double* now,
                    const double* rail_s, double gam, double dt, double* x2, double* y2);

/* ---- Monte Carlo of the interstellar hydrogen (mc_oracle.c) ---- */
double ora_sensor_draw(int s[3]);                                      /* sensor.cpp:11-31 */
void ora_mc_seed_primary(unsigned long long seed, long long k, int s[3]);
void ora_mc_velosity(double u, double cp, double o[3]);                /* Setka.cpp:15830-15866 */
void ora_mc_int(double u, double cp, double sig_a, double o[3]);       /* MKmethod.cpp:2741-2858 */
/* MKmethod::Change_Velosity4 with I = 0 (MKmethod.cpp:2270-2531): in = Ur Uthe Uphi Vr
 * Vthe Vphi cp (velocities in units of cp), out = Wr Wthe Wphi mu; returns mu */
int ora_mc_exchange(int s[3], const double in[7], double sig_a, double out[4]);
double ora_mc_inflow(double v_inf);
/* The tracer on a polar grid (Setka::MK_start_new with Fly_exchenge_Imit_Korol,
 * Setka.cpp:11989-13981, in the form of gel_mc_run): prim = the state (rho, p, u, v the
 * first four of the variable-major fields), pop = population of a secondary born in each
 * cell; out = GEL_MC_NFIELDS x ni*nj as gel_mc_run writes it. */
int ora_mc_run(int ni, int nj, const double* x, const double* y, const double* prim, const int8_t* pop,
               const gel_mc_params* prm, double* out, gel_mc_stats* st);
/* The same on a general mesh (the reference's own grid), the atoms walking each cell's
 * face slots (Setka.cpp:12893-12960): coff/corners the contours, soff/sA/sB the slots in
 * Cell::Grans order, scode each slot's way on (the next cell; -1 out of the grid, where
 * primaries enter; -2 across the inner sphere; -3 an Axis slot, skipped). */
int ora_mc_run_mesh(int nc, int nn, const double* x, const double* y, const int* coff, const int* corners,
                    const int* soff, const int* sA, const int* sB, const int* scode, const double* prim,
                    const int8_t* pop, const gel_mc_params* prm, double* out, gel_mc_stats* st);

#ifdef __cplusplus
}
#endif

#endif
