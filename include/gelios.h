/*
 * gelios.h - C-ABI of the MI355X Godunov flux-and-update sweep.
 *
 * The reference (SergeyMSU/Gelios_2D) has no plugin or FFI layer: its sweep is the
 * C++ member Setka::Go_5_komponent__MK2 (Setka.h:186, Setka.cpp:10121-10482), which
 * loops over cells with OpenMP and calls Gran/Cell helpers and the face flux
 * Setka::HLLC_2d_Korolkov_b_s (Setka.h:187-189, Setka.cpp:10484-10700).  The entry
 * points below are what a host that keeps the reference's Setka/Cell/Gran objects
 * calls instead of that OpenMP loop.  Each declaration names the reference code it
 * replaces.
 *
 * Conventions
 *   - Every call returns an int status, GEL_OK (0) on success; nothing calls exit().
 *     gel_strerror() turns a status into text; gel_last_error() gives the detail of
 *     the last failure on the calling thread.
 *   - A context owns its device buffers; the caller owns every host array passed in.
 *   - One context is driven by one host thread at a time.
 *   - Results are deterministic: fixed summation order, no floating-point atomics
 *     (the global time step is reduced with an integer min on the bit pattern of a
 *     positive double).  The one exception is the Monte Carlo (gel_mc_run), whose
 *     per-cell tallies are FP64 atomic sums: runs agree to rounding, not bit for bit.
 *
 * Data layout (host side and device side alike)
 *   - Grid: logically structured polar grid of ni (along a ray, index i) by nj
 *     (across rays, index j) quadrilateral cells.  Node (i,j), i in [0,ni],
 *     j in [0,nj], lives at index i*(nj+1)+j.  x is the symmetry-axis coordinate,
 *     y >= 0 the distance from the axis (the reference's Point::x, Point::y).
 *     Cell (i,j) has contour (i,j),(i+1,j),(i+1,j+1),(i,j+1), counter-clockwise.
 *     Faces with j = 0 and j = nj lie on the axis (Gran_type Axis), the faces at
 *     i = 0 are the inner sphere (Inner_sphere), the faces at i = ni are Input or
 *     Extern as given by outer_type.
 *   - State: primitive variables, structure of arrays, variable-major:
 *     prim[v*ncells + i*nj + j].  Gas (5 variables): ro, p, u, v, Q in the order of
 *     the reference struct Parametr (Cell.h:19-24).  MHD (9 variables): ro, p, u, v,
 *     w, bx, by, bz, Q (w and bz are the azimuthal velocity and field).
 */
#ifndef GELIOS_H
#define GELIOS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum gel_status {
    GEL_OK = 0,
    GEL_E_ARG = 1,          /* invalid argument (null pointer, bad size, bad enum) */
    GEL_E_HIP = 2,          /* a HIP runtime call failed */
    GEL_E_NOMEM = 3,        /* device allocation failed */
    GEL_E_UNSUPPORTED = 4,  /* combination of options not built */
    GEL_E_STATE = 5,        /* call out of order (e.g. stepping before upload) */
    GEL_E_NUMERIC = 6       /* non-positive reconstructed density/pressure where the
                               reference would exit(-1) (Gran.cpp:444-455) */
};

/* Face boundary types: same enumerators, same order as Gran_type (Gran.h:7-16). */
enum gel_face_type {
    GEL_F_USUALY = 0, GEL_F_INNER_SPHERE = 1, GEL_F_EXTERN = 2,
    GEL_F_INPUT = 3, GEL_F_AXIS = 4, GEL_F_UPPER_WALL = 5
};

/* Cell types: same enumerators, same order as Cell_type (Cell.h:8-16).  C_1 cells
 * get the radially rescaled reconstruction (Setka.cpp:10219-10249), C_centr cells
 * are frozen (Setka.cpp:10461-10468). */
enum gel_cell_type {
    GEL_C_1 = 0, GEL_C_2 = 1, GEL_C_3 = 2, GEL_C_4 = 3, GEL_C_5 = 4,
    GEL_C_CENTR = 5, GEL_C_NO = 6
};

/* Face flux: the `metod` argument of HLLC_2d_Korolkov_b_s (0 = HLL, 1 = HLLC). */
enum gel_flux { GEL_FLUX_HLL = 0, GEL_FLUX_HLLC = 1 };

/* GAS: 5 primitives (ro, p, u, v, Q), face flux of HLLC_2d_Korolkov_b_s.
 * MHD: 9 primitives (ro, p, u, v, w, bx, by, bz, Q), HLLD face flux after
 * Solvers::HLLDQ_Korolkov metod 3 (Solvers.cpp:1980-2185) for any in-plane normal and
 * moving faces; B in Gaussian units as there (b = B/sqrt(4 pi)).
 * GAS4H: the plasma and four hydrogen fluids of Setka::Go_5_komponent_2
 * (Setka.cpp:8951-9777): 21 primitives in the order of Parametr (Cell.h:19-24):
 * ro, p, u, v, Q, then ro_Hk, p_Hk, u_Hk, v_Hk for k = 1..4; every fluid takes the
 * HLL/HLLC face flux, the fluids exchange mass, momentum and energy by charge exchange
 * (Setka.cpp:9400-9617); FP64, no separate source mode. */
enum gel_physics { GEL_PHYS_GAS = 0, GEL_PHYS_MHD = 1, GEL_PHYS_GAS4H = 2, GEL_PHYS_GAS_INNER = 3 };
/* GAS_INNER: the 5 gas primitives advanced by the static inner-region solver
 * Setka::Go_stationary_5_komponent_inner_MK2 (Setka.cpp:6999-7212): both sides of every
 * interior face rescaled radially to the face (radial reconstruction where both cells
 * are C_1), ghost values on both sides of the inner sphere, HLL on the axis, no
 * r_skip, C_centr cells skipped, the tracer losing a bare dt per step
 * (Setka.cpp:7173).  FP64, static grids. */

/* FP64: every operation in FP64 (the reference's arithmetic).
 * FP32_BULK: state stored in FP32 and face reconstruction + Riemann solver in FP32;
 * geometry (normals, areas, swept volumes), the accumulation of face fluxes into cells
 * and the update with primitive recovery stay FP64. */
enum gel_precision { GEL_PREC_FP64 = 0, GEL_PREC_FP32_BULK = 1 };

enum gel_source_mode {
    GEL_SRC_NONE = 0,
    /* Charge exchange with 4 frozen hydrogen populations, Cell::Get_Sourse_MK1
       (Cell.cpp:498-665); moments n,u,v,T per population and k_u,k_v,k_T per cell. */
    GEL_SRC_MK1 = 1,
    /* Per-cell source table (mass, momentum x, momentum y, energy, tracer per unit
       r-weighted volume), set with gel_set_source_table or filled on the device by
       gel_surrogate_refresh. */
    GEL_SRC_TABLE = 2,
    /* The learned source with its analytic fallback: cells the last refresh's judge
       flagged take the GEL_SRC_MK1 closure at every step, the others the table.
       Needs the hydrogen moments and the table; the flags start cleared. */
    GEL_SRC_HYBRID = 3
};

/* Number of per-cell source inputs for GEL_SRC_MK1, variable-major like the state:
 * H_n[4], H_u[4], H_v[4], H_T[4], k_u, k_v, k_T. */
#define GEL_MK1_NFIELDS 19

typedef struct gel_params {
    double gamma;       /* adiabatic index (reference macro ga/ggg = 5/3, Help.h:39-40) */
    double cfl;         /* Courant number (reference macro kurant = 0.1, Help.h:57) */
    double dt0;         /* time step of the first step (T0 = 1e-8, Setka.cpp:10128-10129) */
    double r_skip;      /* cells whose centre radius is below it are skipped (R11_,
                           Setka.cpp:10169-10172); 0 disables */
    int flux;           /* enum gel_flux */
    int physics;        /* enum gel_physics */
    int source_mode;    /* enum gel_source_mode */
    int precision;      /* enum gel_precision */
    /* Inner_sphere ghost state at face centre distance d (Gran.cpp:234-265):
       ro = 1/(in_rho_div*d*d), p = ro*in_p_a*in_p_a/in_p_den,
       (u,v) = in_u*(x,y)/d, Q = ro*in_q. */
    double in_rho_div, in_p_a, in_p_den, in_u, in_q;
    /* Input ghost state (Gran.cpp:276-282): ro, p, u, v, Q. */
    double inf_rho, inf_p, inf_u, inf_v, inf_q;
    /* Extern: copy of the cell, with u >= 0 replaced by this value (Gran.cpp:212-220). */
    double extern_u;
    /* Charge-exchange constants (Help.h:60-64, Cell.cpp:620-659):
       coefficient n_p/Kn, and sigma(U) = (sig_a - sig_b*ln U)^2. */
    double src_coef, sig_a, sig_b;
    /* MHD Input ghost: azimuthal velocity and field (bx, by, bphi); the inner-sphere
       wind carries no field. */
    double inf_w, inf_bx, inf_by, inf_bz;
    /* GAS4H: coefficient n_H/Kn of the hydrogen sources (Setka.cpp:9528-9547); the
       exponent H_pow of the radial rescaling of ro_H1 (Help.h:90, Setka.cpp:10216); the
       H1 inner-sphere ghost ro, p (velocity as the wind's, Gran.cpp:234-265); the H4
       Input ghost ro, p, u, v (Gran.cpp:276-282). */
    double src_coef_h, h_pow, h1_in_rho, h1_in_p, h4_inf_rho, h4_inf_p, h4_inf_u, h4_inf_v;
} gel_params;

/* Fill *p with the reference's own constants (Help.h, Gran.cpp) and HLL flux. */
void gel_params_reference(gel_params* p);

typedef struct gel_grid {
    int ni, nj;                 /* cells along and across the rays */
    const double* x;            /* (ni+1)*(nj+1) node axis coordinates */
    const double* y;            /* (ni+1)*(nj+1) node distances from the axis (>= 0) */
    const int8_t* cell_type;    /* ni*nj enum gel_cell_type, or NULL for all C_2 */
    const int8_t* outer_type;   /* nj face types of the i = ni boundary, or NULL for
                                   Input where the face centre has x > 0, else Extern */
} gel_grid;

typedef struct gel_stats {
    double dt_last;         /* time step used by the last step taken */
    double dt_next;         /* time step the next step will use (global CFL minimum) */
    long long steps;        /* steps taken by this context */
    long long n_rho_clamp;  /* cumulative density clamps (Setka.cpp:10432-10436) */
    long long n_p_clamp;    /* cumulative pressure clamps (Setka.cpp:10450-10453) */
    long long n_numeric;    /* faces where the reference would have exited */
} gel_stats;

typedef struct gel_ctx gel_ctx;

/* Create a context for one grid: uploads node coordinates and the grid tables.
 * Replaces the set-up the reference does before its sweep (Setka::TVD_prepare,
 * Setka.cpp:1294-1492, builds the face stencils that are implicit here). */
int gel_ctx_create(gel_ctx** out, const gel_grid* grid, const gel_params* prm);

/* The reference's own grid: any polygon mesh with the face tables Setka::TVD_prepare
 * leaves (Setka.cpp:1294-1434) -- the block-structured heliosphere grid of Setka(N1..M4)
 * (Setka.cpp:19-1076: rails of four kinds, fitted-surface rows, C_centr wedges), or one
 * read from a checkpoint by Setka::Download_Setka_ALL_ALPHA_2_0 (Setka.cpp:2899-3258).
 * Cell c's corners are corners[corner_offset[c] .. corner_offset[c+1]) (Cell::contour),
 * its faces the slots face_offset[c] .. face_offset[c+1]) in Cell::Grans order, each
 * with the face's nodes A, B as the cell sees them (Gran::A, B), Gran::type, Sosed,
 * Sosed_down and Sosed_up (-1 where absent).  Interior faces must come as a face and its
 * copy (nodes reversed, Sosed_down and Sosed_up exchanged), as the reference builds
 * them.  point_type: 0, 1 P_Contact, 2 P_Inner_shock, 3 P_Outer_shock (NULL: none); a
 * face between two nodes of one fitted surface takes the exact solver, a face with one
 * such node the nul_potok flux (Setka.cpp:10287-10370).  The mesh context runs the
 * gas sweep (HLL / HLLC; GEL_SRC_NONE, GEL_SRC_MK1 or GEL_SRC_TABLE) in FP64 or in FP32 bulk (state
 * stored in FP32, the Riemann solver of the faces other than the fitted-surface and
 * nul_potok ones in FP32; reconstruction, geometry and update in FP64; fitted-surface
 * tracking needs FP64) with gel_step (static) and
 * gel_step_moving (next node positions from the caller, as Move_Setka_Calculate_2 leaves
 * them); state and sources as for polar grids, n_cells cells in the mesh's order. */
typedef struct gel_mesh {
    int n_nodes, n_cells;
    const double* x;
    const double* y;
    const int8_t* point_type;
    const int* corner_offset;
    const int* corners;
    const int8_t* cell_type;
    const int* face_offset;
    const int* face_a;
    const int* face_b;
    const int* face_type;
    const int* face_sosed;
    const int* face_down;
    const int* face_up;
} gel_mesh;
int gel_ctx_create_mesh(gel_ctx** out, const gel_mesh* mesh, const gel_params* params);
void gel_ctx_destroy(gel_ctx* ctx);

/* Copy the primitive state (variable-major SoA) to the device; resets the time step
 * to prm->dt0 and the step counter.  Replaces Cell::par[now1] as the sweep input. */
int gel_upload_state(gel_ctx* ctx, const double* prim);
/* Copy the current primitive state back (Cell::par[now2] after the sweep). */
int gel_download_state(gel_ctx* ctx, double* prim);

/* Per-cell source inputs, GEL_MK1_NFIELDS variable-major arrays of ncells doubles
 * (Parametr::H_n/H_u/H_v/H_T/k_u/k_v/k_T, Cell.h:55-66).  Kept on the device in the
 * state's precision: FP32-bulk contexts round them to FP32 and evaluate the analytic
 * closure in FP32. */
int gel_set_sources(gel_ctx* ctx, const double* fields);

/* Per-cell source table for GEL_SRC_TABLE: 5 variable-major arrays of ncells doubles
 * (mass, momentum x, momentum y, energy, tracer), applied like the sources q2_1, q2_2,
 * q3 of the update (Setka.cpp:10444-10446).  Clears the fallback flags.  Kept on the
 * device in the state's precision (FP32-bulk contexts round it to FP32). */
int gel_set_source_table(gel_ctx* ctx, const double* q);
/* Copy the source table [5][ncells] back to the host (after a refresh, for tests). */
int gel_get_source_table(gel_ctx* ctx, double* q);
/* Copy the analytic-fallback flags [ncells] (1: the cell takes the GEL_SRC_MK1 closure
 * in GEL_SRC_HYBRID) back to the host.  No reference counterpart: the flags belong to
 * the learned source, which replaces the reference's per-cell source fields. */
int gel_get_source_flags(gel_ctx* ctx, int* flags);

/* Advance nsteps steps: replaces the body of the step loop of
 * Setka::Go_5_komponent__MK2 (Setka.cpp:10131-10481) on a static grid.
 * Steps run back to back on the context's stream without host round trips; the
 * lagged CFL step (Setka.cpp:10388-10393) stays on the device.  out may be NULL. */
int gel_step(gel_ctx* ctx, int nsteps, gel_stats* out);

/* One step on a moving grid: x2,y2 are the host node positions after the step, as
 * Setka::Move_Setka_Calculate_2 leaves them in Point::x2, y2 (Setka.cpp:10145); face
 * speeds and swept volumes come from old and new positions (Setka.cpp:10187-10209),
 * and the step ends with Point::x = x2 (Setka.cpp:10473-10480). */
int gel_step_moving(gel_ctx* ctx, const double* x2, const double* y2, gel_stats* out);

/* Prescribed node motion (no counterpart in the reference, whose Move_Setka_Calculate_2
 * redistributes rails, Setka.cpp:4324-4759): every node sits at its grid position
 * times s_k = 1 + amp*sin(2 pi k/period + phase), k = step mod period, so radii
 * breathe by amp; face speeds and swept volumes follow (Setka.cpp:10187-10209).
 * gel_step then runs moving steps. */
int gel_set_motion(gel_ctx* ctx, double amp, int period, double phase);

/* Fitted-surface tracking (Setka::Move_surface, Setka.cpp:3378-3844, and the A rails of
 * Setka::Move_Setka_Calculate_2, Setka.cpp:4324-4381) on the polar grid.  Three node
 * rows are fitted surfaces: the inner shock, the contact and the outer shock, in that
 * radial order, as the rails' key points (Rail.cpp:19-75).  Every node column c is a
 * rail, the ray at angle rail_angle[c] from the origin.  Each step then:
 *   1. solves the exact Riemann problem on every surface face from the one-sided
 *      reconstructions of both sides (Gran.cpp:345-426) and moves the surface nodes
 *      with a fraction of a wave speed (contact: koef*k_contact*contact speed; inner
 *      shock: koef*k_inner*left wave; outer shock: koef*k_outer*right wave), scattered
 *      to the face's nodes by the reference's rules and end-node corrections;
 *   2. moves every rail's key points along the rail by dt times that speed and
 *      redistributes the rail's other nodes between them (power law near the inner
 *      boundary and beyond the outer shock, uniform between the surfaces);
 *   3. takes the moving step (Setka.cpp:10148-10470) with the exact solver on the
 *      surface faces (mass, momentum, energy; the tracer from the HLL/HLLC flux,
 *      Setka.cpp:10335-10370) and the nul_potok HLLC on faces with one surface node
 *      (Setka.cpp:10299-10327), and commits the nodes (Setka.cpp:10473-10480).
 * Gas physics in FP64 only.  The rows must satisfy inner_rail_nodes <= row_inner_shock <
 * row_contact < row_outer_shock <= ni - 2, and the grid's rows 0..row_inner_shock
 * must lie on the rails (the step places them by the rail formulas). */
typedef struct gel_tracking {
    int row_inner_shock;    /* node rows of the fitted surfaces (P_Inner_shock, */
    int row_contact;        /* P_Contact, P_Outer_shock; Line_Inner, Line_Contact, */
    int row_outer_shock;    /* Line_Outer) */
    int outer_faces;        /* outer-shock faces j < outer_faces move it (Setka.cpp:3746,
                               n_outer_shock + N1 + N2 + 1 on the reference grid) */
    int average_lines;      /* 0: node speeds averaged as the reference does (the contact
                               line's nodes after each of the three loops,
                               Setka.cpp:3522-3530, :3718-3731, :3829-3837, so the
                               shocks' nodes keep sums); 1: each line averages its own */
    int inner_rail_nodes;   /* rail nodes in the inner block (the macro n_inner = 20, Help.h:66) */
    double r_inner, r_mid, r_outer;          /* R1_, R111_, R5_ (Help.h:76-83) */
    double p_inner, p_mid, p_outer, s_outer; /* rail spacing: 1.3, 1.05, 1.76, 1.5708
                                                (Setka.cpp:4349-4380) */
    double koef;                             /* 0.1 (Setka.cpp:3380) */
    double k_contact, k_inner, k_outer;      /* 0.3, 0.1, 1.0 */
    double tang_inner, tang_outer;           /* tangential-speed thresholds 0.1, 0.01 */
} gel_tracking;

/* Fill *t with the reference's constants; rows -1 (to be set), all outer faces. */
void gel_tracking_reference(gel_tracking* t);
/* Turn tracking on for gel_step (t != NULL; rail_angle[nj+1]) or off (t == NULL). */
int gel_set_tracking(gel_ctx* ctx, const gel_tracking* t, const double* rail_angle);
/* Fitted-surface tracking on a general-mesh context (the reference's own block grid,
 * gel_ctx_create_mesh): every gel_step first runs Setka::Move_surface over the three
 * lines (Setka.cpp:3378-3844) and Setka::Move_Setka_Calculate_2 over the four rail
 * families (Setka.cpp:4324-4759) on the device, then the sweep from the node positions
 * before to those after, then commits the nodes (Setka.cpp:10473-10480).  The rails
 * and lines are the reference's Setka::A_Rails, B_Rails, C_Rails, D_Rails (Rail.h:
 * type, s, M1..M4, All_point, Key_point) and Line_Contact, Line_Inner, Line_Outer
 * (Setka.h), given by node and cell index. */
typedef struct gel_mesh_rails {
    int n_rails;            /* A_Rails, then B_Rails, C_Rails, D_Rails, each in its order */
    const int* kind;        /* [n_rails] Rail::type: 0 A, 1 B, 2 C, 3 D */
    const int* m;           /* [n_rails][4] Rail::M1, M2, M3, M4 */
    const double* s;        /* [n_rails] Rail::s (the ray's angle, or the rail's coordinate) */
    const int* all_offset;  /* [n_rails + 1] offsets into all */
    const int* all;         /* Rail::All_point of every rail, node indices */
    const int* key;         /* [n_rails][4] Rail::Key_point (C rails: two, the rest -1) */
    int n_line[3];          /* faces of Line_Contact, Line_Inner, Line_Outer */
    const int* line;        /* [n_line[0] + n_line[1] + n_line[2]][4] per face, in the
                               lists' order: Gran::A, B (nodes), Master, Sosed (cells) */
    double left;            /* Left_ (Help.h:85): the C rails' outer end */
    int d_chain;            /* n_outer_shock (Help.h:68): D rails from this index on take
                               the previous D rail's outer-shock position */
} gel_mesh_rails;
/* t: the constants of gel_tracking_reference (rows unused; outer_faces = n_outer_shock
 * + N1 + N2 + 1 on the reference's grid).  NULL t turns tracking off. */
int gel_set_mesh_tracking(gel_ctx* ctx, const gel_tracking* t, const gel_mesh_rails* r);

/* Node positions now (the reference's Point::x, y) into x, y[(ni+1)*(nj+1)]. */
int gel_get_nodes(gel_ctx* ctx, double* x, double* y);

/* Override the time step the next step uses (T[now1], Setka.cpp:10143-10145). */
int gel_set_dt(gel_ctx* ctx, double dt);

/* Rows of a ray each wavefront marches over (launch shape; the autotuner sets it). */
int gel_set_rows_per_wave(gel_ctx* ctx, int rows);

/* Block until queued work is done and fill *out (may be NULL). */
int gel_sync(gel_ctx* ctx, gel_stats* out);

/* Average device time in milliseconds per step over the steps timed since the last
 * call (HIP events on the context's stream); resets the accumulator.  *n_launches
 * gets the number of steps averaged. */
int gel_kernel_timing(gel_ctx* ctx, double* avg_ms, long long* n_launches);
/* Event timing: 0 off (default); 1 events around every step's launch (each interval
 * also holds that launch's dispatch latency); 2 events around every gel_step call, so
 * that a call of n back-to-back steps gives the per-step time the stream sustains. */
int gel_enable_timing(gel_ctx* ctx, int on);

/* Batched device evaluation of the face flux, for unit tests against fixtures:
 * same arguments and outputs as Setka::HLLC_2d_Korolkov_b_s (Setka.h:187-189),
 * one face per index.  Host arrays in and out; L and R are [n][5] as
 * (ro, Q, p, u, v); out P is [n][4], PQ, Vl, Vc, Vp, dt are [n]. */
int gel_flux_hllc2d_batch(int n, const double* L, const double* R, const double* W,
                          const double* n1, const double* n2, const double* rad,
                          double cfl, double gamma, int metod, int nul_potok,
                          double* P, double* PQ, double* Vl, double* Vc, double* Vp,
                          double* dt);

/* Batched device evaluation of the MHD face flux (HLLD after Solvers::HLLDQ_Korolkov
 * metod 3, Solvers.cpp:1642-2185, any in-plane normal, face speed W).  L, R: [n][9]
 * primitives (ro, p, u, v, w, bx, by, bz, Q); P: [n][9] = mass, momentum x/y/phi,
 * energy, B x/y/phi, tracer; smax: [n] max(|SL|,|SR|). */
int gel_flux_hlld_batch(int n, const double* L, const double* R, const double* W, const double* n1,
                        const double* n2, double gamma, double* P, double* smax);

/* The FP64 division and square root of the sweep's kernels, on the device, for checking
 * them against IEEE arithmetic (no counterpart in the reference, which divides and takes
 * square roots with the host's IEEE operations): q = a / b, q2[k] = a[(k + 1) % n] / b[k]
 * from b[k]'s reciprocal shared as the sweep shares it, s = sqrt(a).  All arrays [n]. */
int gel_fp64_div_sqrt(int n, const double* a, const double* b, double* q, double* q2, double* s);

/* The same flux with no device allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory, such as serving loops and harnesses
 * that forbid hipMalloc.  The inputs and outputs are staged through a fixed device
 * array of GEL_NOALLOC_MAX_FACES faces (8 MB) that each device holds for the life of
 * the process; n above that limit gives GEL_E_ARG.  Calls that overlap on one device
 * share the array and race: one call at a time per device. */
#define GEL_NOALLOC_MAX_FACES 32768
int gel_flux_hlld_batch_noalloc(int n, const double* L, const double* R, const double* W, const double* n1,
                                const double* n2, double gamma, double* P, double* smax);

/* Batched device evaluation of the exact Riemann solver for a gamma-gas with a moving
 * face, Solvers::Godunov_Solver_Alexashov with newton / devtwo / lev (Solvers.h:23-25,
 * Solvers.cpp:693-1321; the reference applies it to fitted-surface faces,
 * Setka.cpp:10335-10370).  L, R: [n][5] = (ro, p, u, v, w); nrm: [n][3] unit normals
 * (tangents (n2, -n1, 0) and (0, 0, 1), as the reference); W, dist: [n].  Out: q [n][5]
 * = mass, momentum x/y/z and energy fluxes relative to the face; speeds [n][3] = (dsl,
 * dsp, dsc); dt [n] = cfl * dist / max|wave speed|; status [n] = 0, or 1 where the
 * reference's pressure iteration would exit (no convergence in 1100 steps, or a
 * non-positive pressure), q then zero. */
int gel_flux_godunov_batch(int n, const double* L, const double* R, const double* nrm, const double* W,
                           const double* dist, double cfl, double gamma, double* q, double* speeds, double* dt,
                           int* status);

/* The same evaluation with one problem per wavefront, the form in which the
 * fitted-surface speeds are solved (the bisection's midpoints spread over the lanes);
 * the same outputs, bit for bit.  No counterpart of its own in the reference. */
int gel_flux_godunov_wave(int n, const double* L, const double* R, const double* nrm, const double* W,
                          const double* dist, double cfl, double gamma, double* q, double* speeds, double* dt,
                          int* status);

/* ---- batched cases (no counterpart in the reference, which runs one grid) ----
 * Many independent contexts (polar grids, not the inner region, same physics /
 * precision / flux / source mode) advanced together: one persistent launch per step
 * walks the (case, strip, chunk) work list (the plasma + hydrogen physics: one source
 * launch and one launch per fluid; fitted-surface tracking, all cases or none: the
 * batched surface-speed and rail launches first); each case keeps its own state,
 * buffers and CFL time step.  The cases are all static, or all move by prescribed breathing (gel_set_motion)
 * with one common period, each with its own amplitude and phase; all have C_1 /
 * C_centr cells (the grid's cell_type: gas, FP64) or none has.  Moving, tracked and
 * typed batches take GEL_SRC_NONE or GEL_SRC_MK1 (GEL_E_UNSUPPORTED otherwise).
 * With use_graph, whole cycles of 6 steps (of a common multiple of 6 and the motion
 * period, and of the refresh interval of an attached network) replay one captured HIP
 * graph and the rest of a call another, both kept for later calls of the same length
 * and refresh phase.  Each context may appear once (GEL_E_ARG otherwise, and for a null
 * entry). */
typedef struct gel_batch gel_batch;
int gel_batch_create(gel_batch** out, int ncases, gel_ctx* const* cases);
void gel_batch_destroy(gel_batch* b);
/* Rows per chunk (0: each case's default) and persistent workgroups (0: automatic). */
int gel_batch_set_launch(gel_batch* b, int rows, int workgroups);
/* Advance all cases nsteps steps; *ms (may be NULL) gets the device time. */
int gel_batch_step(gel_batch* b, int nsteps, int use_graph, double* ms);
/* Capture (without running anything) the graphs that gel_batch_step(b, nsteps, 1, .)
 * replays next, so that the call itself only launches them. */
int gel_batch_prepare(gel_batch* b, int nsteps);
/* Time candidate launch shapes on the cases (states restored) and keep the fastest. */
int gel_batch_autotune(gel_batch* b, int trial_steps, int* best_rows, int* best_workgroups, double* ms_per_step);

/* ---- learned charge-exchange source (no counterpart in the reference, whose
 * sources come from the Monte Carlo of Setka::MK_start_new, Setka.cpp:11989-12535,
 * through the closure Cell::Get_Sourse_MK1, Cell.cpp:498-665) ----
 * A 14 -> 4 x 128 SiLU -> 5 MLP with BF16 weights and activations on MFMA and FP32
 * accumulation.  Parameters are packed per layer as W [out][in] then b [out]. */
typedef struct gel_mlp gel_mlp;
int gel_mlp_create(gel_mlp** out, uint64_t seed);
void gel_mlp_destroy(gel_mlp* m);
int gel_mlp_param_count(void);
int gel_mlp_get_params(gel_mlp* m, float* p);
int gel_mlp_set_params(gel_mlp* m, const float* p);
/* y [n][5] = net(x [n][14]) on host arrays (normalised features). */
int gel_mlp_forward(gel_mlp* m, int n, const float* x, float* y);
/* Mean squared error and its gradient for a batch (n a multiple of 16). */
int gel_mlp_grad(gel_mlp* m, int n, const float* x, const float* t, float* grad, float* loss);
int gel_mlp_set_dataset(gel_mlp* m, int n, const float* x, const float* t);
/* Adam over the stored set, batch rows in order; loss_out[steps] may be NULL. */
int gel_mlp_train(gel_mlp* m, int steps, int batch, float lr, float* loss_out);
int gel_mlp_set_scaling(gel_mlp* m, const float* mu, const float* sd, const float* ymu, const float* ysd);
/* What the source model's outputs y (after the scaling) are.  GEL_MLP_OUT_SOURCE: the
 * five sources in units of src_coef * ro * n_H (n_H summed over the populations).
 * GEL_MLP_OUT_FACTOR: factors on the closure of Cell::Get_Sourse_MK1 taken without the
 * k factors, source = closure x (1 + y) -- the network stands in for the Monte Carlo's
 * k_u, k_v, k_T (Setka.cpp:12395-12449) as functions of the local state; its judge
 * compares with the closure times the k factors the context holds.  No reference
 * counterpart (the reference's k factors come from the last Monte Carlo run). */
enum gel_mlp_output { GEL_MLP_OUT_SOURCE = 0, GEL_MLP_OUT_FACTOR = 1 };
int gel_mlp_set_output_mode(gel_mlp* m, int mode);
const char* gel_mlp_last_error(void);
/* Raw features [ncells][14] and analytic sources [5][ncells] of a context's current
 * state and hydrogen moments (gel_set_sources). */
int gel_surrogate_features(gel_ctx* ctx, float* x);
int gel_source_analytic(gel_ctx* ctx, double* q);
/* Fill the context's source table (GEL_SRC_TABLE) from the network; with judge != 0
 * cells whose error against the analytic source exceeds tol fall back to it.  One
 * fused launch (features, five MFMA layers, judge) over tiles of 256 cells.
 * judge = 2 (GEL_SRC_HYBRID contexts only): cells already flagged keep the analytic
 * closure and are not evaluated again; only the others run the network and the judge
 * (a tile packs them, so that the network runs over ceil(live / 64) rows of 64).
 * n_fallback counts every flagged cell, kept or new. */
int gel_surrogate_refresh(gel_ctx* ctx, gel_mlp* m, double tol, int judge, long long* n_fallback);
/* The same for every case of a batch at once (their current states). */
int gel_batch_refresh(gel_batch* b, gel_mlp* m, double tol, int judge, long long* n_fallback);
/* Attach a network to a batch of GEL_SRC_TABLE cases: gel_batch_step then refreshes
 * every case's source table after every `every`-th step the batch advances from now
 * on, counted over all calls whatever their lengths (also inside the captured
 * graphs).  every = 0 detaches. */
int gel_batch_set_source_model(gel_batch* b, gel_mlp* m, int every, double tol, int judge);
/* Refreshes run by gel_batch_step and the cells that fell back in them, since the
 * last call (both reset). */
int gel_batch_source_stats(gel_batch* b, long long* refreshes, long long* fallback_cells);

/* ---- Monte Carlo of the interstellar hydrogen --------------------------------------
 * The trajectory tracer that gives the sweep its charge-exchange inputs: replaces
 * Setka::MK_start_new (Setka.cpp:11989-12535) with its tracer
 * Setka::Fly_exchenge_Imit_Korol (:12814-13981) and the moments' post-processing
 * (:12241-12449).  Atoms of a drifting Maxwellian (density 1, bulk velocity v_inf along
 * x, thermal speed 1) enter through the sphere around the grid; each flies straight,
 * cell by cell, and in every cell it crosses gives up the weight fraction
 * 1 - exp(-kappa) of the charge-exchange optical depth kappa to a secondary atom born at
 * a sampled point with a velocity drawn from the plasma's exchange kernel (the parent
 * flies on with the rest).  Secondaries lighter than w_min x a primary's weight survive
 * with probability weight / (w_min x primary) at that weight.  Population of a
 * secondary: 0 (C_1, C_centr cells), 1 (other), 2 (C_4), 3 (C_5); primaries are 3. */
typedef struct gel_mc_params {
    double v_inf;           /* hydrogen bulk velocity along x at infinity (Velosity_inf, Help.h:53) */
    double kn;              /* Knudsen number (Kn_, Help.h:60) */
    double sig_a;           /* a_2 of the cross section (1 - a_2 ln u)^2 (Help.h:63-68) */
    double n_p_lism;        /* scale of the source integrals (n_p_LISM_, Help.h:64) */
    double w_min;           /* roulette weight of secondaries, in primary weights */
    long long n_primary;    /* primary atoms */
    unsigned long long seed;
    int max_segments;       /* bound on the cell segments of one atom */
} gel_mc_params;

/* The reference's constants; w_min 0.25 (weight_, Help.h:93, the roulette of
 * Setka::Fly_exchenge, Setka.cpp:12654-12665), n_primary 1e6, seed 1, max_segments 1e6. */
void gel_mc_params_reference(gel_mc_params* p);

typedef struct gel_mc_stats {
    long long primaries_in;  /* primaries that reached the grid */
    long long atoms;         /* trajectories traced on the grid (primaries + secondaries kept) */
    long long segments;      /* cell segments flown */
    long long lost;          /* atoms whose way out of a cell could not be resolved */
    long long dropped;       /* secondaries dropped for want of queue room (must be 0) */
    int generations;         /* secondary generations traced */
    double ms;               /* device time */
} gel_mc_stats;

/* Per-cell outputs of gel_mc_run, GEL_MC_NFIELDS variable-major arrays of ncells:
 *   0..18   H_n[4], H_u[4], H_v[4], H_T[4], k_u, k_v, k_T: the GEL_MK1_NFIELDS source
 *           inputs of the sweep (Cell.h:55-66), as Setka.cpp:12247-12449 leaves them;
 *   19..21  I_u, I_v, I_T: momentum and energy sources from the rate integrals;
 *   22..24  II_u, II_v, II_T: the same from the sampled exchanges;
 *   25..28  F_n, F_u, F_v, F_T: moments of all populations together. */
#define GEL_MC_NFIELDS 29

/* Run the Monte Carlo on a context's grid and plasma state (a polar grid, gas or MHD, or
 * a general mesh -- the reference's own grid -- at rest: rho, p, u, v are read).  On a
 * mesh the atoms walk each cell's face slots as Setka::Fly_exchenge_Imit_Korol walks
 * Cell::Grans (Setka.cpp:12893-12960): Axis faces skipped, a face with a Sosed leads to
 * it, an Inner_sphere face with none is crossed in a straight line to another, any other
 * boundary face leaves the grid (primaries enter by those).  out (NULL: none) receives
 * GEL_MC_NFIELDS x ncells; a GEL_SRC_MK1 context also gets fields 0..18 as its source
 * inputs, as from gel_set_sources. */
int gel_mc_run(gel_ctx* ctx, const gel_mc_params* p, double* out, gel_mc_stats* st);

/* Point-wise pieces of the Monte Carlo on the device, for checks against the
 * reference.  Rates at n (u, cp): out[n][6] = Setka::Velosity_1..3 (Setka.cpp:15830-15866)
 * and MKmethod::int_1..3 (MKmethod.cpp:2741-2858; here from the integrals' definition,
 * 0 where u / cp > 7). */
int gel_mc_rates(int n, const double* u, const double* cp, double* out);
/* The exchange-velocity sampler of MKmethod::Change_Velosity4 with no splitting zones
 * (MKmethod.cpp:2270-2531, I = 0) driven by the reference's random sensor
 * (sensor.cpp:11-31): in[n][7] = Ur Uthe Uphi Vr Vthe Vphi cp (velocities in units of
 * cp), seeds[n][3]; reps draws per input into out[n][reps][4] = Wr Wthe Wphi mu, and the
 * sensors' final states into state[n][3]. */
int gel_mc_exchange(int n, const double* in, const int* seeds, int reps, double* out, int* state);

const char* gel_strerror(int status);
const char* gel_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GELIOS_H */
