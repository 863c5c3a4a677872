/* spade_amd: MI355X-native right-hand-side hot path of SPADE.
 *
 * C ABI of libspade_amd.so.  Plain pointers, sizes and status codes only.
 * Every entry point returns SPD_OK (0) or an SPD_ERR_* code; the message of
 * the last failure on the calling thread is spd_last_error().
 *
 * Host arrays passed to upload/download use the reference storage layout:
 * AoS with the variable index fastest, ghost layers stored, blocks in x-fastest
 * order:  offset = v + 5*((i+ng) + Nx*((j+ng) + Ny*((k+ng) + Nz*lb))),
 * N* = n* + 2 ng  (SPADE core/mem_map.h:466-497, 654-676).
 * Primitive order (p, T, u, v, w); flux/residual order (continuity, energy,
 * x-, y-, z-momentum) (navier-stokes/fluid_state.h:11-77).
 */
#ifndef SPADE_AMD_H
#define SPADE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SPD_OK = 0,
    SPD_ERR_ARG = 1,          /* bad argument / size / configuration (reference: except::sp_exception, core/except.h:12-30) */
    SPD_ERR_HIP = 2,          /* HIP runtime or launch failure (reference: CUDA error check after every launch, dispatch/execute.h:85-96) */
    SPD_ERR_UNSUPPORTED = 3,  /* combination the library does not implement */
    SPD_ERR_MEMORY = 4,       /* device allocation failure (reference: mem_exception, dispatch/device_allocator.h:30-44) */
    SPD_ERR_IO = 5            /* a file could not be opened, read or written (the reference does not check its streams) */
};

typedef struct spd_ctx_s*        spd_ctx_t;
typedef struct spd_grid_s*       spd_grid_t;
typedef struct spd_array_s*      spd_array_t;
typedef struct spd_integrator_s* spd_integrator_t;

/* Message of the last failure on this thread ("" if none). */
const char* spd_last_error(void);

/* ---- context: one device, one HIP stream; every call is stream-ordered ---- */
int spd_ctx_create(int device, spd_ctx_t* out);
int spd_ctx_destroy(spd_ctx_t ctx);
int spd_ctx_sync(spd_ctx_t ctx);
/* The hipStream_t all work of this context is issued on. */
void* spd_ctx_stream(spd_ctx_t ctx);
/* HIP-event timer around work on the context stream (milliseconds). */
int spd_timer_start(spd_ctx_t ctx);
int spd_timer_stop(spd_ctx_t ctx, float* ms);

/* ---- grid: cartesian_grid_t + cartesian_blocks_t
 *      (grid/cartesian_grid.h:84-96, grid/cartesian_blocks.h:38-107) ---- */
enum { SPD_MAP_IDENTITY = 0, SPD_MAP_TANH = 1 };
typedef struct {
    int    kind;        /* SPD_MAP_* */
    double p0, p1;      /* computational interval the map stretches over */
    double beta;        /* tanh stretching factor: x = c + h tanh(beta s)/tanh(beta) */
} spd_map1d;

typedef struct {
    int       num_blocks[3];
    int       cells_in_block[3];
    int       num_exch;          /* ghost layers (same on every axis) */
    double    lo[3], hi[3];      /* computational bounds */
    int       periodic[3];
    spd_map1d map[3];            /* identity or diagonal (1-D per axis) coordinates */
} spd_grid_desc;

int spd_grid_create(spd_ctx_t ctx, const spd_grid_desc* desc, spd_grid_t* out);
int spd_grid_destroy(spd_grid_t grid);

/* ---- grid_array: 5-variable FP64 cell-centred array with ghosts
 *      (grid/grid_array.h:182-422).  Device layout is block-major SoA
 *      [block][var][k][j][i] with ghosts. ---- */
int spd_array_create(spd_grid_t grid, spd_array_t* out);
int spd_array_destroy(spd_array_t arr);
int spd_array_upload(spd_array_t arr, const double* host);    /* reference layout */
int spd_array_download(spd_array_t arr, double* host);        /* reference layout */
/* arr = value on every stored cell (grid_array::operator=(scalar), grid_array.h:289-395) */
int spd_array_fill(spd_array_t arr, double value);
int spd_array_copy(spd_array_t dst, spd_array_t src);
/* Number of doubles of the reference-layout host image of an array. */
long spd_array_host_size(spd_grid_t grid);
/* Number of doubles one array occupies on the device: per block and
 * variable the stored cells (interior-x rows, then the x-ghost sections),
 * blocks padded to 256 bytes; no counterpart in the reference. */
long spd_array_device_size(spd_grid_t grid);

/* ---- flux functors (composed as F0 + (F1 + ...), omni/compose.h:7-45) ---- */
enum {
    SPD_TERM_KEEP = 1,    /* cent_keep<order>            navier-stokes/convective.h:96-192 */
    SPD_TERM_TOTANI,      /* totani_lr                   convective.h:54-94 */
    SPD_TERM_PDISS,       /* pressure_diss_lr            convective.h:194-246 */
    SPD_TERM_WENO,        /* weno_t<rusanov|phys_flux>   convective.h:255-333, flux_funcs.h */
    SPD_TERM_FWENO,       /* fweno_t                     convective.h:335-497 */
    SPD_TERM_FIRST,       /* first_order_t (rusanov)     convective.h:14-33 */
    SPD_TERM_VISC,        /* visc_lr                     navier-stokes/viscous.h:13-113 */
    SPD_TERM_HYBRID       /* hybrid_scheme_t: cent_keep<order> blended with weno_t / fweno_t by a face
                             sensor (navier-stokes/hybrid_scheme.h:14-48, state_sensor.h:21-60); the
                             descriptor is defined, but spd_flux_div refuses it (SPD_ERR_UNSUPPORTED):
                             the library has no device code for it */ ,
    SPD_TERM_WENO_FDS     /* weno_fds_t<rusanov_fds_t>      convective.h:499-555, flux_funcs.h:54-95 (smooth
                             selects enable_smooth / disable_smooth) */
};
enum { SPD_SENSOR_CONST = 0, SPD_SENSOR_DUCROS = 1 };   /* state_sensor.h:21-60 */
enum { SPD_SPLIT_RUSANOV = 0, SPD_SPLIT_PHYS = 1 };     /* flux_funcs.h:8-126 */
enum { SPD_LAW_CONST = 0, SPD_LAW_POWER = 1 };          /* viscous_laws.h:62-135 */
enum { SPD_SGS_NONE = 0, SPD_SGS_WALE = 1 };            /* subgrid_scale.h:24-86 */
#define SPD_MAX_TERMS 4

typedef struct {
    int    kind;
    int    order;                    /* KEEP: even nominal order (2 or 4 with 2 ghosts) */
    int    sensor;                   /* PDISS */
    double sensor_param;             /* const: alpha; Ducros: epsilon */
    double eps_p, eps_T;             /* PDISS */
    int    split;                    /* WENO */
    int    smooth;                   /* WENO/FWENO/WENO_FDS: 1 = nonlinear weights (enable_smooth) */
    int    law;                      /* VISC */
    double mu, T_ref, power, prandtl;/* VISC: constant mu, or power law mu_ref (T/T_ref)^power */
    int    diss_kind;                /* HYBRID: dissipative scheme, SPD_TERM_WENO or SPD_TERM_FWENO (with
                                        split / smooth as for those terms); order = the central cent_keep order;
                                        sensor / sensor_param = the blending sensor */
    int    full;                     /* HYBRID: 1 = full_flux ((1 - b) F_central + b F_diss),
                                        0 = diss_flux (F_central + b F_diss) (hybrid_scheme.h:7-12, 36-41) */
    int    sgs;                      /* VISC: SPD_SGS_WALE = sgs_visc_t(law, wale_t(gas, sgs_cw, sgs_delta,
                                        sgs_prt)) (viscous_laws.h:174-216, subgrid_scale.h:24-86): mu + mu_t,
                                        beta - 0.66666666667 mu_t, alpha + mu_t / sgs_prt */
    double sgs_cw, sgs_delta, sgs_prt;
} spd_flux_term;

typedef struct {
    int           nterms;
    spd_flux_term terms[SPD_MAX_TERMS];
    double        gamma, R;          /* ideal_gas_t (navier-stokes/gas.h:35-49) */
} spd_flux_desc;

enum { SPD_OVERWRITE = 0, SPD_INCREMENT = 1 };   /* pde-algs/pde_traits.h:7-20 */

/* rhs (+)= -div F on interior cells: pde_algs::flux_div(prims, rhs, flux, traits)
 * (pde-algs/flux-div/flux_div.h:16-56, flux_div_basic.h:17-76).  prims must have
 * current ghosts.  Overwrite mode writes every interior cell of rhs; its ghost
 * cells are left as they are (the reference's rhs = 0 also zeroes them; no
 * kernel here reads residual ghosts). */
int spd_flux_div(spd_array_t prims, spd_array_t rhs, const spd_flux_desc* flux, int mode);

/* Which kernel spd_flux_div would run for this array and flux: *kind = 1 for
 * the LDS-staged face-once kernel (cent_keep<2> [+pressure_diss_lr]
 * [+visc_lr], identity coordinates, 16x16 block cross-section), 2 for the
 * one-thread-per-cell kernel specialised for such a composition on any other
 * grid (mapped coordinates included), 0 for the generic one-thread-per-cell
 * kernel.  No counterpart in the reference (it has
 * a single algorithm per trait); for tests and benchmarks. */
int spd_flux_div_kernel(spd_array_t prims, const spd_flux_desc* flux, int* kind);

/* Ghost exchange between blocks (create_exchange + exchange,
 * grid/exchange_pattern.h:304-424): periodic axes wrap, domain faces of
 * non-periodic axes are left to the boundary fill. */
int spd_exchange(spd_array_t arr);

/* boundary_fill with the no-slip isothermal mirror kernel
 * ghost = (p, 2 T_wall - T, -u, -v, -w) of cell -1-i / 2n-(i+1)
 * (grid/boundary_fill.h:32-133).  face = 2*dir + (0: min, 1: max). */
int spd_boundary_fill_wall(spd_array_t arr, int face, double t_wall);

/* Faces of a boundary_fill call: bit (2*dir + pm), pm 0 = min, 1 = max; the
 * faces are filled in face order xmin, xmax, ymin, ymax, zmin, zmax (later
 * faces read the ghosts earlier ones wrote at edges and corners), on the
 * blocks whose face lies on the domain boundary, every ghost layer beyond the
 * face over the whole stored tangential range (grid/boundary_fill.h:32-60;
 * the boundary::identifier_t of boundary_fill.h:9-15). */
enum { SPD_XMIN = 1, SPD_XMAX = 2, SPD_YMIN = 4, SPD_YMAX = 8, SPD_ZMIN = 16, SPD_ZMAX = 32 };

/* boundary_fill(arr, faces, boundary::extrapolate<order>) (boundary_fill.h:
 * 17-26, 63-98): Lagrange extrapolation along the face normal through the
 * order + 1 interior cells nearest a min face / farthest cells ending at n - 1
 * of a max face.  Needs order + 1 <= cells per block along each filled axis. */
int spd_boundary_fill_extrap(spd_array_t arr, int face_mask, int order);

/* boundary_fill(arr, faces, kernel) with a user kernel (boundary_fill.h:
 * 99-127), kern(q_image, q_ghost, x, idir) or kern(q_image, idir).  An
 * arbitrary callable cannot cross a C ABI, so the kernel is given as data:
 * per primitive variable v (p, T, u, v, w)
 *     ghost[v] = dom[v] * q_image[v] + fill[v] * q_ghost[v] + add[v]
 * where q_image is the mirror cell (-1 - i at a min face, 2n - (i + 1) at a
 * max face) and q_ghost the ghost's current value; the velocity component
 * normal to the face takes dom_normal / fill_normal / add_normal instead.  A
 * term with a zero coefficient is left out; the others are rounded in the
 * order written (so e.g. the no-slip isothermal wall, the slip wall, a fixed
 * state or a zero-gradient copy give the bits of the same kernel written as a
 * reference lambda).  Kernels that depend on the ghost's coordinates x have no
 * descriptor. */
typedef struct
{
    double dom[5], fill[5], add[5];
    double dom_normal, fill_normal, add_normal;
} spd_fill_kernel;
int spd_boundary_fill(spd_array_t arr, int face_mask, const spd_fill_kernel* kern);

/* ---- time integration: integrator_t(axis, scheme, data, rhs, bc, trans).advance()
 *      (time-integration/integrator.h:18-59) with rhs = flux_div(overwrite),
 *      bc = exchange + walls, trans = prim<->cons state transform. ---- */
enum { SPD_SCHEME_SSPRK3 = 0,       /* ssprk3_t, transform-aware advance, advance.h:233-280 */
       SPD_SCHEME_SSPRK3_OPT = 1 }; /* ssprk3_opt low storage, advance.h:354-402 */
typedef struct { int wall_mask; double t_wall; } spd_bc_desc;  /* bit (2*dir+pm): isothermal wall */

int spd_integrator_create(spd_array_t q, const spd_flux_desc* flux, const spd_bc_desc* bc,
                          int scheme, double dt, spd_integrator_t* out);
int spd_integrator_advance(spd_integrator_t it, int nsteps);
int spd_integrator_destroy(spd_integrator_t it);
/* Residual array k_i of the integrator (0-based), for inspection. */
spd_array_t spd_integrator_residual(spd_integrator_t it, int i);
/* *fused = 1 when spd_integrator_advance runs each SSP-RK3 stage as one
 * launch of flux_div fused with the stage update (the fast kernel's
 * compositions on 16 x 16 x nz blocks, nz >= 2 ng, ng <= 2, every axis
 * periodic or walled on both faces; the step then ping-pongs between the
 * state and a second buffer the integrator owns, and the array handle holds
 * the new state after each step), 0 when it runs separate flux_div and stage
 * launches.  The results are the same bits either way.  No counterpart in
 * the reference (integrator.h:18-59 has one path). */
int spd_integrator_fused(spd_integrator_t it, int* fused);

/* ---- explicit Runge-Kutta with any Butcher table and caller-supplied
 *      residual / boundary callbacks: integrator_t(axis, scheme, data, rhs, bc,
 *      trans).advance() (time-integration/integrator.h:18-59, advance.h:107-280,
 *      explicit.h:27-116) ---- */
#define SPD_RK_MAX_STAGES 6
typedef struct { long num, den; } spd_ratio;       /* std::ratio<num, den>; den = 0 reads as 0 */
typedef struct {
    int       stages;
    int       high_storage;                         /* rk_t<..., true>: a second solution copy (var_size 2) */
    spd_ratio a[SPD_RK_MAX_STAGES][SPD_RK_MAX_STAGES];/* table rows: a[i][j], j < i */
    spd_ratio b[SPD_RK_MAX_STAGES];                 /* accumulation (accum_type) */
    spd_ratio c[SPD_RK_MAX_STAGES];                 /* stage times (dt_type) */
} spd_rk_table;
/* The tables explicit.h defines: rk2_t, rk2hs_t, rk4_t, ssprk34_t, ssprk3_t, ssprk3hs_t, rk38r_t. */
enum { SPD_RK2 = 0, SPD_RK2HS, SPD_RK4, SPD_SSPRK34, SPD_SSPRK3_TABLE, SPD_SSPRK3HS, SPD_RK38R };
int spd_rk_table_builtin(int which, spd_rk_table* out);

/* The state transform of the integrator: state_transform_t(cons_t, gas) (primitive
 * storage, conserved stage arithmetic; navier-stokes/state_transformation.h:37-66),
 * which selects the transform-aware advance (advance.h:233-280), or
 * identity_transform (integrator.h:8-14), which selects the generic advance
 * (advance.h:107-230) on whatever variables the array stores. */
enum { SPD_TRANS_PRIM_CONS = 0, SPD_TRANS_IDENTITY = 1 };
/* rhs(r&, const q&, t) and bc(q&, t) of integrator_t; a non-zero return stops the
 * step with SPD_ERR_ARG.  They are called on the integrator's own host thread and
 * may call any function of this library on the arrays they are given. */
typedef int (*spd_rhs_fn)(spd_array_t resid, spd_array_t sol, double t, void* user);
typedef int (*spd_bc_fn)(spd_array_t sol, double t, void* user);
typedef struct {
    spd_rk_table        table;
    int                 trans;          /* SPD_TRANS_* */
    double              gamma, R;       /* gas of SPD_TRANS_PRIM_CONS */
    double              t0, dt;         /* time_axis_t(t0, dt) */
    spd_rhs_fn          rhs;            /* NULL: flux_div(sol, resid, *flux, overwrite) */
    spd_bc_fn           bc;             /* NULL: exchange + the isothermal walls of *walls (walls may be NULL) */
    void*               user;
    const spd_flux_desc* flux;          /* for the default rhs */
    const spd_bc_desc*  walls;          /* for the default bc */
} spd_rk_desc;
/* An integrator on q (advanced in place) with stages residual arrays (and a second
 * solution copy for a high-storage table with the identity transform);
 * spd_integrator_advance / _destroy / _residual apply to it. */
int spd_rk_integrator_create(spd_array_t q, const spd_rk_desc* desc, spd_integrator_t* out);
/* Current time of the integrator's axis (t0 + steps * dt, summed as the reference does). */
int spd_integrator_time(spd_integrator_t it, double* t);

/* One transform-aware stage update (transform_advance_to, advance.h:56-102) on the
 * interior cells of prims: cons = p2c(q); cons += coef[j] * resid[j] for each
 * non-zero coef[j] in order; q = c2p(cons), where coef[j] is the caller's
 * (a_ij - a_{i-1,j}) dt.  push_ghosts = 1 also writes each updated cell into the
 * ghost cells of the blocks it feeds (what spd_exchange would copy afterwards; the
 * domain faces of non-periodic axes are left for the boundary fill). */
int spd_rk_stage(spd_array_t prims, const spd_array_t* resid, const double* coef, int n, double gamma, double R,
                 int push_ghosts);

/* Element-wise array operations of grid_array (grid/grid_array.h:286-396): scale
 * every stored cell (operator*=(scalar)); dst += src / dst -= src on the interior
 * cells (operator+= / -=, support_of(exclude_exchanges)). */
int spd_array_scale(spd_array_t arr, double s);
int spd_array_add(spd_array_t dst, spd_array_t src);
int spd_array_sub(spd_array_t dst, spd_array_t src);
/* convert_state in place (navier-stokes/fluid_state.h:102-135): SPD_TO_CONS
 * (prim -> cons) or SPD_TO_PRIM (cons -> prim), on every stored cell
 * (include_ghosts = 1, the state transform's include_exchanges) or the interior. */
enum { SPD_TO_CONS = 0, SPD_TO_PRIM = 1 };
int spd_array_convert(spd_array_t arr, int direction, double gamma, double R, int include_ghosts);

/* ---- reductions on interior cells (transform_reduce, algs/transform_reduce.h:53-190) ---- */
/* max(|u| + sqrt(gamma R T)): the CFL wave speed (development/cuda-tgv/main.cc:155-172) */
int spd_max_wavespeed(spd_array_t prims, double gamma, double R, double* out);
/* The same wave speed left in device memory, allocating nothing: no counterpart
 * in the reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  d_out: one double of device memory on the
 * array's device; the call is stream-ordered on the context's stream and
 * returns without waiting.  The per-block partials live in a static device
 * array of 32768 doubles (256 KiB per device, held until the process ends):
 * grids of more than 32768 blocks are refused (SPD_ERR_ARG).  All calls on one
 * device share that array, so calls that overlap on one device (other streams,
 * other threads) race; calls ordered on one stream do not. */
int spd_max_wavespeed_noalloc(spd_array_t prims, double gamma, double R, double* d_out);

/* Volume averages over interior cells of a primitive array with current ghosts:
 * kinetic energy <rho |u|^2 / 2> (rho = p/(R T)) and enstrophy <|omega|^2 / 2>,
 * omega from second-order central differences.  Summed per block in a fixed
 * order, then over blocks in block order (deterministic).  No counterpart in
 * the reference (it has no such diagnostics). */
int spd_diagnostics(spd_array_t prims, double R, double* kinetic_energy, double* enstrophy);

/* ---- batched engine: many independent periodic cases stepped together ------
 * FP32 face fluxes about each case's reference state, FP64 conserved state,
 * FP64 residual sum and SSP-RK3 (Shu-Osher) stage combination.  One persistent
 * kernel per RK stage walks a work list of 16^3 blocks of all active cases; a
 * whole step is replayed from a hipGraph.  No counterpart in the reference (it
 * steps one grid at a time); it serves the same physics as
 * cent_keep<2> + visc_lr on identity coordinates. */
typedef struct spd_batch_s* spd_batch_t;
enum { SPD_CASE_TGV = 0, SPD_CASE_VORTEX = 1 };
typedef struct {
    int    kind;            /* SPD_CASE_TGV: [0,2pi]^3 Taylor-Green; SPD_CASE_VORTEX: [0,10]^3 isentropic vortex */
    int    cells[3];        /* cells per axis, positive multiples of 16 */
    double mach, reynolds;
    double gamma, prandtl;
    int    law;             /* SPD_LAW_CONST or SPD_LAW_POWER (must match the batch) */
    double power;           /* power-law exponent */
    double cfl;             /* dt = cfl dx / max(|u| + c) of the initial state, fixed */
} spd_case_desc;

/* Pool of block slots sized by hbm_budget_bytes (3 FP64 state copies per block). */
int spd_batch_create(spd_ctx_t ctx, long long hbm_budget_bytes, int law, spd_batch_t* out);
int spd_batch_destroy(spd_batch_t b);
/* Blocks a case needs, and the slots still free. */
int spd_batch_capacity(spd_batch_t b, const spd_case_desc* d, long long* blocks_needed, long long* blocks_free);
/* Adds a case (initial state computed on the device); fails with SPD_ERR_ARG when it does not fit. */
int spd_batch_add_case(spd_batch_t b, const spd_case_desc* d, int* case_id);
int spd_batch_remove_case(spd_batch_t b, int case_id);
/* Advance every active case by nsteps SSP-RK3 steps. */
int spd_batch_step(spd_batch_t b, int nsteps);
int spd_batch_case_info(spd_batch_t b, int case_id, long long* cells, int* steps_done, double* dt);
/* Health of a case: *diverged_step = -1 while every recorded kinetic energy is
 * finite, else the first step whose starting state had a non-finite one (the
 * reference's driver stops on a NaN wave speed, development/cuda-tgv/main.cc:
 * 250-259; here the case can be retired while the others go on: cases never
 * read each other's blocks). */
int spd_batch_case_status(spd_batch_t b, int case_id, int* diverged_step);
/* Conserved state (rho, rhoE, rho u, rho v, rho w), [block][var][k][j][i], 16^3 per block. */
int spd_batch_case_download(spd_batch_t b, int case_id, double* host);
int spd_batch_case_upload(spd_batch_t b, int case_id, const double* host);
/* Per-step history of the volume-averaged kinetic energy and enstrophy of the
 * state at the start of each step; *n = number of entries written (<= max). */
int spd_batch_case_diag(spd_batch_t b, int case_id, double* ke, double* enstrophy, int max, int* n);
/* Autotuner: times the launch shapes on the current batch (state is restored)
 * and keeps the fastest: stage kernel shape (rows of a block's 16x16
 * cross-section per workgroup 16|8, planes of state in flight 1|2, registers
 * sized for 3|2 waves per SIMD), persistent workgroups (1..4 blocks' worth per
 * CU) and work order.  Returns the chosen persistent grid and work order; the
 * kernel shape is read with spd_batch_get_shape.  Every shape gives the same
 * bits. */
int spd_batch_tune(spd_batch_t b, int* grid, int* interleaved);
/* 1: replay each step from a hipGraph (default); 0: launch the kernels directly. */
int spd_batch_use_graph(spd_batch_t b, int enable);
/* Launch shape in use (persistent workgroups, work order). */
int spd_batch_set_shape(spd_batch_t b, int grid, int interleaved);
/* Stage kernel shape in use: rows = rows of a block's 16x16 cross-section per
 * workgroup (16: one 256-thread workgroup per block, 8: two 128-thread ones),
 * prefetch = planes of state in flight ahead of the LDS window (1 or 2),
 * occupancy = waves per SIMD its registers are sized for (3 or 2).  Other
 * values are an argument error.  The persistent grid (spd_batch_set_shape)
 * counts workgroups of the chosen size. */
int spd_batch_set_kernel_shape(spd_batch_t b, int rows, int prefetch, int occupancy);
/* The launch shape in use (after spd_batch_tune, the chosen one). */
int spd_batch_get_shape(spd_batch_t b, int* grid, int* interleaved, int* rows, int* prefetch, int* occupancy);
/* One candidate the last spd_batch_tune timed. */
typedef struct
{
    int grid, interleaved, rows, prefetch, occupancy;
    float ms_per_step;
} spd_tune_entry;
/* The candidates of the last spd_batch_tune in the order they were timed;
 * *n = number written (<= max). */
int spd_batch_tune_log(spd_batch_t b, spd_tune_entry* out, int max, int* n);

/* ---- on-disk formats: spade::io (io/io_vtk.h, io/io_native.h) ---- */

/* io::output_vtk(out_dir, basename, array) (io/io_vtk.h:276-288): writes
 * <out_dir>/<basename>.pvts (a PStructuredGrid whose pieces stack the blocks
 * along z, io_vtk.h:131-174) and one StructuredGrid file per block,
 * <out_dir>/data_<basename>/b<9-digit block>.vts (io_vtk.h:176-273): interior
 * cell data per variable (i fastest) and the (n+1)^3 node coordinates x =
 * map(box min + i dx), Float64, inline binary as base64 of a 4-byte byte
 * count followed by base64 of the data (core/base_64.h:63-69).  names: the 5
 * variable names, NULL for the primitive names P,T,U,V,W (prim_t::name,
 * navier-stokes/fluid_state.h:28-32).  Creates the two directories if
 * missing. */
int spd_io_output_vtk(spd_array_t arr, const char* out_dir, const char* basename, const char* const* names);
/* The same output for an array image in host memory (reference layout, as
 * spd_array_download writes it) on the grid desc describes: a CPU-resident
 * grid_array in the reference.  Needs no device. */
int spd_io_output_vtk_host(const spd_grid_desc* desc, const double* host, const char* out_dir, const char* basename,
                           const char* const* names);
/* io::binary_write(filename, array) (io/io_native.h:19-45): every block's
 * stored cells, ghosts included, at byte offset lb * block_bytes, i.e. the
 * array's host image in the reference layout. */
int spd_io_binary_write(spd_array_t arr, const char* filename);
/* io::binary_read(filename, array) (io/io_native.h:47-55); fails with
 * SPD_ERR_IO unless the file holds exactly the array's host image size. */
int spd_io_binary_read(spd_array_t arr, const char* filename);

#ifdef __cplusplus
}
#endif
#endif
