/* C interface of the CPU oracle (test infrastructure only; see
 * spade_oracle.cpp).  Arrays are in the SPADE reference layout: AoS with the
 * variable index fastest and ghost layers stored, blocks x-fastest. */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_MAP_IDENTITY = 0, ORA_MAP_TANH = 1 };

typedef struct
{
    int kind;           /* ORA_MAP_* */
    double p0, p1;      /* computational interval the map is stretched over */
    double beta;        /* tanh stretching factor */
} ora_map1d;

typedef struct
{
    int nb[3];          /* blocks per axis */
    int nc[3];          /* cells per block per axis */
    int ng;             /* ghost layers */
    double lo[3], hi[3];
    int periodic[3];
    ora_map1d map[3];
} ora_grid;

typedef struct { double gamma, R; } ora_gas;

enum { ORA_TERM_KEEP = 1, ORA_TERM_TOTANI, ORA_TERM_PDISS, ORA_TERM_WENO,
       ORA_TERM_FWENO, ORA_TERM_FIRST, ORA_TERM_VISC, ORA_TERM_HYBRID, ORA_TERM_WENO_FDS };
enum { ORA_SENSOR_CONST = 0, ORA_SENSOR_DUCROS = 1 };
enum { ORA_SPLIT_RUSANOV = 0, ORA_SPLIT_PHYS = 1 };
enum { ORA_LAW_CONST = 0, ORA_LAW_POWER = 1 };
enum { ORA_SGS_NONE = 0, ORA_SGS_WALE = 1 };
#define ORA_MAX_TERMS 4

typedef struct
{
    int kind;
    int order;                  /* KEEP */
    int sensor;                 /* PDISS */
    double sensor_param;        /* const alpha, or Ducros epsilon */
    double eps_p, eps_T;        /* PDISS */
    int split;                  /* WENO */
    int smooth;                 /* WENO, FWENO, WENO_FDS: 1 = nonlinear weights (enable_smooth) */
    int law;                    /* VISC */
    double mu, T_ref, power, prandtl;
    int diss_kind;              /* HYBRID: dissipative scheme, ORA_TERM_WENO or ORA_TERM_FWENO */
    int full;                   /* HYBRID: 1 = (1-b) central + b dissipative, 0 = central + b dissipative */
    int sgs;                    /* VISC: ORA_SGS_WALE = sgs_visc_t(law, wale_t(cw, delta, prt)) */
    double sgs_cw, sgs_delta, sgs_prt;
} ora_term;

typedef struct { int nterms; ora_term terms[ORA_MAX_TERMS]; } ora_flux;

typedef struct { int wall_mask; double t_wall; } ora_bc;

int ora_array_size(const ora_grid* g);
/* threads of the OpenMP loops over blocks (flux_div); returns the count now in use */
int ora_set_threads(int n);
int ora_flux_div(const ora_grid* g, const ora_gas* gas, const ora_flux* fl,
                 const double* prim, double* rhs, int increment);
/* flux_div on the interior cells of the listed blocks only (the others are
 * left untouched): the same per-cell arithmetic as ora_flux_div */
int ora_flux_div_blocks(const ora_grid* g, const ora_gas* gas, const ora_flux* fl, const double* prim, double* rhs,
                        int increment, const int* blocks, int nblocks);
int ora_exchange(const ora_grid* g, double* a);
int ora_boundary_fill_wall(const ora_grid* g, double* a, int face, double t_wall);
/* boundary_fill(arr, faces, extrapolate<order>); face_mask bit 2*dir+pm */
int ora_boundary_fill_extrap(const ora_grid* g, double* a, int face_mask, int order);
/* boundary_fill(arr, faces, kernel) for an affine kernel: per primitive v,
 * ghost = dom[v] image + fill[v] ghost + add[v]; the face-normal velocity
 * component takes the *_normal coefficients */
typedef struct
{
    double dom[5], fill[5], add[5];
    double dom_normal, fill_normal, add_normal;
} ora_fill_kernel;
int ora_boundary_fill_affine(const ora_grid* g, double* a, int face_mask, const ora_fill_kernel* kern);
int ora_prim_to_cons(const ora_gas* gas, const double* prim, double* cons, long ncell);
int ora_cons_to_prim(const ora_gas* gas, const double* cons, double* prim, long ncell);
int ora_ssprk3_step(const ora_grid* g, const ora_gas* gas, const ora_flux* fl, const ora_bc* bc,
                    double dt, double* q, double* r0, double* r1, double* r2);
int ora_ssprk3opt_step(const ora_grid* g, const ora_gas* gas, const ora_flux* fl, const ora_bc* bc,
                       double dt, double* q, double* r0, double* r1);

/* explicit Butcher table (time-integration/explicit.h:27-116) as ratios
 * {num, den}: a[i][j] (j < i), b[j], c[i]; a zero den means 0 */
#define ORA_RK_MAX 6
typedef struct
{
    int stages;
    int high_storage;            /* two solution copies (var_size 2) */
    long a[ORA_RK_MAX][ORA_RK_MAX][2];
    long b[ORA_RK_MAX][2];
    long c[ORA_RK_MAX][2];
} ora_rk_table;
/* one step, prim<->cons transform (advance.h:233-280); res: stages arrays */
int ora_rk_transform_step(const ora_grid* g, const ora_gas* gas, const ora_flux* fl, const ora_bc* bc,
                          const ora_rk_table* tb, double dt, double* q, double* const* res);
/* one step, identity transform on conserved storage (advance.h:107-230) */
int ora_rk_identity_step(const ora_grid* g, const ora_gas* gas, const ora_flux* fl, const ora_rk_table* tb,
                         double dt, double* q, double* q2, double* const* res, double* prim);

#ifdef __cplusplus
}
#endif
