// Internal types shared by the HIP translation units of libspade_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <cstdint>
#include "../../include/spade_amd.h"

namespace spd {

void set_error(const std::string& msg);

struct HipError { hipError_t code; std::string where; };

#define SPD_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) throw ::spd::HipError{e_, std::string(#call) + " @ " +    \
                                                    __FILE__ + ":" + std::to_string(__LINE__)}; \
    } while (0)

#define SPD_LAUNCH_CHECK() SPD_HIP(hipGetLastError())

struct ArgError { std::string msg; };

// Geometry as the kernels see it.  Per-block dx / 1/dx / box minimum are
// tabulated on the host exactly as the reference computes them
// (grid/cartesian_grid.h:116-145) because dx = ((min+bsize)-min)/n differs in
// the last bit between blocks.
struct DevGeom
{
    int nc[3];        // cells per block
    int ng;           // ghost layers
    int N[3];         // stored extent nc + 2 ng
    int nb[3];        // blocks per axis
    int nblk;
    long long sv;     // doubles between variables of one block (N0 N1 N2)
    long long sb;     // doubles between blocks (5 sv, padded to 256 bytes)
    long long xh;     // start of the x-halo part of a variable (N1 N2 nc0)
    int identity;     // all three maps identity
    int map_kind[3];
    double map_c[3], map_h[3], map_beta[3], map_tanhb[3];
    const double* dx;      // [3*lb + d]
    const double* inv_dx;  // [3*lb + d]
    const double* bmin;    // [3*lb + d]
};

struct Ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // two pinned host buffers for the binary checkpoint, double-buffered
    // against the device conversion and copies; allocated on first use
    double* h_io[2] = {nullptr, nullptr};
    size_t h_io_doubles = 0;
    hipEvent_t io_ev[2] = {nullptr, nullptr};
};

struct Grid
{
    Ctx* ctx = nullptr;
    spd_grid_desc desc{};
    DevGeom geom{};
    double* d_tables = nullptr;   // dx, inv_dx, bmin
    int* d_nbr = nullptr;         // [27*lb + e] source block for ghost region e, -1 = not exchanged
    int* d_push = nullptr;        // [27*lb + e] block whose ghost region e block lb feeds, -1 = none
    int* d_start = nullptr;       // [lb] first plane the LDS-staged flux kernel left to the fix-up kernel
    double* d_stage = nullptr;    // host-layout staging for upload / download (stage_blocks blocks), lazily allocated
    int stage_blocks = 0;
    double* d_red = nullptr;      // reduction partials, 2 nblk + 1 doubles, lazily allocated
    double* h_red = nullptr;      // pinned host copy of them
    std::vector<double> h_dx, h_inv_dx, h_bmin;
};

struct Array
{
    Grid* grid = nullptr;
    double* base = nullptr;       // allocation
    double* d = nullptr;          // device data, geom.sb * nblk doubles (= base)
};


// ---- device helpers shared by the kernels ----

// pins x in a register at this point of the program: keeps the compiler from
// sinking its computation past a following barrier
__device__ inline void keep_here(double x) { asm volatile("" : : "v"(x)); }
__device__ inline void keep_here(float x) { asm volatile("" : : "v"(x)); }

// Workgroup barrier that orders LDS only.  For kernels whose threads exchange
// data through LDS alone (global memory read-only, or written only at cells
// no other thread of the launch reads): the barrier does not wait for
// outstanding global loads, so a prefetch stays in flight across it, which
// __syncthreads (a fence over all address spaces) would drain.
__device__ inline void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The first plane of block lb left to the fix-up kernels by the LDS-staged
// flux kernel (nz when none): that kernel may split a block's planes between
// up to kFixupParts workgroups, each recording the first plane it left in its
// own entry (d_start holds kFixupParts ints per block)
constexpr int kFixupParts = 4;
__host__ __device__ inline int fixup_start(const int* start, int lb)
{
    int m = start[kFixupParts*lb];
    for (int e = 1; e < kFixupParts; ++e) m = start[kFixupParts*lb + e] < m ? start[kFixupParts*lb + e] : m;
    return m;
}

// ---- kernels' host launchers ----
void launch_flux_div_generic(const Grid& g, const double* prims, double* rhs, const spd_flux_desc& f,
                             bool increment, hipStream_t s);
void launch_flux_div_fixup(const Grid& g, const double* prims, double* rhs, const spd_flux_desc& f, bool increment,
                           hipStream_t s);
bool launch_flux_div_fast(const Grid& g, const double* prims, double* rhs, const spd_flux_desc& f,
                          bool increment, hipStream_t s);
bool flux_div_fast_eligible(const Grid& g, const spd_flux_desc& f);
// One transform-aware RK stage fused with flux_div on the fast kernel:
// qout <- c2p(p2c(qin) + sum_{j < nr-1} coef[j] r[j] + coef[nr-1] L(qin)) on
// interior cells; L(qin) is also written to rhs when store_rhs (and, for the
// planes left to the fix-up, always).  push: the ghosts the result feeds are
// written into qout (else qout's ghost cells are left alone).
// read_neighbours: L(qin) takes its halo from the neighbour blocks' interiors
// instead of qin's ghost cells (x and z periodic, y periodic or isothermal
// walls on both faces, whose ghosts the kernel forms at t_wall:
// stage_reads_neighbours_ok), so qin's ghosts need not be current.
// Bit-identical to run_flux_div into rhs followed by launch_rk_transform_stage
// on qin (whose ghosts hold what an exchange copies).  Needs a fast-kernel
// composition and 16 x 16 x nz blocks (nz >= 2 ng, ng <= 2); false when not
// eligible (nothing launched).  qin and qout must be distinct arrays.
bool launch_flux_stage_fast(const Grid& g, const double* qin, double* qout, double* rhs, const spd_flux_desc& f,
                            const double* const* rprev, const double* coef, int nr, bool store_rhs,
                            double gamma, double R, hipStream_t s, bool read_neighbours, bool push,
                            int wall_mask = 0, double t_wall = 0.0);
bool flux_stage_fast_eligible(const Grid& g, const spd_flux_desc& f);
bool stage_reads_neighbours_ok(const Grid& g, int wall_mask);
// non-zero when the composition has a compile-time specialised per-cell kernel
int flux_div_spec_code(const spd_flux_desc& f);
// the composition contains a hybrid_scheme_t term
bool flux_has_hybrid(const spd_flux_desc& f);
void launch_exchange(const Grid& g, double* a, hipStream_t s);
// The fix-up of a fused stage in one launch for the compositions with a
// specialised per-cell kernel: per block with d_start[lb] < nz, (pull) its
// ghosts and (wall_y) wall slabs, the residual of its planes from there into
// rhs, and their stage qin -> qout with r[0..nr-1], r[nr-1] = rhs; false
// (nothing launched) for the other compositions
bool launch_stage_fixup_spec(const Grid& g, double* qin, double* qout, double* rhs, const spd_flux_desc& f,
                             const double* const* r, const double* coef, int nr, double gamma, double R,
                             bool pull, int wall_y, double t_wall, hipStream_t s);
// the exchange of the blocks with d_start[lb] < nz only (the others exit at
// once); wall_y: those at a y wall then fill its isothermal ghost slab at t_wall
void launch_exchange_flagged(const Grid& g, double* a, hipStream_t s, int wall_y = 0, double t_wall = 0.0);
void launch_wall_fill(const Grid& g, double* a, int face, double t_wall, hipStream_t s);
// boundary_fill over the faces of face_mask (bit 2*dir+pm), one launch per face in face order
void launch_fill_extrap(const Grid& g, double* a, int face_mask, int order, hipStream_t s);
void launch_fill_affine(const Grid& g, double* a, int face_mask, const spd_fill_kernel& k, hipStream_t s);
void launch_fill(const Grid& g, double* a, double value, hipStream_t s);
void launch_aos_to_soa(const Grid& g, const double* stage, double* a, int lb0, int nb, hipStream_t s);
void launch_soa_to_aos(const Grid& g, const double* a, double* stage, int lb0, int nb, hipStream_t s);
// RK stage: q <- c2p(p2c(q) + sum_j coef[j] * r_j) on interior cells
// The stage kernels also write every updated cell into the ghost cells of the
// blocks it feeds (d_push): the exchange of the stage result, fused.
// push = false leaves every ghost cell alone (interior update only)
void launch_rk_transform_stage(const Grid& g, double* q, const double* const* r, const double* coef, int nr,
                               double gamma, double R, hipStream_t s, bool push = true);
// the same stage, qin -> qout (ghosts pushed into qout), for the planes
// k >= d_start[lb] of each block only (the cells a fused launch left to the
// fix-up); 16 x 16 block cross-sections
void launch_stage_fixup(const Grid& g, const double* qin, double* qout, const double* const* r, const double* coef,
                        int nr, double gamma, double R, hipStream_t s);
// grid_array element-wise operations: scale every stored double; a (+|-)= b on
// interior cells; convert_state in place (every stored cell or the interior)
void launch_array_scale(const Grid& g, double* a, double sc, hipStream_t s);
void launch_array_incr(const Grid& g, double* a, const double* b, bool sub, hipStream_t s);
void launch_state_convert(const Grid& g, double* a, bool to_prim, double gamma, double R, bool ghosts, hipStream_t s);
void launch_rk_opt_s0(const Grid& g, double* q, const double* r0, double dt, double gamma, double R, hipStream_t s);
void launch_rk_opt_s1(const Grid& g, double* q, double* r0, const double* r1, double dt, double gamma, double R, hipStream_t s);
void launch_rk_opt_s2(const Grid& g, double* q, const double* r0, const double* r1, double dt, double gamma, double R, hipStream_t s);
// both keep their reduction buffers with the grid (allocated on first use)
void diagnostics(Grid& g, const double* q, double R, double& ke, double& en, hipStream_t s);
double reduce_max_wavespeed(Grid& g, const double* q, double gamma, double R, hipStream_t s);
// The CFL reduction with its partials in the device's static scratch, result
// left in device memory d_out; allocates nothing.  init_device_scratch looks
// up the scratch's address on `device` (which must be current), once.
void init_device_scratch(int device);
void max_wavespeed_noalloc(Grid& g, const double* q, double gamma, double R, double* d_out, hipStream_t s);

// Device layout of one variable of one block ("core + x-halo"):
//   core:   every stored row (j, k in [-ng, n+ng)), ghost rows included, holds
//           its nc0 interior-x cells (i = 0 .. nc0-1); rows in (j, k) order,
//           pitch nc0 -- with nc0 = 16 a row is one 128-byte line;
//   x-halo: at offset xh, one section per ghost layer and side (layer 0 low:
//           i = -1, layer 0 high: i = nc0, layer 1 low: i = -2, ...), each
//           holding one cell of every row in the same row order.
// Interior rows are whole, aligned lines (row stores that are not 64-byte
// aligned run at a third of the aligned rate on MI355X HBM), the x-ghost
// cells of consecutive rows are contiguous (the exchange writes whole lines,
// a one-layer stencil reads only the first section), and nothing is padding.
__host__ __device__ inline long long stored_offset(const DevGeom& g, int i, int j, int k)
{
    const long long row = (long long)(k + g.ng)*g.N[1] + (j + g.ng);
    if (i >= 0 && i < g.nc[0]) return row*g.nc[0] + i;
    const int section = (i < 0) ? 2*(-1 - i) : 2*(i - g.nc[0]) + 1;
    return g.xh + (long long)section*g.N[1]*g.N[2] + row;
}

// d x / d xi of a diagonal map along axis d at computational coordinate xi
// (core/coord_system.h:103-140): 1 for identity axes, for the tanh
// stretching x = c + h tanh(beta s)/tanh(beta), s = (xi - c)/h:
// beta (1 - tanh^2(beta s))/tanh(beta)
__device__ inline double map_deriv(const DevGeom& G, int d, double xi)
{
    if (G.map_kind[d] == SPD_MAP_TANH)
    {
        const double s = (xi - G.map_c[d])/G.map_h[d];
        const double t = tanh(G.map_beta[d]*s);
        return G.map_beta[d]*(1.0 - t*t)/G.map_tanhb[d];
    }
    return 1.0;
}

// The hardware hands workgroup w of a launch to XCD w mod 8.  A launch that
// takes its work in XCD runs permutes its workgroup numbers with this map so
// that each XCD takes one contiguous run of work items instead (XCD c: the
// c-th eighth, in order): blocks next to each other then run on the same XCD
// at about the same time and share the L2 lines of the halos they read from
// each other's interiors.  A bijection of [0, n) for every n.
__device__ inline int xcd_run_order(int w, int n)
{
    constexpr int X = 8;
    const int c = w % X, s = w / X, q = n / X, rem = n % X;
    return c < rem ? c*(q + 1) + s : rem*(q + 1) + (c - rem)*q + s;
}

// Index of a stored cell (i, j, k in [-ng, n+ng)) of variable v in block lb.
__host__ __device__ inline long long cell_index(const DevGeom& g, int lb, int v, int i, int j, int k)
{
    return (long long)lb*g.sb + (long long)v*g.sv + stored_offset(g, i, j, k);
}

// The same for a cell with 0 <= i < nc0 (core part only).
__host__ __device__ inline long long core_index(const DevGeom& g, int lb, int v, int i, int j, int k)
{
    return (long long)lb*g.sb + (long long)v*g.sv + ((long long)(k + g.ng)*g.N[1] + (j + g.ng))*g.nc[0] + i;
}

} // namespace spd
