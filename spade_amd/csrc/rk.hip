// SSP-RK3 stage updates on interior cells, in conserved variables with the
// prim<->cons transform around them (navier-stokes/fluid_state.h:102-135).
//   * transform-aware stage (time-integration/advance.h:56-102):
//       cons = p2c(q); cons += c_j * r_j (j in order, nonzero c_j only); q = c2p(cons)
//   * ssprk3_opt stages s0, s1, s2 (advance.h:285-353)
// Compiled without contraction: the combination keeps the reference's
// rounding sequence.
#include "spd_internal.hpp"
#include <algorithm>
#include <string>

namespace spd {
namespace {

struct CellIt
{
    int i, j, k, lb;
};

// grid: (cells of a block / 256 rounded up, blocks); 32-bit index arithmetic
__device__ inline bool interior_cell(const DevGeom& G, CellIt& c)
{
    const unsigned per_blk = (unsigned)(G.nc[0]*G.nc[1]*G.nc[2]);
    const unsigned r = blockIdx.x*blockDim.x + threadIdx.x;
    if (r >= per_blk) return false;
    c.lb = blockIdx.y;
    const unsigned n0 = (unsigned)G.nc[0], n01 = (unsigned)(G.nc[0]*G.nc[1]);
    c.k = (int)(r / n01);
    const unsigned rr = r - (unsigned)c.k*n01;
    c.j = (int)(rr / n0);
    c.i = (int)(rr - (unsigned)c.j*n0);
    return true;
}

__device__ inline void p2c(double gamma, double R, const double q[5], double c[5])
{
    const double rho   = q[0]/(R*q[1]);
    const double rhoU2 = rho*(q[2]*q[2] + q[3]*q[3] + q[4]*q[4]);
    const double rhoE  = 0.5*rhoU2 + (q[0]/((gamma - 1.0)));
    c[0] = rho; c[1] = rhoE; c[2] = rho*q[2]; c[3] = rho*q[3]; c[4] = rho*q[4];
}

__device__ inline void c2p(double gamma, double R, const double c[5], double q[5])
{
    const double rho = c[0];
    const double invrho = 1.0/rho;
    const double u = invrho*c[2], v = invrho*c[3], w = invrho*c[4];
    const double rhoU2 = rho*(u*u + v*v + w*w);
    const double p = (gamma - 1.0)*(c[1] - 0.5*rhoU2);
    const double T = p/(R*rho);
    q[0] = p; q[1] = T; q[2] = u; q[3] = v; q[4] = w;
}

constexpr int kMaxStages = SPD_RK_MAX_STAGES;
struct Coefs { const double* r[kMaxStages]; double c[kMaxStages]; int n; };

// The exchange of the stage result, fused: interior cell (i, j, k) of block lb
// is written into every ghost cell that holds it in another block.  Ghost
// region e of block D is fed by block nbr[D][e] (grid/exchange_pattern.h:
// 382-424); push[27 lb + e] = D names the block that region of lb's cells
// goes to, at (i + ex nc0, j + ey nc1, k + ez nc2).  A cell within ng of the
// low face feeds regions with e = +1 on that axis, within ng of the high face
// regions with e = -1.  The values are the ones just stored, so the ghosts
// equal what the exchange would copy.
__device__ inline void push_one(const DevGeom& G, const int* __restrict__ push, double* __restrict__ q,
                                const CellIt& c, const double qq[5], int ex, int ey, int ez)
{
    const int D = push[27*c.lb + (ex + 1) + 3*(ey + 1) + 9*(ez + 1)];
    if (D < 0) return;
    for (int v = 0; v < 5; ++v)
        q[cell_index(G, D, v, c.i + ex*G.nc[0], c.j + ey*G.nc[1], c.k + ez*G.nc[2])] = qq[v];
}

__device__ inline void push_ghosts(const DevGeom& G, const int* __restrict__ push, double* __restrict__ q,
                                   const CellIt& c, const double qq[5], bool skip_x = false)
{
    const int ng = G.ng;
    const bool lo0 = c.i < ng, lo1 = c.j < ng, lo2 = c.k < ng;
    const bool hi0 = c.i >= G.nc[0] - ng, hi1 = c.j >= G.nc[1] - ng, hi2 = c.k >= G.nc[2] - ng;
    if (!(lo0 | lo1 | lo2 | hi0 | hi1 | hi2)) return;
    if (G.nc[0] >= 2*ng && G.nc[1] >= 2*ng && G.nc[2] >= 2*ng)
    {
        // at most one direction per axis: the cell's own face set, 1 to 7 regions;
        // skip_x leaves every region across x to the staged plane flush
        const int ex = (lo0 ? 1 : (hi0 ? -1 : 0))*(skip_x ? 0 : 1);
        const int ey = lo1 ? 1 : (hi1 ? -1 : 0), ez = lo2 ? 1 : (hi2 ? -1 : 0);
        if (ex)             push_one(G, push, q, c, qq, ex, 0, 0);
        if (ey)             push_one(G, push, q, c, qq, 0, ey, 0);
        if (ez)             push_one(G, push, q, c, qq, 0, 0, ez);
        if (ex && ey)       push_one(G, push, q, c, qq, ex, ey, 0);
        if (ex && ez)       push_one(G, push, q, c, qq, ex, 0, ez);
        if (ey && ez)       push_one(G, push, q, c, qq, 0, ey, ez);
        if (ex && ey && ez) push_one(G, push, q, c, qq, ex, ey, ez);
        return;
    }
    // blocks narrower than two ghost layers: a cell can feed both sides of an axis
    for (int ez = -1; ez <= 1; ++ez)
    {
        if ((ez == 1 && !lo2) || (ez == -1 && !hi2)) continue;
        for (int ey = -1; ey <= 1; ++ey)
        {
            if ((ey == 1 && !lo1) || (ey == -1 && !hi1)) continue;
            for (int ex = -1; ex <= 1; ++ex)
            {
                if ((ex == 1 && !lo0) || (ex == -1 && !hi0)) continue;
                if (ex == 0 && ey == 0 && ez == 0) continue;
                push_one(G, push, q, c, qq, ex, ey, ez);
            }
        }
    }
}

// Regions across x written from LDS.  An x-ghost section holds one cell per
// stored row, so a wave pushing its own cells writes 8-byte pieces spread over
// a section; with the plane's cells within ng of an x face staged first, each
// (section, variable) receives 16 consecutive rows from consecutive threads.
// Needs a workgroup of 256 threads on one 16 x 16 plane and blocks at least
// 2 ng wide (one direction per axis).  Slot s < ng holds cells i = s (fed to
// ex = +1 at i + 16), slot ng + s holds i = 15 - s (ex = -1 at i - 16).
struct XStage { double v[4][5][16]; };

__device__ inline bool plane_staged(const DevGeom& G)
{
    return G.nc[0] == 16 && G.nc[1] == 16 && G.ng <= 2 && G.nc[2] >= 2*G.ng;
}

__device__ inline void stage_x(const DevGeom& G, XStage& xs, const CellIt& c, const double qq[5])
{
    int slot = -1;
    if (c.i < G.ng) slot = c.i;
    else if (c.i >= 16 - G.ng) slot = G.ng + (15 - c.i);
    if (slot >= 0)
        for (int v = 0; v < 5; ++v) xs.v[slot][v][c.j] = qq[v];
}

__device__ inline void flush_x(const DevGeom& G, const int* __restrict__ push, double* __restrict__ q,
                               const XStage& xs, int lb, int k)
{
    const int ng = G.ng;
    const int ez_k = k < ng ? 1 : (k >= G.nc[2] - ng ? -1 : 0);
    const int n = 2*ng*5*16;
    for (int t = threadIdx.x; t < n; t += blockDim.x)
    {
        const int slot = t/80, v = (t/16) % 5, j = t % 16;
        const bool lo = slot < ng;
        const int i = lo ? slot : 15 - (slot - ng);
        const int ex = lo ? 1 : -1;
        const int ey_j = j < ng ? 1 : (j >= 16 - ng ? -1 : 0);
        const double val = xs.v[slot][v][j];
        for (int a = 0; a < 2; ++a)
        {
            const int ey = a ? ey_j : 0;
            if (a && !ey) continue;
            for (int b = 0; b < 2; ++b)
            {
                const int ez = b ? ez_k : 0;
                if (b && !ez) continue;
                const int D = push[27*lb + (ex + 1) + 3*(ey + 1) + 9*(ez + 1)];
                if (D < 0) continue;
                q[cell_index(G, D, v, i + ex*16, j + ey*16, k + ez*G.nc[2])] = val;
            }
        }
    }
}

// The stage result of this thread's cell into the ghosts it feeds: directly,
// or, on staged planes, everything across x from LDS after a workgroup barrier
// (every thread of the workgroup reaches it: a staged plane has no idle thread)
__device__ inline void push_result(const DevGeom& G, const int* __restrict__ push, double* __restrict__ q,
                                   const CellIt& c, const double qq[5], XStage& xs)
{
    if (!push) return;
    const bool staged = plane_staged(G);
    push_ghosts(G, push, q, c, qq, staged);
    if (!staged) return;
    stage_x(G, xs, c, qq);
    lds_barrier();
    flush_x(G, push, q, xs, c.lb, c.k);
}

__global__ void __launch_bounds__(256)
k_rk_transform(DevGeom G, double* __restrict__ q, Coefs C, double gamma, double R, const int* __restrict__ push)
{
    __shared__ XStage xs;
    CellIt ci;
    if (!interior_cell(G, ci)) return;
    double qq[5], cons[5];
    long long idx[5];
    for (int v = 0; v < 5; ++v) { idx[v] = core_index(G, ci.lb, v, ci.i, ci.j, ci.k); qq[v] = q[idx[v]]; }
    p2c(gamma, R, qq, cons);
    #pragma unroll
    for (int s = 0; s < kMaxStages; ++s)      // static indices: C stays in scalar registers
    {
        if (s >= C.n || C.c[s] == 0.0) continue;
        double inc[5];
        for (int v = 0; v < 5; ++v) inc[v] = C.c[s]*C.r[s][idx[v]];
        for (int v = 0; v < 5; ++v) cons[v] += inc[v];
    }
    c2p(gamma, R, cons, qq);
    for (int v = 0; v < 5; ++v) q[idx[v]] = qq[v];
    push_result(G, push, q, ci, qq, xs);
}

__global__ void __launch_bounds__(256)
k_rk_opt_s0(DevGeom G, double* __restrict__ q, const double* __restrict__ r0, double dt, double gamma, double R,
            const int* __restrict__ push)
{
    __shared__ XStage xs;
    CellIt ci;
    if (!interior_cell(G, ci)) return;
    double qq[5], c[5]; long long idx[5];
    for (int v = 0; v < 5; ++v) { idx[v] = core_index(G, ci.lb, v, ci.i, ci.j, ci.k); qq[v] = q[idx[v]]; }
    p2c(gamma, R, qq, c);
    for (int v = 0; v < 5; ++v) c[v] += dt*r0[idx[v]];
    c2p(gamma, R, c, qq);
    for (int v = 0; v < 5; ++v) q[idx[v]] = qq[v];
    push_result(G, push, q, ci, qq, xs);
}

__global__ void __launch_bounds__(256)
k_rk_opt_s1(DevGeom G, double* __restrict__ q, double* __restrict__ r0, const double* __restrict__ r1, double dt, double gamma, double R,
            const int* __restrict__ push)
{
    __shared__ XStage xs;
    CellIt ci;
    if (!interior_cell(G, ci)) return;
    double qq[5], c[5], old[5], nr[5]; long long idx[5];
    for (int v = 0; v < 5; ++v) { idx[v] = core_index(G, ci.lb, v, ci.i, ci.j, ci.k); qq[v] = q[idx[v]]; }
    p2c(gamma, R, qq, c);
    const double sixth_dt = (1.0/6.0)*dt;
    for (int v = 0; v < 5; ++v) { old[v] = r0[idx[v]]; nr[v] = sixth_dt*(old[v] + r1[idx[v]]); r0[idx[v]] = nr[v]; }
    for (int v = 0; v < 5; ++v) c[v] += (3.0/2.0)*nr[v];
    for (int v = 0; v < 5; ++v) c[v] -= dt*old[v];
    c2p(gamma, R, c, qq);
    for (int v = 0; v < 5; ++v) q[idx[v]] = qq[v];
    push_result(G, push, q, ci, qq, xs);
}

__global__ void __launch_bounds__(256)
k_rk_opt_s2(DevGeom G, double* __restrict__ q, const double* __restrict__ r0, const double* __restrict__ r1, double dt, double gamma, double R,
            const int* __restrict__ push)
{
    __shared__ XStage xs;
    CellIt ci;
    if (!interior_cell(G, ci)) return;
    double qq[5], c[5]; long long idx[5];
    for (int v = 0; v < 5; ++v) { idx[v] = core_index(G, ci.lb, v, ci.i, ci.j, ci.k); qq[v] = q[idx[v]]; }
    p2c(gamma, R, qq, c);
    const double two_thirds_dt = dt*(2.0/3.0);
    for (int v = 0; v < 5; ++v) c[v] -= (1.0/2.0)*r0[idx[v]];
    for (int v = 0; v < 5; ++v) c[v] += two_thirds_dt*r1[idx[v]];
    c2p(gamma, R, c, qq);
    for (int v = 0; v < 5; ++v) q[idx[v]] = qq[v];
    push_result(G, push, q, ci, qq, xs);
}

// grid_array::operator*=(scalar) (grid/grid_array.h:369-379): every stored double
__global__ void __launch_bounds__(256)
k_array_scale(double* __restrict__ a, long long n, double s)
{
    for (long long e = blockIdx.x*(long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x*blockDim.x)
        a[e] *= s;
}

// grid_array::operator+= / -= (grid_array.h:286-318): a[v] += b[v] (or -=) on
// the interior cells, variable by variable
__global__ void __launch_bounds__(256)
k_array_incr(DevGeom G, double* __restrict__ a, const double* __restrict__ b, int sub)
{
    CellIt ci;
    if (!interior_cell(G, ci)) return;
    for (int v = 0; v < 5; ++v)
    {
        const long long e = core_index(G, ci.lb, v, ci.i, ci.j, ci.k);
        if (sub) a[e] -= b[e]; else a[e] += b[e];
    }
}

// convert_state in place (navier-stokes/fluid_state.h:102-135) over every
// stored cell of a block (ghosts = 1) or its interior cells
__global__ void __launch_bounds__(256)
k_state_convert(DevGeom G, double* __restrict__ a, int to_prim, double gamma, double R, int ghosts)
{
    const int ext0 = ghosts ? G.N[0] : G.nc[0], ext1 = ghosts ? G.N[1] : G.nc[1], ext2 = ghosts ? G.N[2] : G.nc[2];
    const int c = blockIdx.x*blockDim.x + threadIdx.x;
    if (c >= ext0*ext1*ext2) return;
    const int off = ghosts ? G.ng : 0;
    const int i = c % ext0 - off, j = (c / ext0) % ext1 - off, k = c / (ext0*ext1) - off;
    const int lb = blockIdx.y;
    long long idx[5];
    double x[5], y[5];
    for (int v = 0; v < 5; ++v) { idx[v] = cell_index(G, lb, v, i, j, k); x[v] = a[idx[v]]; }
    if (to_prim) c2p(gamma, R, x, y);
    else         p2c(gamma, R, x, y);
    for (int v = 0; v < 5; ++v) a[idx[v]] = y[v];
}

inline dim3 interior_blocks(const DevGeom& G)
{
    const int n = G.nc[0]*G.nc[1]*G.nc[2];
    return dim3((unsigned)((n + 255)/256), (unsigned)G.nblk);
}

} // namespace

void launch_rk_transform_stage(const Grid& g, double* q, const double* const* r, const double* coef, int nr,
                               double gamma, double R, hipStream_t s, bool push)
{
    if (nr < 1 || nr > kMaxStages) throw ArgError{"a stage combines 1.." + std::to_string(kMaxStages) + " residuals"};
    Coefs C{};
    C.n = nr;
    for (int j = 0; j < nr; ++j) { C.r[j] = r[j]; C.c[j] = coef[j]; }
    hipLaunchKernelGGL(k_rk_transform, interior_blocks(g.geom), dim3(256), 0, s, g.geom, q, C, gamma, R,
                       push ? g.d_push : nullptr);
    SPD_LAUNCH_CHECK();
}

void launch_array_scale(const Grid& g, double* a, double sc, hipStream_t s)
{
    const long long n = g.geom.sb*(long long)g.geom.nblk;
    const unsigned wg = (unsigned)std::min<long long>((n + 255)/256, 8192);
    hipLaunchKernelGGL(k_array_scale, dim3(wg), dim3(256), 0, s, a, n, sc);
    SPD_LAUNCH_CHECK();
}

void launch_array_incr(const Grid& g, double* a, const double* b, bool sub, hipStream_t s)
{
    hipLaunchKernelGGL(k_array_incr, interior_blocks(g.geom), dim3(256), 0, s, g.geom, a, b, sub ? 1 : 0);
    SPD_LAUNCH_CHECK();
}

void launch_state_convert(const Grid& g, double* a, bool to_prim, double gamma, double R, bool ghosts, hipStream_t s)
{
    const DevGeom& G = g.geom;
    const int n = ghosts ? G.N[0]*G.N[1]*G.N[2] : G.nc[0]*G.nc[1]*G.nc[2];
    hipLaunchKernelGGL(k_state_convert, dim3((unsigned)((n + 255)/256), (unsigned)G.nblk), dim3(256), 0, s, G, a,
                       to_prim ? 1 : 0, gamma, R, ghosts ? 1 : 0);
    SPD_LAUNCH_CHECK();
}

void launch_rk_opt_s0(const Grid& g, double* q, const double* r0, double dt, double gamma, double R, hipStream_t s)
{
    hipLaunchKernelGGL(k_rk_opt_s0, interior_blocks(g.geom), dim3(256), 0, s, g.geom, q, r0, dt, gamma, R, g.d_push);
    SPD_LAUNCH_CHECK();
}

void launch_rk_opt_s1(const Grid& g, double* q, double* r0, const double* r1, double dt, double gamma, double R, hipStream_t s)
{
    hipLaunchKernelGGL(k_rk_opt_s1, interior_blocks(g.geom), dim3(256), 0, s, g.geom, q, r0, r1, dt, gamma, R, g.d_push);
    SPD_LAUNCH_CHECK();
}

void launch_rk_opt_s2(const Grid& g, double* q, const double* r0, const double* r1, double dt, double gamma, double R, hipStream_t s)
{
    hipLaunchKernelGGL(k_rk_opt_s2, interior_blocks(g.geom), dim3(256), 0, s, g.geom, q, r0, r1, dt, gamma, R, g.d_push);
    SPD_LAUNCH_CHECK();
}

} // namespace spd
