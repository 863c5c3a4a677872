// SSP-RK3 stage updates on interior cells, in conserved variables with the
// prim<->cons transform around them (navier-stokes/fluid_state.h:102-135).
//   * transform-aware stage (time-integration/advance.h:56-102):
//       cons = p2c(q); cons += c_j * r_j (j in order, nonzero c_j only); q = c2p(cons)
//   * ssprk3_opt stages s0, s1, s2 (advance.h:285-353)
// Compiled without contraction: the combination keeps the reference's
// rounding sequence.
#include "spd_internal.hpp"
#include "spd_stage.hpp"
#include <algorithm>
#include <string>

namespace spd {
namespace {

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

// The stage of the cells a fused flux + stage launch left to the fix-up: the
// planes k >= start[lb] of each block, whose residual (C.r[C.n - 1]) the
// per-cell fix-up kernel has just written.  Reads qin, writes qout and its
// pushed ghosts, with k_rk_transform's arithmetic.  One 256-thread workgroup
// per 16 x 16 block cross-section, planes in order; blocks with nothing to
// redo exit at once.
__global__ void __launch_bounds__(256)
k_stage_fixup(DevGeom G, const double* __restrict__ qin, double* __restrict__ qout, Coefs C, double gamma,
              double R, const int* __restrict__ push, const int* __restrict__ start)
{
    __shared__ XStage xs;
    const int lb = blockIdx.x;
    for (int k = fixup_start(start, lb); k < G.nc[2]; ++k)
    {
        const CellIt ci{(int)(threadIdx.x & 15), (int)(threadIdx.x >> 4), k, lb};
        double qq[5], cons[5];
        long long idx[5];
        for (int v = 0; v < 5; ++v) { idx[v] = core_index(G, lb, v, ci.i, ci.j, k); qq[v] = qin[idx[v]]; }
        p2c(gamma, R, qq, cons);
        #pragma unroll
        for (int s = 0; s < kMaxStages; ++s)
        {
            if (s >= C.n || C.c[s] == 0.0) continue;
            double inc[5];
            for (int v = 0; v < 5; ++v) inc[v] = C.c[s]*C.r[s][idx[v]];
            for (int v = 0; v < 5; ++v) cons[v] += inc[v];
        }
        c2p(gamma, R, cons, qq);
        for (int v = 0; v < 5; ++v) qout[idx[v]] = qq[v];
        push_ghosts(G, push, qout, ci, qq, true);
        stage_x(G, xs, ci, qq);
        lds_barrier();
        flush_x(G, push, qout, xs, lb, k);
        lds_barrier();
    }
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

void launch_stage_fixup(const Grid& g, const double* qin, double* qout, const double* const* r, const double* coef,
                        int nr, double gamma, double R, hipStream_t s)
{
    if (nr < 1 || nr > kMaxStages) throw ArgError{"a stage combines 1.." + std::to_string(kMaxStages) + " residuals"};
    Coefs C{};
    C.n = nr;
    for (int j = 0; j < nr; ++j) { C.r[j] = r[j]; C.c[j] = coef[j]; }
    hipLaunchKernelGGL(k_stage_fixup, dim3((unsigned)g.geom.nblk), dim3(256), 0, s, g.geom, qin, qout, C, gamma, R,
                       g.d_push, g.d_start);
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
