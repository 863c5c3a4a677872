// Per-cell pieces of a transform-aware RK stage shared by the stage kernels
// (rk.hip), the fused flux + stage kernel (flux_fast.hip) and its fix-up
// (flux_generic.hip): the prim<->cons transforms, the fused exchange that
// writes a stage result into the ghost cells of the neighbour blocks, the
// exchange of one block's ghost cells and the isothermal wall mirror.  Compiled without contraction, so
// both kernels round every expression alike.
#pragma once
#include "spd_internal.hpp"

namespace spd {

struct CellIt
{
    int i, j, k, lb;
};

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

constexpr int kMaxStages = SPD_RK_MAX_STAGES;
// stage coefficients: qout = c2p(p2c(qin) + c[0] r[0] + .. + c[n-1] r[n-1])
struct Coefs { const double* r[kMaxStages]; double c[kMaxStages]; int n; };

// Ghost exchange (grid/exchange_pattern.h:382-424): every ghost cell copies
// the cell of the neighbour block that owns it (26-neighbour table built on
// the host, grid/cartesian_blocks.h:73-105; region (ex+1) + 3(ey+1) + 9(ez+1),
// e_d = -1 below the interior, +1 above, 0 inside; -1 = not exchanged).  The
// items of a block, in storage order, one per thread (grid: items / 256 x
// blocks; every thread's copy is independent of the others', so the loads of
// all items are in flight together):
//   (i) the ghost rows of the core part -- rows (j, k) with j or k outside
//       the interior: the 2 ng planes below / above in k, then the 2 ng rows
//       below / above in j of each interior plane -- one double of a row per
//       item, so a wave copies whole rows (nc0 contiguous doubles);
//  (ii) the x-halo: one item per (side, row) copies the row's ng x-ghost
//       cells on that side, whose sources are adjacent interior cells of one
//       row of the x-neighbour (i = -1-l <- nc0-1-l, i = nc0+l <- l);
//       consecutive items write consecutive doubles of a section.
__host__ __device__ inline int exchange_items(const DevGeom& G)
{
    return (2*G.ng*G.N[1] + 2*G.ng*G.nc[2])*G.nc[0] + 2*G.N[1]*G.N[2];
}

// Exchange item t of block lb (storage order of its ghost cells, see above)
__device__ inline void exchange_item(const DevGeom& G, const int* __restrict__ nbr, double* __restrict__ a,
                                     int lb, int t)
{
    const int n0 = G.nc[0], n1 = G.nc[1], n2 = G.nc[2], ng = G.ng;
    const int N1 = G.N[1];
    const int rows = N1*G.N[2];
    const long long sv = G.sv;
    const int* nb = nbr + 27*lb;
    double* __restrict__ blk = a + (long long)lb*G.sb;
    auto region = [](int x, int n) { return x < 0 ? -1 : (x >= n ? 1 : 0); };

    const int plane_rows = 2*ng*N1;
    const int core_items = (plane_rows + 2*ng*n2)*n0;
    if (t < core_items)
    {
        const int i = t % n0, g = t / n0;
        int j, k;
        if (g < plane_rows)
        {
            const int kk = g / N1;
            j = g - kk*N1 - ng;
            k = kk < ng ? kk - ng : n2 + kk - ng;
        }
        else
        {
            const int h = g - plane_rows;
            const int jj = h % (2*ng);
            k = h / (2*ng);
            j = jj < ng ? jj - ng : n1 + jj - ng;
        }
        const int ey = region(j, n1), ez = region(k, n2);
        const int src = nb[1 + 3*(ey + 1) + 9*(ez + 1)];
        if (src < 0) return;
        const int r = (k + ng)*N1 + (j + ng);
        const int sr = (k - ez*n2 + ng)*N1 + (j - ey*n1 + ng);
        const double* s = a + (long long)src*G.sb + (long long)sr*n0 + i;
        double* d = blk + (long long)r*n0 + i;
        double q[5];
        #pragma unroll
        for (int v = 0; v < 5; ++v) q[v] = s[v*sv];
        #pragma unroll
        for (int v = 0; v < 5; ++v) d[v*sv] = q[v];
        return;
    }
    const int x = t - core_items;
    if (x >= 2*rows) return;
    const int side = x / rows, r = x - side*rows;
    const int j = r % N1 - ng, k = r / N1 - ng;
    const int ex = side ? 1 : -1, ey = region(j, n1), ez = region(k, n2);
    const int src = nb[(ex + 1) + 3*(ey + 1) + 9*(ez + 1)];
    if (src < 0) return;
    const int sr = (k - ez*n2 + ng)*N1 + (j - ey*n1 + ng);
    const double* s = a + (long long)src*G.sb + (long long)sr*n0;
    for (int l = 0; l < ng; ++l)
    {
        const int si = side ? l : n0 - 1 - l;
        double* d = blk + G.xh + (long long)(2*l + side)*rows + r;
        double q[5];
        #pragma unroll
        for (int v = 0; v < 5; ++v) q[v] = s[si + v*sv];
        #pragma unroll
        for (int v = 0; v < 5; ++v) d[v*sv] = q[v];
    }
}

// grid (workgroups per block) x blocks.  With xcd_runs the linear workgroup
// number is permuted so that each XCD takes a contiguous run of blocks: a
// block's two x-neighbours both gather x-halo cells from its rows (16 bytes
// of each 128-byte row for each of them), and in runs they do so on the same
// XCD at about the same time, so the second gather finds the lines in L2
// boundary_fill with the isothermal no-slip mirror (grid/boundary_fill.h:
// 104-130): ghost (normal index < 0 or >= n, tangential over the full stored
// range) = (p, 2 Tw - T, -u, -v, -w) of the mirror cell.  Cell c of the ghost
// slab of face (idir, pm) of block lb.
__device__ inline void wall_cell(const DevGeom& G, double* __restrict__ a, int lb, int idir, int pm, int c,
                                 double t_wall)
{
    int ext[3] = {G.N[0], G.N[1], G.N[2]};
    ext[idir] = G.ng;
    if (c >= ext[0]*ext[1]*ext[2]) return;
    int idx[3] = {c % ext[0] - G.ng, (c / ext[0]) % ext[1] - G.ng, c / (ext[0]*ext[1]) - G.ng};
    if (pm == 0) idx[idir] = -G.ng + (idx[idir] + G.ng);                 // [-ng, 0)
    else         idx[idir] = G.nc[idir] + (idx[idir] + G.ng);            // [n, n+ng)
    int im[3] = {idx[0], idx[1], idx[2]};
    im[idir] = (pm == 0) ? (-1 - idx[idir]) : (2*G.nc[idir] - (idx[idir] + 1));
    const double p = a[cell_index(G, lb, 0, im[0], im[1], im[2])];
    const double T = a[cell_index(G, lb, 1, im[0], im[1], im[2])];
    const double u = a[cell_index(G, lb, 2, im[0], im[1], im[2])];
    const double v = a[cell_index(G, lb, 3, im[0], im[1], im[2])];
    const double w = a[cell_index(G, lb, 4, im[0], im[1], im[2])];
    a[cell_index(G, lb, 0, idx[0], idx[1], idx[2])] = p;
    a[cell_index(G, lb, 1, idx[0], idx[1], idx[2])] = 2.0*t_wall - T;
    a[cell_index(G, lb, 2, idx[0], idx[1], idx[2])] = -u;
    a[cell_index(G, lb, 3, idx[0], idx[1], idx[2])] = -v;
    a[cell_index(G, lb, 4, idx[0], idx[1], idx[2])] = -w;
}

} // namespace spd
