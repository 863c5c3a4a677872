// Per-cell pieces of a transform-aware RK stage shared by the stage kernels
// (rk.hip) and the fused flux + stage kernel (flux_fast.hip): the prim<->cons
// transforms, and the fused exchange that writes a stage result into the
// ghost cells of the neighbour blocks.  Compiled without contraction, so
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

} // namespace spd
