// Ghost exchange, wall boundary fill, array fill and reductions.
#include "spd_internal.hpp"
#include "spd_stage.hpp"
#include <cmath>
#include <algorithm>
#include <mutex>
#include <string>

namespace spd {
namespace {

// max that passes NaN through from either operand.  The reference's
// utils::max (t1 > t2 ? t1 : t2, core/utils.h:87-90) returns a NaN second
// operand, and its driver stops on isnan(umax)
// (development/cuda-tgv/main.cc:251); here the result is NaN whenever any
// cell's wave speed is, whatever the reduction order, so a diverged state
// gives a NaN wave speed (fmax would return the max of the finite cells)
__device__ inline double max_nan(double a, double b)
{
    return (a > b || a != a) ? a : b;
}

__global__ void __launch_bounds__(256)
k_exchange(DevGeom G, const int* __restrict__ nbr, double* __restrict__ a, int xcd_runs)
{
    const int per = gridDim.x;
    int lb = blockIdx.y, part = blockIdx.x;
    if (xcd_runs)
    {
        const int w = xcd_run_order(blockIdx.x + per*blockIdx.y, per*gridDim.y);
        lb = w / per; part = w - lb*per;
    }
    exchange_item(G, nbr, a, lb, part*blockDim.x + threadIdx.x);
}

// The exchange of the blocks the LDS-staged flux kernel left to the fix-up
// (only[lb] < nz): one workgroup per block, the others exit at once.  With
// wall_y (isothermal walls on both y faces), a flagged block at a wall then
// fills that wall's ghost slab, as the wall fill after an exchange would: the
// slab mirrors rows of the block, x- and z-ghost cells included, which this
// workgroup has just written (visible to it after the barrier)
__global__ void __launch_bounds__(256)
k_exchange_flagged(DevGeom G, const int* __restrict__ nbr, double* __restrict__ a, const int* __restrict__ only,
                   int wall_y, double t_wall)
{
    const int lb = blockIdx.x;
    if (fixup_start(only, lb) >= G.nc[2]) return;
    const int items = exchange_items(G);
    for (int t = threadIdx.x; t < items; t += blockDim.x) exchange_item(G, nbr, a, lb, t);
    if (!wall_y) return;
    const int by = (lb / G.nb[0]) % G.nb[1];
    const bool low = by == 0, high = by == G.nb[1] - 1;
    if (!low && !high) return;                       // the same for the whole workgroup
    __syncthreads();
    const int n = G.N[0]*G.ng*G.N[2];
    for (int pm = 0; pm < 2; ++pm)
    {
        if (pm == 0 ? !low : !high) continue;
        for (int c = threadIdx.x; c < n; c += blockDim.x) wall_cell(G, a, lb, 1, pm, c, t_wall);
    }
}

__global__ void __launch_bounds__(256)
k_wall_fill(DevGeom G, double* __restrict__ a, int idir, int pm, double t_wall)
{
    // blockIdx.y enumerates the blocks on the face: (b[d1], b[d2]), d1 < d2 the other axes
    const int d1 = idir == 0 ? 1 : 0, d2 = idir == 2 ? 1 : 2;
    int b[3];
    b[idir] = pm == 0 ? 0 : G.nb[idir] - 1;
    b[d1] = blockIdx.y % G.nb[d1];
    b[d2] = blockIdx.y / G.nb[d1];
    const int lb = b[0] + G.nb[0]*(b[1] + G.nb[1]*b[2]);
    wall_cell(G, a, lb, idir, pm, blockIdx.x*blockDim.x + threadIdx.x, t_wall);
}

// Layout conversion on the device for nb blocks starting at block lb0:
// stage holds them in the reference's host layout (AoS, variable fastest,
// stored cells including ghosts, mem_map.h:484-496), a is the device array.
__device__ inline long long stored_offset(const DevGeom& G, int c)
{
    const int i = c % G.N[0], r = c / G.N[0];     // r = (j + ng) + N1 (k + ng)
    return stored_offset(G, i - G.ng, r % G.N[1] - G.ng, r / G.N[1] - G.ng);
}

__global__ void __launch_bounds__(256)
k_aos_to_soa(DevGeom G, const double* __restrict__ stage, double* __restrict__ a, int lb0)
{
    const int cells = G.N[0]*G.N[1]*G.N[2];
    const int c = blockIdx.x*blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const int b = blockIdx.y;
    const double* e = stage + 5*((long long)b*cells + c);
    double* d = a + (long long)(lb0 + b)*G.sb + stored_offset(G, c);
    #pragma unroll
    for (int v = 0; v < 5; ++v) d[v*G.sv] = e[v];
}

__global__ void __launch_bounds__(256)
k_soa_to_aos(DevGeom G, const double* __restrict__ a, double* __restrict__ stage, int lb0)
{
    const int cells = G.N[0]*G.N[1]*G.N[2];
    const int c = blockIdx.x*blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const int b = blockIdx.y;
    double* e = stage + 5*((long long)b*cells + c);
    const double* d = a + (long long)(lb0 + b)*G.sb + stored_offset(G, c);
    #pragma unroll
    for (int v = 0; v < 5; ++v) e[v] = d[v*G.sv];
}

__global__ void __launch_bounds__(256)
k_fill(double* __restrict__ a, long long n, double value)
{
    const long long c = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (c < n) a[c] = value;
}

__global__ void __launch_bounds__(1024)
k_max_wavespeed(DevGeom G, const double* __restrict__ q, double gamma, double R, double* __restrict__ part)
{
    __shared__ double red[1024];
    const int per_blk = G.nc[0]*G.nc[1]*G.nc[2];
    const int lb = blockIdx.y;
    double m = 0.0;
    for (int c = blockIdx.x*blockDim.x + threadIdx.x; c < per_blk; c += gridDim.x*blockDim.x)
    {
        const int i = c % G.nc[0], j = (c / G.nc[0]) % G.nc[1], k = c / (G.nc[0]*G.nc[1]);
        const double T = q[cell_index(G, lb, 1, i, j, k)];
        const double u = q[cell_index(G, lb, 2, i, j, k)];
        const double v = q[cell_index(G, lb, 3, i, j, k)];
        const double w = q[cell_index(G, lb, 4, i, j, k)];
        const double s = sqrt(gamma*R*T) + sqrt(u*u + v*v + w*w);
        m = max_nan(s, m);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1)
    {
        if (threadIdx.x < w) red[threadIdx.x] = max_nan(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.y*gridDim.x + blockIdx.x] = red[0];
}

// Per-block sums of rho |u|^2 / 2 and |omega|^2 / 2 (omega from central
// differences; needs current ghosts).  One workgroup of 1024 threads per block
// (four times the loads in flight of 256), fixed-order tree.
__global__ void __launch_bounds__(1024)
k_diag(DevGeom G, const double* __restrict__ q, double R, double* __restrict__ part)
{
    __shared__ double red[2][1024];
    const int lb = blockIdx.x;
    const int n0 = G.nc[0], n1 = G.nc[1], ng = G.ng, N1 = G.N[1];
    const int per_blk = n0*n1*G.nc[2];
    const long long sv = G.sv, rows = (long long)N1*G.N[2];
    const double* __restrict__ B = q + (long long)lb*G.sb;
    const double ix = G.inv_dx[3*lb], iy = G.inv_dx[3*lb + 1], iz = G.inv_dx[3*lb + 2];
    // neighbours in j and k are core cells (ghost rows included) at -+n0 and
    // -+N1 n0; in i, core cells inside the block, else the first x-halo
    // section of that side (i = -1: section 0, i = n0: section 1)
    const long long dj = n0, dk = (long long)N1*n0;
    double ke = 0.0, en = 0.0;
    for (int c = threadIdx.x; c < per_blk; c += blockDim.x)
    {
        const int i = c % n0, j = (c / n0) % n1, k = c / (n0*n1);
        const long long row = (long long)(k + ng)*N1 + (j + ng);
        const long long o = row*n0 + i;
        const long long om = i > 0 ? o - 1 : G.xh + row;
        const long long op = i + 1 < n0 ? o + 1 : G.xh + rows + row;
        auto Q = [&](int v, long long off) { return B[v*sv + off]; };
        const double p = Q(0, o), T = Q(1, o);
        const double u = Q(2, o), v = Q(3, o), w = Q(4, o);
        ke += 0.5*(p/(R*T))*(u*u + v*v + w*w);
        const double ox = 0.5*((Q(4, o + dj) - Q(4, o - dj))*iy - (Q(3, o + dk) - Q(3, o - dk))*iz);
        const double oy = 0.5*((Q(2, o + dk) - Q(2, o - dk))*iz - (Q(4, op) - Q(4, om))*ix);
        const double oz = 0.5*((Q(3, op) - Q(3, om))*ix - (Q(2, o + dj) - Q(2, o - dj))*iy);
        en += 0.5*(ox*ox + oy*oy + oz*oz);
    }
    red[0][threadIdx.x] = ke; red[1][threadIdx.x] = en;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) { red[0][threadIdx.x] += red[0][threadIdx.x + s]; red[1][threadIdx.x] += red[1][threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2*lb] = red[0][0]; part[2*lb + 1] = red[1][0]; }
}

// max of n partials by one workgroup (order-free: max is exact)
__global__ void __launch_bounds__(1024)
k_max_final(const double* __restrict__ part, int n, double* __restrict__ out)
{
    __shared__ double red[1024];
    double m = 0.0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) m = max_nan(part[t], m);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1)
    {
        if (threadIdx.x < w) red[threadIdx.x] = max_nan(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// Per-grid reduction buffers, allocated on first use and kept with the grid:
// partials on the device (2 nblk + 1 doubles) and a pinned host copy, so a
// reduction allocates nothing and copies back through pinned memory.
void ensure_red(Grid& g)
{
    if (!g.d_red) SPD_HIP(hipMalloc((void**)&g.d_red, sizeof(double)*(2*(size_t)g.geom.nblk + 1)));
    if (!g.h_red) SPD_HIP(hipHostMalloc((void**)&g.h_red, sizeof(double)*(2*(size_t)g.geom.nblk + 1)));
}

} // namespace

void diagnostics(Grid& g, const double* q, double R, double& ke, double& en, hipStream_t s)
{
    const DevGeom& G = g.geom;
    ensure_red(g);
    hipLaunchKernelGGL(k_diag, dim3(G.nblk), dim3(1024), 0, s, G, q, R, g.d_red);
    SPD_LAUNCH_CHECK();
    SPD_HIP(hipMemcpyAsync(g.h_red, g.d_red, sizeof(double)*2*G.nblk, hipMemcpyDeviceToHost, s));
    SPD_HIP(hipStreamSynchronize(s));
    const double* h = g.h_red;
    double a = 0.0, b = 0.0;
    for (int lb = 0; lb < G.nblk; ++lb) { a += h[2*lb]; b += h[2*lb + 1]; }
    const double n = (double)G.nc[0]*G.nc[1]*G.nc[2]*G.nblk;
    ke = a/n; en = b/n;
}

// Ghost cell c of face (idir, pm) of face block blockIdx.y: its block and
// index (normal index outside [0, n), tangential indices over [-ng, n+ng)).
// false for threads past the face's ghost cells.
__device__ inline bool face_ghost(const DevGeom& G, int idir, int pm, int& lb, int idx[3])
{
    const int d1 = idir == 0 ? 1 : 0, d2 = idir == 2 ? 1 : 2;
    int b[3];
    b[idir] = pm == 0 ? 0 : G.nb[idir] - 1;
    b[d1] = blockIdx.y % G.nb[d1];
    b[d2] = blockIdx.y / G.nb[d1];
    lb = b[0] + G.nb[0]*(b[1] + G.nb[1]*b[2]);
    int ext[3] = {G.N[0], G.N[1], G.N[2]};
    ext[idir] = G.ng;
    const int c = blockIdx.x*blockDim.x + threadIdx.x;
    if (c >= ext[0]*ext[1]*ext[2]) return false;
    idx[0] = c % ext[0] - G.ng; idx[1] = (c / ext[0]) % ext[1] - G.ng; idx[2] = c / (ext[0]*ext[1]) - G.ng;
    if (pm == 1) idx[idir] = G.nc[idir] + (idx[idir] + G.ng);           // [n, n+ng); pm = 0: [-ng, 0)
    return true;
}

// boundary_fill with extrap_t<order> (grid/boundary_fill.h:17-26, 63-98):
// Lagrange extrapolation along the face normal through cells i0 .. i0+order
// (i0 = 0 at a min face, n - order - 1 at a max face), the weights built on
// integer cell indices as l_j = l_j (x - x_i) / (x_j - x_i) over ascending
// i != j, the ghost accumulated as g = g + l_j q_j from 0.
__global__ void __launch_bounds__(256)
k_fill_extrap(DevGeom G, double* __restrict__ a, int idir, int pm, int order)
{
    int lb, idx[3];
    if (!face_ghost(G, idir, pm, lb, idx)) return;
    const int i0 = pm*(G.nc[idir] - order - 1);
    double gv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int j = 0; j <= order; ++j)
    {
        double lj = 1.0;
        for (int i = 0; i <= order; ++i)
        {
            if (i == j) continue;
            const int x = idx[idir], xj = i0 + j, xi = i0 + i;
            lj = lj*(x - xi)/(xj - xi);
        }
        int src[3] = {idx[0], idx[1], idx[2]};
        src[idir] = i0 + j;
        #pragma unroll
        for (int v = 0; v < 5; ++v) gv[v] = gv[v] + lj*a[cell_index(G, lb, v, src[0], src[1], src[2])];
    }
    #pragma unroll
    for (int v = 0; v < 5; ++v) a[cell_index(G, lb, v, idx[0], idx[1], idx[2])] = gv[v];
}

// boundary_fill with a kernel of the affine family (boundary_fill.h:99-127):
// ghost[v] = dom[v] q_image[v] + fill[v] q_ghost[v] + add[v], the velocity
// component normal to the face taking the *_normal coefficients; terms with a
// zero coefficient are left out, the others rounded in that order.  The image
// cell mirrors the ghost across the face (-1 - i, or 2n - (i + 1)).
__global__ void __launch_bounds__(256)
k_fill_affine(DevGeom G, double* __restrict__ a, int idir, int pm, spd_fill_kernel K)
{
    int lb, idx[3];
    if (!face_ghost(G, idir, pm, lb, idx)) return;
    int im[3] = {idx[0], idx[1], idx[2]};
    im[idir] = (pm == 0) ? (-1 - idx[idir]) : (2*G.nc[idir] - (idx[idir] + 1));
    double out[5];
    #pragma unroll
    for (int v = 0; v < 5; ++v)
    {
        const bool nrm = (v == 2 + idir);
        const double cd = nrm ? K.dom_normal : K.dom[v];
        const double cf = nrm ? K.fill_normal : K.fill[v];
        const double ca = nrm ? K.add_normal : K.add[v];
        double r = 0.0;
        bool any = false;
        if (cd != 0.0) { r = cd*a[cell_index(G, lb, v, im[0], im[1], im[2])]; any = true; }
        if (cf != 0.0) { const double t = cf*a[cell_index(G, lb, v, idx[0], idx[1], idx[2])]; r = any ? r + t : t; any = true; }
        if (ca != 0.0) { r = any ? r + ca : ca; }
        out[v] = r;
    }
    #pragma unroll
    for (int v = 0; v < 5; ++v) a[cell_index(G, lb, v, idx[0], idx[1], idx[2])] = out[v];
}

void launch_exchange(const Grid& g, double* a, hipStream_t s)
{
    const int items = exchange_items(g.geom);
    // runs by default: on the 256^3 Taylor-Green grid 0.465 -> 0.401 ms per
    // exchange, the same array; SPD_EXCH_XCD=0|1 forces the choice
    int xr = 1;
    if (const char* e = getenv("SPD_EXCH_XCD")) xr = e[0] == '1' ? 1 : 0;
    hipLaunchKernelGGL(k_exchange, dim3((items + 255)/256, g.geom.nblk), dim3(256), 0, s, g.geom, g.d_nbr, a, xr);
    SPD_LAUNCH_CHECK();
}

void launch_exchange_flagged(const Grid& g, double* a, hipStream_t s, int wall_y, double t_wall)
{
    hipLaunchKernelGGL(k_exchange_flagged, dim3(g.geom.nblk), dim3(256), 0, s, g.geom, g.d_nbr, a,
                       (const int*)g.d_start, wall_y, t_wall);
    SPD_LAUNCH_CHECK();
}

void launch_wall_fill(const Grid& g, double* a, int face, double t_wall, hipStream_t s)
{
    const DevGeom& G = g.geom;
    const int idir = face/2, pm = face % 2;
    int ext[3] = {G.N[0], G.N[1], G.N[2]};
    ext[idir] = G.ng;
    const int n = ext[0]*ext[1]*ext[2];
    const int d1 = idir == 0 ? 1 : 0, d2 = idir == 2 ? 1 : 2;
    const int face_blocks = G.nb[d1]*G.nb[d2];
    hipLaunchKernelGGL(k_wall_fill, dim3((n + 255)/256, face_blocks), dim3(256), 0, s, G, a, idir, pm, t_wall);
    SPD_LAUNCH_CHECK();
}

// grid of one boundary face: ghost cells per face block x face blocks
static void face_grid(const DevGeom& G, int idir, dim3& grid)
{
    int ext[3] = {G.N[0], G.N[1], G.N[2]};
    ext[idir] = G.ng;
    const int n = ext[0]*ext[1]*ext[2];
    const int d1 = idir == 0 ? 1 : 0, d2 = idir == 2 ? 1 : 2;
    grid = dim3((n + 255)/256, G.nb[d1]*G.nb[d2]);
}

void launch_fill_extrap(const Grid& g, double* a, int face_mask, int order, hipStream_t s)
{
    for (int face = 0; face < 6; ++face)
    {
        if (!(face_mask & (1 << face))) continue;
        dim3 grid;
        face_grid(g.geom, face/2, grid);
        hipLaunchKernelGGL(k_fill_extrap, grid, dim3(256), 0, s, g.geom, a, face/2, face % 2, order);
        SPD_LAUNCH_CHECK();
    }
}

void launch_fill_affine(const Grid& g, double* a, int face_mask, const spd_fill_kernel& k, hipStream_t s)
{
    for (int face = 0; face < 6; ++face)
    {
        if (!(face_mask & (1 << face))) continue;
        dim3 grid;
        face_grid(g.geom, face/2, grid);
        hipLaunchKernelGGL(k_fill_affine, grid, dim3(256), 0, s, g.geom, a, face/2, face % 2, k);
        SPD_LAUNCH_CHECK();
    }
}

void launch_aos_to_soa(const Grid& g, const double* stage, double* a, int lb0, int nb, hipStream_t s)
{
    const DevGeom& G = g.geom;
    const int cells = G.N[0]*G.N[1]*G.N[2];
    hipLaunchKernelGGL(k_aos_to_soa, dim3((cells + 255)/256, nb), dim3(256), 0, s, G, stage, a, lb0);
    SPD_LAUNCH_CHECK();
}

void launch_soa_to_aos(const Grid& g, const double* a, double* stage, int lb0, int nb, hipStream_t s)
{
    const DevGeom& G = g.geom;
    const int cells = G.N[0]*G.N[1]*G.N[2];
    hipLaunchKernelGGL(k_soa_to_aos, dim3((cells + 255)/256, nb), dim3(256), 0, s, G, a, stage, lb0);
    SPD_LAUNCH_CHECK();
}

void launch_fill(const Grid& g, double* a, double value, hipStream_t s)
{
    const long long n = g.geom.sb*g.geom.nblk;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255)/256)), dim3(256), 0, s, a, n, value);
    SPD_LAUNCH_CHECK();
}

namespace {

// max(|u| + c) over the interior: one workgroup of 1024 per block writes one
// partial per block into part[0 .. nblk), then one workgroup writes their max
// to out (device memory)
void launch_max_wavespeed(const DevGeom& G, const double* q, double gamma, double R, double* part, double* out,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_max_wavespeed, dim3(1, G.nblk), dim3(1024), 0, s, G, q, gamma, R, part);
    SPD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(1024), 0, s, part, G.nblk, out);
    SPD_LAUNCH_CHECK();
}

} // namespace

// Scratch of the allocation-free CFL reduction (spd_max_wavespeed_noalloc):
// one partial per block, at most kScratchBlocks blocks (256 KiB per device,
// held until the process ends).  Every call on one device shares that
// device's copy and overwrites the partials it reads, so nothing needs
// clearing; calls that overlap on one device race.  The array's name is a
// host-side stand-in: host code uses the per-device address below, looked up
// once per device at context creation with that device current.
constexpr int kScratchBlocks = 32768;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchBlocks];
static double* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

void init_device_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices)
        throw ArgError{"device index " + std::to_string(device) + " beyond " + std::to_string(kMaxDevices)};
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    SPD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double*>(ptr);
}

void max_wavespeed_noalloc(Grid& g, const double* q, double gamma, double R, double* d_out, hipStream_t s)
{
    const DevGeom& G = g.geom;
    if (G.nblk > kScratchBlocks)
        throw ArgError{"the allocation-free reduction takes at most " + std::to_string(kScratchBlocks) +
                       " blocks (grid has " + std::to_string(G.nblk) + ")"};
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        part = g_scratch_ptr[g.ctx->device];
    }
    if (!part) throw ArgError{"device scratch not set up for this context's device"};
    launch_max_wavespeed(G, q, gamma, R, part, d_out, s);
}

double reduce_max_wavespeed(Grid& g, const double* q, double gamma, double R, hipStream_t s)
{
    const DevGeom& G = g.geom;
    ensure_red(g);
    launch_max_wavespeed(G, q, gamma, R, g.d_red, g.d_red + 2*(size_t)G.nblk, s);
    SPD_HIP(hipMemcpyAsync(g.h_red, g.d_red + 2*(size_t)G.nblk, sizeof(double), hipMemcpyDeviceToHost, s));
    SPD_HIP(hipStreamSynchronize(s));
    return g.h_red[0];
}

} // namespace spd
