// C ABI of libspade_amd (declared in include/spade_amd.h).
#include "api_internal.hpp"
#include "batch_host.hpp"
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <memory>
#include <utility>

namespace spd {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

} // namespace spd

using namespace spd;

struct spd_batch_s { std::unique_ptr<Batch> b; };

namespace spd {

void validate_flux(const spd_flux_desc* f, int ng)
{
    require(f != nullptr, "flux descriptor is null");
    require(f->nterms >= 1 && f->nterms <= SPD_MAX_TERMS, "flux descriptor needs 1..4 terms");
    require(f->gamma > 1.0 && f->R > 0.0, "gas needs gamma > 1 and R > 0");
    for (int m = 0; m < f->nterms; ++m)
    {
        const spd_flux_term& t = f->terms[m];
        switch (t.kind)
        {
            case SPD_TERM_KEEP:
                require(t.order >= 2 && t.order % 2 == 0 && t.order <= 8, "cent_keep order must be even, 2..8");
                require(t.order/2 <= ng, "cent_keep order needs order/2 ghost layers");
                require(t.order <= 4, "cent_keep orders above 4 are not implemented");
                break;
            case SPD_TERM_WENO_FDS:
                require(ng >= 2, "WENO stencils need 2 ghost layers");
                break;
            case SPD_TERM_WENO: case SPD_TERM_FWENO:
                require(ng >= 2, "WENO stencils need 2 ghost layers");
                require(t.kind == SPD_TERM_FWENO || t.split == SPD_SPLIT_RUSANOV || t.split == SPD_SPLIT_PHYS,
                        "WENO split must be SPD_SPLIT_RUSANOV or SPD_SPLIT_PHYS");
                break;
            case SPD_TERM_PDISS:
                require(ng >= 1, "flux needs at least one ghost layer");
                require(t.sensor == SPD_SENSOR_CONST || t.sensor == SPD_SENSOR_DUCROS,
                        "pressure_diss sensor must be SPD_SENSOR_CONST or SPD_SENSOR_DUCROS");
                break;
            case SPD_TERM_VISC:
                require(ng >= 1, "flux needs at least one ghost layer");
                require(t.law == SPD_LAW_CONST || t.law == SPD_LAW_POWER, "viscous law must be SPD_LAW_CONST or SPD_LAW_POWER");
                require(t.prandtl > 0.0, "viscous law needs a Prandtl number > 0");
                require(t.law == SPD_LAW_CONST || t.T_ref > 0.0, "power law needs T_ref > 0");
                require(t.sgs == SPD_SGS_NONE || t.sgs == SPD_SGS_WALE, "subgrid model must be SPD_SGS_NONE or SPD_SGS_WALE");
                require(t.sgs == SPD_SGS_NONE || (t.sgs_prt > 0.0 && std::isfinite(t.sgs_cw) && std::isfinite(t.sgs_delta)),
                        "WALE needs a turbulent Prandtl number > 0 and finite cw, delta");
                break;
            case SPD_TERM_TOTANI: case SPD_TERM_FIRST:
                require(ng >= 1, "flux needs at least one ghost layer");
                break;
            case SPD_TERM_HYBRID:
                require(t.order == 2 || t.order == 4, "hybrid central scheme must be cent_keep<2> or cent_keep<4>");
                require(t.diss_kind == SPD_TERM_WENO || t.diss_kind == SPD_TERM_FWENO,
                        "hybrid dissipative scheme must be SPD_TERM_WENO or SPD_TERM_FWENO");
                require(ng >= 2, "hybrid WENO stencils need 2 ghost layers");
                require(t.diss_kind == SPD_TERM_FWENO || t.split == SPD_SPLIT_RUSANOV || t.split == SPD_SPLIT_PHYS,
                        "WENO split must be SPD_SPLIT_RUSANOV or SPD_SPLIT_PHYS");
                require(t.sensor == SPD_SENSOR_CONST || t.sensor == SPD_SENSOR_DUCROS,
                        "hybrid sensor must be SPD_SENSOR_CONST or SPD_SENSOR_DUCROS");
                require(t.full == 0 || t.full == 1, "hybrid full must be 0 (diss_flux) or 1 (full_flux)");
                break;
            default:
                throw ArgError{"unknown flux term kind " + std::to_string(t.kind)};
        }
    }
    // a well-formed hybrid_scheme_t descriptor is refused: the library has no
    // device code for it (an earlier device build of it was followed by a GPU
    // memory fault whose cause was never established)
    if (flux_has_hybrid(*f))
        throw UnsupportedError{"hybrid_scheme_t is not offered on the device"};
}

// Host arrays in the reference layout (AoS, variable fastest) cross the
// boundary through a device staging buffer of a few blocks; the layout is
// converted by a kernel, so the host only copies contiguous bytes.
double* ensure_stage(Grid& g)
{
    if (!g.d_stage)
    {
        const long long per_block = g.geom.sb;                     // doubles of one block (>= 5 stored cells x 5)
        const long long want = (64LL << 20)/(8*per_block);        // about 64 MiB of staging
        g.stage_blocks = (int)std::max(1LL, std::min<long long>(want, g.geom.nblk));
        SPD_HIP(hipMalloc((void**)&g.d_stage, sizeof(double)*per_block*g.stage_blocks));
    }
    return g.d_stage;
}

// The boundary step after a stage (exchange + wall fill): the stage kernels
// have already written the exchanged ghosts (push_ghosts), so only the walls
// remain; they mirror cells that include exchanged x-ghosts, hence after.
void apply_bc(spd_integrator_s* it, double* a)
{
    Grid& g = *it->q->a.grid;
    for (int face = 0; face < 6; ++face)
        if (it->bc.wall_mask & (1 << face)) launch_wall_fill(g, a ? a : it->q->a.d, face, it->bc.t_wall, g.ctx->stream);
}

// The SSP-RK3 step runs as three fused flux + stage launches (ping-pong
// between the state and a second buffer) when the fast kernel takes the
// composition and every ghost cell of a stage's output is rewritten by that
// stage: each axis periodic (the fused exchange) or walled on both faces
// (the wall fill covers its whole ghost slab).  SPD_DISABLE_FUSED=1 keeps
// the separate flux_div and stage launches.
static bool fused_step_ok(spd_integrator_s* it)
{
    if (const char* e = getenv("SPD_DISABLE_FUSED")) if (e[0] == '1') return false;
    Grid& g = *it->q->a.grid;
    for (int a = 0; a < 3; ++a)
    {
        const bool walls = (it->bc.wall_mask & (1 << (2*a))) && (it->bc.wall_mask & (1 << (2*a + 1)));
        if (!g.desc.periodic[a] && !walls) return false;
    }
    return flux_stage_fast_eligible(g, it->flux);
}

spd_array_s* new_array(Grid& g)
{
    auto* A = new spd_array_s;
    std::unique_ptr<spd_array_s> guard(A);
    A->a.grid = &g;
    const size_t bytes = sizeof(double)*(size_t)g.geom.sb*g.geom.nblk;
    SPD_HIP(hipMalloc((void**)&A->a.base, bytes));
    A->a.d = A->a.base;
    SPD_HIP(hipMemsetAsync(A->a.base, 0, bytes, g.ctx->stream));
    return guard.release();
}

void free_array(spd_array_s* A)
{
    if (!A) return;
    SPD_HIP(hipStreamSynchronize(A->a.grid->ctx->stream));
    SPD_HIP(hipFree(A->a.base));
    delete A;
}

void run_flux_div(const Grid& g, const double* prims, double* rhs, const spd_flux_desc& f, bool increment)
{
    if (!launch_flux_div_fast(g, prims, rhs, f, increment, g.ctx->stream))
        launch_flux_div_generic(g, prims, rhs, f, increment, g.ctx->stream);
}

} // namespace spd

extern "C" {

const char* spd_last_error(void) { return g_last_error.c_str(); }

int spd_ctx_create(int device, spd_ctx_t* out)
{
    return guarded([&]
    {
        require(out != nullptr, "out is null");
        int n = 0;
        SPD_HIP(hipGetDeviceCount(&n));
        require(device >= 0 && device < n, "device index out of range");
        auto* c = new spd_ctx_s;
        c->c.device = device;
        SPD_HIP(hipSetDevice(device));
        init_device_scratch(device);
        SPD_HIP(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking));
        SPD_HIP(hipEventCreate(&c->c.ev0));
        SPD_HIP(hipEventCreate(&c->c.ev1));
        *out = c;
    });
}

int spd_ctx_destroy(spd_ctx_t ctx)
{
    return guarded([&]
    {
        if (!ctx) return;
        SPD_HIP(hipStreamSynchronize(ctx->c.stream));
        for (int b = 0; b < 2; ++b)
        {
            if (ctx->c.h_io[b]) SPD_HIP(hipHostFree(ctx->c.h_io[b]));
            if (ctx->c.io_ev[b]) SPD_HIP(hipEventDestroy(ctx->c.io_ev[b]));
        }
        SPD_HIP(hipEventDestroy(ctx->c.ev0));
        SPD_HIP(hipEventDestroy(ctx->c.ev1));
        SPD_HIP(hipStreamDestroy(ctx->c.stream));
        delete ctx;
    });
}

int spd_ctx_sync(spd_ctx_t ctx)
{
    return guarded([&] { require(ctx != nullptr, "ctx is null"); SPD_HIP(hipStreamSynchronize(ctx->c.stream)); });
}

void* spd_ctx_stream(spd_ctx_t ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

int spd_timer_start(spd_ctx_t ctx)
{
    return guarded([&] { require(ctx != nullptr, "ctx is null"); SPD_HIP(hipEventRecord(ctx->c.ev0, ctx->c.stream)); });
}

int spd_timer_stop(spd_ctx_t ctx, float* ms)
{
    return guarded([&]
    {
        require(ctx != nullptr && ms != nullptr, "null argument");
        SPD_HIP(hipEventRecord(ctx->c.ev1, ctx->c.stream));
        SPD_HIP(hipEventSynchronize(ctx->c.ev1));
        SPD_HIP(hipEventElapsedTime(ms, ctx->c.ev0, ctx->c.ev1));
    });
}

int spd_grid_create(spd_ctx_t ctx, const spd_grid_desc* d, spd_grid_t* out)
{
    return guarded([&]
    {
        require(ctx && out, "null argument");
        validate_grid_desc(d);
        auto* G = new spd_grid_s;
        std::unique_ptr<spd_grid_s> guard(G);
        Grid& g = G->g;
        g.ctx = &ctx->c;
        g.desc = *d;
        DevGeom& dg = g.geom;
        dg.ng = d->num_exch;
        dg.nblk = 1;
        for (int a = 0; a < 3; ++a)
        {
            dg.nc[a] = d->cells_in_block[a];
            dg.N[a] = dg.nc[a] + 2*dg.ng;
            dg.nb[a] = d->num_blocks[a];
            dg.nblk *= dg.nb[a];
            dg.map_kind[a] = d->map[a].kind;
            dg.map_c[a] = 0.5*(d->map[a].p1 + d->map[a].p0);
            dg.map_h[a] = 0.5*(d->map[a].p1 - d->map[a].p0);
            dg.map_beta[a] = d->map[a].beta;
            dg.map_tanhb[a] = d->map[a].kind == SPD_MAP_TANH ? std::tanh(d->map[a].beta) : 1.0;
        }
        dg.identity = (dg.map_kind[0] == SPD_MAP_IDENTITY && dg.map_kind[1] == SPD_MAP_IDENTITY && dg.map_kind[2] == SPD_MAP_IDENTITY);
        dg.sv = (long long)dg.N[0]*dg.N[1]*dg.N[2];
        dg.xh = (long long)dg.N[1]*dg.N[2]*dg.nc[0];
        dg.sb = ((5*dg.sv + 31)/32)*32;   // 256-byte aligned blocks
        // geometry tables (grid/cartesian_blocks.h:38-71, cartesian_grid.h:116-145)
        g.h_dx.resize(3*dg.nblk); g.h_inv_dx.resize(3*dg.nblk); g.h_bmin.resize(3*dg.nblk);
        std::vector<int> nbr(27*dg.nblk, -1);
        for (int lb = 0; lb < dg.nblk; ++lb)
        {
            const int b[3] = {lb % dg.nb[0], (lb / dg.nb[0]) % dg.nb[1], lb / (dg.nb[0]*dg.nb[1])};
            for (int a = 0; a < 3; ++a)
            {
                block_box(*d, lb, a, g.h_bmin[3*lb + a], g.h_dx[3*lb + a]);
                g.h_inv_dx[3*lb + a] = 1.0/g.h_dx[3*lb + a];
            }
            // ghost region e of block lb is fed by block lb+e (wrapped); skipped
            // when that crosses a non-periodic domain face
            for (int ez = -1; ez <= 1; ++ez) for (int ey = -1; ey <= 1; ++ey) for (int ex = -1; ex <= 1; ++ex)
            {
                if (!ex && !ey && !ez) continue;
                const int e[3] = {ex, ey, ez};
                int s[3];
                bool skip = false;
                for (int a = 0; a < 3; ++a)
                {
                    s[a] = b[a] + e[a];
                    if (s[a] < 0 || s[a] >= dg.nb[a])
                    {
                        if (!d->periodic[a]) skip = true;
                        s[a] = (s[a] + dg.nb[a]) % dg.nb[a];
                    }
                }
                if (!skip) nbr[27*lb + (ex + 1) + 3*(ey + 1) + 9*(ez + 1)] = s[0] + dg.nb[0]*(s[1] + dg.nb[1]*s[2]);
            }
        }
        SPD_HIP(hipSetDevice(ctx->c.device));
        SPD_HIP(hipMalloc((void**)&g.d_tables, sizeof(double)*9*dg.nblk));
        // the same table read backwards: ghost region e of block D is fed by
        // S = nbr[D][e], so S pushes its cells of that region to D
        std::vector<int> push(27*dg.nblk, -1);
        for (int D = 0; D < dg.nblk; ++D)
            for (int e = 0; e < 27; ++e)
                if (nbr[27*D + e] >= 0) push[27*nbr[27*D + e] + e] = D;
        SPD_HIP(hipMalloc((void**)&g.d_nbr, sizeof(int)*27*dg.nblk));
        SPD_HIP(hipMalloc((void**)&g.d_push, sizeof(int)*27*dg.nblk));
        SPD_HIP(hipMalloc((void**)&g.d_start, sizeof(int)*kFixupParts*dg.nblk));
        SPD_HIP(hipMemcpy(g.d_tables, g.h_dx.data(), sizeof(double)*3*dg.nblk, hipMemcpyHostToDevice));
        SPD_HIP(hipMemcpy(g.d_tables + 3*dg.nblk, g.h_inv_dx.data(), sizeof(double)*3*dg.nblk, hipMemcpyHostToDevice));
        SPD_HIP(hipMemcpy(g.d_tables + 6*dg.nblk, g.h_bmin.data(), sizeof(double)*3*dg.nblk, hipMemcpyHostToDevice));
        SPD_HIP(hipMemcpy(g.d_nbr, nbr.data(), sizeof(int)*27*dg.nblk, hipMemcpyHostToDevice));
        SPD_HIP(hipMemcpy(g.d_push, push.data(), sizeof(int)*27*dg.nblk, hipMemcpyHostToDevice));
        const std::vector<int> complete((size_t)kFixupParts*dg.nblk, dg.nc[2]);
        SPD_HIP(hipMemcpy(g.d_start, complete.data(), sizeof(int)*kFixupParts*dg.nblk, hipMemcpyHostToDevice));
        dg.dx = g.d_tables;
        dg.inv_dx = g.d_tables + 3*dg.nblk;
        dg.bmin = g.d_tables + 6*dg.nblk;
        *out = guard.release();
    });
}

int spd_grid_destroy(spd_grid_t grid)
{
    return guarded([&]
    {
        if (!grid) return;
        SPD_HIP(hipStreamSynchronize(grid->g.ctx->stream));
        SPD_HIP(hipFree(grid->g.d_tables));
        SPD_HIP(hipFree(grid->g.d_nbr));
        SPD_HIP(hipFree(grid->g.d_push));
        SPD_HIP(hipFree(grid->g.d_start));
        if (grid->g.d_stage) SPD_HIP(hipFree(grid->g.d_stage));
        if (grid->g.d_red) SPD_HIP(hipFree(grid->g.d_red));
        if (grid->g.h_red) SPD_HIP(hipHostFree(grid->g.h_red));
        delete grid;
    });
}

long spd_array_host_size(spd_grid_t grid)
{
    if (!grid) return 0;
    const DevGeom& G = grid->g.geom;
    return 5L*G.N[0]*G.N[1]*G.N[2]*G.nblk;
}

long spd_array_device_size(spd_grid_t grid)
{
    if (!grid) return 0;
    const DevGeom& G = grid->g.geom;
    return (long)G.sb*G.nblk;
}

int spd_array_create(spd_grid_t grid, spd_array_t* out)
{
    return guarded([&]
    {
        require(grid && out, "null argument");
        *out = new_array(grid->g);
    });
}

int spd_array_destroy(spd_array_t arr)
{
    return guarded([&]
    {
        free_array(arr);
    });
}

int spd_array_upload(spd_array_t arr, const double* host)
{
    return guarded([&]
    {
        require(arr && host, "null argument");
        upload_blocks(arr, 0, arr->a.grid->geom.nblk, host);
    });
}

int spd_array_download(spd_array_t arr, double* host)
{
    return guarded([&]
    {
        require(arr && host, "null argument");
        download_blocks(arr, 0, arr->a.grid->geom.nblk, host);
    });
}

int spd_array_fill(spd_array_t arr, double value)
{
    return guarded([&]
    {
        require(arr != nullptr, "null argument");
        launch_fill(*arr->a.grid, arr->a.d, value, arr->a.grid->ctx->stream);
    });
}

int spd_array_copy(spd_array_t dst, spd_array_t src)
{
    return guarded([&]
    {
        require(dst && src, "null argument");
        require(dst->a.grid == src->a.grid, "arrays live on different grids");
        const Grid& g = *dst->a.grid;
        SPD_HIP(hipMemcpyAsync(dst->a.d, src->a.d, sizeof(double)*(size_t)g.geom.sb*g.geom.nblk,
                               hipMemcpyDeviceToDevice, g.ctx->stream));
    });
}

int spd_flux_div(spd_array_t prims, spd_array_t rhs, const spd_flux_desc* flux, int mode)
{
    return guarded([&]
    {
        require(prims && rhs, "null array");
        require(prims->a.grid == rhs->a.grid, "prims and rhs live on different grids");
        require(prims != rhs, "prims and rhs must be different arrays");
        require(mode == SPD_OVERWRITE || mode == SPD_INCREMENT, "mode must be SPD_OVERWRITE or SPD_INCREMENT");
        const Grid& g = *prims->a.grid;
        validate_flux(flux, g.geom.ng);
        run_flux_div(g, prims->a.d, rhs->a.d, *flux, mode == SPD_INCREMENT);
    });
}

int spd_flux_div_kernel(spd_array_t prims, const spd_flux_desc* flux, int* kind)
{
    return guarded([&]
    {
        require(prims && kind, "null argument");
        validate_flux(flux, prims->a.grid->geom.ng);
        *kind = flux_div_fast_eligible(*prims->a.grid, *flux) ? 1 : (flux_div_spec_code(*flux) ? 2 : 0);
    });
}

int spd_exchange(spd_array_t arr)
{
    return guarded([&]
    {
        require(arr != nullptr, "null argument");
        launch_exchange(*arr->a.grid, arr->a.d, arr->a.grid->ctx->stream);
    });
}

int spd_boundary_fill_wall(spd_array_t arr, int face, double t_wall)
{
    return guarded([&]
    {
        require(arr != nullptr, "null argument");
        require(face >= 0 && face < 6, "face must be 0..5");
        launch_wall_fill(*arr->a.grid, arr->a.d, face, t_wall, arr->a.grid->ctx->stream);
    });
}

int spd_boundary_fill_extrap(spd_array_t arr, int face_mask, int order)
{
    return guarded([&]
    {
        require(arr != nullptr, "null argument");
        require(face_mask >= 0 && face_mask < 64, "face mask must be a subset of the six faces (0..63)");
        require(order >= 0, "extrapolation order must be >= 0");
        const DevGeom& G = arr->a.grid->geom;
        for (int d = 0; d < 3; ++d)
            if (face_mask & (3 << (2*d)))
                require(order + 1 <= G.nc[d], "extrapolation needs order + 1 <= cells per block along the face normal");
        launch_fill_extrap(*arr->a.grid, arr->a.d, face_mask, order, arr->a.grid->ctx->stream);
    });
}

int spd_boundary_fill(spd_array_t arr, int face_mask, const spd_fill_kernel* kern)
{
    return guarded([&]
    {
        require(arr != nullptr && kern != nullptr, "null argument");
        require(face_mask >= 0 && face_mask < 64, "face mask must be a subset of the six faces (0..63)");
        const DevGeom& G = arr->a.grid->geom;
        for (int d = 0; d < 3; ++d)
            if (face_mask & (3 << (2*d)))
                require(G.ng <= G.nc[d], "mirror fill needs ng <= cells per block along the face normal");
        launch_fill_affine(*arr->a.grid, arr->a.d, face_mask, *kern, arr->a.grid->ctx->stream);
    });
}

int spd_integrator_create(spd_array_t q, const spd_flux_desc* flux, const spd_bc_desc* bc,
                          int scheme, double dt, spd_integrator_t* out)
{
    return guarded([&]
    {
        require(q && out, "null argument");
        require(scheme == SPD_SCHEME_SSPRK3 || scheme == SPD_SCHEME_SSPRK3_OPT, "unknown scheme");
        require(dt > 0.0 && std::isfinite(dt), "dt must be positive and finite");
        validate_flux(flux, q->a.grid->geom.ng);
        auto* it = new spd_integrator_s;
        std::unique_ptr<spd_integrator_s> guard(it);
        it->q = q;
        it->flux = *flux;
        if (bc) it->bc = *bc;
        if (bc)
            for (int face = 0; face < 6; ++face)
                if (bc->wall_mask & (1 << face))
                    require(!q->a.grid->desc.periodic[face/2], "walls are only allowed on non-periodic axes");
        it->scheme = scheme;
        it->dt = dt;
        const int nres = scheme == SPD_SCHEME_SSPRK3 ? 3 : 2;
        try
        {
            for (int i = 0; i < nres; ++i) it->r[i] = new_array(*q->a.grid);
        }
        catch (...)
        {
            for (auto*& r : it->r) { if (r) { (void)hipFree(r->a.base); delete r; r = nullptr; } }
            throw;
        }
        *out = guard.release();
    });
}

int spd_integrator_advance(spd_integrator_t it, int nsteps)
{
    return guarded([&]
    {
        require(it != nullptr, "null integrator");
        require(nsteps >= 0, "nsteps must be >= 0");
        if (it->generic)
        {
            for (int n = 0; n < nsteps; ++n) advance_generic(it);
            return;
        }
        Grid& g = *it->q->a.grid;
        double* q = it->q->a.d;
        const double dt = it->dt;
        const double gamma = it->flux.gamma, R = it->flux.R;
        for (int n = 0; n < nsteps; ++n)
        {
            if (it->scheme == SPD_SCHEME_SSPRK3 && fused_step_ok(it))
            {
                // the stages of the branch below, each fused with its flux_div:
                // q -> alt -> q -> alt, then the two buffers trade places so the
                // array handle holds the new state (the handle is opaque).  When
                // x and z are periodic and y is periodic or walled on both
                // faces, a stage reads its halo from the neighbour blocks'
                // interiors (forming the wall ghosts itself), so only the last
                // stage of this call writes ghost cells and fills the walls
                // (the first reads the caller's, as the separate launches would)
                if (!it->alt) it->alt = new_array(g);
                const bool nb = stage_reads_neighbours_ok(g, it->bc.wall_mask);
                const int wm = it->bc.wall_mask;
                const double tw = it->bc.t_wall;
                double* a = q;
                double* b = it->alt->a.d;
                double* r[3] = {it->r[0]->a.d, it->r[1]->a.d, it->r[2]->a.d};
                const double* rp[2] = {r[0], r[1]};
                const bool first = n == 0, last = n + 1 == nsteps;
                { const double c[1] = {(1.0/1.0)*dt};
                  launch_flux_stage_fast(g, a, b, r[0], it->flux, rp, c, 1, true, gamma, R, g.ctx->stream,
                                         nb && !first, !nb, wm, tw); }
                if (!nb) apply_bc(it, b);
                { const double c[2] = {(-3.0/4.0)*dt, (1.0/4.0)*dt};
                  launch_flux_stage_fast(g, b, a, r[1], it->flux, rp, c, 2, true, gamma, R, g.ctx->stream,
                                         nb, !nb, wm, tw); }
                if (!nb) apply_bc(it, a);
                { const double c[3] = {(-1.0/12.0)*dt, (-1.0/12.0)*dt, (2.0/3.0)*dt};
                  launch_flux_stage_fast(g, a, b, r[2], it->flux, rp, c, 3, false, gamma, R, g.ctx->stream,
                                         nb, !nb || last, wm, tw); }
                if (!nb || last) apply_bc(it, b);
                std::swap(it->q->a.base, it->alt->a.base);
                std::swap(it->q->a.d, it->alt->a.d);
                q = it->q->a.d;
            }
            else if (it->scheme == SPD_SCHEME_SSPRK3)
            {
                // rows a = {(0,0,0), (1,0,0), (1/4,1/4,0)}, b = (1/6,1/6,2/3)
                // (time-integration/explicit.h:80-88); stage coefficients are
                // the reduced row differences times dt (advance.h:66-95)
                double* r[3] = {it->r[0]->a.d, it->r[1]->a.d, it->r[2]->a.d};
                const double* rc[3] = {r[0], r[1], r[2]};
                run_flux_div(g, q, r[0], it->flux, false);
                { const double c[1] = {(1.0/1.0)*dt};
                  launch_rk_transform_stage(g, q, rc, c, 1, gamma, R, g.ctx->stream); }
                apply_bc(it);
                run_flux_div(g, q, r[1], it->flux, false);
                { const double c[2] = {(-3.0/4.0)*dt, (1.0/4.0)*dt};
                  launch_rk_transform_stage(g, q, rc, c, 2, gamma, R, g.ctx->stream); }
                apply_bc(it);
                run_flux_div(g, q, r[2], it->flux, false);
                { const double c[3] = {(-1.0/12.0)*dt, (-1.0/12.0)*dt, (2.0/3.0)*dt};
                  launch_rk_transform_stage(g, q, rc, c, 3, gamma, R, g.ctx->stream); }
                apply_bc(it);
            }
            else
            {
                double* r0 = it->r[0]->a.d; double* r1 = it->r[1]->a.d;
                run_flux_div(g, q, r0, it->flux, false);
                launch_rk_opt_s0(g, q, r0, dt, gamma, R, g.ctx->stream);
                apply_bc(it);
                run_flux_div(g, q, r1, it->flux, false);
                launch_rk_opt_s1(g, q, r0, r1, dt, gamma, R, g.ctx->stream);
                apply_bc(it);
                run_flux_div(g, q, r1, it->flux, false);
                launch_rk_opt_s2(g, q, r0, r1, dt, gamma, R, g.ctx->stream);
                apply_bc(it);
            }
        }
    });
}

int spd_integrator_fused(spd_integrator_t it, int* fused)
{
    return guarded([&]
    {
        require(it && fused, "null argument");
        *fused = (!it->generic && it->scheme == SPD_SCHEME_SSPRK3 && fused_step_ok(it)) ? 1 : 0;
    });
}

int spd_integrator_destroy(spd_integrator_t it)
{
    return guarded([&]
    {
        if (!it) return;
        for (auto* r : it->r) free_array(r);
        for (auto* r : it->res) free_array(r);
        free_array(it->q2);
        free_array(it->alt);
        delete it;
    });
}

spd_array_t spd_integrator_residual(spd_integrator_t it, int i)
{
    if (!it) return nullptr;
    if (it->generic) return (i >= 0 && i < it->table.stages) ? it->res[i] : nullptr;
    if (i < 0 || i > 2) return nullptr;
    return it->r[i];
}

int spd_max_wavespeed(spd_array_t prims, double gamma, double R, double* out)
{
    return guarded([&]
    {
        require(prims && out, "null argument");
        Grid& g = *prims->a.grid;
        *out = reduce_max_wavespeed(g, prims->a.d, gamma, R, g.ctx->stream);
    });
}

int spd_max_wavespeed_noalloc(spd_array_t prims, double gamma, double R, double* d_out)
{
    return guarded([&]
    {
        require(prims && d_out, "null argument");
        Grid& g = *prims->a.grid;
        hipPointerAttribute_t attr{};
        const hipError_t e = hipPointerGetAttributes(&attr, d_out);
        if (e != hipSuccess) (void)hipGetLastError();     // a host pointer: clear the error it leaves
        require(e == hipSuccess && attr.type == hipMemoryTypeDevice && attr.device == g.ctx->device,
                "d_out must point to device memory of the array's device");
        max_wavespeed_noalloc(g, prims->a.d, gamma, R, d_out, g.ctx->stream);
    });
}

int spd_diagnostics(spd_array_t prims, double R, double* ke, double* en)
{
    return guarded([&]
    {
        require(prims && ke && en, "null argument");
        require(R > 0.0, "R must be positive");
        Grid& g = *prims->a.grid;
        diagnostics(g, prims->a.d, R, *ke, *en, g.ctx->stream);
    });
}

int spd_batch_create(spd_ctx_t ctx, long long budget, int law, spd_batch_t* out)
{
    return guarded([&]
    {
        require(ctx && out, "null argument");
        auto* b = new spd_batch_s;
        std::unique_ptr<spd_batch_s> guard(b);
        b->b.reset(new Batch(&ctx->c, budget, law, 64));
        *out = guard.release();
    });
}

int spd_batch_destroy(spd_batch_t b)
{
    return guarded([&] { delete b; });
}

int spd_batch_capacity(spd_batch_t b, const spd_case_desc* d, long long* need_blocks, long long* free_blocks)
{
    return guarded([&]
    {
        require(b && d && need_blocks && free_blocks, "null argument");
        *need_blocks = Batch::footprint(*d);
        *free_blocks = (long long)b->b->free_slots.size();
    });
}

int spd_batch_add_case(spd_batch_t b, const spd_case_desc* d, int* id)
{
    return guarded([&]
    {
        require(b && d && id, "null argument");
        *id = b->b->add_case(*d);
    });
}

int spd_batch_remove_case(spd_batch_t b, int id)
{
    return guarded([&] { require(b != nullptr, "null argument"); b->b->remove_case(id); });
}

int spd_batch_step(spd_batch_t b, int nsteps)
{
    return guarded([&] { require(b != nullptr, "null argument"); b->b->step(nsteps); });
}

int spd_batch_case_info(spd_batch_t b, int id, long long* cells, int* steps, double* dt)
{
    return guarded([&]
    {
        require(b && cells && steps && dt, "null argument");
        const BatchCase& c = b->b->get(id);
        *cells = 4096LL*(long long)c.slots.size();
        *steps = c.steps_done;
        *dt = c.dt;
    });
}

int spd_batch_case_status(spd_batch_t b, int id, int* diverged_step)
{
    return guarded([&]
    {
        require(b && diverged_step, "null argument");
        *diverged_step = b->b->get(id).diverged_step;
    });
}

int spd_batch_case_download(spd_batch_t b, int id, double* host)
{
    return guarded([&] { require(b && host, "null argument"); b->b->download(id, host); });
}

int spd_batch_case_upload(spd_batch_t b, int id, const double* host)
{
    return guarded([&] { require(b && host, "null argument"); b->b->upload(id, host); });
}

int spd_batch_case_diag(spd_batch_t b, int id, double* ke, double* en, int max, int* n)
{
    return guarded([&]
    {
        require(b && ke && en && n, "null argument");
        const BatchCase& c = b->b->get(id);
        const int m = std::min<int>(max, (int)c.ke.size());
        for (int t = 0; t < m; ++t) { ke[t] = c.ke[t]; en[t] = c.en[t]; }
        *n = m;
    });
}

int spd_batch_tune(spd_batch_t b, int* grid, int* interleaved)
{
    return guarded([&]
    {
        require(b && grid && interleaved, "null argument");
        *grid = b->b->tune();
        *interleaved = b->b->order_interleaved ? 1 : 0;
    });
}

int spd_batch_use_graph(spd_batch_t b, int enable)
{
    return guarded([&] { require(b != nullptr, "null argument"); b->b->use_graph = enable != 0; });
}

int spd_batch_set_shape(spd_batch_t b, int grid, int interleaved)
{
    return guarded([&]
    {
        require(b != nullptr, "null argument");
        require(grid >= 1 && grid <= 65536, "grid must be in 1..65536");
        b->b->grid = grid;
        if (b->b->order_interleaved != (interleaved != 0)) { b->b->order_interleaved = interleaved != 0; b->b->worklist_dirty = true; }
    });
}

int spd_batch_set_kernel_shape(spd_batch_t b, int rows, int prefetch, int occupancy)
{
    return guarded([&]
    {
        require(b != nullptr, "null argument");
        require(rows == 16 || rows == 8, "rows must be 16 or 8");
        require(prefetch == 1 || prefetch == 2, "prefetch must be 1 or 2");
        require(occupancy == 2 || occupancy == 3, "occupancy must be 2 or 3");
        b->b->variant = spd::batch_variant(rows, prefetch, occupancy);
    });
}

int spd_batch_tune_log(spd_batch_t b, spd_tune_entry* out, int max, int* n)
{
    return guarded([&]
    {
        require(b && n && (out || max == 0), "null argument");
        require(max >= 0, "max must be >= 0");
        const auto& log = b->b->tune_log;
        const int m = std::min<int>(max, (int)log.size());
        for (int e = 0; e < m; ++e)
        {
            const int v = log[e].variant;
            out[e] = spd_tune_entry{log[e].grid, log[e].interleaved, spd::batch_variant_rows(v),
                                    spd::batch_variant_prefetch(v), spd::batch_variant_occupancy(v), log[e].ms_per_step};
        }
        *n = m;
    });
}

int spd_batch_get_shape(spd_batch_t b, int* grid, int* interleaved, int* rows, int* prefetch, int* occupancy)
{
    return guarded([&]
    {
        require(b && grid && interleaved && rows && prefetch && occupancy, "null argument");
        *grid = b->b->grid;
        *interleaved = b->b->order_interleaved ? 1 : 0;
        *rows = spd::batch_variant_rows(b->b->variant);
        *prefetch = spd::batch_variant_prefetch(b->b->variant);
        *occupancy = spd::batch_variant_occupancy(b->b->variant);
    });
}

} // extern "C"

namespace spd {

void validate_grid_desc(const spd_grid_desc* d)
{
    require(d != nullptr, "grid description is null");
    for (int a = 0; a < 3; ++a)
    {
        require(d->num_blocks[a] >= 1, "num_blocks must be >= 1");
        require(d->cells_in_block[a] >= 1, "cells_in_block must be >= 1");
        require(d->hi[a] > d->lo[a], "bounds must satisfy hi > lo");
        require(d->map[a].kind == SPD_MAP_IDENTITY || d->map[a].kind == SPD_MAP_TANH, "unknown coordinate map");
        if (d->map[a].kind == SPD_MAP_TANH)
            require(d->map[a].p1 > d->map[a].p0 && d->map[a].beta > 0.0, "tanh map needs p1 > p0 and beta > 0");
    }
    require(d->num_exch >= 1 && d->num_exch <= 4, "num_exch must be in 1..4");
    for (int a = 0; a < 3; ++a) require(d->cells_in_block[a] >= d->num_exch, "blocks must be at least num_exch cells wide");
    // per-block kernels put the block index in the second grid dimension
    require((long long)d->num_blocks[0]*d->num_blocks[1]*d->num_blocks[2] <= 65535, "at most 65535 blocks per grid");
}

double map_coord(const spd_map1d& m, double xi)
{
    if (m.kind != SPD_MAP_TANH) return xi;
    const double c = 0.5*(m.p1 + m.p0), h = 0.5*(m.p1 - m.p0);
    return c + h*std::tanh(m.beta*((xi - c)/h))/std::tanh(m.beta);
}

void upload_blocks(spd_array_s* arr, int lb0, int nb, const double* host, bool sync)
{
    Grid& g = *arr->a.grid;
    require(lb0 >= 0 && nb >= 0 && lb0 + nb <= g.geom.nblk, "block range outside the grid");
    double* stage = ensure_stage(g);
    const long long cells = (long long)g.geom.N[0]*g.geom.N[1]*g.geom.N[2];
    for (int b0 = lb0; b0 < lb0 + nb; b0 += g.stage_blocks)
    {
        const int n = std::min(g.stage_blocks, lb0 + nb - b0);
        SPD_HIP(hipMemcpyAsync(stage, host + 5*cells*(b0 - lb0), sizeof(double)*5*cells*n, hipMemcpyHostToDevice,
                               g.ctx->stream));
        launch_aos_to_soa(g, stage, arr->a.d, b0, n, g.ctx->stream);
    }
    if (sync) SPD_HIP(hipStreamSynchronize(g.ctx->stream));
}

void download_blocks(const spd_array_s* arr, int lb0, int nb, double* host, bool sync)
{
    Grid& g = *arr->a.grid;
    require(lb0 >= 0 && nb >= 0 && lb0 + nb <= g.geom.nblk, "block range outside the grid");
    double* stage = ensure_stage(g);
    const long long cells = (long long)g.geom.N[0]*g.geom.N[1]*g.geom.N[2];
    for (int b0 = lb0; b0 < lb0 + nb; b0 += g.stage_blocks)
    {
        const int n = std::min(g.stage_blocks, lb0 + nb - b0);
        launch_soa_to_aos(g, arr->a.d, stage, b0, n, g.ctx->stream);
        SPD_HIP(hipMemcpyAsync(host + 5*cells*(b0 - lb0), stage, sizeof(double)*5*cells*n, hipMemcpyDeviceToHost,
                               g.ctx->stream));
    }
    if (sync) SPD_HIP(hipStreamSynchronize(g.ctx->stream));
}

} // namespace spd
