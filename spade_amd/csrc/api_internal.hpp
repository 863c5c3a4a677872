// Handles and error plumbing shared by the translation units of the C ABI.
#pragma once
#include "spd_internal.hpp"

struct spd_ctx_s   { spd::Ctx c; };
struct spd_grid_s  { spd::Grid g; };
struct spd_array_s { spd::Array a; };

// An integrator: the fixed SSP-RK3 forms (scheme SPD_SCHEME_*, residuals r),
// or a general explicit table (generic = true) with its callbacks
struct spd_integrator_s
{
    spd_array_s* q = nullptr;
    spd_flux_desc flux{};
    spd_bc_desc bc{};
    int scheme = 0;
    double dt = 0;
    spd_array_s* r[3] = {nullptr, nullptr, nullptr};
    // general table
    bool generic = false;
    spd_rk_table table{};
    int trans = SPD_TRANS_PRIM_CONS;
    double gamma = 0, R = 0, t = 0;
    spd_rhs_fn rhs = nullptr;
    spd_bc_fn bcfn = nullptr;
    void* user = nullptr;
    bool has_flux = false, has_walls = false;
    spd_array_s* res[SPD_RK_MAX_STAGES] = {};
    spd_array_s* q2 = nullptr;          // second solution copy (high storage, identity transform)
    spd_array_s* alt = nullptr;         // fused SSP-RK3 stages: the other state buffer (ping-pong)
};

namespace spd {

// a file could not be opened, read or written
struct IoError { std::string msg; };
// a valid request the library does not implement
struct UnsupportedError { std::string msg; };

void set_error(const std::string& msg);

inline void require(bool ok, const std::string& msg) { if (!ok) throw ArgError{msg}; }

// Runs fn and maps what it throws to the SPD_* status codes (the message is
// kept for spd_last_error()).
template <typename F> int guarded(F&& fn)
{
    try
    {
        fn();
        return SPD_OK;
    }
    catch (const ArgError& e) { set_error(e.msg); return SPD_ERR_ARG; }
    catch (const IoError& e) { set_error(e.msg); return SPD_ERR_IO; }
    catch (const UnsupportedError& e) { set_error(e.msg); return SPD_ERR_UNSUPPORTED; }
    catch (const HipError& e)
    {
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return e.code == hipErrorOutOfMemory ? SPD_ERR_MEMORY : SPD_ERR_HIP;
    }
    catch (const std::bad_alloc&) { set_error("host allocation failed"); return SPD_ERR_MEMORY; }
    catch (const std::exception& e) { set_error(e.what()); return SPD_ERR_ARG; }
}

// Checks a grid description (sizes, bounds, maps, ghost layers).
void validate_grid_desc(const spd_grid_desc* d);
// Checks a flux descriptor for a grid with ng ghost layers.
void validate_flux(const spd_flux_desc* f, int ng);
// Device arrays of a grid (zeroed) and their release (after the stream drains).
spd_array_s* new_array(Grid& g);
void free_array(spd_array_s* A);
// flux_div with the kernel the library selects
void run_flux_div(const Grid& g, const double* prims, double* rhs, const spd_flux_desc& f, bool increment);
// The wall fills of the fixed integrators after a stage (the exchange is fused into the stage)
void apply_bc(spd_integrator_s* it, double* a = nullptr);   // a: the array to fill (default: the state)
// One step of a general-table integrator
void advance_generic(spd_integrator_s* it);

// Box minimum and cell size of block lb along axis a, as cartesian_blocks_t
// and grid_geometry_t compute them (grid/cartesian_blocks.h:46-71,
// grid/grid_geometry.h:74-85): bsize = (hi - lo)/nb, min = lo + b bsize,
// max = min + bsize, dx = (max - min)/nc.
inline void block_box(const spd_grid_desc& d, int lb, int a, double& mn, double& dx)
{
    const int b[3] = {lb % d.num_blocks[0], (lb / d.num_blocks[0]) % d.num_blocks[1],
                      lb / (d.num_blocks[0]*d.num_blocks[1])};
    const double bsize = (d.hi[a] - d.lo[a])/d.num_blocks[a];
    mn = d.lo[a] + b[a]*bsize;
    const double mx = mn + bsize;
    dx = (mx - mn)/d.cells_in_block[a];
}

// Physical coordinate of computational coordinate xi along axis a: identity,
// or x = c + h tanh(beta (xi - c)/h)/tanh(beta) with c, h the centre and
// half-width of [p0, p1].
double map_coord(const spd_map1d& m, double xi);

// Blocks [lb0, lb0 + nb) of an array from / to host memory in the reference
// layout (AoS, variable fastest, ghosts stored), through the grid's staging
// buffer; both return when the host memory may be reused.
// sync = false leaves the copies in flight on the context's stream (the host
// buffer must then stay untouched until the stream has passed them)
void upload_blocks(spd_array_s* arr, int lb0, int nb, const double* host, bool sync = true);
void download_blocks(const spd_array_s* arr, int lb0, int nb, double* host, bool sync = true);

} // namespace spd
