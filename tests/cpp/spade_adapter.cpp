// The adapter a SPADE user adds to run the right-hand-side hot path through
// spade_amd (the C++ facade include/spade_amd.hpp over the C ABI): a device
// twin of a SPADE grid, flux_div of a CPU-resident SPADE grid_array on the GPU,
// and an integrator whose residual and boundary are SPADE-style callables.
//
// It is compiled and linked against the SPADE headers where they lie by
// tests/test_spade_adapter.py (CPU, where the SPADE sources are present); run
// without arguments it builds a SPADE grid and array and checks, without a
// GPU, that the device description and the host layout agree with SPADE's:
// the grid's blocks, cells, bounds and periodicity, and the size of the
// array's data (AoS, variable fastest, ghosts stored: mem_map.h:654-676).
#include <array>
#include <cstdio>
#include <cstdlib>
#include <spade.h>
#include "spade_amd.hpp"

namespace adapter {

// Device description of a SPADE cartesian grid (grid/cartesian_grid.h:84-96).
template <typename grid_t>
spade_amd::grid_desc device_grid_desc(const grid_t& g, int nexch, std::array<bool, 3> periodic)
{
    spade_amd::grid_desc d;
    const auto bnd = g.get_bounds();
    for (int a = 0; a < 3; ++a)
    {
        d.num_blocks[a]     = g.get_blocks().num_blocks[a];
        d.cells_in_block[a] = g.get_num_cells(a);
        d.lo[a] = bnd.min(a);
        d.hi[a] = bnd.max(a);
        d.periodic[a] = periodic[a];
    }
    d.num_exch = nexch;
    return d;
}

template <typename grid_t>
spade_amd::grid make_device_grid(spade_amd::context& ctx, const grid_t& g, int nexch, std::array<bool, 3> periodic)
{
    return spade_amd::grid(ctx, device_grid_desc(g, nexch, periodic));
}

// pde_algs::flux_div(prims, rhs, flux) for a CPU-resident SPADE array, on the
// GPU: the host data already has the library's boundary layout.
template <typename arr_t>
void flux_div_gpu(spade_amd::grid& dg, const arr_t& prims, arr_t& rhs, const spade_amd::flux& f, bool increment)
{
    spade_amd::grid_array q(dg), r(dg);
    q.upload(prims.data.data());
    if (increment) r.upload(rhs.data.data());
    spade_amd::flux_div(q, r, f, increment);
    r.download(rhs.data.data());
}

// integrator_t(axis, rk4_t, data, rhs, bc, state_transform) with the state on
// the device: the residual is flux_div of the stored primitives, the boundary
// the exchange (the defaults), or any callables given here.
inline spade_amd::rk_integrator_t make_rk4(spade_amd::grid_array& q, double dt, const spade_amd::ideal_gas_t& gas,
                                           const spade_amd::flux& f)
{
    return spade_amd::rk_integrator_t(q, spade_amd::rk_table(SPD_RK4), dt, spade_amd::state_transform, gas,
                                      [&f](spade_amd::grid_array& r, const spade_amd::grid_array& s, double)
                                      { spade_amd::flux_div(s, r, f, false); });
}

} // namespace adapter

int main(int argc, char** argv)
{
    spade::parallel::mpi_t group(&argc, &argv);
    spade::coords::identity<double> coords;
    spade::ctrs::array<int, 3> num_blocks(2, 3, 2), cells(4, 8, 4), nexch(2, 2, 2);
    spade::bound_box_t<double, 3> bounds;
    for (int d = 0; d < 3; ++d) { bounds.min(d) = 0.5*d; bounds.max(d) = 1.0 + d; }
    spade::grid::cartesian_blocks_t blocks(num_blocks, bounds);
    spade::grid::cartesian_grid_t grid(cells, blocks, coords, group);
    spade::fluid_state::prim_t<double> fill = 0.0;
    spade::grid::grid_array prims(grid, fill, nexch, spade::device::cpu);

    const auto d = adapter::device_grid_desc(grid, 2, {true, false, true});
    int bad = 0;
    for (int a = 0; a < 3; ++a)
    {
        bad += d.num_blocks[a] != num_blocks[a];
        bad += d.cells_in_block[a] != cells[a];
        bad += d.lo[a] != bounds.min(a) || d.hi[a] != bounds.max(a);
    }
    bad += d.periodic[0] != 1 || d.periodic[1] != 0 || d.periodic[2] != 1 || d.num_exch != 2;
    const std::size_t stored = 5*std::size_t(2*3*2)*(4 + 4)*(8 + 4)*(4 + 4);
    bad += prims.data.size() != stored;
    std::printf("adapter: %d mismatches, %zu doubles in the array\n", bad, prims.data.size());
    (void)&adapter::flux_div_gpu<decltype(prims)>;
    (void)&adapter::make_device_grid<decltype(grid)>;
    (void)&adapter::make_rk4;
    return bad == 0 ? 0 : 1;
}
