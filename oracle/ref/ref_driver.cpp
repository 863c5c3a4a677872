// Golden-vector driver for the SPADE reference (C++20, header-only).
//
// This file is NOT part of the product.  It is compiled against the reference
// headers where they lie (see oracle/ref/Makefile) and writes small binary
// fixtures that pin the CPU oracle in oracle/spade_oracle.cpp:
//   * flux_div (basic, overwrite) for every convective / viscous functor on
//     the hot path, on a non-cubic periodic multi-block grid;
//   * the periodic 26-neighbour ghost exchange and a wall boundary_fill;
//   * the transform-aware SSP-RK3 step and the low-storage ssprk3_opt step.
//
// Output files are raw little-endian float64, interior cells only, in the
// reference's memory order (lb, k, j, i, var) unless the name ends in
// "_ghost" (then the full ghosted block, same order).  A text manifest lists
// every file with its shape and the run parameters.
#include <cstdio>
#include <cmath>
#include <string>
#include <vector>
#include <fstream>
#include <chrono>
#include <algorithm>
#include "spade.h"

using real_t = double;
using prim_t = spade::fluid_state::prim_t<real_t>;
using flux_t = spade::fluid_state::flux_t<real_t>;
using cons_t = spade::fluid_state::cons_t<real_t>;

// Driver-side power law: the reference power_law_t has no get_all(), which
// visc_lr requires, so the plugin adds it from the law's own three getters.
template <typename dtype> struct power_law_all_t : public spade::viscous_laws::power_law_t<dtype>
{
    using base = spade::viscous_laws::power_law_t<dtype>;
    using result_type = spade::viscous_laws::visc_result_t<dtype>;
    power_law_all_t(dtype mu_ref, dtype t_ref, dtype n, dtype pr) : base(mu_ref, t_ref, n, pr) {}
    result_type get_all(const auto& input) const
    {
        return result_type{this->get_visc(input), this->get_beta(input), this->get_diffuse(input)};
    }
};

// cartesian_blocks_t leaves the false domain-boundary flags uninitialised
// (grid/cartesian_blocks.h:51-64 only ever writes true), which makes the
// exchange skip rule and boundary_fill read garbage on non-periodic grids.
// The driver rewrites every flag with the rule the constructor intends.
template <typename blocks_t> static void fix_boundary_flags(blocks_t& blocks)
{
    for (std::size_t lb = 0; lb < blocks.block_is_domain_boundary.size(); ++lb)
    {
        int b[3] = {int(lb % blocks.num_blocks[0]), int((lb / blocks.num_blocks[0]) % blocks.num_blocks[1]),
                    int(lb / (blocks.num_blocks[0]*blocks.num_blocks[1]))};
        for (int d = 0; d < 3; ++d)
        {
            blocks.block_is_domain_boundary[lb].min(d) = (b[d] == 0);
            blocks.block_is_domain_boundary[lb].max(d) = (b[d] == blocks.num_blocks[d] - 1);
        }
    }
}

static std::string g_outdir;
static std::ofstream g_manifest;

template <typename arr_t>
static void dump(const std::string& name, const arr_t& arr, bool ghosts = false)
{
    const auto& grid = arr.get_grid();
    const auto img   = arr.image();
    const int ng     = ghosts ? arr.get_num_exchange(0) : 0;
    const int nx = grid.get_num_cells(0), ny = grid.get_num_cells(1), nz = grid.get_num_cells(2);
    const int nlb = grid.get_num_local_blocks();
    std::vector<double> buf;
    for (int lb = 0; lb < nlb; ++lb)
    for (int k = -ng; k < nz + ng; ++k)
    for (int j = -ng; j < ny + ng; ++j)
    for (int i = -ng; i < nx + ng; ++i)
    {
        spade::grid::cell_idx_t idx(i, j, k, lb);
        const auto e = img.get_elem(idx);
        for (int v = 0; v < 5; ++v) buf.push_back(e[v]);
    }
    const std::string fn = g_outdir + "/" + name + ".bin";
    std::FILE* f = std::fopen(fn.c_str(), "wb");
    std::fwrite(buf.data(), sizeof(double), buf.size(), f);
    std::fclose(f);
    g_manifest << "file " << name << " " << nlb << " " << (nz + 2*ng) << " " << (ny + 2*ng) << " " << (nx + 2*ng) << " 5\n";
}

// Deterministic roughness added to the interior (exercises the WENO weights
// and every gradient component; the perturbed field itself is stored, so
// nothing downstream needs to regenerate it).
static real_t rough(int i, int j, int k, int v)
{
    const double s = std::sin(12.9898*i + 78.233*j + 37.719*k + 4.581*v) * 43758.5453;
    return 2.0*(s - std::floor(s)) - 1.0;
}

// Timing mode: flux_div (basic) with cent_keep<2> + visc_lr(power law) on a
// TGV grid of nb^3 blocks of 16^3 cells, repeated reps times on one core.
// Prints one JSON line with the throughput in Mcell-RHS/s.
static int run_bench(spade::parallel::mpi_t& group, int nb, int reps)
{
    const real_t gamma = 1.4, mach = 0.1, u0 = 1.0, rho0 = 1.0, T0 = 1.0;
    const real_t p0 = rho0*(u0/mach)*(u0/mach)/gamma, r_gas = p0/(rho0*T0);
    spade::fluid_state::ideal_gas_t<real_t> air(gamma, r_gas);
    spade::coords::identity<real_t> coords;
    spade::ctrs::array<int, 3> num_blocks(nb, nb, nb), cells(16, 16, 16), nexch(2, 2, 2);
    spade::bound_box_t<real_t, 3> bounds;
    for (int d = 0; d < 3; ++d) { bounds.min(d) = 0.0; bounds.max(d) = 2.0*spade::consts::pi; }
    spade::grid::cartesian_blocks_t blocks(num_blocks, bounds);
    fix_boundary_flags(blocks);
    spade::grid::cartesian_grid_t grid(cells, blocks, coords, group);
    prim_t fill1 = 0.0;
    flux_t fill2 = 0.0;
    spade::grid::grid_array prim(grid, fill1, nexch, spade::device::cpu);
    spade::grid::grid_array rhs (grid, fill2, nexch, spade::device::cpu);
    spade::algs::fill_array(prim, [&](const spade::coords::point_t<real_t>& x) -> prim_t
    {
        prim_t q;
        q.p() = p0 + (rho0*u0*u0/16.0)*(std::cos(2.0*x[0]) + std::cos(2.0*x[1]))*(std::cos(2.0*x[2]) + 2.0);
        q.T() = T0;
        q.u() =  u0*std::sin(x[0])*std::cos(x[1])*std::cos(x[2]);
        q.v() = -u0*std::cos(x[0])*std::sin(x[1])*std::cos(x[2]);
        q.w() = 0.0;
        return q;
    });
    power_law_all_t<real_t> visc_pow(rho0*u0/1600.0, T0, 0.76, 0.72);
    spade::viscous::visc_lr vsc_pow(visc_pow, air);
    const auto flx = spade::omni::compose(spade::convective::cent_keep<2>(air), vsc_pow);
    rhs = 0.0;
    spade::pde_algs::flux_div(prim, rhs, flx);   // warm-up
    std::vector<double> times;
    for (int r = 0; r < reps; ++r)
    {
        const auto t0 = std::chrono::steady_clock::now();
        rhs = 0.0;
        spade::pde_algs::flux_div(prim, rhs, flx);
        const auto t1 = std::chrono::steady_clock::now();
        times.push_back(std::chrono::duration<double>(t1 - t0).count());
    }
    std::sort(times.begin(), times.end());
    const double med = times[times.size()/2];
    const double cells_tot = double(nb)*nb*nb*4096.0;
    std::printf("{\"kind\": \"reference\", \"cells\": %.0f, \"reps\": %d, \"median_s\": %.6f, \"mcell_rhs_per_s\": %.4f}\n",
                cells_tot, reps, med, cells_tot/med*1e-6);
    return 0;
}

int main(int argc, char** argv)
{
    spade::parallel::mpi_t group(&argc, &argv);
    if (argc > 1 && std::string(argv[1]) == "--bench")
        return run_bench(group, argc > 2 ? std::atoi(argv[2]) : 2, argc > 3 ? std::atoi(argv[3]) : 5);
    g_outdir = (argc > 1) ? argv[1] : ".";
    g_manifest.open(g_outdir + "/manifest.txt");
    g_manifest.precision(17);

    const real_t gamma = 1.4, mach = 0.1, u0 = 1.0, rho0 = 1.0, T0 = 1.0, L = 1.0;
    const real_t c0 = u0/mach;
    const real_t p0 = rho0*c0*c0/gamma;
    const real_t r_gas = p0/(rho0*T0);
    const real_t reynolds = 1600.0, prandtl = 0.72, power = 0.76;
    const real_t mu_ref = rho0*u0*L/reynolds;
    g_manifest << "param gamma " << gamma << "\nparam R " << r_gas << "\nparam p0 " << p0
               << "\nparam mu_ref " << mu_ref << "\nparam T_ref " << T0 << "\nparam power " << power
               << "\nparam prandtl " << prandtl << "\n";

    spade::fluid_state::ideal_gas_t<real_t> air(gamma, r_gas);
    spade::coords::identity<real_t> coords;
    const real_t twopi = 2.0*spade::consts::pi;

    auto tgv = [&](const spade::coords::point_t<real_t>& x) -> prim_t
    {
        prim_t q;
        q.p() = p0 + (rho0*u0*u0/16.0)*(std::cos(2.0*x[0]/L) + std::cos(2.0*x[1]/L))*(std::cos(2.0*x[2]/L) + 2.0);
        q.T() = T0;
        q.u() =  u0*std::sin(x[0]/L)*std::cos(x[1]/L)*std::cos(x[2]/L);
        q.v() = -u0*std::cos(x[0]/L)*std::sin(x[1]/L)*std::cos(x[2]/L);
        q.w() = 0.0;
        return q;
    };

    //------------------------------------------------------------------
    // Grid G1: 2x3x2 blocks of 4x8x4 cells, ng=2, [0,2pi]^3, periodic.
    //------------------------------------------------------------------
    {
        spade::ctrs::array<int, 3> num_blocks(2, 3, 2);
        spade::ctrs::array<int, 3> cells_in_block(4, 8, 4);
        spade::ctrs::array<int, 3> nexch(2, 2, 2);
        spade::bound_box_t<real_t, 3> bounds;
        for (int d = 0; d < 3; ++d) { bounds.min(d) = 0.0; bounds.max(d) = twopi; }
        spade::grid::cartesian_blocks_t blocks(num_blocks, bounds);
        fix_boundary_flags(blocks);
        spade::grid::cartesian_grid_t grid(cells_in_block, blocks, coords, group);
        prim_t fill1 = 0.0;
        flux_t fill2 = 0.0;
        spade::grid::grid_array prim(grid, fill1, nexch, spade::device::cpu);
        spade::grid::grid_array rhs (grid, fill2, nexch, spade::device::cpu);
        spade::algs::fill_array(prim, tgv);
        {
            auto img = prim.image();
            for (int lb = 0; lb < (int)grid.get_num_local_blocks(); ++lb)
            {
                const auto box = grid.get_blocks().get_block_box(lb);
                for (int k = 0; k < 4; ++k) for (int j = 0; j < 8; ++j) for (int i = 0; i < 4; ++i)
                {
                    // global cell index from the block box (x fastest block order checked below)
                    const int gi = int(std::lround(box.min(0)/(twopi/8.0))) + i;
                    const int gj = int(std::lround(box.min(1)/(twopi/24.0))) + j;
                    const int gk = int(std::lround(box.min(2)/(twopi/8.0))) + k;
                    spade::grid::cell_idx_t idx(i, j, k, lb);
                    auto q = img.get_elem(idx);
                    q.p() += 0.05*p0*0.01*rough(gi, gj, gk, 0);
                    q.T() += 0.01*T0*rough(gi, gj, gk, 1);
                    q.u() += 0.05*u0*rough(gi, gj, gk, 2);
                    q.v() += 0.05*u0*rough(gi, gj, gk, 3);
                    q.w() += 0.05*u0*rough(gi, gj, gk, 4);
                    img.set_elem(idx, q);
                }
                g_manifest << "g1_block " << lb << " " << box.min(0) << " " << box.min(1) << " " << box.min(2) << "\n";
            }
        }
        auto handle = spade::grid::create_exchange(prim, group, spade::ctrs::array<bool,3>(true, true, true));
        handle.exchange(prim);
        dump("g1_prim", prim);
        dump("g1_prim_ghost", prim, true);
        // the reference's parallel VTK StructuredGrid output of the same array
        // (io/io_vtk.h:276-288): g1.pvts plus data_g1/b<lb>.vts
        spade::io::output_vtk(g_outdir + "/vtk_g1", "g1", prim);

        spade::viscous_laws::constant_viscosity_t<real_t> visc_const(mu_ref, prandtl);
        power_law_all_t<real_t> visc_pow(mu_ref, T0, power, prandtl);
        spade::viscous::visc_lr vsc_const(visc_const, air);
        spade::viscous::visc_lr vsc_pow(visc_pow, air);
        spade::state_sensor::const_sensor_t<real_t> s_const(0.7);
        spade::state_sensor::ducros_t<real_t> s_ducros(1e-3);
        spade::convective::pressure_diss_lr pd_const(air, s_const, 0.3, 0.2);
        spade::convective::pressure_diss_lr pd_ducros(air, s_ducros, 0.3, 0.2);
        spade::convective::rusanov_t<decltype(air)> rus(air);
        spade::convective::phys_flux_t<decltype(air)> phys(air);

        auto run = [&](const std::string& name, const auto& flx)
        {
            rhs = 0.0;
            spade::pde_algs::flux_div(prim, rhs, flx);
            dump("g1_rhs_" + name, rhs);
        };
        run("keep2",    spade::convective::cent_keep<2>(air));
        run("keep4",    spade::convective::cent_keep<4>(air));
        run("totani",   spade::convective::totani_lr<decltype(air)>(air));
        run("visc_const", vsc_const);
        run("visc_pow", vsc_pow);
        run("keep2_visc_pow", spade::omni::compose(spade::convective::cent_keep<2>(air), vsc_pow));
        run("pdiss_const", pd_const);
        run("pdiss_ducros", pd_ducros);
        run("weno_rus", spade::convective::weno_t<decltype(rus)>(rus));
        run("weno_rus_lin", spade::convective::weno_t<decltype(rus), spade::convective::disable_smooth>(rus));
        run("weno_phys", spade::convective::weno_t<decltype(phys)>(phys));
        run("fweno", spade::convective::fweno_t<decltype(air)>(air));
        run("fweno_lin", spade::convective::fweno_t<decltype(air), spade::convective::disable_smooth>(air));
        run("first_order", spade::convective::first_order_t<decltype(air)>(air));
        // face-reconstructed WENO with the Rusanov flux-difference splitting
        // (convective.h:499-556, flux_funcs.h:54-95)
        spade::convective::rusanov_fds_t<decltype(air)> rus_fds(air);
        run("weno_fds", spade::convective::weno_fds_t<decltype(rus_fds)>(rus_fds));
        run("weno_fds_lin", spade::convective::weno_fds_t<decltype(rus_fds), spade::convective::disable_smooth>(rus_fds));
        run("keep2_pdiss_visc_const", spade::omni::compose(spade::convective::cent_keep<2>(air), pd_ducros, vsc_const));
        // sgs_visc_t: a laminar law plus the WALE subgrid viscosity
        // (viscous_laws.h:174-216, subgrid_scale.h:24-86), C_w 0.5, filter
        // width 0.3, turbulent Prandtl number 0.9
        spade::subgrid_scale::wale_t<real_t, decltype(air)> wale(air, 0.5, 0.3, 0.9);
        spade::viscous_laws::sgs_visc_t sgs_const(visc_const, wale);
        spade::viscous_laws::sgs_visc_t sgs_pow(visc_pow, wale);
        spade::viscous::visc_lr vsc_sgs_const(sgs_const, air);
        spade::viscous::visc_lr vsc_sgs_pow(sgs_pow, air);
        run("visc_sgs_const", vsc_sgs_const);
        run("keep2_visc_sgs_pow", spade::omni::compose(spade::convective::cent_keep<2>(air), vsc_sgs_pow));
        // central / WENO blends (navier-stokes/hybrid_scheme.h)
        run("hybrid_keep2_weno_ducros_full",
            spade::convective::hybrid_scheme_t(spade::convective::cent_keep<2>(air), spade::convective::weno_t<decltype(rus)>(rus),
                                               s_ducros, spade::convective::full_flux));
        run("hybrid_keep2_fweno_const_diss",
            spade::convective::hybrid_scheme_t(spade::convective::cent_keep<2>(air), spade::convective::fweno_t<decltype(air)>(air),
                                               s_const, spade::convective::diss_flux));
        run("hybrid_keep4_weno_ducros_full_visc",
            spade::omni::compose(spade::convective::hybrid_scheme_t(spade::convective::cent_keep<4>(air),
                                                                    spade::convective::weno_t<decltype(rus)>(rus),
                                                                    s_ducros, spade::convective::full_flux), vsc_const));

        // increment mode: rhs accumulates onto its existing content
        {
            auto img = rhs.image();
            for (int lb = 0; lb < (int)grid.get_num_local_blocks(); ++lb)
            for (int k = 0; k < 4; ++k) for (int j = 0; j < 8; ++j) for (int i = 0; i < 4; ++i)
            {
                spade::grid::cell_idx_t idx(i, j, k, lb);
                flux_t f;
                for (int v = 0; v < 5; ++v) f[v] = 0.125*rough(i + 4*lb, j, k, v + 7);
                img.set_elem(idx, f);
            }
            dump("g1_rhs_incr_init", rhs);
            spade::pde_algs::flux_div(prim, rhs, spade::convective::cent_keep<2>(air));
            dump("g1_rhs_incr_keep2", rhs);
        }
    }

    //------------------------------------------------------------------
    // Grid G3: same layout, periodic in x/z only, wall ghosts in y by
    // boundary_fill (no-slip isothermal mirror).  Checks the exchange
    // skipping domain faces, the mirror index rule, and flux_div on top.
    //------------------------------------------------------------------
    {
        spade::ctrs::array<int, 3> num_blocks(2, 3, 2);
        spade::ctrs::array<int, 3> cells_in_block(4, 8, 4);
        spade::ctrs::array<int, 3> nexch(2, 2, 2);
        spade::bound_box_t<real_t, 3> bounds;
        for (int d = 0; d < 3; ++d) { bounds.min(d) = 0.0; bounds.max(d) = twopi; }
        spade::grid::cartesian_blocks_t blocks(num_blocks, bounds);
        fix_boundary_flags(blocks);
        spade::grid::cartesian_grid_t grid(cells_in_block, blocks, coords, group);
        prim_t fill1 = 0.0;
        flux_t fill2 = 0.0;
        spade::grid::grid_array prim(grid, fill1, nexch, spade::device::cpu);
        spade::grid::grid_array rhs (grid, fill2, nexch, spade::device::cpu);
        // Every stored cell (ghosts too) gets a value that names it, so a
        // ghost that is wrongly left alone or wrongly copied shows up.
        {
            auto img = prim.image();
            for (int lb = 0; lb < (int)grid.get_num_local_blocks(); ++lb)
            for (int k = -2; k < 6; ++k) for (int j = -2; j < 10; ++j) for (int i = -2; i < 6; ++i)
            {
                spade::grid::cell_idx_t idx(i, j, k, lb);
                prim_t q;
                q.p() = p0 + 0.01*(lb + 1) + 0.001*(i + 2) + 0.0001*(j + 2) + 0.00001*(k + 2);
                q.T() = T0 + 0.02*rough(i, j, k + 16*lb, 1);
                q.u() = 0.3*rough(i, j, k + 16*lb, 2);
                q.v() = 0.3*rough(i, j, k + 16*lb, 3);
                q.w() = 0.3*rough(i, j, k + 16*lb, 4);
                img.set_elem(idx, q);
            }
        }
        dump("g3_prim_init_ghost", prim, true);
        auto handle = spade::grid::create_exchange(prim, group, spade::ctrs::array<bool,3>(true, false, true));
        handle.exchange(prim);
        dump("g3_prim_exch_ghost", prim, true);
        const real_t t_wall = 0.9;
        auto wall = [=](const prim_t& q_dom, const int& idir) -> prim_t
        {
            prim_t g;
            g.p() = q_dom.p();
            g.T() = 2.0*t_wall - q_dom.T();
            g.u() = -q_dom.u();
            g.v() = -q_dom.v();
            g.w() = -q_dom.w();
            return g;
        };
        g_manifest << "param t_wall " << t_wall << "\n";
        spade::algs::boundary_fill(prim, spade::boundary::ymin, wall);
        spade::algs::boundary_fill(prim, spade::boundary::ymax, wall);
        dump("g3_prim_wall_ghost", prim, true);
        spade::viscous_laws::constant_viscosity_t<real_t> visc_const(mu_ref, prandtl);
        spade::viscous::visc_lr vsc_const(visc_const, air);
        rhs = 0.0;
        spade::pde_algs::flux_div(prim, rhs, spade::omni::compose(spade::convective::cent_keep<2>(air), vsc_const));
        dump("g3_rhs_keep2_visc_const", rhs);
    }

    //------------------------------------------------------------------
    // Grid G4: same layout, no periodic axis.  After the exchange (which
    // leaves every domain face alone), boundary_fill with Lagrange
    // extrapolation on all six faces (face order xmin .. zmax matters at
    // edges and corners), extrapolation of order 3 on two faces, and a
    // four-argument user kernel (slip wall: normal velocity mirrored,
    // temperature isothermal, pressure blended with the ghost's own value).
    //------------------------------------------------------------------
    {
        spade::ctrs::array<int, 3> num_blocks(2, 3, 2);
        spade::ctrs::array<int, 3> cells_in_block(4, 8, 4);
        spade::ctrs::array<int, 3> nexch(2, 2, 2);
        spade::bound_box_t<real_t, 3> bounds;
        for (int d = 0; d < 3; ++d) { bounds.min(d) = 0.0; bounds.max(d) = twopi; }
        spade::grid::cartesian_blocks_t blocks(num_blocks, bounds);
        fix_boundary_flags(blocks);
        spade::grid::cartesian_grid_t grid(cells_in_block, blocks, coords, group);
        prim_t fill1 = 0.0;
        spade::grid::grid_array prim(grid, fill1, nexch, spade::device::cpu);
        auto init = [&]()
        {
            auto img = prim.image();
            for (int lb = 0; lb < (int)grid.get_num_local_blocks(); ++lb)
            for (int k = -2; k < 6; ++k) for (int j = -2; j < 10; ++j) for (int i = -2; i < 6; ++i)
            {
                spade::grid::cell_idx_t idx(i, j, k, lb);
                prim_t q;
                q.p() = p0 + 0.01*(lb + 1) + 0.001*(i + 2) + 0.0001*(j + 2) + 0.00001*(k + 2);
                q.T() = T0 + 0.02*rough(i, j, k + 16*lb, 1);
                q.u() = 0.3*rough(i, j, k + 16*lb, 2);
                q.v() = 0.3*rough(i, j, k + 16*lb, 3);
                q.w() = 0.3*rough(i, j, k + 16*lb, 4);
                img.set_elem(idx, q);
            }
            auto handle = spade::grid::create_exchange(prim, group, spade::ctrs::array<bool,3>(false, false, false));
            handle.exchange(prim);
        };
        const spade::boundary::identifier_t all_faces(true, true, true, true, true, true);
        init();
        dump("g4_prim_exch_ghost", prim, true);
        spade::algs::boundary_fill(prim, all_faces, spade::boundary::extrapolate<2>);
        dump("g4_prim_extrap2_ghost", prim, true);
        init();
        spade::algs::boundary_fill(prim, all_faces, spade::boundary::extrapolate<0>);
        dump("g4_prim_extrap0_ghost", prim, true);
        init();
        const spade::boundary::identifier_t xmin_ymax(true, false, false, true, false, false);
        spade::algs::boundary_fill(prim, xmin_ymax, spade::boundary::extrapolate<3>);
        dump("g4_prim_extrap3_xmin_ymax_ghost", prim, true);
        init();
        const real_t t_wall = 0.9;
        auto slip = [=](const prim_t& q_dom, const prim_t& q_fill, const spade::coords::point_t<real_t>& x, const int& idir) -> prim_t
        {
            prim_t g = q_dom;
            g.p() = 0.25*q_fill.p() + q_dom.p();
            g.T() = 2.0*t_wall - q_dom.T();
            g.u(idir) = -q_dom.u(idir);
            return g;
        };
        spade::algs::boundary_fill(prim, all_faces, slip);
        dump("g4_prim_slip_ghost", prim, true);
    }

    //------------------------------------------------------------------
    // Grid G2: TGV 16^3 as 2x2x2 blocks of 8^3, ng=2: cent_keep<2> +
    // visc_lr(power law), SSP-RK3 (transform-aware) 1 and 10 steps,
    // and the low-storage ssprk3_opt for 3 steps.
    //------------------------------------------------------------------
    {
        spade::ctrs::array<int, 3> num_blocks(2, 2, 2);
        spade::ctrs::array<int, 3> cells_in_block(8, 8, 8);
        spade::ctrs::array<int, 3> nexch(2, 2, 2);
        spade::bound_box_t<real_t, 3> bounds;
        for (int d = 0; d < 3; ++d) { bounds.min(d) = 0.0; bounds.max(d) = twopi; }
        spade::grid::cartesian_blocks_t blocks(num_blocks, bounds);
        fix_boundary_flags(blocks);
        spade::grid::cartesian_grid_t grid(cells_in_block, blocks, coords, group);
        prim_t fill1 = 0.0;
        flux_t fill2 = 0.0;
        spade::grid::grid_array prim(grid, fill1, nexch, spade::device::cpu);
        spade::grid::grid_array rhs (grid, fill2, nexch, spade::device::cpu);
        spade::algs::fill_array(prim, tgv);
        auto handle = spade::grid::create_exchange(prim, group, spade::ctrs::array<bool,3>(true, true, true));
        handle.exchange(prim);
        dump("g2_prim0", prim);

        power_law_all_t<real_t> visc_pow(mu_ref, T0, power, prandtl);
        spade::viscous::visc_lr vsc_pow(visc_pow, air);
        const auto flx = spade::omni::compose(spade::convective::cent_keep<2>(air), vsc_pow);

        // dt = CFL * dx / max(|u| + c) over interior cells at t = 0
        real_t umax = 0.0;
        {
            const auto img = prim.image();
            for (int lb = 0; lb < 8; ++lb)
            for (int k = 0; k < 8; ++k) for (int j = 0; j < 8; ++j) for (int i = 0; i < 8; ++i)
            {
                const auto q = img.get_elem(spade::grid::cell_idx_t(i, j, k, lb));
                const real_t s = std::sqrt(gamma*r_gas*q.T()) + std::sqrt(q.u()*q.u() + q.v()*q.v() + q.w()*q.w());
                umax = std::max(umax, s);
            }
        }
        const real_t dx = grid.get_dx(0, 0);
        const real_t dt = 0.5*dx/umax;
        g_manifest << "param g2_dt " << dt << "\n";

        rhs = 0.0;
        spade::pde_algs::flux_div(prim, rhs, flx);
        dump("g2_rhs0", rhs);

        auto bc = [&](auto& q, const auto& t) { handle.exchange(q); };
        auto calc_rhs = [&](auto& rhs_in, const auto& q, const auto& t) -> void
        {
            rhs_in = 0.0;
            spade::pde_algs::flux_div(q, rhs_in, flx);
        };
        cons_t transform_state;
        spade::fluid_state::state_transform_t trans(transform_state, air);
        {
            spade::time_integration::time_axis_t axis(0.0, dt);
            spade::time_integration::ssprk3_t alg;
            spade::time_integration::integrator_data_t q(prim, rhs, alg);
            spade::time_integration::integrator_t time_int(axis, alg, q, calc_rhs, bc, trans);
            for (int n = 0; n < 10; ++n)
            {
                time_int.advance();
                if (n == 0) dump("g2_ssprk3_step1", time_int.solution());
            }
            dump("g2_ssprk3_step10", time_int.solution());
        }
        {
            spade::time_integration::time_axis_t axis(0.0, dt);
            spade::time_integration::integrator_data_t q(prim, rhs, spade::time_integration::ssprk3_opt);
            spade::time_integration::integrator_t time_int(axis, spade::time_integration::ssprk3_opt, q, calc_rhs, bc, trans);
            for (int n = 0; n < 3; ++n) time_int.advance();
            dump("g2_ssprk3opt_step3", time_int.solution());
        }
    }
    //------------------------------------------------------------------
    // Grid G2 again: the other Butcher tables of explicit.h.
    //  * transform-aware advance (advance.h:233-280), primitive storage:
    //    rk2_t, rk4_t, ssprk34_t, rk38r_t, 2 steps each;
    //  * generic advance (advance.h:107-230) with the identity transform on
    //    conserved storage (the residual callback converts to primitives):
    //    rk4_t and rk38r_t (one solution copy), rk2hs_t and ssprk3hs_t (two),
    //    2 steps each.
    //------------------------------------------------------------------
    {
        spade::ctrs::array<int, 3> num_blocks(2, 2, 2);
        spade::ctrs::array<int, 3> cells_in_block(8, 8, 8);
        spade::ctrs::array<int, 3> nexch(2, 2, 2);
        spade::bound_box_t<real_t, 3> bounds;
        for (int d = 0; d < 3; ++d) { bounds.min(d) = 0.0; bounds.max(d) = twopi; }
        spade::grid::cartesian_blocks_t blocks(num_blocks, bounds);
        fix_boundary_flags(blocks);
        spade::grid::cartesian_grid_t grid(cells_in_block, blocks, coords, group);
        prim_t fill1 = 0.0;
        flux_t fill2 = 0.0;
        cons_t fill3 = 0.0;
        spade::grid::grid_array prim0(grid, fill1, nexch, spade::device::cpu);
        spade::grid::grid_array rhs (grid, fill2, nexch, spade::device::cpu);
        spade::algs::fill_array(prim0, tgv);
        auto handle = spade::grid::create_exchange(prim0, group, spade::ctrs::array<bool,3>(true, true, true));
        handle.exchange(prim0);
        power_law_all_t<real_t> visc_pow(mu_ref, T0, power, prandtl);
        spade::viscous::visc_lr vsc_pow(visc_pow, air);
        const auto flx = spade::omni::compose(spade::convective::cent_keep<2>(air), vsc_pow);
        // dt as for the G2 runs above (CFL 0.5 at t = 0)
        real_t umax = 0.0;
        {
            const auto img = prim0.image();
            for (int lb = 0; lb < 8; ++lb)
            for (int k = 0; k < 8; ++k) for (int j = 0; j < 8; ++j) for (int i = 0; i < 8; ++i)
            {
                const auto q = img.get_elem(spade::grid::cell_idx_t(i, j, k, lb));
                const real_t s = std::sqrt(gamma*r_gas*q.T()) + std::sqrt(q.u()*q.u() + q.v()*q.v() + q.w()*q.w());
                umax = std::max(umax, s);
            }
        }
        const real_t dt = 0.5*grid.get_dx(0, 0)/umax;
        auto calc_rhs = [&](auto& rhs_in, const auto& q, const auto& t) -> void
        {
            rhs_in = 0.0;
            spade::pde_algs::flux_div(q, rhs_in, flx);
        };
        auto bc = [&](auto& q, const auto& t) { handle.exchange(q); };
        cons_t transform_state;
        spade::fluid_state::state_transform_t trans(transform_state, air);
        auto run_trans = [&](const std::string& name, const auto& alg)
        {
            spade::grid::grid_array q = prim0;
            spade::time_integration::time_axis_t axis(0.0, dt);
            spade::time_integration::integrator_data_t data(q, rhs, alg);
            spade::time_integration::integrator_t time_int(axis, alg, data, calc_rhs, bc, trans);
            for (int n = 0; n < 2; ++n) time_int.advance();
            dump("g2_" + name + "_step2", time_int.solution());
        };
        run_trans("rk2", spade::time_integration::rk2_t());
        run_trans("rk4", spade::time_integration::rk4_t());
        run_trans("ssprk34", spade::time_integration::ssprk34_t());
        run_trans("rk38r", spade::time_integration::rk38r_t());

        // conserved storage, identity transform
        spade::grid::grid_array cons0(grid, fill3, nexch, spade::device::cpu);
        spade::grid::grid_array pscr (grid, fill1, nexch, spade::device::cpu);
        {
            const auto pimg = prim0.image();
            auto cimg = cons0.image();
            for (int lb = 0; lb < 8; ++lb)
            for (int k = -2; k < 10; ++k) for (int j = -2; j < 10; ++j) for (int i = -2; i < 10; ++i)
            {
                spade::grid::cell_idx_t idx(i, j, k, lb);
                cons_t c;
                spade::fluid_state::convert_state(pimg.get_elem(idx), c, air);
                cimg.set_elem(idx, c);
            }
        }
        auto cons_rhs = [&](auto& rhs_in, const auto& qc, const auto& t) -> void
        {
            const auto cimg = qc.image();
            auto pimg = pscr.image();
            for (int lb = 0; lb < 8; ++lb)
            for (int k = -2; k < 10; ++k) for (int j = -2; j < 10; ++j) for (int i = -2; i < 10; ++i)
            {
                spade::grid::cell_idx_t idx(i, j, k, lb);
                prim_t p;
                spade::fluid_state::convert_state(cimg.get_elem(idx), p, air);
                pimg.set_elem(idx, p);
            }
            rhs_in = 0.0;
            spade::pde_algs::flux_div(pscr, rhs_in, flx);
        };
        auto cons_bc = [&](auto& qc, const auto& t) { handle.exchange(qc); };
        auto run_ident = [&](const std::string& name, const auto& alg)
        {
            spade::grid::grid_array q = cons0;
            spade::time_integration::time_axis_t axis(0.0, dt);
            spade::time_integration::integrator_data_t data(q, rhs, alg);
            spade::time_integration::integrator_t time_int(axis, alg, data, cons_rhs, cons_bc);
            for (int n = 0; n < 2; ++n) time_int.advance();
            dump("g2c_" + name + "_step2", time_int.solution(), true);
        };
        dump("g2c_cons0", cons0, true);
        run_ident("rk4", spade::time_integration::rk4_t());
        run_ident("rk38r", spade::time_integration::rk38r_t());
        run_ident("rk2hs", spade::time_integration::rk2hs_t());
        run_ident("ssprk3hs", spade::time_integration::ssprk3hs_t());
    }
    g_manifest.close();
    return 0;
}
