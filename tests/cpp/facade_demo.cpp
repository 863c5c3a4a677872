// A SPADE-style driver written against the C++ facade (include/spade_amd.hpp):
// TGV on 2x2x2 blocks of 16^3 with 2 ghost layers, cent_keep<2> + visc_lr with
// the power law; writes the initial primitives, the residual of one flux_div,
// and the state after 3 SSP-RK3 steps (reference layout, raw float64) into the
// folder given as argv[1].  tests/test_gpu_facade.py checks the files against
// the CPU oracle.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>
#include "spade_amd.hpp"

static void write(const std::string& path, const std::vector<double>& v)
{
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

int main(int argc, char** argv)
{
    const std::string out = argc > 1 ? argv[1] : ".";
    try
    {
        namespace sa = spade_amd;
        sa::context ctx(0);
        sa::grid_desc d;
        const double L = 2.0*M_PI;
        for (int a = 0; a < 3; ++a) { d.num_blocks[a] = 2; d.cells_in_block[a] = 16; d.lo[a] = 0.0; d.hi[a] = L; }
        d.num_exch = 2;
        sa::grid g(ctx, d);
        const double gamma = 1.4, mach = 0.1, p0 = 1.0/(mach*mach)/gamma;
        const sa::ideal_gas_t air{gamma, p0};
        // initial condition, interior cells (ghosts by the exchange)
        std::vector<double> host(g.host_size(), 0.0);
        const int N = 20, ng = 2;
        for (int lb = 0; lb < 8; ++lb)
        {
            const int b[3] = {lb % 2, (lb/2) % 2, lb/4};
            for (int k = 0; k < 16; ++k) for (int j = 0; j < 16; ++j) for (int i = 0; i < 16; ++i)
            {
                const double x = (b[0]*16 + i + 0.5)*L/32, y = (b[1]*16 + j + 0.5)*L/32, z = (b[2]*16 + k + 0.5)*L/32;
                double* e = &host[5*((size_t)(i + ng) + N*((size_t)(j + ng) + N*((size_t)(k + ng) + (size_t)N*lb)))];
                e[0] = p0 + (1.0/16.0)*(std::cos(2*x) + std::cos(2*y))*(std::cos(2*z) + 2.0);
                e[1] = 1.0 + 0.01*std::sin(x + 2*y);
                e[2] =  std::sin(x)*std::cos(y)*std::cos(z);
                e[3] = -std::cos(x)*std::sin(y)*std::cos(z);
                e[4] = 0.05*std::sin(z);
            }
        }
        sa::grid_array q(g), rhs(g);
        q.upload(host.data());
        sa::create_exchange(q).exchange(q);
        write(out + "/facade_prim0.bin", q.download());
        const auto flux = sa::compose({sa::cent_keep(2, air), sa::visc_lr(sa::power_law_t{1.0/1600.0, 1.0, 0.76, 0.72}, air)});
        rhs = 0.0;
        sa::flux_div(q, rhs, flux);
        write(out + "/facade_rhs.bin", rhs.download());
        const double dt = 0.5*(L/32)/sa::max_wavespeed(q, air);
        // rk4_t with the transform-aware advance and a caller residual that
        // counts its calls around flux_div (integrator_t with a rhs callable)
        {
            sa::grid_array q4(g);
            q4.upload(host.data());
            sa::create_exchange(q4).exchange(q4);
            int calls = 0;
            sa::rk_integrator_t rk(q4, sa::rk_table(SPD_RK4), dt, sa::state_transform, air,
                                   [&](sa::grid_array& r, const sa::grid_array& s, double) { ++calls; sa::flux_div(s, r, flux, false); });
            rk.advance(2);
            if (calls != 8) throw std::runtime_error("rk4 residual calls " + std::to_string(calls));
            write(out + "/facade_rk4_step2.bin", q4.download());
        }
        sa::integrator_t it(q, flux, dt, sa::ssprk3);
        it.advance(3);
        write(out + "/facade_step3.bin", q.download());
        // io::binary_write / binary_read round trip and io::output_vtk
        sa::io::binary_write(out + "/facade_step3_checkpoint.bin", q);
        sa::grid_array back(g);
        back = 0.0;
        sa::io::binary_read(out + "/facade_step3_checkpoint.bin", back);
        write(out + "/facade_step3_readback.bin", back.download());
        sa::io::output_vtk(out + "/vtk", "step3", q);
        std::printf("dt %.17g\n", dt);
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
