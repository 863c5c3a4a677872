// Explicit Runge-Kutta with any Butcher table and caller-supplied residual and
// boundary callbacks (time-integration/integrator.h:18-59, advance.h:107-280,
// explicit.h:27-116), the stand-alone transform-aware stage update and the
// grid_array element-wise operations the callbacks build on.
//
// Coefficients are taken as the reference computes them: a table entry
// std::ratio<n, d> is numeric_t(n)/numeric_t(d) (advance.h:33-46), and a row
// difference (a_ij - a_{i-1,j}) is formed exactly as a ratio before that
// division (ratio_diff_t, advance.h:50-54).
#include "api_internal.hpp"
#include <cmath>
#include <cstdlib>
#include <memory>
#include <string>

using namespace spd;

namespace spd {
namespace {

long gcd_l(long a, long b)
{
    a = std::labs(a); b = std::labs(b);
    while (b) { const long t = a % b; a = b; b = t; }
    return a;
}

// value of the ratio n/d (d = 0 reads as 0): the reduced ratio's num/den in double
double ratio_value(long n, long d)
{
    if (d == 0 || n == 0) return 0.0;
    const long g = gcd_l(n, d);
    n /= g; d /= g;
    if (d < 0) { n = -n; d = -d; }
    return double(n)/double(d);
}

// a_ij - a_{i-1,j} as a ratio, then its value (row `stages` is the b row)
double row_diff(const spd_rk_table& tb, int curr, int prev, int j)
{
    auto entry = [&](int r, long& n, long& d)
    {
        const spd_ratio& x = (r < tb.stages) ? tb.a[r][j] : tb.b[j];
        n = x.num; d = x.den;
        if (d == 0) { n = 0; d = 1; }
    };
    long n0, d0, n1, d1;
    entry(curr, n0, d0); entry(prev, n1, d1);
    return ratio_value(d1*n0 - d0*n1, d0*d1);
}

void check_table(const spd_rk_table& tb)
{
    require(tb.stages >= 1 && tb.stages <= SPD_RK_MAX_STAGES,
            "table needs 1.." + std::to_string(SPD_RK_MAX_STAGES) + " stages");
    for (int i = 0; i < tb.stages; ++i)
    {
        for (int j = 0; j < SPD_RK_MAX_STAGES; ++j)
        {
            const spd_ratio& x = tb.a[i][j];
            require(x.den >= 0, "table ratios need a non-negative denominator");
            require(j < i || x.num == 0 || x.den == 0, "explicit tables have a[i][j] = 0 for j >= i");
        }
        require(tb.b[i].den >= 0 && tb.c[i].den >= 0, "table ratios need a non-negative denominator");
    }
}

void set_ratio(spd_ratio& r, long n, long d) { r.num = n; r.den = d; }

// The boundary of a stage: the caller's callback, else the exchange and the
// isothermal walls of the descriptor
void stage_bc(spd_integrator_s* it, spd_array_s* sol)
{
    if (it->bcfn)
    {
        if (it->bcfn(sol, it->t, it->user) != 0) throw ArgError{"boundary callback returned non-zero"};
        return;
    }
    Grid& g = *sol->a.grid;
    launch_exchange(g, sol->a.d, g.ctx->stream);
    if (it->has_walls)
        for (int face = 0; face < 6; ++face)
            if (it->bc.wall_mask & (1 << face)) launch_wall_fill(g, sol->a.d, face, it->bc.t_wall, g.ctx->stream);
}

void stage_rhs(spd_integrator_s* it, spd_array_s* r, spd_array_s* sol)
{
    if (it->rhs)
    {
        if (it->rhs(r, sol, it->t, it->user) != 0) throw ArgError{"residual callback returned non-zero"};
        return;
    }
    run_flux_div(*sol->a.grid, sol->a.d, r->a.d, it->flux, false);
}

// Evaluate residual i at t + c_i dt (the axis time moves there and back, as
// axis.time() += time_coeff*dt; ...; axis.time() -= time_coeff*dt)
void eval_stage(spd_integrator_s* it, int i, spd_array_s* sol, bool with_bc)
{
    const double tc = ratio_value(it->table.c[i].num, it->table.c[i].den);
    it->t += tc*it->dt;
    if (with_bc) stage_bc(it, sol);
    stage_rhs(it, it->res[i], sol);
    it->t -= tc*it->dt;
}

} // namespace

void advance_generic(spd_integrator_s* it)
{
    const spd_rk_table& tb = it->table;
    const int S = tb.stages;
    const double dt = it->dt;
    spd_array_s* q = it->q;
    Grid& g = *q->a.grid;
    hipStream_t st = g.ctx->stream;
    if (it->trans == SPD_TRANS_PRIM_CONS)
    {
        // transform-aware advance (advance.h:233-280)
        auto advance_to = [&](int prev, int curr)
        {
            const double* rr[SPD_RK_MAX_STAGES];
            double coef[SPD_RK_MAX_STAGES];
            for (int j = 0; j < S; ++j)
            {
                rr[j] = it->res[j]->a.d;
                coef[j] = row_diff(tb, curr, prev, j)*dt;   // zero entries are skipped by the kernel
            }
            // the ghosts are left to the boundary step: with the caller's own
            // bc the exchange may not be what follows
            launch_rk_transform_stage(g, q->a.d, rr, coef, S, it->gamma, it->R, st, false);
        };
        eval_stage(it, 0, q, false);
        for (int i = 1; i < S; ++i)
        {
            advance_to(i - 1, i);
            eval_stage(it, i, q, true);
        }
        advance_to(S - 1, S);
        it->t += dt;
        stage_bc(it, q);
        return;
    }
    // generic advance with the identity transform (advance.h:107-230)
    auto a_val = [&](int i, int j) { return ratio_value(tb.a[i][j].num, tb.a[i][j].den); };
    for (int i = 0; i < S; ++i)
    {
        spd_array_s* sol = tb.high_storage ? it->q2 : q;
        if (tb.high_storage)
            SPD_HIP(hipMemcpyAsync(it->q2->a.d, q->a.d, sizeof(double)*(size_t)g.geom.sb*g.geom.nblk,
                                   hipMemcpyDeviceToDevice, st));
        for (int j = 0; j < i; ++j)
        {
            const double c = a_val(i, j);
            if (c == 0.0) continue;
            launch_array_scale(g, it->res[j]->a.d, dt*c, st);
            launch_array_incr(g, sol->a.d, it->res[j]->a.d, false, st);
            launch_array_scale(g, it->res[j]->a.d, 1.0/(dt*c), st);
        }
        eval_stage(it, i, sol, i > 0);
        if (!tb.high_storage)
        {
            for (int j = 0; j < i; ++j)
            {
                const double c = a_val(i, j);
                if (c == 0.0) continue;
                launch_array_scale(g, it->res[j]->a.d, dt*c, st);
                launch_array_incr(g, sol->a.d, it->res[j]->a.d, true, st);
                launch_array_scale(g, it->res[j]->a.d, 1.0/(dt*c), st);
            }
        }
    }
    for (int i = 0; i < S; ++i)
    {
        const double c = ratio_value(tb.b[i].num, tb.b[i].den);
        if (c == 0.0) continue;
        launch_array_scale(g, it->res[i]->a.d, c*dt, st);
        launch_array_incr(g, q->a.d, it->res[i]->a.d, false, st);
        launch_array_scale(g, it->res[i]->a.d, 1.0/(c*dt), st);
    }
    it->t += dt;
    stage_bc(it, q);
}

} // namespace spd

extern "C" {

int spd_rk_table_builtin(int which, spd_rk_table* out)
{
    return guarded([&]
    {
        require(out != nullptr, "null argument");
        spd_rk_table t{};
        auto row = [&](int i, std::initializer_list<std::pair<long, long>> a)
        {
            int j = 0;
            for (auto& x : a) set_ratio(t.a[i][j++], x.first, x.second);
        };
        auto vec = [&](spd_ratio* v, std::initializer_list<std::pair<long, long>> a)
        {
            int j = 0;
            for (auto& x : a) set_ratio(v[j++], x.first, x.second);
        };
        switch (which)
        {
            case SPD_RK2: case SPD_RK2HS:               // explicit.h:41-61
                t.stages = 2; t.high_storage = which == SPD_RK2HS;
                row(1, {{1, 2}});
                vec(t.b, {{0, 1}, {1, 1}}); vec(t.c, {{0, 1}, {1, 2}});
                break;
            case SPD_RK4:                               // explicit.h:63-72
                t.stages = 4;
                row(1, {{1, 2}}); row(2, {{0, 1}, {1, 2}}); row(3, {{0, 1}, {0, 1}, {1, 1}});
                vec(t.b, {{1, 6}, {1, 3}, {1, 3}, {1, 6}}); vec(t.c, {{0, 1}, {1, 2}, {1, 2}, {1, 1}});
                break;
            case SPD_SSPRK34:                           // explicit.h:74-83
                t.stages = 4;
                row(1, {{1, 2}}); row(2, {{1, 2}, {1, 2}}); row(3, {{1, 6}, {1, 6}, {1, 6}});
                vec(t.b, {{1, 6}, {1, 6}, {1, 6}, {1, 2}}); vec(t.c, {{0, 1}, {1, 2}, {1, 1}, {1, 2}});
                break;
            case SPD_SSPRK3_TABLE: case SPD_SSPRK3HS:   // explicit.h:85-104
                t.stages = 3; t.high_storage = which == SPD_SSPRK3HS;
                row(1, {{1, 1}}); row(2, {{1, 4}, {1, 4}});
                vec(t.b, {{1, 6}, {1, 6}, {2, 3}}); vec(t.c, {{0, 1}, {1, 1}, {1, 2}});
                break;
            case SPD_RK38R:                             // explicit.h:106-115
                t.stages = 4;
                row(1, {{1, 3}}); row(2, {{-1, 3}, {1, 1}}); row(3, {{1, 1}, {-1, 1}, {1, 1}});
                vec(t.b, {{1, 8}, {3, 8}, {3, 8}, {1, 8}}); vec(t.c, {{0, 1}, {1, 3}, {2, 3}, {1, 1}});
                break;
            default:
                throw ArgError{"unknown built-in table " + std::to_string(which)};
        }
        *out = t;
    });
}

int spd_rk_integrator_create(spd_array_t q, const spd_rk_desc* d, spd_integrator_t* out)
{
    return guarded([&]
    {
        require(q && d && out, "null argument");
        check_table(d->table);
        require(d->trans == SPD_TRANS_PRIM_CONS || d->trans == SPD_TRANS_IDENTITY, "unknown state transform");
        require(d->dt > 0.0 && std::isfinite(d->dt) && std::isfinite(d->t0), "dt must be positive and finite");
        if (d->trans == SPD_TRANS_PRIM_CONS) require(d->gamma > 1.0 && d->R > 0.0, "gas needs gamma > 1 and R > 0");
        if (!d->rhs)
        {
            require(d->trans == SPD_TRANS_PRIM_CONS,
                    "the default residual (flux_div of the stored primitives) needs SPD_TRANS_PRIM_CONS");
            validate_flux(d->flux, q->a.grid->geom.ng);
        }
        if (!d->bc && d->walls)
            for (int face = 0; face < 6; ++face)
                if (d->walls->wall_mask & (1 << face))
                    require(!q->a.grid->desc.periodic[face/2], "walls are only allowed on non-periodic axes");
        auto* it = new spd_integrator_s;
        std::unique_ptr<spd_integrator_s, int (*)(spd_integrator_t)> guard(it, spd_integrator_destroy);
        it->generic = true;
        it->q = q;
        it->table = d->table;
        it->trans = d->trans;
        it->gamma = d->gamma; it->R = d->R;
        it->t = d->t0; it->dt = d->dt;
        it->rhs = d->rhs; it->bcfn = d->bc; it->user = d->user;
        if (d->flux) { it->flux = *d->flux; it->has_flux = true; }
        if (d->walls) { it->bc = *d->walls; it->has_walls = true; }
        Grid& g = *q->a.grid;
        for (int i = 0; i < d->table.stages; ++i) it->res[i] = new_array(g);
        if (d->table.high_storage && d->trans == SPD_TRANS_IDENTITY) it->q2 = new_array(g);
        *out = guard.release();
    });
}

int spd_integrator_time(spd_integrator_t it, double* t)
{
    return guarded([&]
    {
        require(it && t, "null argument");
        *t = it->t;
    });
}

int spd_rk_stage(spd_array_t prims, const spd_array_t* resid, const double* coef, int n, double gamma, double R,
                 int push_ghosts)
{
    return guarded([&]
    {
        require(prims && resid && coef, "null argument");
        require(n >= 1 && n <= SPD_RK_MAX_STAGES, "a stage combines 1.." + std::to_string(SPD_RK_MAX_STAGES) + " residuals");
        require(gamma > 1.0 && R > 0.0, "gas needs gamma > 1 and R > 0");
        const double* rr[SPD_RK_MAX_STAGES];
        for (int j = 0; j < n; ++j)
        {
            require(resid[j] != nullptr, "null residual array");
            require(resid[j]->a.grid == prims->a.grid, "residual arrays must be on the array's grid");
            require(resid[j] != prims, "a residual array cannot be the updated array");
            require(std::isfinite(coef[j]), "stage coefficients must be finite");
            rr[j] = resid[j]->a.d;
        }
        Grid& g = *prims->a.grid;
        launch_rk_transform_stage(g, prims->a.d, rr, coef, n, gamma, R, g.ctx->stream, push_ghosts != 0);
    });
}

int spd_array_scale(spd_array_t arr, double s)
{
    return guarded([&]
    {
        require(arr != nullptr, "null argument");
        Grid& g = *arr->a.grid;
        launch_array_scale(g, arr->a.d, s, g.ctx->stream);
    });
}

static int array_incr(spd_array_t dst, spd_array_t src, bool sub)
{
    return guarded([&]
    {
        require(dst && src, "null argument");
        require(dst->a.grid == src->a.grid, "arrays must be on the same grid");
        Grid& g = *dst->a.grid;
        launch_array_incr(g, dst->a.d, src->a.d, sub, g.ctx->stream);
    });
}

int spd_array_add(spd_array_t dst, spd_array_t src) { return array_incr(dst, src, false); }
int spd_array_sub(spd_array_t dst, spd_array_t src) { return array_incr(dst, src, true); }

int spd_array_convert(spd_array_t arr, int direction, double gamma, double R, int include_ghosts)
{
    return guarded([&]
    {
        require(arr != nullptr, "null argument");
        require(direction == SPD_TO_CONS || direction == SPD_TO_PRIM, "direction must be SPD_TO_CONS or SPD_TO_PRIM");
        require(gamma > 1.0 && R > 0.0, "gas needs gamma > 1 and R > 0");
        Grid& g = *arr->a.grid;
        launch_state_convert(g, arr->a.d, direction == SPD_TO_PRIM, gamma, R, include_ghosts != 0, g.ctx->stream);
    });
}

} // extern "C"
