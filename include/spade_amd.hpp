// C++ facade over the libspade_amd C ABI (include/spade_amd.h), with the names
// of the SPADE types and calls it stands in for.  Header-only; link with
// -lspade_amd.  Failures throw spade_amd::error (SPADE throws
// except::sp_exception, core/except.h:12-30).
#pragma once
#include <cstring>
#include <functional>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#include "spade_amd.h"

namespace spade_amd {

struct error : std::runtime_error
{
    int code;
    error(int c, const std::string& m) : std::runtime_error("spade_amd: " + m), code(c) {}
};

inline void check(int rc)
{
    if (rc != SPD_OK) throw error(rc, spd_last_error());
}

// One device and one stream (every call is ordered on it).
class context
{
    spd_ctx_t h_ = nullptr;
public:
    explicit context(int device = 0) { check(spd_ctx_create(device, &h_)); }
    ~context() { if (h_) spd_ctx_destroy(h_); }
    context(const context&) = delete;
    context& operator=(const context&) = delete;
    void sync() const { check(spd_ctx_sync(h_)); }
    spd_ctx_t handle() const { return h_; }
};

// cartesian_blocks_t + cartesian_grid_t (grid/cartesian_blocks.h:38, grid/cartesian_grid.h:84)
struct grid_desc : spd_grid_desc
{
    grid_desc()
    {
        std::memset(static_cast<spd_grid_desc*>(this), 0, sizeof(spd_grid_desc));
        num_exch = 2;
        for (int a = 0; a < 3; ++a) { periodic[a] = 1; map[a].kind = SPD_MAP_IDENTITY; }
    }
};

class grid
{
    spd_grid_t h_ = nullptr;
    grid_desc d_;
public:
    grid(const context& ctx, const grid_desc& d) : d_(d) { check(spd_grid_create(ctx.handle(), &d_, &h_)); }
    ~grid() { if (h_) spd_grid_destroy(h_); }
    grid(const grid&) = delete;
    grid& operator=(const grid&) = delete;
    spd_grid_t handle() const { return h_; }
    const grid_desc& desc() const { return d_; }
    // doubles in the reference-layout host image of one array
    long host_size() const { return spd_array_host_size(h_); }
    long device_size() const { return spd_array_device_size(h_); }
    int num_blocks() const { return d_.num_blocks[0]*d_.num_blocks[1]*d_.num_blocks[2]; }
};

// grid_array with 5 FP64 variables (grid/grid_array.h:182-422)
class grid_array
{
    spd_array_t h_ = nullptr;
    const grid* g_;
    bool owner_ = true;
public:
    explicit grid_array(const grid& g) : g_(&g) { check(spd_array_create(g.handle(), &h_)); }
    // a non-owning view of an array the library hands to a callback
    grid_array(const grid& g, spd_array_t view) : h_(view), g_(&g), owner_(false) {}
    ~grid_array() { if (h_ && owner_) spd_array_destroy(h_); }
    grid_array(const grid_array&) = delete;
    grid_array& operator=(const grid_array&) = delete;
    grid_array& operator=(double v) { check(spd_array_fill(h_, v)); return *this; }
    // element-wise operations (grid/grid_array.h:286-396)
    grid_array& operator*=(double s) { check(spd_array_scale(h_, s)); return *this; }
    grid_array& operator+=(const grid_array& o) { check(spd_array_add(h_, o.h_)); return *this; }
    grid_array& operator-=(const grid_array& o) { check(spd_array_sub(h_, o.h_)); return *this; }
    void upload(const double* host) { check(spd_array_upload(h_, host)); }
    void download(double* host) const { check(spd_array_download(h_, host)); }
    std::vector<double> download() const { std::vector<double> v(g_->host_size()); download(v.data()); return v; }
    spd_array_t handle() const { return h_; }
    const grid& get_grid() const { return *g_; }
};

// ideal_gas_t (navier-stokes/gas.h:35-49)
struct ideal_gas_t { double gamma, R; };

// A flux functor or a composition of them (omni/compose.h:41-45)
class flux
{
    spd_flux_desc d_;
public:
    flux() { std::memset(&d_, 0, sizeof(d_)); }
    flux(const spd_flux_term& t, const ideal_gas_t& gas) : flux()
    {
        d_.nterms = 1; d_.terms[0] = t; d_.gamma = gas.gamma; d_.R = gas.R;
    }
    const spd_flux_desc& desc() const { return d_; }
    // appends the terms of another flux: this + (other ...)
    void append(const flux& p)
    {
        for (int m = 0; m < p.d_.nterms; ++m)
        {
            if (d_.nterms == SPD_MAX_TERMS) throw error(SPD_ERR_ARG, "at most 4 composed terms");
            d_.terms[d_.nterms++] = p.d_.terms[m];
        }
        d_.gamma = p.d_.gamma; d_.R = p.d_.R;
    }
};

// F0 + (F1 + ...) evaluated per face
inline flux compose(std::initializer_list<flux> parts)
{
    flux out;
    for (const flux& p : parts) out.append(p);
    return out;
}

inline spd_flux_term term(int kind) { spd_flux_term t; std::memset(&t, 0, sizeof(t)); t.kind = kind; return t; }

// convective.h:187-192
inline flux cent_keep(int order, const ideal_gas_t& gas) { auto t = term(SPD_TERM_KEEP); t.order = order; return flux(t, gas); }
// convective.h:54-94
inline flux totani_lr(const ideal_gas_t& gas) { return flux(term(SPD_TERM_TOTANI), gas); }

struct const_sensor_t { double alpha; };      // state_sensor.h:45-60
struct ducros_t { double epsilon; };          // state_sensor.h:21-43

// convective.h:194-246
inline flux pressure_diss_lr(const ideal_gas_t& gas, const const_sensor_t& s, double eps_p, double eps_T)
{
    auto t = term(SPD_TERM_PDISS); t.sensor = SPD_SENSOR_CONST; t.sensor_param = s.alpha; t.eps_p = eps_p; t.eps_T = eps_T;
    return flux(t, gas);
}
inline flux pressure_diss_lr(const ideal_gas_t& gas, const ducros_t& s, double eps_p, double eps_T)
{
    auto t = term(SPD_TERM_PDISS); t.sensor = SPD_SENSOR_DUCROS; t.sensor_param = s.epsilon; t.eps_p = eps_p; t.eps_T = eps_T;
    return flux(t, gas);
}
// convective.h:255-333 with rusanov_t (split = 0) or phys_flux_t (split = 1)
inline flux weno(const ideal_gas_t& gas, int split, bool smooth = true)
{
    auto t = term(SPD_TERM_WENO); t.split = split; t.smooth = smooth; return flux(t, gas);
}
// convective.h:335-497
inline flux fweno(const ideal_gas_t& gas, bool smooth = true) { auto t = term(SPD_TERM_FWENO); t.smooth = smooth; return flux(t, gas); }
// convective.h:14-33
inline flux first_order(const ideal_gas_t& gas) { return flux(term(SPD_TERM_FIRST), gas); }
// weno_fds_t<rusanov_fds_t> (convective.h:499-555, flux_funcs.h:54-95)
inline flux weno_fds(const ideal_gas_t& gas, bool smooth = true)
{
    auto t = term(SPD_TERM_WENO_FDS); t.smooth = smooth; return flux(t, gas);
}

// hybrid_scheme_t(scheme0, scheme1, blender, full_flux | diss_flux) (navier-stokes/hybrid_scheme.h:14-48):
// scheme0 = cent_keep<2|4>, scheme1 = weno(...) or fweno(...).  flux_div refuses the
// result (SPD_ERR_UNSUPPORTED): the device library has no code for this scheme.
enum hybrid_tag { diss_flux = 0, full_flux = 1 };
inline spd_flux_term hybrid_term(const flux& central, const flux& diss, hybrid_tag tag)
{
    const spd_flux_desc& c = central.desc();
    const spd_flux_desc& d = diss.desc();
    if (c.nterms != 1 || c.terms[0].kind != SPD_TERM_KEEP)
        throw error(SPD_ERR_ARG, "hybrid_scheme_t: scheme0 must be cent_keep");
    if (d.nterms != 1 || (d.terms[0].kind != SPD_TERM_WENO && d.terms[0].kind != SPD_TERM_FWENO))
        throw error(SPD_ERR_ARG, "hybrid_scheme_t: scheme1 must be weno or fweno");
    spd_flux_term t = d.terms[0];
    t.diss_kind = d.terms[0].kind;
    t.kind = SPD_TERM_HYBRID;
    t.order = c.terms[0].order;
    t.full = tag;
    return t;
}
inline flux hybrid_scheme_t(const flux& central, const flux& diss, const const_sensor_t& s, hybrid_tag tag)
{
    auto t = hybrid_term(central, diss, tag); t.sensor = SPD_SENSOR_CONST; t.sensor_param = s.alpha;
    return flux(t, ideal_gas_t{central.desc().gamma, central.desc().R});
}
inline flux hybrid_scheme_t(const flux& central, const flux& diss, const ducros_t& s, hybrid_tag tag)
{
    auto t = hybrid_term(central, diss, tag); t.sensor = SPD_SENSOR_DUCROS; t.sensor_param = s.epsilon;
    return flux(t, ideal_gas_t{central.desc().gamma, central.desc().R});
}

// viscous_laws.h:62-100 and 102-135
struct constant_viscosity_t { double mu, prandtl; };
struct power_law_t { double mu_ref, T_ref, power, prandtl; };

// viscous.h:13-113
inline flux visc_lr(const constant_viscosity_t& law, const ideal_gas_t& gas)
{
    auto t = term(SPD_TERM_VISC); t.law = SPD_LAW_CONST; t.mu = law.mu; t.prandtl = law.prandtl; return flux(t, gas);
}
inline flux visc_lr(const power_law_t& law, const ideal_gas_t& gas)
{
    auto t = term(SPD_TERM_VISC); t.law = SPD_LAW_POWER; t.mu = law.mu_ref; t.T_ref = law.T_ref;
    t.power = law.power; t.prandtl = law.prandtl; return flux(t, gas);
}

// subgrid_scale.h:24-86 and viscous_laws.h:174-216: a laminar law plus the
// WALE subgrid viscosity (the gas of the flux gives R)
struct wale_t { double cw, delta, prt; };
template <typename laminar_t> struct sgs_visc_t { laminar_t lam; wale_t turb; };
template <typename laminar_t> inline flux visc_lr(const sgs_visc_t<laminar_t>& law, const ideal_gas_t& gas)
{
    flux f = visc_lr(law.lam, gas);
    spd_flux_term t = f.desc().terms[0];
    t.sgs = SPD_SGS_WALE; t.sgs_cw = law.turb.cw; t.sgs_delta = law.turb.delta; t.sgs_prt = law.turb.prt;
    return flux(t, gas);
}

// pde_algs::flux_div (pde-algs/flux-div/flux_div.h:16-56); default increment as in the reference
inline void flux_div(const grid_array& prims, grid_array& rhs, const flux& f, bool increment = true)
{
    check(spd_flux_div(prims.handle(), rhs.handle(), &f.desc(), increment ? SPD_INCREMENT : SPD_OVERWRITE));
}

// create_exchange(...).exchange(arr) (grid/exchange_pattern.h:304-424)
struct exchange_handle { void exchange(grid_array& a) const { check(spd_exchange(a.handle())); } };
inline exchange_handle create_exchange(const grid_array&) { return exchange_handle{}; }

// boundary_fill with the isothermal no-slip wall mirror (grid/boundary_fill.h:32-133)
inline void boundary_fill_wall(grid_array& a, int face, double t_wall) { check(spd_boundary_fill_wall(a.handle(), face, t_wall)); }

// algs::boundary_fill(arr, boundaries, kern) (grid/boundary_fill.h:32-133):
// boundaries = boundary::xmin | boundary::ymax | ...; kern = boundary::
// extrapolate<order> or a fill_kernel (a user kernel given as data)
namespace boundary
{
constexpr int xmin = SPD_XMIN, xmax = SPD_XMAX, ymin = SPD_YMIN, ymax = SPD_YMAX, zmin = SPD_ZMIN, zmax = SPD_ZMAX;
template <int order> struct extrap_t { static constexpr int value = order; };
template <int order> constexpr extrap_t<order> extrapolate{};
}
using fill_kernel = spd_fill_kernel;
template <int order> inline void boundary_fill(grid_array& a, int faces, boundary::extrap_t<order>)
{
    check(spd_boundary_fill_extrap(a.handle(), faces, order));
}
inline void boundary_fill(grid_array& a, int faces, const fill_kernel& k) { check(spd_boundary_fill(a.handle(), faces, &k)); }

enum scheme { ssprk3 = SPD_SCHEME_SSPRK3, ssprk3_opt = SPD_SCHEME_SSPRK3_OPT };

// integrator_t(axis, scheme, data, rhs = flux_div, bc = exchange + walls,
// trans = prim<->cons).advance() (time-integration/integrator.h:18-59)
class integrator_t
{
    spd_integrator_t h_ = nullptr;
public:
    integrator_t(grid_array& q, const flux& f, double dt, scheme s = ssprk3, int wall_mask = 0, double t_wall = 0.0)
    {
        spd_bc_desc bc{wall_mask, t_wall};
        check(spd_integrator_create(q.handle(), &f.desc(), &bc, s, dt, &h_));
    }
    ~integrator_t() { if (h_) spd_integrator_destroy(h_); }
    integrator_t(const integrator_t&) = delete;
    integrator_t& operator=(const integrator_t&) = delete;
    void advance(int nsteps = 1) { check(spd_integrator_advance(h_, nsteps)); }
};

// integrator_t with any explicit table and the caller's residual and boundary
// (time-integration/integrator.h:18-59, advance.h:107-280): trans =
// state_transform (primitive storage) or identity_transform.  An empty rhs is
// flux_div(sol, resid, f); an empty bc the exchange plus the walls.
enum transform { state_transform = SPD_TRANS_PRIM_CONS, identity_transform = SPD_TRANS_IDENTITY };
inline spd_rk_table rk_table(int which) { spd_rk_table t; check(spd_rk_table_builtin(which, &t)); return t; }
class rk_integrator_t
{
public:
    using rhs_fn = std::function<void(grid_array& resid, const grid_array& sol, double t)>;
    using bc_fn = std::function<void(grid_array& sol, double t)>;
private:
    spd_integrator_t h_ = nullptr;
    const grid* g_;
    rhs_fn rhs_;
    bc_fn bc_;
    static int call_rhs(spd_array_t r, spd_array_t q, double t, void* u)
    {
        auto* self = static_cast<rk_integrator_t*>(u);
        try { grid_array rr(*self->g_, r), qq(*self->g_, q); self->rhs_(rr, qq, t); return 0; }
        catch (...) { return 1; }
    }
    static int call_bc(spd_array_t q, double t, void* u)
    {
        auto* self = static_cast<rk_integrator_t*>(u);
        try { grid_array qq(*self->g_, q); self->bc_(qq, t); return 0; }
        catch (...) { return 1; }
    }
public:
    rk_integrator_t(grid_array& q, const spd_rk_table& table, double dt, transform trans, const ideal_gas_t& gas,
                    rhs_fn rhs = {}, bc_fn bc = {}, const flux* f = nullptr, double t0 = 0.0)
        : g_(&q.get_grid()), rhs_(std::move(rhs)), bc_(std::move(bc))
    {
        spd_rk_desc d;
        std::memset(&d, 0, sizeof(d));
        d.table = table; d.trans = trans; d.gamma = gas.gamma; d.R = gas.R; d.t0 = t0; d.dt = dt;
        d.rhs = rhs_ ? &call_rhs : nullptr;
        d.bc = bc_ ? &call_bc : nullptr;
        d.user = this;
        d.flux = f ? &f->desc() : nullptr;
        check(spd_rk_integrator_create(q.handle(), &d, &h_));
    }
    ~rk_integrator_t() { if (h_) spd_integrator_destroy(h_); }
    rk_integrator_t(const rk_integrator_t&) = delete;
    rk_integrator_t& operator=(const rk_integrator_t&) = delete;
    void advance(int nsteps = 1) { check(spd_integrator_advance(h_, nsteps)); }
    double time() const { double t = 0.0; check(spd_integrator_time(h_, &t)); return t; }
};

inline double max_wavespeed(const grid_array& q, const ideal_gas_t& gas)
{
    double m = 0.0;
    check(spd_max_wavespeed(q.handle(), gas.gamma, gas.R, &m));
    return m;
}

// spade::io (io/io_vtk.h:276-288, io/io_native.h:39-55)
namespace io
{
    // names: the 5 variable names; empty for P,T,U,V,W (prim_t)
    inline void output_vtk(const std::string& out_dir, const std::string& basename, const grid_array& arr,
                           const std::vector<std::string>& names = {})
    {
        const char* n[5];
        if (!names.empty())
        {
            if (names.size() != 5) throw error(SPD_ERR_ARG, "output_vtk needs 5 variable names");
            for (int v = 0; v < 5; ++v) n[v] = names[v].c_str();
        }
        check(spd_io_output_vtk(arr.handle(), out_dir.c_str(), basename.c_str(), names.empty() ? nullptr : n));
    }
    inline void binary_write(const std::string& filename, const grid_array& arr)
    {
        check(spd_io_binary_write(arr.handle(), filename.c_str()));
    }
    inline void binary_read(const std::string& filename, grid_array& arr)
    {
        check(spd_io_binary_read(arr.handle(), filename.c_str()));
    }
}

} // namespace spade_amd
