// CPU oracle: a plain C++ restatement of SPADE's right-hand-side hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (spade_amd/) links, loads
// or calls this file; only tests/, the smoke check and the CPU-baseline leg of
// bench.py do, and only as the checker.
//
// What it restates (all paths relative to the SPADE source tree):
//   * storage: AoS, variable index fastest, ghosts stored
//       offset = v + 5*((i+ng) + Nx*((j+ng) + Ny*((k+ng) + Nz*lb)))
//       (core/mem_map.h:466-497, 654-676; grid/grid_array.h:244-260)
//   * block geometry: box = lo + b*bsize, dx = box/n, inv_dx = 1/dx, cell
//       centre xi = box.min + (i+1/2)dx (grid/cartesian_blocks.h:38-107,
//       grid/cartesian_grid.h:116-145, grid/grid_geometry.h:57-99,
//       grid/grid_index_types.h:207-214); block order x fastest
//       (core/ctrs.h:341-353)
//   * flux_div_basic: per interior cell, dirs 0..2, pm 0..1:
//       accum = F(face); accum *= jac*(1-2pm)*inv_dx; rhs += accum
//       (pde-algs/flux-div/flux_div_basic.h:17-76), overwrite/increment
//       (pde-algs/pde_traits.h:7-20), compose = F0 + (F1 + ...)
//       (omni/compose.h:7-45)
//   * face data: value = (qL+qR)*0.5 (omni/infos/info_value.h:151-163),
//       gradient four-point rule (omni/infos/info_gradient.h:19-86),
//       stencil offsets -3,-1,1,3 -> cells i-2,i-1,i,i+1 (omni/geom.h:87-145)
//   * fluxes: totani_lr, cent_keep<order>, pressure_diss_lr, weno_t, fweno_t,
//       weno_fds_t, first_order_t (navier-stokes/convective.h:14-556);
//       rusanov_t, rusanov_fds_t,
//       phys_flux_t (navier-stokes/flux_funcs.h:8-126); visc_lr
//       (navier-stokes/viscous.h:39-112); constant_viscosity_t and a power law
//       with the three getters of power_law_t (navier-stokes/viscous_laws.h:
//       62-135); const_sensor_t, ducros_t (navier-stokes/state_sensor.h:21-60)
//   * prim<->cons (navier-stokes/fluid_state.h:102-135)
//   * ghost exchange: 26 neighbours with periodic wrap, non-periodic domain
//       faces skipped (grid/cartesian_blocks.h:73-105,
//       grid/exchange_pattern.h:382-424, grid/get_transaction.h:55-85)
//   * boundary_fill with a mirror kernel (grid/boundary_fill.h:104-130)
//   * SSP-RK3 transform-aware advance (time-integration/advance.h:56-102,
//       233-280; explicit.h:80-88) and ssprk3_opt (advance.h:285-402)
//
// Curvilinear (diagonal) coordinates do not compile in the reference, so
// for them this file states its own formulation (computational coordinates
// throughout, gradient scaled by 1/x'(xi) per axis); those paths are pinned by
// closed-form checks in the tests, not by reference vectors.
#include <omp.h>
#include <cmath>
#include <cstring>
#include <cstdint>
#include <vector>
#include <algorithm>
#include "spade_oracle.h"

namespace {

struct Geo
{
    ora_grid g;
    int N[3];                 // stored extent per axis (n + 2 ng)
    int nblk;
    std::vector<double> bmin; // [lb*3+d]
    std::vector<double> dx;   // [lb*3+d]
    std::vector<double> idx;  // [lb*3+d] = 1/dx
    std::vector<int> bpos;    // [lb*3+d]

    explicit Geo(const ora_grid* gin) : g(*gin)
    {
        for (int d = 0; d < 3; ++d) N[d] = g.nc[d] + 2*g.ng;
        nblk = g.nb[0]*g.nb[1]*g.nb[2];
        bmin.resize(3*nblk); dx.resize(3*nblk); idx.resize(3*nblk); bpos.resize(3*nblk);
        double bsize[3];
        for (int d = 0; d < 3; ++d) bsize[d] = (g.hi[d] - g.lo[d])/g.nb[d];
        for (int lb = 0; lb < nblk; ++lb)
        {
            int b[3] = {lb % g.nb[0], (lb / g.nb[0]) % g.nb[1], lb / (g.nb[0]*g.nb[1])};
            for (int d = 0; d < 3; ++d)
            {
                const double mn = g.lo[d] + b[d]*bsize[d];
                const double mx = mn + bsize[d];
                bpos[3*lb+d] = b[d];
                bmin[3*lb+d] = mn;
                dx[3*lb+d]   = (mx - mn)/g.nc[d];
                idx[3*lb+d]  = 1.0/dx[3*lb+d];
            }
        }
    }
    size_t cell(int i, int j, int k, int lb) const
    {
        return 5*((size_t)(i + g.ng) + (size_t)N[0]*((size_t)(j + g.ng) + (size_t)N[1]*((size_t)(k + g.ng) + (size_t)N[2]*lb)));
    }
    // computational coordinate of cell (or face, when half = true in dir d)
    double comp(int lb, int d, double iref) const { return bmin[3*lb+d] + iref*dx[3*lb+d]; }
    // x'(xi) of axis d
    double deriv(int d, double xi) const
    {
        const ora_map1d& m = g.map[d];
        switch (m.kind)
        {
            case ORA_MAP_TANH:
            {
                // y = c0 + h * tanh(beta*s)/tanh(beta),  s = (xi - c0)/h
                const double h = 0.5*(m.p1 - m.p0), c = 0.5*(m.p1 + m.p0);
                const double s = (xi - c)/h;
                const double t = std::tanh(m.beta*s);
                return m.beta*(1.0 - t*t)/std::tanh(m.beta);
            }
            default: return 1.0;
        }
    }
    bool identity() const { return g.map[0].kind == ORA_MAP_IDENTITY && g.map[1].kind == ORA_MAP_IDENTITY && g.map[2].kind == ORA_MAP_IDENTITY; }
};

struct Q { double p, T, u, v, w; double operator[](int n) const { return n==0?p:n==1?T:n==2?u:n==3?v:w; } double vel(int d) const { return d==0?u:d==1?v:w; } };

inline Q load(const Geo& G, const double* a, int i, int j, int k, int lb)
{
    const double* e = a + G.cell(i, j, k, lb);
    return Q{e[0], e[1], e[2], e[3], e[4]};
}

// Face index: dir d, (i,j,k) is the cell on the right (face i is the low face
// of cell i).  Cell at offset o (half-cells, odd) is i + floor(o/2).
struct Face
{
    int d, i, j, k, lb;
    void at(int o, int& ci, int& cj, int& ck) const
    {
        const int s = (o >= 0) ? o/2 : -((-o + 1)/2);
        ci = i; cj = j; ck = k;
        if (d == 0) ci += s; else if (d == 1) cj += s; else ck += s;
    }
};

// metric (area) vector of direction d at computational point xi:
// identity -> unit vector; diagonal -> e_d / (m_d * J), J = 1/(m0 m1 m2)
// (core/coord_system.h:206-217, 250-267)
inline void metric(const Geo& G, int d, const double xi[3], double n[3])
{
    if (G.identity()) { n[0] = (d==0); n[1] = (d==1); n[2] = (d==2); return; }
    const double m0 = G.deriv(0, xi[0]), m1 = G.deriv(1, xi[1]), m2 = G.deriv(2, xi[2]);
    const double jac = 1.0/(m0*m1*m2);
    n[0] = 0.0; n[1] = 0.0; n[2] = 0.0;
    n[d] = 1.0;
    n[0] /= m0*jac; n[1] /= m1*jac; n[2] /= m2*jac;
}

inline void cell_xi(const Geo& G, int lb, int i, int j, int k, double xi[3])
{
    xi[0] = G.comp(lb, 0, i + 0.5); xi[1] = G.comp(lb, 1, j + 0.5); xi[2] = G.comp(lb, 2, k + 0.5);
}
inline void face_xi(const Geo& G, const Face& f, double xi[3])
{
    double r[3] = {f.i + 0.5, f.j + 0.5, f.k + 0.5};
    r[f.d] -= 0.5;
    for (int d = 0; d < 3; ++d) xi[d] = G.comp(f.lb, d, r[d]);
}

inline void cell_metric(const Geo& G, const Face& f, int o, double n[3])
{
    int ci, cj, ck; f.at(o, ci, cj, ck);
    double xi[3]; cell_xi(G, f.lb, ci, cj, ck, xi);
    metric(G, f.d, xi, n);
}
inline void face_metric(const Geo& G, const Face& f, double n[3])
{
    double xi[3]; face_xi(G, f, xi);
    metric(G, f.d, xi, n);
}

// face gradient of all 5 primitives: out[dir][var] (info_gradient.h:45-81);
// diagonal coordinates additionally divide component c by x_c'(xi_face).
inline void face_grad(const Geo& G, const double* a, const Face& f, double out[3][5])
{
    const int d0 = f.d, d1 = (f.d + 1) % 3, d2 = (f.d + 2) % 3;
    const int dirs[3] = {d0, d1, d2};
    const double* inv = &G.idx[3*f.lb];
    for (int c = 0; c < 3; ++c) for (int v = 0; v < 5; ++v) out[c][v] = 0.0;
    int ic[3]; f.at(-1, ic[0], ic[1], ic[2]);
    auto apply = [&](int iset, double coeff)
    {
        const double* e = a + G.cell(ic[0], ic[1], ic[2], f.lb);
        for (int v = 0; v < 5; ++v) out[iset][v] += coeff*e[v]*inv[iset];
    };
    apply(dirs[0], -1.0);
    for (int ii = 1; ii < 3; ++ii)
    {
        ic[dirs[ii]] += 1; apply(dirs[ii],  0.25);
        ic[dirs[ii]] -= 2; apply(dirs[ii], -0.25);
        ic[dirs[ii]] += 1;
    }
    ic[dirs[0]] += 1;
    apply(dirs[0], 1.0);
    for (int ii = 1; ii < 3; ++ii)
    {
        ic[dirs[ii]] += 1; apply(dirs[ii],  0.25);
        ic[dirs[ii]] -= 2; apply(dirs[ii], -0.25);
        ic[dirs[ii]] += 1;
    }
    if (!G.identity())
    {
        double xi[3]; face_xi(G, f, xi);
        for (int c = 0; c < 3; ++c)
        {
            const double m = G.deriv(c, xi[c]);
            for (int v = 0; v < 5; ++v) out[c][v] /= m;
        }
    }
}

inline Q face_value(const Geo& G, const double* a, const Face& f)
{
    int li, lj, lk, ri, rj, rk;
    f.at(-1, li, lj, lk); f.at(1, ri, rj, rk);
    const double* l = a + G.cell(li, lj, lk, f.lb);
    const double* r = a + G.cell(ri, rj, rk, f.lb);
    double o[5];
    for (int v = 0; v < 5; ++v) { o[v] = l[v]; o[v] += r[v]; o[v] *= 0.5; }
    return Q{o[0], o[1], o[2], o[3], o[4]};
}

inline Q cell_at(const Geo& G, const double* a, const Face& f, int o)
{
    int ci, cj, ck; f.at(o, ci, cj, ck);
    return load(G, a, ci, cj, ck, f.lb);
}

struct Ctx
{
    const Geo& G;
    const double* a;
    const ora_gas& gas;
};

// centered_finite_diff_t<idx, order>::value (core/finite_diff.h:25-34)
inline double cfd_coeff(int idx, int order)
{
    auto fact = [](int n) { double r = 1.0; for (int q = 2; q <= n; ++q) r *= q; return r; };
    const int h = order/2;
    const double d0 = fact(h)/fact(h + idx);
    const double d1 = fact(h)/fact(h - idx);
    const int sgn = ((1 + idx) % 2 == 0) ? 1 : -1;
    return d0*d1*sgn/double(idx);
}

// totani_lr (convective.h:68-93)
void f_totani(const Ctx& C, const Face& f, double out[5])
{
    const Q ql = cell_at(C.G, C.a, f, -1), qr = cell_at(C.G, C.a, f, 1);
    double nl[3], nr[3]; cell_metric(C.G, f, -1, nl); cell_metric(C.G, f, 1, nr);
    const double un_l = nl[0]*ql.u + nl[1]*ql.v + nl[2]*ql.w;
    const double un_r = nr[0]*qr.u + nr[1]*qr.v + nr[2]*qr.w;
    const double rho_l = ql.p/(C.gas.R*ql.T);
    const double rho_r = qr.p/(C.gas.R*qr.T);
    const double e_l = ql.p/(rho_l*(C.gas.gamma - 1.0));
    const double e_r = qr.p/(rho_r*(C.gas.gamma - 1.0));
    const double c = 0.25*(rho_l + rho_r)*(un_l + un_r);
    out[0] = c;
    out[1] = 0.5*c*(e_l + e_r + ql.u*qr.u + ql.v*qr.v + ql.w*qr.w) + 0.5*(un_l*qr.p + un_r*ql.p);
    out[2] = 0.5*c*(ql.u + qr.u) + 0.5*(nl[0]*ql.p + nr[0]*qr.p);
    out[3] = 0.5*c*(ql.v + qr.v) + 0.5*(nl[1]*ql.p + nr[1]*qr.p);
    out[4] = 0.5*c*(ql.w + qr.w) + 0.5*(nl[2]*ql.p + nr[2]*qr.p);
}

// cent_keep_scheme_t<gas, order> (convective.h:117-184)
void f_keep(const Ctx& C, const Face& f, int order, double out[5])
{
    const int hw = order/2;
    double nface[3]; face_metric(C.G, f, nface);
    double un[8], rho[8], eng[8];
    Q q[8];
    for (int ii = 0; ii < order; ++ii)
    {
        const int o = -(2*hw - 1) + 2*ii;
        q[ii] = cell_at(C.G, C.a, f, o);
        double n[3]; cell_metric(C.G, f, o, n);
        un[ii]  = q[ii].u*n[0] + q[ii].v*n[1] + q[ii].w*n[2];
        rho[ii] = q[ii].p/(C.gas.R*q[ii].T);
        eng[ii] = q[ii].p/(rho[ii]*(C.gas.gamma - 1.0));
    }
    double c = 0.0, mx = 0.0, my = 0.0, mz = 0.0, g = 0.0, k = 0.0, ie = 0.0, pp = 0.0;
    for (int ii = 1; ii <= hw; ++ii)
    {
        const double a = cfd_coeff(ii, order);
        for (int jj = 0; jj < ii; ++jj)
        {
            const int i0 = hw - 1 - jj, i1 = i0 + ii;
            const Q& q0 = q[i0]; const Q& q1 = q[i1];
            const double c_loc = a*(0.25*(rho[i0] + rho[i1])*(un[i1] + un[i0]));
            c  += c_loc;
            mx += c_loc*(0.5*(q0.u + q1.u));
            my += c_loc*(0.5*(q0.v + q1.v));
            mz += c_loc*(0.5*(q0.w + q1.w));
            g  += a*(0.5*(q0.p + q1.p));
            k  += 0.5*c_loc*(q0.u*q1.u + q0.v*q1.v + q0.w*q1.w);
            ie += c_loc*(0.5*(eng[i0] + eng[i1]));
            pp += a*(0.5*(un[i0]*q1.p + un[i1]*q0.p));
        }
    }
    out[0] = 2.0*c;
    out[1] = 2.0*(k + ie + pp);
    out[2] = 2.0*(mx + g*nface[0]);
    out[3] = 2.0*(my + g*nface[1]);
    out[4] = 2.0*(mz + g*nface[2]);
}

// ducros_t::get_sensor (state_sensor.h:32-42)
double ducros(const double gr[3][5], double eps)
{
    double theta_sq = gr[0][2] + gr[1][3] + gr[2][4];
    theta_sq *= theta_sq;
    const double v0 = gr[1][4] - gr[2][3];
    const double v1 = gr[2][2] - gr[0][4];
    const double v2 = gr[0][3] - gr[1][2];
    const double vort_sq = v0*v0 + v1*v1 + v2*v2;
    return theta_sq/(theta_sq + vort_sq + eps);
}

// pressure_diss_lr (convective.h:212-245)
void f_pdiss(const Ctx& C, const Face& f, const ora_term& t, double out[5])
{
    const Q ql = cell_at(C.G, C.a, f, -1), qr = cell_at(C.G, C.a, f, 1);
    const double delta_p = ql.p - qr.p;
    const double delta_T = ql.T - qr.T;
    const double pavg = 0.5*(ql.p + qr.p);
    const double Tavg = 0.5*(ql.T + qr.T);
    const double inv_RT = 1.0/(0.5*(C.gas.R + C.gas.R)*Tavg);
    const double p_inv_rt2 = pavg*inv_RT/Tavg;
    const double gamma = 0.5*(C.gas.gamma + C.gas.gamma);
    const double k = 0.5*(ql.u*qr.u + ql.v*qr.v + ql.w*qr.w);
    const double u = 0.5*(ql.u + qr.u);
    const double v = 0.5*(ql.v + qr.v);
    const double w = 0.5*(ql.w + qr.w);
    double scl = t.sensor_param;
    if (t.sensor == ORA_SENSOR_DUCROS)
    {
        double gr[3][5]; face_grad(C.G, C.a, f, gr);
        scl = ducros(gr, t.sensor_param);
    }
    out[0] = scl*t.eps_p*delta_p*inv_RT;
    out[1] = scl*t.eps_p*delta_p*(k*inv_RT + 1.0/gamma);
    out[2] = scl*t.eps_p*delta_p*u*inv_RT;
    out[3] = scl*t.eps_p*delta_p*v*inv_RT;
    out[4] = scl*t.eps_p*delta_p*w*inv_RT;
    out[0] -= scl*t.eps_T*delta_T*p_inv_rt2;
    out[1] -= scl*t.eps_T*delta_T*p_inv_rt2*(0.5*k);
    out[2] -= scl*t.eps_T*delta_T*p_inv_rt2*u;
    out[3] -= scl*t.eps_T*delta_T*p_inv_rt2*v;
    out[4] -= scl*t.eps_T*delta_T*p_inv_rt2*w;
}

// rusanov_t (flux_funcs.h:19-52): fu, fd
void split_rusanov(const ora_gas& gas, const Q& q, const double nv[3], double fu[5], double fd[5])
{
    const double u_n    = nv[0]*q.u + nv[1]*q.v + nv[2]*q.w;
    const double u_norm = std::sqrt(q.u*q.u + q.v*q.v + q.w*q.w);
    const double rho    = q.p/(gas.R*q.T);
    const double pdiff  = q.p/rho;
    const double engy   = 0.5*(q.u*q.u + q.v*q.v + q.w*q.w) + q.p/(rho*(gas.gamma - 1.0));
    fu[0] = 0.5*rho*u_n;
    fu[1] = 0.5*rho*u_n*(engy + pdiff);
    fu[2] = 0.5*rho*q.u*u_n + 0.5*q.p*nv[0];
    fu[3] = 0.5*rho*q.v*u_n + 0.5*q.p*nv[1];
    fu[4] = 0.5*rho*q.w*u_n + 0.5*q.p*nv[2];
    for (int n = 0; n < 5; ++n) fd[n] = fu[n];
    const double sigma = 0.5*(u_norm + std::sqrt(gas.gamma*gas.R*q.T));
    fd[0] -= sigma*rho;
    fd[1] -= sigma*rho*engy;
    fd[2] -= sigma*rho*q.u;
    fd[3] -= sigma*rho*q.v;
    fd[4] -= sigma*rho*q.w;
    fu[0] += sigma*rho;
    fu[1] += sigma*rho*engy;
    fu[2] += sigma*rho*q.u;
    fu[3] += sigma*rho*q.v;
    fu[4] += sigma*rho*q.w;
}

// phys_flux_t (flux_funcs.h:107-124)
void split_phys(const ora_gas& gas, const Q& q, const double nv[3], double fu[5], double fd[5])
{
    const double u_n  = nv[0]*q.u + nv[1]*q.v + nv[2]*q.w;
    const double rho  = q.p/(gas.R*q.T);
    const double enth = 0.5*(q.u*q.u + q.v*q.v + q.w*q.w) + q.p/rho + q.p/(rho*(gas.gamma - 1.0));
    fu[0] = 0.5*rho*u_n;
    fu[1] = 0.5*rho*u_n*enth;
    fu[2] = 0.5*rho*q.u*u_n + 0.5*q.p*nv[0];
    fu[3] = 0.5*rho*q.v*u_n + 0.5*q.p*nv[1];
    fu[4] = 0.5*rho*q.w*u_n + 0.5*q.p*nv[2];
    for (int n = 0; n < 5; ++n) fd[n] = fu[n];
}

// weno_t (convective.h:275-332)
void f_weno(const Ctx& C, const Face& f, const ora_term& t, double out[5])
{
    double fu[4][5], fd[4][5];
    for (int c = 0; c < 4; ++c)
    {
        const int o = -3 + 2*c;
        const Q q = cell_at(C.G, C.a, f, o);
        double n[3]; cell_metric(C.G, f, o, n);
        if (t.split == ORA_SPLIT_PHYS) split_phys(C.gas, q, n, fu[c], fd[c]);
        else                           split_rusanov(C.gas, q, n, fu[c], fd[c]);
    }
    for (int k = 0; k < 5; ++k)
    {
        const double r0 = -0.5*fu[0][k] + 1.5*fu[1][k];
        const double r1 =  0.5*fu[1][k] + 0.5*fu[2][k];
        const double r2 =  0.5*fd[1][k] + 0.5*fd[2][k];
        const double r3 =  1.5*fd[2][k] - 0.5*fd[3][k];
        if (!t.smooth)
        {
            out[k] = (1.0/3.0)*r0 + (2.0/3.0)*r1 + (2.0/3.0)*r2 + (1.0/3.0)*r3;
        }
        else
        {
            const double be0 = (fu[0][k] - fu[1][k])*(fu[0][k] - fu[1][k]);
            const double be1 = (fu[1][k] - fu[2][k])*(fu[1][k] - fu[2][k]);
            const double be2 = (fd[1][k] - fd[2][k])*(fd[1][k] - fd[2][k]);
            const double be3 = (fd[2][k] - fd[3][k])*(fd[2][k] - fd[3][k]);
            const double eps = 1e-16;
            const double a0 = (1.0/3.0)/((be0 + eps)*(be0 + eps));
            const double a1 = (2.0/3.0)/((be1 + eps)*(be1 + eps));
            const double a2 = (2.0/3.0)/((be2 + eps)*(be2 + eps));
            const double a3 = (1.0/3.0)/((be3 + eps)*(be3 + eps));
            const double w0 = a0/(a0 + a1);
            const double w1 = 1.0 - w0;
            const double w2 = a2/(a2 + a3);
            const double w3 = 1.0 - w2;
            out[k] = w0*r0 + w1*r1 + w2*r2 + w3*r3;
        }
    }
}

// fweno_t (convective.h:355-496)
void f_fweno(const Ctx& C, const Face& f, const ora_term& t, double out[5])
{
    Q q[4]; double nv[4][3];
    double rho[4], hsr[4], fluxv[4], dissv[4];
    for (int i = 0; i < 4; ++i)
    {
        const int o = -3 + 2*i;
        q[i] = cell_at(C.G, C.a, f, o);
        cell_metric(C.G, f, o, nv[i]);
        rho[i] = C.gas.R*q[i].T;
        hsr[i] = q[i].u*q[i].u;
        hsr[i] += q[i].v*q[i].v;
        hsr[i] += q[i].w*q[i].w;
        hsr[i] = std::sqrt(hsr[i]);
        hsr[i] += std::sqrt(rho[i]*C.gas.gamma);
        rho[i] = q[i].p/rho[i];
        hsr[i] *= rho[i];
        hsr[i] *= 0.5;
    }
    auto apply_weno = [&](int k)
    {
        const double f0u = fluxv[0] + dissv[0], f0d = fluxv[0] - dissv[0];
        const double f1u = fluxv[1] + dissv[1], f1d = fluxv[1] - dissv[1];
        const double f2u = fluxv[2] + dissv[2], f2d = fluxv[2] - dissv[2];
        const double f3u = fluxv[3] + dissv[3], f3d = fluxv[3] - dissv[3];
        (void)f3u; (void)f0d;
        const double r0 = -0.5*f0u + 1.5*f1u;
        const double r1 =  0.5*f1u + 0.5*f2u;
        const double r2 =  0.5*f1d + 0.5*f2d;
        const double r3 =  1.5*f2d - 0.5*f3d;
        if (!t.smooth)
        {
            out[k] = (1.0/3.0)*r0 + (2.0/3.0)*r1 + (2.0/3.0)*r2 + (1.0/3.0)*r3;
        }
        else
        {
            const double eps = 1e-16;
            double as[4] = {f0u - f1u, f1u - f2u, f1d - f2d, f2d - f3d};
            for (int n = 0; n < 4; ++n) { as[n] *= as[n]; as[n] += eps; as[n] *= as[n]; }
            as[0] += as[0]; as[0] += as[1]; as[0] = as[1]/as[0];
            as[3] += as[3]; as[3] += as[2]; as[3] = as[2]/as[3];
            as[1] = 1.0 - as[0];
            as[2] = 1.0 - as[3];
            out[k] = as[0]*r0 + as[1]*r1 + as[2]*r2 + as[3]*r3;
        }
    };
    for (int i = 0; i < 4; ++i)
    {
        fluxv[i]  = q[i].u*nv[i][0];
        fluxv[i] += q[i].v*nv[i][1];
        fluxv[i] += q[i].w*nv[i][2];
        fluxv[i] *= 0.5*rho[i];
        dissv[i]  = hsr[i];
    }
    apply_weno(0);
    for (int i = 0; i < 4; ++i)
    {
        double engy = q[i].u*q[i].u;
        engy += q[i].v*q[i].v;
        engy += q[i].w*q[i].w;
        engy *= 0.5;
        const double pdiff = C.gas.R*q[i].T;
        engy += pdiff/(C.gas.gamma - 1.0);
        fluxv[i] *= (engy + pdiff);
        dissv[i] *= engy;
    }
    apply_weno(1);
    for (int dr = 0; dr < 3; ++dr)
    {
        for (int i = 0; i < 4; ++i)
        {
            fluxv[i]  = q[i].u*nv[i][0];
            fluxv[i] += q[i].v*nv[i][1];
            fluxv[i] += q[i].w*nv[i][2];
            fluxv[i] *= rho[i];
            fluxv[i] *= q[i].vel(dr);
            fluxv[i] += q[i].p*nv[i][dr];
            fluxv[i] *= 0.5;
            dissv[i]  = hsr[i];
            dissv[i] *= q[i].vel(dr);
        }
        apply_weno(2 + dr);
    }
}

// weno_fds_t<rusanov_fds_t> (convective.h:499-555, flux_funcs.h:54-95): the
// primitive states are reconstructed at the face from cells i-2 .. i+1
// (offsets -3, -1, 1, 3), each component as the reference's prim_t arithmetic
// rounds it: with nonlinear weights enabled the reference keeps the central
// candidates ql1 = qr0 = 0.5 q1 + 0.5 q2 ("nothing for now"); with linear
// weights ql = 2/3 ql1 + 1/3 ql0, qr = 2/3 qr0 + 1/3 qr1.  The Rusanov
// flux-difference form then takes the face value and face metric:
// sigma = u_n(face) + c(face) and F = sigma (U(qL) - U(qR)) with U =
// (rho, rho e_tot, rho u, rho v, rho w), no physical-flux part.
void f_weno_fds(const Ctx& C, const Face& f, const ora_term& t, double out[5])
{
    const Q q0 = cell_at(C.G, C.a, f, -3), q1 = cell_at(C.G, C.a, f, -1);
    const Q q2 = cell_at(C.G, C.a, f, 1), q3 = cell_at(C.G, C.a, f, 3);
    double ql[5], qr[5];
    for (int n = 0; n < 5; ++n)
    {
        const double ql0 = -0.5*q0[n] + 1.5*q1[n];
        const double ql1 =  0.5*q1[n] + 0.5*q2[n];
        const double qr0 =  0.5*q1[n] + 0.5*q2[n];
        const double qr1 =  1.5*q2[n] - 0.5*q3[n];
        ql[n] = ql1;
        qr[n] = qr0;
        if (!t.smooth)
        {
            ql[n] *= 2.0/3.0;
            qr[n] *= 2.0/3.0;
            ql[n] += (1.0/3.0)*ql0;
            qr[n] += (1.0/3.0)*qr1;
        }
    }
    const Q qf = face_value(C.G, C.a, f);
    double nv[3]; face_metric(C.G, f, nv);
    const double R = C.gas.R, gam = C.gas.gamma;
    const double u_n   = nv[0]*qf.u + nv[1]*qf.v + nv[2]*qf.w;
    const double rhoL  = ql[0]/(R*ql[1]);
    const double engyL = 0.5*(ql[2]*ql[2] + ql[3]*ql[3] + ql[4]*ql[4]) + ql[0]/(rhoL*(gam - 1.0));
    const double rhoR  = qr[0]/(R*qr[1]);
    const double engyR = 0.5*(qr[2]*qr[2] + qr[3]*qr[3] + qr[4]*qr[4]) + qr[0]/(rhoR*(gam - 1.0));
    const double sigma = u_n + std::sqrt(gam*R*qf.T);
    out[0] = sigma*rhoL;
    out[1] = sigma*rhoL*engyL;
    out[2] = sigma*rhoL*ql[2];
    out[3] = sigma*rhoL*ql[3];
    out[4] = sigma*rhoL*ql[4];
    out[0] -= sigma*rhoR;
    out[1] -= sigma*rhoR*engyR;
    out[2] -= sigma*rhoR*qr[2];
    out[3] -= sigma*rhoR*qr[3];
    out[4] -= sigma*rhoR*qr[4];
}

// first_order_t (convective.h:27-32): fu(L) + fd(R) with rusanov_t
void f_first(const Ctx& C, const Face& f, double out[5])
{
    double fuL[5], fdL[5], fuR[5], fdR[5], nl[3], nr[3];
    cell_metric(C.G, f, -1, nl); cell_metric(C.G, f, 1, nr);
    split_rusanov(C.gas, cell_at(C.G, C.a, f, -1), nl, fuL, fdL);
    split_rusanov(C.gas, cell_at(C.G, C.a, f,  1), nr, fuR, fdR);
    for (int n = 0; n < 5; ++n) out[n] = fuL[n] + fdR[n];
}

// WALE subgrid viscosity at a face (subgrid_scale.h:44-84, after Nicoud and
// Ducros 1999): with g_ij = d u_i / d x_j from the face gradient,
//   S_ij = (g_ij + g_ji)/2,  Sd_ij = ((g^2)_ij + (g^2)_ji)/2 - delta_ij tr(g^2)/3,
//   mu_t = rho cw^2 delta^2 |Sd|^2 |Sd| / (eps + |S|^2 |S|^2 |S| + |Sd|^2 |Sd|^(1/2))
// with |A|^2 = sum A_ij^2, eps = 1e-8, rho = p/(R T) of the face value; every
// sum in the reference's order (columns outer, rows inner for |A|^2).
double wale_mu_t(const ora_term& t, const Q& qf, const double g[3][5], double rgas)
{
    const double rho = qf.p/(rgas*qf.T);
    auto gij = [&](int i, int j) { return g[j][2 + i]; };
    double g2[3][3], sd[3][3], s[3][3];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
            g2[i][j] = gij(i, 0)*gij(0, j) + gij(i, 1)*gij(1, j) + gij(i, 2)*gij(2, j);
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
        {
            sd[i][j] = 0.5*(g2[i][j] + g2[j][i]);
            s[i][j]  = 0.5*(gij(i, j) + gij(j, i));
        }
    double tr = 0.0;
    for (int d = 0; d < 3; ++d) tr += 0.33333333333333333*g2[d][d];
    for (int d = 0; d < 3; ++d) sd[d][d] -= tr;
    auto sq = [](const double a[3][3])
    {
        double acc = 0.0;
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) acc += a[i][j]*a[i][j];
        return acc;
    };
    const double ssd = sq(sd), ss = sq(s);
    const double r0 = std::sqrt(ssd), r1 = std::sqrt(r0), r2 = std::sqrt(ss);
    return rho*t.sgs_cw*t.sgs_cw*t.sgs_delta*t.sgs_delta*ssd*r0/(1e-8 + ss*ss*r2 + ssd*r1);
}

// viscous law: mu, beta, alpha (viscous_laws.h:81-84, 121-134); with a
// subgrid model, sgs_visc_t::get_all (viscous_laws.h:186-196) adds mu_t,
// -0.66666666667 mu_t and mu_t / Pr_t
void visc_law(const ora_term& t, const Q& qf, const double g[3][5], double rgas,
              double& mu, double& beta, double& alpha)
{
    if (t.law == ORA_LAW_POWER)
    {
        mu    = t.mu*std::pow(qf.T/t.T_ref, t.power);
        beta  = -0.66666666667*(t.mu*std::pow(qf.T/t.T_ref, t.power));
        alpha = (t.mu*std::pow(qf.T/t.T_ref, t.power))/t.prandtl;
    }
    else
    {
        mu    = t.mu;
        beta  = double(-2.0*t.mu/3.0);
        alpha = t.mu*(1.0/t.prandtl);
    }
    if (t.sgs == ORA_SGS_WALE)
    {
        const double mu_t = wale_mu_t(t, qf, g, rgas);
        mu    = mu + mu_t;
        beta  = beta - 0.66666666667*mu_t;
        alpha = alpha + mu_t/t.sgs_prt;
    }
}

// visc_lr (viscous.h:39-112)
void f_visc(const Ctx& C, const Face& f, const ora_term& t, double out[5])
{
    const Q qf = face_value(C.G, C.a, f);
    double g[3][5]; face_grad(C.G, C.a, f, g);
    double n[3]; face_metric(C.G, f, n);
    double visc, visc2, alpha;
    visc_law(t, qf, g, C.gas.R, visc, visc2, alpha);
    enum { P = 0, T = 1, U = 2, V = 3, W = 4 };
    const double div = g[0][U] + g[1][V] + g[2][W];
    double tau[3][3];
    tau[0][0] = visc*(g[0][U] + g[0][U]) + visc2*div;
    tau[0][1] = visc*(g[0][V] + g[1][U]);
    tau[0][2] = visc*(g[0][W] + g[2][U]);
    tau[1][2] = visc*(g[1][W] + g[2][V]);
    tau[1][1] = visc*(g[1][V] + g[1][V]) + visc2*div;
    tau[1][0] = tau[0][1];
    tau[2][0] = tau[0][2];
    tau[2][1] = tau[1][2];
    tau[2][2] = visc*(g[2][W] + g[2][W]) + visc2*div;
    const double gam = C.gas.gamma, rgas = C.gas.R;
    const double spec_heat = gam*rgas/(gam - 1.0);
    const double cond = spec_heat*alpha;
    double ht[3];
    for (int r = 0; r < 3; ++r) ht[r] = qf.u*tau[r][0] + qf.v*tau[r][1] + qf.w*tau[r][2] + cond*g[r][T];
    out[0] = 0.0;
    out[1] = -(n[0]*ht[0] + n[1]*ht[1] + n[2]*ht[2]);
    out[2] = -(n[0]*tau[0][0] + n[1]*tau[0][1] + n[2]*tau[0][2]);
    out[3] = -(n[0]*tau[1][0] + n[1]*tau[1][1] + n[2]*tau[1][2]);
    out[4] = -(n[0]*tau[2][0] + n[1]*tau[2][1] + n[2]*tau[2][2]);
}

// hybrid_scheme_t (navier-stokes/hybrid_scheme.h:14-48): central cent_keep<order>
// blended with a WENO flux by the blender's face sensor b:
// full_flux: (1 - b) F0 + b F1; diss_flux: F0 + b F1 (coefficient 1 applied)
void f_hybrid(const Ctx& C, const Face& f, const ora_term& t, double out[5])
{
    double blend = t.sensor_param;
    if (t.sensor == ORA_SENSOR_DUCROS)
    {
        double gr[3][5]; face_grad(C.G, C.a, f, gr);
        blend = ducros(gr, t.sensor_param);
    }
    double f0[5], f1[5];
    f_keep(C, f, t.order, f0);
    if (t.diss_kind == ORA_TERM_FWENO) f_fweno(C, f, t, f1);
    else                               f_weno(C, f, t, f1);
    double coeff0 = 1.0 - blend;
    const double coeff1 = blend;
    if (!t.full) coeff0 = 1.0;
    for (int n = 0; n < 5; ++n) f0[n] *= coeff0;
    for (int n = 0; n < 5; ++n) f1[n] *= coeff1;
    for (int n = 0; n < 5; ++n) out[n] = f0[n] + f1[n];
}

void eval_term(const Ctx& C, const Face& f, const ora_term& t, double out[5])
{
    switch (t.kind)
    {
        case ORA_TERM_KEEP:   f_keep(C, f, t.order, out); break;
        case ORA_TERM_TOTANI: f_totani(C, f, out); break;
        case ORA_TERM_PDISS:  f_pdiss(C, f, t, out); break;
        case ORA_TERM_WENO:   f_weno(C, f, t, out); break;
        case ORA_TERM_FWENO:  f_fweno(C, f, t, out); break;
        case ORA_TERM_FIRST:  f_first(C, f, out); break;
        case ORA_TERM_VISC:   f_visc(C, f, t, out); break;
        case ORA_TERM_HYBRID: f_hybrid(C, f, t, out); break;
        case ORA_TERM_WENO_FDS: f_weno_fds(C, f, t, out); break;
        default: for (int n = 0; n < 5; ++n) out[n] = 0.0;
    }
}

// compose(t0, t1, ..., tn) = t0 + (t1 + (... + tn)) (omni/compose.h:19-24)
void eval_flux(const Ctx& C, const Face& f, const ora_flux& fl, double out[5])
{
    eval_term(C, f, fl.terms[fl.nterms - 1], out);
    for (int m = fl.nterms - 2; m >= 0; --m)
    {
        double z[5];
        eval_term(C, f, fl.terms[m], z);
        for (int n = 0; n < 5; ++n) { z[n] += out[n]; out[n] = z[n]; }
    }
}

double cell_jac(const Geo& G, int lb, int i, int j, int k)
{
    if (G.identity()) return 1.0;
    double xi[3]; cell_xi(G, lb, i, j, k, xi);
    return 1.0/(G.deriv(0, xi[0])*G.deriv(1, xi[1])*G.deriv(2, xi[2]));
}

// prim -> cons, cons -> prim (fluid_state.h:102-135)
inline void p2c(const ora_gas& gas, const double* q, double* c)
{
    const double rho   = q[0]/(gas.R*q[1]);
    const double rhoU2 = rho*(q[2]*q[2] + q[3]*q[3] + q[4]*q[4]);
    const double rhoE  = 0.5*rhoU2 + (q[0]/((gas.gamma - 1.0)));
    c[0] = rho; c[1] = rhoE; c[2] = rho*q[2]; c[3] = rho*q[3]; c[4] = rho*q[4];
}
inline void c2p(const ora_gas& gas, const double* c, double* q)
{
    const double rho = c[0];
    const double invrho = 1.0/rho;
    const double u = invrho*c[2], v = invrho*c[3], w = invrho*c[4];
    const double rhoU2 = rho*(u*u + v*v + w*w);
    const double p = (gas.gamma - 1.0)*(c[1] - 0.5*rhoU2);
    const double T = p/(gas.R*rho);
    q[0] = p; q[1] = T; q[2] = u; q[3] = v; q[4] = w;
}

template <typename F> void for_interior(const Geo& G, F&& fn)
{
    #pragma omp parallel for schedule(static)
    for (int lb = 0; lb < G.nblk; ++lb)
        for (int k = 0; k < G.g.nc[2]; ++k)
            for (int j = 0; j < G.g.nc[1]; ++j)
                for (int i = 0; i < G.g.nc[0]; ++i) fn(i, j, k, lb);
}

// rhs of interior cell (i, j, k) of block lb, flux_div_basic's loop body
// (flux_div_basic.h:47-72), accumulating onto the current rhs value
void flux_div_cell(const Geo& G, const Ctx& C, const ora_flux& fl, double* rhs, int i, int j, int k, int lb)
{
    const double jac = cell_jac(G, lb, i, j, k);
    double* r = rhs + G.cell(i, j, k, lb);
    double relem[5] = {r[0], r[1], r[2], r[3], r[4]};
    for (int idir = 0; idir < 3; ++idir)
    {
        const double inv_dx = 1.0/G.dx[3*lb + idir];
        for (int pm = 0; pm < 2; ++pm)
        {
            Face f{idir, i, j, k, lb};
            if (pm == 1) { if (idir == 0) f.i++; else if (idir == 1) f.j++; else f.k++; }
            double acc[5] = {0, 0, 0, 0, 0};
            double fx[5];
            eval_flux(C, f, fl, fx);
            for (int n = 0; n < 5; ++n) acc[n] += fx[n];
            const double s = jac*(1.0 - 2.0*pm)*inv_dx;
            for (int n = 0; n < 5; ++n) { acc[n] *= s; relem[n] += acc[n]; }
        }
    }
    for (int n = 0; n < 5; ++n) r[n] = relem[n];
}

void flux_div(const Geo& G, const double* prim, double* rhs, const ora_flux& fl, const ora_gas& gas, int increment)
{
    if (!increment)
    {
        const size_t tot = (size_t)5*G.N[0]*G.N[1]*G.N[2]*G.nblk;
        std::fill(rhs, rhs + tot, 0.0);
    }
    const Ctx C{G, prim, gas};
    for_interior(G, [&](int i, int j, int k, int lb) { flux_div_cell(G, C, fl, rhs, i, j, k, lb); });
}

// 26-neighbour exchange (exchange_pattern.h:382-424, get_transaction.h:55-85)
void exchange(const Geo& G, double* a)
{
    const int ng = G.g.ng;
    for (int lb = 0; lb < G.nblk; ++lb)
    {
        const int b[3] = {G.bpos[3*lb], G.bpos[3*lb+1], G.bpos[3*lb+2]};
        for (int ez = -1; ez <= 1; ++ez)
        for (int ey = -1; ey <= 1; ++ey)
        for (int ex = -1; ex <= 1; ++ex)
        {
            if (ex == 0 && ey == 0 && ez == 0) continue;
            const int e[3] = {ex, ey, ez};
            bool skip = false;
            int nb[3];
            for (int d = 0; d < 3; ++d)
            {
                const bool lo = (b[d] == 0), hi = (b[d] == G.g.nb[d] - 1);
                if (!G.g.periodic[d] && ((lo && e[d] == -1) || (hi && e[d] == 1))) skip = true;
                nb[d] = b[d] + e[d];
                if (nb[d] < 0) nb[d] += G.g.nb[d];
                if (nb[d] >= G.g.nb[d]) nb[d] -= G.g.nb[d];
            }
            if (skip) continue;
            const int lt = nb[0] + G.g.nb[0]*(nb[1] + G.g.nb[1]*nb[2]);
            int smin[3], smax[3], dmin[3];
            for (int d = 0; d < 3; ++d)
            {
                const int n = G.g.nc[d];
                if (e[d] == -1)     { smin[d] = 0;      smax[d] = ng; dmin[d] = n; }
                else if (e[d] == 0) { smin[d] = 0;      smax[d] = n;  dmin[d] = 0; }
                else                { smin[d] = n - ng; smax[d] = n;  dmin[d] = -ng; }
            }
            for (int k = smin[2]; k < smax[2]; ++k)
            for (int j = smin[1]; j < smax[1]; ++j)
            for (int i = smin[0]; i < smax[0]; ++i)
            {
                const double* s = a + G.cell(i, j, k, lb);
                double* dd = a + G.cell(dmin[0] + i - smin[0], dmin[1] + j - smin[1], dmin[2] + k - smin[2], lt);
                for (int v = 0; v < 5; ++v) dd[v] = s[v];
            }
        }
    }
}

} // namespace

// boundary_fill over the faces in face_mask, in face order xmin, xmax, ymin,
// ymax, zmin, zmax (grid/boundary_fill.h:32-133): on every block whose face
// lies on the domain boundary, every ghost cell beyond that face (normal index
// outside [0, n), tangential indices over the whole stored range) gets
// ghost(fill index, image cell value, current ghost value).
template <typename F> void fill_faces(const Geo& G, double* a, int face_mask, F&& ghost)
{
    const int ng = G.g.ng;
    for (int face = 0; face < 6; ++face)
    {
        if (!(face_mask & (1 << face))) continue;
        const int idir = face/2, pm = face % 2;
        for (int lb = 0; lb < G.nblk; ++lb)
        {
            const int bp = G.bpos[3*lb + idir];
            if (pm == 0 && bp != 0) continue;
            if (pm == 1 && bp != G.g.nb[idir] - 1) continue;
            int ll[3], ur[3];
            for (int d = 0; d < 3; ++d) { ll[d] = -ng; ur[d] = G.g.nc[d] + ng; }
            if (pm == 0) ur[idir] = 0;
            if (pm == 1) ll[idir] = G.g.nc[idir];
            for (int k = ll[2]; k < ur[2]; ++k)
            for (int j = ll[1]; j < ur[1]; ++j)
            for (int i = ll[0]; i < ur[0]; ++i)
            {
                const int fi[3] = {i, j, k};
                ghost(lb, idir, pm, fi, a + G.cell(i, j, k, lb));
            }
        }
    }
}

extern "C" {

int ora_array_size(const ora_grid* g)
{
    const Geo G(g);
    return 5*G.N[0]*G.N[1]*G.N[2]*G.nblk;
}

int ora_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

int ora_flux_div(const ora_grid* g, const ora_gas* gas, const ora_flux* fl, const double* prim, double* rhs, int increment)
{
    if (!g || !gas || !fl || !prim || !rhs || fl->nterms < 1 || fl->nterms > ORA_MAX_TERMS) return -1;
    const Geo G(g);
    flux_div(G, prim, rhs, *fl, *gas, increment);
    return 0;
}

int ora_flux_div_blocks(const ora_grid* g, const ora_gas* gas, const ora_flux* fl, const double* prim, double* rhs,
                        int increment, const int* blocks, int nblocks)
{
    if (!g || !gas || !fl || !prim || !rhs || !blocks || fl->nterms < 1 || fl->nterms > ORA_MAX_TERMS) return -1;
    const Geo G(g);
    for (int b = 0; b < nblocks; ++b) if (blocks[b] < 0 || blocks[b] >= G.nblk) return -1;
    const Ctx C{G, prim, *gas};
    for (int b = 0; b < nblocks; ++b)
    {
        const int lb = blocks[b];
        for (int k = 0; k < G.g.nc[2]; ++k)
            for (int j = 0; j < G.g.nc[1]; ++j)
                for (int i = 0; i < G.g.nc[0]; ++i)
                {
                    if (!increment) { double* r = rhs + G.cell(i, j, k, lb); for (int n = 0; n < 5; ++n) r[n] = 0.0; }
                    flux_div_cell(G, C, *fl, rhs, i, j, k, lb);
                }
    }
    return 0;
}

int ora_exchange(const ora_grid* g, double* a)
{
    if (!g || !a) return -1;
    const Geo G(g);
    exchange(G, a);
    return 0;
}

// boundary_fill with a no-slip isothermal mirror kernel: ghost =
// (p, 2 T_wall - T, -u, -v, -w) of the mirror cell -1-i / 2n-(i+1)
// (grid/boundary_fill.h:104-130).  face = 2*dir + pm.
int ora_boundary_fill_wall(const ora_grid* g, double* a, int face, double t_wall)
{
    if (!g || !a || face < 0 || face > 5) return -1;
    const Geo G(g);
    const int idir = face/2, pm = face % 2, ng = G.g.ng;
    for (int lb = 0; lb < G.nblk; ++lb)
    {
        const int bp = G.bpos[3*lb + idir];
        if (pm == 0 && bp != 0) continue;
        if (pm == 1 && bp != G.g.nb[idir] - 1) continue;
        int ll[3], ur[3];
        for (int d = 0; d < 3; ++d) { ll[d] = -ng; ur[d] = G.g.nc[d] + ng; }
        if (pm == 0) ur[idir] = 0;
        if (pm == 1) ll[idir] = G.g.nc[idir];
        for (int k = ll[2]; k < ur[2]; ++k)
        for (int j = ll[1]; j < ur[1]; ++j)
        for (int i = ll[0]; i < ur[0]; ++i)
        {
            int fi[3] = {i, j, k}, im[3] = {i, j, k};
            if (pm == 0) im[idir] = -1 - fi[idir];
            else         im[idir] = 2*G.g.nc[idir] - (fi[idir] + 1);
            const double* s = a + G.cell(im[0], im[1], im[2], lb);
            double* dd = a + G.cell(fi[0], fi[1], fi[2], lb);
            const double q[5] = {s[0], 2.0*t_wall - s[1], -s[2], -s[3], -s[4]};
            for (int v = 0; v < 5; ++v) dd[v] = q[v];
        }
    }
    return 0;
}

// extrap_t<order> (boundary_fill.h:17-26, 63-98): Lagrange extrapolation
// through the order + 1 cells i0 .. i0 + order along the face normal (i0 = 0
// at a min face, n - order - 1 at a max face), basis weights built as
// l_j = l_j (x - x_i) / (x_j - x_i) over ascending i != j on integer cell
// indices, the ghost accumulated as g = g + l_j q_j from g = 0.
int ora_boundary_fill_extrap(const ora_grid* g, double* a, int face_mask, int order)
{
    if (!g || !a || order < 0) return -1;
    const Geo G(g);
    for (int d = 0; d < 3; ++d) if (order + 1 > G.g.nc[d] && (face_mask & (3 << (2*d)))) return -1;
    fill_faces(G, a, face_mask, [&](int lb, int idir, int pm, const int fi[3], double* dst)
    {
        const int i0 = pm*(G.g.nc[idir] - order - 1);
        double gv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int j = 0; j <= order; ++j)
        {
            double lj = 1.0;
            for (int i = 0; i <= order; ++i)
            {
                if (i == j) continue;
                const int x = fi[idir], xj = i0 + j, xi = i0 + i;
                lj = lj*(x - xi)/(xj - xi);
            }
            int src[3] = {fi[0], fi[1], fi[2]};
            src[idir] = i0 + j;
            const double* e = a + G.cell(src[0], src[1], src[2], lb);
            for (int v = 0; v < 5; ++v) gv[v] = gv[v] + lj*e[v];
        }
        for (int v = 0; v < 5; ++v) dst[v] = gv[v];
    });
    return 0;
}

// A boundary_fill kernel (boundary_fill.h:99-127) of the affine family
// ghost[v] = dom[v] q_image[v] + fill[v] q_ghost[v] + add[v], the velocity
// component normal to the face taking the *_normal coefficients; a term with
// a zero coefficient is left out, the others are rounded in that order.
int ora_boundary_fill_affine(const ora_grid* g, double* a, int face_mask, const ora_fill_kernel* kern)
{
    if (!g || !a || !kern) return -1;
    const Geo G(g);
    fill_faces(G, a, face_mask, [&](int lb, int idir, int pm, const int fi[3], double* dst)
    {
        int im[3] = {fi[0], fi[1], fi[2]};
        if (pm == 0) im[idir] = -1 - fi[idir];
        else         im[idir] = 2*G.g.nc[idir] - (fi[idir] + 1);
        const double* d = a + G.cell(im[0], im[1], im[2], lb);
        double out[5];
        for (int v = 0; v < 5; ++v)
        {
            const bool nrm = (v == 2 + idir);
            const double cd = nrm ? kern->dom_normal : kern->dom[v];
            const double cf = nrm ? kern->fill_normal : kern->fill[v];
            const double ca = nrm ? kern->add_normal : kern->add[v];
            double r = 0.0;
            bool any = false;
            if (cd != 0.0) { r = cd*d[v]; any = true; }
            if (cf != 0.0) { const double t = cf*dst[v]; r = any ? r + t : t; any = true; }
            if (ca != 0.0) { r = any ? r + ca : ca; any = true; }
            out[v] = r;
        }
        for (int v = 0; v < 5; ++v) dst[v] = out[v];
    });
    return 0;
}

int ora_prim_to_cons(const ora_gas* gas, const double* prim, double* cons, long ncell)
{
    for (long c = 0; c < ncell; ++c) p2c(*gas, prim + 5*c, cons + 5*c);
    return 0;
}

int ora_cons_to_prim(const ora_gas* gas, const double* cons, double* prim, long ncell)
{
    for (long c = 0; c < ncell; ++c) c2p(*gas, cons + 5*c, prim + 5*c);
    return 0;
}

// Boundary conditions applied after every stage: exchange, then walls on the
// non-periodic faces listed in bc->wall_mask.
static void apply_bc(const Geo& G, double* q, const ora_bc* bc)
{
    exchange(G, q);
    if (bc)
        for (int face = 0; face < 6; ++face)
            if (bc->wall_mask & (1 << face)) ora_boundary_fill_wall(&G.g, q, face, bc->t_wall);
}

// One transform-aware SSP-RK3 step (advance.h:233-280 with
// transform_advance_to, advance.h:56-102; table explicit.h:80-88).
// q: primitive array (ghosts current); r0,r1,r2: residual scratch arrays.
int ora_ssprk3_step(const ora_grid* g, const ora_gas* gas, const ora_flux* fl, const ora_bc* bc,
                    double dt, double* q, double* r0, double* r1, double* r2)
{
    const Geo G(g);
    double* res[3] = {r0, r1, r2};
    // rows of the table (a_ij), last row = b
    const double A[4][3][2] = {
        {{0,1},{0,1},{0,1}},
        {{1,1},{0,1},{0,1}},
        {{1,4},{1,4},{0,1}},
        {{1,6},{1,6},{2,3}}};
    auto advance_to = [&](int prev, int curr, int nupd)
    {
        for_interior(G, [&](int i, int j, int k, int lb)
        {
            double* e = q + G.cell(i, j, k, lb);
            double cons[5]; p2c(*gas, e, cons);
            for (int s = 0; s < nupd; ++s)
            {
                // ratio difference a_curr - a_prev, reduced (std::ratio normalises)
                long n0 = (long)A[curr][s][0], d0 = (long)A[curr][s][1];
                long n1 = (long)A[prev][s][0], d1 = (long)A[prev][s][1];
                long num = d1*n0 - d0*n1, den = d0*d1;
                if (num == 0) continue;
                long a = std::labs(num), b = den;
                while (b) { long t = a % b; a = b; b = t; }
                num /= a; den /= a;
                const double tc = (double(num)/double(den))*dt;
                const double* rr = res[s] + G.cell(i, j, k, lb);
                double inc[5];
                for (int v = 0; v < 5; ++v) inc[v] = tc*rr[v];
                for (int v = 0; v < 5; ++v) cons[v] += inc[v];
            }
            c2p(*gas, cons, e);
        });
    };
    flux_div(G, q, res[0], *fl, *gas, 0);
    for (int st = 1; st < 3; ++st)
    {
        advance_to(st - 1, st, 3);
        apply_bc(G, q, bc);
        flux_div(G, q, res[st], *fl, *gas, 0);
    }
    advance_to(2, 3, 3);
    apply_bc(G, q, bc);
    return 0;
}

// value of the reduced ratio n/d as the reference's coeff_value_t computes it
// (numeric_t(num)/numeric_t(denum), time-integration/advance.h:33-46)
static double ratio_value(long n, long d)
{
    long a = std::labs(n), b = std::labs(d);
    while (b) { long t = a % b; a = b; b = t; }
    if (a == 0) return 0.0;
    n /= a; d /= a;
    if (d < 0) { n = -n; d = -d; }
    return double(n)/double(d);
}

// One step of an explicit table with the prim<->cons transform, advance.h:
// 233-280: stage 0's residual from q; stage i advances q in place by the row
// differences (a_i - a_{i-1}) dt (transform_advance_to, :56-102), then the
// boundary and residual i; the last update takes (b - a_{s-1}) dt.  res holds
// one array per stage.
int ora_rk_transform_step(const ora_grid* g, const ora_gas* gas, const ora_flux* fl, const ora_bc* bc,
                          const ora_rk_table* tb, double dt, double* q, double* const* res)
{
    if (!g || !gas || !fl || !tb || !q || !res || tb->stages < 1 || tb->stages > ORA_RK_MAX) return -1;
    const Geo G(g);
    const int S = tb->stages;
    auto row = [&](int r, int j, long& n, long& d)
    {
        if (r < S) { n = tb->a[r][j][0]; d = tb->a[r][j][1]; }
        else       { n = tb->b[j][0];    d = tb->b[j][1]; }
        if (d == 0) { n = 0; d = 1; }
    };
    auto advance_to = [&](int prev, int curr)
    {
        double tc[ORA_RK_MAX];
        bool nz[ORA_RK_MAX];
        for (int j = 0; j < S; ++j)
        {
            long n0, d0, n1, d1;
            row(curr, j, n0, d0); row(prev, j, n1, d1);
            const long num = d1*n0 - d0*n1, den = d0*d1;
            nz[j] = num != 0;
            tc[j] = nz[j] ? ratio_value(num, den)*dt : 0.0;
        }
        for_interior(G, [&](int i, int j, int k, int lb)
        {
            double* e = q + G.cell(i, j, k, lb);
            double cons[5]; p2c(*gas, e, cons);
            for (int st = 0; st < S; ++st)
            {
                if (!nz[st]) continue;
                const double* rr = res[st] + G.cell(i, j, k, lb);
                double inc[5];
                for (int v = 0; v < 5; ++v) inc[v] = tc[st]*rr[v];
                for (int v = 0; v < 5; ++v) cons[v] += inc[v];
            }
            c2p(*gas, cons, e);
        });
    };
    flux_div(G, q, res[0], *fl, *gas, 0);
    for (int st = 1; st < S; ++st)
    {
        advance_to(st - 1, st);
        apply_bc(G, q, bc);
        flux_div(G, q, res[st], *fl, *gas, 0);
    }
    advance_to(S - 1, S);
    apply_bc(G, q, bc);
    return 0;
}

// One step of an explicit table with the identity transform, advance.h:
// 107-230, the state stored in conserved variables: the residual of a stage
// converts every stored cell of the stage solution to primitives and takes
// flux_div of them; the boundary is the exchange of the conserved array.
// Residuals are scaled in place by dt a_ij, added to (and, with one solution
// copy, afterwards subtracted from) the stage solution and scaled back by
// 1/(dt a_ij), as the reference does.  q2: the second solution copy of a
// high-storage table (unused otherwise); prim: scratch for the primitives.
int ora_rk_identity_step(const ora_grid* g, const ora_gas* gas, const ora_flux* fl, const ora_rk_table* tb,
                         double dt, double* q, double* q2, double* const* res, double* prim)
{
    if (!g || !gas || !fl || !tb || !q || !res || !prim || tb->stages < 1 || tb->stages > ORA_RK_MAX) return -1;
    if (tb->high_storage && !q2) return -1;
    const Geo G(g);
    const int S = tb->stages;
    const size_t tot = (size_t)5*G.N[0]*G.N[1]*G.N[2]*G.nblk;
    auto scale = [&](double* a, double c) { for (size_t n = 0; n < tot; ++n) a[n] *= c; };
    auto incr = [&](double* a, const double* b, bool sub)
    {
        for_interior(G, [&](int i, int j, int k, int lb)
        {
            double* x = a + G.cell(i, j, k, lb);
            const double* y = b + G.cell(i, j, k, lb);
            for (int v = 0; v < 5; ++v) { if (sub) x[v] -= y[v]; else x[v] += y[v]; }
        });
    };
    auto rhs = [&](double* r, const double* sol)
    {
        for (size_t c = 0; c < tot/5; ++c)
        {
            const double* cc = sol + 5*c;
            double* pp = prim + 5*c;
            // convert_state(cons -> prim), navier-stokes/fluid_state.h:118-135
            const double rho = cc[0];
            const double invrho = 1.0/rho;
            const double u = invrho*cc[2], v = invrho*cc[3], w = invrho*cc[4];
            const double rhoU2 = rho*(u*u + v*v + w*w);
            const double p = (gas->gamma - 1.0)*(cc[1] - 0.5*rhoU2);
            const double T = p/(gas->R*rho);
            pp[0] = p; pp[1] = T; pp[2] = u; pp[3] = v; pp[4] = w;
        }
        flux_div(G, prim, r, *fl, *gas, 0);
    };
    auto a_val = [&](int i, int j) { return tb->a[i][j][1] == 0 ? 0.0 : ratio_value(tb->a[i][j][0], tb->a[i][j][1]); };
    for (int i = 0; i < S; ++i)
    {
        double* sol = tb->high_storage ? q2 : q;
        if (tb->high_storage) std::memcpy(q2, q, tot*sizeof(double));
        for (int j = 0; j < i; ++j)
        {
            const double c = a_val(i, j);
            if (c == 0.0) continue;
            scale(res[j], dt*c);
            incr(sol, res[j], false);
            scale(res[j], 1.0/(dt*c));
        }
        if (i > 0) exchange(G, sol);
        rhs(res[i], sol);
        if (!tb->high_storage)
        {
            for (int j = 0; j < i; ++j)
            {
                const double c = a_val(i, j);
                if (c == 0.0) continue;
                scale(res[j], dt*c);
                incr(sol, res[j], true);
                scale(res[j], 1.0/(dt*c));
            }
        }
    }
    for (int i = 0; i < S; ++i)
    {
        const double c = tb->b[i][1] == 0 ? 0.0 : ratio_value(tb->b[i][0], tb->b[i][1]);
        if (c == 0.0) continue;
        scale(res[i], c*dt);
        incr(q, res[i], false);
        scale(res[i], 1.0/(c*dt));
    }
    exchange(G, q);
    return 0;
}

// One low-storage step, ssprk3_opt (advance.h:285-402).
int ora_ssprk3opt_step(const ora_grid* g, const ora_gas* gas, const ora_flux* fl, const ora_bc* bc,
                       double dt, double* q, double* r0, double* r1)
{
    const Geo G(g);
    flux_div(G, q, r0, *fl, *gas, 0);
    for_interior(G, [&](int i, int j, int k, int lb)
    {
        double* e = q + G.cell(i, j, k, lb);
        const double* a = r0 + G.cell(i, j, k, lb);
        double c[5]; p2c(*gas, e, c);
        for (int v = 0; v < 5; ++v) c[v] += dt*a[v];
        c2p(*gas, c, e);
    });
    apply_bc(G, q, bc);
    flux_div(G, q, r1, *fl, *gas, 0);
    for_interior(G, [&](int i, int j, int k, int lb)
    {
        double* e = q + G.cell(i, j, k, lb);
        double* a = r0 + G.cell(i, j, k, lb);
        const double* b = r1 + G.cell(i, j, k, lb);
        double c[5]; p2c(*gas, e, c);
        double old[5], nr[5];
        for (int v = 0; v < 5; ++v) { old[v] = a[v]; nr[v] = (1.0/6.0)*dt*(a[v] + b[v]); a[v] = nr[v]; }
        for (int v = 0; v < 5; ++v) c[v] += (3.0/2.0)*nr[v];
        for (int v = 0; v < 5; ++v) c[v] -= dt*old[v];
        c2p(*gas, c, e);
    });
    apply_bc(G, q, bc);
    flux_div(G, q, r1, *fl, *gas, 0);
    for_interior(G, [&](int i, int j, int k, int lb)
    {
        double* e = q + G.cell(i, j, k, lb);
        const double* a = r0 + G.cell(i, j, k, lb);
        const double* b = r1 + G.cell(i, j, k, lb);
        double c[5]; p2c(*gas, e, c);
        for (int v = 0; v < 5; ++v) c[v] -= (1.0/2.0)*a[v];
        for (int v = 0; v < 5; ++v) c[v] += dt*(2.0/3.0)*b[v];
        c2p(*gas, c, e);
    });
    apply_bc(G, q, bc);
    return 0;
}

} // extern "C"
