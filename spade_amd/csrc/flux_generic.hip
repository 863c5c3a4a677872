// Generic flux_div: one thread per interior cell, six face evaluations per
// cell in the order of the reference's flux_div_basic
// (pde-algs/flux-div/flux_div_basic.h:47-72): for idir in 0..2, pm in 0..1,
// accum = F(face); accum *= jac*(1-2pm)*inv_dx; rhs += accum.
//
// This kernel serves every flux functor the library supports (including the
// 4-cell WENO stencils and the 4th-order KEEP stencil).  It is compiled without
// floating-point contraction and keeps the reference's expression order, so
// for identity coordinates its result equals the reference bit for bit except
// where a libm routine (pow, tanh) is involved.  The hot combination
// (KEEP-2 with pressure diffusion and viscous terms) has its own LDS-staged
// kernel in flux_fast.hip for identity coordinates, and elsewhere (mapped
// coordinates) a compile-time specialised form of this kernel, k_flux_div_spec.
#include "spd_internal.hpp"
#include "spd_stage.hpp"
#include <cmath>
#include <algorithm>

namespace spd {
namespace {

struct Prim { double p, T, u, v, w; };

__device__ inline double vel(const Prim& q, int d) { return d == 0 ? q.u : (d == 1 ? q.v : q.w); }

struct Cell
{
    const DevGeom* G;
    const double* __restrict__ B;   // block base (var 0, stored cell (0,0,0))
    int lb;
    __device__ inline double val(int v, int i, int j, int k) const
    {
        return B[(long long)v*G->sv + stored_offset(*G, i, j, k)];
    }
    __device__ inline Prim load(int i, int j, int k) const
    {
        return Prim{val(0, i, j, k), val(1, i, j, k), val(2, i, j, k), val(3, i, j, k), val(4, i, j, k)};
    }
};

// Face d at index (i,j,k): face i is the low face of cell i; the cell at
// half-cell offset o (odd) is i + floor(o/2) (omni/geom.h:87-145).
struct FaceIdx
{
    int d, i, j, k;
    __device__ inline void at(int o, int& ci, int& cj, int& ck) const
    {
        const int s = (o >= 0) ? o/2 : -((-o + 1)/2);
        ci = i; cj = j; ck = k;
        if (d == 0) ci += s; else if (d == 1) cj += s; else ck += s;
    }
};

__device__ inline void metric(const DevGeom& G, int d, const double xi[3], double n[3])
{
    if (G.identity) { n[0] = (d == 0); n[1] = (d == 1); n[2] = (d == 2); return; }
    const double m0 = map_deriv(G, 0, xi[0]), m1 = map_deriv(G, 1, xi[1]), m2 = map_deriv(G, 2, xi[2]);
    const double jac = 1.0/(m0*m1*m2);
    n[0] = 0.0; n[1] = 0.0; n[2] = 0.0;
    n[d] = 1.0;
    n[0] /= m0*jac; n[1] /= m1*jac; n[2] /= m2*jac;
}

__device__ inline void cell_xi(const DevGeom& G, int lb, int i, int j, int k, double xi[3])
{
    xi[0] = G.bmin[3*lb + 0] + (i + 0.5)*G.dx[3*lb + 0];
    xi[1] = G.bmin[3*lb + 1] + (j + 0.5)*G.dx[3*lb + 1];
    xi[2] = G.bmin[3*lb + 2] + (k + 0.5)*G.dx[3*lb + 2];
}

__device__ inline void face_xi(const DevGeom& G, int lb, const FaceIdx& f, double xi[3])
{
    double r[3] = {f.i + 0.5, f.j + 0.5, f.k + 0.5};
    r[f.d] -= 0.5;
    for (int d = 0; d < 3; ++d) xi[d] = G.bmin[3*lb + d] + r[d]*G.dx[3*lb + d];
}

__device__ inline void cell_metric(const Cell& C, const FaceIdx& f, int o, double n[3])
{
    int ci, cj, ck; f.at(o, ci, cj, ck);
    double xi[3]; cell_xi(*C.G, C.lb, ci, cj, ck, xi);
    metric(*C.G, f.d, xi, n);
}

__device__ inline void face_metric(const Cell& C, const FaceIdx& f, double n[3])
{
    double xi[3]; face_xi(*C.G, C.lb, f, xi);
    metric(*C.G, f.d, xi, n);
}

__device__ inline Prim cell_at(const Cell& C, const FaceIdx& f, int o)
{
    int ci, cj, ck; f.at(o, ci, cj, ck);
    return C.load(ci, cj, ck);
}

// Face gradient of the 5 primitives, component-major out[dir][var]: the
// four-point rule of omni/infos/info_gradient.h:45-81 in the same order.
__device__ inline void face_grad(const Cell& C, const FaceIdx& f, double out[3][5])
{
    const int dirs[3] = {f.d, (f.d + 1) % 3, (f.d + 2) % 3};
    const double* inv = C.G->inv_dx + 3*C.lb;
    for (int c = 0; c < 3; ++c) for (int v = 0; v < 5; ++v) out[c][v] = 0.0;
    int ic[3]; f.at(-1, ic[0], ic[1], ic[2]);
    auto apply = [&](int iset, double coeff)
    {
        for (int v = 0; v < 5; ++v) out[iset][v] += coeff*C.val(v, ic[0], ic[1], ic[2])*inv[iset];
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
    if (!C.G->identity)
    {
        double xi[3]; face_xi(*C.G, C.lb, f, xi);
        for (int c = 0; c < 3; ++c)
        {
            const double m = map_deriv(*C.G, c, xi[c]);
            for (int v = 0; v < 5; ++v) out[c][v] /= m;
        }
    }
}

__device__ inline double cfd_coeff(int idx, int order)
{
    double fh = 1.0, fp = 1.0, fm = 1.0;
    const int h = order/2;
    for (int q = 2; q <= h; ++q) fh *= q;
    for (int q = 2; q <= h + idx; ++q) fp *= q;
    for (int q = 2; q <= h - idx; ++q) fm *= q;
    const double d0 = fh/fp, d1 = fh/fm;
    const int sgn = ((1 + idx) % 2 == 0) ? 1 : -1;
    return d0*d1*sgn/double(idx);
}

// totani_lr (navier-stokes/convective.h:68-93)
__device__ __forceinline__ void f_totani(const Cell& C, const FaceIdx& f, double R, double gam, double out[5])
{
    const Prim ql = cell_at(C, f, -1), qr = cell_at(C, f, 1);
    double nl[3], nr[3]; cell_metric(C, f, -1, nl); cell_metric(C, f, 1, nr);
    const double un_l = nl[0]*ql.u + nl[1]*ql.v + nl[2]*ql.w;
    const double un_r = nr[0]*qr.u + nr[1]*qr.v + nr[2]*qr.w;
    const double rho_l = ql.p/(R*ql.T);
    const double rho_r = qr.p/(R*qr.T);
    const double e_l = ql.p/(rho_l*(gam - 1.0));
    const double e_r = qr.p/(rho_r*(gam - 1.0));
    const double c = 0.25*(rho_l + rho_r)*(un_l + un_r);
    out[0] = c;
    out[1] = 0.5*c*(e_l + e_r + ql.u*qr.u + ql.v*qr.v + ql.w*qr.w) + 0.5*(un_l*qr.p + un_r*ql.p);
    out[2] = 0.5*c*(ql.u + qr.u) + 0.5*(nl[0]*ql.p + nr[0]*qr.p);
    out[3] = 0.5*c*(ql.v + qr.v) + 0.5*(nl[1]*ql.p + nr[1]*qr.p);
    out[4] = 0.5*c*(ql.w + qr.w) + 0.5*(nl[2]*ql.p + nr[2]*qr.p);
}

// cent_keep_scheme_t<order> (convective.h:117-184)
__device__ __forceinline__ void f_keep(const Cell& C, const FaceIdx& f, int order, double R, double gam, double out[5])
{
    const int hw = order/2;
    double nface[3]; face_metric(C, f, nface);
    double un[4], rho[4], eng[4];
    Prim q[4];
    for (int ii = 0; ii < order; ++ii)
    {
        const int o = -(2*hw - 1) + 2*ii;
        q[ii] = cell_at(C, f, o);
        double n[3]; cell_metric(C, f, o, n);
        un[ii]  = q[ii].u*n[0] + q[ii].v*n[1] + q[ii].w*n[2];
        rho[ii] = q[ii].p/(R*q[ii].T);
        eng[ii] = q[ii].p/(rho[ii]*(gam - 1.0));
    }
    double c = 0.0, mx = 0.0, my = 0.0, mz = 0.0, g = 0.0, k = 0.0, ie = 0.0, pp = 0.0;
    for (int ii = 1; ii <= hw; ++ii)
    {
        const double a = cfd_coeff(ii, order);
        for (int jj = 0; jj < ii; ++jj)
        {
            const int i0 = hw - 1 - jj, i1 = i0 + ii;
            const Prim& q0 = q[i0]; const Prim& q1 = q[i1];
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

// ducros_t (state_sensor.h:32-42)
__device__ inline double ducros(const double gr[3][5], double eps)
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
__device__ __forceinline__ void f_pdiss(const Cell& C, const FaceIdx& f, const spd_flux_term& t, double R, double gam_, double out[5])
{
    const Prim ql = cell_at(C, f, -1), qr = cell_at(C, f, 1);
    const double delta_p = ql.p - qr.p;
    const double delta_T = ql.T - qr.T;
    const double pavg = 0.5*(ql.p + qr.p);
    const double Tavg = 0.5*(ql.T + qr.T);
    const double inv_RT = 1.0/(0.5*(R + R)*Tavg);
    const double p_inv_rt2 = pavg*inv_RT/Tavg;
    const double gamma = 0.5*(gam_ + gam_);
    const double k = 0.5*(ql.u*qr.u + ql.v*qr.v + ql.w*qr.w);
    const double u = 0.5*(ql.u + qr.u);
    const double v = 0.5*(ql.v + qr.v);
    const double w = 0.5*(ql.w + qr.w);
    double scl = t.sensor_param;
    if (t.sensor == SPD_SENSOR_DUCROS)
    {
        double gr[3][5]; face_grad(C, f, gr);
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

// rusanov_t (flux_funcs.h:19-52)
__device__ inline void split_rusanov(double R, double gam, const Prim& q, const double nv[3], double fu[5], double fd[5])
{
    const double u_n    = nv[0]*q.u + nv[1]*q.v + nv[2]*q.w;
    const double u_norm = sqrt(q.u*q.u + q.v*q.v + q.w*q.w);
    const double rho    = q.p/(R*q.T);
    const double pdiff  = q.p/rho;
    const double engy   = 0.5*(q.u*q.u + q.v*q.v + q.w*q.w) + q.p/(rho*(gam - 1.0));
    fu[0] = 0.5*rho*u_n;
    fu[1] = 0.5*rho*u_n*(engy + pdiff);
    fu[2] = 0.5*rho*q.u*u_n + 0.5*q.p*nv[0];
    fu[3] = 0.5*rho*q.v*u_n + 0.5*q.p*nv[1];
    fu[4] = 0.5*rho*q.w*u_n + 0.5*q.p*nv[2];
    for (int n = 0; n < 5; ++n) fd[n] = fu[n];
    const double sigma = 0.5*(u_norm + sqrt(gam*R*q.T));
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
__device__ inline void split_phys(double R, double gam, const Prim& q, const double nv[3], double fu[5], double fd[5])
{
    const double u_n  = nv[0]*q.u + nv[1]*q.v + nv[2]*q.w;
    const double rho  = q.p/(R*q.T);
    const double enth = 0.5*(q.u*q.u + q.v*q.v + q.w*q.w) + q.p/rho + q.p/(rho*(gam - 1.0));
    fu[0] = 0.5*rho*u_n;
    fu[1] = 0.5*rho*u_n*enth;
    fu[2] = 0.5*rho*q.u*u_n + 0.5*q.p*nv[0];
    fu[3] = 0.5*rho*q.v*u_n + 0.5*q.p*nv[1];
    fu[4] = 0.5*rho*q.w*u_n + 0.5*q.p*nv[2];
    for (int n = 0; n < 5; ++n) fd[n] = fu[n];
}

// weno_t (convective.h:275-332)
__device__ __forceinline__ void f_weno(const Cell& C, const FaceIdx& f, const spd_flux_term& t, double R, double gam, double out[5])
{
    double fu[4][5], fd[4][5];
    for (int c = 0; c < 4; ++c)
    {
        const int o = -3 + 2*c;
        const Prim q = cell_at(C, f, o);
        double n[3]; cell_metric(C, f, o, n);
        if (t.split == SPD_SPLIT_PHYS) split_phys(R, gam, q, n, fu[c], fd[c]);
        else                           split_rusanov(R, gam, q, n, fu[c], fd[c]);
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

// weno_fds_t<rusanov_fds_t> (convective.h:499-555, flux_funcs.h:54-95): the
// face states are reconstructed from cells i-2 .. i+1 component by component
// (enable_smooth keeps the central candidates, disable_smooth blends them
// with the one-sided ones at 2/3 : 1/3); the Rusanov flux difference then
// takes the face value and face metric: sigma = u_n + c at the face and
// F = sigma (U(qL) - U(qR)), U = (rho, rho e_tot, rho u, rho v, rho w).
__device__ __forceinline__ void f_weno_fds(const Cell& C, const FaceIdx& f, const spd_flux_term& t, double R,
                                           double gam, double out[5])
{
    const Prim a0 = cell_at(C, f, -3), a1 = cell_at(C, f, -1), a2 = cell_at(C, f, 1), a3 = cell_at(C, f, 3);
    const double q0[5] = {a0.p, a0.T, a0.u, a0.v, a0.w}, q1[5] = {a1.p, a1.T, a1.u, a1.v, a1.w};
    const double q2[5] = {a2.p, a2.T, a2.u, a2.v, a2.w}, q3[5] = {a3.p, a3.T, a3.u, a3.v, a3.w};
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
    // face value (omni/infos/info_value.h:151-163): (q(-1) + q(1)) * 0.5
    double qf[5];
    for (int n = 0; n < 5; ++n) { qf[n] = q1[n]; qf[n] += q2[n]; qf[n] *= 0.5; }
    double nv[3]; face_metric(C, f, nv);
    const double u_n   = nv[0]*qf[2] + nv[1]*qf[3] + nv[2]*qf[4];
    const double rhoL  = ql[0]/(R*ql[1]);
    const double engyL = 0.5*(ql[2]*ql[2] + ql[3]*ql[3] + ql[4]*ql[4]) + ql[0]/(rhoL*(gam - 1.0));
    const double rhoR  = qr[0]/(R*qr[1]);
    const double engyR = 0.5*(qr[2]*qr[2] + qr[3]*qr[3] + qr[4]*qr[4]) + qr[0]/(rhoR*(gam - 1.0));
    const double sigma = u_n + sqrt(gam*R*qf[1]);
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

// fweno_t (convective.h:355-496)
__device__ __forceinline__ void f_fweno(const Cell& C, const FaceIdx& f, const spd_flux_term& t, double R, double gam, double out[5])
{
    Prim q[4]; double nv[4][3];
    double rho[4], hsr[4], fluxv[4], dissv[4];
    for (int i = 0; i < 4; ++i)
    {
        const int o = -3 + 2*i;
        q[i] = cell_at(C, f, o);
        cell_metric(C, f, o, nv[i]);
        rho[i] = R*q[i].T;
        hsr[i] = q[i].u*q[i].u;
        hsr[i] += q[i].v*q[i].v;
        hsr[i] += q[i].w*q[i].w;
        hsr[i] = sqrt(hsr[i]);
        hsr[i] += sqrt(rho[i]*gam);
        rho[i] = q[i].p/rho[i];
        hsr[i] *= rho[i];
        hsr[i] *= 0.5;
    }
    auto apply_weno = [&](int k)
    {
        const double f0u = fluxv[0] + dissv[0];
        const double f1u = fluxv[1] + dissv[1], f1d = fluxv[1] - dissv[1];
        const double f2u = fluxv[2] + dissv[2], f2d = fluxv[2] - dissv[2];
        const double f3d = fluxv[3] - dissv[3];
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
        const double pdiff = R*q[i].T;
        engy += pdiff/(gam - 1.0);
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
            fluxv[i] *= vel(q[i], dr);
            fluxv[i] += q[i].p*nv[i][dr];
            fluxv[i] *= 0.5;
            dissv[i]  = hsr[i];
            dissv[i] *= vel(q[i], dr);
        }
        apply_weno(2 + dr);
    }
}

// first_order_t (convective.h:27-32)
__device__ __forceinline__ void f_first(const Cell& C, const FaceIdx& f, double R, double gam, double out[5])
{
    double fuL[5], fdL[5], fuR[5], fdR[5], nl[3], nr[3];
    cell_metric(C, f, -1, nl); cell_metric(C, f, 1, nr);
    split_rusanov(R, gam, cell_at(C, f, -1), nl, fuL, fdL);
    split_rusanov(R, gam, cell_at(C, f,  1), nr, fuR, fdR);
    for (int n = 0; n < 5; ++n) out[n] = fuL[n] + fdR[n];
}

// WALE subgrid viscosity at a face (subgrid_scale.h:44-84): with
// g_ij = d u_i / d x_j of the face gradient, S = (g + g^T)/2,
// Sd = (g^2 + (g^2)^T)/2 - tr(g^2)/3 I, |A|^2 = sum of squares (columns outer),
//   mu_t = rho cw^2 delta^2 |Sd|^2 |Sd| / (1e-8 + |S|^4 |S| + |Sd|^2 |Sd|^(1/2)),
// rho = p / (R T) of the face value; each sum in the reference's order
__device__ inline double wale_mu_t(const spd_flux_term& t, const Prim& qf, const double g[3][5], double R)
{
    const double rho = qf.p/(R*qf.T);
    double g2[3][3], sd[3][3], s[3][3];
    #pragma unroll
    for (int j = 0; j < 3; ++j)
        #pragma unroll
        for (int i = 0; i < 3; ++i)
            g2[i][j] = g[0][2 + i]*g[j][2] + g[1][2 + i]*g[j][3] + g[2][2 + i]*g[j][4];
    #pragma unroll
    for (int j = 0; j < 3; ++j)
        #pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            sd[i][j] = 0.5*(g2[i][j] + g2[j][i]);
            s[i][j]  = 0.5*(g[j][2 + i] + g[i][2 + j]);
        }
    double tr = 0.0;
    #pragma unroll
    for (int d = 0; d < 3; ++d) tr += 0.33333333333333333*g2[d][d];
    #pragma unroll
    for (int d = 0; d < 3; ++d) sd[d][d] -= tr;
    double ssd = 0.0, ss = 0.0;
    #pragma unroll
    for (int j = 0; j < 3; ++j)
        #pragma unroll
        for (int i = 0; i < 3; ++i) ssd += sd[i][j]*sd[i][j];
    #pragma unroll
    for (int j = 0; j < 3; ++j)
        #pragma unroll
        for (int i = 0; i < 3; ++i) ss += s[i][j]*s[i][j];
    const double r0 = sqrt(ssd), r1 = sqrt(r0), r2 = sqrt(ss);
    return rho*t.sgs_cw*t.sgs_cw*t.sgs_delta*t.sgs_delta*ssd*r0/(1e-8 + ss*ss*r2 + ssd*r1);
}

// visc_lr (viscous.h:39-112) with constant_viscosity_t or the power law
// (viscous_laws.h:62-135), optionally inside sgs_visc_t with the WALE model
// (viscous_laws.h:186-196)
__device__ __forceinline__ void f_visc(const Cell& C, const FaceIdx& f, const spd_flux_term& t, double R, double gam, double out[5])
{
    const Prim ql = cell_at(C, f, -1), qr = cell_at(C, f, 1);
    Prim qf;
    qf.p = (ql.p + qr.p)*0.5; qf.T = (ql.T + qr.T)*0.5;
    qf.u = (ql.u + qr.u)*0.5; qf.v = (ql.v + qr.v)*0.5; qf.w = (ql.w + qr.w)*0.5;
    double g[3][5]; face_grad(C, f, g);
    double n[3]; face_metric(C, f, n);
    double visc, visc2, alpha;
    if (t.law == SPD_LAW_POWER)
    {
        const double m = t.mu*pow(qf.T/t.T_ref, t.power);
        visc  = m;
        visc2 = -0.66666666667*m;
        alpha = m/t.prandtl;
    }
    else
    {
        visc  = t.mu;
        visc2 = double(-2.0*t.mu/3.0);
        alpha = t.mu*(1.0/t.prandtl);
    }
    if (t.sgs == SPD_SGS_WALE)
    {
        const double mu_t = wale_mu_t(t, qf, g, R);
        visc  = visc + mu_t;
        visc2 = visc2 - 0.66666666667*mu_t;
        alpha = alpha + mu_t/t.sgs_prt;
    }
    const double div = g[0][2] + g[1][3] + g[2][4];
    double tau[3][3];
    tau[0][0] = visc*(g[0][2] + g[0][2]) + visc2*div;
    tau[0][1] = visc*(g[0][3] + g[1][2]);
    tau[0][2] = visc*(g[0][4] + g[2][2]);
    tau[1][2] = visc*(g[1][4] + g[2][3]);
    tau[1][1] = visc*(g[1][3] + g[1][3]) + visc2*div;
    tau[1][0] = tau[0][1];
    tau[2][0] = tau[0][2];
    tau[2][1] = tau[1][2];
    tau[2][2] = visc*(g[2][4] + g[2][4]) + visc2*div;
    const double spec_heat = gam*R/(gam - 1.0);
    const double cond = spec_heat*alpha;
    double ht[3];
    for (int r = 0; r < 3; ++r) ht[r] = qf.u*tau[r][0] + qf.v*tau[r][1] + qf.w*tau[r][2] + cond*g[r][1];
    out[0] = 0.0;
    out[1] = -(n[0]*ht[0] + n[1]*ht[1] + n[2]*ht[2]);
    out[2] = -(n[0]*tau[0][0] + n[1]*tau[0][1] + n[2]*tau[0][2]);
    out[3] = -(n[0]*tau[1][0] + n[1]*tau[1][1] + n[2]*tau[1][2]);
    out[4] = -(n[0]*tau[2][0] + n[1]*tau[2][1] + n[2]*tau[2][2]);
}

// The run-time term dispatch of the generic kernel.  SPD_TERM_HYBRID has no
// device code: validate_flux refuses it before any launch.
__device__ __forceinline__ void eval_term(const Cell& C, const FaceIdx& f, const spd_flux_term& t, double R, double gam, double out[5])
{
    switch (t.kind)
    {
        case SPD_TERM_KEEP:   f_keep(C, f, t.order, R, gam, out); break;
        case SPD_TERM_TOTANI: f_totani(C, f, R, gam, out); break;
        case SPD_TERM_PDISS:  f_pdiss(C, f, t, R, gam, out); break;
        case SPD_TERM_WENO:   f_weno(C, f, t, R, gam, out); break;
        case SPD_TERM_FWENO:  f_fweno(C, f, t, R, gam, out); break;
        case SPD_TERM_FIRST:  f_first(C, f, R, gam, out); break;
        case SPD_TERM_VISC:   f_visc(C, f, t, R, gam, out); break;
        case SPD_TERM_WENO_FDS: f_weno_fds(C, f, t, R, gam, out); break;
        default: for (int n = 0; n < 5; ++n) out[n] = 0.0;
    }
}

// rhs of cell (i, j, k) of block lb in the reference's per-cell order
// (flux_div_basic.h:47-72)
__device__ inline void generic_cell(const DevGeom& G, const double* __restrict__ prims, double* __restrict__ rhs,
                                    const spd_flux_desc& F, int increment, int i, int j, int k, int lb)
{
    Cell C{&G, prims + (long long)lb*G.sb, lb};
    double jac = 1.0;
    if (!G.identity)
    {
        double xi[3]; cell_xi(G, lb, i, j, k, xi);
        jac = 1.0/(map_deriv(G, 0, xi[0])*map_deriv(G, 1, xi[1])*map_deriv(G, 2, xi[2]));
    }
    double relem[5];
    for (int v = 0; v < 5; ++v) relem[v] = increment ? rhs[cell_index(G, lb, v, i, j, k)] : 0.0;
    for (int idir = 0; idir < 3; ++idir)
    {
        const double inv_dx = 1.0/G.dx[3*lb + idir];
        for (int pm = 0; pm < 2; ++pm)
        {
            FaceIdx f{idir, i, j, k};
            if (pm == 1) { if (idir == 0) f.i++; else if (idir == 1) f.j++; else f.k++; }
            // compose: F0 + (F1 + (... + Fn))
            double acc[5];
            eval_term(C, f, F.terms[F.nterms - 1], F.R, F.gamma, acc);
            for (int m = F.nterms - 2; m >= 0; --m)
            {
                double z[5];
                eval_term(C, f, F.terms[m], F.R, F.gamma, z);
                for (int n = 0; n < 5; ++n) { z[n] += acc[n]; acc[n] = z[n]; }
            }
            const double s = jac*(1.0 - 2.0*pm)*inv_dx;
            for (int n = 0; n < 5; ++n) { acc[n] *= s; relem[n] += acc[n]; }
        }
    }
    for (int v = 0; v < 5; ++v) rhs[cell_index(G, lb, v, i, j, k)] = relem[v];
}

__global__ void __launch_bounds__(256) __attribute__((flatten))
k_flux_div_generic(DevGeom G, const double* __restrict__ prims, double* __restrict__ rhs, spd_flux_desc F, int increment)
{
    const int tiles_x = (G.nc[0] + 15)/16, tiles_y = (G.nc[1] + 15)/16;
    long long w = blockIdx.x;
    const int tx = (int)(w % tiles_x); w /= tiles_x;
    const int ty = (int)(w % tiles_y); w /= tiles_y;
    const int k  = (int)(w % G.nc[2]); w /= G.nc[2];
    const int lb = (int)w;
    const int i = tx*16 + (threadIdx.x & 15);
    const int j = ty*16 + (threadIdx.x >> 4);
    // geometry and flux descriptor in LDS: the per-cell code indexes them
    // with run-time values, which on kernel arguments would make every lane
    // copy them to private (scratch) memory first
    __shared__ DevGeom sG;
    __shared__ spd_flux_desc sF;
    if (threadIdx.x == 0) { sG = G; sF = F; }
    __syncthreads();
    if (i >= G.nc[0] || j >= G.nc[1]) return;
    generic_cell(sG, prims, rhs, sF, increment, i, j, k, lb);
}

// Completes blocks the LDS-staged kernel left at plane start[lb] < nz (a face
// temperature outside its pow series range): planes start[lb] .. nz-1 of the
// block, one workgroup per block.  Planes below start[lb] are complete and
// planes from start[lb] on were not written, so increment mode stays exact.
__global__ void __launch_bounds__(256) __attribute__((flatten))
k_flux_fixup(DevGeom G, const double* __restrict__ prims, double* __restrict__ rhs, spd_flux_desc F, int increment,
             const int* __restrict__ start)
{
    const int nz = G.nc[2];
    // a grid of at most 1024 workgroups walks the blocks.  The usual case is
    // that none of this workgroup's blocks needs completing: every thread reads
    // the same flags, so the whole workgroup leaves before staging anything
    bool any = false;
    for (int lb = blockIdx.x; lb < G.nblk; lb += gridDim.x) any |= fixup_start(start, lb) < nz;
    if (!any) return;
    __shared__ DevGeom sG;                // see k_flux_div_generic
    __shared__ spd_flux_desc sF;
    if (threadIdx.x == 0) { sG = G; sF = F; }
    __syncthreads();
    const int per_plane = G.nc[0]*G.nc[1];
    for (int lb = blockIdx.x; lb < G.nblk; lb += gridDim.x)
    {
        int s = fixup_start(start, lb);
        if (s >= nz) continue;
        if (s < 0) s = 0;
        const int n = per_plane*(nz - s);
        for (int c = threadIdx.x; c < n; c += blockDim.x)
        {
            const int i = c % G.nc[0], j = (c / G.nc[0]) % G.nc[1], k = s + c/per_plane;
            generic_cell(sG, prims, rhs, sF, increment, i, j, k, lb);
        }
    }
}

// The per-cell kernel specialised for one composition of up to three terms
// (cent_keep<2> / <4>, totani_lr, pressure_diss_lr, weno_t, fweno_t,
// first_order_t, visc_lr), on any geometry, in particular mapped coordinates,
// which the LDS-staged kernel does not take.  The terms are known at compile
// time and the kernel is flattened: every term and helper is inlined, so the
// small arrays the term functions fill (stencils, gradients, metrics) stay in
// registers instead of the per-lane stack that the run-time dispatch of
// k_flux_div_generic needs.  The arithmetic is the same code in the same order
// as generic_cell, so the two give the same bits.
constexpr int TERM_KEEP4 = 19;  // cent_keep<4>; SPD_TERM_KEEP stands for cent_keep<2> here

template <int K>
__device__ inline void eval_static(const Cell& C, const FaceIdx& f, const spd_flux_term& t, double R, double gam,
                                   double out[5])
{
    if constexpr (K == SPD_TERM_KEEP) f_keep(C, f, 2, R, gam, out);
    else if constexpr (K == TERM_KEEP4) f_keep(C, f, 4, R, gam, out);
    else if constexpr (K == SPD_TERM_TOTANI) f_totani(C, f, R, gam, out);
    else if constexpr (K == SPD_TERM_PDISS) f_pdiss(C, f, t, R, gam, out);
    else if constexpr (K == SPD_TERM_WENO) f_weno(C, f, t, R, gam, out);
    else if constexpr (K == SPD_TERM_FWENO) f_fweno(C, f, t, R, gam, out);
    else if constexpr (K == SPD_TERM_FIRST) f_first(C, f, R, gam, out);
    else if constexpr (K == SPD_TERM_WENO_FDS) f_weno_fds(C, f, t, R, gam, out);
    else f_visc(C, f, t, R, gam, out);
}

template <int K>
__device__ inline void add_static(const Cell& C, const FaceIdx& f, const spd_flux_term& t, double R, double gam,
                                  double acc[5])
{
    double z[5];
    eval_static<K>(C, f, t, R, gam, z);
    for (int n = 0; n < 5; ++n) { z[n] += acc[n]; acc[n] = z[n]; }
}

// F0 + (F1 + F2): the last term first, then each earlier one plus the sum
template <int T0, int T1, int T2>
__device__ inline void spec_face(const Cell& C, const FaceIdx& f, const spd_flux_desc& F, double acc[5])
{
    if constexpr (T2 != 0)
    {
        eval_static<T2>(C, f, F.terms[2], F.R, F.gamma, acc);
        add_static<T1>(C, f, F.terms[1], F.R, F.gamma, acc);
        add_static<T0>(C, f, F.terms[0], F.R, F.gamma, acc);
    }
    else if constexpr (T1 != 0)
    {
        eval_static<T1>(C, f, F.terms[1], F.R, F.gamma, acc);
        add_static<T0>(C, f, F.terms[0], F.R, F.gamma, acc);
    }
    else eval_static<T0>(C, f, F.terms[0], F.R, F.gamma, acc);
}

template <int T0, int T1, int T2>
__device__ inline void spec_cell(const DevGeom& G, const double* __restrict__ prims, double* __restrict__ rhs,
                                 const spd_flux_desc& F, int increment, int i, int j, int k, int lb)
{
    Cell C{&G, prims + (long long)lb*G.sb, lb};
    double jac = 1.0;
    if (!G.identity)
    {
        double xi[3]; cell_xi(G, lb, i, j, k, xi);
        jac = 1.0/(map_deriv(G, 0, xi[0])*map_deriv(G, 1, xi[1])*map_deriv(G, 2, xi[2]));
    }
    double relem[5];
    for (int v = 0; v < 5; ++v) relem[v] = increment ? rhs[cell_index(G, lb, v, i, j, k)] : 0.0;
    for (int idir = 0; idir < 3; ++idir)
    {
        const double inv_dx = 1.0/G.dx[3*lb + idir];
        for (int pm = 0; pm < 2; ++pm)
        {
            FaceIdx f{idir, i, j, k};
            if (pm == 1) { if (idir == 0) f.i++; else if (idir == 1) f.j++; else f.k++; }
            double acc[5];
            spec_face<T0, T1, T2>(C, f, F, acc);
            const double s = jac*(1.0 - 2.0*pm)*inv_dx;
            for (int n = 0; n < 5; ++n) { acc[n] *= s; relem[n] += acc[n]; }
        }
    }
    for (int v = 0; v < 5; ++v) rhs[cell_index(G, lb, v, i, j, k)] = relem[v];
}

template <int T0, int T1, int T2>
__global__ void __launch_bounds__(256) __attribute__((flatten))
k_flux_div_spec(DevGeom G, const double* __restrict__ prims, double* __restrict__ rhs, spd_flux_desc F, int increment)
{
    // same cell mapping as k_flux_div_generic
    const int tiles_x = (G.nc[0] + 15)/16, tiles_y = (G.nc[1] + 15)/16;
    long long w = blockIdx.x;
    const int tx = (int)(w % tiles_x); w /= tiles_x;
    const int ty = (int)(w % tiles_y); w /= tiles_y;
    const int k  = (int)(w % G.nc[2]); w /= G.nc[2];
    const int lb = (int)w;
    const int i = tx*16 + (threadIdx.x & 15);
    const int j = ty*16 + (threadIdx.x >> 4);
    __shared__ DevGeom sG;                // see k_flux_div_generic
    __shared__ spd_flux_desc sF;
    if (threadIdx.x == 0) { sG = G; sF = F; }
    __syncthreads();
    if (i >= G.nc[0] || j >= G.nc[1]) return;
    spec_cell<T0, T1, T2>(sG, prims, rhs, sF, increment, i, j, k, lb);
}

// k_flux_fixup with the specialised per-cell code (the compositions the
// LDS-staged kernel takes all have one)
template <int T0, int T1, int T2>
__global__ void __launch_bounds__(256) __attribute__((flatten))
k_flux_fixup_spec(DevGeom G, const double* __restrict__ prims, double* __restrict__ rhs, spd_flux_desc F,
                  int increment, const int* __restrict__ start)
{
    const int nz = G.nc[2];
    bool any = false;
    for (int lb = blockIdx.x; lb < G.nblk; lb += gridDim.x) any |= fixup_start(start, lb) < nz;
    if (!any) return;
    __shared__ DevGeom sG;
    __shared__ spd_flux_desc sF;
    if (threadIdx.x == 0) { sG = G; sF = F; }
    __syncthreads();
    const int per_plane = G.nc[0]*G.nc[1];
    for (int lb = blockIdx.x; lb < G.nblk; lb += gridDim.x)
    {
        int s = fixup_start(start, lb);
        if (s >= nz) continue;
        if (s < 0) s = 0;
        const int n = per_plane*(nz - s);
        for (int c = threadIdx.x; c < n; c += blockDim.x)
        {
            const int i = c % G.nc[0], j = (c / G.nc[0]) % G.nc[1], k = s + c/per_plane;
            spec_cell<T0, T1, T2>(sG, prims, rhs, sF, increment, i, j, k, lb);
        }
    }
}

// The fix-up of a fused stage in one launch, one workgroup per block (the
// blocks the LDS-staged kernel completed exit at once).  For a block left
// from plane s on: with pull, its ghost cells as the exchange writes them,
// then (wall_y) its isothermal wall slabs from those cells; the residual of
// planes s.. by the composition's per-cell code (k_flux_fixup_spec); then
// the stage of those planes, qin -> qout, with the ghosts they feed
// (k_stage_fixup).  Each phase reads only what this workgroup wrote before
// the barrier between them, so the result equals the three launches'.
template <int T0, int T1, int T2>
__global__ void __launch_bounds__(256) __attribute__((flatten))
k_stage_fixup_spec(DevGeom G, double* __restrict__ qin, double* __restrict__ qout, double* __restrict__ rhs,
                   spd_flux_desc F, Coefs C, double gamma, double R, const int* __restrict__ push,
                   const int* __restrict__ nbr, const int* __restrict__ start, int pull, int wall_y, double t_wall)
{
    const int lb = blockIdx.x, nz = G.nc[2];
    int s = fixup_start(start, lb);
    if (s >= nz) return;
    if (s < 0) s = 0;
    __shared__ DevGeom sG;
    __shared__ spd_flux_desc sF;
    __shared__ XStage xs;
    if (threadIdx.x == 0) { sG = G; sF = F; }
    if (pull)
    {
        const int items = exchange_items(G);
        for (int t = threadIdx.x; t < items; t += blockDim.x) exchange_item(G, nbr, qin, lb, t);
        const int by = (lb / G.nb[0]) % G.nb[1];
        const bool low = by == 0, high = by == G.nb[1] - 1;
        if (wall_y && (low || high))
        {
            __syncthreads();
            const int n = G.N[0]*G.ng*G.N[2];
            for (int pm = 0; pm < 2; ++pm)
            {
                if (pm == 0 ? !low : !high) continue;
                for (int c = threadIdx.x; c < n; c += blockDim.x) wall_cell(G, qin, lb, 1, pm, c, t_wall);
            }
        }
    }
    __syncthreads();
    const int per_plane = G.nc[0]*G.nc[1];
    const int n = per_plane*(nz - s);
    for (int c = threadIdx.x; c < n; c += blockDim.x)
    {
        const int i = c % G.nc[0], j = (c / G.nc[0]) % G.nc[1], k = s + c/per_plane;
        spec_cell<T0, T1, T2>(sG, qin, rhs, sF, 0, i, j, k, lb);
    }
    __syncthreads();
    for (int k = s; k < nz; ++k)
    {
        const CellIt ci{(int)(threadIdx.x & 15), (int)(threadIdx.x >> 4), k, lb};
        double qq[5], cons[5];
        long long idx[5];
        for (int v = 0; v < 5; ++v) { idx[v] = core_index(G, lb, v, ci.i, ci.j, k); qq[v] = qin[idx[v]]; }
        p2c(gamma, R, qq, cons);
        #pragma unroll
        for (int t = 0; t < kMaxStages; ++t)
        {
            if (t >= C.n || C.c[t] == 0.0) continue;
            double inc[5];
            for (int v = 0; v < 5; ++v) inc[v] = C.c[t]*C.r[t][idx[v]];
            for (int v = 0; v < 5; ++v) cons[v] += inc[v];
        }
        c2p(gamma, R, cons, qq);
        for (int v = 0; v < 5; ++v) qout[idx[v]] = qq[v];
        push_ghosts(G, push, qout, ci, qq, true);
        stage_x(G, xs, ci, qq);
        lds_barrier();
        flush_x(G, push, qout, xs, lb, k);
        lds_barrier();
    }
}

} // namespace

// Compositions with a specialised kernel, as term codes T0 T1 T2 (0: none);
// the first list is the LDS-staged kernel's set (its fix-up launches)
#define SPD_SPEC_FAST_LIST(X) \
    X(SPD_TERM_KEEP, 0, 0) X(SPD_TERM_PDISS, 0, 0) X(SPD_TERM_KEEP, SPD_TERM_PDISS, 0) \
    X(SPD_TERM_VISC, 0, 0) X(SPD_TERM_KEEP, SPD_TERM_VISC, 0) X(SPD_TERM_PDISS, SPD_TERM_VISC, 0) \
    X(SPD_TERM_KEEP, SPD_TERM_PDISS, SPD_TERM_VISC)
#define SPD_SPEC_LIST(X) SPD_SPEC_FAST_LIST(X) \
    X(TERM_KEEP4, 0, 0) X(SPD_TERM_TOTANI, 0, 0) X(SPD_TERM_WENO, 0, 0) X(SPD_TERM_FWENO, 0, 0) \
    X(SPD_TERM_FIRST, 0, 0) X(TERM_KEEP4, SPD_TERM_VISC, 0) X(SPD_TERM_TOTANI, SPD_TERM_VISC, 0) \
    X(SPD_TERM_WENO, SPD_TERM_VISC, 0) X(SPD_TERM_FWENO, SPD_TERM_VISC, 0) X(SPD_TERM_FIRST, SPD_TERM_VISC, 0) \
    X(SPD_TERM_WENO_FDS, 0, 0) X(SPD_TERM_WENO_FDS, SPD_TERM_VISC, 0)

bool flux_has_hybrid(const spd_flux_desc& f)
{
    for (int m = 0; m < f.nterms && m < SPD_MAX_TERMS; ++m) if (f.terms[m].kind == SPD_TERM_HYBRID) return true;
    return false;
}

int flux_div_spec_code(const spd_flux_desc& f)
{
    // SPD_DISABLE_SPEC=1 sends every composition to the generic kernel (for
    // timing comparisons)
    if (const char* e = getenv("SPD_DISABLE_SPEC")) if (e[0] == '1') return 0;
    if (f.nterms < 1 || f.nterms > 3) return 0;
    int code = 0;
    for (int m = 0; m < f.nterms; ++m)
    {
        const spd_flux_term& t = f.terms[m];
        int k = t.kind;
        if (k == SPD_TERM_KEEP) { if (t.order == 4) k = TERM_KEEP4; else if (t.order != 2) return 0; }
        else if ((k < SPD_TERM_TOTANI || k > SPD_TERM_VISC) && k != SPD_TERM_WENO_FDS) return 0;
        code = code*100 + k;
    }
    for (int m = f.nterms; m < 3; ++m) code *= 100;
#define SPD_SPEC_MATCH(a, b, c) if (code == (a)*10000 + (b)*100 + (c)) return code;
    SPD_SPEC_LIST(SPD_SPEC_MATCH)
#undef SPD_SPEC_MATCH
    return 0;
}

void launch_flux_div_generic(const Grid& g, const double* prims, double* rhs, const spd_flux_desc& f,
                             bool increment, hipStream_t s)
{
    const DevGeom& G = g.geom;
    const long long nwg = (long long)((G.nc[0] + 15)/16)*((G.nc[1] + 15)/16)*G.nc[2]*G.nblk;
    const int inc = increment ? 1 : 0;
    switch (flux_div_spec_code(f))
    {
#define SPD_SPEC_LAUNCH(a, b, c) \
        case (a)*10000 + (b)*100 + (c): \
            hipLaunchKernelGGL((k_flux_div_spec<a, b, c>), dim3((unsigned)nwg), dim3(256), 0, s, G, prims, rhs, f, inc); \
            break;
        SPD_SPEC_LIST(SPD_SPEC_LAUNCH)
#undef SPD_SPEC_LAUNCH
        default:
            hipLaunchKernelGGL(k_flux_div_generic, dim3((unsigned)nwg), dim3(256), 0, s, G, prims, rhs, f, inc);
    }
    SPD_LAUNCH_CHECK();
}

void launch_flux_div_fixup(const Grid& g, const double* prims, double* rhs, const spd_flux_desc& f, bool increment,
                           hipStream_t s)
{
    const dim3 grid((unsigned)std::min(g.geom.nblk, 1024));
    const int inc = increment ? 1 : 0;
    switch (flux_div_spec_code(f))
    {
#define SPD_FIXUP_LAUNCH(a, b, c) \
        case (a)*10000 + (b)*100 + (c): \
            hipLaunchKernelGGL((k_flux_fixup_spec<a, b, c>), grid, dim3(256), 0, s, g.geom, prims, rhs, f, inc, \
                               g.d_start); \
            break;
        SPD_SPEC_FAST_LIST(SPD_FIXUP_LAUNCH)
#undef SPD_FIXUP_LAUNCH
        default:
            hipLaunchKernelGGL(k_flux_fixup, grid, dim3(256), 0, s, g.geom, prims, rhs, f, inc, g.d_start);
    }
    SPD_LAUNCH_CHECK();
}

bool launch_stage_fixup_spec(const Grid& g, double* qin, double* qout, double* rhs, const spd_flux_desc& f,
                             const double* const* r, const double* coef, int nr, double gamma, double R,
                             bool pull, int wall_y, double t_wall, hipStream_t s)
{
    if (nr < 1 || nr > kMaxStages) throw ArgError{"stage fix-up: 1..SPD_RK_MAX_STAGES coefficients"};
    Coefs C{};
    C.n = nr;
    for (int j = 0; j < nr; ++j) { C.r[j] = r[j]; C.c[j] = coef[j]; }
    const dim3 grid((unsigned)g.geom.nblk);
    const int p = pull ? 1 : 0;
    switch (flux_div_spec_code(f))
    {
#define SPD_STAGE_FIXUP_LAUNCH(a, b, c) \
        case (a)*10000 + (b)*100 + (c): \
            hipLaunchKernelGGL((k_stage_fixup_spec<a, b, c>), grid, dim3(256), 0, s, g.geom, qin, qout, rhs, f, C, \
                               gamma, R, g.d_push, g.d_nbr, g.d_start, p, wall_y, t_wall); \
            break;
        SPD_SPEC_FAST_LIST(SPD_STAGE_FIXUP_LAUNCH)
#undef SPD_STAGE_FIXUP_LAUNCH
        default:
            return false;
    }
    SPD_LAUNCH_CHECK();
    return true;
}

} // namespace spd
