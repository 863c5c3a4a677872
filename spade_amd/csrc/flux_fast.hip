// LDS-staged flux_div for the hot combinations on gfx950:
//   cent_keep<2> [+ pressure_diss_lr (const | Ducros)] [+ visc_lr (const | power law)]
// on identity coordinates and blocks with a 16x16 (i,j) cross-section.
//
// One workgroup (256 threads, thread (i,j)) owns one block and marches over k.
// Every face flux is evaluated once: thread (i,j) evaluates the low x-face
// i-1/2, the low y-face j-1/2 and the high z-face k+1/2 of its cell; the 16
// block-boundary x-faces and y-faces of a plane are evaluated by 16 lanes of
// waves 0 and 1.  Face products P = F * (1/dx) are exchanged through LDS
// (y-faces, boundary x-faces) and ds_bpermute (x-faces); the z-face product
// stays in registers for the next plane.
//
// LDS holds, for the three planes k-1, k, k+1, the velocities (u,v,w) on the
// 18x18 plane with one halo layer (ring of three slots), and for plane k the
// p, T, rho, e needed as left/right states of the in-plane faces.
//
// Arithmetic follows the reference's expression order exactly
// (navier-stokes/convective.h:117-184, 212-245; viscous.h:39-112;
// omni/infos/info_gradient.h:45-81; pde-algs/flux-div/flux_div_basic.h:47-72)
// and the file is compiled without contraction, so the result equals the
// reference bit for bit except through pow() of the power-law viscosity
// (a binomial series within one ulp of x^n, see pow_series.hpp).
// Per-cell quantities the reference recomputes for every face (rho = p/(R T),
// e = p/(rho (gamma-1))) are computed once per cell: same operands, same
// rounding.
#include "spd_internal.hpp"
#include "pow_series.hpp"

namespace spd {
namespace {

constexpr int HX = 18, HP = 18*18;

// planes of primitives in flight ahead of the one entering the LDS window:
// enough bytes in flight per CU to cover HBM latency at full bandwidth
#ifndef SPD_FAST_PREFETCH
#define SPD_FAST_PREFETCH 2
#endif
constexpr int PF = SPD_FAST_PREFETCH;

struct FastCfg
{
    double R, gm1;                  // gas constant, gamma - 1
    double spec_heat;               // gamma R / (gamma - 1)
    double half_RR, inv_gamma;      // 0.5 (R + R), 1 / (0.5 (gamma + gamma))
    // viscous law, pre-multiplied by the composition scale s (8 with cent_keep<2>, else 1)
    double mu_s, visc2_s, cond_s;   // constant law: s mu, s beta, s cp alpha
    double mu_ref_s, T_ref, inv_T_ref, power, prandtl, inv_prandtl;
    int    t_ref_one;               // T_ref == 1: T/T_ref == T exactly
    PowSeries ps;                   // (T/T_ref)^power near T = T_ref
    // pressure diffusion, sensor pre-multiplied by s
    double scl_s, ducros_eps, ducros_s, eps_p, eps_T;
};

struct CellV { double p, T, u, v, w, rho, e; };

template <int D> __device__ inline double vel(const CellV& c) { return D == 0 ? c.u : (D == 1 ? c.v : c.w); }

__device__ inline CellV make_cell(const FastCfg& C, double p, double T, double u, double v, double w)
{
    CellV c{p, T, u, v, w, 0.0, 0.0};
    c.rho = p/(C.R*T);
    c.e   = p/(c.rho*C.gm1);
    return c;
}

// a / b, correctly rounded, for b a constant with y = RN(1/b) (Markstein's
// correction step: q0 = a y, r = a - b q0 exactly, q = q0 + r y)
__device__ inline double div_const(double a, double b, double y)
{
    const double q0 = a*y;
    const double r = __builtin_fma(-q0, b, a);
    return __builtin_fma(r, y, q0);
}

// library pow, kept out of line: it only runs for T/T_ref far from 1
__device__ __attribute__((noinline)) double pow_far(double x, double n) { return pow(x, n); }

// x^n of the power law: binomial series near 1, library pow otherwise
__device__ inline double pow_law(const FastCfg& C, double x)
{
    if (C.ps.in_range(x)) return C.ps.eval(x);
    return pow_far(x, C.power);
}

// ducros_t::get_sensor (state_sensor.h:32-42); g[c][q], q = u, v, w
__device__ inline double ducros(const double g[3][3], double eps)
{
    double theta_sq = g[0][0] + g[1][1] + g[2][2];
    theta_sq *= theta_sq;
    const double v0 = g[1][2] - g[2][1];
    const double v1 = g[2][0] - g[0][2];
    const double v2 = g[0][1] - g[1][0];
    const double vort_sq = v0*v0 + v1*v1 + v2*v2;
    return theta_sq/(theta_sq + vort_sq + eps);
}

// Sums shared by the terms of one face: SU = uL+uR (= 2 u_face) etc.
struct FaceSums { double SU, SV, SW, SP, ST, K; };

__device__ inline FaceSums face_sums(const CellV& L, const CellV& R)
{
    FaceSums f;
    f.SU = L.u + R.u; f.SV = L.v + R.v; f.SW = L.w + R.w;
    f.SP = L.p + R.p; f.ST = L.T + R.T;
    f.K  = L.u*R.u + L.v*R.v + L.w*R.w;
    return f;
}

// pressure_diss_lr (convective.h:212-245), times the composition scale s:
// every output is a product chain that starts with the sensor value, so a
// sensor pre-multiplied by s (a power of two) scales every output exactly.
template <int SENSOR>
__device__ inline void pdiss_s(const FastCfg& C, const CellV& L, const CellV& R, const FaceSums& f,
                               const double g[3][3], double F[5])
{
    const double delta_p = L.p - R.p;
    const double delta_T = L.T - R.T;
    const double pavg = 0.5*f.SP;
    const double Tavg = 0.5*f.ST;
    const double inv_RT = 1.0/(C.half_RR*Tavg);
    const double p_inv_rt2 = pavg*inv_RT/Tavg;
    const double k = 0.5*f.K;
    const double u = 0.5*f.SU, v = 0.5*f.SV, w = 0.5*f.SW;
    const double scl = (SENSOR == SPD_SENSOR_DUCROS) ? C.ducros_s*ducros(g, C.ducros_eps) : C.scl_s;
    const double a = scl*C.eps_p*delta_p;
    const double b = scl*C.eps_T*delta_T*p_inv_rt2;
    F[0] = a*inv_RT - b;
    F[1] = a*(k*inv_RT + C.inv_gamma) - b*(0.5*k);
    F[2] = a*u*inv_RT - b*u;
    F[3] = a*v*inv_RT - b*v;
    F[4] = a*w*inv_RT - b*w;
}

// visc_lr row D for the unit normal of identity coordinates (viscous.h:39-112),
// times the composition scale s (carried by mu, beta and the conductivity).
template <int D, int LAW>
__device__ inline void visc_s(const FastCfg& C, const FaceSums& f, const double g[3][3], double gT, double F[5])
{
    double mu, visc2, cond;
    if (LAW == SPD_LAW_POWER)
    {
        const double Tf = f.ST*0.5;
        const double x = C.t_ref_one ? Tf : div_const(Tf, C.T_ref, C.inv_T_ref);
        mu    = C.mu_ref_s*pow_law(C, x);
        visc2 = -0.66666666667*mu;
        cond  = C.spec_heat*div_const(mu, C.prandtl, C.inv_prandtl);
    }
    else
    {
        mu = C.mu_s; visc2 = C.visc2_s; cond = C.cond_s;
    }
    const double div = g[0][0] + g[1][1] + g[2][2];
    double t[3];
    #pragma unroll
    for (int b = 0; b < 3; ++b)
    {
        if (b == D) t[b] = mu*(g[D][D] + g[D][D]) + visc2*div;
        else if (D < b) t[b] = mu*(g[D][b] + g[b][D]);
        else            t[b] = mu*(g[b][D] + g[D][b]);
    }
    // (uf t0 + vf t1 + wf t2) with uf = SU/2: the halving is exact
    const double hsum = 0.5*(f.SU*t[0] + f.SV*t[1] + f.SW*t[2]);
    const double ht = hsum + cond*gT;
    F[0] = 0.0;
    F[1] = -ht;
    F[2] = -t[0];
    F[3] = -t[1];
    F[4] = -t[2];
}

template <int KEEP, int PDISS, int VISC> struct Spec
{
    static constexpr bool keep = KEEP != 0;
    static constexpr bool pd = PDISS != 0;           // 1 + sensor kind
    static constexpr int sensor = PDISS - 1;
    static constexpr bool vs = VISC != 0;             // 1 + law
    static constexpr int law = VISC - 1;
    static constexpr bool full_grad = pd && sensor == SPD_SENSOR_DUCROS;
};

// Face product P = F / dx with F = keep + (pdiss + visc) composed in the
// reference's order (omni/compose.h:19-24) and flux_div_basic's scaling
// (flux_div_basic.h:66-68).  cent_keep<2> (convective.h:117-184) is written
// with its power-of-two factors pulled out: with CX = (rhoL+rhoR)(unR+unL),
// continuity = CX/4 and the other outputs are 1/8 of
//   energy:   CX K + CX (eL+eR) + 4 (unL pR + unR pL)
//   momentum: CX (qL+qR) + 4 (pL+pR) n
// Every factor is a power of two, so each rounding equals the reference's;
// the other terms are evaluated pre-multiplied by 8 and the factor is folded
// into 1/dx.
template <int D, class S>
__device__ inline void face_product(const FastCfg& C, const CellV& L, const CellV& R, const double g[3][3], double gT,
                                    double idx, double P[5])
{
    const FaceSums f = face_sums(L, R);
    double s8[5];
    if (S::vs) visc_s<D, S::law>(C, f, g, gT, s8);
    if (S::pd)
    {
        double z[5];
        pdiss_s<S::sensor>(C, L, R, f, g, z);
        if (S::vs) { for (int n = 0; n < 5; ++n) s8[n] = z[n] + s8[n]; }
        else       { for (int n = 0; n < 5; ++n) s8[n] = z[n]; }
    }
    if (S::keep)
    {
        const double unL = vel<D>(L), unR = vel<D>(R);
        const double CX = (L.rho + R.rho)*(unR + unL);
        const double Ep = CX*f.K + CX*(L.e + R.e) + 4.0*(unL*R.p + unR*L.p);
        double M[3] = {CX*f.SU, CX*f.SV, CX*f.SW};
        M[D] = M[D] + 4.0*f.SP;
        const double i4 = 0.25*idx, i8 = 0.125*idx;
        if (S::vs || S::pd)
        {
            P[0] = S::pd ? (CX + 0.5*s8[0])*i4 : CX*i4;
            P[1] = (Ep + s8[1])*i8;
            P[2] = (M[0] + s8[2])*i8;
            P[3] = (M[1] + s8[3])*i8;
            P[4] = (M[2] + s8[4])*i8;
        }
        else
        {
            P[0] = CX*i4; P[1] = Ep*i8; P[2] = M[0]*i8; P[3] = M[1]*i8; P[4] = M[2]*i8;
        }
    }
    else
    {
        #pragma unroll
        for (int n = 0; n < 5; ++n) P[n] = s8[n]*idx;
    }
}

// 0.25-weighted four-point tangential gradient (info_gradient.h:63-79):
// the reference sums r(0.25 q it) terms; 0.25 factors out exactly.
__device__ inline double tang(double a, double b, double c, double d, double it)
{
    return 0.25*(((a*it - b*it) + c*it) - d*it);
}

// normal gradient (info_gradient.h:63, 73): -qL/dx + qR/dx
__device__ inline double normal_grad(double qL, double qR, double id)
{
    return qR*id - qL*id;
}

// slot s of the velocity ring: s[pos] = u, s[HP + pos] = v, s[2*HP + pos] = w

template <class S>
__device__ inline void grads_x(const double* sm, const double* s0, const double* sp, int pL, int pR,
                               const CellV& L, const CellV& R, double ix, double iy, double iz,
                               double g[3][3], double& gT)
{
    g[0][0] = normal_grad(L.u, R.u, ix);
    g[0][1] = normal_grad(L.v, R.v, ix);
    g[0][2] = normal_grad(L.w, R.w, ix);
    gT      = normal_grad(L.T, R.T, ix);
    g[1][0] = tang(s0[pL + HX], s0[pL - HX], s0[pR + HX], s0[pR - HX], iy);
    g[1][1] = tang(s0[HP + pL + HX], s0[HP + pL - HX], s0[HP + pR + HX], s0[HP + pR - HX], iy);
    g[1][2] = S::full_grad ? tang(s0[2*HP + pL + HX], s0[2*HP + pL - HX], s0[2*HP + pR + HX], s0[2*HP + pR - HX], iy) : 0.0;
    g[2][0] = tang(sp[pL], sm[pL], sp[pR], sm[pR], iz);
    g[2][1] = S::full_grad ? tang(sp[HP + pL], sm[HP + pL], sp[HP + pR], sm[HP + pR], iz) : 0.0;
    g[2][2] = tang(sp[2*HP + pL], sm[2*HP + pL], sp[2*HP + pR], sm[2*HP + pR], iz);
}

template <class S>
__device__ inline void grads_y(const double* sm, const double* s0, const double* sp, int pL, int pR,
                               const CellV& L, const CellV& R, double ix, double iy, double iz,
                               double g[3][3], double& gT)
{
    g[1][0] = normal_grad(L.u, R.u, iy);
    g[1][1] = normal_grad(L.v, R.v, iy);
    g[1][2] = normal_grad(L.w, R.w, iy);
    gT      = normal_grad(L.T, R.T, iy);
    g[2][0] = S::full_grad ? tang(sp[pL], sm[pL], sp[pR], sm[pR], iz) : 0.0;
    g[2][1] = tang(sp[HP + pL], sm[HP + pL], sp[HP + pR], sm[HP + pR], iz);
    g[2][2] = tang(sp[2*HP + pL], sm[2*HP + pL], sp[2*HP + pR], sm[2*HP + pR], iz);
    g[0][0] = tang(s0[pL + 1], s0[pL - 1], s0[pR + 1], s0[pR - 1], ix);
    g[0][1] = tang(s0[HP + pL + 1], s0[HP + pL - 1], s0[HP + pR + 1], s0[HP + pR - 1], ix);
    g[0][2] = S::full_grad ? tang(s0[2*HP + pL + 1], s0[2*HP + pL - 1], s0[2*HP + pR + 1], s0[2*HP + pR - 1], ix) : 0.0;
}

template <class S>
__device__ inline void grads_z(const double* sL, const double* sR, int p,
                               const CellV& L, const CellV& R, double ix, double iy, double iz,
                               double g[3][3], double& gT)
{
    g[2][0] = normal_grad(L.u, R.u, iz);
    g[2][1] = normal_grad(L.v, R.v, iz);
    g[2][2] = normal_grad(L.w, R.w, iz);
    gT      = normal_grad(L.T, R.T, iz);
    g[0][0] = tang(sL[p + 1], sL[p - 1], sR[p + 1], sR[p - 1], ix);
    g[0][1] = S::full_grad ? tang(sL[HP + p + 1], sL[HP + p - 1], sR[HP + p + 1], sR[HP + p - 1], ix) : 0.0;
    g[0][2] = tang(sL[2*HP + p + 1], sL[2*HP + p - 1], sR[2*HP + p + 1], sR[2*HP + p - 1], ix);
    g[1][0] = S::full_grad ? tang(sL[p + HX], sL[p - HX], sR[p + HX], sR[p - HX], iy) : 0.0;
    g[1][1] = tang(sL[HP + p + HX], sL[HP + p - HX], sR[HP + p + HX], sR[HP + p - HX], iy);
    g[1][2] = tang(sL[2*HP + p + HX], sL[2*HP + p - HX], sR[2*HP + p + HX], sR[2*HP + p - HX], iy);
}

__device__ inline CellV cell_from_lds(const double* sS, const double* s0, int p)
{
    // sS: [p, T, rho, e][HP] of plane k; s0: ring slot of plane k
    return CellV{sS[p], sS[HP + p], s0[p], s0[HP + p], s0[2*HP + p], sS[2*HP + p], sS[3*HP + p]};
}

__device__ inline int lpos(int i, int j) { return (j + 1)*HX + (i + 1); }

struct Raw { double p, T, u, v, w; };

template <int KEEP, int PDISS, int VISC>
__global__ void __launch_bounds__(256, 2)
k_flux_fast(DevGeom G, const double* __restrict__ prims, double* __restrict__ rhs, FastCfg C, int increment)
{
    using S = Spec<KEEP, PDISS, VISC>;
    __shared__ double sU[3][3*HP];        // u, v, w of planes k-1, k, k+1 (ring)
    __shared__ double sS[2][4*HP];        // p, T, rho, e of planes k, k+1 (by parity)
    __shared__ double sFx[5][16*17];      // x-face products, faces i = 0..16 of row j
    __shared__ double sFy[5][17*16];      // y-face products, faces j = 0..16 of column i

    const int lb = blockIdx.x;
    const int tid = threadIdx.x;
    const int i = tid & 15, j = tid >> 4;
    const double* __restrict__ B = prims + (long long)lb*G.sb;
    double* __restrict__ Rb = rhs + (long long)lb*G.sb;
    const double ix = G.inv_dx[3*lb + 0], iy = G.inv_dx[3*lb + 1], iz = G.inv_dx[3*lb + 2];
    const int nz = G.nc[2];
    const long long sv = G.sv;
    const int N0 = G.N[0], N1 = G.N[1], ng = G.ng;

    // halo cell of this thread: waves 0 and 1 own the 64 edge cells and the
    // 4 corners of the 18x18 plane (waves 2 and 3 evaluate the block-boundary faces)
    int hi = 0, hj = 0;
    const bool has_halo = tid < 68;
    if (tid < 16)      { hi = -1;       hj = tid; }
    else if (tid < 32) { hi = 16;       hj = tid - 16; }
    else if (tid < 48) { hi = tid - 32; hj = -1; }
    else if (tid < 64) { hi = tid - 48; hj = 16; }
    else if (tid < 68) { const int c = tid - 64; hi = (c & 1) ? 16 : -1; hj = (c & 2) ? 16 : -1; }
    const int pown = lpos(i, j), phalo = lpos(hi, hj);

    auto gidx = [&](int ci, int cj, int ck) -> long long
    {
        return ((long long)(ck + ng)*N1 + (cj + ng))*N0 + (ci + ng);
    };
    auto load_raw = [&](int ci, int cj, int ck) -> Raw
    {
        const long long o = gidx(ci, cj, ck);
        return Raw{B[o], B[sv + o], B[2*sv + o], B[3*sv + o], B[4*sv + o]};
    };
    auto cell = [&](const Raw& r) -> CellV { return make_cell(C, r.p, r.T, r.u, r.v, r.w); };
    auto slot = [](int k) { return (k + 3) % 3; };
    // a plane entering the window: velocities to the ring, scalars to its parity buffer
    auto put = [&](int k, int p, const CellV& c)
    {
        double* s = sU[slot(k)];
        s[p] = c.u; s[HP + p] = c.v; s[2*HP + p] = c.w;
        double* t = sS[k & 1];
        t[p] = c.p; t[HP + p] = c.T; t[2*HP + p] = c.rho; t[3*HP + p] = c.e;
    };

    // prologue: planes -1, 0, 1 resident, planes 2 .. 1 + PF in flight
    const CellV prev = cell(load_raw(i, j, -1));
    CellV cur = cell(load_raw(i, j, 0));
    CellV nxt = cell(load_raw(i, j, 1));
    Raw nraw[PF], hraw[PF];
    #pragma unroll
    for (int d = 0; d < PF; ++d)
    {
        nraw[d] = Raw{};
        hraw[d] = Raw{};
        if (2 + d <= nz) nraw[d] = load_raw(i, j, 2 + d);
    }
    {
        double* s = sU[slot(-1)];
        s[pown] = prev.u; s[HP + pown] = prev.v; s[2*HP + pown] = prev.w;
    }
    put(0, pown, cur); put(1, pown, nxt);
    if (has_halo)
    {
        const CellV hp = cell(load_raw(hi, hj, -1));
        double* s = sU[slot(-1)];
        s[phalo] = hp.u; s[HP + phalo] = hp.v; s[2*HP + phalo] = hp.w;
        put(0, phalo, cell(load_raw(hi, hj, 0)));
        put(1, phalo, cell(load_raw(hi, hj, 1)));
        #pragma unroll
        for (int d = 0; d < PF; ++d)
            if (2 + d <= nz) hraw[d] = load_raw(hi, hj, 2 + d);
    }
    __syncthreads();

    double zlo[5];
    {
        double g[3][3], gT;
        grads_z<S>(sU[slot(-1)], sU[slot(0)], pown, prev, cur, ix, iy, iz, g, gT);
        face_product<2, S>(C, prev, cur, g, gT, iz, zlo);
    }

    for (int k = 0; k < nz; ++k)
    {
        const double* sm = sU[slot(k - 1)];
        const double* s0 = sU[slot(k)];
        const double* sp = sU[slot(k + 1)];
        const double* sk = sS[k & 1];
        // low x-face of (i,j): L = (i-1, j)
        {
            const int pL = pown - 1;
            const CellV L = cell_from_lds(sk, s0, pL);
            double g[3][3], gT, P[5];
            grads_x<S>(sm, s0, sp, pL, pown, L, cur, ix, iy, iz, g, gT);
            face_product<0, S>(C, L, cur, g, gT, ix, P);
            for (int n = 0; n < 5; ++n) sFx[n][j*17 + i] = P[n];
        }
        // low y-face of (i,j): L = (i, j-1)
        {
            const int pL = pown - HX;
            const CellV L = cell_from_lds(sk, s0, pL);
            double g[3][3], gT, P[5];
            grads_y<S>(sm, s0, sp, pL, pown, L, cur, ix, iy, iz, g, gT);
            face_product<1, S>(C, L, cur, g, gT, iy, P);
            for (int n = 0; n < 5; ++n) sFy[n][j*16 + i] = P[n];
        }
        // block-boundary x-faces (16 - 1/2, jj) on wave 2
        if (tid >= 128 && tid < 144)
        {
            const int jj = tid - 128;
            const int pL = lpos(15, jj), pR = lpos(16, jj);
            const CellV L = cell_from_lds(sk, s0, pL), Rr = cell_from_lds(sk, s0, pR);
            double g[3][3], gT, P[5];
            grads_x<S>(sm, s0, sp, pL, pR, L, Rr, ix, iy, iz, g, gT);
            face_product<0, S>(C, L, Rr, g, gT, ix, P);
            for (int n = 0; n < 5; ++n) sFx[n][jj*17 + 16] = P[n];
        }
        // block-boundary y-faces (ii, 16 - 1/2) on wave 3
        if (tid >= 192 && tid < 208)
        {
            const int ii = tid - 192;
            const int pL = lpos(ii, 15), pR = lpos(ii, 16);
            const CellV L = cell_from_lds(sk, s0, pL), Rr = cell_from_lds(sk, s0, pR);
            double g[3][3], gT, P[5];
            grads_y<S>(sm, s0, sp, pL, pR, L, Rr, ix, iy, iz, g, gT);
            face_product<1, S>(C, L, Rr, g, gT, iy, P);
            for (int n = 0; n < 5; ++n) sFy[n][16*16 + ii] = P[n];
        }
        // high z-face k+1/2: L = (i,j,k), R = (i,j,k+1)
        double zhi[5];
        {
            double g[3][3], gT;
            grads_z<S>(s0, sp, pown, cur, nxt, ix, iy, iz, g, gT);
            face_product<2, S>(C, cur, nxt, g, gT, iz, zhi);
        }
        __syncthreads();

        // rhs of cell (i,j,k): faces in the order x-, x+, y-, y+, z-, z+
        // (flux_div_basic.h:54-72; the high faces enter as -P)
        const long long o = gidx(i, j, k);
        #pragma unroll
        for (int n = 0; n < 5; ++n)
        {
            double r = increment ? Rb[n*sv + o] : 0.0;
            r += sFx[n][j*17 + i];
            r -= sFx[n][j*17 + i + 1];
            r += sFy[n][j*16 + i];
            r -= sFy[n][(j + 1)*16 + i];
            r += zlo[n];
            r -= zhi[n];
            Rb[n*sv + o] = r;
            zlo[n] = zhi[n];
        }

        if (k + 1 < nz)
        {
            // plane k+2 (<= nz) enters from the prefetched values into the ring
            // slot of plane k-1 and the parity buffer of plane k; plane k+2+PF is requested
            cur = nxt;
            nxt = cell(nraw[0]);
            put(k + 2, pown, nxt);
            #pragma unroll
            for (int d = 0; d + 1 < PF; ++d) nraw[d] = nraw[d + 1];
            if (k + 2 + PF <= nz) nraw[PF - 1] = load_raw(i, j, k + 2 + PF);
            if (has_halo)
            {
                put(k + 2, phalo, cell(hraw[0]));
                #pragma unroll
                for (int d = 0; d + 1 < PF; ++d) hraw[d] = hraw[d + 1];
                if (k + 2 + PF <= nz) hraw[PF - 1] = load_raw(hi, hj, k + 2 + PF);
            }
        }
        __syncthreads();
    }
}

template <int KEEP, int PDISS, int VISC>
void launch_fast(const Grid& g, const double* prims, double* rhs, const FastCfg& C, bool increment, hipStream_t s)
{
    hipLaunchKernelGGL((k_flux_fast<KEEP, PDISS, VISC>), dim3(g.geom.nblk), dim3(256), 0, s,
                       g.geom, prims, rhs, C, increment ? 1 : 0);
    SPD_LAUNCH_CHECK();
}

template <int KEEP, int PDISS>
bool dispatch_visc(int visc, const Grid& g, const double* prims, double* rhs, const FastCfg& C, bool inc, hipStream_t s)
{
    switch (visc)
    {
        case 0: if (KEEP || PDISS) { launch_fast<KEEP, PDISS, 0>(g, prims, rhs, C, inc, s); return true; } return false;
        case 1: launch_fast<KEEP, PDISS, 1>(g, prims, rhs, C, inc, s); return true;
        case 2: launch_fast<KEEP, PDISS, 2>(g, prims, rhs, C, inc, s); return true;
    }
    return false;
}

} // namespace

namespace {

// Selects the fast kernel for (grid, flux); false when the generic kernel must run.
bool select_fast(const DevGeom& G, const spd_flux_desc& f, int& keep, int& pd, int& vs, FastCfg& C)
{
    if (!G.identity || G.nc[0] != 16 || G.nc[1] != 16 || G.ng < 1) return false;
    if (const char* e = getenv("SPD_DISABLE_FAST")) if (e[0] == '1') return false;
    // terms must be a subsequence of [keep2, pdiss, visc] in that order
    keep = 0; pd = 0; vs = 0;
    int stage = 0;
    C = FastCfg{};
    C.R = f.R;
    C.gm1 = f.gamma - 1.0;
    C.spec_heat = f.gamma*f.R/(f.gamma - 1.0);
    C.half_RR = 0.5*(f.R + f.R);
    C.inv_gamma = 1.0/(0.5*(f.gamma + f.gamma));
    bool has_keep = false;
    for (int m = 0; m < f.nterms; ++m) has_keep |= f.terms[m].kind == SPD_TERM_KEEP;
    const double sc = has_keep ? 8.0 : 1.0;   // composition scale (see face_product)
    for (int m = 0; m < f.nterms; ++m)
    {
        const spd_flux_term& t = f.terms[m];
        if (t.kind == SPD_TERM_KEEP && t.order == 2 && stage < 1) { keep = 1; stage = 1; }
        else if (t.kind == SPD_TERM_PDISS && stage < 2)
        {
            pd = 1 + t.sensor; stage = 2;
            C.scl_s = sc*t.sensor_param; C.ducros_eps = t.sensor_param; C.ducros_s = sc;
            C.eps_p = t.eps_p; C.eps_T = t.eps_T;
        }
        else if (t.kind == SPD_TERM_VISC && stage < 3)
        {
            vs = 1 + t.law; stage = 3;
            // constant law (viscous_laws.h:77-99): beta = -2 mu / 3, alpha = mu / Pr
            C.mu_s = sc*t.mu;
            C.visc2_s = sc*double(-2.0*t.mu/3.0);
            C.cond_s = C.spec_heat*(sc*(t.mu*(1.0/t.prandtl)));
            // power law (viscous_laws.h:121-134)
            C.mu_ref_s = sc*t.mu; C.T_ref = t.T_ref; C.power = t.power; C.prandtl = t.prandtl;
            C.inv_T_ref = 1.0/t.T_ref; C.inv_prandtl = 1.0/t.prandtl;
            C.t_ref_one = (t.T_ref == 1.0);
            C.ps = PowSeries::make(t.power);
            if (t.law == SPD_LAW_POWER && !(t.T_ref > 0.0 && t.prandtl > 0.0)) return false;
        }
        else return false;
    }
    return true;
}

} // namespace

bool flux_div_fast_eligible(const Grid& g, const spd_flux_desc& f)
{
    int keep, pd, vs;
    FastCfg C;
    return select_fast(g.geom, f, keep, pd, vs, C);
}

bool launch_flux_div_fast(const Grid& g, const double* prims, double* rhs, const spd_flux_desc& f,
                          bool increment, hipStream_t s)
{
    int keep, pd, vs;
    FastCfg C;
    if (!select_fast(g.geom, f, keep, pd, vs, C)) return false;
    switch (keep*3 + pd)
    {
        case 0: return dispatch_visc<0, 0>(vs, g, prims, rhs, C, increment, s);
        case 1: return dispatch_visc<0, 1>(vs, g, prims, rhs, C, increment, s);
        case 2: return dispatch_visc<0, 2>(vs, g, prims, rhs, C, increment, s);
        case 3: return dispatch_visc<1, 0>(vs, g, prims, rhs, C, increment, s);
        case 4: return dispatch_visc<1, 1>(vs, g, prims, rhs, C, increment, s);
        case 5: return dispatch_visc<1, 2>(vs, g, prims, rhs, C, increment, s);
    }
    return false;
}

} // namespace spd
