// LDS-staged flux_div for the hot combinations on gfx950:
//   cent_keep<2> [+ pressure_diss_lr (const | Ducros)] [+ visc_lr (const | power law)]
// on identity coordinates and blocks with a 16x16 (i,j) cross-section.
//
// One workgroup (256 threads, thread (i,j)) owns one block and marches over k.
// Every face flux is evaluated once: thread (i,j) evaluates the low x-face
// i-1/2, the low y-face j-1/2 and the high z-face k+1/2 of its cell; the 16
// block-boundary x-faces and the 16 y-faces of a plane are evaluated by 16
// lanes of waves 2 and 3.  In-plane face products P = F * (1/dx) are exchanged
// through LDS; the z-face product stays in registers for the next plane.
//
// LDS holds, for the three planes k-1, k, k+1, the velocities (u,v,w) on the
// 18x18 plane with one halo layer (ring of three slots), and for planes k and
// k+1 the p, T, rho, e needed as left/right states of the faces.
//
// Power law: x^n comes from a binomial series valid near x = 1.  A block
// with a face temperature outside its range stops at that plane, and
// k_flux_fixup (flux_generic.hip) completes it with the library pow.
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
#include "spd_stage.hpp"

namespace spd {
namespace {

constexpr int HX = 18, HP = 18*18;


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

// x^n of the power law by the branch-free binomial series; `bad` records a
// lane whose x is outside the series range (its block is then completed by
// the fix-up kernel with the library pow, from that plane on)
__device__ inline double pow_law(const FastCfg& C, double x, bool& bad)
{
    bad |= !C.ps.in_range(x);
    return C.ps.eval(x);
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
template <int D, int LAW, bool MAP>
__device__ inline void visc_s(const FastCfg& C, const FaceSums& f, const double g[3][3], double gT, double nf,
                              double F[5], bool& bad)
{
    double mu, visc2, cond;
    if (LAW == SPD_LAW_POWER)
    {
        const double Tf = f.ST*0.5;
        const double x = C.t_ref_one ? Tf : div_const(Tf, C.T_ref, C.inv_T_ref);
        mu    = C.mu_ref_s*pow_law(C, x, bad);
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
    if constexpr (MAP)
    {
        // -(n . ht), -(n . tau) with the area-weighted face normal n = nf e_D
        F[1] = -(nf*ht);
        F[2] = -(nf*t[0]);
        F[3] = -(nf*t[1]);
        F[4] = -(nf*t[2]);
    }
    else
    {
        F[1] = -ht;
        F[2] = -t[0];
        F[3] = -t[1];
        F[4] = -t[2];
    }
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
// On mapped (diagonal) coordinates (MAP) the normals enter: mt = {n_D of the
// left cell, n_D of the right cell, n_D of the face} (metric(), the area
// factors of coord_system.h:250-267), u_n = u_D n_D (the other components of a
// diagonal normal are zero); the result is then the flux scaled by the
// composition scale, not yet by 1/dx: the cells apply their own
// jac (1 - 2 pm)/dx (flux_div_basic.h:66-68).
template <int D, class S, bool MAP>
__device__ inline void face_product(const FastCfg& C, const CellV& L, const CellV& R, const double g[3][3], double gT,
                                    double idx, const double mt[3], double P[5], bool& bad)
{
    const FaceSums f = face_sums(L, R);
    double s8[5];
    if (S::vs) visc_s<D, S::law, MAP>(C, f, g, gT, mt[2], s8, bad);
    if (S::pd)
    {
        double z[5];
        pdiss_s<S::sensor>(C, L, R, f, g, z);
        if (S::vs) { for (int n = 0; n < 5; ++n) s8[n] = z[n] + s8[n]; }
        else       { for (int n = 0; n < 5; ++n) s8[n] = z[n]; }
    }
    if (S::keep)
    {
        double unL = vel<D>(L), unR = vel<D>(R);
        if constexpr (MAP) { unL = unL*mt[0]; unR = unR*mt[1]; }
        const double CX = (L.rho + R.rho)*(unR + unL);
        const double Ep = CX*f.K + CX*(L.e + R.e) + 4.0*(unL*R.p + unR*L.p);
        double M[3] = {CX*f.SU, CX*f.SV, CX*f.SW};
        if constexpr (MAP) M[D] = M[D] + 4.0*(f.SP*mt[2]);
        else               M[D] = M[D] + 4.0*f.SP;
        if constexpr (MAP)
        {
            // 8 F: continuity 2 CX (+ s8), the others as below; the cells scale by 1/8
            if (S::vs || S::pd)
            {
                P[0] = S::pd ? 2.0*(CX + 0.5*s8[0]) : 2.0*CX;
                P[1] = Ep + s8[1]; P[2] = M[0] + s8[2]; P[3] = M[1] + s8[3]; P[4] = M[2] + s8[4];
            }
            else
            {
                P[0] = 2.0*CX; P[1] = Ep; P[2] = M[0]; P[3] = M[1]; P[4] = M[2];
            }
            return;
        }
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
        for (int n = 0; n < 5; ++n) P[n] = MAP ? s8[n] : s8[n]*idx;
    }
}

// info::gradient on mapped coordinates (omni/infos/info_gradient.h, the
// coordinate scaling of face_grad): component a of every gradient divided by
// d x_a / d xi_a at the face point (mn for the face normal axis D, the cell
// centre values mt1 / mt2 for the other two); identity axes divide by 1.
template <int D>
__device__ inline void grads_map(const DevGeom& G, double g[3][3], double& gT, double mD, double mA, double mB)
{
    constexpr int A = (D + 1) % 3, B = (D + 2) % 3;
    if (G.map_kind[D] != SPD_MAP_IDENTITY) { for (int b = 0; b < 3; ++b) g[D][b] = g[D][b]/mD; gT = gT/mD; }
    if (G.map_kind[A] != SPD_MAP_IDENTITY) for (int b = 0; b < 3; ++b) g[A][b] = g[A][b]/mA;
    if (G.map_kind[B] != SPD_MAP_IDENTITY) for (int b = 0; b < 3; ++b) g[B][b] = g[B][b]/mB;
}

// metric() of a diagonal map at a point with derivatives m0, m1, m2: jac =
// 1/(m0 m1 m2), normal component D = 1/(m_D jac) (core/coord_system.h:250-302)
__device__ inline double jac_at(double m0, double m1, double m2) { return 1.0/(m0*m1*m2); }
__device__ inline double nrm_at(double mD, double jac) { return 1.0/(mD*jac); }

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

// The RK stage a fused launch applies to each cell once its residual L is
// known (transform-aware, time-integration/advance.h:56-102):
//   qout = c2p(p2c(q) + c[0] r[0] + .. + c[n-2] r[n-2] + c[n-1] L)
// in k_rk_transform's order (nonzero coefficients only), then pushed into the
// ghosts of qout it feeds.
struct StageArgs
{
    double* qout;
    const double* r[2];          // residuals of the earlier stages, rhs layout
    double c[3];
    int n;                       // 1..3 coefficients, the last one for L
    int store_rhs;               // also write L to rhs
    double gamma, R;
    const int* push;             // push table (d_push)
    int push_ghosts;             // write the ghosts this stage's result feeds
    const int* nbr;              // source table (d_nbr), read halos from neighbour interiors (STG == 2)
    double t_wall;               // isothermal wall temperature (wall_y)
    int wall_y;                  // STG == 2: y faces without a source block are isothermal walls
};

// STG: 0 flux_div only; 1 fused stage, halo from this block's ghost cells;
// 2 fused stage, halo from the neighbour blocks' interiors (every ghost
// region has a source block, or, in the mapped kernel, the y regions without
// one are isothermal walls whose ghosts are formed from the mirrored cells)
template <int KEEP, int PDISS, int VISC, bool INC, bool MAP, int STG = 0>
__global__ void __launch_bounds__(256, 2)
k_flux_fast(DevGeom G, const double* __restrict__ prims, double* __restrict__ rhs, FastCfg C,
            int* __restrict__ start, StageArgs A, int split)
{
    using S = Spec<KEEP, PDISS, VISC>;
    static_assert(!(STG && INC), "a fused stage writes a fresh residual");
    // x-ghost pushes of a fused stage: the plane's cells within ng of an x
    // face, staged for whole-row writes (flushed after the next barrier)
    __shared__ double sXs[STG ? sizeof(XStage)/sizeof(double) : 1];
    XStage& xs = *reinterpret_cast<XStage*>(sXs);
    __shared__ int sBad[2];               // a face of plane k is outside the pow series range (by parity)
    // mapped coordinates: d x_a / d xi_a of this block at cell centres -1 .. 16
    // and at faces 0 .. 16 of each axis, from the same map_deriv as the
    // per-cell kernels (so the same bits)
    __shared__ double sMc[MAP ? 3 : 1][18], sMf[MAP ? 3 : 1][17];
    // mapped y only (x and z identity, so their d x / d xi are 1.0 exactly):
    // every face normal and Jacobian depends on the row alone, tabulated per
    // block with the same expressions: sY[0][r] = jac at cell row r - 1,
    // sY[1][r] = its normal factor 1/(1 jac), sY[2][r] = the y-face normal
    // factor of a cell row, sY[3][f] = the y-face normal factor at face f
    __shared__ double sY[MAP ? 4 : 1][18];
    __shared__ double sU[3][3*HP];        // u, v, w of planes k-1, k, k+1 (ring)
    __shared__ double sS[2][4*HP];        // p, T, rho, e of planes k, k+1 (by parity)
    __shared__ double sFx[5][16*17];      // x-face products, faces i = 0..16 of row j
    __shared__ double sFy[5][17*16];      // y-face products, faces j = 0..16 of column i

    // zsplit 2 or 4: that many workgroups per block, each marching over its
    // share of the planes [k0, k1) with its own window prologue (fewer, shorter tail rounds when
    // the block count is not a multiple of the resident workgroups); each
    // records the first plane it left to the fix-up in its own d_start entry.
    // The fix-up redoes the block from the first of them to the end, which
    // overwrites whatever the other half wrote there: the same planes the
    // one-workgroup launch leaves, so the same bits (overwrite mode only)
    // split: workgroups per block in the low byte, bit 8 set for XCD runs
    const int zsplit = split & 0xff;
    const int w = (split & 0x100) ? xcd_run_order(blockIdx.x, G.nblk*zsplit) : blockIdx.x;
    const int lb = w/zsplit, half = w - lb*zsplit;
    const int k0 = half*(G.nc[2]/zsplit), k1 = k0 + G.nc[2]/zsplit;
    int* __restrict__ my_start = start + kFixupParts*lb + half;
    const int tid = threadIdx.x;
    if (half == 0 && tid < kFixupParts && tid >= zsplit) start[kFixupParts*lb + tid] = G.nc[2];
    const int i = tid & 15, j = tid >> 4;
    const double ix = G.inv_dx[3*lb + 0], iy = G.inv_dx[3*lb + 1], iz = G.inv_dx[3*lb + 2];
    const bool yonly = MAP && G.map_kind[0] == SPD_MAP_IDENTITY && G.map_kind[2] == SPD_MAP_IDENTITY;
    if constexpr (MAP)
    {
        if (tid < 54)
        {
            const int a = tid/18, c = tid % 18 - 1;
            const double m = map_deriv(G, a, G.bmin[3*lb + a] + (c + 0.5)*G.dx[3*lb + a]);
            sMc[a][c + 1] = m;
            if (a == 1)
            {
                const double jc = jac_at(1.0, m, 1.0);
                sY[0][c + 1] = jc; sY[1][c + 1] = nrm_at(1.0, jc); sY[2][c + 1] = nrm_at(m, jc);
            }
        }
        else if (tid < 105)
        {
            const int a = (tid - 54)/17, c = (tid - 54) % 17;
            const double m = map_deriv(G, a, G.bmin[3*lb + a] + ((c + 0.5) - 0.5)*G.dx[3*lb + a]);
            sMf[a][c] = m;
            if (a == 1) sY[3][c] = nrm_at(m, jac_at(1.0, m, 1.0));
        }
    }
    const int nz = G.nc[2];
    const long long sv = G.sv;
    const int N1 = G.N[1];
    // plane (k) strides of a core column and of an x-halo column
    const long long plane = (long long)G.nc[0]*N1, xplane = (long long)N1;

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

    // column bases at plane k = 0; plane k is base + k*stride (the own column
    // and the residual are core columns, a halo column with hi = -1 or 16 is
    // an x-halo column)
    const long long blk = (long long)lb*G.sb;
    const double* __restrict__ Bown = prims + blk + stored_offset(G, i, j, 0);
    const double* __restrict__ Bhalo = prims + blk + stored_offset(G, hi, hj, 0);
    const long long hplane = (hi >= 0 && hi < 16) ? plane : xplane;
    // STG == 2: the halo is read from the neighbour blocks' interiors (d_nbr:
    // the block that feeds each ghost region), never from this block's ghost
    // cells; the halo column is then a core column of its x/y neighbour, and
    // planes -1 and nz come from the z neighbours (the same values an
    // exchange would have copied into the ghosts)
    const int hex = hi < 0 ? -1 : (hi > 15 ? 1 : 0), hey = hj < 0 ? -1 : (hj > 15 ? 1 : 0);
    const int* __restrict__ nbl = STG == 2 ? A.nbr + 27*lb : nullptr;
    // a y-halo cell across an isothermal wall (no source block in y): what
    // the wall fill writes there, i.e. the cell of the mirrored row (0 for the
    // low wall, nc1 - 1 for the high one) in the same column and plane,
    // itself found through the source table without the y step, with
    // T -> 2 T_wall - T and the velocity negated (boundary_fill.h)
    // (built into the mapped variant only, where the walled channel runs)
    const bool hwall = STG == 2 && MAP && has_halo && A.wall_y && hey != 0 && nbl[(hex + 1) + 3*(hey + 1) + 9] < 0;
    const int hey_s = hwall ? 0 : hey;
    const int hj_s = hwall ? (hey < 0 ? -1 - hj : 2*G.nc[1] - (hj + 1)) : hj - 16*hey;
    // (a mirrored cell of the block's own columns lies in the block itself:
    // the source table has no entry for the block's own region)
    const int src_n = (STG == 2 && has_halo)
        ? ((hex == 0 && hey_s == 0) ? lb : nbl[(hex + 1) + 3*(hey_s + 1) + 9]) : 0;
    const double* __restrict__ Bhalo_n = (STG == 2 && has_halo)
        ? prims + (long long)src_n*G.sb + stored_offset(G, hi - 16*hex, hj_s, 0)
        : Bhalo;
    auto load_ptr = [&](const double* e) -> Raw { return Raw{e[0], e[sv], e[2*sv], e[3*sv], e[4*sv]}; };
    auto load_own = [&](int ck) -> Raw
    {
        if (STG != 2 || (ck >= 0 && ck < nz)) return load_ptr(Bown + ck*plane);
        const int ez = ck < 0 ? -1 : 1;
        const long long S = nbl[1 + 3 + 9*(ez + 1)];
        return load_ptr(prims + S*G.sb + stored_offset(G, i, j, ck - ez*nz));
    };
    auto load_halo = [&](int ck) -> Raw
    {
        if (STG != 2) return load_ptr(Bhalo + ck*hplane);
        Raw r;
        if (ck >= 0 && ck < nz) r = load_ptr(Bhalo_n + ck*plane);
        else
        {
            const int ez = ck < 0 ? -1 : 1;
            const long long S = nbl[(hex + 1) + 3*(hey_s + 1) + 9*(ez + 1)];
            r = load_ptr(prims + S*G.sb + stored_offset(G, hi - 16*hex, hj_s, ck - ez*nz));
        }
        if (MAP && hwall) r = Raw{r.p, 2.0*A.t_wall - r.T, -r.u, -r.v, -r.w};
        return r;
    };
    const long long own_off = blk + stored_offset(G, i, j, 0);   // core column (i, j) at k = 0
    double* __restrict__ Rown = rhs + own_off;

    auto cell = [&](const Raw& r) -> CellV { return make_cell(C, r.p, r.T, r.u, r.v, r.w); };
    // LDS window: ring slots of velocities (rotated by pointer), scalar buffers by parity
    double* ring[3] = {sU[0], sU[1], sU[2]};      // planes k-1, k, k+1
    auto put_vel = [&](double* s, int p, const CellV& c) { s[p] = c.u; s[HP + p] = c.v; s[2*HP + p] = c.w; };
    auto put_sca = [&](double* t, int p, const CellV& c)
    {
        t[p] = c.p; t[HP + p] = c.T; t[2*HP + p] = c.rho; t[3*HP + p] = c.e;
    };

    // increment mode: the residual of the next plane to be written is requested
    // before that iteration's primitive prefetch, so waiting for it never waits
    // for the prefetch
    double rin[5];
    #pragma unroll
    for (int n = 0; n < 5; ++n) rin[n] = INC ? Rown[k0*plane + n*sv] : 0.0;

    // prologue: planes k0-1, k0, k0+1 resident; planes k0+2 and k0+3 in flight
    // (two raw register sets used alternately, so a plane's load is never
    // copied and waiting for it never waits for the later one)
    const CellV prev = cell(load_own(k0 - 1));
    const CellV cur0 = cell(load_own(k0));
    const CellV nxt0 = cell(load_own(k0 + 1));
    Raw nA{}, nB{}, hA{}, hB{};
    if (k0 + 2 <= k1) nA = load_own(k0 + 2);
    if (k0 + 3 <= k1) nB = load_own(k0 + 3);
    put_vel(ring[0], pown, prev);
    put_vel(ring[1], pown, cur0); put_sca(sS[0], pown, cur0);
    put_vel(ring[2], pown, nxt0); put_sca(sS[1], pown, nxt0);
    if (tid == 0) { sBad[0] = 0; sBad[1] = 0; }
    if (has_halo)
    {
        put_vel(ring[0], phalo, cell(load_halo(k0 - 1)));
        const CellV h0 = cell(load_halo(k0)), h1 = cell(load_halo(k0 + 1));
        put_vel(ring[1], phalo, h0); put_sca(sS[0], phalo, h0);
        put_vel(ring[2], phalo, h1); put_sca(sS[1], phalo, h1);
        if (k0 + 2 <= k1) hA = load_halo(k0 + 2);
        if (k0 + 3 <= k1) hB = load_halo(k0 + 3);
    }
    lds_barrier();

    // mapped coordinates: this column's d x / d xi along x and y (cell centres)
    const double m0i = MAP ? sMc[0][i + 1] : 1.0, m1j = MAP ? sMc[1][j + 1] : 1.0;
    double zlo[5];
    bool bad = false;                     // carried into plane 0's check
    {
        double g[3][3], gT, mt[3] = {1.0, 1.0, 1.0};
        grads_z<S>(ring[0], ring[1], pown, prev, cur0, ix, iy, iz, g, gT);
        if constexpr (MAP)
        {
            // face k0 - 1/2: L = plane k0 - 1, R = plane k0
            const double mL = sMc[2][k0], mR = sMc[2][k0 + 1], mf = sMf[2][k0];
            if (yonly) { mt[0] = mt[1] = mt[2] = sY[1][j + 1]; }
            else
            {
                mt[0] = nrm_at(mL, jac_at(m0i, m1j, mL));
                mt[1] = nrm_at(mR, jac_at(m0i, m1j, mR));
                mt[2] = nrm_at(mf, jac_at(m0i, m1j, mf));
            }
            grads_map<2>(G, g, gT, mf, m0i, m1j);
        }
        face_product<2, S, MAP>(C, prev, cur0, g, gT, iz, mt, zlo, bad);
    }

    // One plane k: every face, the residual of its cells, and the entry of
    // plane k+2 (from nr / hr, which are then refilled with plane k+4).
    // Returns false when the block must be left to the fix-up kernel.
    auto plane_step = [&](int k, Raw& nr, Raw& hr) -> bool
    {
        // the x pushes staged by plane k-1 (visible since its last barrier;
        // the staging buffer is refilled only after this plane's first barrier)
        if (STG && A.push_ghosts && k > k0) flush_x(G, A.push, A.qout, xs, lb, k - 1);
        const double* sm = ring[0];
        const double* s0 = ring[1];
        const double* sp = ring[2];
        const double* sk = sS[k & 1];
        // the cells of this column in planes k and k+1, from the window
        const CellV cur = cell_from_lds(sk, s0, pown);
        const CellV nxt = cell_from_lds(sS[(k + 1) & 1], sp, pown);
        double zhi[5];
        // this cell's own low x- and y-face products (the LDS copies serve
        // the neighbour cells that share those faces); the residual reads
        // them from registers where that fits the register budget (plain
        // flux_div on identity coordinates without pressure dissipation)
        double px[5], py[5];
        constexpr bool own_reg = STG == 0 && !MAP && !S::pd;
        // mapped coordinates: plane k's d z / d zeta and this cell's Jacobian
        const double m2k = MAP ? sMc[2][k + 1] : 1.0;
        const double jc = MAP ? (yonly ? sY[0][j + 1] : jac_at(m0i, m1j, m2k)) : 1.0;
        // the three faces of all lanes, then the block-boundary x-faces
        // (wave 2) and y-faces (wave 3); the series pow has no branch, so the
        // three faces form one straight-line block
        {
            // low x-face of (i,j): L = (i-1, j)
            const int pL = pown - 1;
            const CellV L = cell_from_lds(sk, s0, pL);
            double g[3][3], gT, mt[3] = {1.0, 1.0, 1.0};
            grads_x<S>(sm, s0, sp, pL, pown, L, cur, ix, iy, iz, g, gT);
            if constexpr (MAP)
            {
                const double mL = sMc[0][i], mf = sMf[0][i];
                if (yonly) { mt[0] = mt[1] = mt[2] = sY[1][j + 1]; }
                else
                {
                    mt[0] = nrm_at(mL, jac_at(mL, m1j, m2k));
                    mt[1] = nrm_at(m0i, jc);
                    mt[2] = nrm_at(mf, jac_at(mf, m1j, m2k));
                }
                grads_map<0>(G, g, gT, mf, m1j, m2k);
            }
            face_product<0, S, MAP>(C, L, cur, g, gT, ix, mt, px, bad);
            for (int n = 0; n < 5; ++n) sFx[n][j*17 + i] = px[n];
        }
        {
            // low y-face of (i,j): L = (i, j-1)
            const int pL = pown - HX;
            const CellV L = cell_from_lds(sk, s0, pL);
            double g[3][3], gT, mt[3] = {1.0, 1.0, 1.0};
            grads_y<S>(sm, s0, sp, pL, pown, L, cur, ix, iy, iz, g, gT);
            if constexpr (MAP)
            {
                const double mL = sMc[1][j], mf = sMf[1][j];
                if (yonly) { mt[0] = sY[2][j]; mt[1] = sY[2][j + 1]; mt[2] = sY[3][j]; }
                else
                {
                    mt[0] = nrm_at(mL, jac_at(m0i, mL, m2k));
                    mt[1] = nrm_at(m1j, jc);
                    mt[2] = nrm_at(mf, jac_at(m0i, mf, m2k));
                }
                grads_map<1>(G, g, gT, mf, m2k, m0i);
            }
            face_product<1, S, MAP>(C, L, cur, g, gT, iy, mt, py, bad);
            for (int n = 0; n < 5; ++n) sFy[n][j*16 + i] = py[n];
        }
        {
            // high z-face k+1/2: L = (i,j,k), R = (i,j,k+1)
            double g[3][3], gT, mt[3] = {1.0, 1.0, 1.0};
            grads_z<S>(s0, sp, pown, cur, nxt, ix, iy, iz, g, gT);
            if constexpr (MAP)
            {
                const double mR = sMc[2][k + 2], mf = sMf[2][k + 1];
                if (yonly) { mt[0] = mt[1] = mt[2] = sY[1][j + 1]; }
                else
                {
                    mt[0] = nrm_at(m2k, jc);
                    mt[1] = nrm_at(mR, jac_at(m0i, m1j, mR));
                    mt[2] = nrm_at(mf, jac_at(m0i, m1j, mf));
                }
                grads_map<2>(G, g, gT, mf, m0i, m1j);
            }
            face_product<2, S, MAP>(C, cur, nxt, g, gT, iz, mt, zhi, bad);
        }
        if (tid >= 128 && tid < 144)
        {
            // block-boundary x-faces (16 - 1/2, jj)
            const int jj = tid - 128;
            const int pL = lpos(15, jj), pR = lpos(16, jj);
            const CellV L = cell_from_lds(sk, s0, pL), Rr = cell_from_lds(sk, s0, pR);
            double g[3][3], gT, P[5], mt[3] = {1.0, 1.0, 1.0};
            grads_x<S>(sm, s0, sp, pL, pR, L, Rr, ix, iy, iz, g, gT);
            if constexpr (MAP)
            {
                const double m1 = sMc[1][jj + 1], mL = sMc[0][16], mR = sMc[0][17], mf = sMf[0][16];
                if (yonly) { mt[0] = mt[1] = mt[2] = sY[1][jj + 1]; }
                else
                {
                    mt[0] = nrm_at(mL, jac_at(mL, m1, m2k));
                    mt[1] = nrm_at(mR, jac_at(mR, m1, m2k));
                    mt[2] = nrm_at(mf, jac_at(mf, m1, m2k));
                }
                grads_map<0>(G, g, gT, mf, m1, m2k);
            }
            face_product<0, S, MAP>(C, L, Rr, g, gT, ix, mt, P, bad);
            for (int n = 0; n < 5; ++n) sFx[n][jj*17 + 16] = P[n];
        }
        if (tid >= 192 && tid < 208)
        {
            // block-boundary y-faces (ii, 16 - 1/2)
            const int ii = tid - 192;
            const int pL = lpos(ii, 15), pR = lpos(ii, 16);
            const CellV L = cell_from_lds(sk, s0, pL), Rr = cell_from_lds(sk, s0, pR);
            double g[3][3], gT, P[5], mt[3] = {1.0, 1.0, 1.0};
            grads_y<S>(sm, s0, sp, pL, pR, L, Rr, ix, iy, iz, g, gT);
            if constexpr (MAP)
            {
                const double m0 = sMc[0][ii + 1], mL = sMc[1][16], mR = sMc[1][17], mf = sMf[1][16];
                if (yonly) { mt[0] = sY[2][16]; mt[1] = sY[2][17]; mt[2] = sY[3][16]; }
                else
                {
                    mt[0] = nrm_at(mL, jac_at(m0, mL, m2k));
                    mt[1] = nrm_at(mR, jac_at(m0, mR, m2k));
                    mt[2] = nrm_at(mf, jac_at(m0, mf, m2k));
                }
                grads_map<1>(G, g, gT, mf, m2k, m0);
            }
            face_product<1, S, MAP>(C, L, Rr, g, gT, iy, mt, P, bad);
            for (int n = 0; n < 5; ++n) sFy[n][16*16 + ii] = P[n];
        }
        #pragma unroll
        for (int n = 0; n < 5; ++n) keep_here(zhi[n]);
        // a fused stage's earlier residuals of cell (i,j,k), requested before
        // the barrier so that their latency overlaps the wait for the other waves
        double rprev[2][5];
        if constexpr (STG)
        {
            const long long co = own_off + k*plane;
            #pragma unroll
            for (int t = 0; t < 2; ++t)
            {
                if (t + 1 >= A.n || A.c[t] == 0.0) continue;
                #pragma unroll
                for (int v = 0; v < 5; ++v) rprev[t][v] = A.r[t][co + v*sv];
            }
        }
        if (bad) sBad[k & 1] = 1;
        lds_barrier();
        if (sBad[k & 1])
        {
            // some face of plane k needs the library pow: planes k .. nz-1 are
            // left to the fix-up kernel (nothing of them was written)
            if (tid == 0) *my_start = k;
            return false;
        }
        if (tid == 0) sBad[(k + 1) & 1] = 0;
        bad = false;

        // rhs of cell (i,j,k): faces in the order x-, x+, y-, y+, z-, z+
        // (flux_div_basic.h:54-72; the high faces enter as -P)
        double* __restrict__ ro = Rown + k*plane;
        // identity coordinates: the face products already carry 1/dx; mapped:
        // each cell scales the (composition-scaled) face fluxes by its own
        // jac/dx, i.e. s = jac (1 - 2 pm) inv_dx (flux_div_basic.h:66-68)
        double fx = 1.0, fy = 1.0, fz = 1.0;
        if constexpr (MAP)
        {
            const double sc = S::keep ? 0.125 : 1.0;
            fx = sc*(jc*ix); fy = sc*(jc*iy); fz = sc*(jc*iz);
        }
        double rr[5];
        #pragma unroll
        for (int n = 0; n < 5; ++n)
        {
            double r = rin[n];
            if constexpr (MAP)
            {
                r += sFx[n][j*17 + i]*fx;
                r -= sFx[n][j*17 + i + 1]*fx;
                r += sFy[n][j*16 + i]*fy;
                r -= sFy[n][(j + 1)*16 + i]*fy;
                r += zlo[n]*fz;
                r -= zhi[n]*fz;
            }
            else
            {
                r += own_reg ? px[n] : sFx[n][j*17 + i];
                r -= sFx[n][j*17 + i + 1];
                r += own_reg ? py[n] : sFy[n][j*16 + i];
                r -= sFy[n][(j + 1)*16 + i];
                r += zlo[n];
                r -= zhi[n];
            }
            // streamed: nothing in this launch reads it again
            if (!STG || A.store_rhs) __builtin_nontemporal_store(r, ro + n*sv);
            rr[n] = r;
            zlo[n] = zhi[n];
        }
        if constexpr (STG)
        {
            // the stage of cell (i,j,k): its stored primitives are the window's
            const long long co = own_off + k*plane;
            double qq[5] = {cur.p, cur.T, cur.u, cur.v, cur.w}, cons[5];
            p2c(A.gamma, A.R, qq, cons);
            #pragma unroll
            for (int t = 0; t < 2; ++t)
            {
                if (t + 1 >= A.n || A.c[t] == 0.0) continue;
                double inc[5];
                #pragma unroll
                for (int v = 0; v < 5; ++v) inc[v] = A.c[t]*rprev[t][v];
                #pragma unroll
                for (int v = 0; v < 5; ++v) cons[v] += inc[v];
            }
            const double cl = A.c[A.n - 1];
            if (cl != 0.0)
            {
                double inc[5];
                #pragma unroll
                for (int v = 0; v < 5; ++v) inc[v] = cl*rr[v];
                #pragma unroll
                for (int v = 0; v < 5; ++v) cons[v] += inc[v];
            }
            c2p(A.gamma, A.R, cons, qq);
            #pragma unroll
            for (int v = 0; v < 5; ++v) __builtin_nontemporal_store(qq[v], A.qout + co + v*sv);
            if (A.push_ghosts)
            {
                const CellIt c{i, j, k, lb};
                push_ghosts(G, A.push, A.qout, c, qq, true);
                stage_x(G, xs, c, qq);
            }
        }

        if (k + 1 < k1)
        {
            // plane k+2 (<= k1) enters into the ring slot of plane k-1 and the
            // parity buffer of plane k; its raw registers take plane k+4
            double* s_new = ring[0];
            ring[0] = ring[1]; ring[1] = ring[2]; ring[2] = s_new;
            double* t_new = sS[k & 1];
            if (INC)
            {
                #pragma unroll
                for (int n = 0; n < 5; ++n) rin[n] = ro[plane + n*sv];
            }
            const CellV c2 = cell(nr);
            put_vel(s_new, pown, c2); put_sca(t_new, pown, c2);
            if (k + 4 <= k1) nr = load_own(k + 4);
            if (has_halo)
            {
                const CellV h = cell(hr);
                put_vel(s_new, phalo, h); put_sca(t_new, phalo, h);
                if (k + 4 <= k1) hr = load_halo(k + 4);
            }
        }
        lds_barrier();
        return true;
    };

    for (int k = k0; k < k1; k += 2)
    {
        if (!plane_step(k, nA, hA)) return;
        if (k + 1 < k1 && !plane_step(k + 1, nB, hB)) return;
    }
    if (STG && A.push_ghosts) flush_x(G, A.push, A.qout, xs, lb, k1 - 1);   // after the last plane's closing barrier
    if (tid == 0) *my_start = nz;
}

// Workgroups per block (1, 2 or 4), each marching over an equal share of the
// block's planes, chosen by rounds of resident workgroups (two per CU):
// ceil(zs nblk / slots) rounds, each costing 1/zs of a block's march plus a
// window prologue of about 0.1 of it.  A share must hold an even number of
// planes.  SPD_FAST_ZSPLIT=1|2|4 forces the choice.
int fast_zsplit(const Grid& g)
{
    const int nz = g.geom.nc[2];
    auto ok = [&](int zs) { return nz % (2*zs) == 0; };
    if (const char* e = getenv("SPD_FAST_ZSPLIT"))
    {
        const int zs = e[0] == '4' ? 4 : (e[0] == '2' ? 2 : 1);
        return ok(zs) ? zs : 1;
    }
    static int cus[64] = {};
    const int dev = g.ctx->device;
    if (dev < 0 || dev >= 64) return 1;
    if (!cus[dev]) SPD_HIP(hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev));
    const long slots = 2L*(cus[dev] > 0 ? cus[dev] : 1), nb = g.geom.nblk;
    int best = 1;
    double best_cost = 1e300;
    for (int zs = 1; zs <= kFixupParts; zs *= 2)
    {
        if (!ok(zs)) continue;
        const double rounds = (double)((zs*nb + slots - 1)/slots);
        const double cost = rounds*(1.0/zs + (zs > 1 ? 0.1 : 0.0));
        if (cost < best_cost - 1e-9) { best_cost = cost; best = zs; }
    }
    return best;
}

// Workgroup order of a launch: contiguous runs of blocks per XCD (1) or the
// launch order (0).  Runs pay where other workgroups read the same lines:
// a block that reads its halo from the neighbour blocks' interiors (fused
// stages on periodic identity grids: on the 256^3 Taylor-Green step they
// halve the lines those stages fetch into L2, 2.97 -> 1.54 GB per launch, and
// save 2% of the step), or a block split between workgroups, whose parts
// re-read the planes next to their shares (the 192x128x96 channel flux_div,
// 4 parts per block: 1-3.5% faster in alternating runs on one GPU).  A whole block that reads its own ghost
// cells shares nothing with its neighbours, and flux_div times the same
// either way.  SPD_FAST_XCD=0|1 forces the choice.
int fast_xcd_runs(bool shared_reads)
{
    if (const char* e = getenv("SPD_FAST_XCD")) return e[0] == '1' ? 1 : 0;
    return shared_reads ? 1 : 0;
}

template <int KEEP, int PDISS, int VISC>
void launch_fast(const Grid& g, const double* prims, double* rhs, const FastCfg& C, bool increment, hipStream_t s)
{
    // increment mode keeps one workgroup per block: the fix-up redoes a block
    // from its first flagged plane on, which would add the other half's
    // already-written planes twice
    const int zs = increment ? 1 : fast_zsplit(g);
    const int xr = fast_xcd_runs(zs > 1) ? 0x100 : 0;
    const bool map = !g.geom.identity;
    if (increment && map)
        hipLaunchKernelGGL((k_flux_fast<KEEP, PDISS, VISC, true, true>), dim3(g.geom.nblk*zs), dim3(256), 0, s,
                           g.geom, prims, rhs, C, g.d_start, StageArgs{}, zs | xr);
    else if (increment)
        hipLaunchKernelGGL((k_flux_fast<KEEP, PDISS, VISC, true, false>), dim3(g.geom.nblk*zs), dim3(256), 0, s,
                           g.geom, prims, rhs, C, g.d_start, StageArgs{}, zs | xr);
    else if (map)
        hipLaunchKernelGGL((k_flux_fast<KEEP, PDISS, VISC, false, true>), dim3(g.geom.nblk*zs), dim3(256), 0, s,
                           g.geom, prims, rhs, C, g.d_start, StageArgs{}, zs | xr);
    else
        hipLaunchKernelGGL((k_flux_fast<KEEP, PDISS, VISC, false, false>), dim3(g.geom.nblk*zs), dim3(256), 0, s,
                           g.geom, prims, rhs, C, g.d_start, StageArgs{}, zs | xr);
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

template <int KEEP, int PDISS, int VISC>
void launch_fast_stage(const Grid& g, const double* qin, double* rhs, const FastCfg& C, const StageArgs& A,
                       bool read_neighbours, hipStream_t s)
{
    const int zs = fast_zsplit(g);
    const int xr = fast_xcd_runs(read_neighbours || zs > 1) ? 0x100 : 0;
    if (!g.geom.identity && read_neighbours)
        hipLaunchKernelGGL((k_flux_fast<KEEP, PDISS, VISC, false, true, 2>), dim3(g.geom.nblk*zs), dim3(256), 0, s,
                           g.geom, qin, rhs, C, g.d_start, A, zs | xr);
    else if (!g.geom.identity)
        hipLaunchKernelGGL((k_flux_fast<KEEP, PDISS, VISC, false, true, 1>), dim3(g.geom.nblk*zs), dim3(256), 0, s,
                           g.geom, qin, rhs, C, g.d_start, A, zs | xr);
    else if (read_neighbours)
        hipLaunchKernelGGL((k_flux_fast<KEEP, PDISS, VISC, false, false, 2>), dim3(g.geom.nblk*zs), dim3(256), 0, s,
                           g.geom, qin, rhs, C, g.d_start, A, zs | xr);
    else
        hipLaunchKernelGGL((k_flux_fast<KEEP, PDISS, VISC, false, false, 1>), dim3(g.geom.nblk*zs), dim3(256), 0, s,
                           g.geom, qin, rhs, C, g.d_start, A, zs | xr);
    SPD_LAUNCH_CHECK();
}

template <int KEEP, int PDISS>
bool dispatch_visc_stage(int visc, const Grid& g, const double* qin, double* rhs, const FastCfg& C,
                         const StageArgs& A, bool nb, hipStream_t s)
{
    switch (visc)
    {
        case 0: if (KEEP || PDISS) { launch_fast_stage<KEEP, PDISS, 0>(g, qin, rhs, C, A, nb, s); return true; } return false;
        case 1: launch_fast_stage<KEEP, PDISS, 1>(g, qin, rhs, C, A, nb, s); return true;
        case 2: launch_fast_stage<KEEP, PDISS, 2>(g, qin, rhs, C, A, nb, s); return true;
    }
    return false;
}

} // namespace

namespace {

// Selects the fast kernel for (grid, flux); false when the generic kernel must run.
bool select_fast(const DevGeom& G, const spd_flux_desc& f, int& keep, int& pd, int& vs, FastCfg& C)
{
    // identity or diagonal (tanh) maps; a 16 x 16 cross-section
    if (G.nc[0] != 16 || G.nc[1] != 16 || G.ng < 1) return false;
    for (int a = 0; a < 3; ++a) if (G.map_kind[a] != SPD_MAP_IDENTITY && G.map_kind[a] != SPD_MAP_TANH) return false;
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
            if (t.sgs != SPD_SGS_NONE) return false;   // the per-cell kernels evaluate subgrid models
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
    bool ok = false;
    switch (keep*3 + pd)
    {
        case 0: ok = dispatch_visc<0, 0>(vs, g, prims, rhs, C, increment, s); break;
        case 1: ok = dispatch_visc<0, 1>(vs, g, prims, rhs, C, increment, s); break;
        case 2: ok = dispatch_visc<0, 2>(vs, g, prims, rhs, C, increment, s); break;
        case 3: ok = dispatch_visc<1, 0>(vs, g, prims, rhs, C, increment, s); break;
        case 4: ok = dispatch_visc<1, 1>(vs, g, prims, rhs, C, increment, s); break;
        case 5: ok = dispatch_visc<1, 2>(vs, g, prims, rhs, C, increment, s); break;
    }
    // power law: blocks with a face temperature outside the pow series range
    // are completed by the per-cell kernel (an early exit per block when none)
    if (ok && vs == 1 + SPD_LAW_POWER) launch_flux_div_fixup(g, prims, rhs, f, increment, s);
    return ok;
}

// The fused stage needs the x-staged push (16 x 16 cross-sections, ng <= 2,
// nz >= 2 ng) beside the fast kernel's own conditions
static bool stage_geometry(const DevGeom& G)
{
    return G.nc[0] == 16 && G.nc[1] == 16 && G.ng >= 1 && G.ng <= 2 && G.nc[2] >= 2*G.ng;
}

// A stage can read its halo from the neighbour blocks' interiors when every
// ghost region has a source block (each axis periodic) or, on diagonal
// (mapped) coordinates, along y only is an isothermal wall the caller fills on
// both faces (wall_mask bits 2 and 3), whose ghosts the kernel forms from the
// mirrored cells (the walled channel; the identity-coordinate kernel is built
// without the wall step)
bool stage_reads_neighbours_ok(const Grid& g, int wall_mask)
{
    const bool xz = g.desc.periodic[0] && g.desc.periodic[2];
    if (g.geom.identity) return xz && g.desc.periodic[1];
    return xz && (g.desc.periodic[1] || ((wall_mask >> 2) & 3) == 3);
}

bool flux_stage_fast_eligible(const Grid& g, const spd_flux_desc& f)
{
    int keep, pd, vs;
    FastCfg C;
    if (!stage_geometry(g.geom) || !select_fast(g.geom, f, keep, pd, vs, C)) return false;
    return keep || pd || vs;
}

bool launch_flux_stage_fast(const Grid& g, const double* qin, double* qout, double* rhs, const spd_flux_desc& f,
                            const double* const* rprev, const double* coef, int nr, bool store_rhs,
                            double gamma, double R, hipStream_t s, bool read_neighbours, bool push,
                            int wall_mask, double t_wall)
{
    int keep, pd, vs;
    FastCfg C;
    if (!stage_geometry(g.geom) || !select_fast(g.geom, f, keep, pd, vs, C)) return false;
    if (nr < 1 || nr > 3) throw ArgError{"a fused stage combines 1..3 residuals"};
    if (qin == qout) throw ArgError{"a fused stage needs distinct input and output arrays"};
    if (read_neighbours && !stage_reads_neighbours_ok(g, wall_mask))
        throw ArgError{"neighbour-read stages need every axis periodic, or on mapped coordinates y walled on both faces"};
    StageArgs A{};
    A.qout = qout;
    for (int t = 0; t + 1 < nr; ++t) A.r[t] = rprev[t];
    for (int t = 0; t < nr; ++t) A.c[t] = coef[t];
    A.n = nr;
    A.store_rhs = store_rhs ? 1 : 0;
    A.gamma = gamma; A.R = R;
    A.push = g.d_push;
    A.push_ghosts = push ? 1 : 0;
    A.nbr = g.d_nbr;
    A.wall_y = (read_neighbours && !g.desc.periodic[1]) ? 1 : 0;
    A.t_wall = t_wall;
    bool ok = false;
    switch (keep*3 + pd)
    {
        case 0: ok = dispatch_visc_stage<0, 0>(vs, g, qin, rhs, C, A, read_neighbours, s); break;
        case 1: ok = dispatch_visc_stage<0, 1>(vs, g, qin, rhs, C, A, read_neighbours, s); break;
        case 2: ok = dispatch_visc_stage<0, 2>(vs, g, qin, rhs, C, A, read_neighbours, s); break;
        case 3: ok = dispatch_visc_stage<1, 0>(vs, g, qin, rhs, C, A, read_neighbours, s); break;
        case 4: ok = dispatch_visc_stage<1, 1>(vs, g, qin, rhs, C, A, read_neighbours, s); break;
        case 5: ok = dispatch_visc_stage<1, 2>(vs, g, qin, rhs, C, A, read_neighbours, s); break;
    }
    if (!ok) return false;
    // power law: blocks with a face temperature outside the pow series range
    // were left from that plane on; their residual (into rhs) and stage
    // (qin -> qout) are completed here (early exits per block when none)
    if (vs == 1 + SPD_LAW_POWER)
    {
        // the per-cell fix-up reads its block's ghost cells: after a
        // neighbour-read stage they may be stale, so the flagged blocks pull
        // theirs first (what an exchange copies; the others exit at once)
        // (and, on the walled channel, fill their wall ghost slabs from the
        // cells just pulled).  One launch does all of it for the compositions
        // with a specialised per-cell kernel (every one the fast kernel
        // takes); SPD_SPLIT_FIXUP=1 keeps the three launches
        const double* r[3] = {rprev && nr > 1 ? rprev[0] : nullptr, rprev && nr > 2 ? rprev[1] : nullptr, nullptr};
        r[nr - 1] = rhs;
        bool merged = false;
        if (const char* e = getenv("SPD_SPLIT_FIXUP"); !(e && e[0] == '1'))
            merged = launch_stage_fixup_spec(g, const_cast<double*>(qin), qout, rhs, f, r, coef, nr, gamma, R,
                                             read_neighbours, A.wall_y, t_wall, s);
        if (!merged)
        {
            if (read_neighbours) launch_exchange_flagged(g, const_cast<double*>(qin), s, A.wall_y, t_wall);
            launch_flux_div_fixup(g, qin, rhs, f, false, s);
            launch_stage_fixup(g, qin, qout, r, coef, nr, gamma, R, s);
        }
    }
    return true;
}

} // namespace spd
