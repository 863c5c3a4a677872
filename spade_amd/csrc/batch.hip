// Batched FP32-flux / FP64-state stepping of many independent periodic cases.
//
// Design (the "chosen form" beside the FP64 parity path):
//   * every case is a set of 16^3 blocks; a block's conserved state
//     (rho, rhoE, rho u, rho v, rho w) lives in FP64, SoA, without ghost
//     layers, in a slot of a device pool; halos are read from the neighbour
//     blocks' interiors (per-slot 27-entry neighbour table, periodic wrap);
//   * one persistent kernel per SSP-RK3 stage walks a work list of slots of
//     all active cases; per block it marches over k with the plane window in
//     LDS (as the FP64 kernel does), evaluates every face flux once in FP32,
//     sums the residual in FP64 and applies the stage combination in FP64:
//       u_out = alpha u0 + beta (u_in + dt L(u_in))   (Shu-Osher SSP-RK3)
//   * FP32 only carries perturbation fluxes: the face states are split about
//     each case's reference state (rho_ref, p_ref, T_ref, e_ref = cv T_ref);
//     the reference parts of the KEEP fluxes reduce to
//       d(rho)  += rho_ref D(U),   d(rhoE) += e_ref d(rho) + p_ref D(U)
//     with D(U) = -div u evaluated per cell in FP64 from FP64 velocities, and
//     the constant p_ref n momentum flux cancels exactly;
//   * kinetic energy and enstrophy of the stage-1 input state are summed per
//     block (fixed-order tree) for the per-case diagnostics history.
#include "spd_internal.hpp"
#include "batch.hpp"
#include <cmath>
#include <algorithm>

namespace spd {
namespace {

constexpr int HXB = 18;


struct CellF { float p, T, u, v, w, rho, e; };   // p, T, rho, e are perturbations

__device__ inline int lposb(int i, int j) { return (j + 1)*HXB + (i + 1); }

template <int D> __device__ inline float velf(const CellF& c) { return D == 0 ? c.u : (D == 1 ? c.v : c.w); }

// FP32 face flux of cent_keep<2> + visc_lr about the reference state.
// Outputs (f32): continuity perturbation, energy perturbation, x/y/z momentum.
// The fused multiply-adds are written out and the compiler forms no others,
// so every inlined copy of a face (interior faces, the 17th x- and y-faces of
// a tile, the z-faces) rounds the same face identically: a face value is the
// same for the two cells it separates whichever code evaluated it, the FP64
// residual sums telescope over a periodic case, and every launch shape gives
// the same bits.
template <int D, int LAW>
__device__ inline void face_f32(const CaseParams& P, const CellF& L, const CellF& R, const float g[3][3], float gT,
                                float F[5])
{
    #pragma clang fp contract(off)
    const float unL = velf<D>(L), unR = velf<D>(R);
    const float U   = 0.5f*(unL + unR);
    const float Fcp = (0.5f*(L.rho + R.rho))*U;              // rho'-part of the mass flux
    const float Fc  = __builtin_fmaf(P.f_rho_ref, U, Fcp);    // full mass flux
    const float ub[3] = {0.5f*(L.u + R.u), 0.5f*(L.v + R.v), 0.5f*(L.w + R.w)};
    const float K   = __builtin_fmaf(L.w, R.w, __builtin_fmaf(L.v, R.v, L.u*R.u));
    const float eb  = 0.5f*(L.e + R.e);
    float Fm[3] = {Fc*ub[0], Fc*ub[1], Fc*ub[2]};
    Fm[D] = __builtin_fmaf(0.5f, L.p + R.p, Fm[D]);
    const float FE = __builtin_fmaf(0.5f*Fc, K, __builtin_fmaf(Fc, eb, 0.5f*__builtin_fmaf(unL, R.p, unR*L.p)));
    // viscous terms
    const float Tf = __builtin_fmaf(0.5f, L.T + R.T, P.f_T_ref);
    float mu;
    // (T/T_ref)^n as exp2(n log2(T/T_ref)) on the hardware log2 / exp2 (about
    // 1e-7 relative for T near T_ref; OCML's powf costs ~120 instructions)
    if (LAW == SPD_LAW_POWER)
        mu = P.f_mu_ref*__builtin_amdgcn_exp2f(P.f_power*__builtin_amdgcn_logf(Tf*P.f_inv_T_law));
    else
        mu = P.f_mu_ref;
    const float visc2 = -0.6666666666666667f*mu;
    const float cond  = mu*P.f_cp_over_pr;
    const float div = (g[0][0] + g[1][1]) + g[2][2];
    float t[3];
    #pragma unroll
    for (int b = 0; b < 3; ++b)
        t[b] = (b == D) ? __builtin_fmaf(mu, 2.0f*g[D][D], visc2*div) : mu*(g[D][b] + g[b][D]);
    const float ht = __builtin_fmaf(cond, gT, __builtin_fmaf(ub[2], t[2], __builtin_fmaf(ub[1], t[1], ub[0]*t[0])));
    F[0] = Fcp;
    F[1] = FE - ht;
    F[2] = Fm[0] - t[0];
    F[3] = Fm[1] - t[1];
    F[4] = Fm[2] - t[2];
}

__device__ inline float tangf(float a, float b, float c, float d, float it)
{
    return (0.25f*it)*((a - b) + (c - d));
}

// slot layout of the f32 velocity ring: s[pos] = u, s[HPB + pos] = v, s[2*HPB + pos] = w
template <int HPB>
__device__ inline void gradsf_x(const float* sm, const float* s0, const float* sp, int pL, int pR,
                                const CellF& L, const CellF& R, float ix, float iy, float iz, float g[3][3], float& gT)
{
    g[0][0] = (R.u - L.u)*ix; g[0][1] = (R.v - L.v)*ix; g[0][2] = (R.w - L.w)*ix; gT = (R.T - L.T)*ix;
    g[1][0] = tangf(s0[pL + HXB], s0[pL - HXB], s0[pR + HXB], s0[pR - HXB], iy);
    g[1][1] = tangf(s0[HPB + pL + HXB], s0[HPB + pL - HXB], s0[HPB + pR + HXB], s0[HPB + pR - HXB], iy);
    g[1][2] = 0.0f;
    g[2][0] = tangf(sp[pL], sm[pL], sp[pR], sm[pR], iz);
    g[2][1] = 0.0f;
    g[2][2] = tangf(sp[2*HPB + pL], sm[2*HPB + pL], sp[2*HPB + pR], sm[2*HPB + pR], iz);
}

template <int HPB>
__device__ inline void gradsf_y(const float* sm, const float* s0, const float* sp, int pL, int pR,
                                const CellF& L, const CellF& R, float ix, float iy, float iz, float g[3][3], float& gT)
{
    g[1][0] = (R.u - L.u)*iy; g[1][1] = (R.v - L.v)*iy; g[1][2] = (R.w - L.w)*iy; gT = (R.T - L.T)*iy;
    g[2][0] = 0.0f;
    g[2][1] = tangf(sp[HPB + pL], sm[HPB + pL], sp[HPB + pR], sm[HPB + pR], iz);
    g[2][2] = tangf(sp[2*HPB + pL], sm[2*HPB + pL], sp[2*HPB + pR], sm[2*HPB + pR], iz);
    g[0][0] = tangf(s0[pL + 1], s0[pL - 1], s0[pR + 1], s0[pR - 1], ix);
    g[0][1] = tangf(s0[HPB + pL + 1], s0[HPB + pL - 1], s0[HPB + pR + 1], s0[HPB + pR - 1], ix);
    g[0][2] = 0.0f;
}

template <int HPB>
__device__ inline void gradsf_z(const float* sL, const float* sR, int p, const CellF& L, const CellF& R,
                                float ix, float iy, float iz, float g[3][3], float& gT)
{
    g[2][0] = (R.u - L.u)*iz; g[2][1] = (R.v - L.v)*iz; g[2][2] = (R.w - L.w)*iz; gT = (R.T - L.T)*iz;
    g[0][0] = tangf(sL[p + 1], sL[p - 1], sR[p + 1], sR[p - 1], ix);
    g[0][1] = 0.0f;
    g[0][2] = tangf(sL[2*HPB + p + 1], sL[2*HPB + p - 1], sR[2*HPB + p + 1], sR[2*HPB + p - 1], ix);
    g[1][0] = 0.0f;
    g[1][1] = tangf(sL[HPB + p + HXB], sL[HPB + p - HXB], sR[HPB + p + HXB], sR[HPB + p - HXB], iy);
    g[1][2] = tangf(sL[2*HPB + p + HXB], sL[2*HPB + p - HXB], sR[2*HPB + p + HXB], sR[2*HPB + p - HXB], iy);
}

struct Cons { double r, E, mu, mv, mw; };
struct PrimD { double rho, u, v, w, p, T; };

__device__ inline PrimD to_prim(const CaseParams& P, const Cons& c)
{
    PrimD q;
    q.rho = c.r;
    // 1/rho: hardware reciprocal and two Newton steps (within an ulp or so;
    // this engine's contract is the FP32-flux tolerance, not bit parity)
    double inv = __builtin_amdgcn_rcp(c.r);
    inv = __builtin_fma(inv, __builtin_fma(-c.r, inv, 1.0), inv);
    inv = __builtin_fma(inv, __builtin_fma(-c.r, inv, 1.0), inv);
    q.u = c.mu*inv; q.v = c.mv*inv; q.w = c.mw*inv;
    q.p = P.gm1*(c.E - 0.5*c.r*(q.u*q.u + q.v*q.v + q.w*q.w));
    q.T = q.p*inv*P.inv_R;
    return q;
}

__device__ inline CellF to_f32(const CaseParams& P, const PrimD& q)
{
    const double Tp = q.T - P.T_ref;
    return CellF{(float)(q.p - P.p_ref), (float)Tp, (float)q.u, (float)q.v, (float)q.w,
                 (float)(q.rho - P.rho_ref), (float)(P.cv*Tp)};
}

__device__ inline double uniform_d(double x)
{
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | lo));
}
__device__ inline float uniform_f(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }

// a wave-uniform copy of a case's constants (every lane reads the same entry)
__device__ inline CaseParams uniform_params(const CaseParams& c)
{
    CaseParams u;
    u.dt = uniform_d(c.dt); u.ix = uniform_d(c.ix); u.iy = uniform_d(c.iy); u.iz = uniform_d(c.iz);
    u.gamma = uniform_d(c.gamma); u.gm1 = uniform_d(c.gm1); u.R = uniform_d(c.R); u.inv_R = uniform_d(c.inv_R);
    u.cv = uniform_d(c.cv);
    u.rho_ref = uniform_d(c.rho_ref); u.p_ref = uniform_d(c.p_ref); u.T_ref = uniform_d(c.T_ref);
    u.e_ref = uniform_d(c.e_ref);
    u.f_rho_ref = uniform_f(c.f_rho_ref); u.f_T_ref = uniform_f(c.f_T_ref); u.f_mu_ref = uniform_f(c.f_mu_ref);
    u.f_inv_T_law = uniform_f(c.f_inv_T_law); u.f_power = uniform_f(c.f_power);
    u.f_cp_over_pr = uniform_f(c.f_cp_over_pr);
    u.law = __builtin_amdgcn_readfirstlane(c.law);
    return u;
}

// Launch shapes the autotuner chooses between:
//   TJ   rows of the block's 16x16 cross-section per workgroup (16: one
//        256-thread workgroup per block; 8: two 128-thread workgroups, one per
//        half, each with its own halo rows: more, smaller workgroups per CU);
//   PFB  planes of conserved state in flight ahead of the one entering the
//        LDS window (1 or 2);
//   OCC  waves per SIMD the register budget is sized for (3: 168 VGPRs with
//        56-172 bytes per lane spilled, 2: up to 256 VGPRs without); a
//        256-thread workgroup holds 48.5 KB of LDS and a 128-thread one
//        26.5 KB, so three or six of them fit a CU.
// Every shape evaluates the same expressions in the same order, and the
// per-block diagnostics are summed over the same two half-block trees (rows
// 0-7 and 8-15, 128 leaves each, added on the host), so all shapes give the
// same bits.
template <int TJ> struct Tile
{
    static constexpr int NT = 16*TJ;                     // threads
    static constexpr int HP = HXB*(TJ + 2);              // LDS plane with halo
    static constexpr int NHALO = 2*TJ + 36;              // halo columns
    // threads that evaluate the 17th x-faces (i = 16) and y-faces (j = TJ)
    static constexpr int XB = TJ == 16 ? 128 : 52;
    static constexpr int YB = TJ == 16 ? 192 : 64;
};

template <int LAW, bool HAS_U0, int TJ, int PFB, int OCC>
__global__ void __launch_bounds__(16*TJ, OCC)
k_stage(BatchDev B, const double* __restrict__ Qin, const double* __restrict__ Q0, double* __restrict__ Qout,
        double alpha, double beta, int flags)
{
    using T = Tile<TJ>;
    // flags: bit 0 per-block diagnostics, bit 1 work items in XCD runs
    const bool want_diag = flags & 1;
    constexpr int HP = T::HP, NTILE = 16/TJ;
    __shared__ float  sU[3][3*HP];        // u, v, w (f32) of planes k-1, k, k+1
    __shared__ float  sS[2][4*HP];        // p', T', rho', e' of planes k, k+1 (parity)
    __shared__ double sV[2][3*HP];        // u, v, w (f64) of planes k, k+1 (parity)
    // face fluxes as computed (FP32); the cells widen them to FP64 when they
    // read them, which is exact, so every shape sums the same doubles
    __shared__ alignas(16) float sFx[5][TJ*17];
    __shared__ float sFy[5][(TJ + 1)*16];
    // the per-block diagnostics tree reuses the x-face buffer after the
    // k-march (the march's last barrier has passed, and the next work item
    // writes face fluxes only after its prologue barrier); the window plus
    // the fluxes then fit three 256-thread or six 128-thread workgroups per CU
    static_assert(sizeof(double)*2*T::NT <= sizeof(sFx), "diagnostics tree exceeds the x-face buffer");
    double (*sRed)[T::NT] = reinterpret_cast<double (*)[T::NT]>(&sFx[0][0]);

    const int tid = threadIdx.x;
    const int i = tid & 15, j = tid >> 4;                // j: row within the tile
    int hi = 0, hj = 0;
    const bool has_halo = tid < T::NHALO;
    if (tid < TJ)             { hi = -1;              hj = tid; }
    else if (tid < 2*TJ)      { hi = 16;              hj = tid - TJ; }
    else if (tid < 2*TJ + 16) { hi = tid - 2*TJ;      hj = -1; }
    else if (tid < 2*TJ + 32) { hi = tid - 2*TJ - 16; hj = TJ; }
    else if (tid < T::NHALO)  { const int c = tid - 2*TJ - 32; hi = (c & 1) ? 16 : -1; hj = (c & 2) ? TJ : -1; }
    const int pown = lposb(i, j), phalo = lposb(hi, hj);
    const int nitems = *B.nwork*NTILE;

    // XCD runs: the workgroups on one XCD take consecutive work items of each
    // round (neighbour blocks of a case, in case-major order) instead of every
    // eighth; every item is still taken once
    const int w0 = (flags & 2) ? xcd_run_order(blockIdx.x, gridDim.x) : blockIdx.x;
    for (int w = w0; w < nitems; w += gridDim.x)
    {
        // slot and case constants are the same for the whole workgroup: held
        // in scalar registers, not reloaded per lane inside the k-march
        const int slot = __builtin_amdgcn_readfirstlane(B.worklist[w/NTILE]);
        const int j0 = (w % NTILE)*TJ;                   // first row of the tile in the block
        const CaseParams P = uniform_params(B.cases[__builtin_amdgcn_readfirstlane(B.slot_case[slot])]);
        const int* nb = B.nbr + 27*slot;
        const float ix = (float)P.ix, iy = (float)P.iy, iz = (float)P.iz;
        const int gj = j0 + j, ghj = j0 + hj;            // rows in the block (may be -1 or 16)
        // the slots this thread's two columns (own, halo) read from below the
        // block, inside it and above it: looked up once per tile
        auto column_slots = [&](int ci, int cj, int out[3])
        {
            const int exy = (ci < 0 ? 0 : (ci > 15 ? 2 : 1)) + 3*(cj < 0 ? 0 : (cj > 15 ? 2 : 1));
            out[0] = nb[exy]; out[1] = nb[exy + 9]; out[2] = nb[exy + 18];
        };
        int own_slot[3], halo_slot[3];
        column_slots(i, gj, own_slot);
        column_slots(hi, ghj, halo_slot);

        auto load_col = [&](const int sl[3], int ci, int cj, int ck) -> Cons
        {
            const int s = ck < 0 ? sl[0] : (ck > 15 ? sl[2] : sl[1]);
            const long long base = (long long)s*5*4096 + (((ck & 15)*16 + (cj & 15))*16 + (ci & 15));
            return Cons{Qin[base], Qin[base + 4096], Qin[base + 2*4096], Qin[base + 3*4096], Qin[base + 4*4096]};
        };
        auto slotk = [](int k) { return (k + 3) % 3; };
        auto put = [&](int k, int p, const PrimD& q, const CellF& c)
        {
            float* s = sU[slotk(k)];
            s[p] = c.u; s[HP + p] = c.v; s[2*HP + p] = c.w;
            float* t = sS[k & 1];
            t[p] = c.p; t[HP + p] = c.T; t[2*HP + p] = c.rho; t[3*HP + p] = c.e;
            double* d = sV[k & 1];
            d[p] = q.u; d[HP + p] = q.v; d[2*HP + p] = q.w;
        };
        auto from_lds = [&](const float* t, const float* s0, int p) -> CellF
        {
            return CellF{t[p], t[HP + p], s0[p], s0[HP + p], s0[2*HP + p], t[2*HP + p], t[3*HP + p]};
        };

        // u0 of the plane to be written next (stages 2 and 3), requested before
        // the state prefetch so that waiting for it never waits for the prefetch
        double u0[5];
        auto load_u0 = [&](int k)
        {
            const long long o = (long long)slot*5*4096 + (k*16 + gj)*16 + i;
            #pragma unroll
            for (int n = 0; n < 5; ++n) u0[n] = HAS_U0 ? Q0[o + n*4096] : 0.0;
        };
        // the three-wave budget (168 VGPRs) requests u0 of plane k just before
        // the plane's first barrier instead of a plane ahead, so that it is not
        // live across the face evaluation
        constexpr bool late_u0 = OCC == 3;
        if (!late_u0) load_u0(0);

        // prologue: planes -1, 0, 1 resident, planes 2 .. 1 + PFB in flight.
        // The own column's conserved state of planes k and k+1 stays in
        // registers (the stage input u_in); velocities of planes k, k+1 are in LDS.
        const PrimD qprev = to_prim(P, load_col(own_slot, i, gj, -1));
        Cons rcur = load_col(own_slot, i, gj, 0);
        Cons rnxt = load_col(own_slot, i, gj, 1);
        const PrimD qcur0 = to_prim(P, rcur), qnxt0 = to_prim(P, rnxt);
        CellF prev = to_f32(P, qprev), cur = to_f32(P, qcur0), nxt = to_f32(P, qnxt0);
        Cons nraw[PFB], hraw[PFB];
        #pragma unroll
        for (int d = 0; d < PFB; ++d) { nraw[d] = load_col(own_slot, i, gj, 2 + d); hraw[d] = Cons{}; }
        {
            float* s = sU[slotk(-1)];
            s[pown] = prev.u; s[HP + pown] = prev.v; s[2*HP + pown] = prev.w;
        }
        put(0, pown, qcur0, cur); put(1, pown, qnxt0, nxt);
        if (has_halo)
        {
            const CellF hp = to_f32(P, to_prim(P, load_col(halo_slot, hi, ghj, -1)));
            float* s = sU[slotk(-1)];
            s[phalo] = hp.u; s[HP + phalo] = hp.v; s[2*HP + phalo] = hp.w;
            const PrimD h0 = to_prim(P, load_col(halo_slot, hi, ghj, 0)), h1 = to_prim(P, load_col(halo_slot, hi, ghj, 1));
            put(0, phalo, h0, to_f32(P, h0));
            put(1, phalo, h1, to_f32(P, h1));
            #pragma unroll
            for (int d = 0; d < PFB; ++d) hraw[d] = load_col(halo_slot, hi, ghj, 2 + d);
        }
        double ke_acc = 0.0, en_acc = 0.0;
        double wprev_u = qprev.u, wprev_v = qprev.v, wprev_w = qprev.w;   // own column, plane k-1 (f64)
        lds_barrier();

        float zlo[5];                    // low z-face flux of the current plane (widened where used)
        {
            float g[3][3], gT, F[5];
            gradsf_z<HP>(sU[slotk(-1)], sU[slotk(0)], pown, prev, cur, ix, iy, iz, g, gT);
            face_f32<2, LAW>(P, prev, cur, g, gT, F);
            #pragma unroll
            for (int n = 0; n < 5; ++n) zlo[n] = F[n];
        }

        for (int k = 0; k < 16; ++k)
        {
            const long long o = (long long)slot*5*4096 + (k*16 + gj)*16 + i;
            const float* sm = sU[slotk(k - 1)];
            const float* s0 = sU[slotk(k)];
            const float* sp = sU[slotk(k + 1)];
            const float* sk = sS[k & 1];
            {
                const int pL = pown - 1;
                const CellF L = from_lds(sk, s0, pL);
                float g[3][3], gT, F[5];
                gradsf_x<HP>(sm, s0, sp, pL, pown, L, cur, ix, iy, iz, g, gT);
                face_f32<0, LAW>(P, L, cur, g, gT, F);
                for (int n = 0; n < 5; ++n) sFx[n][j*17 + i] = F[n];
            }
            {
                const int pL = pown - HXB;
                const CellF L = from_lds(sk, s0, pL);
                float g[3][3], gT, F[5];
                gradsf_y<HP>(sm, s0, sp, pL, pown, L, cur, ix, iy, iz, g, gT);
                face_f32<1, LAW>(P, L, cur, g, gT, F);
                for (int n = 0; n < 5; ++n) sFy[n][j*16 + i] = F[n];
            }
            if (tid >= T::XB && tid < T::XB + TJ)
            {
                const int jj = tid - T::XB;
                const int pL = lposb(15, jj), pR = lposb(16, jj);
                const CellF L = from_lds(sk, s0, pL), Rr = from_lds(sk, s0, pR);
                float g[3][3], gT, F[5];
                gradsf_x<HP>(sm, s0, sp, pL, pR, L, Rr, ix, iy, iz, g, gT);
                face_f32<0, LAW>(P, L, Rr, g, gT, F);
                for (int n = 0; n < 5; ++n) sFx[n][jj*17 + 16] = F[n];
            }
            if (tid >= T::YB && tid < T::YB + 16)
            {
                const int ii = tid - T::YB;
                const int pL = lposb(ii, TJ - 1), pR = lposb(ii, TJ);
                const CellF L = from_lds(sk, s0, pL), Rr = from_lds(sk, s0, pR);
                float g[3][3], gT, F[5];
                gradsf_y<HP>(sm, s0, sp, pL, pR, L, Rr, ix, iy, iz, g, gT);
                face_f32<1, LAW>(P, L, Rr, g, gT, F);
                for (int n = 0; n < 5; ++n) sFy[n][TJ*16 + ii] = F[n];
            }
            float zhi[5];
            {
                float g[3][3], gT;
                gradsf_z<HP>(s0, sp, pown, cur, nxt, ix, iy, iz, g, gT);
                face_f32<2, LAW>(P, cur, nxt, g, gT, zhi);
            }
            // FP64 velocity differences of the cell (planes k-1..k+1)
            const double* vk = sV[k & 1];
            const double dux = vk[pown + 1] - vk[pown - 1];                       // u(i+1) - u(i-1)
            const double dvy = vk[HP + pown + HXB] - vk[HP + pown - HXB];          // v(j+1) - v(j-1)
            const double* vn = sV[(k + 1) & 1];
            const double dwz = vn[2*HP + pown] - wprev_w;                          // w(k+1) - w(k-1)
            const double DU = -0.5*__builtin_fma(dux, P.ix, __builtin_fma(dvy, P.iy, dwz*P.iz));   // -div u
            if (want_diag)
            {
                const double dwy = vk[2*HP + pown + HXB] - vk[2*HP + pown - HXB];
                const double duy = vk[pown + HXB] - vk[pown - HXB];
                const double dvx = vk[HP + pown + 1] - vk[HP + pown - 1];
                const double dwx = vk[2*HP + pown + 1] - vk[2*HP + pown - 1];
                const double dvz = vn[HP + pown] - wprev_v, duz = vn[pown] - wprev_u;
                const double ox = 0.5*(dwy*P.iy - dvz*P.iz);
                const double oy = 0.5*(duz*P.iz - dwx*P.ix);
                const double oz = 0.5*(dvx*P.ix - duy*P.iy);
                const double uc = vk[pown], vc = vk[HP + pown], wc = vk[2*HP + pown];
                ke_acc += 0.5*rcur.r*(uc*uc + vc*vc + wc*wc);
                en_acc += 0.5*(ox*ox + oy*oy + oz*oz);
            }
            if (late_u0) load_u0(k);
            lds_barrier();

            // FP64 residual sum and stage combination for cell (i,j,k); the
            // difference of two FP32 face fluxes is exact in FP64, so each
            // direction is differenced first and scaled once (fused
            // multiply-adds written out, the same in every launch shape)
            double r[5];
            #pragma unroll
            for (int n = 0; n < 5; ++n)
            {
                const double zh = (double)zhi[n];
                r[n] = __builtin_fma((double)sFx[n][j*17 + i] - (double)sFx[n][j*17 + i + 1], P.ix,
                       __builtin_fma((double)sFy[n][j*16 + i] - (double)sFy[n][(j + 1)*16 + i], P.iy,
                                     ((double)zlo[n] - zh)*P.iz));
                zlo[n] = zhi[n];
            }
            const double d_rho = __builtin_fma(P.rho_ref, DU, r[0]);
            const double d_E   = __builtin_fma(P.p_ref, DU, __builtin_fma(P.e_ref, d_rho, r[1]));
            const double res[5] = {d_rho, d_E, r[2], r[3], r[4]};
            #pragma unroll
            for (int n = 0; n < 5; ++n)
            {
                const double uin = n == 0 ? rcur.r : n == 1 ? rcur.E : n == 2 ? rcur.mu : n == 3 ? rcur.mv : rcur.mw;
                const double stage = __builtin_fma(P.dt, res[n], uin);
                __builtin_nontemporal_store(HAS_U0 ? __builtin_fma(alpha, u0[n], beta*stage) : beta*stage,
                                            Qout + o + n*4096);
            }

            // advance the window
            wprev_u = vk[pown]; wprev_v = vk[HP + pown]; wprev_w = vk[2*HP + pown];
            if (k + 1 < 16)
            {
                if (!late_u0) load_u0(k + 1);
                rcur = rnxt; rnxt = nraw[0]; cur = nxt;
                const PrimD q2 = to_prim(P, rnxt);
                nxt = to_f32(P, q2);
                put(k + 2, pown, q2, nxt);
                #pragma unroll
                for (int d = 0; d + 1 < PFB; ++d) nraw[d] = nraw[d + 1];
                if (k + 2 + PFB <= 16) nraw[PFB - 1] = load_col(own_slot, i, gj, k + 2 + PFB);
                if (has_halo)
                {
                    const PrimD hq = to_prim(P, hraw[0]);
                    put(k + 2, phalo, hq, to_f32(P, hq));
                    #pragma unroll
                    for (int d = 0; d + 1 < PFB; ++d) hraw[d] = hraw[d + 1];
                    if (k + 2 + PFB <= 16) hraw[PFB - 1] = load_col(halo_slot, hi, ghj, k + 2 + PFB);
                }
            }
            lds_barrier();
        }
        if (want_diag)
        {
            // one 128-leaf tree per half block (rows 0-7, 8-15), the same in every shape
            sRed[0][tid] = ke_acc; sRed[1][tid] = en_acc;
            lds_barrier();
            for (int s = 64; s > 0; s >>= 1)
            {
                if ((tid & 127) < s) { sRed[0][tid] += sRed[0][tid + s]; sRed[1][tid] += sRed[1][tid + s]; }
                lds_barrier();
            }
            if ((tid & 127) == 0)
            {
                const int half = (j0 + j) >> 3;
                const int h = *B.step % B.hist_len;
                B.diag[(((long long)h*B.max_slots + slot)*2 + half)*2 + 0] = sRed[0][tid];
                B.diag[(((long long)h*B.max_slots + slot)*2 + half)*2 + 1] = sRed[1][tid];
            }
            lds_barrier();
        }
    }
}

__global__ void k_step_count(int* step) { if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1; }

// Initial conditions for the blocks of one case (TGV or isentropic vortex), in
// conserved FP64 variables.
__global__ void __launch_bounds__(256)
k_init_case(const int* __restrict__ stage, int nslots, CaseParams P, CaseIC ic, double* __restrict__ Q,
            int* __restrict__ nbr, int* __restrict__ slot_case)
{
    // stage: per block of the case, {slot, case table, 27 neighbour slots}
    const int b = blockIdx.y;
    if (b >= nslots) return;
    const int* e = stage + 29*b;
    const int slot = e[0];
    if (blockIdx.x == 0 && threadIdx.x < 27) nbr[27*(long long)slot + threadIdx.x] = e[2 + threadIdx.x];
    if (blockIdx.x == 0 && threadIdx.x == 27) slot_case[slot] = e[1];
    const int bx = b % ic.nb[0], by = (b / ic.nb[0]) % ic.nb[1], bz = b / (ic.nb[0]*ic.nb[1]);
    for (int c = threadIdx.x + blockIdx.x*blockDim.x; c < 4096; c += blockDim.x*gridDim.x)
    {
        const int i = c & 15, j = (c >> 4) & 15, k = c >> 8;
        const double x = (bx*16 + i + 0.5)*ic.dx[0];
        const double y = (by*16 + j + 0.5)*ic.dx[1];
        const double z = (bz*16 + k + 0.5)*ic.dx[2];
        double p, T, u, v, w;
        if (ic.kind == SPD_CASE_TGV)
        {
            p = ic.p0 + (ic.rho0*ic.u0*ic.u0/16.0)*(cos(2.0*x) + cos(2.0*y))*(cos(2.0*z) + 2.0);
            T = ic.T0;
            u =  ic.u0*sin(x)*cos(y)*cos(z);
            v = -ic.u0*cos(x)*sin(y)*cos(z);
            w = 0.0;
        }
        else
        {
            const double xr = x - ic.xc, yr = y - ic.yc;
            const double r2 = xr*xr + yr*yr;
            const double f = exp(0.5*(1.0 - r2));
            u = ic.u0 - ic.beta/(2.0*M_PI)*yr*f;
            v = ic.beta/(2.0*M_PI)*xr*f;
            w = 0.0;
            T = ic.T0 - ic.dT*exp(1.0 - r2);
            const double rho = ic.rho0*pow(T/ic.T0, 1.0/(P.gamma - 1.0));
            p = rho*P.R*T;
        }
        const double rho = p/(P.R*T);
        const long long o = (long long)slot*5*4096 + c;
        Q[o]          = rho;
        Q[o + 4096]   = 0.5*rho*(u*u + v*v + w*w) + p/(P.gamma - 1.0);
        Q[o + 2*4096] = rho*u;
        Q[o + 3*4096] = rho*v;
        Q[o + 4*4096] = rho*w;
    }
}

} // namespace

namespace {

// shape variant v: PFB = 1 + (v & 1), OCC = (v & 2) ? 2 : 3, TJ = (v & 4) ? 8 : 16
template <int LAW, bool HAS_U0, int TJ, int PFB, int OCC>
void launch_one(int grid, hipStream_t s, const BatchDev& B, const double* Qin, const double* Q0, double* Qout,
                double alpha, double beta, int want_diag)
{
    hipLaunchKernelGGL((k_stage<LAW, HAS_U0, TJ, PFB, OCC>), dim3(grid), dim3(16*TJ), 0, s,
                       B, Qin, Q0, Qout, alpha, beta, want_diag);
}

template <int LAW, bool HAS_U0>
void launch_shape(int v, const BatchDev& B, const double* Qin, const double* Q0, double* Qout,
                  double alpha, double beta, int want_diag, int grid, hipStream_t s)
{
    switch (v)
    {
    case 0:  launch_one<LAW, HAS_U0, 16, 1, 3>(grid, s, B, Qin, Q0, Qout, alpha, beta, want_diag); break;
    case 1:  launch_one<LAW, HAS_U0, 16, 2, 3>(grid, s, B, Qin, Q0, Qout, alpha, beta, want_diag); break;
    case 2:  launch_one<LAW, HAS_U0, 16, 1, 2>(grid, s, B, Qin, Q0, Qout, alpha, beta, want_diag); break;
    case 3:  launch_one<LAW, HAS_U0, 16, 2, 2>(grid, s, B, Qin, Q0, Qout, alpha, beta, want_diag); break;
    case 4:  launch_one<LAW, HAS_U0, 8, 1, 3>(grid, s, B, Qin, Q0, Qout, alpha, beta, want_diag); break;
    case 5:  launch_one<LAW, HAS_U0, 8, 2, 3>(grid, s, B, Qin, Q0, Qout, alpha, beta, want_diag); break;
    case 6:  launch_one<LAW, HAS_U0, 8, 1, 2>(grid, s, B, Qin, Q0, Qout, alpha, beta, want_diag); break;
    default: launch_one<LAW, HAS_U0, 8, 2, 2>(grid, s, B, Qin, Q0, Qout, alpha, beta, want_diag); break;
    }
}

template <int LAW, int TJ, int PFB, int OCC>
int occ_one()
{
    int n = 0;
    SPD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_stage<LAW, true, TJ, PFB, OCC>, 16*TJ, 0));
    return n;
}

template <int LAW>
int occupancy(int v)
{
    switch (v)
    {
    case 0:  return occ_one<LAW, 16, 1, 3>();
    case 1:  return occ_one<LAW, 16, 2, 3>();
    case 2:  return occ_one<LAW, 16, 1, 2>();
    case 3:  return occ_one<LAW, 16, 2, 2>();
    case 4:  return occ_one<LAW, 8, 1, 3>();
    case 5:  return occ_one<LAW, 8, 2, 3>();
    case 6:  return occ_one<LAW, 8, 1, 2>();
    default: return occ_one<LAW, 8, 2, 2>();
    }
}

} // namespace

// Work items in XCD runs (1, the default) or every eighth per XCD (0).  A
// block reads its halo from the neighbour slots' interiors, so with runs the
// neighbours of a case share L2 lines: the 300-case TGV / vortex ensemble
// steps 11.4 -> 12.7-12.9 G cell-updates/s (the tuner then picks
// three-wave shapes).  SPD_BATCH_XCD=0|1 forces the choice.
static int batch_xcd_runs()
{
    if (const char* e = getenv("SPD_BATCH_XCD")) return e[0] == '1' ? 1 : 0;
    return 1;
}

void launch_batch_stage(const BatchDev& B, int law, int variant, const double* Qin, const double* Q0, double* Qout,
                        double alpha, double beta, int want_diag, int grid, hipStream_t s)
{
    if (variant < 0 || variant >= kBatchShapeVariants) throw ArgError{"batch kernel shape variant out of range"};
    want_diag = (want_diag ? 1 : 0) | (batch_xcd_runs() ? 2 : 0);
    // stage 1 of the Shu-Osher form (alpha = 0) reads no u0
    const bool u0 = alpha != 0.0;
    if (law == SPD_LAW_POWER)
    {
        if (u0) launch_shape<SPD_LAW_POWER, true>(variant, B, Qin, Q0, Qout, alpha, beta, want_diag, grid, s);
        else    launch_shape<SPD_LAW_POWER, false>(variant, B, Qin, Q0, Qout, alpha, beta, want_diag, grid, s);
    }
    else
    {
        if (u0) launch_shape<SPD_LAW_CONST, true>(variant, B, Qin, Q0, Qout, alpha, beta, want_diag, grid, s);
        else    launch_shape<SPD_LAW_CONST, false>(variant, B, Qin, Q0, Qout, alpha, beta, want_diag, grid, s);
    }
    SPD_LAUNCH_CHECK();
}

int batch_stage_blocks_per_cu(int law, int variant)
{
    if (variant < 0 || variant >= kBatchShapeVariants) throw ArgError{"batch kernel shape variant out of range"};
    const int n = law == SPD_LAW_POWER ? occupancy<SPD_LAW_POWER>(variant) : occupancy<SPD_LAW_CONST>(variant);
    return std::max(1, n);
}

void launch_batch_step_count(const BatchDev& B, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_count, dim3(1), dim3(64), 0, s, B.step);
    SPD_LAUNCH_CHECK();
}

void launch_batch_init(const BatchDev& B, const int* d_stage, int nslots, const CaseParams& P, const CaseIC& ic,
                       double* Q, hipStream_t s)
{
    hipLaunchKernelGGL(k_init_case, dim3(4, nslots), dim3(256), 0, s, d_stage, nslots, P, ic, Q,
                       const_cast<int*>(B.nbr), const_cast<int*>(B.slot_case));
    SPD_LAUNCH_CHECK();
}

} // namespace spd
