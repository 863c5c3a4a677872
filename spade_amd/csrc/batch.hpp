// Device-side tables of the batched engine (see batch.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace spd {

// Per-case constants; FP64 for the state arithmetic, f_* copies for the FP32 fluxes.
struct CaseParams
{
    double dt;
    double ix, iy, iz;                       // 1 / dx per direction
    double gamma, gm1, R, inv_R, cv;         // ideal gas
    double rho_ref, p_ref, T_ref, e_ref;     // reference state the FP32 fluxes are split about
    float  f_rho_ref, f_T_ref;
    float  f_mu_ref, f_inv_T_law, f_power, f_cp_over_pr;
    int    law;
};

// Initial condition of a case (TGV or isentropic vortex)
struct CaseIC
{
    int    kind;                             // SPD_CASE_*
    int    nb[3];                            // blocks per axis
    double dx[3];
    double rho0, u0, p0, T0;
    double xc, yc, beta, dT;                 // vortex centre, strength, temperature dip
};

struct BatchDev
{
    const int* worklist;                     // slots to advance
    const int* nwork;                        // number of entries in worklist (device-side)
    const int* slot_case;                    // slot -> case table index
    const int* nbr;                          // [27*slot + e] neighbour slot (periodic within case)
    const CaseParams* cases;
    double* diag;                            // [hist][slot][half][2]: kinetic energy, enstrophy sums per half block
    int* step;                               // device step counter (history row)
    int max_slots;
    int hist_len;
};

// Kernel shape variants of the stage kernel: variant v keeps 1 + (v & 1)
// planes of state in flight, sizes registers for (v & 2) ? 2 : 3 waves per
// SIMD, and gives each workgroup (v & 4) ? 8 : 16 rows of a block's 16x16
// cross-section (two 128-thread or one 256-thread workgroup per block).
constexpr int kBatchShapeVariants = 8;
inline int batch_variant_prefetch(int v) { return 1 + (v & 1); }
inline int batch_variant_occupancy(int v) { return (v & 2) ? 2 : 3; }
inline int batch_variant_rows(int v) { return (v & 4) ? 8 : 16; }
inline int batch_variant(int rows, int prefetch, int occupancy)
{
    return (prefetch - 1) | (occupancy == 2 ? 2 : 0) | (rows == 8 ? 4 : 0);
}

void launch_batch_stage(const BatchDev& B, int law, int variant, const double* Qin, const double* Q0, double* Qout,
                        double alpha, double beta, int want_diag, int grid, hipStream_t s);
void launch_batch_step_count(const BatchDev& B, hipStream_t s);
// Workgroups of the stage kernel (for law and shape variant) resident per CU
// at once: the default persistent grid is this many per CU
int batch_stage_blocks_per_cu(int law, int variant);
// d_stage: per block {slot, case table, 27 neighbour slots}; also fills B.nbr and B.slot_case
void launch_batch_init(const BatchDev& B, const int* d_stage, int nslots, const CaseParams& P, const CaseIC& ic,
                       double* Q, hipStream_t s);

} // namespace spd
