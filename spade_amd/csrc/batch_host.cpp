// Host side of the batched engine: slot pool sized by an HBM budget, case
// registry, work list, per-step diagnostics history, hipGraph of one SSP-RK3
// step, and the launch-shape autotuner.
#include "spd_internal.hpp"
#include "batch.hpp"
#include "batch_host.hpp"
#include <cmath>
#include <algorithm>
#include <chrono>

namespace spd {

namespace {

constexpr size_t kSlotDoubles = 5*4096;
constexpr double kPi = 3.14159265358979323846;

inline void need(bool ok, const std::string& msg) { if (!ok) throw ArgError{msg}; }

} // namespace

Batch::Batch(Ctx* c, long long budget_bytes, int law_in, int hist) : ctx(c), law(law_in), hist_len(hist)
{
    need(law == SPD_LAW_CONST || law == SPD_LAW_POWER, "unknown viscosity law");
    need(hist_len >= 1 && hist_len <= 4096, "history length must be in 1..4096");
    const size_t per_slot = 3*kSlotDoubles*sizeof(double) + 27*sizeof(int) + 2*sizeof(int)
                          + (size_t)hist_len*4*sizeof(double);
    need(budget_bytes > (long long)per_slot, "HBM budget smaller than one block");
    max_slots = (int)std::min<long long>(budget_bytes/(long long)per_slot, 1 << 22);
    max_cases = std::max(1, max_slots);
    SPD_HIP(hipSetDevice(ctx->device));
    SPD_HIP(hipMalloc((void**)&Q[0], 3*kSlotDoubles*sizeof(double)*(size_t)max_slots));
    Q[1] = Q[0] + kSlotDoubles*(size_t)max_slots;
    Q[2] = Q[1] + kSlotDoubles*(size_t)max_slots;
    SPD_HIP(hipMalloc((void**)&d_worklist, sizeof(int)*max_slots));
    SPD_HIP(hipMalloc((void**)&d_slot_case, sizeof(int)*max_slots));
    SPD_HIP(hipMalloc((void**)&d_nbr, sizeof(int)*27*(size_t)max_slots));
    SPD_HIP(hipMalloc((void**)&d_small, sizeof(int)*64));
    SPD_HIP(hipMalloc((void**)&d_cases, sizeof(CaseParams)*(size_t)max_cases));
    SPD_HIP(hipMalloc((void**)&d_diag, sizeof(double)*4*(size_t)hist_len*max_slots));
    SPD_HIP(hipMalloc((void**)&d_init_slots, sizeof(int)*29*(size_t)max_slots));
    SPD_HIP(hipMemsetAsync(d_small, 0, sizeof(int)*64, ctx->stream));
    SPD_HIP(hipMemsetAsync(d_diag, 0, sizeof(double)*4*(size_t)hist_len*max_slots, ctx->stream));
    h_slot_case.assign(max_slots, -1);
    h_nbr.assign(27*(size_t)max_slots, 0);
    h_cases.resize(max_cases);
    case_of_table.assign(max_cases, -1);
    for (int s = max_slots - 1; s >= 0; --s) free_slots.push_back(s);
    for (int t = max_cases - 1; t >= 0; --t) free_tables.push_back(t);
    int cus = 0;
    SPD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    num_cu = std::max(1, cus);
    // as many persistent workgroups as fit at once (spd_batch_tune may pick another)
    grid = batch_stage_blocks_per_cu(law, variant)*num_cu;
    dev = BatchDev{d_worklist, d_small + 0, d_slot_case, d_nbr, d_cases, d_diag, d_small + 1, max_slots, hist_len};
}

Batch::~Batch()
{
    if (gexec) (void)hipGraphExecDestroy(gexec);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(Q[0]); (void)hipFree(d_worklist); (void)hipFree(d_slot_case); (void)hipFree(d_nbr);
    (void)hipFree(d_small); (void)hipFree(d_cases); (void)hipFree(d_diag); (void)hipFree(d_init_slots);
}

long long Batch::footprint(const spd_case_desc& d)
{
    long long blocks = 1;
    for (int a = 0; a < 3; ++a) blocks *= d.cells[a]/16;
    return blocks;
}

void Batch::validate(const spd_case_desc& d) const
{
    need(d.kind == SPD_CASE_TGV || d.kind == SPD_CASE_VORTEX, "unknown case kind");
    for (int a = 0; a < 3; ++a)
        need(d.cells[a] >= 16 && d.cells[a] % 16 == 0, "case cells per axis must be positive multiples of 16");
    need(d.mach > 0.0 && d.mach < 5.0, "mach must be in (0, 5)");
    need(d.reynolds > 0.0, "reynolds must be positive");
    need(d.gamma > 1.0 && d.prandtl > 0.0, "gamma must exceed 1 and prandtl be positive");
    need(d.cfl > 0.0 && d.cfl <= 2.0, "cfl must be in (0, 2]");
    need(d.law == law, "case viscosity law differs from the batch's");
}

// Non-dimensional set-up of a case.
//   TGV: [0, 2 pi]^3, rho0 = u0 = T0 = 1, c0 = u0 / M, p0 = rho0 c0^2 / gamma,
//        R = p0 / (rho0 T0), mu_ref = rho0 u0 L / Re with L = 1.
//   Isentropic vortex: [0, 10]^3, free stream u0 = M c0 along x with
//        rho0 = T0 = 1, R = 1 / gamma (c0 = 1), strength beta = 5 M, centre
//        (5, 5), T = T0 - (gamma-1) beta^2 / (8 gamma R pi^2) exp(1 - r^2)
//        (cp dT balances the vortex kinetic energy), rho = (T/T0)^(1/(gamma-1)).
void Batch::setup_case(const spd_case_desc& d, CaseParams& P, CaseIC& ic, double& dt) const
{
    ic = CaseIC{};
    ic.kind = d.kind;
    for (int a = 0; a < 3; ++a) ic.nb[a] = d.cells[a]/16;
    const double g = d.gamma;
    double L[3], R, umax, cmax, mu_ref;
    if (d.kind == SPD_CASE_TGV)
    {
        for (int a = 0; a < 3; ++a) L[a] = 2.0*kPi;
        ic.rho0 = 1.0; ic.u0 = 1.0; ic.T0 = 1.0;
        const double c0 = ic.u0/d.mach;
        ic.p0 = ic.rho0*c0*c0/g;
        R = ic.p0/(ic.rho0*ic.T0);
        mu_ref = ic.rho0*ic.u0*1.0/d.reynolds;
        umax = ic.u0;
        // T is uniform; the pressure field bounds c from above by its maximum
        const double pmax = ic.p0 + 4.0*3.0*ic.rho0*ic.u0*ic.u0/16.0;
        cmax = std::sqrt(g*pmax/(ic.p0/(R*ic.T0)));
    }
    else
    {
        for (int a = 0; a < 3; ++a) L[a] = 10.0;
        ic.rho0 = 1.0; ic.T0 = 1.0;
        R = 1.0/g;
        const double c0 = std::sqrt(g*R*ic.T0);
        ic.u0 = d.mach*c0;
        ic.beta = 5.0*d.mach;
        ic.xc = 5.0; ic.yc = 5.0;
        ic.dT = (g - 1.0)*ic.beta*ic.beta/(8.0*g*R*kPi*kPi);
        ic.p0 = ic.rho0*R*ic.T0;
        mu_ref = ic.rho0*ic.u0*1.0/d.reynolds;
        umax = ic.u0 + ic.beta/(2.0*kPi);
        cmax = c0;
    }
    for (int a = 0; a < 3; ++a) ic.dx[a] = L[a]/d.cells[a];
    const double dxmin = std::min(ic.dx[0], std::min(ic.dx[1], ic.dx[2]));
    dt = d.cfl*dxmin/(umax + cmax);
    P = CaseParams{};
    P.dt = dt;
    P.ix = 1.0/ic.dx[0]; P.iy = 1.0/ic.dx[1]; P.iz = 1.0/ic.dx[2];
    P.gamma = g; P.gm1 = g - 1.0; P.R = R; P.inv_R = 1.0/R; P.cv = R/(g - 1.0);
    P.rho_ref = ic.rho0; P.T_ref = ic.T0; P.p_ref = ic.rho0*R*ic.T0; P.e_ref = P.cv*ic.T0;
    P.f_rho_ref = (float)P.rho_ref; P.f_T_ref = (float)P.T_ref;
    P.f_mu_ref = (float)mu_ref; P.f_inv_T_law = (float)(1.0/ic.T0); P.f_power = (float)d.power;
    P.f_cp_over_pr = (float)((g*R/(g - 1.0))/d.prandtl);
    P.law = d.law;
}

int Batch::add_case(const spd_case_desc& d)
{
    validate(d);
    const long long nblk = footprint(d);
    need((long long)free_slots.size() >= nblk && !free_tables.empty(), "not enough free HBM slots for this case");
    BatchCase c;
    c.id = next_id++;
    c.desc = d;
    setup_case(d, c.P, c.ic, c.dt);
    c.table = free_tables.back(); free_tables.pop_back();
    c.slots.resize(nblk);
    for (long long b = 0; b < nblk; ++b) { c.slots[b] = free_slots.back(); free_slots.pop_back(); }
    const int* nb = c.ic.nb;
    for (long long b = 0; b < nblk; ++b)
    {
        const int bx = b % nb[0], by = (b / nb[0]) % nb[1], bz = b / (nb[0]*nb[1]);
        const int s = c.slots[b];
        h_slot_case[s] = c.table;
        for (int ez = -1; ez <= 1; ++ez) for (int ey = -1; ey <= 1; ++ey) for (int ex = -1; ex <= 1; ++ex)
        {
            const int x = (bx + ex + nb[0]) % nb[0], y = (by + ey + nb[1]) % nb[1], z = (bz + ez + nb[2]) % nb[2];
            h_nbr[27*(size_t)s + (ex + 1) + 3*(ey + 1) + 9*(ez + 1)] = c.slots[x + nb[0]*(y + nb[1]*z)];
        }
    }
    h_cases[c.table] = c.P;
    case_of_table[c.table] = c.id;
    // tables and initial state, in one staged copy (on the stream, ordered
    // before any later step; pageable copies leave the host buffer free on return)
    std::vector<int> stage(29*(size_t)nblk);
    for (long long b = 0; b < nblk; ++b)
    {
        const int s = c.slots[b];
        stage[29*b] = s;
        stage[29*b + 1] = h_slot_case[s];
        for (int e = 0; e < 27; ++e) stage[29*b + 2 + e] = h_nbr[27*(size_t)s + e];
    }
    SPD_HIP(hipMemcpyAsync(d_cases + c.table, &h_cases[c.table], sizeof(CaseParams), hipMemcpyHostToDevice, ctx->stream));
    SPD_HIP(hipMemcpyAsync(d_init_slots, stage.data(), sizeof(int)*stage.size(), hipMemcpyHostToDevice, ctx->stream));
    launch_batch_init(dev, d_init_slots, (int)nblk, c.P, c.ic, Q[0], ctx->stream);
    c.steps_done = 0;
    cases.emplace(c.id, std::move(c));
    worklist_dirty = true;
    return next_id - 1;
}

void Batch::remove_case(int id)
{
    auto it = cases.find(id);
    need(it != cases.end(), "unknown case id");
    for (int s : it->second.slots) { h_slot_case[s] = -1; free_slots.push_back(s); }
    case_of_table[it->second.table] = -1;
    free_tables.push_back(it->second.table);
    cases.erase(it);
    worklist_dirty = true;
}

void Batch::upload_worklist()
{
    std::vector<int> wl;
    if (order_interleaved)
    {
        // round-robin over cases: neighbouring work items belong to different cases
        size_t maxb = 0;
        for (auto& kv : cases) maxb = std::max(maxb, kv.second.slots.size());
        for (size_t b = 0; b < maxb; ++b)
            for (auto& kv : cases) if (b < kv.second.slots.size()) wl.push_back(kv.second.slots[b]);
    }
    else
    {
        for (auto& kv : cases) wl.insert(wl.end(), kv.second.slots.begin(), kv.second.slots.end());
    }
    h_worklist = wl;
    nwork = (int)wl.size();
    if (nwork) SPD_HIP(hipMemcpyAsync(d_worklist, h_worklist.data(), sizeof(int)*nwork, hipMemcpyHostToDevice, ctx->stream));
    SPD_HIP(hipMemcpyAsync(d_small + 0, &nwork, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    SPD_HIP(hipStreamSynchronize(ctx->stream));
    worklist_dirty = false;
}

void Batch::enqueue_step(hipStream_t s)
{
    // Shu-Osher SSP-RK3: u1 = u0 + dt L(u0); u2 = 3/4 u0 + 1/4 (u1 + dt L(u1));
    // u3 = 1/3 u0 + 2/3 (u2 + dt L(u2)).  Q0 -> Q1 -> Q2 -> Q0.
    launch_batch_stage(dev, law, variant, Q[0], Q[0], Q[1], 0.0, 1.0, 1, grid, s);
    launch_batch_stage(dev, law, variant, Q[1], Q[0], Q[2], 0.75, 0.25, 0, grid, s);
    launch_batch_stage(dev, law, variant, Q[2], Q[0], Q[0], 1.0/3.0, 2.0/3.0, 0, grid, s);
    launch_batch_step_count(dev, s);
}

void Batch::build_graph()
{
    if (gexec) { SPD_HIP(hipGraphExecDestroy(gexec)); gexec = nullptr; }
    hipGraph_t g = nullptr;
    SPD_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(ctx->stream);
    SPD_HIP(hipStreamEndCapture(ctx->stream, &g));
    SPD_HIP(hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0));
    SPD_HIP(hipGraphDestroy(g));
    graph_grid = grid;
    graph_variant = variant;
}

void Batch::collect_diag(int first_step, int nsteps)
{
    // rows first_step .. first_step+nsteps-1 of the history ring, each cut to
    // the slots below the highest one in use (freed slots are reused first,
    // so that bound stays near the number of resident blocks, not the pool's)
    size_t hi = 0;
    for (auto& kv : cases)
        for (int s : kv.second.slots) hi = std::max(hi, (size_t)s + 1);
    // per slot: {ke, enstrophy} of rows 0-7, then of rows 8-15
    std::vector<double> rows((size_t)nsteps*4*hi);
    if (hi > 0)
    {
        for (int t = 0; t < nsteps; ++t)
        {
            const int h = (first_step + t) % hist_len;
            SPD_HIP(hipMemcpyAsync(rows.data() + (size_t)t*4*hi, d_diag + (size_t)h*4*max_slots,
                                   sizeof(double)*4*hi, hipMemcpyDeviceToHost, ctx->stream));
        }
    }
    SPD_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& kv : cases)
    {
        BatchCase& c = kv.second;
        const double ncell = 4096.0*c.slots.size();
        for (int t = 0; t < nsteps; ++t)
        {
            double ke = 0.0, en = 0.0;
            for (int s : c.slots)   // fixed (case block) order: identical alone or in a mix
            {
                const double* r = rows.data() + (size_t)t*4*hi + 4*(size_t)s;
                ke += r[0] + r[2];
                en += r[1] + r[3];
            }
            c.ke.push_back(ke/ncell);
            c.en.push_back(en/ncell);
            if (c.diverged_step < 0 && !std::isfinite(ke)) c.diverged_step = c.steps_done + t;
        }
        c.steps_done += nsteps;
    }
}

void Batch::step(int nsteps)
{
    need(nsteps >= 0, "nsteps must be >= 0");
    if (worklist_dirty) upload_worklist();
    int done = 0;
    while (done < nsteps)
    {
        const int chunk = std::min(nsteps - done, hist_len);
        if (use_graph && (!gexec || graph_grid != grid || graph_variant != variant)) build_graph();
        for (int t = 0; t < chunk; ++t)
        {
            if (use_graph) SPD_HIP(hipGraphLaunch(gexec, ctx->stream));
            else           enqueue_step(ctx->stream);
        }
        collect_diag(host_step, chunk);
        host_step += chunk;
        done += chunk;
    }
}

const BatchCase& Batch::get(int id) const
{
    auto it = cases.find(id);
    need(it != cases.end(), "unknown case id");
    return it->second;
}

// Calls f(b, e) for each run slots[b..e) of consecutive slot indices, so that
// blocks move in one copy per run (a case admitted into a fresh pool is one run).
template <class F>
static void for_slot_runs(const std::vector<int>& slots, F f)
{
    for (size_t b = 0, e; b < slots.size(); b = e)
    {
        for (e = b + 1; e < slots.size() && slots[e] == slots[e - 1] + 1; ++e) {}
        f(b, e);
    }
}

void Batch::download(int id, double* host) const
{
    const BatchCase& c = get(id);
    for_slot_runs(c.slots, [&](size_t b, size_t e) {
        SPD_HIP(hipMemcpyAsync(host + b*kSlotDoubles, Q[0] + (size_t)c.slots[b]*kSlotDoubles,
                               sizeof(double)*kSlotDoubles*(e - b), hipMemcpyDeviceToHost, ctx->stream));
    });
    SPD_HIP(hipStreamSynchronize(ctx->stream));
}

void Batch::upload(int id, const double* host)
{
    const BatchCase& c = get(id);
    for_slot_runs(c.slots, [&](size_t b, size_t e) {
        SPD_HIP(hipMemcpyAsync(Q[0] + (size_t)c.slots[b]*kSlotDoubles, host + b*kSlotDoubles,
                               sizeof(double)*kSlotDoubles*(e - b), hipMemcpyHostToDevice, ctx->stream));
    });
    SPD_HIP(hipStreamSynchronize(ctx->stream));
}

// Autotuner: times one step for each candidate launch shape on the current
// batch and keeps the fastest.  Candidates: the stage kernel's shape variant
// (16 or 8 rows of a block's cross-section per workgroup; 1 or 2 planes of
// state in flight; registers sized for 3 or 2 waves per SIMD), the persistent
// grid (1..4 blocks' worth of workgroups per CU) and the work order
// (case-major or interleaved).  State is saved and restored around the trials
// so tuning does not advance any case.
int Batch::tune()
{
    if (worklist_dirty) upload_worklist();
    if (nwork == 0) return grid;
    // save the step-start state (Q0) of the active slots
    double* save = nullptr;
    SPD_HIP(hipMalloc((void**)&save, sizeof(double)*kSlotDoubles*(size_t)nwork));
    const std::vector<int> saved_slots = h_worklist;
    for_slot_runs(saved_slots, [&](size_t b, size_t e) {
        SPD_HIP(hipMemcpyAsync(save + b*kSlotDoubles, Q[0] + (size_t)saved_slots[b]*kSlotDoubles,
                               sizeof(double)*kSlotDoubles*(e - b), hipMemcpyDeviceToDevice, ctx->stream));
    });
    int saved_step = 0;
    SPD_HIP(hipMemcpyAsync(&saved_step, d_small + 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SPD_HIP(hipStreamSynchronize(ctx->stream));
    float best = 1e30f; int best_grid = grid, best_variant = variant; bool best_order = order_interleaved;
    tune_log.clear();
    for (int order = 0; order < 2; ++order)
    {
        order_interleaved = (order == 1);
        upload_worklist();
        for (int v = 0; v < kBatchShapeVariants; ++v)
        {
            variant = v;
            // 1..4 block-sized workgroups per CU (two half-block workgroups count as one)
            const int per_block = 16/batch_variant_rows(v);
            for (int per_cu = 1; per_cu <= 4; ++per_cu)
            {
                grid = per_cu*per_block*num_cu;
                enqueue_step(ctx->stream);              // warm-up
                SPD_HIP(hipEventRecord(ctx->ev0, ctx->stream));
                for (int r = 0; r < 2; ++r) enqueue_step(ctx->stream);
                SPD_HIP(hipEventRecord(ctx->ev1, ctx->stream));
                SPD_HIP(hipEventSynchronize(ctx->ev1));
                float ms = 0.0f;
                SPD_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
                tune_log.push_back({grid, order, v, ms/2.0f});
                if (ms < best) { best = ms; best_grid = grid; best_order = order_interleaved; best_variant = v; }
            }
        }
    }
    grid = best_grid;
    variant = best_variant;
    order_interleaved = best_order;
    upload_worklist();
    for_slot_runs(saved_slots, [&](size_t b, size_t e) {
        SPD_HIP(hipMemcpyAsync(Q[0] + (size_t)saved_slots[b]*kSlotDoubles, save + b*kSlotDoubles,
                               sizeof(double)*kSlotDoubles*(e - b), hipMemcpyDeviceToDevice, ctx->stream));
    });
    SPD_HIP(hipMemcpyAsync(d_small + 1, &saved_step, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    SPD_HIP(hipStreamSynchronize(ctx->stream));
    SPD_HIP(hipFree(save));
    return grid;
}

} // namespace spd
