// Host-side batched engine (see batch_host.cpp).
#pragma once
#include <map>
#include <vector>
#include <string>
#include "spd_internal.hpp"
#include "batch.hpp"

namespace spd {

struct BatchCase
{
    int id = -1;
    spd_case_desc desc{};
    CaseParams P{};
    CaseIC ic{};
    double dt = 0.0;
    int table = -1;
    std::vector<int> slots;          // slot of each block, block order x-fastest
    int steps_done = 0;
    std::vector<double> ke, en;      // per step: volume-averaged kinetic energy and enstrophy
    int diverged_step = -1;          // first step whose starting state had a non-finite kinetic energy
};

struct TuneEntry { int grid; int interleaved; int variant; float ms_per_step; };

struct Batch
{
    Ctx* ctx;
    int law;
    int hist_len;
    int max_slots = 0, max_cases = 0;
    int num_cu = 1;
    int grid = 0, graph_grid = -1;
    int variant = 0, graph_variant = -1;   // stage kernel shape (batch.hpp)
    bool use_graph = true;
    bool order_interleaved = false;
    bool worklist_dirty = true;
    int nwork = 0;
    int next_id = 0;
    int host_step = 0;
    double* Q[3] = {nullptr, nullptr, nullptr};
    int* d_worklist = nullptr;
    int* d_slot_case = nullptr;
    int* d_nbr = nullptr;
    int* d_small = nullptr;          // [0] = nwork, [1] = step counter
    int* d_init_slots = nullptr;
    CaseParams* d_cases = nullptr;
    double* d_diag = nullptr;
    BatchDev dev{};
    hipGraphExec_t gexec = nullptr;
    std::vector<int> free_slots, free_tables, h_slot_case, h_nbr, h_worklist, case_of_table;
    std::vector<CaseParams> h_cases;
    std::map<int, BatchCase> cases;
    std::vector<TuneEntry> tune_log;

    Batch(Ctx* c, long long budget_bytes, int law, int hist);
    ~Batch();
    static long long footprint(const spd_case_desc& d);   // blocks
    void validate(const spd_case_desc& d) const;
    void setup_case(const spd_case_desc& d, CaseParams& P, CaseIC& ic, double& dt) const;
    int add_case(const spd_case_desc& d);
    void remove_case(int id);
    void upload_worklist();
    void enqueue_step(hipStream_t s);
    void build_graph();
    void collect_diag(int first_step, int nsteps);
    void step(int nsteps);
    const BatchCase& get(int id) const;
    void download(int id, double* host) const;
    void upload(int id, const double* host);
    int tune();
};

} // namespace spd
