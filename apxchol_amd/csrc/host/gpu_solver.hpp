// gpu_solver: the MI355X counterpart of the reference's cpu_solver
// (include/apxchol/solver/solve.h:64-131) -- factor once on the host, upload,
// then PCG-solve any number of right-hand sides on one GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <vector>

#include "../device.hpp"
#include "../kernels/launch.hpp"
#include "factor.hpp"
#include "trsv_prep.hpp"

namespace apxchol_amd {

enum class sweep_precision { fp64 = 0, fp32 = 1 };   // vector / accumulation width of the two sweeps
enum class sweep_schedule { dataflow = 0, levels = 1, hybrid = 2 };

// solve.h:13-25 plus the device knobs.
struct solve_options {
    double tol = 1e-8;
    int max_iter = 200;
    int stagnation_window = 50;
    bool keep_factor_values = false;
    factor_options factor_opts;
    sweep_precision precision = sweep_precision::fp64;
    bool fp16_storage = false;
    sweep_schedule schedule = sweep_schedule::hybrid;
    int tail_rows = 0;            // hybrid: levels below this many rows go to the one-workgroup tail (0: default)
    int tail_batches = 0;         // hybrid: ... and of at most this many batches (0: APXCHOL_AMD_TAIL_BATCHES or 48; <= 64)
    int tail_cut = 0;             // hybrid: a run ends before a level needing this many overflow groups per lane
                                  // (0: APXCHOL_AMD_TAIL_CUT or 1; < 0: never)
    double drop_rel = -1.0;       // < 0: APXCHOL_FACTOR_DROP / the 1e-4 default
    int iters_per_check = 2;      // iterations enqueued between completion tests
    int sweep_chunk = 0;          // batches per dataflow ticket (0: APXCHOL_AMD_SWEEP_CHUNK or 8)
    int sweep_blocks = 0;         // dataflow grid (0: APXCHOL_AMD_SWEEP_BLOCKS or the resident grid)
    bool graph = true;            // replay each PCG iteration as one captured HIP graph
    bool relabel = true;          // sort every level's rows by original vertex (APXCHOL_AMD_RELABEL=0: off)
    bool depth_levels = true;     // with relabel: levels = topological depth (APXCHOL_AMD_DEPTH_LEVELS=0: rounds)
    int lane_batches = 0;         // launched levels of <= this many batches use the lane layout
                                  // (0: APXCHOL_AMD_LANE_BATCHES or 4096; < 0: never)
};

// Setup / solve phases in milliseconds (host wall clock).
struct solve_timings {
    double factorize_ms = 0.0;      // setup/factorize (0 when the factor was handed in)
    double sptrsv_setup_ms = 0.0;   // setup/sptrsv_setup: trsv preparation, layouts, device uploads
    double pcg_setup_ms = 0.0;      // setup/gpu_pcg_setup: the permuted operator (CSR + sliced ELL)
    double pcg_ms = 0.0;            // pcg/gpu_pcg_loop: the solve itself, final synchronization included
};

// solve.h:27-42. status: 0 running/none, 1 converged, 2 pAp <= 0 breakdown,
// 3 stagnation, 4 max_iter, 5 zero right-hand side.
struct solve_result {
    std::vector<double> x;
    std::int64_t iterations = 0;
    double residual = 0.0;
    std::int64_t lumped_offdiag = 0;
    double solve_vram_mb = -1.0;
    int status = 0;
    // solve_result::timings (the reference's checkpoint tree, solve.h:27-42;
    // labels of src/solve.cpp:681-737 and preconditioner.h:238-286): the
    // solver's setup phases (every solve of a solver reports the same) and
    // this solve's PCG, host wall clock in milliseconds.
    solve_timings timings;
};

struct solver_stats {
    std::int64_t n = 0, m = 0;
    bool sddm = false;
    std::uint64_t factor_nnz = 0, stored_nnz = 0, offdiag = 0, dropped = 0;
    std::int64_t rounds = 0, levels = 0, fwd_batches = 0, bck_batches = 0;
    std::int64_t op_nnz = 0;
    bool op_fp32 = false;
    int spmv_lanes = 0, dataflow_grid = 0;
    std::uint64_t device_bytes = 0;
    // Per direction (0 forward, 1 back) of the hybrid / level schedule: the
    // algorithmic HBM bytes of one sweep's level launches and tail runs, and
    // their kernel counts. Bytes of a level: off-diagonals x (4 index + value
    // width) + rows x (4 pointer + 4 diagonal + input + output width [+ 4
    // inverse scale on the fp16 back sweep]) + 12 per batch (batch table).
    std::uint64_t level_bytes[2] = {0, 0}, tail_bytes[2] = {0, 0};
    std::int64_t level_launches[2] = {0, 0}, tail_runs[2] = {0, 0};
    std::int64_t lane_levels[2] = {0, 0};   // launched levels on the static lane layout
    int spmv_format = 0;                    // single SpMV: 0 CSR lanes, 1 sliced ELL
    std::uint64_t spmv_sell_slots = 0;      // padded entries of the sliced-ELL form (0: none)
    std::int64_t split_couplings[2] = {0, 0};   // tail runs whose coupling is split, per direction
};

struct block_sched;

// What a grouped solve (solver_group) needs of one member solver: its device
// arrays, its solve-state defaults and the steps of its two sweeps (launched
// levels and tail runs, in order; host copies of the device descriptors).
struct group_member_view {
    group_member mem;                  // st is left null: the group owns the states
    const unsigned* perm = nullptr;    // original vertex -> permuted position (device)
    bool sddm = true;
    vec_prec prec = vec_prec::fp64;
    bool fp16 = false;
    std::vector<group_unit> steps[2];  // forward, back (st left null)
    int stagnation_window = 50;
    int center_k = 10;
    std::int64_t lumped_offdiag = 0;
};

class gpu_solver {
public:
    explicit gpu_solver(const csc_view& L, const solve_options& opts = {});
    gpu_solver(const csc_view& L, factorization F, const solve_options& opts = {});
    ~gpu_solver();
    gpu_solver(const gpu_solver&) = delete;
    gpu_solver& operator=(const gpu_solver&) = delete;

    // Host-vector forms (solve.h:83-103).
    solve_result solve(const std::vector<double>& b, double tol = -1.0, int max_iter = -1,
                       const std::vector<double>* x0 = nullptr) const;
    std::vector<double> apply(const std::vector<double>& r) const;

    // Device-vector forms: every pointer is device memory of n doubles; the
    // result's x is left empty.
    solve_result solve_device(const double* d_b, double* d_x, double tol, int max_iter, const double* d_x0,
                              hipStream_t s) const;
    void apply_device(const double* d_r, double* d_z, hipStream_t s) const;
    // k right-hand sides in blocked PCG (no counterpart in the reference:
    // k independent cpu_solver::solve(b) calls, solve.h:83-100, run as one):
    // column c of B / X is n doubles at d_B + c * n / d_X + c * n. Every
    // column keeps its own convergence test and stops like its own solve;
    // the sweeps and SpMM give each column the single-vector kernels' bits.
    std::vector<solve_result> solve_batch_device(int k, const double* d_B, double* d_X, double tol, int max_iter,
                                                 hipStream_t s) const;
    std::vector<solve_result> solve_batch(int k, const std::vector<double>& B, std::vector<double>& X,
                                          double tol = -1.0, int max_iter = -1) const;
    // Serving sessions: one running block of `slots` columns (rounded up to
    // a power of two <= 64) that admits right-hand sides into free slots
    // between iteration groups and releases finished ones. A column's
    // arithmetic does not depend on the other columns or on its slot, so a
    // request gets the same bits alone and in a full block of this width.
    // Not to be mixed with other calls on this solver while a session runs.
    struct session_column {
        int slot;
        const double* d_b;   // n doubles (device), caller order
        double tol;
        int max_iter;
    };
    struct session_slot {
        bool done;
        int status;
        std::int64_t iterations;
        double residual;
    };
    // which: 0 or 1, two independent sessions (their own block buffers,
    // tables and graphs), which may step side by side on two streams.
    int session_begin(int slots, hipStream_t s, int which = 0) const;
    void session_admit(const std::vector<session_column>& cols, hipStream_t s, int which = 0) const;
    void session_step(int iterations, hipStream_t s, int which = 0) const;
    void session_read(std::vector<session_slot>& out, hipStream_t s, int which = 0) const;
    void session_collect(const std::vector<int>& slots, const std::vector<double*>& d_x, hipStream_t s,
                         int which = 0) const;
    // A single-vector solve in steps (a serving loop's single lane runs one
    // beside its block): solve_device is single_begin, single_step until
    // single_poll says done, then single_collect, so a request run this way
    // gets the bits of solve_device with the same tol and max_iter. It uses
    // the solver's single-vector buffers and state, which a session leaves
    // alone; its steps may run on another stream than the session's.
    // single_prepare captures the iteration graph beforehand (single_begin
    // does it on first use otherwise).
    void single_prepare() const;
    // Split couplings (on by default): a tail run's coupling over the entries
    // before the previous run runs on a side stream beside that run (a
    // parallel branch of the captured iteration). A serving loop turns them
    // off for the solvers it owns -- their extra branches contend for the
    // process's few hardware queues with the other factors' streams -- and
    // back on when it lets them go. Same bits either way; a change drops the
    // captured iteration graph.
    void set_split_couplings(bool on) const;
    void single_begin(const double* d_b, const double* d_x0, double tol, int max_iter, hipStream_t s) const;
    void single_step(int iterations, hipStream_t s) const;
    bool single_poll(session_slot& out, hipStream_t s) const;
    void single_collect(double* d_x, solve_result& res, hipStream_t s) const;

    // Autotuner: measures on this device, for this factor and operator, the
    // SpMV lanes per row (powers of two; the single SpMV and the blocked
    // SpMM then use the winner) and the block width of blocked solves (the
    // width with the most columns swept per second; used by servers that
    // ask for automatic slots). Captured graphs are rebuilt afterwards.
    struct autotune_result {
        int spmv_lanes = 0;
        double spmv_us[7] = {};           // lanes 1, 2, ..., 64
        int block_width = 0;
        double sweep_pair_us[4] = {};     // block widths 8, 16, 32, 64: forward + back sweep
        int spmv_sell = 0;                // 1: the single SpMV now runs on the sliced-ELL form
        double spmv_sell_us = 0.0;        // its time (0: no sliced-ELL form)
    };
    autotune_result autotune(int max_block_width, hipStream_t s) const;
    int preferred_block_width() const { return tuned_block_width_; }

    // k preconditioner applications at once (k calls of cpu_solver::apply,
    // solve.h:103, counting as ONE application of the center-k schedule).
    void apply_batch_device(int k, const double* d_R, double* d_Z, hipStream_t s) const;
    // One sweep in factor order with fp64 vectors (rhs, out: m doubles).
    void sweep_device(bool backward, const double* d_rhs, double* d_out, hipStream_t s) const;
    // The pair y = L^-T L^-1 x of cuda_sptrsv::solve_LLt_dev (cuda.h:480-487,
    // 736-785) in factor order, m doubles at the boundary, the two sweeps in
    // the solver's vector precision (fp32 vectors when precision is fp32).
    // d_x == d_y is allowed.
    void solve_LLt_device(const double* d_x, double* d_y, hipStream_t s) const;
    // y = A x in the caller's order (n doubles each) with the single SpMV.
    void spmv_device(const double* d_x, double* d_y, hipStream_t s) const;
    // Y = A X for k columns (column c at d_X + c n) with the blocked SpMM,
    // blocks of at most 64 columns; every column gets spmv_device's bits.
    void spmv_batch_device(int k, const double* d_X, double* d_Y, hipStream_t s) const;

    void reset_apply_count() const { apply_count_ = 0; }

    // Per-kernel-class timing with HIP events on the launching stream
    // (classes: 0 forward sweep, 1 back sweep, 2 SpMV, 3 vector passes, then
    // single kernels 4 sweep_levels, 5 sweep_tail, 6 sweep_couple). Solves run
    // eagerly while timing is on. Only launches that did work are counted:
    // those of iterations after convergence return at once and are excluded.
    static constexpr int kTimingClasses = 7;
    void set_timing(bool on) const;
    // ms[kTimingClasses], launches[kTimingClasses] since the last call; resets them.
    void take_timing(double* ms, std::int64_t* launches) const;
    // A solver serves one server at a time and, while it does, takes no other
    // device call (the serving session owns its block buffers, its stream's
    // graphs and the capture stream). claim_serving throws if it is already
    // serving; every other device entry point throws while it is.
    void claim_serving() const;
    void release_serving() const { serving_.store(false); }
    bool serving() const { return serving_.load(); }
    const solve_timings& setup_timings() const { return setup_times_; }
    // The member view of a grouped solve (see group_member_view). Throws if the
    // solver is serving or its sweeps are not the hybrid schedule.
    group_member_view group_view() const;
    const factorization& factor() const { return F_; }
    const trsv_factor& trsv() const { return T_; }
    const solver_stats& stats() const { return stats_; }
    std::int64_t rows() const { return n_; }

private:
    struct device_state;
    struct timer;
    void setup(const csc_view& L);
    void check_sweep_error() const;
    void not_serving(const char* what) const;
    // Enqueue PCG iteration `it` (SpMV, x/r update, both sweeps, the
    // preconditioned dot and the p update); step labels the timing records.
    void enqueue_iteration(int step, hipStream_t s) const;
    // y = A x with the solver's single-vector SpMV (sliced ELL or CSR lanes).
    void spmv(const double* x, double* y, pcg_state* st, bool want_dot, hipStream_t s, const launch_hook& hook) const;
    void enqueue_precondition(bool first, int step, hipStream_t s) const;
    std::vector<solve_result> solve_block(int k, const double* d_B, double* d_X, double tol, int max_iter,
                                          hipStream_t s) const;
    int prepare_block(int k, int which = 0) const;
    void replay_iterations_k(int kp, int count, hipStream_t s, int which = 0) const;
    void fresh_state(pcg_state& h, double tol, int max_iter) const;
    const block_sched* block_scheds(int kp, int which = 0) const;
    void spmv_block(int k, const double* d_X, double* d_Y, hipStream_t s) const;
    void apply_block(int k, const double* d_R, double* d_Z, bool center, hipStream_t s) const;
    void enqueue_iteration_k(int kp, hipStream_t s, int which = 0) const;
    void enqueue_precondition_k(bool first, int kp, hipStream_t s, int which = 0) const;

    solve_options opts_;
    std::vector<std::uint32_t> relabel_;   // factor position -> storage position (empty: identity)
    mutable int tuned_block_width_ = 0;
    factorization F_;
    trsv_factor T_;
    mutable solver_stats stats_;
    std::int64_t n_ = 0;
    int center_k_ = 10;
    mutable std::uint64_t apply_count_ = 0;
    mutable std::atomic<bool> serving_{false};
    solve_timings setup_times_;
    std::unique_ptr<device_state> dev_;
    std::unique_ptr<timer> timer_;
};

// One-shot solve (solve.h:136-138).
solve_result solve(const csc_view& L, const std::vector<double>& b, const solve_options& opts = {});

// Batch tables of one sweep direction (host side, exposed for tests).
struct batch_table {
    std::vector<std::uint32_t> pos;      // nb+1
    std::vector<std::uint64_t> mask;     // nb
    std::vector<int> level_first;        // levels+1
};
batch_table build_batches(const trsv_factor& T, bool backward);
void validate_sweep_layout(const trsv_factor& T, const batch_table& bt, bool backward);

// Wave table of a blocked sweep over kp columns. Row of `len` off-diagonals
// needs U = G(len) * kp lanes (lane = column * G + chain). A wave with
// mask != 0 packs whole rows (bit = first lane of a row's aligned U-lane
// group, rows at consecutive positions from pos); a wave with mask == 0
// holds columns [col0, col0 + 64 / G) of the row at pos (U > 64).
struct block_batch_table {
    int kp = 0;
    std::vector<std::uint32_t> pos;
    std::vector<std::uint64_t> mask;
    std::vector<std::int32_t> col0;
    std::vector<int> level_first;     // levels + 1
};
block_batch_table build_block_batches(const trsv_factor& T, bool backward, int kp);
void validate_block_batches(const trsv_factor& T, const block_batch_table& bt, bool backward);

// Static tail layout of one direction (device.hpp's tail_desc), run by run.
struct tail_tables {
    std::vector<tail_lane> lane;
    std::vector<tail_lane> lane_recip;   // the same lanes with 1 / diag (fp32-vector sweeps)
    std::vector<tail_pair> ent;
    std::vector<std::int32_t> ovf;
    std::vector<entry32> ovf_ent;
    std::vector<tail_walk> walk;   // per tail batch
    int batches = 0;
};
// Appends the run of batches [bf, bl) (whole levels) and returns its first
// tail batch. Throws if a lane's chain does not have the shape the tail
// kernel relies on (entries before the run form a prefix of the chain).
int append_tail_run(const trsv_factor& T, const batch_table& bt, bool backward, int bf, int bl, tail_tables& tt);
// Appends the walk entries of the run of levels [l, e) (level_first: first
// batch of each level): with rotate, wave w of the tail workgroup takes the
// run's batches w, w + 16, ...; without, batches w, w + 16, ... of every
// level. Returns the run's barrier phases (with merge, a one-batch level after
// a one-batch level of the same wave joins its phase).
int append_tail_walk(const std::vector<int>& level_first, int l, int e, tail_tables& tt, bool merge = true,
                     bool rotate = true);

// Static lane layout of launched levels (device.hpp's lane_rec / entry32
// planes), level by level.
struct lane_tables {
    std::vector<lane_rec> rec;
    std::vector<entry32> ent;
    long long batches = 0;
};
// Appends the batches [bf, bl) of one level and returns its first lane batch.
// Throws on a row longer than 64 * kEpl or a dependency inside the level.
long long append_lane_level(const trsv_factor& T, const batch_table& bt, bool backward, int bf, int bl,
                            lane_tables& lt);

}  // namespace apxchol_amd
