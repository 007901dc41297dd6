/* apxchol_amd -- C ABI of the MI355X (gfx950) solve path of apxchol.
 *
 * Plain C types only: host arrays are host pointers, "d_" arguments are
 * device pointers (HIP device memory of the current device), `stream` is a
 * hipStream_t passed as void* (NULL = the default stream). No call throws;
 * every call returns an apx_status and apx_last_error() holds the message of
 * the last failure on the calling thread.
 *
 * Each entry point names the reference interface it replaces (paths are in
 * the reference tree AlgOptGroup/apxchol).
 */
#ifndef APXCHOL_AMD_H
#define APXCHOL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    APX_OK = 0,
    APX_ERR_INVALID = 1, /* std::invalid_argument in the reference (bad input / misuse) */
    APX_ERR_RUNTIME = 2, /* std::runtime_error in the reference (device / numerical failure) */
    APX_ERR_MEMORY = 3
} apx_status;

typedef struct apx_factor_s* apx_factor_t; /* host factorization (factorization struct) */
typedef struct apx_solver_s* apx_solver_t; /* factor + operator resident on one GPU */

/* solve_options / factor_options (include/apxchol/solver/solve.h:13-25,
 * include/apxchol/solver/factor_options.h:56-156) plus the device knobs. */
typedef struct {
    double tol;              /* default 1e-8 */
    int32_t max_iter;        /* default 200 */
    int32_t stagnation_window; /* default 50; 0 disables */
    uint32_t seed;           /* factor seed (reference default 42) */
    int32_t precision;       /* sweep vectors: 0 = fp64, 1 = fp32 (factor values are fp32 or fp16) */
    int32_t fp16_storage;    /* 1 = per-column-scaled fp16 factor values */
    int32_t schedule;        /* 0 = sync-free dataflow, 1 = level-set launches, 2 = hybrid (default):
                                big levels launched, runs of small levels walked by one workgroup */
    double drop_rel;         /* < 0: APXCHOL_FACTOR_DROP or 1e-4; 0 disables the drop */
    int32_t keep_factor;     /* keep the host factor values for apx_solver_export_factor */
    int32_t iters_per_check; /* PCG iterations enqueued between completion tests (default 2) */
    int32_t sweep_chunk;     /* batches claimed per dataflow ticket (0 = default) */
    int32_t sweep_blocks;    /* dataflow workgroups / level-launch cap (0 = the device's resident grid) */
    int32_t tail_rows;       /* hybrid: levels with fewer rows go to the one-workgroup tail (0 = default) */
    int32_t graph;           /* 1 (default): each PCG iteration replays one captured HIP graph;
                                0: eager launches (also used while kernel timing is on) */
    int32_t relabel;         /* 1 (default): device vectors order every level's rows by original
                                vertex (same factor, local gathers); 0: the factor order */
    int32_t lane_batches;    /* launched levels of at most this many batches read the static lane
                                layout (0 = default 4096, < 0 = never: batch tables only) */
    int32_t depth_levels;    /* 1 (default, with relabel): sweep levels = topological depth of the
                                dependency graph; 0: the elimination rounds (round-as-level) */
    int32_t tail_batches;    /* hybrid: a tail level holds at most this many batches (0 = 48, max 64;
                                the tail workgroup's 16 waves take a run's batches round robin) */
    int32_t tail_cut;        /* hybrid: a tail run ends before a level one of whose lanes would read
                                at least this many groups of 8 entries inside the run from HBM
                                (rows of thousands of entries); that level's entries are then summed
                                by the device-wide coupling launch (0 = 1, < 0 = never) */
} apx_options;

/* Fill `o` with the defaults above. */
void apx_options_default(apx_options* o);

/* solve_result (solve.h:27-42). status: 1 converged, 2 pAp <= 0 breakdown,
 * 3 stagnation, 4 max_iter reached, 5 zero right-hand side. */
typedef struct {
    int64_t iterations;
    double residual; /* recurrence residual ||r|| / ||b|| */
    int64_t lumped_offdiag;
    double solve_vram_mb;
    int32_t status;
    /* solve_result::timings (the checkpoint tree of src/solve.cpp:681-737 and
     * preconditioner.h:238-286), host wall clock in milliseconds. The setup
     * phases are the solver's (the same for every solve of one solver);
     * factorize_ms is 0 when the solver was created with a given factor. */
    double factorize_ms;     /* setup/factorize */
    double sptrsv_setup_ms;  /* setup/sptrsv_setup: trsv preparation, device layouts and uploads */
    double pcg_setup_ms;     /* setup/gpu_pcg_setup: the permuted operator (CSR + sliced ELL) */
    double pcg_ms;           /* pcg/gpu_pcg_loop: this solve, its final synchronization included
                                (a served request: its submit -> result latency) */
} apx_solve_info;

typedef struct {
    int64_t n, m;
    int32_t sddm;
    uint64_t factor_nnz, stored_nnz, offdiag, dropped;
    int64_t rounds, levels, fwd_batches, bck_batches, op_nnz;
    int32_t op_fp32, spmv_lanes, dataflow_grid, fp16_storage;
    uint64_t device_bytes;
    /* Per direction [0 forward, 1 back], level / hybrid schedules: algorithmic
     * HBM bytes of one sweep's level launches and tail runs and their kernel
     * counts (off-diagonals x (4 + value width) + rows x (8 + input + output
     * width) + 12 per batch). */
    uint64_t level_bytes[2], tail_bytes[2];
    int64_t level_launches[2], tail_runs[2];
    int64_t lane_levels[2];  /* launched levels that read the static lane layout */
    int32_t spmv_format;     /* single SpMV: 0 CSR lanes, 1 sliced ELL (APXCHOL_AMD_SPMV=csr|sell forces) */
    uint64_t spmv_sell_slots; /* padded entries of the sliced-ELL form (0: none built) */
    int64_t split_couplings[2]; /* per direction: tail runs whose coupling is split in two launches */
} apx_solver_info;

const char* apx_last_error(void);

/* ---- host factorization -------------------------------------------------
 * apxchol::factorize(L, storage, opts, cp) (include/apxchol/solver/factorization.h:156-159,
 * src/factorization.cpp:125-175) with vec_pool storage and block_greedy: the
 * reference's one-thread factor, bit for bit, whatever number of host threads
 * runs it (APXCHOL_AMD_FACTOR_THREADS, else the CPUs of this process's
 * affinity capped by OMP_NUM_THREADS and 16). A: n x n symmetric Laplacian / SDDM M-matrix in CSC with BOTH
 * triangles stored (what Eigen::SparseMatrix<double> holds). */
apx_status apx_factorize(int64_t n, const int64_t* colptr, const int32_t* rowidx, const double* vals,
                         uint32_t seed, apx_factor_t* out);
/* out[0..6] = n, nnz, rounds, sddm, lumped_offdiag, host threads used, 0 */
apx_status apx_factor_info(apx_factor_t f, int64_t* out);
/* factorization::L (CSC, diagonal first; outer n+1, inner/vals nnz), perm (n),
 * per-round IS sizes (rounds) -- factorization.h:22-56. */
apx_status apx_factor_export(apx_factor_t f, uint64_t* outer, uint32_t* inner, float* vals, uint32_t* perm,
                             uint64_t* round_is);
void apx_factor_destroy(apx_factor_t f);

/* ---- fp16 storage conversions (no GPU) -----------------------------------
 * fp16_t(float) and fp16_t::to_float (include/apxchol/lowprec.h:111-201):
 * IEEE binary16 round-to-nearest-even with subnormals, overflow to infinity
 * and NaN kept; the widening is exact. (The factor storage additionally
 * flushes stored subnormals to signed zero, omp.h:353-366.) */
uint16_t apx_fp16_round(float x);
float apx_fp16_widen(uint16_t h);

/* ---- host SpTRSV preparation (no GPU) ------------------------------------
 * omp_sptrsv::setup / cuda_sptrsv::setup host part (include/apxchol/solver/sptrsv/omp.h:505-861,
 * cuda.h:233-459, cuda_host.h:62-232): L11, compacting drop, fp16 narrowing,
 * the forward (rows of L) and back (columns of L) off-diagonal lists and the
 * round-as-level boundaries. sizes[0..5] = m, offdiag, levels, nnz_factor, nnz_stored, fp16. */
apx_status apx_trsv_sizes(apx_factor_t f, int32_t fp16, double drop_rel, int64_t* sizes);
apx_status apx_trsv_export(apx_factor_t f, int32_t fp16, double drop_rel, uint32_t* f_ptr, int32_t* f_idx,
                           float* f_val, uint16_t* f_h, uint32_t* b_ptr, int32_t* b_idx, float* b_val, uint16_t* b_h,
                           float* diag, float* inv_scale, uint32_t* level_ptr);

/* ---- GPU solver ----------------------------------------------------------
 * cpu_solver(L, opts) / cpu_solver(L, F, opts) (solve.h:66-76, src/solve.cpp:277-318):
 * factorize (or adopt `f`, which is not consumed), prepare and upload. */
apx_status apx_solver_create(int64_t n, const int64_t* colptr, const int32_t* rowidx, const double* vals,
                             const apx_options* opts, apx_solver_t* out);
apx_status apx_solver_create_with_factor(int64_t n, const int64_t* colptr, const int32_t* rowidx,
                                         const double* vals, apx_factor_t f, const apx_options* opts,
                                         apx_solver_t* out);
void apx_solver_destroy(apx_solver_t s);
apx_status apx_solver_info_get(apx_solver_t s, apx_solver_info* info);

/* cpu_solver::solve(b, x, tol, max_iter, x0) (solve.h:83-100, src/solve.cpp:478-654),
 * on device vectors of n doubles. tol < 0 / max_iter < 0 use the options;
 * d_x0 may be NULL. The PCG runs on the GPU (cuda_pcg::solve, pcg_cuda.h:157-261). */
apx_status apx_solver_solve(apx_solver_t s, const double* d_b, double* d_x, const double* d_x0, double tol,
                            int32_t max_iter, apx_solve_info* info, void* stream);
/* Host-vector convenience form of apx_solver_solve (b, x, x0 host arrays of n). */
apx_status apx_solver_solve_host(apx_solver_t s, const double* b, double* x, const double* x0, double tol,
                                 int32_t max_iter, apx_solve_info* info);

/* k right-hand sides in one blocked PCG. No counterpart in the reference: it
 * stands for k calls of cpu_solver::solve(b) (solve.h:83-100) run together.
 * Column c of d_B / d_X is n doubles at d_B + c*n / d_X + c*n (device
 * memory). Every column keeps its own convergence test (tol, max_iter,
 * stagnation) and stops as its own solve would; the sweeps and the SpMM give
 * each column the bits of the single-vector kernels, the per-column dot
 * products are reduced in another fixed order. infos: k entries, or NULL. */
apx_status apx_solver_solve_batch(apx_solver_t s, int32_t k, const double* d_B, double* d_X, double tol,
                                  int32_t max_iter, apx_solve_info* infos, void* stream);
/* Host-array form of apx_solver_solve_batch (B, X: k * n doubles). */
apx_status apx_solver_solve_batch_host(apx_solver_t s, int32_t k, const double* B, double* X, double tol,
                                       int32_t max_iter, apx_solve_info* infos);

/* k preconditioner applications z_c = M^-1 r_c at once: k calls of
 * cpu_solver::apply (solve.h:103) counted as ONE application of the center-k
 * schedule (no counterpart in the reference). Column c at d_R + c*n /
 * d_Z + c*n. Non-centring applications give each column the single apply's
 * bits. */
apx_status apx_solver_apply_batch(apx_solver_t s, int32_t k, const double* d_R, double* d_Z, void* stream);

/* Autotuner (no counterpart in the reference): measures, for this solver on
 * this device, the SpMV lanes per row (1..64; the solver then uses the
 * fastest, for the single SpMV and the blocked SpMM alike; the single SpMV
 * keeps the sliced-ELL form instead when that is faster still) and the block
 * width of blocked solves up to max_block_width (the most columns swept per
 * second; a server created with slots = 0 uses it). Times in microseconds:
 * spmv_us[i] for 2^i lanes, sweep_pair_us[i] for block width 8 << i. */
typedef struct {
    int32_t spmv_lanes, block_width;
    double spmv_us[7];
    double sweep_pair_us[4];
    int32_t spmv_sell;       /* 1: the single SpMV now runs on the sliced-ELL form (faster than any lanes) */
    double spmv_sell_us;     /* its time (0: the solver has no sliced-ELL form) */
} apx_autotune_result;
apx_status apx_solver_autotune(apx_solver_t s, int32_t max_block_width, apx_autotune_result* out, void* stream);

/* ---- serving (no counterpart in the reference) ---------------------------
 * A request loop over resident factors: each factor keeps one blocked PCG of
 * `slots` columns running on its own stream; requests are admitted into free
 * columns every `iters_per_check` iterations and released when they finish
 * (continuous batching), and factors are admitted against `memory_budget`
 * bytes of HBM (0: 90% of the free memory). A request run in a block gets
 * the same bits alone or among others in a block of that width. Each factor
 * also has a narrow block (apx_server_set_narrow_block) filled before the
 * main one, and a single lane (on by default): one request at a time runs
 * there as a single-vector solve, side by side with the blocks on a stream
 * of its own, and the oldest waiting request takes it whenever it is free;
 * such a request gets the bits of apx_solver_solve with its tol and
 * max_iter. A solver added to a server is used by the server only (the
 * server keeps it alive after apx_solver_destroy). */
typedef struct apx_server_s* apx_server_t;
typedef struct {
    int64_t submitted, completed, factors;
    double p50_ms, p99_ms, mean_ms; /* submit -> completion latency */
    double solves_per_s;            /* completed / (last completion - first submission) */
    uint64_t memory_used, memory_budget;
    int64_t single_lane_completed;  /* completed requests that ran on a single lane */
} apx_server_stats;
apx_status apx_server_create(uint64_t memory_budget, int32_t slots, int32_t iters_per_check, apx_server_t* out);
void apx_server_destroy(apx_server_t srv);
/* Single lanes on (1, the default) or off (0); a request already on one finishes there. */
apx_status apx_server_set_single_lane(apx_server_t srv, int32_t on);
/* Columns of the narrow block of factors added afterwards (default 4; 0: none;
 * used only when narrower than `slots`): a second block per factor, on its
 * own stream, that takes requests before the main block does. */
apx_status apx_server_set_narrow_block(apx_server_t srv, int32_t slots);
apx_status apx_server_add(apx_server_t srv, apx_solver_t solver, int32_t* factor_id);
/* b: n host doubles (copied); tol / max_iter as in apx_solver_solve. */
apx_status apx_server_submit(apx_server_t srv, int32_t factor, const double* b, double tol, int32_t max_iter,
                             int64_t* request_id);
/* Waits up to timeout_s seconds (< 0: no limit); *ready = 1 and x (n host
 * doubles) / info filled when the request is done (it is then forgotten).
 * A request whose factor lane ended on a device error returns
 * APX_ERR_RUNTIME with that error; submissions to such a lane are refused
 * the same way. */
apx_status apx_server_wait(apx_server_t srv, int64_t request_id, double* x, apx_solve_info* info, double timeout_s,
                           int32_t* ready);
apx_status apx_server_get_stats(apx_server_t srv, apx_server_stats* st);

/* ---- many small factors in one launch (no counterpart in the reference) ---
 * A group runs the PCG solves of several resident solvers together: every
 * kernel kind of an iteration (SpMV, vector passes, reductions, each step of
 * the two sweeps) is ONE launch over all members, each member's share being
 * the launch its own solve would make. Meant for many small factors, none of
 * which fills the GPU alone. Members: 1..64 solvers of one vector precision
 * and factor storage on the hybrid schedule (the default); the group keeps
 * them alive and uses their device arrays while it solves (a member runs no
 * other solve meanwhile; a serving solver is refused). */
typedef struct apx_group_s* apx_group_t;
apx_status apx_group_create(const apx_solver_t* solvers, int32_t count, apx_group_t* out);
void apx_group_destroy(apx_group_t g);
/* Members and the launches of one grouped PCG iteration (all members). */
apx_status apx_group_info(apx_group_t g, int32_t* members, int32_t* launches_per_iteration);
/* One right-hand side per member: d_b[i], d_x[i] are member i's n doubles
 * (device, its caller order); tol / max_iter as in apx_solver_solve. Every
 * member's x and info are the bits of its own apx_solver_solve(b, x, NULL,
 * tol, max_iter). infos: count entries, or NULL. */
apx_status apx_group_solve(apx_group_t g, const double* const* d_b, double* const* d_x, double tol, int32_t max_iter,
                           apx_solve_info* infos, void* stream);

/* cpu_solver::apply(r) = apx_cholesky::_solve_impl (solve.h:103, preconditioner.h:344-353):
 * z = M^-1 r with the center-k schedule (every K-th application centres). */
apx_status apx_solver_apply(apx_solver_t s, const double* d_r, double* d_z, void* stream);
apx_status apx_solver_apply_host(apx_solver_t s, const double* r, double* z);
/* apx_cholesky::reset_apply_count (preconditioner.h:327) */
apx_status apx_solver_reset_apply_count(apx_solver_t s);

/* One triangular sweep in factor order with fp64 vectors: forward L y = x
 * (backward = 0) or back L^T z = y (backward = 1) on the stored L11 -- the
 * device boundary of dataflow_solve (include/apxchol/solver/sptrsv/cuda_dataflow.h:185-208)
 * and omp_sptrsv::forward_solve / transpose_solve (omp.h:1048-1058). d_rhs, d_out: m doubles. */
apx_status apx_solver_sweep(apx_solver_t s, int32_t backward, const double* d_rhs, double* d_out, void* stream);

/* The sweep pair y = L^-T L^-1 x on the stored L11 in factor order (m
 * doubles each; d_x == d_y allowed), the two sweeps in the solver's vector
 * precision -- fp32 vectors and accumulation when the solver was created with
 * precision = 1, fp64 otherwise -- with fp64 at the boundary:
 * cuda_sptrsv::solve_LLt_dev (include/apxchol/solver/sptrsv/cuda.h:480-487,
 * 736-785), the cast kernels folded into the boundary permutations. */
apx_status apx_solver_solve_LLt(apx_solver_t s, const double* d_x, double* d_y, void* stream);

/* y = L x in the caller's order with the solver's single-vector SpMV (sliced
 * ELL or CSR lanes, fp64 accumulation) -- the device boundary of
 * pcg_cuda::spmv_pAp (include/apxchol/solver/pcg_cuda_kernels.h:75-93) without
 * the dot product. d_x, d_y: n doubles. */
apx_status apx_solver_spmv(apx_solver_t s, const double* d_x, double* d_y, void* stream);
/* Y = L X for k columns (column c at d_X + c*n / d_Y + c*n) with the blocked
 * SpMM (no counterpart in the reference); every column gets apx_solver_spmv's
 * bits. */
apx_status apx_solver_spmv_batch(apx_solver_t s, int32_t k, const double* d_X, double* d_Y, void* stream);

/* Kernel timing with HIP events on the launching stream (no reference
 * counterpart; the reference's APXCHOL_PCG_TRACE per-component ms/iter,
 * src/solve.cpp:578-653, is the closest). While timing is on, solves launch
 * eagerly. Classes (APX_TIMING_CLASSES): 0 forward sweep, 1 back sweep, 2 SpMV,
 * 3 vector passes (whole sweeps / passes, gaps between their kernels included);
 * then single kernels: 4 sweep_levels, 5 sweep_tail, 6 sweep_couple (both
 * directions). Launches of iterations after convergence are excluded.
 * apx_solver_take_timing returns and resets the totals. */
#define APX_TIMING_CLASSES 7
apx_status apx_solver_set_timing(apx_solver_t s, int32_t on);
apx_status apx_solver_take_timing(apx_solver_t s, double* ms /* APX_TIMING_CLASSES */,
                                  int64_t* launches /* APX_TIMING_CLASSES */);

/* Host copies of the factor kept with keep_factor = 1 (factorization::L, perm, rounds). */
apx_status apx_solver_export_factor(apx_solver_t s, uint64_t* outer, uint32_t* inner, float* vals, uint32_t* perm);

/* apxchol::solve(L, b, opts) one-shot (solve.h:136-138, src/solve.cpp:681-737), host vectors. */
apx_status apx_solve(int64_t n, const int64_t* colptr, const int32_t* rowidx, const double* vals, const double* b,
                     double* x, const apx_options* opts, apx_solve_info* info);

/* ---- stand-alone PCG pieces on caller-owned device arrays ------------------
 * The device boundary of the reference's GPU PCG
 * (include/apxchol/solver/pcg_cuda_kernels.h:43-93, src/cuda_pcg_kernels.cu:18-152):
 * every reduction writes one partial per workgroup of its producer into the
 * caller's `d_part` (at least APX_PCG_MAX_BLOCKS doubles), over a grid that
 * is a pure function of the size (apx_pcg_blocks), and apx_pcg_reduce_partials
 * sums them in a fixed order -- bit-identical run to run. */
#define APX_PCG_MAX_BLOCKS 2048
/* pcg_blocks(units): min(ceil(units / 256), APX_PCG_MAX_BLOCKS), at least 1. */
int32_t apx_pcg_blocks(int64_t units);
/* spmv_lanes_for(avg): the power of two nearest avg_nnz_per_row in log2, in [1, 32]. */
int32_t apx_pcg_spmv_lanes_for(double avg_nnz_per_row);
/* spmv_pAp: y = A x over a full CSR (d_rowptr n+1 non-negative int32,
 * d_colidx nnz int32, d_vals nnz fp64, or fp32 when vals_fp32 = 1, each
 * product promoted to fp64), `lanes` threads per row (a power of two <= 64),
 * and the partials of x.y into d_part[0 .. apx_pcg_blocks(n * lanes)). */
apx_status apx_pcg_spmv_pAp(int32_t n, const int32_t* d_rowptr, const int32_t* d_colidx, const void* d_vals,
                            int32_t vals_fp32, const double* d_x, double* d_y, double* d_part, int32_t lanes,
                            void* stream);
/* update_xr: x += alpha p; r -= alpha Ap; partials of r.r (the updated r)
 * into d_part[0 .. apx_pcg_blocks(n)). */
apx_status apx_pcg_update_xr(int64_t n, double* d_x, const double* d_p, double* d_r, const double* d_Ap, double alpha,
                             double* d_part, void* stream);
/* dot: partials of a.b into d_part[0 .. apx_pcg_blocks(n)). */
apx_status apx_pcg_dot(int64_t n, const double* d_a, const double* d_b, double* d_part, void* stream);
/* update_p: p = z + beta p. */
apx_status apx_pcg_update_p(int64_t n, double* d_p, const double* d_z, double beta, void* stream);
/* reduce_partials: d_out[0] = the fixed-order sum of d_part[0 .. blocks). */
apx_status apx_pcg_reduce_partials(const double* d_part, int32_t blocks, double* d_out, void* stream);
/* spmv_pAp_noalloc: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops and
 * harnesses that forbid hipMalloc). y = A x as apx_pcg_spmv_pAp, and
 * d_pAp[0] = x.y, the bits of apx_pcg_spmv_pAp followed by
 * apx_pcg_reduce_partials. The partials live in a library-owned device array
 * of APX_PCG_MAX_BLOCKS doubles (16 KiB), one per GPU, held until the process
 * ends; it covers every n (the producer grid never exceeds it), and a grid
 * beyond it is refused with APX_ERR_INVALID. Allocates nothing. Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (streams
 * with no order between them) race. n = 0 writes d_pAp[0] = 0. */
apx_status apx_pcg_spmv_pAp_noalloc(int32_t n, const int32_t* d_rowptr, const int32_t* d_colidx, const void* d_vals,
                                    int32_t vals_fp32, const double* d_x, double* d_y, double* d_pAp, int32_t lanes,
                                    void* stream);

/* ---- MatrixMarket files and the reading of an input matrix -----------------
 * The reference CLI's input layer (src/main.cpp:14-98 with the
 * fast_matrix_market reader; src/mtx_input.h:36-120): read a .mtx file, decide
 * whether it holds an assembled Laplacian/SDDM operator or a graph adjacency
 * matrix, and assemble L = D - A for the latter. Host-only. */
typedef struct apx_matrix_s* apx_matrix_t; /* host CSC matrix, rows ascending in each column */
/* Reads a `matrix coordinate` file (real / integer / pattern, general /
 * symmetric; symmetric files mirrored, duplicates summed). pattern_field: 1 for
 * a `pattern` file (may be null). */
apx_status apx_mtx_read(const char* path, apx_matrix_t* out, int32_t* pattern_field);
apx_status apx_matrix_create(int64_t n, const int64_t* colptr, const int32_t* rowidx, const double* vals,
                             apx_matrix_t* out);
apx_status apx_matrix_size(apx_matrix_t m, int64_t* rows, int64_t* cols, int64_t* nnz);
apx_status apx_matrix_export(apx_matrix_t m, int64_t* colptr, int32_t* rowidx, double* vals);
void apx_matrix_destroy(apx_matrix_t m);
/* Writes m (symmetric = 1: lower triangle with a `symmetric` header), 17 digits. */
apx_status apx_mtx_write(const char* path, apx_matrix_t m, int32_t symmetric);
/* Dense vectors (`matrix array` n x 1, or coordinate n x 1). Reading: *n gets
 * the length; x (capacity cap) is filled when cap >= *n. */
apx_status apx_mtx_read_vector(const char* path, double* x, int64_t cap, int64_t* n);
apx_status apx_mtx_write_vector(const char* path, const double* x, int64_t n);

/* input_scan (src/mtx_input.h:46-73). */
typedef struct {
    int64_t n, offdiag_nnz, offdiag_pos, offdiag_neg, upper_nnz, lower_nnz, diag_pos, deficient_rows;
    double worst_deficit;
    int64_t worst_row, excess_rows;
} apx_input_scan;
/* Input kinds: 0 automatic, 1 laplacian (assembled operator), 2 adjacency. */
apx_status apx_input_scan_get(apx_matrix_t m, apx_input_scan* s);
/* resolve_input_kind (mtx_input.h:76-90): *kind = 1 or 2 and the reason
 * (truncated to reason_len bytes, NUL included); a one-sided or ambiguous
 * matrix returns APX_ERR_RUNTIME with the explanation in apx_last_error(). */
apx_status apx_input_resolve(int32_t requested, const apx_input_scan* s, int32_t pattern_field, int32_t* kind,
                             char* reason, int64_t reason_len);
/* adjacency_to_laplacian (mtx_input.h:92-97): *out = D - |A| with self-loops
 * dropped and an explicit diagonal entry for every vertex. */
apx_status apx_adjacency_to_laplacian(apx_matrix_t a, apx_matrix_t* out);
/* describe_input / diagnose_failure (mtx_input.h:99-120), truncated to out_len. */
apx_status apx_input_describe(int32_t kind, const apx_input_scan* s, const char* reason, char* out, int64_t out_len);
apx_status apx_input_diagnose(int32_t kind, const apx_input_scan* s, char* out, int64_t out_len);

#ifdef __cplusplus
}
#endif
#endif /* APXCHOL_AMD_H */
