// Host-callable launchers of the gfx950 kernels (defined in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../device.hpp"

namespace apxchol_amd {

enum class vec_prec { fp64 = 0, fp32 = 1 };

// Persistent grid (workgroups) of the dataflow sweep for this device.
int sweep_dataflow_grid(vec_prec p, bool fp16);

// Rows a one-workgroup tail run may hold (its outputs live in LDS).
int sweep_tail_max_rows(vec_prec p);
// Whether a tail run of `rows` rows, `batches` batches and `levels` levels fits the tail workgroup.
bool sweep_tail_fits(vec_prec p, int rows, int batches, int levels);
// LDS bytes of the walk of a tail run of `rows` rows and `batches` batches.
std::size_t sweep_tail_lds(vec_prec p, int rows, int batches);

// How one sweep direction is scheduled.
//   dataflow: one persistent launch of `grid` workgroups (polling);
//   levels:   one launch per level;
//   hybrid:   per segment, either one launch of a single big level
//             (tail == 0) or one single-workgroup launch walking a run of
//             small levels (tail == 1).
struct sweep_sched {
    int kind = 0;                          // 0 dataflow, 1 levels, 2 hybrid
    const int* level_first = nullptr;      // host: first batch of each level (levels+1)
    const int* level_first_dev = nullptr;  // device copy (tail kernel)
    int nlevels = 0;
    const int* seg_l0 = nullptr;           // hybrid segments (host)
    const int* seg_l1 = nullptr;
    const int* seg_tail = nullptr;
    const int* seg_p0 = nullptr;           // tail segments: sweep positions [p0, p1)
    const int* seg_p1 = nullptr;
    const int* seg_base = nullptr;         // tail segments: first tail batch
    const int* seg_phases = nullptr;       // tail segments: barrier phases of the walk
    // Tail segments whose coupling is split (host; -1 or a position ps): the
    // entries before ps are summed on `side` while the segment before runs,
    // forked and joined through the two events; only the rest after it.
    const int* seg_split = nullptr;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int nseg = 0;
    int grid = 0;                          // dataflow grid / cap of a level launch
    tail_desc tail;                        // static layout of the tail runs (device)
    int tail_batches = 0;                  // tail batches of this direction
    void* tail_val = nullptr;              // coupling output, tail_val<T> per (tail batch, lane) (device)
    size_t tail_val_bytes = 0;
    const int* level_lb = nullptr;         // host: first lane batch of a launched level, or -1 (batch table)
    const int* level_pos = nullptr;        // host: first sweep position of each level (nlevels + 1)
    const unsigned char* level_diag = nullptr;  // host: 1 when no row of the level has an entry
    const lane_rec* lane_rec_dev = nullptr;  // static lane layout of the launched levels (device)
    const entry32* lane_ent_dev = nullptr;
    long long lane_batches = 0;
    int lane_bpw2 = 0;                     // lane levels of >= this many batches: 2 per wave (0: never)
    int lane_bpw4 = 0;                     // ... >= this many: 4 per wave (0: never)
};

// Kernel timing: for kernel `kind` (0 sweep_levels, 1 sweep_tail,
// 2 sweep_couple, 3 sweep_dataflow, 4 spmv_pAp) the hook may hand out a
// start / stop event pair, which the launch carries (hipExtLaunchKernelGGL:
// the events take the kernel's own start and end times).
struct launch_hook {
    bool (*fn)(void* ctx, int kind, hipEvent_t* ev) = nullptr;
    void* ctx = nullptr;
    bool events(int kind, hipEvent_t* ev) const { return fn && fn(ctx, kind, ev); }
};

// One sweep. rhs: fp64 r (forward) or the forward output (back, fp64 or fp32
// as `p`); out: fp64 or fp32 as `p` (every slot pending on entry for the
// dataflow schedule).
void launch_sweep(const sweep_desc& d, vec_prec p, bool fp16, const void* rhs, void* out, const pcg_state* st,
                  int* err, const sweep_sched& sc, hipStream_t s, const launch_hook& hook = {});

// z past the sweep rows = 0; with y given, y and z pending (dataflow only).
void launch_fill_outputs(vec_prec p, int n, int m, void* y, void* z, const pcg_state* st, hipStream_t s);

void launch_spmv(int n, const unsigned* ptr, const int* idx, const void* val, bool val32, int lanes, const double* x,
                 double* y, pcg_state* st, double* part, bool want_dot, hipStream_t s,
                 const launch_hook& hook = {});

// y = A x (and x.y into st->pAp when want_dot) over the sliced-ELL form.
void launch_spmv_sell(const sell_desc& a, bool val32, const double* x, double* y, pcg_state* st, double* part,
                      bool want_dot, hipStream_t s, const launch_hook& hook = {});

void launch_scatter_init(int n, const unsigned* perm, const double* b, const double* x0, double* r, double* x,
                         hipStream_t s);
void launch_norm2_b(int n, const double* b, pcg_state* st, double* part, hipStream_t s);
void launch_init_residual(int n, double* r, const double* ax0, pcg_state* st, double* part, hipStream_t s);
// x, r update and its reductions; y, z as in launch_fill_outputs (y null: no sentinels).
void launch_update_xr(vec_prec p, int n, int m, double* x, const double* pv, double* r, const double* ap, void* y,
                      void* z, pcg_state* st, double* part, hipStream_t s);
void launch_dot_center(vec_prec p, int n, const double* r, const void* out, pcg_state* st, double* part, bool first, hipStream_t s);
void launch_update_p(vec_prec p, int n, double* pv, const void* out, const pcg_state* st, bool first, hipStream_t s);
void launch_sum_vec(int n, const double* a, pcg_state* st, double* part, hipStream_t s);
void launch_gather_shift(int n, const unsigned* perm, const double* src, double* dst, const pcg_state* st,
                         bool use_mean, hipStream_t s);

// Stand-alone application pieces.
void launch_sum_shift(int n, const double* r, pcg_state* st, double* part, bool center,
                      hipStream_t s);
void launch_apply_scatter(vec_prec p, int n, int m, const unsigned* perm, const double* r, double* rp, void* y,
                          void* z, hipStream_t s);
void launch_sum_out(vec_prec p, int n, const void* a, pcg_state* st, double* part, hipStream_t s);
void launch_apply_gather(vec_prec p, int n, const unsigned* perm, const void* out, double* z, const pcg_state* st,
                         bool center, hipStream_t s);
void launch_fill_pending(vec_prec p, void* a, int n, hipStream_t s);
// Between factor order and storage order (Q: factor position -> storage
// position; null: the identity): in, dst[Q[q]] = src[q]; out, dst[q] = src[Q[q]].
void launch_permute(int m, const unsigned* Q, const double* src, double* dst, bool in, hipStream_t s);
// dst[q] = double(src[Q[q]]) from a sweep output of the given precision.
void launch_permute_out(vec_prec p, int m, const unsigned* Q, const void* src, double* dst, hipStream_t s);

// ---- blocked solves: kp columns (a power of two <= 64), row-major blocks ----

// Wave table of a blocked sweep for one block width (device arrays) and its
// level boundaries (host: first wave of each level, nlevels + 1 entries).
struct block_sched {
    int kp = 0;
    const unsigned* pos = nullptr;
    const unsigned long long* mask = nullptr;
    const int* col0 = nullptr;
    const int* level_first = nullptr;       // host
    const int* level_first_dev = nullptr;   // device copy (tail runs)
    const int* level_pos = nullptr;         // host: first sweep position of each level (nlevels + 1)
    const int* level_gmax = nullptr;        // host: lanes of the level's widest row (0: a row past 64 lanes)
    int rows_max_g = 16;                    // levels whose widest row has at most this many lanes (1..16)
                                            // run one thread per (row, column)
    int nlevels = 0;
    // Segments: levels [seg_l0, seg_l1), tail flag (one workgroup walks the
    // run with its outputs in LDS) or one launch per level.
    const int* seg_l0 = nullptr;
    const int* seg_l1 = nullptr;
    const int* seg_tail = nullptr;
    int nseg = 0;
    // When set: the single-vector schedule whose levels this sweep follows,
    // its tail runs walked column by column (one tail workgroup per column)
    // with the coupling values in col_tail_val (kp * 64 tail_val per run batch).
    const sweep_sched* cols = nullptr;
    void* col_tail_val = nullptr;
    size_t col_tail_val_bytes = 0;
};
// One blocked sweep as one launch per level.
void launch_sweep_k(const sweep_desc& d, vec_prec p, bool fp16, const block_sched& bsc, const void* rhs, void* out,
                    const pcg_state* st, const batch_state* bs, hipStream_t s);
// Rows a blocked tail run of block width kp may hold (LDS).
int sweep_block_tail_max_rows(vec_prec p, int kp);
void launch_spmm_k(int n, int kp, int lanes, const unsigned* ptr, const int* idx, const void* val, bool val32,
                   const double* x, double* y, pcg_state* st, batch_state* bs, double* part, bool want_dot, hipStream_t s);
void launch_scatter_init_k(int n, int k, int kp, const unsigned* perm, const double* B, double* r, double* x,
                           hipStream_t s);
// The blocked SpMM over the sliced-ELL form (each column = spmv_sell's bits).
void launch_spmm_sell_k(const sell_desc& a, bool val32, int kp, const double* x, double* y, pcg_state* st,
                        batch_state* bs, double* part, bool want_dot, hipStream_t s);
void launch_init_residual_k(int n, int kp, const double* r, pcg_state* st, batch_state* bs, double* part, hipStream_t s);
void launch_update_xr_k(int n, int kp, double* x, const double* p, double* r, const double* ap, pcg_state* st,
                        batch_state* bs, double* part, hipStream_t s);
void launch_refresh_active_k(int kp, const pcg_state* st, batch_state* bs, hipStream_t s);
void launch_dot_center_k(vec_prec p, int n, int kp, const double* r, const void* out, pcg_state* st, batch_state* bs,
                         double* part, bool first, hipStream_t s);
void launch_update_p_k(vec_prec p, int n, int kp, double* pv, const void* out, const pcg_state* st, batch_state* bs,
                       bool first, hipStream_t s);
void launch_zero_tail_k(vec_prec p, int n, int m, int kp, void* a, hipStream_t s);
// x: fp64 (p == fp64) or fp32 block.
void launch_sum_cols_k(vec_prec p, int n, int kp, const void* x, pcg_state* st, double* part, hipStream_t s);
void launch_shift_from_sum_k(int kp, pcg_state* st, bool center, hipStream_t s);
void launch_end_iteration_k(int kp, pcg_state* st, batch_state* bs, hipStream_t s);
// cols, srcs / dsts: device arrays of ncol entries.
void launch_admit_scatter_k(int n, int kp, const unsigned* perm, int ncol, const int* cols,
                            const double* const* srcs, double* r, double* x, hipStream_t s);
void launch_gather_cols_k(int n, int kp, const unsigned* perm, int ncol, const int* cols, double* const* dsts,
                          const double* x, const pcg_state* st, bool use_mean, hipStream_t s);
void launch_gather_k(vec_prec p, int n, int k, int kp, const unsigned* perm, const void* x, double* X,
                     const pcg_state* st, bool use_mean, hipStream_t s);

// ---- grouped launches: several factors' PCG iterations, one launch per kernel kind ----
//
// A member is one factor's solver state (its own vectors, partials, solve
// state and SpMV form); a unit is one step of one member's sweep (a launched
// level, or a tail run = its coupling launch + its one-workgroup walk). A
// grouped launch concatenates the members' (units') own grids; starts[i] is
// the first workgroup of entry i, starts[count] the launch's grid (device
// arrays).
struct group_member {
    int n = 0, m = 0;
    double* r = nullptr;
    double* x = nullptr;
    double* p = nullptr;
    double* ap = nullptr;
    void* y = nullptr;                      // sweep outputs (the solver's vector precision)
    void* z = nullptr;
    pcg_state* st = nullptr;
    double* part = nullptr;
    int use_sell = 0, val32 = 0, lanes = 1;
    sell_desc sell;
    const unsigned* optr = nullptr;         // CSR operator (use_sell == 0)
    const int* oidx = nullptr;
    const void* oval = nullptr;
};
inline constexpr int kGroupLane = 0, kGroupLevels = 1, kGroupTail = 2;
// Workgroups of one level launch from the batch tables: one wave per batch, at most this many;
// waves per workgroup of the level launches.
inline constexpr int kLevelLaunchMaxBlocks = 16384;
inline constexpr int kLevelWavesPerBlock = 4;
struct group_unit {
    int kind = kGroupLane;
    sweep_desc d;
    int lb0 = 0, nlb = 0;                   // lane level: lane batches [lb0, lb0 + nlb)
    const lane_rec* rec = nullptr;
    const entry32* ent = nullptr;
    int b0 = 0, b1 = 0;                     // batch-table level / tail run: batches [b0, b1)
    int p0 = 0;                             // tail run: its first sweep position
    tail_desc td;                           // tail run: the direction's tail layout
    int base = 0, nlev = 0, zslot = 0, row0 = 0;
    void* tv = nullptr;                     // tail run: its coupling values
    const void* rhs = nullptr;
    void* out = nullptr;
    const pcg_state* st = nullptr;
    int* err = nullptr;
    int grid = 0;                           // workgroups of the unit's own launch (couple: batches)
    std::size_t lds = 0;                    // tail run: LDS bytes of its walk
};
// The own grid of a member's single-vector SpMV and of its vector passes.
int group_spmv_blocks(const group_member& g);
int group_vec_blocks(int n);
void launch_group_spmv(const group_member* mem, const int* starts, int count, int blocks, hipStream_t s);
void launch_group_finish(int op, const group_member* mem, const int* starts, int count, hipStream_t s);
void launch_group_update_xr(vec_prec p, const group_member* mem, const int* starts, int count, int blocks,
                            hipStream_t s);
void launch_group_fill_outputs(vec_prec p, const group_member* mem, const int* starts, int count, int blocks,
                               hipStream_t s);
void launch_group_dot_center(vec_prec p, const group_member* mem, const int* starts, int count, int blocks,
                             hipStream_t s);
void launch_group_update_p(vec_prec p, const group_member* mem, const int* starts, int count, int blocks, bool first,
                           hipStream_t s);
// One step of the members' sweeps in one direction: the launched levels, the
// coupling launches of the tail runs, then the tail runs (count units each).
void launch_group_sweep_level(vec_prec p, bool fp16, bool backward, const group_unit* units, const int* starts,
                              int count, int blocks, hipStream_t s);
void launch_group_sweep_couple(vec_prec p, bool fp16, bool backward, const group_unit* units, const int* starts,
                               int count, int blocks, hipStream_t s);
void launch_group_sweep_tail(vec_prec p, const group_unit* units, int count, std::size_t lds, hipStream_t s);
// The reduction-stage op codes of pcg_finish (kFinPAp ...), for the group host.
int group_fin_op(const char* name);

// ---- stand-alone PCG pieces (pcg_cuda_kernels.h:43-93): caller-owned partials,
// one double per producer workgroup, pcg_blocks(units) of them ----
int pcg_blocks(long long units);
// y = A x over a full CSR (lanes per row: a power of two <= 64) and the
// partials of x.y, pcg_blocks(n * lanes).
void launch_piece_spmv_pAp(int n, const unsigned* ptr, const int* idx, const void* val, bool val32, int lanes,
                           const double* x, double* y, double* part, hipStream_t s);
// x += alpha p; r -= alpha Ap; partials of r.r (pcg_blocks(n)).
void launch_piece_update_xr(int n, double* x, const double* p, double* r, const double* ap, double alpha, double* part,
                            hipStream_t s);
// Partials of a.b (pcg_blocks(n)).
void launch_piece_dot(int n, const double* a, const double* b, double* part, hipStream_t s);
// p = z + beta p.
void launch_piece_update_p(int n, double* p, const double* z, double beta, hipStream_t s);
// out[0] = the fixed-order sum of part[0..nb).
void launch_piece_reduce(const double* part, int nb, double* out, hipStream_t s);
// launch_piece_spmv_pAp then launch_piece_reduce with the partials in the
// library's per-device scratch array (allocates nothing): y = A x and
// pAp[0] = x.y, the bits of the two-step form. Calls that overlap on one
// device race on the scratch.
void launch_piece_spmv_pAp_scratch(int n, const unsigned* ptr, const int* idx, const void* val, bool val32, int lanes,
                                   const double* x, double* y, double* pAp, hipStream_t s);

}  // namespace apxchol_amd
