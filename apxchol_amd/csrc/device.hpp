// Descriptors shared by the host code and the gfx950 kernels. Plain structs of
// device pointers, passed by value as kernel arguments.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace apxchol_amd {

// Off-diagonal entry of the fp32 storage: column/row id and value bits side
// by side, so one 8-byte load brings both (coalesced across a lane group).
struct alignas(8) entry32 {
    std::int32_t idx;
    float val;
};

// Entries per lane per chunk of the sweep: a row of len off-diagonals gets
// G = pow2ceil(ceil(len / kEpl)) lanes (1..64); rows longer than 64 * kEpl
// are walked by the whole wave in chunks of 64 * kEpl.
inline constexpr int kEpl = 4;
inline constexpr int kWave = 64;

// Sentinels marking "not yet computed" in the sweep output vectors: NaN
// payloads that IEEE arithmetic never produces (it returns the canonical
// quiet NaN), so a finished row can never look unfinished.
inline constexpr std::uint64_t kPending64 = 0xFFF4D00DFEEDF00DULL;
inline constexpr std::uint32_t kPending32 = 0xFFA0D00DU;

// One triangular sweep over m rows in SWEEP ORDER (position t = row t for the
// forward sweep, row m-1-t for the back sweep). Rows are packed into batches
// of consecutive positions; batch b covers positions [batch_pos[b],
// batch_pos[b+1]) and bit k of batch_mask[b] marks the first lane of a row's
// lane group (groups are aligned to their power-of-two size).
struct sweep_desc {
    int m = 0;
    int backward = 0;                       // 0: forward on rows of L, 1: back on columns of L
    const std::uint32_t* ptr = nullptr;     // m+1, per ROW id
    const entry32* ent = nullptr;           // fp32 storage
    const std::int32_t* idx = nullptr;      // fp16 storage: ids
    const std::uint16_t* h = nullptr;       // fp16 storage: values
    const float* diag = nullptr;            // m
    const float* inv_diag = nullptr;        // m, fp32 1 / diag (the fp32-vector sweeps multiply by it)
    const float* inv_scale = nullptr;       // m, fp16 back sweep input scale (else null)
    const std::uint32_t* batch_pos = nullptr;  // nb+1
    const std::uint64_t* batch_mask = nullptr; // nb
    int nb = 0;
    int chunk = 1;                          // batches claimed per ticket (dataflow schedule)
    int* ctrl = nullptr;                    // 2 ints: ticket, finished waves (zero between sweeps)
};

// Sliced-ELL form of the PCG operator (the single-vector SpMV). Slice s holds
// 64 rows (lane = row rowmap[s*64 + lane]; -1: none) and soff[s+1] - soff[s]
// entry columns:
//   idx/val[(soff[s] + k)*64 + lane]   entry k of the lane's row, CSR order
//                                      (padding: the row itself / 0, value 0)
// Rows too long for a slice are listed in longrows and read from the CSR
// arrays (ptr, cidx, cval) by a whole wave.
struct sell_desc {
    int n = 0;
    int nslices = 0;
    int nlong = 0;
    int max_blocks = 2048;                  // workgroup cap of the launch (partials buffer size)
    const std::uint32_t* soff = nullptr;
    const std::int32_t* idx = nullptr;
    const void* val = nullptr;              // fp64, or fp32 when every value is fp32-exact
    const std::int32_t* rowmap = nullptr;
    const std::int32_t* longrows = nullptr;
    const std::uint32_t* ptr = nullptr;
    const std::int32_t* cidx = nullptr;
    const void* cval = nullptr;
};

// Device-resident PCG state of one solve (all reductions land here; the host
// reads it only to test completion).
struct pcg_state {
    double rz, pAp, rr, rs, bnorm, tol, prev_check, residual;
    double shift;        // input-centring shift of the next preconditioner application
    double out_sum;      // sum of the last application's output
    double mu;           // output-centring shift of the last application
    double beta;
    double pad[4];
    int iter, done, max_iter, stag_window;
    int apply_count, center_k, laplacian, center_now;
    int status;          // 0 running, 1 converged, 2 breakdown (pAp <= 0), 3 stagnation, 4 max_iter,
                         // 5 zero right-hand side
    int n;
    int phase;           // blocked solves: 0 iterating, 1 joining (first application pending)
    int pad_i;
};

// Bookkeeping of a blocked (k right-hand sides) solve: the number of columns
// still iterating; every blocked kernel returns at once when it is zero.
struct batch_state {
    int active;
    int pad[15];
};

// Static layout of the TAIL RUNS: maximal runs of small levels that one
// workgroup walks with the run's outputs in LDS. Everything a lane needs for a
// batch of a run sits at an address computed from (batch, lane) alone, so the
// next batches' loads are independent and stay in flight across the level
// barriers. Tail batches of a direction are numbered consecutively, run by run.
//   lane[t*64 + lane]                what the lane does in tail batch t
//   ent[(t*kTailPairs + k)*64 + lane] its first 2*kTailPairs entries inside
//                                    the run, in pairs (padding: zero slot, 0)
//   ovf[t]                           first overflow slot row of batch t
//   ovf_ent[(ovf[t] + s)*64 + lane]  entry 2*kTailPairs+s of a lane (rows
//                                    longer than 64 * 2 * kTailPairs)
// A lane's entries inside the run are the suffix of its sweep_batch chain (in
// chain order); `idx` is the dependency's position minus the run's first.
inline constexpr int kTailPairs = 4;
inline constexpr int kTailBlock = 1024;                 // threads of the tail workgroup
inline constexpr int kTailWaves = kTailBlock / kWave;   // waves of the tail workgroup
inline constexpr int kTailMaxBatches = 4 * kTailWaves;  // max batches of a tail level
inline constexpr int kBlockTailMaxWaves = 4 * kTailWaves;  // max table waves of a blocked tail level
struct alignas(16) tail_lane {
    std::int32_t row;    // output row the lane writes, or -1
    std::int32_t slot;   // the row's LDS slot (position minus the run's first)
    std::int32_t gc;     // G | (entries inside the run) << 8
    float diag;          // the row's diagonal operand: diag (tail_desc::lane) or 1 / diag
                         // (tail_desc::lane_recip); 1 for lanes that do not write
};
struct alignas(16) tail_pair {
    std::int32_t i0;
    float v0;
    std::int32_t i1;
    float v1;
};
// The walk of a run: for every tail batch (index inside its run) the same
// wave's next batch in walk order (or -1) and the batch's barrier phase inside
// the run, its wave, and whether it is that wave's first batch. Phases follow
// the levels, except that a level of one batch that follows another level of
// one batch of the same wave shares its phase: one wave's LDS writes and reads
// complete in issue order, so no workgroup barrier separates them.
inline constexpr int kTailWalkWaveShift = 20;
inline constexpr int kTailWalkFirst = 1 << 24;
inline constexpr int kTailWalkMaxLevels = 1 << kTailWalkWaveShift;
struct alignas(8) tail_walk {
    std::int32_t next;
    std::int32_t info;   // phase | wave << kTailWalkWaveShift | kTailWalkFirst
};
struct tail_desc {
    const tail_lane* lane = nullptr;
    const tail_lane* lane_recip = nullptr;   // the same lanes with 1 / diag (fp32-vector sweeps)
    const tail_pair* ent = nullptr;
    const std::int32_t* ovf = nullptr;
    const entry32* ovf_ent = nullptr;
    const tail_walk* walk = nullptr;
};
// Static lane layout of a launched level (sweep_lanes): everything a lane of
// a batch needs sits at an address computed from (lane batch, lane), so one
// round of coalesced loads brings a wave's whole batch and the gather follows
// directly (two dependent loads per level instead of the batch table -> row
// pointers -> entries -> values chain).
//   rec[b*64 + lane]              the lane's row and chain shape
//   ent[(b*kEpl + k)*64 + lane]   entry sub + k*G of the row (k < nvalid;
//                                 padding (0, 0), never gathered)
// fp16 storage keeps its widened (exact fp32) values here.
struct alignas(8) lane_rec {
    std::int32_t row;    // the lane's row, or -1 for an idle lane
    std::uint32_t info;  // log2 G | sub << 3 | nvalid << 9
};

// Per (tail batch, lane), written by the coupling kernel before a run: the
// lane's fma chain over the entries before the run, and the row's right-hand
// side (writer lanes).
template <class T>
struct alignas(2 * sizeof(T)) tail_val {
    T part0;
    T r;
};

}  // namespace apxchol_amd
