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
    const float* inv_scale = nullptr;       // m, fp16 back sweep input scale (else null)
    const std::uint32_t* batch_pos = nullptr;  // nb+1
    const std::uint64_t* batch_mask = nullptr; // nb
    int nb = 0;
    int chunk = 1;                          // batches claimed per ticket (dataflow schedule)
    int* ctrl = nullptr;                    // 2 ints: ticket, finished waves (zero between sweeps)
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
    int status;          // 0 running, 1 converged, 2 breakdown (pAp <= 0), 3 stagnation, 4 max_iter
    int n;
    int pad_i[2];
};

}  // namespace apxchol_amd
