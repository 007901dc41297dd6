// solve_server: the serving loop of the MI355X solve path (no counterpart in
// the reference, whose solvers answer one call at a time).
//
// Factors stay resident in HBM (admitted against a memory budget); every
// factor runs one blocked PCG session of `slots` columns on its own stream
// and worker thread, so the work of different factors overlaps on the device.
// Requests are admitted into free columns of the running block between
// iteration groups (continuous batching): a request that arrives while others
// are iterating joins them at the next group boundary, and a finished column
// is released and refilled without stopping the others. A column's
// arithmetic does not depend on its neighbours or its slot, so a request that
// runs in the block gets the same bits alone or in a full block.
//
// Narrow block (4 columns by default): a second session of the factor, on the
// main block's stream, that takes the waiting requests while the main block is idle
// and they fit it, so a few requests in flight iterate at a narrow block's
// cost rather than the main block's (under load the running main block takes
// them: its cost does not grow with its columns). A request's bits are those
// of the block that took it.
//
// Single lane (on by default): beside its blocks, each factor runs one request
// at a time as a single-vector solve on a second stream. The oldest waiting
// request takes it whenever it is free, so at light load a request costs a
// single-vector iteration instead of a whole block's; its result is the bits
// of the solver's own solve with that tol and max_iter. The block and the
// single lane iterate side by side, in the same iteration groups. With the
// single lane on, a request's bits are therefore those of its own solve or of
// its block column, whichever took it; both meet its tolerance.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gpu_solver.hpp"

namespace apxchol_amd {

struct server_options {
    std::uint64_t memory_budget = 0;   // bytes of HBM for factors + blocks (0: 90% of the free memory)
    int slots = 32;                    // columns per factor block (power of two <= 64 after rounding;
                                       // 0: each factor's autotuned block width)
    int iters_per_check = 2;           // PCG iterations between admissions / completions
    bool single_lane = true;           // one request per factor at a time as a single-vector solve
    int small_slots = 4;               // columns of a second, narrower block filled before the main
                                       // one (0: none; used only when narrower than `slots`)
};

struct server_stats {
    std::int64_t submitted = 0, completed = 0, factors = 0;
    double p50_ms = 0.0, p99_ms = 0.0, mean_ms = 0.0;
    double solves_per_s = 0.0;          // completed / (last completion - first submission)
    std::uint64_t memory_used = 0, memory_budget = 0;
    std::int64_t single_lane_completed = 0;
};

class solve_server {
public:
    explicit solve_server(const server_options& opts = {});
    ~solve_server();
    solve_server(const solve_server&) = delete;
    solve_server& operator=(const solve_server&) = delete;

    // Registers a resident factor; throws std::runtime_error when its memory
    // (solver + block buffers) does not fit the budget. The server uses the
    // solver exclusively from now on.
    int add(std::shared_ptr<gpu_solver> solver);
    // Queues one solve of b (n host doubles, copied). Returns the request id.
    std::int64_t submit(int factor, const double* b, double tol, int max_iter);
    // Waits for a request (timeout < 0: no limit); copies x (n host doubles)
    // and its result, and forgets the request. False on timeout.
    bool wait(std::int64_t id, double* x, solve_result* res, double timeout_s);
    server_stats stats() const;
    // Turns the single lanes on or off (requests already on one finish there).
    void set_single_lane(bool on);
    // Columns of the narrow block of factors added from now on (0: none).
    void set_small_slots(int slots);

private:
    struct request {
        std::int64_t id = 0;
        int factor = 0;
        std::vector<double> b, x;
        double tol = 1e-8;
        int max_iter = 200;
        solve_result res;
        bool done = false;
        std::chrono::steady_clock::time_point t_submit, t_done;
    };
    // One session of a factor: a block of kp columns.
    struct block_lane {
        int which = 0;                                // the solver's session index
        int kp = 0;
        hipStream_t stream = nullptr;                 // the lane's block stream
        std::vector<std::shared_ptr<request>> slot;   // request in each column
        void* d_b = nullptr;                          // kp x n staging (device)
        void* d_x = nullptr;
        std::vector<gpu_solver::session_slot> states;
    };
    struct lane {
        std::shared_ptr<gpu_solver> solver;
        std::thread worker;
        std::deque<std::shared_ptr<request>> queue;
        std::vector<block_lane> blocks;               // filled in order: the narrow block, then the main one
        hipStream_t block_stream = nullptr;           // the blocks' stream
        std::shared_ptr<request> single;              // request on the single lane
        hipStream_t single_stream = nullptr;
        void* d_sb = nullptr;                         // single lane's b and x (device, n doubles each)
        void* d_sx = nullptr;
        void* h_stage = nullptr;                      // n doubles, pinned host memory
        std::int64_t n = 0;
        std::string error;                            // set when a device failure ended the lane
    };
    void run(lane& ln);
    void finish(lane& ln, const std::vector<std::shared_ptr<request>>& done);

    server_options opts_;
    mutable std::mutex mu_;
    std::condition_variable work_cv_, done_cv_;
    bool stop_ = false;
    int device_ = 0;
    std::vector<std::unique_ptr<lane>> lanes_;
    std::map<std::int64_t, std::shared_ptr<request>> requests_;
    std::int64_t next_id_ = 1;
    std::uint64_t used_ = 0;
    std::int64_t submitted_ = 0;
    std::int64_t single_completed_ = 0;
    std::vector<double> latencies_ms_;
    std::chrono::steady_clock::time_point first_submit_, last_done_;
    bool any_submit_ = false;
};

}  // namespace apxchol_amd
