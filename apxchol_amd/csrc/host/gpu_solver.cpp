// gpu_solver: host orchestration of the GPU solve path.
//
// Setup: host factorization (factorize.cpp) -> L11 prep (trsv_prep.cpp) ->
// upload of the two sweep layouts, their batch tables, the permuted operator
// and the PCG workspace. Solve: the PCG recurrence runs in the factor's
// permuted space on the device (pcg.hip), with the two sweeps (sweep.hip) as
// the preconditioner; the host only enqueues work and reads the completion
// flag every iters_per_check iterations.
#include "gpu_solver.hpp"

#include <chrono>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../device.hpp"
#include "../kernels/launch.hpp"

namespace apxchol_amd {

namespace {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("apxchol_amd: ") + what + ": " + hipGetErrorString(e));
}

class dbuf {
public:
    dbuf() = default;
    explicit dbuf(std::size_t bytes) { alloc(bytes); }
    ~dbuf() { if (p_) (void)hipFree(p_); }
    dbuf(const dbuf&) = delete;
    dbuf& operator=(const dbuf&) = delete;
    void alloc(std::size_t bytes) {
        if (p_) { (void)hipFree(p_); p_ = nullptr; }
        bytes_ = bytes;
        if (bytes == 0) return;
        hip_check(hipMalloc(&p_, bytes), "hipMalloc");
    }
    template <class T>
    void upload(const std::vector<T>& v) {
        alloc(v.size() * sizeof(T));
        if (!v.empty()) hip_check(hipMemcpy(p_, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    }
    void zero() { if (p_) hip_check(hipMemset(p_, 0, bytes_), "hipMemset"); }
    template <class T> T* as() const { return static_cast<T*>(p_); }
    void* get() const { return p_; }
    std::size_t bytes() const { return bytes_; }
private:
    void* p_ = nullptr;
    std::size_t bytes_ = 0;
};

// Lanes of a row of `len` off-diagonals (the kernels' lanes_for).
int lanes_for_host(std::uint32_t len) {
    std::uint32_t g = (len + kEpl - 1) / kEpl;
    if (g <= 1) return 1;
    std::uint32_t p = 1;
    while (p < g) p <<= 1;
    return p > 64 ? 64 : static_cast<int>(p);
}

// Workgroup cap of the vector kernels (their partial buffers are sized by it).
constexpr int kMaxBlocksHost = 2048;
// Workgroup cap of the sliced-ELL SpMV (the partials buffer holds this many),
// and the XCD count its grid is a multiple of.
constexpr int kSellMaxBlocks = 8192;
constexpr int kXcdsHost = 8;

int env_int(const char* name, int dflt) {
    const char* e = std::getenv(name);
    if (!e || !*e) return dflt;
    const int v = std::atoi(e);
    return v > 0 ? v : dflt;
}

// An integer knob that may be 0 (unset or empty: dflt).
int env_flag(const char* name, int dflt) {
    const char* e = std::getenv(name);
    return (e && *e) ? std::atoi(e) : dflt;
}

int center_k_from_env() {
    int k = 10;
    if (const char* e = std::getenv("APXCHOL_CENTER_K"); e && *e) {
        const int v = std::atoi(e);
        if (v >= 1) k = v;
    }
    if (const char* e = std::getenv("APXCHOL_GROUND"); e && *e && std::strcmp(e, "center") == 0) k = 1;
    return k;
}

}  // namespace

batch_table build_batches(const trsv_factor& T, bool backward) {
    batch_table bt;
    const std::uint32_t m = T.m;
    const auto& ptr = backward ? T.b_ptr : T.f_ptr;
    const std::size_t L = T.level_ptr.size() - 1;
    std::uint64_t mask = 0;
    int lane = 0;
    bool open = false;
    auto close = [&]() {
        if (!open) return;
        bt.mask.push_back(mask);
        open = false;
        mask = 0;
        lane = 0;
    };
    auto open_at = [&](std::uint32_t t) {
        bt.pos.push_back(t);
        open = true;
        mask = 0;
        lane = 0;
    };
    for (std::size_t li = 0; li < L; ++li) {
        // Level li in sweep order; positions [p0, p1).
        std::uint32_t p0, p1;
        if (!backward) { p0 = T.level_ptr[li]; p1 = T.level_ptr[li + 1]; }
        else { p0 = m - T.level_ptr[L - li]; p1 = m - T.level_ptr[L - li - 1]; }
        close();
        bt.level_first.push_back(static_cast<int>(bt.mask.size()));
        for (std::uint32_t t = p0; t < p1; ++t) {
            const std::uint32_t row = backward ? m - 1 - t : t;
            const std::uint32_t len = ptr[row + 1] - ptr[row];
            if (len > static_cast<std::uint32_t>(kWave * kEpl)) {
                close();
                open_at(t);
                mask = 1;
                close();
                continue;
            }
            const int G = lanes_for_host(len);
            int a = (lane + G - 1) / G * G;
            if (open && a + G > kWave) { close(); a = 0; }
            if (!open) { open_at(t); a = 0; }
            mask |= 1ull << a;
            lane = a + G;
        }
    }
    close();
    bt.level_first.push_back(static_cast<int>(bt.mask.size()));
    bt.pos.push_back(m);
    return bt;
}

// Every invariant the sweep kernels rely on, checked on the host before any
// upload: entry ids point strictly backwards in sweep order and inside [0, m);
// batches tile [0, m) in order; a batch's mask has one bit per row; every
// row's aligned lane group fits the wave without overlap; a row longer than
// 64 * kEpl is alone in its batch; levels start on batch boundaries.
void validate_sweep_layout(const trsv_factor& T, const batch_table& bt, bool backward) {
    const std::uint32_t m = T.m;
    const auto& ptr = backward ? T.b_ptr : T.f_ptr;
    const auto& idx = backward ? T.b_idx : T.f_idx;
    auto fail = [](const std::string& w) { throw std::runtime_error("apxchol_amd: invalid sweep layout: " + w); };
    if (ptr.size() != static_cast<std::size_t>(m) + 1 || ptr[0] != 0 || ptr[m] != idx.size()) fail("ptr");
    for (std::uint32_t r = 0; r < m; ++r) {
        if (ptr[r + 1] < ptr[r]) fail("ptr order");
        for (std::uint32_t p = ptr[r]; p < ptr[r + 1]; ++p) {
            const std::int64_t j = idx[p];
            if (j < 0 || j >= static_cast<std::int64_t>(m)) fail("index range");
            if (backward ? j <= static_cast<std::int64_t>(r) : j >= static_cast<std::int64_t>(r)) fail("dependency direction");
        }
    }
    const std::size_t nb = bt.mask.size();
    if (bt.pos.size() != nb + 1 || bt.pos[0] != 0 || bt.pos[nb] != m) fail("batch bounds");
    for (std::size_t b = 0; b < nb; ++b) {
        const std::uint32_t p0 = bt.pos[b], p1 = bt.pos[b + 1];
        if (p1 <= p0) fail("empty batch");
        if (static_cast<std::uint32_t>(__builtin_popcountll(bt.mask[b])) != p1 - p0 || !(bt.mask[b] & 1ull))
            fail("batch mask");
        std::uint64_t used = 0;
        int k = 0;
        for (int lane = 0; lane < kWave; ++lane) {
            if (!((bt.mask[b] >> lane) & 1ull)) continue;
            const std::uint32_t t = p0 + static_cast<std::uint32_t>(k++);
            const std::uint32_t row = backward ? m - 1 - t : t;
            const std::uint32_t len = ptr[row + 1] - ptr[row];
            const int G = lanes_for_host(len);
            if (len > static_cast<std::uint32_t>(kWave * kEpl) && (p1 - p0 != 1 || G != kWave)) fail("long row not alone");
            if (lane % G != 0 || lane + G > kWave) fail("lane group alignment");
            const std::uint64_t g = (G == 64 ? ~0ull : ((1ull << G) - 1ull)) << lane;
            if (used & g) fail("lane group overlap");
            used |= g;
            // Lanes up to the next row start map to this row (padding lanes idle).
            const std::uint64_t after = bt.mask[b] >> lane >> 1;
            const int next = after ? lane + 1 + __builtin_ctzll(after) : kWave;
            if (next < lane + G) fail("next group inside this group");
        }
    }
    const std::size_t L = T.level_ptr.size() - 1;
    if (bt.level_first.size() != L + 1 || bt.level_first[0] != 0 || bt.level_first[L] != static_cast<int>(nb))
        fail("level batch bounds");
    for (std::size_t l = 0; l < L; ++l) {
        if (bt.level_first[l + 1] < bt.level_first[l]) fail("level order");
        const std::uint32_t lp0 = backward ? m - T.level_ptr[L - l] : T.level_ptr[l];
        if (bt.level_first[l] < static_cast<int>(nb) && bt.pos[bt.level_first[l]] != lp0) fail("level start");
    }
    // Every dependency lies in an earlier level (what the level-set and tail
    // schedules rely on).
    std::vector<std::uint32_t> level_of(m);
    for (std::size_t l = 0; l < L; ++l)
        for (int b = bt.level_first[l]; b < bt.level_first[l + 1]; ++b)
            for (std::uint32_t t = bt.pos[b]; t < bt.pos[b + 1]; ++t) level_of[t] = static_cast<std::uint32_t>(l);
    for (std::uint32_t t = 0; t < m; ++t) {
        const std::uint32_t row = backward ? m - 1 - t : t;
        for (std::uint32_t p = ptr[row]; p < ptr[row + 1]; ++p) {
            const std::uint32_t pj = backward ? m - 1 - static_cast<std::uint32_t>(idx[p]) : static_cast<std::uint32_t>(idx[p]);
            if (level_of[pj] >= level_of[t]) fail("dependency inside its own level");
        }
    }
}

block_batch_table build_block_batches(const trsv_factor& T, bool backward, int kp) {
    block_batch_table bt;
    bt.kp = kp;
    const std::uint32_t m = T.m;
    const auto& ptr = backward ? T.b_ptr : T.f_ptr;
    const std::size_t L = T.level_ptr.size() - 1;
    std::uint64_t mask = 0;
    int lane = 0;
    bool open = false;
    auto close = [&]() {
        if (!open) return;
        bt.mask.push_back(mask);
        bt.col0.push_back(0);
        open = false;
    };
    for (std::size_t li = 0; li < L; ++li) {
        std::uint32_t p0, p1;
        if (!backward) { p0 = T.level_ptr[li]; p1 = T.level_ptr[li + 1]; }
        else { p0 = m - T.level_ptr[L - li]; p1 = m - T.level_ptr[L - li - 1]; }
        close();
        bt.level_first.push_back(static_cast<int>(bt.mask.size()));
        for (std::uint32_t t = p0; t < p1; ++t) {
            const std::uint32_t row = backward ? m - 1 - t : t;
            const int G = lanes_for_host(ptr[row + 1] - ptr[row]);
            const int U = G * kp;
            if (U > kWave) {
                close();
                for (int c0 = 0; c0 < kp; c0 += kWave / G) {
                    bt.pos.push_back(t);
                    bt.mask.push_back(0);
                    bt.col0.push_back(c0);
                }
                continue;
            }
            int a = (lane + U - 1) / U * U;
            if (open && a + U > kWave) { close(); a = 0; }
            if (!open) { bt.pos.push_back(t); open = true; mask = 0; a = 0; }
            mask |= 1ull << a;
            lane = a + U;
        }
    }
    close();
    bt.level_first.push_back(static_cast<int>(bt.mask.size()));
    bt.pos.push_back(m);
    return bt;
}

// The blocked table's invariants, checked before upload: every position of
// every level is covered once per column, groups are aligned and disjoint,
// split waves cover the columns exactly, levels start on wave boundaries.
void validate_block_batches(const trsv_factor& T, const block_batch_table& bt, bool backward) {
    auto fail = [](const std::string& w) { throw std::runtime_error("apxchol_amd: invalid blocked sweep table: " + w); };
    const std::uint32_t m = T.m;
    const auto& ptr = backward ? T.b_ptr : T.f_ptr;
    const int kp = bt.kp;
    if (kp < 1 || kp > kWave || (kp & (kp - 1))) fail("block width");
    const std::size_t nw = bt.mask.size();
    if (bt.pos.size() != nw + 1 || bt.col0.size() != nw || bt.pos[nw] != m) fail("sizes");
    std::vector<std::uint32_t> seen(m, 0);   // columns covered per position
    for (std::size_t w = 0; w < nw; ++w) {
        const std::uint32_t p = bt.pos[w];
        if (p >= m) fail("position range");
        if (bt.mask[w] == 0) {
            const int G = lanes_for_host(ptr[(backward ? m - 1 - p : p) + 1] - ptr[backward ? m - 1 - p : p]);
            if (G * kp <= kWave || bt.col0[w] < 0 || bt.col0[w] % (kWave / G) || bt.col0[w] >= kp) fail("split wave");
            seen[p] += static_cast<std::uint32_t>(kWave / G);
            continue;
        }
        std::uint64_t used = 0;
        int k = 0;
        for (int lane = 0; lane < kWave; ++lane) {
            if (!((bt.mask[w] >> lane) & 1ull)) continue;
            const std::uint32_t t = p + static_cast<std::uint32_t>(k++);
            if (t >= m) fail("row past m");
            const std::uint32_t row = backward ? m - 1 - t : t;
            const int U = lanes_for_host(ptr[row + 1] - ptr[row]) * kp;
            if (U > kWave || lane % U != 0 || lane + U > kWave) fail("group alignment");
            const std::uint64_t g = (U == 64 ? ~0ull : ((1ull << U) - 1ull)) << lane;
            if (used & g) fail("group overlap");
            used |= g;
            const std::uint64_t after = bt.mask[w] >> lane >> 1;
            const int next = after ? lane + 1 + __builtin_ctzll(after) : kWave;
            if (next < lane + U) fail("next group inside this group");
            seen[t] += static_cast<std::uint32_t>(kp);
        }
    }
    for (std::uint32_t t = 0; t < m; ++t)
        if (seen[t] != static_cast<std::uint32_t>(kp)) fail("coverage");
    const std::size_t L = T.level_ptr.size() - 1;
    if (bt.level_first.size() != L + 1 || bt.level_first[0] != 0 || bt.level_first[L] != static_cast<int>(nw))
        fail("level bounds");
    for (std::size_t l = 0; l < L; ++l) {
        if (bt.level_first[l + 1] < bt.level_first[l]) fail("level order");
        const std::uint32_t lp0 = backward ? m - T.level_ptr[L - l] : T.level_ptr[l];
        const std::uint32_t lp1 = backward ? m - T.level_ptr[L - l - 1] : T.level_ptr[l + 1];
        for (int w = bt.level_first[l]; w < bt.level_first[l + 1]; ++w)
            if (bt.pos[w] < lp0 || bt.pos[w] >= lp1) fail("wave outside its level");
    }
}

namespace {

// What one lane of batch b does (the kernels' lane mapping).
struct host_lane {
    std::uint32_t pos = 0, row = 0, rb = 0, len = 0;
    int G = 1, sub = 0;
    bool has_row = false, active = false;
};
host_lane map_lane_host(const trsv_factor& T, const batch_table& bt, bool backward, int b, int lane) {
    const auto& ptr = backward ? T.b_ptr : T.f_ptr;
    host_lane x;
    const std::uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const std::uint64_t mu = bt.mask[b] & upto;
    const int rib = __builtin_popcountll(mu) - 1;
    x.sub = lane - (63 - __builtin_clzll(mu));
    x.pos = bt.pos[b] + static_cast<std::uint32_t>(rib);
    x.has_row = x.pos < bt.pos[b + 1];
    if (x.has_row) {
        x.row = backward ? T.m - 1 - x.pos : x.pos;
        x.rb = ptr[x.row];
        x.len = ptr[x.row + 1] - x.rb;
    }
    x.G = lanes_for_host(x.len);
    x.active = x.has_row && x.sub < x.G;
    return x;
}

}  // namespace

int append_tail_run(const trsv_factor& T, const batch_table& bt, bool backward, int bf, int bl, tail_tables& tt) {
    const auto& idx = backward ? T.b_idx : T.f_idx;
    const auto& val = backward ? T.b_val : T.f_val;
    const auto& hv = backward ? T.b_h : T.f_h;
    auto fail = [](const std::string& w) { throw std::runtime_error("apxchol_amd: invalid tail run: " + w); };
    if (bf < 0 || bl <= bf || static_cast<std::size_t>(bl) > bt.mask.size()) fail("batch range");
    const std::uint32_t p0 = bt.pos[bf];
    // Padding entries read the zero slot past the run's outputs.
    const std::int32_t zslot = static_cast<std::int32_t>(bt.pos[bl] - p0);
    const int base = tt.batches;
    std::vector<entry32> chain[kWave];
    for (int b = bf; b < bl; ++b) {
        int maxc = 0;
        for (int lane = 0; lane < kWave; ++lane) {
            const host_lane x = map_lane_host(T, bt, backward, b, lane);
            const bool longrow = x.len > static_cast<std::uint32_t>(kWave * kEpl);
            auto& ch = chain[lane];
            ch.clear();
            bool inside = false;
            auto take = [&](std::uint32_t q) {
                const std::uint32_t p = x.rb + q;
                const std::uint32_t j = static_cast<std::uint32_t>(idx[p]);
                const std::uint32_t pj = backward ? T.m - 1 - j : j;
                if (pj < p0) {
                    if (inside) fail("a lane's entries before the run follow entries inside it");
                    return;
                }
                if (pj >= x.pos) fail("dependency not before its row");
                inside = true;
                ch.push_back({static_cast<std::int32_t>(pj - p0), T.fp16 ? fp16_widen(hv[p]) : val[p]});
            };
            if (longrow) {
                for (std::uint32_t q = static_cast<std::uint32_t>(lane); q < x.len; q += kWave) take(q);
            } else if (x.active) {
                for (int k = 0; k < kEpl; ++k) {
                    const std::uint32_t q = static_cast<std::uint32_t>(x.sub + k * x.G);
                    if (q < x.len) take(q);
                }
            }
            const bool writer = x.active && x.sub == 0;
            const int c = static_cast<int>(ch.size());
            if (c >= (1 << 22)) fail("row too long");
            const tail_lane tl{writer ? static_cast<std::int32_t>(x.row) : -1,
                               writer ? static_cast<std::int32_t>(x.pos - p0) : 0, x.G | (c << 8),
                               writer ? T.diag[x.row] : 1.0f};
            tt.lane.push_back(tl);
            tt.lane_recip.push_back({tl.row, tl.slot, tl.gc, writer ? 1.0f / T.diag[x.row] : 1.0f});
            maxc = std::max(maxc, c);
        }
        for (int k = 0; k < kTailPairs; ++k)
            for (int lane = 0; lane < kWave; ++lane) {
                const auto& ch = chain[lane];
                const std::size_t a = 2 * static_cast<std::size_t>(k), c = a + 1;
                tail_pair e{zslot, 0.0f, zslot, 0.0f};
                if (a < ch.size()) { e.i0 = ch[a].idx; e.v0 = ch[a].val; }
                if (c < ch.size()) { e.i1 = ch[c].idx; e.v1 = ch[c].val; }
                tt.ent.push_back(e);
            }
        constexpr int slots = 2 * kTailPairs;
        if (maxc > slots) {
            const std::size_t first = tt.ovf_ent.size() / kWave;
            if (first > static_cast<std::size_t>(INT32_MAX)) fail("overflow table too large");
            tt.ovf.push_back(static_cast<std::int32_t>(first));
            for (int s = 0; s < maxc - slots; ++s)
                for (int lane = 0; lane < kWave; ++lane) {
                    const auto& ch = chain[lane];
                    const std::size_t q = static_cast<std::size_t>(slots + s);
                    tt.ovf_ent.push_back(q < ch.size() ? ch[q] : entry32{zslot, 0.0f});
                }
        } else {
            tt.ovf.push_back(0);
        }
    }
    tt.batches += bl - bf;
    return base;
}

int append_tail_walk(const std::vector<int>& level_first, int l, int e, tail_tables& tt, bool merge, bool rotate) {
    const int b0 = level_first[l], nb = level_first[e] - b0;
    if (nb <= 0 || e - l > kTailWalkMaxLevels) throw std::runtime_error("apxchol_amd: invalid tail walk");
    const std::size_t at = tt.walk.size();
    tt.walk.resize(at + static_cast<std::size_t>(nb));
    int prev[kTailWaves];
    std::fill(prev, prev + kTailWaves, -1);
    // The wave of batch b of level q: with rotate, batches are dealt round
    // robin over the whole run (wave w takes run batches w, w + 16, ...), so a
    // wave's next batch lies levels ahead and its prefetched layout has that
    // long to arrive; without, every level starts again at wave 0.
    const auto wave_of = [&](int q, int b) { return (rotate ? b - b0 : b - level_first[q]) % kTailWaves; };
    const auto single = [&](int q) { return level_first[q + 1] - level_first[q] == 1; };
    int phase = -1;
    for (int q = l; q < e; ++q) {
        // Every level of a phase holds one batch, all of the same wave.
        if (q == l || !merge || !(single(q) && single(q - 1) && wave_of(q, level_first[q]) == wave_of(q - 1, level_first[q - 1])))
            ++phase;
        for (int b = level_first[q]; b < level_first[q + 1]; ++b) {
            const int i = b - b0, w = wave_of(q, b);
            tail_walk& x = tt.walk[at + static_cast<std::size_t>(i)];
            x.next = -1;
            x.info = phase | (w << kTailWalkWaveShift) | (prev[w] < 0 ? kTailWalkFirst : 0);
            if (prev[w] >= 0) tt.walk[at + static_cast<std::size_t>(prev[w])].next = i;
            prev[w] = i;
        }
    }
    return phase + 1;
}

long long append_lane_level(const trsv_factor& T, const batch_table& bt, bool backward, int bf, int bl,
                            lane_tables& lt) {
    const auto& idx = backward ? T.b_idx : T.f_idx;
    const auto& val = backward ? T.b_val : T.f_val;
    const auto& hv = backward ? T.b_h : T.f_h;
    auto fail = [](const std::string& w) { throw std::runtime_error("apxchol_amd: invalid lane layout: " + w); };
    if (bf < 0 || bl <= bf || static_cast<std::size_t>(bl) > bt.mask.size()) fail("batch range");
    const long long base = lt.batches;
    for (int b = bf; b < bl; ++b) {
        entry32 e[kEpl][kWave];
        for (int lane = 0; lane < kWave; ++lane) {
            const host_lane x = map_lane_host(T, bt, backward, b, lane);
            if (x.len > static_cast<std::uint32_t>(kWave * kEpl)) fail("long row");
            int nv = 0;
            for (int k = 0; k < kEpl; ++k) {
                e[k][lane] = {0, 0.0f};
                const std::uint32_t q = static_cast<std::uint32_t>(x.sub + k * x.G);
                if (!x.active || q >= x.len) continue;
                const std::uint32_t p = x.rb + q;
                const std::int32_t j = idx[p];
                const std::uint32_t pj = backward ? T.m - 1 - static_cast<std::uint32_t>(j) : static_cast<std::uint32_t>(j);
                if (j < 0 || static_cast<std::uint32_t>(j) >= T.m || pj >= bt.pos[bf]) fail("dependency not before the level");
                if (nv != k) fail("valid slots not a prefix");
                e[k][lane] = {j, T.fp16 ? fp16_widen(hv[p]) : val[p]};
                ++nv;
            }
            int lg = 0;
            while ((1 << lg) < x.G) ++lg;
            if ((1 << lg) != x.G || x.sub < 0 || x.sub >= kWave) fail("lane group");
            const std::uint32_t info = static_cast<std::uint32_t>(lg) | (static_cast<std::uint32_t>(x.active ? x.sub : 0) << 3) |
                                       (static_cast<std::uint32_t>(nv) << 9);
            lt.rec.push_back({x.active ? static_cast<std::int32_t>(x.row) : -1, x.active ? info : 0u});
        }
        for (int k = 0; k < kEpl; ++k)
            for (int lane = 0; lane < kWave; ++lane) lt.ent.push_back(e[k][lane]);
    }
    lt.batches += bl - bf;
    return base;
}

struct gpu_solver::device_state {
    vec_prec prec = vec_prec::fp64;
    bool fp16 = false;
    int grid = 0;
    dbuf f_ptr, f_ent, f_idx, f_h, b_ptr, b_ent, b_idx, b_h, diag, inv_diag, inv_scale;
    dbuf f_pos, f_mask, b_pos, b_mask;
    std::vector<int> f_levels, b_levels;
    dbuf f_levels_dev, b_levels_dev, tail_val;
    dbuf f_tlane, f_tent, f_tovf, f_tovf_ent, f_twalk, b_tlane, b_tent, b_tovf, b_tovf_ent, b_twalk;
    dbuf f_tlane_r, b_tlane_r;   // tail lanes with 1 / diag (fp32-vector sweeps)
    dbuf f_lrec, f_lent, b_lrec, b_lent;
    std::vector<int> f_level_lb, b_level_lb;   // first lane batch of each level, or -1
    std::vector<int> f_level_pos, b_level_pos;   // first sweep position of each level
    std::vector<unsigned char> f_level_diag, b_level_diag;   // levels without entries
    std::vector<int> f_seg[8], b_seg[8];   // hybrid segments: l0, l1, tail flag, p0, p1, first tail batch, barrier
                                           // phases, coupling split position (-1: none)
    // Side stream and fork / join events of the split couplings (both sweep
    // directions; their sweeps never overlap).
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    sweep_sched f_sched, b_sched;
    dbuf perm, op_ptr, op_idx, op_val;
    bool op32 = false;
    dbuf sell_soff, sell_idx, sell_val, sell_map, sell_long;
    sell_desc sell;
    bool has_sell = false;   // the sliced-ELL form exists
    bool use_sell = false;   // the single SpMV runs on it (else CSR lanes)
    int lanes = 8;
    dbuf r, x, p, ap, rp, y, z, xout;
    dbuf relabel;   // factor position -> storage position (empty: identity)
    dbuf part, st, st_apply, ctrl, err;
    pcg_state* host_st = nullptr;
    int* host_err = nullptr;
    sweep_desc fwd, bck;
    // One PCG iteration captured as a graph on a private stream (default
    // priority) and replayed on the caller's stream.
    hipStream_t cap = nullptr;
    hipGraph_t iter_graph = nullptr;
    hipGraphExec_t iter_exec = nullptr;
    // The single-vector solve in progress (single_begin .. single_collect).
    bool single_graph = false;
    int single_enqueued = 0, single_last = 1, single_max_iter = 0;
    // Blocked solves: the block buffers (allocated on first use for the
    // widest block seen) and the sweep tables per block width.
    struct block_tables {
        dbuf f_pos, f_mask, f_col0, b_pos, b_mask, b_col0, f_lf_dev, b_lf_dev;
        std::vector<int> f_lf, b_lf, f_lpos, b_lpos, f_lgmax, b_lgmax;
        std::vector<int> f_seg[3], b_seg[3];   // l0, l1, tail flag
        dbuf col_tv;                            // coupling values of the column-by-column tail runs
        block_sched f, b;
    };
    struct block_bufs {
        int kp = 0;
        std::map<int, std::unique_ptr<block_tables>> tables;   // per block width
        dbuf r, x, p, ap, y, z, st, bs, part;
        pcg_state* host_st = nullptr;
        batch_state* host_bs = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int graph_kp = 0;
        int session_kp = 0;             // block width of a running serving session
        unsigned char* host_adm = nullptr;   // pinned: 64 slot ids + 64 pointers
        dbuf adm_dev;
        void drop_graph() {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            exec = nullptr;
            graph = nullptr;
            graph_kp = 0;
        }
        ~block_bufs() {
            drop_graph();
            if (host_st) (void)hipHostFree(host_st);
            if (host_bs) (void)hipHostFree(host_bs);
            if (host_adm) (void)hipHostFree(host_adm);
        }
    } blk, blk2;   // blk: batches and a session's main block; blk2: a session's second block
    ~device_state() {
        if (iter_exec) (void)hipGraphExecDestroy(iter_exec);
        if (iter_graph) (void)hipGraphDestroy(iter_graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (side) (void)hipStreamDestroy(side);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (host_st) (void)hipHostFree(host_st);
        if (host_err) (void)hipHostFree(host_err);
    }
};

namespace {
// Block buffers `which` (0 or 1) of a solver's device state.
template <class D>
auto& block_in(D& d, int which) {
    if (which != 0 && which != 1) throw std::invalid_argument("gpu_solver: no such block");
    return which == 0 ? d.blk : d.blk2;
}
}  // namespace

// HIP-event timing of kernel classes. Events are pooled and reused; each
// record carries the PCG step it belongs to, so launches after convergence
// (which return at once) can be dropped when the solve ends.
struct gpu_solver::timer {
    static constexpr int K = gpu_solver::kTimingClasses;
    bool on = false;
    struct rec { int cls; int step; hipEvent_t a, b; };
    std::vector<hipEvent_t> pool;
    std::size_t used = 0;
    std::vector<rec> recs;
    double ms[K] = {};
    std::int64_t launches[K] = {};
    // Single-kernel records (launch_hook): the launch carries the event pair,
    // which then holds the kernel's own start and end.
    int step = 0;
    static bool hook_fn(void* ctx, int kind, hipEvent_t* ev) {
        auto* t = static_cast<timer*>(ctx);
        if (!t->on) return false;
        ev[0] = t->get();
        ev[1] = t->get();
        const int cls = kind == 1 ? 5 : kind == 2 ? 6 : kind == 4 ? 2 : 4;
        t->recs.push_back({cls, t->step, ev[0], ev[1]});
        return true;
    }
    launch_hook hook(int st) {
        step = st;
        return launch_hook{on ? &hook_fn : nullptr, this};
    }
    ~timer() { for (auto e : pool) (void)hipEventDestroy(e); }
    hipEvent_t get() {
        if (used == pool.size()) {
            hipEvent_t e;
            hip_check(hipEventCreate(&e), "hipEventCreate");
            pool.push_back(e);
        }
        return pool[used++];
    }
    template <class F>
    void time(int cls, int step, hipStream_t s, F&& launch) {
        if (!on) { launch(); return; }
        rec r{cls, step, get(), get()};
        hip_check(hipEventRecord(r.a, s), "hipEventRecord");
        launch();
        hip_check(hipEventRecord(r.b, s), "hipEventRecord");
        recs.push_back(r);
    }
    // Fold the records of a finished (synchronized) solve: keep steps <= last.
    void collect(int last_step) {
        for (const auto& r : recs) {
            if (r.step > last_step) continue;
            float t = 0.0f;
            hip_check(hipEventElapsedTime(&t, r.a, r.b), "hipEventElapsedTime");
            ms[r.cls] += t;
            launches[r.cls] += 1;
        }
        recs.clear();
        used = 0;
    }
};

void gpu_solver::set_timing(bool on) const { timer_->on = on; }

void gpu_solver::take_timing(double* ms, std::int64_t* launches) const {
    for (int k = 0; k < kTimingClasses; ++k) {
        ms[k] = timer_->ms[k];
        launches[k] = timer_->launches[k];
        timer_->ms[k] = 0.0;
        timer_->launches[k] = 0;
    }
}

namespace {
// Graph captures of all solvers of the process, one at a time: several
// serving lanes capture their first iterations from their own threads, and
// the runtime's capture bookkeeping is process-wide state.
std::mutex g_capture_mu;

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}
}  // namespace

gpu_solver::gpu_solver(const csc_view& L, const solve_options& opts) : opts_(opts) {
    const auto t0 = std::chrono::steady_clock::now();
    F_ = factorize(L, opts_.factor_opts);
    setup_times_.factorize_ms = ms_since(t0);
    setup(L);
}

void gpu_solver::claim_serving() const {
    bool expect = false;
    if (!serving_.compare_exchange_strong(expect, true))
        throw std::invalid_argument("gpu_solver: this solver already serves a server");
}

void gpu_solver::not_serving(const char* what) const {
    if (serving_.load())
        throw std::invalid_argument(std::string("gpu_solver::") + what +
                                    ": the solver is serving a server, which owns its device state until the server is destroyed");
}

gpu_solver::gpu_solver(const csc_view& L, factorization F, const solve_options& opts) : opts_(opts) {
    if (static_cast<std::int64_t>(F.n) != L.n)
        throw std::invalid_argument("gpu_solver: factorization dimension mismatch");
    if (F.nnz() > 0 && F.vals.empty())
        throw std::invalid_argument("gpu_solver: factorization values were released; pass a freshly computed factorization");
    F_ = std::move(F);
    setup(L);
}

gpu_solver::~gpu_solver() = default;

void gpu_solver::setup(const csc_view& L) {
    const auto t_setup = std::chrono::steady_clock::now();
    timer_ = std::make_unique<timer>();
    n_ = L.n;
    if (n_ > (1ll << 31) - 2) throw std::invalid_argument("gpu_solver: n exceeds the 32-bit device index range");
    center_k_ = center_k_from_env();
    T_ = prepare_trsv(F_, opts_.fp16_storage, opts_.drop_rel);
    // Storage order of the device vectors: the factor order, or with the
    // within-level renumbering (default) the factor order with every level's
    // rows sorted by original vertex, so a vertex's neighbours -- gathered by
    // the sweeps and the SpMV -- sit close together.
    std::vector<node_t> sperm = F_.perm;
    relabel_.clear();
    if (opts_.relabel && env_flag("APXCHOL_AMD_RELABEL", 1) != 0) {
        relabel_ = relabel_within_levels(T_, F_.perm, opts_.depth_levels && env_flag("APXCHOL_AMD_DEPTH_LEVELS", 1) != 0);
        for (auto& q : sperm)
            if (q < T_.m) q = relabel_[q];
    }
    const auto t_op = std::chrono::steady_clock::now();
    const permuted_operator op = build_permuted_operator(L, sperm);
    setup_times_.pcg_setup_ms = ms_since(t_op);
    if (op.ptr.back() > static_cast<std::int64_t>(UINT32_MAX))
        throw std::invalid_argument("gpu_solver: operator nnz exceeds the 32-bit device index range");
    for (const std::int32_t c : op.idx)
        if (c < 0 || c >= n_) throw std::runtime_error("apxchol_amd: operator index out of range");

    dev_ = std::make_unique<device_state>();
    auto& d = *dev_;
    d.prec = opts_.precision == sweep_precision::fp32 ? vec_prec::fp32 : vec_prec::fp64;
    d.fp16 = T_.fp16;
    const int m = static_cast<int>(T_.m);
    const int n = static_cast<int>(n_);

    // Factor layout: per direction a CSR of the off-diagonals; fp32 values
    // interleaved with their ids (entry32), fp16 values in their own array.
    d.f_ptr.upload(T_.f_ptr);
    d.b_ptr.upload(T_.b_ptr);
    // Entry arrays hold at least one element: lanes without an entry read
    // slot 0 (and discard it) so their loads need no branch.
    if (T_.fp16) {
        auto pad_i = [](std::vector<std::int32_t> v) { if (v.empty()) v.push_back(0); return v; };
        auto pad_h = [](std::vector<std::uint16_t> v) { if (v.empty()) v.push_back(0); return v; };
        d.f_idx.upload(pad_i(T_.f_idx));
        d.f_h.upload(pad_h(T_.f_h));
        d.b_idx.upload(pad_i(T_.b_idx));
        d.b_h.upload(pad_h(T_.b_h));
        d.inv_scale.upload(T_.inv_scale);
    } else {
        std::vector<entry32> fe(std::max<std::size_t>(T_.f_idx.size(), 1), entry32{0, 0.0f}),
            be(std::max<std::size_t>(T_.b_idx.size(), 1), entry32{0, 0.0f});
        for (std::size_t i = 0; i < T_.f_idx.size(); ++i) fe[i] = {T_.f_idx[i], T_.f_val[i]};
        for (std::size_t i = 0; i < T_.b_idx.size(); ++i) be[i] = {T_.b_idx[i], T_.b_val[i]};
        d.f_ent.upload(fe);
        d.b_ent.upload(be);
    }
    d.diag.upload(T_.diag);
    {
        // fp32 reciprocals of the diagonal: the fp32-vector sweeps multiply
        // by them (the tail layout stores the same values).
        std::vector<float> inv(T_.diag.size());
        for (std::size_t i = 0; i < inv.size(); ++i) inv[i] = 1.0f / T_.diag[i];
        d.inv_diag.upload(inv);
    }
    const batch_table fb = build_batches(T_, false), bb = build_batches(T_, true);
    validate_sweep_layout(T_, fb, false);
    validate_sweep_layout(T_, bb, true);
    if (T_.diag.size() != static_cast<std::size_t>(m) || F_.perm.size() != static_cast<std::size_t>(n))
        throw std::runtime_error("apxchol_amd: invalid factor layout");
    {
        std::vector<std::uint8_t> seen(n, 0);
        for (const node_t q : sperm) {
            if (q >= static_cast<node_t>(n) || seen[q]) throw std::runtime_error("apxchol_amd: perm is not a permutation");
            seen[q] = 1;
        }
    }
    d.f_pos.upload(fb.pos);
    d.f_mask.upload(fb.mask);
    d.b_pos.upload(bb.pos);
    d.b_mask.upload(bb.mask);
    d.f_levels = fb.level_first;
    d.b_levels = bb.level_first;

    std::vector<std::uint32_t> perm(sperm.begin(), sperm.end());
    d.perm.upload(perm);
    if (!relabel_.empty()) d.relabel.upload(relabel_);
    std::vector<std::uint32_t> optr(op.ptr.begin(), op.ptr.end());
    d.op_ptr.upload(optr);
    d.op_idx.upload(op.idx);
    d.op32 = op.fp32_exact;
    if (d.op32) d.op_val.upload(op.val32); else d.op_val.upload(op.val);
    // Lanes per SpMV row: power of two nearest the mean row length (<= 64).
    const double avg = n > 0 ? static_cast<double>(op.ptr.back()) / n : 1.0;
    int lanes = 1;
    while (lanes < 64 && lanes * 1.5 < avg) lanes <<= 1;
    d.lanes = lanes;
    // Sliced-ELL form of the operator for the single SpMV. Slices take the
    // rows in original vertex order (a row map), so a band of slices -- one
    // XCD's share, see spmv_sell -- gathers from a band of vertices and its
    // neighbours; unsorted when 64-row slices pad to <= 1.2 x the entries,
    // else sorted by length inside 4096-row windows (rows longer than
    // max(64, 8 mean) walked by a wave) when that pads to <= 1.5 x; otherwise
    // CSR lanes only. APXCHOL_AMD_SPMV=csr / sell forces the choice;
    // APXCHOL_AMD_SELL_ORDER=storage slices in storage order instead.
    {
        const std::int64_t nnz = op.ptr.back();
        const std::int64_t long_len = std::max<std::int64_t>(64, 8 * static_cast<std::int64_t>(std::ceil(avg)));
        const char* fe = std::getenv("APXCHOL_AMD_SPMV");
        const std::string force = fe ? fe : "";
        const char* oe = std::getenv("APXCHOL_AMD_SELL_ORDER");
        const bool storage_order = oe && std::strcmp(oe, "storage") == 0;
        if (n > 0 && force != "csr") {
            std::vector<std::int32_t> rows(static_cast<std::size_t>(n));
            for (int v = 0; v < n; ++v) rows[v] = storage_order ? v : static_cast<std::int32_t>(sperm[v]);
            std::int64_t long_nnz = 0;
            for (int r = 0; r < n; ++r)
                if (op.ptr[r + 1] - op.ptr[r] > long_len) long_nnz += op.ptr[r + 1] - op.ptr[r];
            const double short_nnz = static_cast<double>(std::max<std::int64_t>(nnz - long_nnz, 1));
            sell_layout sl;
            bool ok = false;
            if (static_cast<double>(sell_slots(op, long_len, rows)) <= 1.2 * short_nnz) {
                sl = build_sell(op, long_len, 1, rows);
                ok = true;
            } else {
                sl = build_sell(op, long_len, 4096, rows);
                ok = force == "sell" || static_cast<double>(sl.slots) <= 1.5 * short_nnz;
            }
            if (ok) {
                d.sell_soff.upload(sl.soff);
                d.sell_idx.upload(sl.idx);
                if (d.op32) d.sell_val.upload(sl.val32); else d.sell_val.upload(sl.val);
                d.sell_map.upload(sl.rowmap);
                d.sell_long.upload(sl.longrows);
                d.sell.n = n;
                d.sell.max_blocks = kSellMaxBlocks;
                if (const char* gb = std::getenv("APXCHOL_AMD_SELL_BLOCKS"); gb && *gb)
                    d.sell.max_blocks = std::clamp(std::atoi(gb), kXcdsHost, kSellMaxBlocks);
                d.sell.nslices = static_cast<int>(sl.soff.size()) - 1;
                d.sell.nlong = static_cast<int>(sl.longrows.size());
                d.sell.soff = d.sell_soff.as<std::uint32_t>();
                d.sell.idx = d.sell_idx.as<std::int32_t>();
                d.sell.val = d.sell_val.get();
                d.sell.rowmap = d.sell_map.as<std::int32_t>();
                d.sell.longrows = d.sell_long.as<std::int32_t>();
                d.sell.ptr = d.op_ptr.as<std::uint32_t>();
                d.sell.cidx = d.op_idx.as<std::int32_t>();
                d.sell.cval = d.op_val.get();
                d.has_sell = d.use_sell = true;
                stats_.spmv_sell_slots = sl.slots;
            }
        }
    }

    const std::size_t vb = static_cast<std::size_t>(std::max(n, 1)) * sizeof(double);
    d.r.alloc(vb); d.x.alloc(vb); d.p.alloc(vb); d.ap.alloc(vb); d.rp.alloc(vb); d.y.alloc(vb); d.z.alloc(vb);
    d.xout.alloc(vb);
    d.part.alloc(2 * static_cast<std::size_t>(std::max(kMaxBlocksHost, kSellMaxBlocks)) * sizeof(double));
    d.st.alloc(sizeof(pcg_state));
    d.st_apply.alloc(sizeof(pcg_state));
    d.st_apply.zero();
    d.ctrl.alloc(4 * sizeof(int));
    d.ctrl.zero();
    d.err.alloc(sizeof(int));
    d.err.zero();
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&d.host_st), sizeof(pcg_state), hipHostMallocDefault), "hipHostMalloc");
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&d.host_err), sizeof(int), hipHostMallocDefault), "hipHostMalloc");

    auto fill_desc = [&](sweep_desc& s, bool backward, const batch_table& bt) {
        s.m = m;
        s.backward = backward ? 1 : 0;
        s.ptr = (backward ? d.b_ptr : d.f_ptr).as<std::uint32_t>();
        if (T_.fp16) {
            s.idx = (backward ? d.b_idx : d.f_idx).as<std::int32_t>();
            s.h = (backward ? d.b_h : d.f_h).as<std::uint16_t>();
            s.inv_scale = backward ? d.inv_scale.as<float>() : nullptr;
        } else {
            s.ent = (backward ? d.b_ent : d.f_ent).as<entry32>();
        }
        s.diag = d.diag.as<float>();
        s.inv_diag = d.inv_diag.as<float>();
        s.batch_pos = (backward ? d.b_pos : d.f_pos).as<std::uint32_t>();
        s.batch_mask = (backward ? d.b_mask : d.f_mask).as<std::uint64_t>();
        s.nb = static_cast<int>(bt.mask.size());
        s.chunk = opts_.sweep_chunk > 0 ? opts_.sweep_chunk : env_int("APXCHOL_AMD_SWEEP_CHUNK", 8);
        s.ctrl = d.ctrl.as<int>() + (backward ? 2 : 0);
    };
    fill_desc(d.fwd, false, fb);
    fill_desc(d.bck, true, bb);
    // Workgroups of the dataflow launch (and cap of a level launch). The
    // dataflow default is 2 workgroups per CU: at the full resident grid (8
    // per CU) its polling traffic grew superlinearly and two 128^3 sweeps
    // exceeded the spin limit (reported, not hung), so it is kept well below.
    d.grid = opts_.sweep_blocks > 0 ? opts_.sweep_blocks : env_int("APXCHOL_AMD_SWEEP_BLOCKS", 0);
    const int validated_grid = std::max(1, sweep_dataflow_grid(d.prec, d.fp16) / 4);
    if (d.grid <= 0) d.grid = validated_grid;
    if (opts_.schedule == sweep_schedule::dataflow && d.grid > validated_grid)
        throw std::invalid_argument("gpu_solver: the dataflow schedule runs at most " + std::to_string(validated_grid) +
                                    " workgroups (2 per CU); larger grids exhausted the spin limit on 128^3 sweeps");
    // Hybrid segments: a level with at least tail_rows rows or more than
    // level_tail_batches batches is launched on its own; every maximal run of the
    // other (small) levels becomes one tail segment.
    const int tail_rows = std::min(opts_.tail_rows > 0 ? opts_.tail_rows : env_int("APXCHOL_AMD_TAIL_ROWS", 2048),
                                   sweep_tail_max_rows(d.prec));
    // A tail run's outputs (and its walk table) live in LDS: a run longer
    // than that holds is split.
    const vec_prec tail_prec = d.prec;
    // Batches a tail level may hold (up to kTailMaxBatches: each of the
    // workgroup's waves takes every kTailWaves-th batch).
    const int level_tail_batches = std::clamp(opts_.tail_batches > 0 ? opts_.tail_batches
                                                              : env_flag("APXCHOL_AMD_TAIL_BATCHES", 3 * kTailWaves),
                                        1, kTailMaxBatches);
    // A run ends before a level whose longest lane would read at least this
    // many overflow groups inside the run (0: runs end only at launched levels
    // and at the LDS capacity).
    const int tail_cut = opts_.tail_cut != 0 ? std::max(opts_.tail_cut, 0) : env_flag("APXCHOL_AMD_TAIL_CUT", 1);
    // Chains of one-batch tail levels share a barrier phase (0: one phase per
    // level).
    const bool merge_phases = env_flag("APXCHOL_AMD_TAIL_PHASES", 1) != 0;
    // Tail batches dealt to the waves round robin over the run (0: by their
    // index inside the level).
    const bool rotate_waves = env_flag("APXCHOL_AMD_TAIL_ROTATE", 1) != 0;
    // Launched levels of at most lane_max batches (and no row longer than a
    // wave's chunk) read the static lane layout.
    const int lane_max = opts_.lane_batches < 0 ? 0
                         : opts_.lane_batches > 0 ? opts_.lane_batches
                                                  : env_int("APXCHOL_AMD_LANE_BATCHES", 4096);
    auto make_sched = [&](sweep_sched& sc, const batch_table& bt, bool backward, const std::vector<int>& lf,
                          dbuf& lf_dev, std::vector<int> (&seg)[8], dbuf* const (&tbuf)[6], std::vector<int>& level_lb,
                          dbuf& lrec, dbuf& lent, std::vector<int>& level_pos, std::vector<unsigned char>& level_diag) {
        lf_dev.upload(lf);
        tail_tables tt;
        const int L = static_cast<int>(lf.size()) - 1;
        const auto pos = [&](int q) { return static_cast<int>(bt.pos[lf[q]]); };
        const auto rows = [&](int q) { return pos(q + 1) - pos(q); };
        // A tail level holds fewer than tail_rows rows in at most
        // level_tail_batches batches (a few per wave of the tail workgroup).
        const auto small = [&](int q) { return rows(q) < tail_rows && lf[q + 1] - lf[q] <= level_tail_batches; };
        // Overflow groups of level q inside a run starting at position p0: a
        // long row's lane holds every 64th entry, and the tail reads a lane's
        // entries past the first 2 * kTailPairs from HBM in dependent groups
        // of 8 while one wave walks the row. A level whose longest such chain
        // has at least tail_cut groups starts a new run instead: its entries
        // then point before the run and are summed by the coupling launch,
        // which spreads the rows over the whole device.
        const auto& sptr = backward ? T_.b_ptr : T_.f_ptr;
        const auto& sidx = backward ? T_.b_idx : T_.f_idx;
        const auto ovf_groups = [&](int q, int p0) {
            int worst = 0;
            for (int t = pos(q); t < pos(q + 1); ++t) {
                const int row = backward ? m - 1 - t : t;
                const std::uint32_t b0 = sptr[row], b1 = sptr[row + 1];
                if (b1 - b0 <= static_cast<std::uint32_t>(kWave * 2 * kTailPairs)) continue;
                // Entries are in ascending sweep position: the inside ones are a suffix.
                const auto first = std::partition_point(sidx.begin() + b0, sidx.begin() + b1, [&](std::int32_t j) {
                    return (backward ? m - 1 - j : j) < p0;
                });
                const std::uint32_t inside = static_cast<std::uint32_t>(sidx.begin() + b1 - first);
                const int per_lane = static_cast<int>((inside + kWave - 1) / kWave);
                worst = std::max(worst, (per_lane - 2 * kTailPairs + 7) / 8);
            }
            return worst;
        };
        for (int l = 0; l < L;) {
            if (!small(l)) {
                seg[0].push_back(l); seg[1].push_back(l + 1); seg[2].push_back(0);
                seg[3].push_back(pos(l)); seg[4].push_back(pos(l + 1)); seg[5].push_back(-1); seg[6].push_back(1);
                ++l;
                continue;
            }
            int e = l;
            while (e < L && small(e) && sweep_tail_fits(tail_prec, pos(e + 1) - pos(l), lf[e + 1] - lf[l], e + 1 - l) &&
                   (e == l || tail_cut <= 0 || ovf_groups(e, pos(l)) < tail_cut))
                ++e;
            if (e == l) ++e;   // a single small level always fits (rows < tail_rows <= cap)
            seg[0].push_back(l); seg[1].push_back(e); seg[2].push_back(1);
            seg[3].push_back(pos(l)); seg[4].push_back(pos(e));
            if (opts_.schedule == sweep_schedule::hybrid) {
                seg[5].push_back(append_tail_run(T_, bt, backward, lf[l], lf[e], tt));
                seg[6].push_back(append_tail_walk(lf, l, e, tt, merge_phases, rotate_waves));
            } else {
                seg[5].push_back(-1);
                seg[6].push_back(e - l);
            }
            l = e;
        }
        const std::size_t tb = static_cast<std::size_t>(tt.batches);
        if (tt.lane.size() != tb * kWave || tt.lane_recip.size() != tb * kWave ||
            tt.ent.size() != tb * kTailPairs * kWave || tt.ovf.size() != tb || tt.ovf_ent.size() % kWave != 0 ||
            tt.walk.size() != tb)
            throw std::runtime_error("apxchol_amd: inconsistent tail layout");
        tbuf[0]->upload(tt.lane);
        tbuf[1]->upload(tt.ent);
        tbuf[2]->upload(tt.ovf);
        tbuf[3]->upload(tt.ovf_ent);
        tbuf[4]->upload(tt.walk);
        tbuf[5]->upload(tt.lane_recip);
        sc.tail.lane = tbuf[0]->as<tail_lane>();
        sc.tail.lane_recip = tbuf[5]->as<tail_lane>();
        sc.tail.ent = tbuf[1]->as<tail_pair>();
        sc.tail.ovf = tbuf[2]->as<std::int32_t>();
        sc.tail.ovf_ent = tbuf[3]->as<entry32>();
        sc.tail.walk = tbuf[4]->as<tail_walk>();
        sc.tail_batches = tt.batches;
        sc.seg_base = seg[5].data();
        sc.seg_phases = seg[6].data();
        // A tail run right after another tail run sums its entries before
        // that run on the side stream while that run's tail walks (one
        // workgroup: the rest of the device is idle), and only the entries
        // into it afterwards -- the same chains (APXCHOL_AMD_COUPLE_SPLIT=0:
        // one coupling launch per run).
        // Only a coupling of at least split_min entries before that run is
        // split: a light one costs less than the extra launch and the join
        // (APXCHOL_AMD_COUPLE_SPLIT_MIN, entries).
        seg[7].assign(seg[0].size(), -1);
        if (opts_.schedule == sweep_schedule::hybrid && env_flag("APXCHOL_AMD_COUPLE_SPLIT", 1) != 0) {
            const std::uint64_t split_min =
                static_cast<std::uint64_t>(std::max(env_flag("APXCHOL_AMD_COUPLE_SPLIT_MIN", 1 << 20), 0));
            for (std::size_t k = 1; k < seg[0].size(); ++k) {
                if (!seg[2][k] || !seg[2][k - 1] || seg[3][k - 1] <= 0) continue;
                const int ps = seg[3][k - 1];
                std::uint64_t before = 0;
                for (int t = seg[3][k]; t < seg[4][k]; ++t) {
                    const int row = backward ? m - 1 - t : t;
                    const auto b0 = sidx.begin() + sptr[row], b1 = sidx.begin() + sptr[row + 1];
                    before += static_cast<std::uint64_t>(
                        std::partition_point(b0, b1, [&](std::int32_t j) { return (backward ? m - 1 - j : j) < ps; }) - b0);
                }
                if (before >= split_min) {
                    seg[7][k] = ps;
                    ++stats_.split_couplings[backward ? 1 : 0];
                }
            }
        }
        sc.seg_split = seg[7].data();
        sc.side = d.side;
        sc.ev_fork = d.ev_fork;
        sc.ev_join = d.ev_join;
        sc.kind = static_cast<int>(opts_.schedule == sweep_schedule::dataflow ? 0
                                   : opts_.schedule == sweep_schedule::levels ? 1 : 2);
        // Static lane layout of the launched levels.
        level_lb.assign(static_cast<std::size_t>(L), -1);
        lane_tables lt;
        if (sc.kind != 0 && lane_max > 0) {
            const auto& ptr = backward ? T_.b_ptr : T_.f_ptr;
            std::vector<char> in_tail(static_cast<std::size_t>(L), 0);
            if (sc.kind == 2)
                for (std::size_t k = 0; k < seg[0].size(); ++k)
                    if (seg[2][k])
                        for (int q = seg[0][k]; q < seg[1][k]; ++q) in_tail[q] = 1;
            for (int q = 0; q < L; ++q) {
                const int nbq = lf[q + 1] - lf[q];
                if (in_tail[q] || nbq <= 0 || nbq > lane_max) continue;
                bool longrow = false;
                for (int t = pos(q); t < pos(q + 1) && !longrow; ++t) {
                    const int row = backward ? m - 1 - t : t;
                    longrow = ptr[row + 1] - ptr[row] > static_cast<std::uint32_t>(kWave * kEpl);
                }
                if (longrow) continue;
                const long long lb = append_lane_level(T_, bt, backward, lf[q], lf[q + 1], lt);
                if (lb + nbq > INT32_MAX) throw std::runtime_error("apxchol_amd: lane layout too large");
                level_lb[q] = static_cast<int>(lb);
                ++stats_.lane_levels[backward ? 1 : 0];
            }
        }
        if (lt.rec.size() != static_cast<std::size_t>(lt.batches) * kWave ||
            lt.ent.size() != static_cast<std::size_t>(lt.batches) * kWave * kEpl)
            throw std::runtime_error("apxchol_amd: inconsistent lane layout");
        lrec.upload(lt.rec);
        lent.upload(lt.ent);
        sc.level_lb = level_lb.data();
        sc.lane_rec_dev = lrec.as<lane_rec>();
        sc.lane_ent_dev = lent.as<entry32>();
        sc.lane_batches = lt.batches;
        // Big launched levels run several batches per wave (fewer waves to
        // dispatch; the same bits).
        sc.lane_bpw2 = env_int("APXCHOL_AMD_LANE_BPW2", 0);
        sc.lane_bpw4 = env_int("APXCHOL_AMD_LANE_BPW4", 0);
        // Algorithmic bytes per level (see solver_stats).
        if (sc.kind != 0) {
            const auto& ptr = backward ? T_.b_ptr : T_.f_ptr;
            const std::uint64_t vf = T_.fp16 ? 2 : 4;
            const std::uint64_t vin = backward ? (d.prec == vec_prec::fp32 ? 4 : 8) : 8;
            const std::uint64_t vout = d.prec == vec_prec::fp32 ? 4 : 8;
            const std::uint64_t row_b = 8 + vin + vout + ((T_.fp16 && backward) ? 4 : 0);
            auto level_b = [&](int q) {
                std::uint64_t off = 0;
                for (int t = pos(q); t < pos(q + 1); ++t) {
                    const int row = backward ? m - 1 - t : t;
                    off += ptr[row + 1] - ptr[row];
                }
                return off * (4 + vf) + static_cast<std::uint64_t>(rows(q)) * row_b +
                       12ull * static_cast<std::uint64_t>(lf[q + 1] - lf[q]);
            };
            const int di = backward ? 1 : 0;
            for (std::size_t k = 0; k < seg[0].size(); ++k) {
                const bool tail = sc.kind == 2 && seg[2][k];
                for (int q = seg[0][k]; q < seg[1][k]; ++q) {
                    if (tail) stats_.tail_bytes[di] += level_b(q);
                    else { stats_.level_bytes[di] += level_b(q); ++stats_.level_launches[di]; }
                }
                if (tail) ++stats_.tail_runs[di];
            }
        }
        sc.level_first = lf.data();
        sc.level_first_dev = lf_dev.as<int>();
        sc.nlevels = L;
        // Levels none of whose rows has an entry (the forward sweep's first
        // level) are launched as one load per row (sweep_diag_level, the same
        // bits); APXCHOL_AMD_DIAG_LEVELS=0 keeps them on the batch kernels.
        {
            const auto& ptr = backward ? T_.b_ptr : T_.f_ptr;
            level_pos.resize(static_cast<std::size_t>(L) + 1);
            for (int q = 0; q <= L; ++q) level_pos[q] = pos(q);
            level_diag.assign(static_cast<std::size_t>(L), 0);
            for (int q = 0; q < L; ++q) {
                bool none = pos(q + 1) > pos(q);
                for (int t = pos(q); t < pos(q + 1) && none; ++t) {
                    const int row = backward ? m - 1 - t : t;
                    none = ptr[row + 1] == ptr[row];
                }
                level_diag[q] = none ? 1 : 0;
            }
            sc.level_pos = level_pos.data();
            sc.level_diag = env_flag("APXCHOL_AMD_DIAG_LEVELS", 1) != 0 ? level_diag.data() : nullptr;
        }
        sc.seg_l0 = seg[0].data();
        sc.seg_l1 = seg[1].data();
        sc.seg_tail = seg[2].data();
        sc.seg_p0 = seg[3].data();
        sc.seg_p1 = seg[4].data();
        sc.nseg = static_cast<int>(seg[0].size());
        sc.grid = d.grid;
    };
    hip_check(hipStreamCreateWithFlags(&d.side, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&d.ev_fork, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&d.ev_join, hipEventDisableTiming), "hipEventCreate");
    dbuf* f_tb[6] = {&d.f_tlane, &d.f_tent, &d.f_tovf, &d.f_tovf_ent, &d.f_twalk, &d.f_tlane_r};
    dbuf* b_tb[6] = {&d.b_tlane, &d.b_tent, &d.b_tovf, &d.b_tovf_ent, &d.b_twalk, &d.b_tlane_r};
    make_sched(d.f_sched, fb, false, d.f_levels, d.f_levels_dev, d.f_seg, f_tb, d.f_level_lb, d.f_lrec, d.f_lent,
               d.f_level_pos, d.f_level_diag);
    make_sched(d.b_sched, bb, true, d.b_levels, d.b_levels_dev, d.b_seg, b_tb, d.b_level_lb, d.b_lrec, d.b_lent,
               d.b_level_pos, d.b_level_diag);
    // Coupling output, shared by the two directions (their sweeps never overlap).
    const std::size_t tail_batches = std::max<std::size_t>(
        1, static_cast<std::size_t>(std::max(d.f_sched.tail_batches, d.b_sched.tail_batches)));
    d.tail_val.alloc(tail_batches * kWave * sizeof(tail_val<double>));
    for (auto* sc : {&d.f_sched, &d.b_sched}) {
        sc->tail_val = d.tail_val.get();
        sc->tail_val_bytes = d.tail_val.bytes();
    }

    stats_.n = n_;
    stats_.m = m;
    stats_.sddm = F_.sddm;
    stats_.factor_nnz = F_.nnz();
    stats_.stored_nnz = T_.drop.nnz_stored;
    stats_.offdiag = T_.offdiag();
    stats_.dropped = T_.drop.dropped;
    stats_.rounds = static_cast<std::int64_t>(F_.rounds.size());
    stats_.levels = static_cast<std::int64_t>(T_.level_ptr.size()) - 1;
    stats_.fwd_batches = d.fwd.nb;
    stats_.bck_batches = d.bck.nb;
    stats_.op_nnz = op.ptr.back();
    stats_.op_fp32 = d.op32;
    stats_.spmv_lanes = lanes;
    stats_.spmv_format = d.use_sell ? 1 : 0;
    stats_.dataflow_grid = d.grid;
    std::uint64_t bytes = 0;
    for (const dbuf* b : {&d.f_ptr, &d.f_ent, &d.f_idx, &d.f_h, &d.b_ptr, &d.b_ent, &d.b_idx, &d.b_h, &d.diag, &d.inv_diag,
                          &d.inv_scale, &d.f_pos, &d.f_mask, &d.b_pos, &d.b_mask, &d.perm, &d.op_ptr, &d.op_idx,
                          &d.op_val, &d.r, &d.x, &d.p, &d.ap, &d.rp, &d.y, &d.z, &d.xout, &d.f_lrec, &d.f_lent,
                          &d.b_lrec, &d.b_lent, &d.f_tlane, &d.f_tent, &d.f_tovf, &d.f_tovf_ent, &d.f_twalk,
                          &d.b_tlane, &d.b_tent, &d.b_tovf, &d.b_tovf_ent, &d.b_twalk, &d.f_tlane_r, &d.b_tlane_r,
                          &d.tail_val, &d.relabel, &d.sell_soff,
                          &d.sell_idx, &d.sell_val, &d.sell_map, &d.sell_long})
        bytes += b->bytes();
    stats_.device_bytes = bytes;

    if (!opts_.keep_factor_values) {
        std::vector<node_t>().swap(F_.inner);
        std::vector<float>().swap(F_.vals);
    }
    hip_check(hipDeviceSynchronize(), "setup");
    // Everything but the operator build is the SpTRSV side (preparation,
    // layouts, uploads of both).
    setup_times_.sptrsv_setup_ms = ms_since(t_setup) - setup_times_.pcg_setup_ms;
}

void gpu_solver::check_sweep_error() const {
    hip_check(hipMemcpy(dev_->host_err, dev_->err.get(), sizeof(int), hipMemcpyDeviceToHost), "read sweep flag");
    if (*dev_->host_err != 0) {
        dev_->err.zero();
        throw std::runtime_error("apxchol_amd: a sweep dependency did not arrive within the spin limit");
    }
}

// Step k: the preconditioner application after iteration k (k = 0 is the one
// before the loop); the SpMV / update of iteration k carry step k - 1, so a
// solve that stops at iteration `it` keeps steps 0 .. it-1.
void gpu_solver::enqueue_precondition(bool first, int step, hipStream_t s) const {
    auto& d = *dev_;
    timer& tm = *timer_;
    const int n = static_cast<int>(n_);
    const int m = static_cast<int>(T_.m);
    pcg_state* st = d.st.as<pcg_state>();
    double* part = d.part.as<double>();
    int* err = d.err.as<int>();
    const vec_prec pr = d.prec;
    // The pending sentinels only for the dataflow schedule (the only one
    // that polls the outputs); the zero past the sweep rows of z always.
    void* sentinel_y = opts_.schedule == sweep_schedule::dataflow ? d.y.get() : nullptr;
    if (first) launch_fill_outputs(pr, n, m, sentinel_y, d.z.get(), st, s);
    tm.time(0, step, s, [&] {
        launch_sweep(d.fwd, pr, d.fp16, d.r.get(), d.y.get(), st, err, d.f_sched, s, tm.hook(step));
    });
    tm.time(1, step, s, [&] {
        launch_sweep(d.bck, pr, d.fp16, d.y.get(), d.z.get(), st, err, d.b_sched, s, tm.hook(step));
    });
    tm.time(3, step, s, [&] {
        launch_dot_center(pr, n, d.r.as<double>(), d.z.get(), st, part, first, s);
        launch_update_p(pr, n, d.p.as<double>(), d.z.get(), st, first, s);
    });
}

void gpu_solver::spmv(const double* x, double* y, pcg_state* st, bool want_dot, hipStream_t s,
                      const launch_hook& hook) const {
    auto& d = *dev_;
    if (d.use_sell)
        launch_spmv_sell(d.sell, d.op32, x, y, st, d.part.as<double>(), want_dot, s, hook);
    else
        launch_spmv(static_cast<int>(n_), d.op_ptr.as<unsigned>(), d.op_idx.as<int>(), d.op_val.get(), d.op32, d.lanes,
                    x, y, st, d.part.as<double>(), want_dot, s, hook);
}

void gpu_solver::enqueue_iteration(int it, hipStream_t s) const {
    auto& d = *dev_;
    timer& tm = *timer_;
    const int n = static_cast<int>(n_);
    const int m = static_cast<int>(T_.m);
    pcg_state* st = d.st.as<pcg_state>();
    double* part = d.part.as<double>();
    spmv(d.p.as<double>(), d.ap.as<double>(), st, true, s, tm.hook(it - 1));
    tm.time(3, it - 1, s, [&] {
        launch_update_xr(d.prec, n, m, d.x.as<double>(), d.p.as<double>(), d.r.as<double>(), d.ap.as<double>(),
                         opts_.schedule == sweep_schedule::dataflow ? d.y.get() : nullptr, d.z.get(), st, part, s);
    });
    enqueue_precondition(false, it, s);
}

solve_result gpu_solver::solve_device(const double* d_b, double* d_x, double tol, int max_iter, const double* d_x0,
                                      hipStream_t s) const {
    not_serving("solve");
    const auto t_solve = std::chrono::steady_clock::now();
    if (tol < 0.0) tol = opts_.tol;
    if (max_iter < 0) max_iter = opts_.max_iter;
    solve_result res;
    res.lumped_offdiag = F_.lumped_offdiag;
    res.timings = setup_times_;
    if (n_ == 0) return res;
    single_begin(d_b, d_x0, tol, max_iter, s);
    const int J = std::max(1, opts_.iters_per_check);
    session_slot state;
    do single_step(J, s);
    while (!single_poll(state, s));
    single_collect(d_x, res, s);
    res.timings.pcg_ms = ms_since(t_solve);
    std::size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess && tot >= fr) res.solve_vram_mb = static_cast<double>(tot - fr) / (1024.0 * 1024.0);
    return res;
}

void gpu_solver::single_begin(const double* d_b, const double* d_x0, double tol, int max_iter, hipStream_t s) const {
    auto& d = *dev_;
    const int n = static_cast<int>(n_);
    if (n == 0) throw std::invalid_argument("gpu_solver: an empty system has no solve steps");
    pcg_state& h = *d.host_st;
    std::memset(&h, 0, sizeof h);
    h.tol = tol;
    h.max_iter = max_iter;
    h.stag_window = opts_.stagnation_window;
    h.center_k = center_k_;
    h.laplacian = F_.sddm ? 0 : 1;
    h.n = n;
    h.prev_check = 1.0;
    pcg_state* st = d.st.as<pcg_state>();
    double* part = d.part.as<double>();
    hip_check(hipMemcpyAsync(st, &h, sizeof h, hipMemcpyHostToDevice, s), "state upload");

    launch_scatter_init(n, d.perm.as<unsigned>(), d_b, d_x0, d.r.as<double>(), d.x.as<double>(), s);
    if (d_x0) {
        launch_norm2_b(n, d_b, st, part, s);
        spmv(d.x.as<double>(), d.ap.as<double>(), st, false, s, {});
        launch_init_residual(n, d.r.as<double>(), d.ap.as<double>(), st, part, s);
    } else {
        launch_init_residual(n, d.r.as<double>(), nullptr, st, part, s);
    }
    enqueue_precondition(true, 0, s);
    // Graph replay of the iteration; eager launches while timing.
    d.single_graph = opts_.graph && !timer_->on;
    if (d.single_graph) single_prepare();
    d.single_enqueued = 0;
    d.single_last = 1;
    d.single_max_iter = max_iter;
}

void gpu_solver::set_split_couplings(bool on) const {
    auto& d = *dev_;
    const int* f = on ? d.f_seg[7].data() : nullptr;
    const int* b = on ? d.b_seg[7].data() : nullptr;
    if (d.f_sched.seg_split == f && d.b_sched.seg_split == b) return;
    d.f_sched.seg_split = f;
    d.b_sched.seg_split = b;
    if (d.iter_exec) { (void)hipGraphExecDestroy(d.iter_exec); d.iter_exec = nullptr; }
    if (d.iter_graph) { (void)hipGraphDestroy(d.iter_graph); d.iter_graph = nullptr; }
}

void gpu_solver::single_prepare() const {
    auto& d = *dev_;
    if (!opts_.graph || d.iter_exec || n_ == 0) return;
    // Every argument of the iteration's kernels is a pointer into this
    // solver's device state or a size, so one capture serves every solve.
    std::lock_guard<std::mutex> capture_lock(g_capture_mu);
    if (!d.cap) hip_check(hipStreamCreateWithFlags(&d.cap, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamBeginCapture(d.cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        enqueue_iteration(0, d.cap);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(d.cap, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    hip_check(hipStreamEndCapture(d.cap, &d.iter_graph), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&d.iter_exec, d.iter_graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

void gpu_solver::single_step(int iterations, hipStream_t s) const {
    auto& d = *dev_;
    for (int j = 0; j < iterations; ++j) {
        if (d.single_graph) hip_check(hipGraphLaunch(d.iter_exec, s), "hipGraphLaunch");
        else enqueue_iteration(d.single_enqueued + j + 1, s);
    }
    d.single_enqueued += iterations;
    d.single_last = std::max(1, iterations);
}

bool gpu_solver::single_poll(session_slot& out, hipStream_t s) const {
    auto& d = *dev_;
    pcg_state& h = *d.host_st;
    hip_check(hipMemcpyAsync(&h, d.st.get(), sizeof h, hipMemcpyDeviceToHost, s), "state read");
    hip_check(hipStreamSynchronize(s), "solve");
    out.done = h.done != 0;
    out.status = h.status;
    out.iterations = h.iter;
    out.residual = h.residual;
    // The enqueued iterations run out once max_iter is passed (the device
    // stops at max_iter; this bounds the host loop whatever the state says).
    return out.done || d.single_enqueued > d.single_max_iter + d.single_last;
}

void gpu_solver::single_collect(double* d_x, solve_result& res, hipStream_t s) const {
    auto& d = *dev_;
    const int n = static_cast<int>(n_);
    pcg_state* st = d.st.as<pcg_state>();
    double* part = d.part.as<double>();
    pcg_state& h = *d.host_st;
    if (!F_.sddm) launch_sum_vec(n, d.x.as<double>(), st, part, s);
    launch_gather_shift(n, d.perm.as<unsigned>(), d.x.as<double>(), d_x, st, !F_.sddm, s);
    hip_check(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, s), "state read");
    hip_check(hipStreamSynchronize(s), "solve");
    if (h.status == 5) {
        hip_check(hipMemsetAsync(d_x, 0, sizeof(double) * n, s), "zero x");
        hip_check(hipStreamSynchronize(s), "solve");
    }
    timer_->collect(h.iter - 1);
    check_sweep_error();
    res.lumped_offdiag = F_.lumped_offdiag;
    res.iterations = h.iter;
    res.residual = h.residual;
    res.status = h.status;
}

// ---- blocked solves -----------------------------------------------------

// Block width for k columns (a power of two <= 64), with the block buffers
// grown to it.
int gpu_solver::prepare_block(int k, int which) const {
    auto& b = block_in(*dev_, which);
    int kp = 1;
    while (kp < k) kp <<= 1;
    if (kp > 64) throw std::invalid_argument("gpu_solver: a block holds at most 64 columns");
    if (b.kp >= kp) return kp;
    b.drop_graph();
    const std::size_t vb = static_cast<std::size_t>(std::max<std::int64_t>(n_, 1)) * kp * sizeof(double);
    b.r.alloc(vb); b.x.alloc(vb); b.p.alloc(vb); b.ap.alloc(vb); b.y.alloc(vb); b.z.alloc(vb);
    b.st.alloc(static_cast<std::size_t>(kp) * sizeof(pcg_state));
    b.bs.alloc(sizeof(batch_state));
    // Partials: 2 values x kp columns per workgroup; the SpMM grid may round
    // kMaxBlocksHost up by at most 15 (a multiple of G kp / 256 <= 16).
    b.part.alloc(static_cast<std::size_t>(kMaxBlocksHost + 16) * 2 * kp * sizeof(double));
    if (b.host_st) (void)hipHostFree(b.host_st);
    b.host_st = nullptr;
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&b.host_st), static_cast<std::size_t>(kp) * sizeof(pcg_state),
                            hipHostMallocDefault), "hipHostMalloc");
    if (!b.host_bs)
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&b.host_bs), sizeof(batch_state), hipHostMallocDefault),
                  "hipHostMalloc");
    if (!b.host_adm) {
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&b.host_adm), 64 * sizeof(int) + 64 * sizeof(void*),
                                hipHostMallocDefault), "hipHostMalloc");
        b.adm_dev.alloc(64 * sizeof(int) + 64 * sizeof(void*));
    }
    b.kp = kp;
    return kp;
}

void gpu_solver::apply_block(int k, const double* d_R, double* d_Z, bool center, hipStream_t s) const {
    auto& d = *dev_;
    auto& b = d.blk;
    const int n = static_cast<int>(n_);
    const int m = static_cast<int>(T_.m);
    const int kp = prepare_block(k);
    for (int c = 0; c < kp; ++c) {
        std::memset(&b.host_st[c], 0, sizeof(pcg_state));
        b.host_st[c].n = n;
    }
    std::memset(b.host_bs, 0, sizeof(batch_state));
    b.host_bs->active = kp;
    pcg_state* st = b.st.as<pcg_state>();
    batch_state* bs = b.bs.as<batch_state>();
    double* part = b.part.as<double>();
    hip_check(hipMemcpyAsync(st, b.host_st, static_cast<std::size_t>(kp) * sizeof(pcg_state), hipMemcpyHostToDevice, s),
              "state upload");
    hip_check(hipMemcpyAsync(bs, b.host_bs, sizeof(batch_state), hipMemcpyHostToDevice, s), "state upload");
    launch_scatter_init_k(n, k, kp, d.perm.as<unsigned>(), d_R, b.r.as<double>(), b.x.as<double>(), s);
    launch_zero_tail_k(d.prec, n, m, kp, b.z.get(), s);
    if (center) {
        launch_sum_cols_k(vec_prec::fp64, n, kp, b.r.get(), st, part, s);
        launch_shift_from_sum_k(kp, st, true, s);
    }
    block_scheds(kp);
    const auto& t = *b.tables.at(kp);
    launch_sweep_k(d.fwd, d.prec, d.fp16, t.f, b.r.get(), b.y.get(), st, bs, s);
    launch_sweep_k(d.bck, d.prec, d.fp16, t.b, b.y.get(), b.z.get(), st, bs, s);
    if (center) launch_sum_cols_k(d.prec, n, kp, b.z.get(), st, part, s);
    launch_gather_k(d.prec, n, k, kp, d.perm.as<unsigned>(), b.z.get(), d_Z, st, center, s);
}

void gpu_solver::spmv_block(int k, const double* d_X, double* d_Y, hipStream_t s) const {
    auto& d = *dev_;
    auto& b = d.blk;
    const int n = static_cast<int>(n_);
    const int kp = prepare_block(k);
    std::memset(b.host_bs, 0, sizeof(batch_state));
    b.host_bs->active = kp;
    batch_state* bs = b.bs.as<batch_state>();
    hip_check(hipMemcpyAsync(bs, b.host_bs, sizeof(batch_state), hipMemcpyHostToDevice, s), "state upload");
    launch_scatter_init_k(n, k, kp, d.perm.as<unsigned>(), d_X, b.p.as<double>(), b.x.as<double>(), s);
    if (d.use_sell)
        launch_spmm_sell_k(d.sell, d.op32, kp, b.p.as<double>(), b.ap.as<double>(), nullptr, bs, b.part.as<double>(),
                           false, s);
    else
        launch_spmm_k(n, kp, d.lanes, d.op_ptr.as<unsigned>(), d.op_idx.as<int>(), d.op_val.get(), d.op32,
                      b.p.as<double>(), b.ap.as<double>(), nullptr, bs, b.part.as<double>(), false, s);
    launch_gather_k(vec_prec::fp64, n, k, kp, d.perm.as<unsigned>(), b.ap.get(), d_Y, b.st.as<pcg_state>(), false, s);
}

void gpu_solver::spmv_batch_device(int k, const double* d_X, double* d_Y, hipStream_t s) const {
    not_serving("spmv_batch");
    if (k < 0) throw std::invalid_argument("gpu_solver::spmv_batch: negative column count");
    if (n_ == 0 || k == 0) return;
    for (int c0 = 0; c0 < k; c0 += 64)
        spmv_block(std::min(64, k - c0), d_X + static_cast<std::size_t>(c0) * n_, d_Y + static_cast<std::size_t>(c0) * n_,
                   s);
}

void gpu_solver::apply_batch_device(int k, const double* d_R, double* d_Z, hipStream_t s) const {
    not_serving("apply_batch");
    if (k < 0) throw std::invalid_argument("gpu_solver::apply_batch: negative column count");
    if (n_ == 0 || k == 0) return;
    // One application of the center-k schedule for all columns.
    const bool center = !F_.sddm && (++apply_count_ % static_cast<std::uint64_t>(center_k_) == 0);
    for (int c0 = 0; c0 < k; c0 += 64)
        apply_block(std::min(64, k - c0), d_R + static_cast<std::size_t>(c0) * n_,
                    d_Z + static_cast<std::size_t>(c0) * n_, center, s);
}

// The blocked sweep tables of block width kp (built, checked and uploaded on
// first use).
const block_sched* gpu_solver::block_scheds(int kp, int which) const {
    auto& b = block_in(*dev_, which);
    auto it = b.tables.find(kp);
    if (it == b.tables.end()) {
        auto t = std::make_unique<device_state::block_tables>();
        for (int dir = 0; dir < 2; ++dir) {
            const bool backward = dir == 1;
            const block_batch_table bt = build_block_batches(T_, backward, kp);
            validate_block_batches(T_, bt, backward);
            dbuf& dp = backward ? t->b_pos : t->f_pos;
            dbuf& dm = backward ? t->b_mask : t->f_mask;
            dbuf& dc = backward ? t->b_col0 : t->f_col0;
            dp.upload(bt.pos);
            dm.upload(bt.mask);
            dc.upload(bt.col0);
            std::vector<int>& lf = backward ? t->b_lf : t->f_lf;
            lf = bt.level_first;
            dbuf& lfd = backward ? t->b_lf_dev : t->f_lf_dev;
            lfd.upload(lf);
            // First sweep position of each level.
            std::vector<int>& lpos = backward ? t->b_lpos : t->f_lpos;
            const std::size_t L = T_.level_ptr.size() - 1;
            lpos.resize(L + 1);
            for (std::size_t l = 0; l <= L; ++l)
                lpos[l] = static_cast<int>(backward ? T_.m - T_.level_ptr[L - l] : T_.level_ptr[l]);
            // Lanes of each level's widest row (the per-(row, column) level
            // kernel takes levels up to 16; 0 marks a row past 64 lanes, -1
            // a level without any entry, run as one load per (row, column)
            // unless APXCHOL_AMD_DIAG_LEVELS=0).
            std::vector<int>& lg = backward ? t->b_lgmax : t->f_lgmax;
            lg.assign(L, 0);
            const auto& rptr = backward ? T_.b_ptr : T_.f_ptr;
            const bool diag_levels = env_flag("APXCHOL_AMD_DIAG_LEVELS", 1) != 0;
            for (std::size_t l = 0; l < L; ++l) {
                int gm = 1;
                std::uint64_t entries = 0;
                for (int q = lpos[l]; q < lpos[l + 1]; ++q) {
                    const std::uint32_t row = backward ? T_.m - 1 - static_cast<std::uint32_t>(q) : static_cast<std::uint32_t>(q);
                    const std::uint32_t len = rptr[row + 1] - rptr[row];
                    gm = len > static_cast<std::uint32_t>(kWave * kEpl) ? 1 << 30 : std::max(gm, lanes_for_host(len));
                    entries += len;
                }
                lg[l] = (diag_levels && entries == 0 && lpos[l + 1] > lpos[l]) ? -1 : gm > kWave ? 0 : gm;
            }
            // Segments: runs of levels of at most block_tail_waves table
            // waves whose rows fit the LDS go to one workgroup; the others
            // are launched.
            std::vector<int>(&seg)[3] = backward ? t->b_seg : t->f_seg;
            const int cap = sweep_block_tail_max_rows(dev_->prec, kp);
            const bool tails = opts_.schedule == sweep_schedule::hybrid && cap > 0;
            // (16 by default: wider levels -- long rows times the block's
            // columns -- ran slower walked by one workgroup than launched.)
            const int block_tail_waves =
                std::clamp(env_flag("APXCHOL_AMD_BLOCK_TAIL_WAVES", kTailWaves), 1, kBlockTailMaxWaves);
            const int nl = static_cast<int>(L);
            for (int l = 0; l < nl;) {
                auto small = [&](int q) { return tails && lf[q + 1] - lf[q] <= block_tail_waves; };
                if (!small(l)) {
                    seg[0].push_back(l); seg[1].push_back(l + 1); seg[2].push_back(0);
                    ++l;
                    continue;
                }
                int e = l;
                while (e < nl && small(e) && lpos[e + 1] - lpos[l] <= cap) ++e;
                if (e == l) {   // one small level wider than the LDS: launch it
                    seg[0].push_back(l); seg[1].push_back(l + 1); seg[2].push_back(0);
                    ++l;
                    continue;
                }
                seg[0].push_back(l); seg[1].push_back(e); seg[2].push_back(1);
                l = e;
            }
            block_sched& sc = backward ? t->b : t->f;
            sc.kp = kp;
            sc.pos = dp.as<unsigned>();
            sc.mask = dm.as<unsigned long long>();
            sc.col0 = dc.as<int>();
            sc.level_first = lf.data();
            sc.level_first_dev = lfd.as<int>();
            sc.level_pos = lpos.data();
            sc.level_gmax = env_flag("APXCHOL_AMD_BLOCK_ROWS", 1) != 0 ? lg.data() : nullptr;
            sc.rows_max_g = env_flag("APXCHOL_AMD_BLOCK_ROWS_G", 16) >= 16 ? 16 : 8;
            sc.nlevels = static_cast<int>(lf.size()) - 1;
            sc.seg_l0 = seg[0].data();
            sc.seg_l1 = seg[1].data();
            sc.seg_tail = seg[2].data();
            sc.nseg = static_cast<int>(seg[0].size());
        }
        // Column-by-column tail runs: the blocked sweep follows the
        // single-vector schedule (its launched levels as blocked level
        // launches, its tail runs with one tail workgroup per column), so
        // the runs' level structure and overflow cuts carry over.
        if (opts_.schedule == sweep_schedule::hybrid && env_flag("APXCHOL_AMD_BLOCK_COL_TAIL", 1) != 0) {
            std::size_t maxb = 0;
            for (const sweep_sched* sc : {&dev_->f_sched, &dev_->b_sched})
                for (int k = 0; k < sc->nseg; ++k)
                    if (sc->seg_tail[k])
                        maxb = std::max<std::size_t>(
                            maxb, static_cast<std::size_t>(sc->level_first[sc->seg_l1[k]] - sc->level_first[sc->seg_l0[k]]));
            const std::size_t tvb = dev_->prec == vec_prec::fp64 ? sizeof(tail_val<double>) : sizeof(tail_val<float>);
            const std::size_t bytes = std::max<std::size_t>(1, maxb) * static_cast<std::size_t>(kp) * kWave * tvb;
            t->col_tv.alloc(bytes);
            for (block_sched* sc : {&t->f, &t->b}) {
                sc->cols = sc == &t->f ? &dev_->f_sched : &dev_->b_sched;
                sc->col_tail_val = t->col_tv.get();
                sc->col_tail_val_bytes = bytes;
            }
        }
        it = b.tables.emplace(kp, std::move(t)).first;
    }
    return &it->second->f;
}

void gpu_solver::enqueue_precondition_k(bool first, int kp, hipStream_t s, int which) const {
    auto& d = *dev_;
    auto& b = block_in(*dev_, which);
    const int n = static_cast<int>(n_);
    pcg_state* st = b.st.as<pcg_state>();
    batch_state* bs = b.bs.as<batch_state>();
    block_scheds(kp, which);
    const auto& t = *b.tables.at(kp);
    launch_sweep_k(d.fwd, d.prec, d.fp16, t.f, b.r.get(), b.y.get(), st, bs, s);
    launch_sweep_k(d.bck, d.prec, d.fp16, t.b, b.y.get(), b.z.get(), st, bs, s);
    launch_dot_center_k(d.prec, n, kp, b.r.as<double>(), b.z.get(), st, bs, b.part.as<double>(), first, s);
    launch_update_p_k(d.prec, n, kp, b.p.as<double>(), b.z.get(), st, bs, first, s);
}

void gpu_solver::enqueue_iteration_k(int kp, hipStream_t s, int which) const {
    auto& d = *dev_;
    auto& b = block_in(*dev_, which);
    const int n = static_cast<int>(n_);
    pcg_state* st = b.st.as<pcg_state>();
    batch_state* bs = b.bs.as<batch_state>();
    if (d.use_sell)
        launch_spmm_sell_k(d.sell, d.op32, kp, b.p.as<double>(), b.ap.as<double>(), st, bs, b.part.as<double>(), true, s);
    else
        launch_spmm_k(n, kp, d.lanes, d.op_ptr.as<unsigned>(), d.op_idx.as<int>(), d.op_val.get(), d.op32,
                      b.p.as<double>(), b.ap.as<double>(), st, bs, b.part.as<double>(), true, s);
    launch_update_xr_k(n, kp, b.x.as<double>(), b.p.as<double>(), b.r.as<double>(), b.ap.as<double>(), st, bs,
                       b.part.as<double>(), s);
    launch_refresh_active_k(kp, st, bs, s);
    enqueue_precondition_k(false, kp, s, which);
    launch_end_iteration_k(kp, st, bs, s);
}

// The iteration graph of block width kp (captured on first use).
void gpu_solver::replay_iterations_k(int kp, int count, hipStream_t s, int which) const {
    auto& d = *dev_;
    auto& b = block_in(*dev_, which);
    if (!opts_.graph) {
        for (int j = 0; j < count; ++j) enqueue_iteration_k(kp, s, which);
        return;
    }
    if (b.graph_kp != kp) {
        std::lock_guard<std::mutex> capture_lock(g_capture_mu);
        b.drop_graph();
        if (!d.cap) hip_check(hipStreamCreateWithFlags(&d.cap, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipStreamBeginCapture(d.cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            enqueue_iteration_k(kp, d.cap, which);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(d.cap, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(d.cap, &b.graph), "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        b.graph_kp = kp;
    }
    for (int j = 0; j < count; ++j) hip_check(hipGraphLaunch(b.exec, s), "hipGraphLaunch");
}

void gpu_solver::fresh_state(pcg_state& h, double tol, int max_iter) const {
    std::memset(&h, 0, sizeof h);
    h.tol = tol;
    h.max_iter = max_iter;
    h.stag_window = opts_.stagnation_window;
    h.center_k = center_k_;
    h.laplacian = F_.sddm ? 0 : 1;
    h.n = static_cast<int>(n_);
    h.prev_check = 1.0;
    h.phase = 1;
}

// ---- autotuner ----------------------------------------------------------

gpu_solver::autotune_result gpu_solver::autotune(int max_block_width, hipStream_t s) const {
    not_serving("autotune");
    auto& d = *dev_;
    autotune_result res;
    const int n = static_cast<int>(n_);
    if (n == 0) return res;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    auto timed = [&](auto&& launch, int reps) {
        launch();   // warm-up
        hip_check(hipEventRecord(e0, s), "hipEventRecord");
        for (int r = 0; r < reps; ++r) launch();
        hip_check(hipEventRecord(e1, s), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return 1e3 * static_cast<double>(ms) / reps;
    };
    try {
        // SpMV lanes (no dot product: no solver state is touched).
        hip_check(hipMemsetAsync(d.p.get(), 0, static_cast<std::size_t>(n) * sizeof(double), s), "memset");
        double best = 1e300;
        for (int k = 0, lanes = 1; lanes <= 64; ++k, lanes <<= 1) {
            res.spmv_us[k] = timed([&] {
                launch_spmv(n, d.op_ptr.as<unsigned>(), d.op_idx.as<int>(), d.op_val.get(), d.op32, lanes,
                            d.p.as<double>(), d.ap.as<double>(), nullptr, nullptr, false, s);
            }, 5);
            if (res.spmv_us[k] < best) { best = res.spmv_us[k]; res.spmv_lanes = lanes; }
        }
        d.lanes = res.spmv_lanes;
        stats_.spmv_lanes = res.spmv_lanes;
        // The sliced-ELL form, when it exists, against the best lane count.
        if (d.has_sell) {
            res.spmv_sell_us = timed([&] {
                launch_spmv_sell(d.sell, d.op32, d.p.as<double>(), d.ap.as<double>(), nullptr, nullptr, false, s);
            }, 5);
            d.use_sell = res.spmv_sell_us < best;
        }
        res.spmv_sell = d.use_sell ? 1 : 0;
        stats_.spmv_format = d.use_sell ? 1 : 0;
        // Block width: one forward + back blocked sweep per candidate.
        double best_rate = 0.0;
        for (int k = 0, kp = 8; kp <= 64 && kp <= std::max(8, max_block_width); ++k, kp <<= 1) {
            prepare_block(kp);
            block_scheds(kp);
            auto& b = d.blk;
            const auto& t = *b.tables.at(kp);
            const std::size_t vb = static_cast<std::size_t>(n) * kp * sizeof(double);
            hip_check(hipMemsetAsync(b.r.get(), 0, vb, s), "memset");
            hip_check(hipMemsetAsync(b.y.get(), 0, vb, s), "memset");
            res.sweep_pair_us[k] = timed([&] {
                launch_sweep_k(d.fwd, d.prec, d.fp16, t.f, b.r.get(), b.y.get(), nullptr, nullptr, s);
                launch_sweep_k(d.bck, d.prec, d.fp16, t.b, b.y.get(), b.z.get(), nullptr, nullptr, s);
            }, 2);
            const double rate = kp / res.sweep_pair_us[k];
            if (rate > best_rate) { best_rate = rate; res.block_width = kp; }
        }
        tuned_block_width_ = res.block_width;
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // Captured graphs hold the old lanes.
    if (d.iter_exec) { (void)hipGraphExecDestroy(d.iter_exec); d.iter_exec = nullptr; }
    if (d.iter_graph) { (void)hipGraphDestroy(d.iter_graph); d.iter_graph = nullptr; }
    d.blk.drop_graph();
    return res;
}

// ---- serving sessions: a running block that admits and releases columns ----

int gpu_solver::session_begin(int slots, hipStream_t s, int which) const {
    auto& d = *dev_;
    auto& b = block_in(*dev_, which);
    const int kp = prepare_block(slots, which);
    block_scheds(kp, which);
    for (int c = 0; c < kp; ++c) {
        std::memset(&b.host_st[c], 0, sizeof(pcg_state));
        b.host_st[c].done = 1;          // empty slot
        b.host_st[c].n = static_cast<int>(n_);
    }
    std::memset(b.host_bs, 0, sizeof(batch_state));
    hip_check(hipMemcpyAsync(b.st.get(), b.host_st, static_cast<std::size_t>(kp) * sizeof(pcg_state),
                             hipMemcpyHostToDevice, s), "state upload");
    hip_check(hipMemcpyAsync(b.bs.get(), b.host_bs, sizeof(batch_state), hipMemcpyHostToDevice, s), "state upload");
    launch_zero_tail_k(d.prec, static_cast<int>(n_), static_cast<int>(T_.m), kp, b.z.get(), s);
    hip_check(hipStreamSynchronize(s), "session begin");
    // The iteration graph, captured now rather than on the first request.
    replay_iterations_k(kp, 0, s, which);
    b.session_kp = kp;
    return kp;
}

void gpu_solver::session_admit(const std::vector<session_column>& cols, hipStream_t s, int which) const {
    auto& d = *dev_;
    auto& b = block_in(*dev_, which);
    const int kp = b.session_kp;
    if (kp == 0) throw std::logic_error("gpu_solver: no session");
    if (cols.empty()) return;
    const int nc = static_cast<int>(cols.size());
    if (nc > kp) throw std::invalid_argument("gpu_solver: more columns than slots");
    // Per-slot states (only the admitted slots' structs are written: the
    // others are live on the device) and the column / source tables.
    int* hc = reinterpret_cast<int*>(b.host_adm);
    const double** hs = reinterpret_cast<const double**>(b.host_adm + 64 * sizeof(int));
    for (int j = 0; j < nc; ++j) {
        const int c = cols[j].slot;
        if (c < 0 || c >= kp) throw std::invalid_argument("gpu_solver: slot out of range");
        fresh_state(b.host_st[c], cols[j].tol, cols[j].max_iter);
        hip_check(hipMemcpyAsync(b.st.as<pcg_state>() + c, &b.host_st[c], sizeof(pcg_state), hipMemcpyHostToDevice, s),
                  "state upload");
        hc[j] = c;
        hs[j] = cols[j].d_b;
    }
    hip_check(hipMemcpyAsync(b.adm_dev.get(), b.host_adm, 64 * sizeof(int) + 64 * sizeof(void*), hipMemcpyHostToDevice, s),
              "admission upload");
    const int* dcols = b.adm_dev.as<int>();
    const double* const* dsrc = reinterpret_cast<const double* const*>(b.adm_dev.as<unsigned char>() + 64 * sizeof(int));
    launch_admit_scatter_k(static_cast<int>(n_), kp, d.perm.as<unsigned>(), nc, dcols, dsrc, b.r.as<double>(),
                           b.x.as<double>(), s);
    launch_init_residual_k(static_cast<int>(n_), kp, b.r.as<double>(), b.st.as<pcg_state>(), b.bs.as<batch_state>(),
                           b.part.as<double>(), s);
    launch_refresh_active_k(kp, b.st.as<pcg_state>(), b.bs.as<batch_state>(), s);
    // The host copies must stay valid until the stream has consumed them.
    hip_check(hipStreamSynchronize(s), "admit");
}

void gpu_solver::session_step(int iterations, hipStream_t s, int which) const {
    replay_iterations_k(block_in(*dev_, which).session_kp, iterations, s, which);
}

void gpu_solver::session_read(std::vector<session_slot>& out, hipStream_t s, int which) const {
    auto& b = block_in(*dev_, which);
    const int kp = b.session_kp;
    hip_check(hipMemcpyAsync(b.host_st, b.st.get(), static_cast<std::size_t>(kp) * sizeof(pcg_state),
                             hipMemcpyDeviceToHost, s), "state read");
    hip_check(hipStreamSynchronize(s), "session read");
    out.resize(static_cast<std::size_t>(kp));
    for (int c = 0; c < kp; ++c) {
        out[c].done = b.host_st[c].done != 0;
        out[c].status = b.host_st[c].status;
        out[c].iterations = b.host_st[c].iter;
        out[c].residual = b.host_st[c].residual;
    }
}

void gpu_solver::session_collect(const std::vector<int>& slots, const std::vector<double*>& d_x, hipStream_t s,
                                 int which) const {
    auto& d = *dev_;
    auto& b = block_in(*dev_, which);
    const int kp = b.session_kp;
    const int nc = static_cast<int>(slots.size());
    if (nc == 0) return;
    if (nc > kp || d_x.size() != slots.size()) throw std::invalid_argument("gpu_solver: collect tables");
    int* hc = reinterpret_cast<int*>(b.host_adm);
    double** hd = reinterpret_cast<double**>(b.host_adm + 64 * sizeof(int));
    for (int j = 0; j < nc; ++j) {
        if (slots[j] < 0 || slots[j] >= kp) throw std::invalid_argument("gpu_solver: slot out of range");
        hc[j] = slots[j];
        hd[j] = d_x[j];
    }
    hip_check(hipMemcpyAsync(b.adm_dev.get(), b.host_adm, 64 * sizeof(int) + 64 * sizeof(void*), hipMemcpyHostToDevice, s),
              "collect upload");
    pcg_state* st = b.st.as<pcg_state>();
    if (!F_.sddm)
        launch_sum_cols_k(vec_prec::fp64, static_cast<int>(n_), kp, b.x.get(), st, b.part.as<double>(), s);
    launch_gather_cols_k(static_cast<int>(n_), kp, d.perm.as<unsigned>(), nc, b.adm_dev.as<int>(),
                         reinterpret_cast<double* const*>(b.adm_dev.as<unsigned char>() + 64 * sizeof(int)),
                         b.x.as<double>(), st, !F_.sddm, s);
    hip_check(hipStreamSynchronize(s), "collect");
}

std::vector<solve_result> gpu_solver::solve_block(int k, const double* d_B, double* d_X, double tol, int max_iter,
                                                  hipStream_t s) const {
    auto& d = *dev_;
    auto& b = d.blk;
    const int n = static_cast<int>(n_);
    const int m = static_cast<int>(T_.m);
    const int kp = prepare_block(k);
    block_scheds(kp);
    b.session_kp = 0;   // a solve ends any serving session on this solver
    // Every column starts joining: the first iteration is its first
    // preconditioner application (SpMM and x / r update skipped).
    for (int c = 0; c < kp; ++c) fresh_state(b.host_st[c], tol, max_iter);
    std::memset(b.host_bs, 0, sizeof(batch_state));
    b.host_bs->active = kp;
    pcg_state* st = b.st.as<pcg_state>();
    batch_state* bs = b.bs.as<batch_state>();
    hip_check(hipMemcpyAsync(st, b.host_st, static_cast<std::size_t>(kp) * sizeof(pcg_state), hipMemcpyHostToDevice, s),
              "state upload");
    hip_check(hipMemcpyAsync(bs, b.host_bs, sizeof(batch_state), hipMemcpyHostToDevice, s), "state upload");
    launch_scatter_init_k(n, k, kp, d.perm.as<unsigned>(), d_B, b.r.as<double>(), b.x.as<double>(), s);
    launch_zero_tail_k(d.prec, n, m, kp, b.z.get(), s);
    launch_init_residual_k(n, kp, b.r.as<double>(), st, bs, b.part.as<double>(), s);
    launch_refresh_active_k(kp, st, bs, s);
    const int J = std::max(1, opts_.iters_per_check);
    int enqueued = 0;
    for (;;) {
        replay_iterations_k(kp, J, s);
        enqueued += J;
        hip_check(hipMemcpyAsync(b.host_bs, bs, sizeof(batch_state), hipMemcpyDeviceToHost, s), "state read");
        hip_check(hipStreamSynchronize(s), "solve");
        if (b.host_bs->active == 0 || enqueued > max_iter + J + 1) break;
    }
    if (!F_.sddm)
        launch_sum_cols_k(vec_prec::fp64, n, kp, b.x.get(), st, b.part.as<double>(), s);
    launch_gather_k(vec_prec::fp64, n, k, kp, d.perm.as<unsigned>(), b.x.get(), d_X, st, !F_.sddm, s);
    hip_check(hipMemcpyAsync(b.host_st, st, static_cast<std::size_t>(kp) * sizeof(pcg_state), hipMemcpyDeviceToHost, s),
              "state read");
    hip_check(hipStreamSynchronize(s), "solve");
    std::vector<solve_result> res(static_cast<std::size_t>(k));
    for (int c = 0; c < k; ++c) {
        res[c].iterations = b.host_st[c].iter;
        res[c].residual = b.host_st[c].residual;
        res[c].status = b.host_st[c].status;
        res[c].lumped_offdiag = F_.lumped_offdiag;
    }
    return res;
}

std::vector<solve_result> gpu_solver::solve_batch_device(int k, const double* d_B, double* d_X, double tol,
                                                         int max_iter, hipStream_t s) const {
    not_serving("solve_batch");
    if (k < 0) throw std::invalid_argument("gpu_solver::solve_batch: negative column count");
    if (tol < 0.0) tol = opts_.tol;
    if (max_iter < 0) max_iter = opts_.max_iter;
    std::vector<solve_result> res;
    if (n_ == 0 || k == 0) { res.resize(static_cast<std::size_t>(k)); return res; }
    for (int c0 = 0; c0 < k; c0 += 64) {
        const int kc = std::min(64, k - c0);
        auto part = solve_block(kc, d_B + static_cast<std::size_t>(c0) * n_, d_X + static_cast<std::size_t>(c0) * n_,
                                tol, max_iter, s);
        res.insert(res.end(), part.begin(), part.end());
    }
    check_sweep_error();
    return res;
}

std::vector<solve_result> gpu_solver::solve_batch(int k, const std::vector<double>& B, std::vector<double>& X,
                                                  double tol, int max_iter) const {
    const std::size_t total = static_cast<std::size_t>(k) * static_cast<std::size_t>(n_);
    if (k < 0 || B.size() != total) throw std::invalid_argument("gpu_solver::solve_batch: B must hold k * n values");
    X.assign(total, 0.0);
    if (total == 0) return solve_batch_device(k, nullptr, nullptr, tol, max_iter, nullptr);
    dbuf db(total * sizeof(double)), dx(total * sizeof(double));
    hip_check(hipMemcpy(db.get(), B.data(), total * sizeof(double), hipMemcpyHostToDevice), "upload B");
    auto res = solve_batch_device(k, db.as<double>(), dx.as<double>(), tol, max_iter, nullptr);
    hip_check(hipMemcpy(X.data(), dx.get(), total * sizeof(double), hipMemcpyDeviceToHost), "download X");
    return res;
}

solve_result gpu_solver::solve(const std::vector<double>& b, double tol, int max_iter,
                               const std::vector<double>* x0) const {
    if (static_cast<std::int64_t>(b.size()) != n_) throw std::invalid_argument("gpu_solver::solve: b length mismatch");
    if (x0 && static_cast<std::int64_t>(x0->size()) != n_)
        throw std::invalid_argument("gpu_solver::solve: x0 length mismatch");
    const std::size_t bytes = static_cast<std::size_t>(n_) * sizeof(double);
    dbuf db(bytes), dx(bytes), dx0(x0 ? bytes : 0);
    hip_check(hipMemcpy(db.get(), b.data(), bytes, hipMemcpyHostToDevice), "upload b");
    if (x0) hip_check(hipMemcpy(dx0.get(), x0->data(), bytes, hipMemcpyHostToDevice), "upload x0");
    solve_result res = solve_device(db.as<double>(), dx.as<double>(), tol, max_iter, x0 ? dx0.as<double>() : nullptr, nullptr);
    res.x.resize(n_);
    hip_check(hipMemcpy(res.x.data(), dx.get(), bytes, hipMemcpyDeviceToHost), "download x");
    return res;
}

void gpu_solver::apply_device(const double* d_r, double* d_z, hipStream_t s) const {
    not_serving("apply");
    auto& d = *dev_;
    const int n = static_cast<int>(n_);
    const int m = static_cast<int>(T_.m);
    if (n == 0) return;
    // The center-k schedule of the reference's _solve_impl
    // (preconditioner.h:344-353, 393-397): every K-th application centres.
    const bool center = !F_.sddm && (++apply_count_ % static_cast<std::uint64_t>(center_k_) == 0);
    pcg_state* st = d.st_apply.as<pcg_state>();
    double* part = d.part.as<double>();
    int* err = d.err.as<int>();
    launch_sum_shift(n, d_r, st, part, center, s);
    launch_apply_scatter(d.prec, n, m, d.perm.as<unsigned>(), d_r, d.rp.as<double>(), d.y.get(), d.z.get(), s);
    launch_sweep(d.fwd, d.prec, d.fp16, d.rp.get(), d.y.get(), st, err, d.f_sched, s);
    launch_sweep(d.bck, d.prec, d.fp16, d.y.get(), d.z.get(), st, err, d.b_sched, s);
    if (center) launch_sum_out(d.prec, n, d.z.get(), st, part, s);
    launch_apply_gather(d.prec, n, d.perm.as<unsigned>(), d.z.get(), d_z, st, center, s);
}

std::vector<double> gpu_solver::apply(const std::vector<double>& r) const {
    if (static_cast<std::int64_t>(r.size()) != n_) throw std::invalid_argument("gpu_solver::apply: r length mismatch");
    const std::size_t bytes = static_cast<std::size_t>(n_) * sizeof(double);
    dbuf dr(bytes), dz(bytes);
    hip_check(hipMemcpy(dr.get(), r.data(), bytes, hipMemcpyHostToDevice), "upload r");
    apply_device(dr.as<double>(), dz.as<double>(), nullptr);
    std::vector<double> z(n_);
    hip_check(hipMemcpy(z.data(), dz.get(), bytes, hipMemcpyDeviceToHost), "download z");
    check_sweep_error();
    return z;
}

void gpu_solver::sweep_device(bool backward, const double* d_rhs, double* d_out, hipStream_t s) const {
    not_serving("sweep");
    auto& d = *dev_;
    const int m = static_cast<int>(T_.m);
    if (m == 0) return;
    if (relabel_.empty()) {
        launch_fill_pending(vec_prec::fp64, d_out, m, s);
        launch_sweep(backward ? d.bck : d.fwd, vec_prec::fp64, d.fp16, d_rhs, d_out, nullptr, d.err.as<int>(),
                     backward ? d.b_sched : d.f_sched, s);
        return;
    }
    // Factor order at the boundary, storage order inside.
    const unsigned* Q = d.relabel.as<unsigned>();
    launch_permute(m, Q, d_rhs, d.rp.as<double>(), true, s);
    launch_fill_pending(vec_prec::fp64, d.xout.get(), m, s);
    launch_sweep(backward ? d.bck : d.fwd, vec_prec::fp64, d.fp16, d.rp.get(), d.xout.get(), nullptr, d.err.as<int>(),
                 backward ? d.b_sched : d.f_sched, s);
    launch_permute(m, Q, d.xout.as<double>(), d_out, false, s);
}

void gpu_solver::solve_LLt_device(const double* d_x, double* d_y, hipStream_t s) const {
    not_serving("solve_LLt");
    auto& d = *dev_;
    const int m = static_cast<int>(T_.m);
    if (m == 0) return;
    // cuda.h:736-785: cast in, forward sweep, back sweep, cast out. The casts
    // are the permutation kernels here (factor order at the boundary, storage
    // order inside); the forward sweep reads fp64 and writes the solver's
    // vector type, so no separate cast pass runs.
    const unsigned* Q = relabel_.empty() ? nullptr : d.relabel.as<unsigned>();
    launch_permute(m, Q, d_x, d.rp.as<double>(), true, s);
    launch_fill_pending(d.prec, d.y.get(), m, s);
    launch_fill_pending(d.prec, d.z.get(), m, s);
    launch_sweep(d.fwd, d.prec, d.fp16, d.rp.get(), d.y.get(), nullptr, d.err.as<int>(), d.f_sched, s);
    launch_sweep(d.bck, d.prec, d.fp16, d.y.get(), d.z.get(), nullptr, d.err.as<int>(), d.b_sched, s);
    launch_permute_out(d.prec, m, Q, d.z.get(), d_y, s);
}

group_member_view gpu_solver::group_view() const {
    not_serving("group");
    const auto& d = *dev_;
    group_member_view v;
    auto& g = v.mem;
    g.n = static_cast<int>(n_);
    g.m = static_cast<int>(T_.m);
    g.r = d.r.as<double>();
    g.x = d.x.as<double>();
    g.p = d.p.as<double>();
    g.ap = d.ap.as<double>();
    g.y = d.y.get();
    g.z = d.z.get();
    g.part = d.part.as<double>();
    g.use_sell = d.use_sell ? 1 : 0;
    g.val32 = d.op32 ? 1 : 0;
    g.lanes = d.lanes;
    g.sell = d.sell;
    g.optr = d.op_ptr.as<unsigned>();
    g.oidx = d.op_idx.as<int>();
    g.oval = d.op_val.get();
    v.perm = d.perm.as<unsigned>();
    v.sddm = F_.sddm;
    v.prec = d.prec;
    v.fp16 = d.fp16;
    v.stagnation_window = opts_.stagnation_window;
    v.center_k = center_k_;
    v.lumped_offdiag = F_.lumped_offdiag;
    const int m = static_cast<int>(T_.m);
    const std::size_t vsz = d.prec == vec_prec::fp64 ? sizeof(double) : sizeof(float);
    for (int dir = 0; dir < 2; ++dir) {
        const sweep_sched& sc = dir ? d.b_sched : d.f_sched;
        const sweep_desc& desc = dir ? d.bck : d.fwd;
        if (desc.nb == 0) continue;
        if (sc.kind != 2) throw std::invalid_argument("gpu_solver::group: a group member's sweeps must be the hybrid schedule");
        for (int k = 0; k < sc.nseg; ++k) {
            if (sc.seg_tail[k]) {
                group_unit u;
                u.kind = kGroupTail;
                u.d = desc;
                u.b0 = sc.level_first[sc.seg_l0[k]];
                u.b1 = sc.level_first[sc.seg_l1[k]];
                u.p0 = sc.seg_p0[k];
                u.td = sc.tail;
                u.base = sc.seg_base[k];
                u.nlev = sc.seg_phases[k];
                u.zslot = sc.seg_p1[k] - sc.seg_p0[k];
                u.row0 = dir ? m - 1 - u.p0 : u.p0;
                u.tv = static_cast<unsigned char*>(sc.tail_val) + static_cast<std::size_t>(u.base) * kWave * 2 * vsz;
                u.rhs = dir ? d.y.get() : d.r.get();
                u.out = dir ? d.z.get() : d.y.get();
                u.err = d.err.as<int>();
                u.grid = u.b1 - u.b0;
                u.lds = sweep_tail_lds(d.prec, u.zslot, u.b1 - u.b0);
                if (u.b1 <= u.b0 || u.base < 0 || u.base + (u.b1 - u.b0) > sc.tail_batches)
                    throw std::runtime_error("gpu_solver::group: tail run outside the tail layout");
                v.steps[dir].push_back(u);
                continue;
            }
            for (int l = sc.seg_l0[k]; l < sc.seg_l1[k]; ++l) {
                group_unit u;
                u.d = desc;
                u.b0 = sc.level_first[l];
                u.b1 = sc.level_first[l + 1];
                if (u.b1 <= u.b0) continue;
                u.rhs = dir ? d.y.get() : d.r.get();
                u.out = dir ? d.z.get() : d.y.get();
                u.err = d.err.as<int>();
                const int lb = sc.level_lb ? sc.level_lb[l] : -1;
                if (lb >= 0) {
                    u.kind = kGroupLane;
                    u.lb0 = lb;
                    u.nlb = u.b1 - u.b0;
                    u.rec = sc.lane_rec_dev;
                    u.ent = sc.lane_ent_dev;
                    if (static_cast<long long>(lb) + u.nlb > sc.lane_batches)
                        throw std::runtime_error("gpu_solver::group: level outside the lane layout");
                    u.grid = (u.nlb + kLevelWavesPerBlock - 1) / kLevelWavesPerBlock;
                } else {
                    u.kind = kGroupLevels;
                    u.grid = std::min((u.b1 - u.b0 + kLevelWavesPerBlock - 1) / kLevelWavesPerBlock, kLevelLaunchMaxBlocks);
                }
                v.steps[dir].push_back(u);
            }
        }
    }
    return v;
}

void gpu_solver::spmv_device(const double* d_x, double* d_y, hipStream_t s) const {
    not_serving("spmv");
    auto& d = *dev_;
    const int n = static_cast<int>(n_);
    if (n == 0) return;
    launch_scatter_init(n, d.perm.as<unsigned>(), d_x, nullptr, d.rp.as<double>(), d.xout.as<double>(), s);
    spmv(d.rp.as<double>(), d.ap.as<double>(), nullptr, false, s, {});
    launch_gather_shift(n, d.perm.as<unsigned>(), d.ap.as<double>(), d_y, nullptr, false, s);
}

solve_result solve(const csc_view& L, const std::vector<double>& b, const solve_options& opts) {
    const gpu_solver slv(L, opts);
    return slv.solve(b);
}

}  // namespace apxchol_amd
