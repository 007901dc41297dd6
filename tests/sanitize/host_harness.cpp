// Host-code harness for the sanitizer builds (AddressSanitizer +
// UndefinedBehaviorSanitizer, ThreadSanitizer) run by
// tests/test_cpu_host_sanitizers.py: the factorization (one thread and a
// team of four), the SpTRSV preparation (fp32 / fp16 storage, with and
// without the compacting drop), both level relabelings, the permuted
// operator and its sliced-ELL forms, on a 3-D grid (SDDM and Laplacian), a
// random multigraph, a hub with a matching, paths, one edge and one vertex;
// then the MatrixMarket reader, input classification, adjacency conversion
// and a write / read round trip on the fixture files named by argv[1].
// Prints one line per case; the sanitizers abort on the first finding.
#include "host/factor.hpp"
#include "host/trsv_prep.hpp"
#include "host/mtx.hpp"
#include <cstdio>
#include <random>
#include <map>
#include <filesystem>
using namespace apxchol_amd;
struct coo { long n; std::map<std::pair<long,long>, double> e; };
static void add_edge(coo& g, long i, long j, double w) { if (i == j) return; g.e[{i,j}] -= w; g.e[{j,i}] -= w; g.e[{i,i}] += w; g.e[{j,j}] += w; }
static csc_matrix to_csc(const coo& g, double excess) {
    csc_matrix A; A.rows = A.cols = g.n; A.colptr.assign(g.n + 1, 0);
    std::vector<std::vector<std::pair<int,double>>> cols(g.n);
    for (auto& [k, v] : g.e) cols[k.second].push_back({(int)k.first, v + (k.first == k.second ? excess : 0.0)});
    for (long c = 0; c < g.n; ++c) { if (cols[c].empty() || cols[c][0].first > c) {} std::sort(cols[c].begin(), cols[c].end()); }
    for (long c = 0; c < g.n; ++c) { for (auto& [r, v] : cols[c]) { A.rowidx.push_back(r); A.vals.push_back(v); } A.colptr[c + 1] = A.rowidx.size(); }
    return A;
}
static void run(const char* name, const csc_matrix& A) {
    csc_view v = A.view();
    auto chk = check_operator(v);
    for (unsigned threads : {1u, 4u}) {
        factor_options fo; fo.seed = 7; fo.threads = (int)threads;
        factorization F;
        try { F = factorize(v, fo); } catch (const std::exception& e) { printf("%s: factorize threw: %s\n", name, e.what()); return; }
        for (bool fp16 : {false, true}) for (double rel : {-1.0, 1e-3}) {
            trsv_factor T = prepare_trsv(F, fp16, rel);
            for (bool depth : {true, false}) {
                trsv_factor T2 = T;
                auto Q = relabel_within_levels(T2, F.perm, depth);
                (void)Q;
            }
        }
        auto op = build_permuted_operator(v, F.perm);
        std::vector<std::int32_t> rows(op.n); for (long i = 0; i < op.n; ++i) rows[i] = i;
        (void)sell_slots(op, 64, rows);
        for (long sigma : {1L, 4096L}) { auto sl = build_sell(op, 64, sigma, rows); (void)sl; }
    }
    printf("%s: n=%ld ok\n", name, (long)A.cols);
}
int main(int argc, char** argv) {
    std::mt19937_64 rng(1);
    { coo g{20*20*20}; auto id = [](int x,int y,int z){return (long)x+20*(y+20L*z);};
      std::uniform_real_distribution<double> U(0.5, 2.0);
      for (int z=0;z<20;++z) for(int y=0;y<20;++y) for(int x=0;x<20;++x){ if(x<19) add_edge(g,id(x,y,z),id(x+1,y,z),U(rng)); if(y<19) add_edge(g,id(x,y,z),id(x,y+1,z),U(rng)); if(z<19) add_edge(g,id(x,y,z),id(x,y,z+1),U(rng)); }
      run("grid3d_20 sddm", to_csc(g, 1e-3)); run("grid3d_20 laplacian", to_csc(g, 0.0)); }
    { coo g{4096}; std::uniform_int_distribution<long> V(0, 4095);
      for (int e = 0; e < 40000; ++e) { long a = V(rng), b = V(rng); long m = std::min(a,b); add_edge(g, m, (a*b) % 4096, 1.0 + (e % 9)); }
      run("random 4096", to_csc(g, 0.0)); }
    { coo g{3000}; for (long i = 1; i < 3000; ++i) add_edge(g, 0, i, 1.0); for (long i = 1; i + 1 < 3000; i += 2) add_edge(g, i, i + 1, 2.0); run("star+matching", to_csc(g, 0.0)); }
    { coo g{500}; for (long i = 0; i + 1 < 500; ++i) add_edge(g, i, i + 1, 1.0 + i % 3); run("path", to_csc(g, 0.0)); run("path sddm", to_csc(g, 0.5)); }
    { coo g{2}; add_edge(g, 0, 1, 1.0); run("edge", to_csc(g, 0.0)); }
    { coo g{1}; g.e[{0,0}] = 3.0; run("single sddm", to_csc(g, 0.0)); }
    for (auto& p : std::filesystem::directory_iterator(argc > 1 ? argv[1] : "tests/golden/fixtures")) {
        if (p.path().extension() != ".mtx") continue;
        try {
            mtx_header h; csc_matrix A = read_mtx(p.path().string(), &h);
            csc_view v = A.view();
            auto s = scan_input(v);
            std::string why;
            for (int k = 0; k < 3; ++k) { auto kind = resolve_input_kind((input_kind)k, s, false, why); (void)describe_input(kind, s, why); (void)diagnose_failure(kind, s); }
            csc_matrix L = adjacency_to_laplacian(v);
            const std::string rt = std::string(argc > 2 ? argv[2] : ".") + "/roundtrip.mtx";
            write_mtx(rt, L, true); csc_matrix B = read_mtx(rt);
            printf("mtx %s: n=%ld nnz=%zu rt_nnz=%zu\n", p.path().filename().c_str(), (long)A.cols, A.rowidx.size(), B.rowidx.size());
        } catch (const std::exception& e) { printf("mtx %s: threw %s\n", p.path().filename().c_str(), e.what()); }
    }
    return 0;
}
