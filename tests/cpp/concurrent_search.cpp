// FAST-LIO's calling pattern on one tree: many host threads issue
// Nearest_Search at once (an OpenMP loop over scan points; the reference
// allows concurrent searches, ikd_Tree.cpp:366-397).  Every thread's results
// must equal the single-threaded ones bit for bit.  Prints "ok" on success.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "ikd_Tree_amd.hpp"

struct Pt {
    float x, y, z;
};

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 20000;
    const int nq = argc > 2 ? std::atoi(argv[2]) : 800;
    const int threads = argc > 3 ? std::atoi(argv[3]) : 8;
    std::mt19937 gen(7);
    std::uniform_real_distribution<float> u(-50.0f, 50.0f);
    KD_TREE<Pt>::PointVector cloud(n), queries(nq);
    for (auto& p : cloud) p = {u(gen), u(gen), u(gen)};
    for (auto& p : queries) p = {u(gen), u(gen), u(gen)};
    KD_TREE<Pt> tree(0.5, 0.6, 0.2);
    tree.Build(cloud);
    const int k = 5;
    std::vector<float> ref(nq * k * 4, -1.0f), got(threads * (size_t)nq * k * 4, -1.0f);
    auto run = [&](int q, float* out) {
        KD_TREE<Pt>::PointVector res;
        std::vector<float> dist;
        tree.Nearest_Search(queries[q], k, res, dist, 20.0);
        for (size_t j = 0; j < res.size(); ++j) {
            out[4 * j] = res[j].x;
            out[4 * j + 1] = res[j].y;
            out[4 * j + 2] = res[j].z;
            out[4 * j + 3] = dist[j];
        }
    };
    for (int q = 0; q < nq; ++q) run(q, &ref[(size_t)q * k * 4]);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            // each thread walks all queries, starting at a different offset
            for (int i = 0; i < nq; ++i) {
                const int q = (i + t * nq / threads) % nq;
                run(q, &got[((size_t)t * nq + q) * k * 4]);
            }
        });
    for (auto& th : pool) th.join();
    for (int t = 0; t < threads; ++t)
        if (std::memcmp(&got[(size_t)t * nq * k * 4], ref.data(), ref.size() * sizeof(float)) != 0) {
            std::printf("thread %d: results differ from the single-threaded run\n", t);
            return 1;
        }
    std::printf("ok %d threads x %d queries\n", threads, nq);
    return 0;
}
