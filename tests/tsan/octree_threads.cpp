/*
 * Driver for the ThreadSanitizer run of the host octree builder
 * (host/Octree.cpp, sph_octree_build): builds the tree of a seeded
 * two-clump cloud with 1 and with 16 threads and checks that the outputs are
 * identical.  Compiled with g++ -fsanitize=thread by tests/test_host.py; a
 * data race makes ThreadSanitizer report and exit non-zero.
 */
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "sph_host.h"

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 100000;
    std::mt19937_64 rng(7);
    std::normal_distribution<float> g(0.f, 1.f);
    std::vector<float> pos(4 * (size_t)n, 0.f);
    for (int i = 0; i < n; ++i) {
        const float c = (i % 9 == 0) ? 6.f : 0.f;
        for (int a = 0; a < 3; ++a) pos[4 * (size_t)i + a] = c + g(rng) * (c > 0 ? 0.5f : 2.f);
    }
    sph_octree one, many;
    if (sph_octree_build(pos.data(), n, 4, 7, 1, &one) != 0) return 2;
    for (int rep = 0; rep < 3; ++rep) {
        if (sph_octree_build(pos.data(), n, 4, 7, 16, &many) != 0) return 3;
        const bool same = one.n_ints == many.n_ints && one.node_count == many.node_count &&
                          one.n_levels == many.n_levels &&
                          std::memcmp(one.octree, many.octree, 4 * (size_t)one.n_ints) == 0 &&
                          std::memcmp(one.level_nodes, many.level_nodes, 4 * (size_t)one.node_count) == 0 &&
                          std::memcmp(one.level_counts, many.level_counts, 4 * (size_t)one.n_levels) == 0;
        sph_octree_free(&many);
        if (!same) {
            std::printf("16-thread tree differs from the 1-thread tree\n");
            return 4;
        }
    }
    std::printf("ok n=%d nodes=%d\n", n, one.node_count);
    sph_octree_free(&one);
    return 0;
}
