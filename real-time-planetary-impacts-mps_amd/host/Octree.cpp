/*
 * Octree.cpp -- CPU octree builder producing the reference's flat layout
 * (octree.cpp:14-141):
 *   branch record  [0, nodeValue, child0 .. child7]   (child = record offset, -1 if empty)
 *   leaf record    [n, nodeValue, p0 .. p_{n-1}]
 * in depth-first pre-order, nodeValue = pre-order node number, plus one list
 * of record offsets per level in the same order.  A node is a leaf when it
 * holds <= max_leaf particles or sits at depth max_depth; otherwise its
 * particles are split at their mean position with a strict `>` per axis,
 * octant = (x>cx) | (y>cy)<<1 | (z>cz)<<2, each octant keeping the parent's
 * order (octree.cpp:68-86).
 *
 * Unlike the reference, which copies a std::vector per node, this builder
 * partitions one permuted SoA copy of (x, y, z, id) in place (a stable
 * counting scatter per node), and can build the subtrees below a split depth
 * on several threads.  The pieces are then stitched back in pre-order, so the
 * output is identical for every thread count.
 */
#include <cmath>
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "sph_host.h"

namespace {

struct Work {
    std::vector<float> x, y, z, tx, ty, tz;
    std::vector<int32_t> id, tid;
    std::vector<uint8_t> oct;
};

struct Piece {
    /* A piece is either one branch record of the top of the tree or a whole
     * subtree built by build_serial. */
    bool top_branch = false;
    int b = 0, e = 0, level = 0;
    int child_piece[8];
    std::vector<int32_t> data;                 /* local offsets / node numbers */
    std::vector<std::vector<int32_t>> levels;  /* absolute level index */
    int32_t nodes = 0;
    int64_t off = 0;
    int32_t node_base = 0;
};

struct Builder {
    int max_leaf, max_depth;
    Work w;

    /* Mean-point split of [b, e): returns per-octant counts and reorders the
     * range in place, keeping the order inside each octant. */
    void split(int b, int e, int cnt[8]) {
        float cx = 0, cy = 0, cz = 0;
        for (int i = b; i < e; ++i) {
            cx += w.x[i];
            cy += w.y[i];
            cz += w.z[i];
        }
        const float nf = (float)(e - b);
        cx /= nf;
        cy /= nf;
        cz /= nf;
        for (int k = 0; k < 8; ++k) cnt[k] = 0;
        for (int i = b; i < e; ++i) {
            const int o = (w.x[i] > cx ? 1 : 0) | (w.y[i] > cy ? 2 : 0) | (w.z[i] > cz ? 4 : 0);
            w.oct[i] = (uint8_t)o;
            cnt[o]++;
        }
        int pos[8];
        int acc = b;
        for (int k = 0; k < 8; ++k) {
            pos[k] = acc;
            acc += cnt[k];
        }
        for (int i = b; i < e; ++i) {
            const int d = pos[w.oct[i]]++;
            w.tx[d] = w.x[i];
            w.ty[d] = w.y[i];
            w.tz[d] = w.z[i];
            w.tid[d] = w.id[i];
        }
        std::memcpy(&w.x[b], &w.tx[b], sizeof(float) * (e - b));
        std::memcpy(&w.y[b], &w.ty[b], sizeof(float) * (e - b));
        std::memcpy(&w.z[b], &w.tz[b], sizeof(float) * (e - b));
        std::memcpy(&w.id[b], &w.tid[b], sizeof(int32_t) * (e - b));
    }

    bool is_leaf(int n, int level) const { return n <= max_leaf || level == max_depth; }

    int32_t build_serial(int b, int e, int level, Piece& out) {
        const int n = e - b;
        if ((int)out.levels.size() <= level) out.levels.resize(level + 1);
        const int32_t start = (int32_t)out.data.size();
        out.levels[level].push_back(start);
        if (is_leaf(n, level)) {
            out.data.push_back(n);
            out.data.push_back(out.nodes++);
            out.data.insert(out.data.end(), w.id.begin() + b, w.id.begin() + e);
            return start;
        }
        out.data.resize(start + 10, -1);
        out.data[start] = 0;
        out.data[start + 1] = out.nodes++;
        int cnt[8];
        split(b, e, cnt);
        int cb = b;
        for (int c = 0; c < 8; ++c) {
            if (cnt[c] > 0) {
                const int32_t child = build_serial(cb, cb + cnt[c], level + 1, out);
                out.data[start + 2 + c] = child;
            }
            cb += cnt[c];
        }
        return start;
    }

    /* Serial top of the tree down to `split_level`; everything at or below it
     * becomes a subtree piece. */
    int top(int b, int e, int level, int split_level, std::vector<Piece>& pieces) {
        const int idx = (int)pieces.size();
        pieces.emplace_back();
        pieces[idx].b = b;
        pieces[idx].e = e;
        pieces[idx].level = level;
        if (level >= split_level || is_leaf(e - b, level)) return idx;
        pieces[idx].top_branch = true;
        for (int k = 0; k < 8; ++k) pieces[idx].child_piece[k] = -1;
        int cnt[8];
        split(b, e, cnt);
        int cb = b;
        for (int c = 0; c < 8; ++c) {
            if (cnt[c] > 0) {
                const int ch = top(cb, cb + cnt[c], level + 1, split_level, pieces);
                pieces[idx].child_piece[c] = ch;
            }
            cb += cnt[c];
        }
        return idx;
    }
};

}  // namespace

extern "C" {

int sph_octree_build(const float* pos4, int32_t n, int32_t max_leaf, int32_t max_depth, int32_t nthreads,
                     sph_octree* out) {
    if (!out) return SPH_E_ARG;
    std::memset(out, 0, sizeof(*out));
    if (n <= 0) return SPH_OK;
    if (!pos4 || max_leaf < 1 || max_depth < 0) return SPH_E_ARG;
    Builder bld;
    bld.max_leaf = max_leaf;
    bld.max_depth = max_depth;
    Work& w = bld.w;
    w.x.resize(n); w.y.resize(n); w.z.resize(n);
    w.tx.resize(n); w.ty.resize(n); w.tz.resize(n);
    w.id.resize(n); w.tid.resize(n); w.oct.resize(n);
    bool finite = true;
    for (int i = 0; i < n; ++i) {
        w.x[i] = pos4[4 * i];
        w.y[i] = pos4[4 * i + 1];
        w.z[i] = pos4[4 * i + 2];
        w.id[i] = i;
        finite = finite && std::isfinite(w.x[i]) && std::isfinite(w.y[i]) && std::isfinite(w.z[i]);
    }
    /* a non-finite coordinate makes every mean NaN, every `>` false and the
     * whole set one depth-limited leaf, whose P2P is quadratic in n */
    if (!finite) return SPH_E_ARG;
    const int threads = nthreads < 1 ? 1 : nthreads;
    const int split_level = threads <= 1 ? 0 : (threads <= 8 ? 2 : 3);
    std::vector<Piece> pieces;
    bld.top(0, n, 0, split_level, pieces);

    /* subtree pieces, in parallel; each touches only its own index range */
    std::vector<int> jobs;
    for (int i = 0; i < (int)pieces.size(); ++i)
        if (!pieces[i].top_branch) jobs.push_back(i);
    /* largest first for balance */
    std::sort(jobs.begin(), jobs.end(), [&](int a, int b) {
        return (pieces[a].e - pieces[a].b) > (pieces[b].e - pieces[b].b);
    });
    std::vector<std::thread> pool;
    std::atomic_int next{0};
    auto worker = [&]() {
        for (;;) {
            const int k = next.fetch_add(1);
            if (k >= (int)jobs.size()) break;
            Piece& p = pieces[jobs[k]];
            bld.build_serial(p.b, p.e, p.level, p);
        }
    };
    const int nt = std::min<int>(threads, (int)jobs.size());
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();

    /* pre-order offsets and node numbers */
    int64_t off = 0;
    int32_t nodes = 0;
    int n_levels = 0;
    for (auto& p : pieces) {
        p.off = off;
        p.node_base = nodes;
        if (p.top_branch) {
            off += 10;
            nodes += 1;
            n_levels = std::max(n_levels, p.level + 1);
        } else {
            off += (int64_t)p.data.size();
            nodes += p.nodes;
            n_levels = std::max(n_levels, (int)p.levels.size());
        }
    }
    out->n_ints = off;
    out->node_count = nodes;
    out->n_levels = n_levels;
    out->octree = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)off);
    out->level_counts = (int32_t*)std::calloc((size_t)n_levels, sizeof(int32_t));
    out->level_nodes = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)nodes);
    if (!out->octree || !out->level_counts || !out->level_nodes) {
        sph_octree_free(out);
        return SPH_E_CAPACITY;
    }
    std::vector<std::vector<int32_t>> levels(n_levels);
    for (auto& p : pieces) {
        int32_t* dst = out->octree + p.off;
        if (p.top_branch) {
            dst[0] = 0;
            dst[1] = p.node_base;
            for (int c = 0; c < 8; ++c) dst[2 + c] = p.child_piece[c] < 0 ? -1 : (int32_t)pieces[p.child_piece[c]].off;
            levels[p.level].push_back((int32_t)p.off);
            continue;
        }
        const int64_t sz = (int64_t)p.data.size();
        for (int64_t k = 0; k < sz;) {
            const int32_t cnt = p.data[k];
            dst[k] = cnt;
            dst[k + 1] = p.data[k + 1] + p.node_base;
            if (cnt == 0) {
                for (int c = 0; c < 8; ++c) {
                    const int32_t ch = p.data[k + 2 + c];
                    dst[k + 2 + c] = ch < 0 ? -1 : (int32_t)(ch + p.off);
                }
                k += 10;
            } else {
                std::memcpy(dst + k + 2, &p.data[k + 2], sizeof(int32_t) * cnt);
                k += 2 + cnt;
            }
        }
        for (size_t l = 0; l < p.levels.size(); ++l)
            for (int32_t o : p.levels[l]) levels[l].push_back((int32_t)(o + p.off));
    }
    int64_t lo = 0;
    for (int l = 0; l < n_levels; ++l) {
        out->level_counts[l] = (int32_t)levels[l].size();
        std::memcpy(out->level_nodes + lo, levels[l].data(), sizeof(int32_t) * levels[l].size());
        lo += (int64_t)levels[l].size();
    }
    return SPH_OK;
}

void sph_octree_free(sph_octree* t) {
    if (!t) return;
    std::free(t->octree);
    std::free(t->level_nodes);
    std::free(t->level_counts);
    std::memset(t, 0, sizeof(*t));
}

}  // extern "C"
