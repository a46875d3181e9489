/*
 * gravity.hip -- FMM self-gravity on the CPU-built octree, for gfx950.
 *
 * Reference: Compute::gravitationalPass (Compute.cpp:328-377), upPass /
 * downPass (gravity.metal:36-143), helpers.h:15-147 and the poles/ kernels.
 *
 * The uploaded tree keeps the reference's flat layout (octree.cpp); a
 * conversion kernel turns it into per-node tables indexed by node id (the
 * pre-order node number the reference calls dataPointer): kind/particle
 * range, 8 child ids, and the slot of the parent in its level list
 * (parentIndexes, which the reference rewrites in every up pass, M2M.metal:105).
 * Unchecked-node lists hold node ids instead of tree offsets; a list slot is
 * `slot in level x max_unchecked_pointers`, -1 terminated, ping-ponged per
 * level as in the reference.
 *
 * Up pass: one launch per level, deepest first, one thread per node (leaf =
 * P2M over its particles, branch = M2M over its children).
 * Down pass: one launch per level, root first, one thread per node: scan the
 * parent's list with the multipole acceptance criterion, M2L the accepted
 * nodes, push the rejected branch children (or force-accept them once the
 * list is full), then L2L to the active children or, at a leaf, L2P plus the
 * intra-leaf P2P.
 */
#include <cfloat>

#include "sph_internal.h"

namespace {

/* poles.h:18-39 index map; the third-order terms are used at P = 3 */
enum {
    eM = 0, eX, eY, eZ, eXX, eXY, eXZ, eYY, eYZ, eZZ, eNT,
    eXXX = eNT, eXXY, eXXZ, eXYY, eXYZ, eXZZ, eYYY, eYYZ, eYZZ, eZZZ, eNT3
};
template <int P>
constexpr int nterms() { return P == 3 ? (int)eNT3 : (int)eNT; }

struct Tree {
    const int32_t* octree;
    const int4* info;        /* (nParticles, treePointer, level, 0) */
    const int32_t* child;
    const int32_t* parent_slot;
    int32_t* sub_end;        /* node id + pre-order subtree size */
    float4* far;             /* 3 per node */
    float4* aux;             /* 3 per node */
    float4* loc;             /* 4 per node */
    float4* oct;             /* 3 per node (P = 3): (XXX XXY XXZ XYY) (XYZ XZZ YYY YYZ) (YZZ ZZZ power3 -) */
    float4* loc3;            /* 3 per node (P = 3): third-order local terms, same order */
    float4* loc_tot;         /* 7 per node: a near-field leaf's complete local expansion (loc, then loc3 layout) */
    /* near-field leaves in two launches: per node, the particles of the
     * rejected source leaves the M2L launch's walk found, in walk order
     * (near_cap per node), and their count (-1: not recorded, the pair
     * launch walks again) */
    int32_t* near_rec;
    int32_t* near_cnt;
    int near_cap;
    /* precision form: each node's centre is far[3 id].xyz + clo[id].xyz, the
     * second part the rounding residual of the first (null in other forms) */
    float4* clo;
};

/* ------------------------------------------------------------ conversion */
__global__ void k_fill_i32(int32_t* __restrict__ a, int64_t count, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = v;
}

struct ConvertArgs {
    const int32_t* octree;
    int64_t n_ints;
    const int32_t* level_tp; /* all level lists, level by level */
    int32_t* level_id;
    /* level lists first_level .. first_level + chunk_levels - 1 of this
     * launch: slot s belongs to the level l with off[l] <= s < off[l + 1]
     * (l relative to first_level) */
    int32_t off[kMaxLevels + 1];
    int first_level;
    int chunk_levels;
    int n_levels;
    int node_count;
    int n_particles;
    int4* info;
    int32_t* child;
    int32_t* parent_slot;
    int32_t* parent_level;
    int32_t* seen;
    uint32_t* flags; /* [0] != 0: the uploaded tree is malformed, gravity is skipped */
};

/* Clears what k_tree_convert counts and links, and the malformed-tree flag. */
__global__ void k_tree_convert_init(int32_t* __restrict__ parent_slot, int32_t* __restrict__ parent_level,
                                    int32_t* __restrict__ seen, int count, uint32_t* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    parent_slot[i] = -1;
    parent_level[i] = -1;
    seen[i] = 0;
    if (i == 0) flags[0] = 0u;
}

/* One thread per level slot, every level in one launch: node id, kind,
 * children and the parent's slot within its level (parentIndexes), with
 * every index checked against its buffer so that a malformed tree disables
 * gravity instead of reading out of bounds.  The levels are independent: a
 * slot writes only its own node's record and its children's parent links. */
__global__ void k_tree_convert(ConvertArgs a) {
    const int s = a.off[0] + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.off[a.chunk_levels]) return;
    int l = 0;
    while (l + 1 < a.chunk_levels && a.off[l + 1] <= s) ++l;
    const int level = a.first_level + l;
    const int tp = a.level_tp[s];
    bool bad = tp < 0 || (int64_t)tp + 2 > a.n_ints;
    const int np = bad ? 0 : a.octree[tp];
    const int id = bad ? -1 : a.octree[tp + 1];
    bad = bad || id < 0 || id >= a.node_count || np < 0;
    bad = bad || (np > 0 && (int64_t)tp + 2 + np > a.n_ints) || (np == 0 && (int64_t)tp + 10 > a.n_ints);
    bad = bad || (np == 0 && level == a.n_levels - 1) || (level == 0 && (tp != 0 || id != 0));
    if (bad) {
        a.flags[0] = 1u;
        a.level_id[s] = 0;
        return;
    }
    a.level_id[s] = id;
    a.info[id] = make_int4(np, tp, level, 0);
    atomicAdd(&a.seen[id], 1);
    if (np > 0) {
        for (int k = 0; k < np; ++k) {
            const int p = a.octree[tp + 2 + k];
            if (p < 0 || p >= a.n_particles) a.flags[0] = 1u;
        }
        return;
    }
    for (int c = 0; c < 8; ++c) {
        const int cp = a.octree[tp + 2 + c];
        int cid = -1;
        if (cp != -1) {
            if (cp < 0 || (int64_t)cp + 2 > a.n_ints) {
                a.flags[0] = 1u;
            } else {
                cid = a.octree[cp + 1];
                if (cid < 0 || cid >= a.node_count) {
                    a.flags[0] = 1u;
                    cid = -1;
                }
            }
        }
        a.child[8 * id + c] = cid;
        if (cid >= 0) {
            a.parent_slot[cid] = s - a.off[l];
            a.parent_level[cid] = level;
        }
    }
}

/* Every node id appears once in the level lists, and every non-root node is
 * the child of a node one level up. */
__global__ void k_tree_check(int node_count, const int4* __restrict__ info, const int32_t* __restrict__ parent_slot,
                             const int32_t* __restrict__ parent_level, const int32_t* __restrict__ seen,
                             uint32_t* __restrict__ flags) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= node_count) return;
    if (seen[id] != 1) {
        flags[0] = 1u;
        return;
    }
    const int level = info[id].z;
    if (level == 0 ? id != 0 : (parent_slot[id] < 0 || parent_level[id] != level - 1)) flags[0] = 1u;
}

/* ---------------------------------------------------------------- up pass */
struct UpArgs {
    Tree t;
    uint32_t* flags;
    const int32_t* level_id;
    int count;
    const float4* pos;
    const float* m;
    const int32_t* next_active;
    const float* gravnorm;
    uint8_t* active;
    const int32_t* time;
    int readme_rule;
    int fix_zz;
    int fix_range;
    int need_preorder; /* the near-field down pass walks subtrees as id ranges */
    int prec;          /* precision form: centres of mass summed relative to the first member and
                        * kept in two parts (far pos + clo), moments taken about their sum */
    int order;         /* expansion order P: 2 or 3 */
    int fix_p3;        /* SPH_FIX_ORDER3: power[3] summed as intended */
};

/* helpers.metal:76-88.  As written the 1/3 and 1/6 weights are integer
 * divisions and three lines assign instead of add: power[3] = |XXX|. */
__device__ __forceinline__ float power3(const float o[10], int fixed) {
    /* o: XXX XXY XXZ XYY XYZ XZZ YYY YYZ YZZ ZZZ */
    if (!fixed) return fabsf(o[0]);
    const float t = 1.f / 3.f, sx = 1.f / 6.f;
    float p3 = o[9] * o[9];
    p3 += t * o[8] * o[8];
    p3 += t * o[7] * o[7];
    p3 += o[6] * o[6];
    p3 += t * o[5] * o[5];
    p3 += sx * o[4] * o[4];
    p3 += t * o[3] * o[3];
    p3 += t * o[2] * o[2];
    p3 += t * o[1] * o[1];
    p3 += o[0] * o[0];
    return sqrtf(p3);
}

__device__ __forceinline__ void store_oct(const UpArgs& a, int id, const float o[10]) {
    a.t.oct[3 * id + 0] = make_float4(o[0], o[1], o[2], o[3]);
    a.t.oct[3 * id + 1] = make_float4(o[4], o[5], o[6], o[7]);
    a.t.oct[3 * id + 2] = make_float4(o[8], o[9], power3(o, a.fix_p3), 0.f);
}

__device__ __forceinline__ float power2(float XX, float XY, float XZ, float YY, float YZ, float ZZ) {
    /* helpers.metal:63-75 */
    float p2 = ZZ * ZZ;
    p2 += 0.5f * YZ * YZ;
    p2 += YY * YY;
    p2 += 0.5f * XZ * XZ;
    p2 += 0.5f * XY * XY;
    p2 += 0.5f * XX * XX;
    return sqrtf(p2);
}

/* Knuth's TwoSum: s = fl(a + b) and e with s + e = a + b exactly, for any
 * a and b (no ordering of their magnitudes needed). */
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

/* P2M.metal:12-107 */
template <bool PREC>
__device__ void p2m(const UpArgs& a, int id, int np, int tp) {
    const int32_t* pl = a.t.octree + tp + 2;
    const int gt = a.time[0];
    float M = 0.f, px = 0.f, py = 0.f, pz = 0.f;
    float minGrav = FLT_MAX;
    bool act = false;
    float mnx = FLT_MAX, mny = FLT_MAX, mnz = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX, mxz = -FLT_MAX;
    /* precision form: sum m (x - x_0) instead of m x, x_0 = first particle */
    const float4 x0 = PREC ? a.pos[pl[0]] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k = 0; k < np; ++k) {
        const int p = pl[k];
        const float4 xa = a.pos[p];
        const float pm = xa.w; /* the position record carries the mass */
        const float4 x = make_float4(xa.x - x0.x, xa.y - x0.y, xa.z - x0.z, 0.f);
        const int na = a.next_active[p];
        const bool isActive = a.readme_rule ? (na <= gt) : (gt <= na);
        act = act || isActive;
        mxx = fmaxf(mxx, xa.x); mxy = fmaxf(mxy, xa.y); mxz = fmaxf(mxz, xa.z);
        mnx = fminf(mnx, xa.x); mny = fminf(mny, xa.y); mnz = fminf(mnz, xa.z);
        M += pm;
        px += x.x * pm;
        py += x.y * pm;
        pz += x.z * pm;
        if (isActive) minGrav = fminf(minGrav, a.gravnorm[p]);
    }
    for (int k = 0; k < np; ++k) a.active[pl[k]] = act ? 1 : 0;
    /* precision form: the centre as x_0 + (offset) in two parts, the second
     * the residual of the first's rounding, so that offsets from the centre
     * keep the bits an FP32 absolute position (|x| ~ 20, ulp ~ 2e-6) drops */
    float clx = 0.f, cly = 0.f, clz = 0.f;
    if constexpr (PREC) {
        two_sum(x0.x, px / M, px, clx);
        two_sum(x0.y, py / M, py, cly);
        two_sum(x0.z, pz / M, pz, clz);
        a.t.clo[id] = make_float4(clx, cly, clz, 0.f);
    } else {
        px = px / M + x0.x;
        py = py / M + x0.y;
        pz = pz / M + x0.z;
    }
    float X = 0.f, Y = 0.f, Z = 0.f, XX = 0.f, XY = 0.f, XZ = 0.f, YY = 0.f, YZ = 0.f, ZZ = 0.f;
    float o[10] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < np; ++k) {
        const int p = pl[k];
        const float4 x = a.pos[p];
        const float pm = x.w;
        /* x - px is exact where x and px are within a factor 2 of each other
         * (every leaf away from the coordinate planes) */
        const float rx = (x.x - px) - clx, ry = (x.y - py) - cly, rz = (x.z - pz) - clz;
        X -= rx * pm;
        Y -= ry * pm;
        Z -= rz * pm;
        XX += pm * rx * rx;
        XY += pm * rx * ry;
        XZ += pm * rx * rz;
        YY += pm * ry * ry;
        YZ += pm * ry * rz;
        ZZ += pm * rz * rz;
        if (a.order == 3) { /* P2M.metal:70-81 */
            o[0] -= pm * rx * rx * rx;
            o[1] -= pm * rx * rx * ry;
            o[2] -= pm * rx * rx * rz;
            o[3] -= pm * rx * ry * ry;
            o[4] -= pm * rx * ry * rz;
            o[5] -= pm * rx * rz * rz;
            o[6] -= pm * ry * ry * ry;
            o[7] -= pm * ry * ry * rz;
            o[8] -= pm * ry * rz * rz;
            o[9] -= pm * rz * rz * rz;
        }
    }
    XX *= 0.5f;
    YY *= 0.5f;
    ZZ *= 0.5f;
    if (a.order == 3) { /* P2M.metal:89-99 */
        const float sixth = 0.166666667f;
        o[0] *= sixth;
        o[6] *= sixth;
        o[9] *= sixth;
        o[1] *= 0.5f;
        o[2] *= 0.5f;
        o[3] *= 0.5f;
        o[5] *= 0.5f;
        o[7] *= 0.5f;
        o[8] *= 0.5f;
        store_oct(a, id, o);
    }
    const float size = fmaxf(fmaxf(mxx - mnx, mxy - mny), mxz - mnz);
    a.t.far[3 * id + 0] = make_float4(px, py, pz, size);
    a.t.far[3 * id + 1] = make_float4(M, XX, XY, XZ);
    a.t.far[3 * id + 2] = make_float4(YY, YZ, ZZ, power2(XX, XY, XZ, YY, YZ, ZZ));
    a.t.aux[3 * id + 0] = make_float4(mnx, mny, mnz, minGrav);
    a.t.aux[3 * id + 1] = make_float4(mxx, mxy, mxz, act ? 1.f : 0.f);
    a.t.aux[3 * id + 2] = make_float4(X, Y, Z, 0.f);
}

/* M2M.metal:75-160 with transformMultipole (M2M.metal:12-73) */
template <bool PREC>
__device__ void m2m(const UpArgs& a, int id, int tp) {
    const int32_t* ch = a.t.child + 8 * id;
    float minGrav = FLT_MAX, mass = 0.f, px = 0.f, py = 0.f, pz = 0.f;
    bool act = false, allLeaves = true;
    float mnx = FLT_MAX, mny = FLT_MAX, mnz = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX, mxz = -FLT_MAX;
    /* precision form: centre of mass relative to the first child's */
    float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (PREC)
        for (int c = 0; c < 8; ++c)
            if (ch[c] >= 0) {
                r0 = a.t.far[3 * ch[c]];
                break;
            }
    for (int c = 0; c < 8; ++c) {
        const int cid = ch[c];
        if (cid < 0) continue;
        float4 f0 = a.t.far[3 * cid];
        const float4 f1 = a.t.far[3 * cid + 1];
        f0.x -= r0.x;
        f0.y -= r0.y;
        f0.z -= r0.z;
        if constexpr (PREC) { /* the child's centre in two parts */
            const float4 cl = a.t.clo[cid];
            f0.x += cl.x;
            f0.y += cl.y;
            f0.z += cl.z;
        }
        const float4 x0 = a.t.aux[3 * cid], x1 = a.t.aux[3 * cid + 1];
        minGrav = fminf(x0.w, minGrav);
        mass += f1.x;
        px += f0.x * f1.x;
        py += f0.y * f1.x;
        pz += f0.z * f1.x;
        act = act || (x1.w != 0.f);
        if (allLeaves && a.t.info[cid].x == 0) allLeaves = false;
        mxx = fmaxf(mxx, x1.x); mxy = fmaxf(mxy, x1.y); mxz = fmaxf(mxz, x1.z);
        mnx = fminf(mnx, x0.x); mny = fminf(mny, x0.y); mnz = fminf(mnz, x0.z);
    }
    const float size = fmaxf(fmaxf(mxx - mnx, mxy - mny), mxz - mnz);
    float clx = 0.f, cly = 0.f, clz = 0.f;
    if constexpr (PREC) {
        two_sum(r0.x, px / mass, px, clx);
        two_sum(r0.y, py / mass, py, cly);
        two_sum(r0.z, pz / mass, pz, clz);
        a.t.clo[id] = make_float4(clx, cly, clz, 0.f);
    } else {
        px = px / mass + r0.x;
        py = py / mass + r0.y;
        pz = pz / mass + r0.z;
    }
    float M = 0.f, XX = 0.f, XY = 0.f, XZ = 0.f, YY = 0.f, YZ = 0.f, ZZ = 0.f;
    float o[10] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < 8; ++c) {
        const int cid = ch[c];
        if (cid < 0) continue;
        const float4 f0 = a.t.far[3 * cid], f1 = a.t.far[3 * cid + 1], f2 = a.t.far[3 * cid + 2];
        float rx = px - f0.x, ry = py - f0.y, rz = pz - f0.z;
        if constexpr (PREC) {
            const float4 cl = a.t.clo[cid];
            rx += clx - cl.x;
            ry += cly - cl.y;
            rz += clz - cl.z;
        }
        const float cM = f1.x;
        /* as written the ZZ shift uses the child's dipole Z (M2M.metal:29) */
        const float zzw = a.fix_zz ? cM : a.t.aux[3 * cid + 2].z;
        M += cM;
        XX += f1.y + 0.5f * rx * rx * cM;
        XY += f1.z + rx * ry * cM;
        XZ += f1.w + rx * rz * cM;
        YY += f2.x + 0.5f * ry * ry * cM;
        YZ += f2.y + ry * rz * cM;
        ZZ += f2.z + 0.5f * rz * rz * zzw;
        if (a.order == 3) { /* M2M.metal:31-71 */
            const float4 q0 = a.t.oct[3 * cid], q1 = a.t.oct[3 * cid + 1], q2 = a.t.oct[3 * cid + 2];
            const float cxx = f1.y, cxy = f1.z, cxz = f1.w, cyy = f2.x, cyz = f2.y, czz = f2.z;
            const float sixth = 0.166666667f;
            o[9] += q2.y + rz * czz + sixth * rz * rz * rz * cM;
            o[8] += q2.x + rz * cyz + ry * czz + 0.5f * ry * rz * rz * cM;
            o[7] += q1.w + rz * cyy + ry * cyz + 0.5f * ry * ry * rz * cM;
            o[6] += q1.z + ry * cyy + sixth * ry * ry * ry * cM;
            o[5] += q1.y + rz * cxz + rx * czz + 0.5f * rx * rz * rz * cM;
            o[4] += q1.x + rz * cxy + ry * cxz + rx * cyz + rx * ry * rz * cM;
            o[3] += q0.w + ry * cxy + rx * cyy + 0.5f * rx * ry * ry * cM;
            o[2] += q0.z + rz * cxx + rx * cxz + 0.5f * rx * rx * rz * cM;
            o[1] += q0.y + ry * cxx + rx * cxy + 0.5f * rx * rx * ry * cM;
            o[0] += q0.x + rx * cxx + sixth * rx * rx * rx * cM;
        }
        if (allLeaves && act && a.t.aux[3 * cid + 1].w == 0.f) {
            /* leaf-sync activation (M2M.metal:145-154); as written the end of
             * the particle range starts from the parent's record */
            a.t.aux[3 * cid + 1].w = 1.f;
            const int4 ci = a.t.info[cid];
            const int cstart = ci.y + 2;
            const int cend = min((a.fix_range ? cstart : tp + 2) + ci.x, cstart + ci.x);
            for (int j = cstart; j < cend; ++j) a.active[a.t.octree[j]] = 1;
        }
    }
    a.t.far[3 * id + 0] = make_float4(px, py, pz, size);
    a.t.far[3 * id + 1] = make_float4(M, XX, XY, XZ);
    a.t.far[3 * id + 2] = make_float4(YY, YZ, ZZ, power2(XX, XY, XZ, YY, YZ, ZZ));
    a.t.aux[3 * id + 0] = make_float4(mnx, mny, mnz, minGrav);
    a.t.aux[3 * id + 1] = make_float4(mxx, mxy, mxz, act ? 1.f : 0.f);
    a.t.aux[3 * id + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (a.order == 3) store_oct(a, id, o);
}

template <bool PREC>
__global__ __launch_bounds__(256) void k_up(UpArgs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.count || a.flags[0] != 0u) return;
    const int id = a.level_id[s];
    const int4 inf = a.t.info[id];
    if (inf.x > 0) {
        p2m<PREC>(a, id, inf.x, inf.y);
        a.t.sub_end[id] = id + 1;
        return;
    }
    m2m<PREC>(a, id, inf.y);
    /* Pre-order numbering (octree.cpp:14-109 numbers nodes as it recurses,
     * children in slot order): the first child is id + 1 and each later
     * child starts where the previous child's subtree ends. */
    int end = id + 1;
    bool ordered = true;
    for (int c = 0; c < 8; ++c) {
        const int cid = a.t.child[8 * id + c];
        if (cid < 0) continue;
        ordered = ordered && cid == end;
        end = a.t.sub_end[cid];
    }
    a.t.sub_end[id] = end;
    if (!ordered && a.need_preorder) a.flags[0] = 2u;
}

/* -------------------------------------------------------------- down pass */
struct DownArgs {
    Tree t;
    const uint32_t* flags;
    const int32_t* level_id;
    int count;
    const int32_t* in;
    int32_t* out;
    int maxu;
    int max_rec;
    float eta;
    float eps;        /* GRAVITY_SMOOTHING_LENGTH */
    float G;
    int fix_soft;
    int near;         /* near-field form: P2P with every rejected source leaf, one softened kernel */
    float* phi;       /* precision form: potential per particle (G included), else null */
    int macp;         /* order of the acceptance test's error estimate (2 or 3) */
    int fix_p3;       /* SPH_FIX_ORDER3 */
    const float4* pos;
    const float* m;
    const uint8_t* active;
    float4* acc;
    float* gravnorm;
    /* README rule: the level's slots whose node is active (k_level_active):
     * the leaves' from the front, act_count[0] of them, the branches' from
     * the back, act_count[kMaxLevels] of them; null: every slot of the level */
    const int32_t* act_slots;
    const int32_t* act_count;
};

struct LocalExp {
    float px, py, pz;
    float lx, ly, lz; /* precision form: the centre's low part (centre = p + l), else 0 */
    float e[eNT3];    /* the first nterms<P>() are used */
};

/* accept.metal:14-55 at P = 2 (the n = 1 term has power[1] = 0). */
__device__ __forceinline__ bool mac_accept(float ax, float ay, float az, float rho_A, float minGrav, float4 b0,
                                           float4 b1, float4 b2, float eta) {
    const float rho_B = b0.w;
    const float dx = ax - b0.x, dy = ay - b0.y, dz = az - b0.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float rho_sum = rho_A + rho_B;
    if (rho_sum * rho_sum > r2) return false;
    const float rho_max = fmaxf(rho_A, rho_B);
    float E = b1.x * (rho_A * rho_A);
    E += b2.w;
    E *= 8.f;
    if (rho_A + rho_B > 0.f) {
        E *= rho_max;
        E /= (rho_A + rho_B);
    }
    return E < eta * minGrav * r2 * r2;
}

/* accept.metal:14-55 at P = 3: E = 8 (M rho_A^3 + 3 power2 rho_A + power3),
 * and r^p = integer_powf(r2, p / 2) = r^2 as written, r^3 fixed. */
__device__ __forceinline__ bool mac_accept3(float ax, float ay, float az, float rho_A, float minGrav, float4 b0,
                                            float4 b1, float4 b2, float p3, float eta, int fix_p3) {
    const float rho_B = b0.w;
    const float dx = ax - b0.x, dy = ay - b0.y, dz = az - b0.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float rho_sum = rho_A + rho_B;
    if (rho_sum * rho_sum > r2) return false;
    const float rho_max = fmaxf(rho_A, rho_B);
    float E = b1.x * (rho_A * rho_A * rho_A);
    E += 3.f * b2.w * rho_A;
    E += p3;
    E *= 8.f;
    if (rho_A + rho_B > 0.f) {
        E *= rho_max;
        E /= (rho_A + rho_B);
    }
    const float r_to_p = fix_p3 ? r2 * sqrtf(r2) : r2;
    return E < eta * minGrav * r_to_p * r2;
}

/* M2L.metal:12-108 + derivatives.metal:59-156: the local coefficients one
 * source node adds at the expansion centre (ten at P = 2; at P = 3 also the
 * octupole's potential, the quadrupole's field and the third derivatives).
 * Each coefficient is handed to `emit(k, value)` as soon as it is final, so
 * that the accumulation (plain or compensated) does not hold all of them. */
template <int P, class Emit>
__device__ __forceinline__ void m2l_emit(float vx, float vy, float vz, float4 b1, float4 b2, const float4* oct,
                                         float eps, Emit emit) {
    /* (vx, vy, vz): the expansion centre minus the source's centre */
    const float r2 = vx * vx + vy * vy + vz * vz;
    /* the hardware square root and reciprocal (about 1 ulp) instead of the
     * correctly rounded library sequences: this is the down pass's
     * per-entry inner loop, and the expansion's own error is orders larger */
    const float r = __builtin_amdgcn_sqrtf(r2);
    const float r_inv = r > 1e-6f ? __builtin_amdgcn_rcpf(r) : 0.f;
    float Dt_1, Dt_2, Dt_3, Dt_4 = 0.f;
    if (r2 < eps * eps) {
        const float eps_inv = 1.f / eps;
        const float u = r * eps_inv;
        float s1 = -3.f * u + 15.f;
        s1 = s1 * u - 28.f; s1 = s1 * u + 21.f; s1 = s1 * u; s1 = s1 * u - 7.f; s1 = s1 * u; s1 = s1 * u + 3.f;
        float s2 = -21.f * u + 90.f;
        s2 = s2 * u - 140.f; s2 = s2 * u + 84.f; s2 = s2 * u; s2 = s2 * u - 14.f; s2 = s2 * u;
        float s3 = -105.f * u + 360.f;
        s3 = s3 * u - 420.f; s3 = s3 * u + 168.f; s3 = s3 * u; s3 = s3 * u;
        const float eps_inv2 = eps_inv * eps_inv;
        Dt_1 = eps_inv * s1;
        Dt_2 = eps_inv2 * s2;
        Dt_3 = eps_inv2 * eps_inv * s3;
        if (P == 3) {
            float s4 = -315.f * u + 720.f;
            s4 = s4 * u - 420.f; s4 = s4 * u; s4 = s4 * u;
            Dt_4 = eps_inv2 * eps_inv2 * s4;
        }
    } else {
        Dt_1 = r_inv;
        Dt_2 = -1.f * Dt_1 * r_inv;
        Dt_3 = -3.f * Dt_2 * r_inv;
        if (P == 3) Dt_4 = -5.f * Dt_3 * r_inv;
    }
    const float rx = vx * r_inv, ry = vy * r_inv, rz = vz * r_inv;
    const float D2 = Dt_2 * r_inv;
    const float M0 = b1.x;
    {
        const float DXX = rx * rx * Dt_3 + D2, DXY = rx * ry * Dt_3, DXZ = rx * rz * Dt_3;
        const float DYY = ry * ry * Dt_3 + D2, DYZ = ry * rz * Dt_3, DZZ = rz * rz * Dt_3 + D2;
        emit(eXX, M0 * DXX);
        emit(eXY, M0 * DXY);
        emit(eXZ, M0 * DXZ);
        emit(eYY, M0 * DYY);
        emit(eYZ, M0 * DYZ);
        emit(eZZ, M0 * DZZ);
        float lm = M0 * Dt_1;
        lm += b1.y * DXX + b2.x * DYY + b2.z * DZZ;
        lm += b1.z * DXY + b1.w * DXZ + b2.y * DYZ;
        if (P != 3) {
            emit(eM, lm);
            emit(eX, M0 * (rx * Dt_2));
            emit(eY, M0 * (ry * Dt_2));
            emit(eZ, M0 * (rz * Dt_2));
            return;
        }
        /* derivatives.metal:140-154, M2L.metal:58-107 */
        const float D3 = Dt_3 * r_inv;
        const float rx2 = rx * rx, ry2 = ry * ry, rz2 = rz * rz;
        const float DXXX = rx2 * rx * Dt_4 + 3.f * rx * D3, DYYY = ry2 * ry * Dt_4 + 3.f * ry * D3;
        const float DZZZ = rz2 * rz * Dt_4 + 3.f * rz * D3;
        const float DXXY = rx2 * ry * Dt_4 + ry * D3, DXXZ = rx2 * rz * Dt_4 + rz * D3;
        const float DXYY = ry2 * rx * Dt_4 + rx * D3, DYYZ = ry2 * rz * Dt_4 + rz * D3;
        const float DXZZ = rz2 * rx * Dt_4 + rx * D3, DYZZ = rz2 * ry * Dt_4 + ry * D3;
        const float DXYZ = rx * ry * rz * Dt_4;
        const float4 o0 = oct[0], o1 = oct[1], o2 = oct[2];
        /* o0 = (XXX XXY XXZ XYY), o1 = (XYZ XZZ YYY YYZ), o2 = (YZZ ZZZ -, -) */
        float lo = o0.x * DXXX + o1.z * DYYY + o2.y * DZZZ;
        lo += o0.y * DXXY + o0.z * DXXZ + o0.w * DXYY;
        lo += o1.w * DYYZ + o1.y * DXZZ + o2.x * DYZZ;
        lo += o1.x * DXYZ;
        emit(eM, lm + lo);
        const float QXX = b1.y, QXY = b1.z, QXZ = b1.w, QYY = b2.x, QYZ = b2.y, QZZ = b2.z;
        emit(eX, M0 * (rx * Dt_2) +
                     (QXX * DXXX + QYY * DXYY + QZZ * DXZZ + (QXY * DXXY + QXZ * DXXZ + QYZ * DXYZ)));
        emit(eY, M0 * (ry * Dt_2) +
                     (QXX * DXXY + QYY * DYYY + QZZ * DYZZ + (QXY * DXYY + QXZ * DXYZ + QYZ * DYYZ)));
        emit(eZ, M0 * (rz * Dt_2) +
                     (QXX * DXXZ + QYY * DYYZ + QZZ * DZZZ + (QXY * DXYZ + QXZ * DXZZ + QYZ * DYZZ)));
        emit(eXXX, M0 * DXXX);
        emit(eXXY, M0 * DXXY);
        emit(eXXZ, M0 * DXXZ);
        emit(eXYY, M0 * DXYY);
        emit(eXYZ, M0 * DXYZ);
        emit(eXZZ, M0 * DXZZ);
        emit(eYYY, M0 * DYYY);
        emit(eYYZ, M0 * DYYZ);
        emit(eYZZ, M0 * DYZZ);
        emit(eZZZ, M0 * DZZZ);
    }
}

template <int P>
__device__ __forceinline__ void m2l_add(const DownArgs& a, LocalExp& L, int node, float4 b0, float4 b1, float4 b2) {
    m2l_emit<P>(L.px - b0.x, L.py - b0.y, L.pz - b0.z, b1, b2, a.t.oct + 3 * node, a.eps,
                [&](int k, float v) { L.e[k] += v; });
}

/* Neumaier's compensated sum (Neumaier 1974): s + c carries the sum of the
 * added terms to about twice the working precision. */
__device__ __forceinline__ void kadd(float& s, float& c, float x) {
    const float t = s + x;
    if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

template <int P>
__device__ __forceinline__ void m2l_add_comp(const DownArgs& a, LocalExp& L, float Lc[eNT3], int node, float4 b0,
                                             float4 b1, float4 b2) {
    /* centres in two parts: the high parts' difference plus the low parts' */
    const float4 cl = a.t.clo[node];
    m2l_emit<P>((L.px - b0.x) + (L.lx - cl.x), (L.py - b0.y) + (L.ly - cl.y), (L.pz - b0.z) + (L.lz - cl.z), b1, b2,
                a.t.oct + 3 * node, a.eps, [&](int k, float v) { kadd(L.e[k], Lc[k], v); });
}

/* Softened 1/r whose derivatives M2L uses (derivatives.metal:12-57,
 * support eps): the pair potential of the near-field form per unit mass. */
__device__ __forceinline__ float pot_m2l(float r, float eps) {
    if (r >= eps) return 1.f / r;
    const float eps_inv = 1.f / eps;
    const float u = r * eps_inv;
    float s1 = -3.f * u + 15.f;
    s1 = s1 * u - 28.f; s1 = s1 * u + 21.f; s1 = s1 * u; s1 = s1 * u - 7.f; s1 = s1 * u; s1 = s1 * u + 3.f;
    return eps_inv * s1;
}

struct NodeA {
    float x, y, z, size, minGrav;
};

/* The acceptance test with the error estimate of order a.macp. */
__device__ __forceinline__ bool mac_ok(const DownArgs& a, const NodeA& A, int node, float4 b0, float4 b1, float4 b2) {
    if (a.macp == 3)
        return mac_accept3(A.x, A.y, A.z, A.size, A.minGrav, b0, b1, b2, a.t.oct[3 * node + 2].z, a.eta, a.fix_p3);
    return mac_accept(A.x, A.y, A.z, A.size, A.minGrav, b0, b1, b2, a.eta);
}

/* checkAndAddLocalExpansion (helpers.h:34-45) */
__device__ __forceinline__ bool check_add(const DownArgs& a, const NodeA& A, LocalExp& L, int node, bool force) {
    const float4 b0 = a.t.far[3 * node], b1 = a.t.far[3 * node + 1], b2 = a.t.far[3 * node + 2];
    if (force || mac_ok(a, A, node, b0, b1, b2)) {
        m2l_add<2>(a, L, node, b0, b1, b2);
        return true;
    }
    return false;
}

/* recursiveScan (helpers.h:83-123) for a leaf */
__device__ void recursive_scan(const DownArgs& a, const NodeA& A, LocalExp& L, int toScan, int self) {
    if (check_add(a, A, L, toScan, a.t.info[toScan].x != 0)) return;
    int nodes[4], idx[4];
    int depth = 0;
    nodes[0] = toScan;
    idx[0] = 0;
    int count = 0;
    while (idx[0] < 8 && count < 2000) {
        count++;
        const int np = a.t.child[8 * nodes[depth] + idx[depth]];
        if (np == self || np < 0) {
            idx[depth]++;
            while (idx[depth] == 8 && depth != 0) {
                depth--;
                idx[depth]++;
            }
            continue;
        }
        if (!check_add(a, A, L, np, a.t.info[np].x != 0 || depth == a.max_rec - 1)) {
            depth++;
            nodes[depth] = np;
            idx[depth] = 0;
            continue;
        }
        idx[depth]++;
        while (idx[depth] == 8 && depth != 0) {
            depth--;
            idx[depth]++;
        }
    }
}

/* dphi_dr (kernels.h:87-102); as written the fractions are integer divisions */
__device__ __forceinline__ float dphi_dr(float r, float h, int fixed) {
    const float h1 = 1.f / fminf(h, 0.16f);
    const float q = r * h1;
    const float c43 = fixed ? 4.f / 3.f : 1.f, c65 = fixed ? 6.f / 5.f : 1.f;
    const float c83 = fixed ? 8.f / 3.f : 2.f, c16 = fixed ? 1.f / 6.f : 0.f, c115 = fixed ? 1.f / 15.f : 0.f;
    if (q <= 1.f) return (h1 * h1) * (c43 * q - c65 * q * q * q + 0.5f * q * q * q * q);
    if (q <= 2.f) {
        const float q1 = 1.f / q;
        return (h1 * h1) * (c83 * q - 3.f * q * q + c65 * q * q * q - c16 * q * q * q * q - c115 * q1 * q1);
    }
    const float r1 = 1.f / r;
    return r1 * r1;
}

/* -d(potential)/dr of a unit mass for the softened potential whose
 * derivatives M2L uses (derivatives.metal:12-57, support eps).  The
 * near-field form sums every pair with this kernel, so a pair's force does
 * not depend on whether it is reached by M2L or by P2P. */
__device__ __forceinline__ float dphi_dr_m2l(float r, float eps) {
    if (r >= eps) return 1.f / (r * r);
    const float eps_inv = 1.f / eps;
    const float u = r * eps_inv;
    float s2 = -21.f * u + 90.f;
    s2 = s2 * u - 140.f; s2 = s2 * u + 84.f; s2 = s2 * u; s2 = s2 * u - 14.f; s2 = s2 * u;
    return -(eps_inv * eps_inv) * s2;
}

/* transformLocal (L2L.metal:12-88) of the node's local expansion to one
 * active child, written to the child's record (L2L.metal:90-104).  PREC:
 * centres in two parts; the child's record holds its centre's high part. */
template <int P, bool PREC = false>
__device__ __forceinline__ void l2l_child(const DownArgs& a, const LocalExp& L, int cid) {
    if (cid < 0) return;
    if (a.t.aux[3 * cid + 1].w == 0.f) return;
    const float4 c0 = a.t.far[3 * cid];
    float rx = c0.x - L.px, ry = c0.y - L.py, rz = c0.z - L.pz;
    if constexpr (PREC) {
        const float4 cl = a.t.clo[cid];
        rx += cl.x - L.lx;
        ry += cl.y - L.ly;
        rz += cl.z - L.lz;
    }
    const float* e = L.e;
    float nM = e[eM];
    nM += rx * e[eX] + ry * e[eY] + rz * e[eZ];
    nM += 0.5f * (rx * rx * e[eXX] + ry * ry * e[eYY] + rz * rz * e[eZZ]);
    nM += rx * ry * e[eXY] + rx * rz * e[eXZ] + ry * rz * e[eYZ];
    float nX = e[eX] + (rx * e[eXX] + ry * e[eXY] + rz * e[eXZ]);
    float nY = e[eY] + (rx * e[eXY] + ry * e[eYY] + rz * e[eYZ]);
    float nZ = e[eZ] + (rx * e[eXZ] + ry * e[eYZ] + rz * e[eZZ]);
    float nXX = e[eXX], nXY = e[eXY], nXZ = e[eXZ], nYY = e[eYY], nYZ = e[eYZ], nZZ = e[eZZ];
    if (P == 3) {
        /* L2L.metal:37-86 */
        const float sixth = 0.166666667f;
        nM += sixth * rx * rx * rx * e[eXXX] + sixth * ry * ry * ry * e[eYYY] + sixth * rz * rz * rz * e[eZZZ];
        nM += 0.5f * rx * rx * ry * e[eXXY] + 0.5f * rx * rx * rz * e[eXXZ] + 0.5f * rx * ry * ry * e[eXYY];
        nM += 0.5f * ry * ry * rz * e[eYYZ] + 0.5f * rx * rz * rz * e[eXZZ] + 0.5f * ry * rz * rz * e[eYZZ];
        nM += rx * ry * rz * e[eXYZ];
        nX += 0.5f * rx * rx * e[eXXX] + 0.5f * ry * ry * e[eXYY] + 0.5f * rz * rz * e[eXZZ];
        nX += rx * ry * e[eXXY] + rx * rz * e[eXXZ] + ry * rz * e[eXYZ];
        nY += 0.5f * rx * rx * e[eXXY] + 0.5f * ry * ry * e[eYYY] + 0.5f * rz * rz * e[eYZZ];
        nY += rx * ry * e[eXYY] + rx * rz * e[eXYZ] + ry * rz * e[eYYZ];
        nZ += 0.5f * rx * rx * e[eXXZ] + 0.5f * ry * ry * e[eYYZ] + 0.5f * rz * rz * e[eZZZ];
        nZ += rx * ry * e[eXYZ] + rx * rz * e[eXZZ] + ry * rz * e[eYZZ];
        nXX += rx * e[eXXX] + ry * e[eXXY] + rz * e[eXXZ];
        nYY += rx * e[eXYY] + ry * e[eYYY] + rz * e[eYYZ];
        nZZ += rx * e[eXZZ] + ry * e[eYZZ] + rz * e[eZZZ];
        nXY += rx * e[eXXY] + ry * e[eXYY] + rz * e[eXYZ];
        nXZ += rx * e[eXXZ] + ry * e[eXYZ] + rz * e[eXZZ];
        nYZ += rx * e[eXYZ] + ry * e[eYYZ] + rz * e[eYZZ];
        a.t.loc3[3 * cid + 0] = make_float4(e[eXXX], e[eXXY], e[eXXZ], e[eXYY]);
        a.t.loc3[3 * cid + 1] = make_float4(e[eXYZ], e[eXZZ], e[eYYY], e[eYYZ]);
        a.t.loc3[3 * cid + 2] = make_float4(e[eYZZ], e[eZZZ], 0.f, 0.f);
    }
    if constexpr (PREC) a.t.loc[4 * cid + 0] = make_float4(c0.x, c0.y, c0.z, nM);
    else a.t.loc[4 * cid + 0] = make_float4(L.px + rx, L.py + ry, L.pz + rz, nM);
    a.t.loc[4 * cid + 1] = make_float4(nX, nY, nZ, nXX);
    a.t.loc[4 * cid + 2] = make_float4(nXY, nXZ, nYY, nYZ);
    a.t.loc[4 * cid + 3] = make_float4(nZZ, 0.f, 0.f, 0.f);
}

/* The node's incoming local expansion (written by its parent's L2L) added
 * to the group's M2L sum. */
template <int P>
__device__ __forceinline__ void add_incoming(const DownArgs& a, int id, float4 l0, LocalExp& L) {
    const float4 l1 = a.t.loc[4 * id + 1], l2 = a.t.loc[4 * id + 2], l3 = a.t.loc[4 * id + 3];
    L.e[eM] = l0.w + L.e[eM];
    L.e[eX] = l1.x + L.e[eX];
    L.e[eY] = l1.y + L.e[eY];
    L.e[eZ] = l1.z + L.e[eZ];
    L.e[eXX] = l1.w + L.e[eXX];
    L.e[eXY] = l2.x + L.e[eXY];
    L.e[eXZ] = l2.y + L.e[eXZ];
    L.e[eYY] = l2.z + L.e[eYY];
    L.e[eYZ] = l2.w + L.e[eYZ];
    L.e[eZZ] = l3.x + L.e[eZZ];
    if (P == 3) {
        const float4 m0 = a.t.loc3[3 * id], m1 = a.t.loc3[3 * id + 1], m2 = a.t.loc3[3 * id + 2];
        L.e[eXXX] = m0.x + L.e[eXXX];
        L.e[eXXY] = m0.y + L.e[eXXY];
        L.e[eXXZ] = m0.z + L.e[eXXZ];
        L.e[eXYY] = m0.w + L.e[eXYY];
        L.e[eXYZ] = m1.x + L.e[eXYZ];
        L.e[eXZZ] = m1.y + L.e[eXZZ];
        L.e[eYYY] = m1.z + L.e[eYYY];
        L.e[eYYZ] = m1.w + L.e[eYYZ];
        L.e[eYZZ] = m2.x + L.e[eYZZ];
        L.e[eZZZ] = m2.y + L.e[eZZZ];
    }
}

/* L2P + intra-leaf P2P for particle k of a leaf (helpers.h:47-81); (nx, ny,
 * nz) is the near-field form's sum over rejected source leaves (else 0).
 * PREC (precision form, near field only): compensated sums, and the
 * potential -G (local expansion at x + sum of the pair potentials) with
 * `nphi` the rejected leaves' part.  INTRA false: (nx, ny, nz, nphi) already
 * hold the intra-leaf pairs (summed over the group's lanes). */
template <bool PREC, int P, bool INTRA = true>
__device__ __forceinline__ void leaf_particle(const DownArgs& a, const LocalExp& L, const int32_t* pl, int np, int k,
                                              float nx = 0.f, float ny = 0.f, float nz = 0.f, float nphi = 0.f) {
    const int p = pl[k];
    if (!a.active[p]) return;
    const float4 xi = a.pos[p];
    /* PREC: the offset from the centre's two parts (xi - px is exact) */
    const float rx = PREC ? (xi.x - L.px) - L.lx : xi.x - L.px;
    const float ry = PREC ? (xi.y - L.py) - L.ly : xi.y - L.py;
    const float rz = PREC ? (xi.z - L.pz) - L.lz : xi.z - L.pz;
    const float* e = L.e;
    /* L2P (L2P.metal:12-32) */
    float ax = e[eX], ay = e[eY], az = e[eZ];
    ax += rx * e[eXX] + ry * e[eXY] + rz * e[eXZ];
    ay += rx * e[eXY] + ry * e[eYY] + rz * e[eYZ];
    az += rx * e[eXZ] + ry * e[eYZ] + rz * e[eZZ];
    if (P == 3) {
        ax += 0.5f * rx * rx * e[eXXX] + 0.5f * ry * ry * e[eXYY] + 0.5f * rz * rz * e[eXZZ];
        ax += rx * ry * e[eXXY] + rx * rz * e[eXXZ] + ry * rz * e[eXYZ];
        ay += 0.5f * rx * rx * e[eXXY] + 0.5f * ry * ry * e[eYYY] + 0.5f * rz * rz * e[eYZZ];
        ay += rx * ry * e[eXYY] + rx * rz * e[eXYZ] + ry * rz * e[eYYZ];
        az += 0.5f * rx * rx * e[eXXZ] + 0.5f * ry * ry * e[eYYZ] + 0.5f * rz * rz * e[eZZZ];
        az += rx * ry * e[eXYZ] + rx * rz * e[eXZZ] + ry * rz * e[eYZZ];
    }
    float cx = 0.f, cy = 0.f, cz = 0.f, ps = 0.f, pc = 0.f;
    if (PREC) {
        ps = e[eM] + rx * e[eX] + ry * e[eY] + rz * e[eZ];
        ps += 0.5f * (rx * rx * e[eXX] + ry * ry * e[eYY] + rz * rz * e[eZZ]);
        ps += rx * ry * e[eXY] + rx * rz * e[eXZ] + ry * rz * e[eYZ];
        if (P == 3) {
            const float sixth = 0.166666667f;
            ps += sixth * (rx * rx * rx * e[eXXX] + ry * ry * ry * e[eYYY] + rz * rz * rz * e[eZZZ]);
            ps += 0.5f * (rx * rx * ry * e[eXXY] + rx * rx * rz * e[eXXZ] + rx * ry * ry * e[eXYY]);
            ps += 0.5f * (ry * ry * rz * e[eYYZ] + rx * rz * rz * e[eXZZ] + ry * rz * rz * e[eYZZ]);
            ps += rx * ry * rz * e[eXYZ];
        }
    }
    for (int j = 0; INTRA && j < np; ++j) {
        if (j == k) continue;
        const int pj = pl[j];
        const float4 xj = a.pos[pj];
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float r = sqrtf(dx * dx + dy * dy + dz * dz);
        if (r == 0.f) continue;
        const float mj = xj.w;
        const float w = mj * (a.near ? dphi_dr_m2l(r, a.eps) : dphi_dr(r, a.eps, a.fix_soft));
        if (PREC) {
            kadd(ax, cx, -(w * (dx / r)));
            kadd(ay, cy, -(w * (dy / r)));
            kadd(az, cz, -(w * (dz / r)));
            kadd(ps, pc, mj * pot_m2l(r, a.eps));
        } else {
            ax -= w * (dx / r);
            ay -= w * (dy / r);
            az -= w * (dz / r);
        }
    }
    if (PREC) {
        kadd(ax, cx, nx);
        kadd(ay, cy, ny);
        kadd(az, cz, nz);
        kadd(ps, pc, nphi);
        ax += cx;
        ay += cy;
        az += cz;
        a.phi[p] = -a.G * (ps + pc);
    } else {
        ax += nx;
        ay += ny;
        az += nz;
    }
    a.acc[p] = make_float4(a.G * ax, a.G * ay, a.G * az, 0.f);
    a.gravnorm[p] = sqrtf(ax * ax + ay * ay + az * az);
}

template <bool PREC, int P>
__device__ __forceinline__ void leaf_field_lanes(const DownArgs& a, const LocalExp& L, int4 inf, int lane, int width) {
    const int32_t* pl = a.t.octree + inf.y + 2;
    for (int k = lane; k < inf.x; k += width) leaf_particle<PREC, P>(a, L, pl, inf.x, k);
}

/* localToLeaf (helpers.h:47-81): L2P plus the intra-leaf P2P */
template <int P>
__device__ void leaf_field(const DownArgs& a, const LocalExp& L, int4 inf) {
    if (a.phi) leaf_field_lanes<true, P>(a, L, inf, 0, 1);
    else leaf_field_lanes<false, P>(a, L, inf, 0, 1);
}

/* downPass (gravity.metal:64-143) for one non-root node */
__global__ __launch_bounds__(256) void k_down(DownArgs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.count || a.flags[0] != 0u) return;
    const int id = a.level_id[s];
    if (a.t.aux[3 * id + 1].w == 0.f) return;
    const int4 inf = a.t.info[id];
    const int np = inf.x;
    uint32_t outp = (uint32_t)s * (uint32_t)a.maxu;
    const uint32_t maxOut = outp + (uint32_t)a.maxu;
    if (np == 0) a.out[outp++] = id;
    const float4 f0 = a.t.far[3 * id];
    NodeA A{f0.x, f0.y, f0.z, f0.w, a.t.aux[3 * id].w};
    LocalExp L;
    {
        const float4 l0 = a.t.loc[4 * id], l1 = a.t.loc[4 * id + 1], l2 = a.t.loc[4 * id + 2],
                     l3 = a.t.loc[4 * id + 3];
        L.px = l0.x; L.py = l0.y; L.pz = l0.z;
        L.lx = L.ly = L.lz = 0.f;
        L.e[eM] = l0.w; L.e[eX] = l1.x; L.e[eY] = l1.y; L.e[eZ] = l1.z; L.e[eXX] = l1.w;
        L.e[eXY] = l2.x; L.e[eXZ] = l2.y; L.e[eYY] = l2.z; L.e[eYZ] = l2.w; L.e[eZZ] = l3.x;
    }
    const int32_t* in = a.in + (int64_t)a.t.parent_slot[id] * a.maxu;
    for (int i = 0; i < a.maxu; ++i) {
        const int toScan = in[i];
        if (toScan == -1) break;
        if (np != 0) {
            recursive_scan(a, A, L, toScan, id);
        } else {
            /* nonRecursiveScan (helpers.h:125-147) */
            if (check_add(a, A, L, toScan, false)) continue;
            for (int c = 0; c < 8; ++c) {
                const int cid = a.t.child[8 * toScan + c];
                if (cid == id || cid < 0) continue;
                if (!check_add(a, A, L, cid, a.t.info[cid].x != 0 || outp == maxOut)) a.out[outp++] = cid;
            }
        }
    }
    if (outp < maxOut && np == 0) a.out[outp] = -1;
    if (np == 0) {
        /* L2L (L2L.metal:90-104) to the active children */
        for (int c = 0; c < 8; ++c) l2l_child<2>(a, L, a.t.child[8 * id + c]);
    } else {
        leaf_field<2>(a, L, inf);
    }
}


/* Group reduction of one value over GS lanes (fixed order: reproducible). */
template <int GS>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int off = GS / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, GS);
    return v;
}

/* Targets of a near-field leaf (at most NEAR_K particles) in registers and
 * their accumulated field (compensations and potential in the precision
 * form). */
template <int NEAR_K>
struct NearTargets {
    float tx[NEAR_K], ty[NEAR_K], tz[NEAR_K];
    float gx[NEAR_K], gy[NEAR_K], gz[NEAR_K];
    float cx[NEAR_K], cy[NEAR_K], cz[NEAR_K], gp[NEAR_K], cp[NEAR_K];
};

/* The pair terms of one source particle (position record xj, mass in w)
 * against the targets, with the M2L softening kernel (the near-field form's
 * single pair force). */
template <bool PREC, int NEAR_K>
__device__ __forceinline__ void near_pair(const DownArgs& a, NearTargets<NEAR_K>& T, int np, float4 xj) {
    {
        const float mj = xj.w;
#pragma unroll
        for (int kk = 0; kk < NEAR_K; ++kk) {
            const float dx = T.tx[kk] - xj.x, dy = T.ty[kk] - xj.y, dz = T.tz[kk] - xj.z;
            if constexpr (PREC) {
                /* correctly rounded square root and one correctly rounded
                 * reciprocal, products for the rest (instead of a division
                 * per component and per kernel term) */
                const float r = sqrtf(dx * dx + dy * dy + dz * dz);
                if (kk >= np || r == 0.f) continue;
                const float ri = 1.f / r;
                float wr, pot;
                if (r >= a.eps) {
                    wr = mj * (ri * ri) * ri;
                    pot = ri;
                } else {
                    wr = mj * dphi_dr_m2l(r, a.eps) * ri;
                    pot = pot_m2l(r, a.eps);
                }
                kadd(T.gx[kk], T.cx[kk], -(wr * dx));
                kadd(T.gy[kk], T.cy[kk], -(wr * dy));
                kadd(T.gz[kk], T.cz[kk], -(wr * dz));
                kadd(T.gp[kk], T.cp[kk], mj * pot);
            } else {
                /* FP32 form: the hardware square root and reciprocal (about
                 * 1 ulp), as in M2L; the precision form keeps the correctly
                 * rounded operations its compensated sums are built on */
                const float r = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
                if (kk >= np || r == 0.f) continue;
                const float ri = __builtin_amdgcn_rcpf(r);
                const float w = mj * (r >= a.eps ? ri * ri : dphi_dr_m2l(r, a.eps));
                T.gx[kk] -= w * (dx * ri);
                T.gy[kk] -= w * (dy * ri);
                T.gz[kk] -= w * (dz * ri);
            }
        }
    }
}

/* P2P of one rejected source leaf against the targets. */
template <bool PREC, int NEAR_K>
__device__ __forceinline__ void near_p2p(const DownArgs& a, NearTargets<NEAR_K>& T, int np, int4 ci) {
    const int32_t* src = a.t.octree + ci.y + 2;
    for (int q = 0; q < ci.x; ++q) near_pair<PREC, NEAR_K>(a, T, np, a.pos[src[q]]);
}

/* A near-field leaf's complete local expansion, between its two launches. */
template <int P>
__device__ __forceinline__ void store_local_total(const DownArgs& a, int id, const LocalExp& L) {
    float4* r = a.t.loc_tot + 7 * id;
    const float* e = L.e;
    r[0] = make_float4(L.px, L.py, L.pz, e[eM]);
    r[1] = make_float4(e[eX], e[eY], e[eZ], e[eXX]);
    r[2] = make_float4(e[eXY], e[eXZ], e[eYY], e[eYZ]);
    r[3] = make_float4(e[eZZ], 0.f, 0.f, 0.f);
    if (P == 3) {
        r[4] = make_float4(e[eXXX], e[eXXY], e[eXXZ], e[eXYY]);
        r[5] = make_float4(e[eXYZ], e[eXZZ], e[eYYY], e[eYYZ]);
        r[6] = make_float4(e[eYZZ], e[eZZZ], 0.f, 0.f);
    }
}

template <int P>
__device__ __forceinline__ void load_local_total(const DownArgs& a, int id, LocalExp& L) {
    const float4* r = a.t.loc_tot + 7 * id;
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    L.px = r0.x;
    L.py = r0.y;
    L.pz = r0.z;
    L.e[eM] = r0.w;
    L.e[eX] = r1.x;
    L.e[eY] = r1.y;
    L.e[eZ] = r1.z;
    L.e[eXX] = r1.w;
    L.e[eXY] = r2.x;
    L.e[eXZ] = r2.y;
    L.e[eYY] = r2.z;
    L.e[eYZ] = r2.w;
    L.e[eZZ] = r3.x;
    if (P == 3) {
        const float4 r4 = r[4], r5 = r[5], r6 = r[6];
        L.e[eXXX] = r4.x;
        L.e[eXXY] = r4.y;
        L.e[eXXZ] = r4.z;
        L.e[eXYY] = r4.w;
        L.e[eXYZ] = r5.x;
        L.e[eXZZ] = r5.y;
        L.e[eYYY] = r5.z;
        L.e[eYYZ] = r5.w;
        L.e[eYZZ] = r6.x;
        L.e[eZZZ] = r6.y;
    }
}

/* Near-field form at a leaf, with the work balanced over the GS lanes: every
 * round each lane takes one node -- list entries first, then nodes from the
 * group's LDS queue -- and either M2Ls it (MAC accepted), sums it pair by
 * pair (rejected leaf) or queues its children (rejected branch).  Queue
 * positions come from ballots in lane order, so the assignment of nodes to
 * lanes, and with it every sum, is reproducible.  When the queue is full a
 * lane walks the branch's subtree itself in pre-order (node ids are
 * pre-order: a subtree is [id, subtree_end[id])).  The targets are the
 * leaf's particles, NEAR_K at a time in registers; the M2L part runs in the
 * first pass only.  Partial sums are reduced over the group, the incoming
 * local expansion is added, and lane k finishes particle k of the pass. */
template <int GS, bool PREC, int P, int QCAP, int PART>
__device__ void down_near_leaf_q(const DownArgs& a, int id, int4 inf, const NodeA& A, LocalExp& L, float4 l0,
                                 const int32_t* in, int lane, int wl, uint64_t gmask, int32_t* q) {
    /* PART 1: the walk's M2L part only; the complete local expansion goes to
     * loc_tot.  PART 2: the walk's pair sums only, with that expansion read
     * back for L2P.  Both make the same walk (the same acceptance decisions,
     * queue order and lane assignment), so together they sum what one walk
     * does, with fewer registers live in each. */
    constexpr int NEAR_K = 4;
    constexpr int NTP = nterms<P>();
    const int np = inf.x;
    const int32_t* pl = a.t.octree + inf.y + 2;
    const uint64_t ltmask = wl == 0 ? 0ull : (~0ull >> (64 - wl));
    float Lc[eNT3];
#pragma unroll
    for (int k = 0; k < NTP; ++k) Lc[k] = 0.f;
    for (int k0 = 0; k0 < np; k0 += NEAR_K) {
        if (PART == 1 && k0 > 0) break;
        const int nk = min(NEAR_K, np - k0);
        NearTargets<NEAR_K> T;
#pragma unroll
        for (int kk = 0; kk < NEAR_K; ++kk) {
            const float4 x = (PART != 1 && kk < nk) ? a.pos[pl[k0 + kk]] : make_float4(0.f, 0.f, 0.f, 0.f);
            T.tx[kk] = x.x;
            T.ty[kk] = x.y;
            T.tz[kk] = x.z;
            T.gx[kk] = T.gy[kk] = T.gz[kk] = 0.f;
            T.cx[kk] = T.cy[kk] = T.cz[kk] = T.gp[kk] = T.cp[kk] = 0.f;
        }
        const bool do_m2l = PART != 2 && k0 == 0;
        /* PART 2 with the M2L launch's record of the rejected source leaves'
         * particles: every lane takes every GS-th source particle, no walk */
        if constexpr (PART == 2) {
            const int rc = a.t.near_cap > 0 ? a.t.near_cnt[id] : -1;
            if (rc >= 0) {
                const int32_t* rec = a.t.near_rec + (int64_t)id * a.t.near_cap;
                /* the recorded source particles, then the leaf's own, as one
                 * sequence over the lanes */
                for (int e = lane; e < rc + np; e += GS)
                    near_pair<PREC, NEAR_K>(a, T, nk, a.pos[e < rc ? rec[e] : pl[e - rc]]);
                goto reduce_sums;
            }
        }
        {
        int nrec = 0;          /* PART 1: source particles recorded so far (group-uniform) */
        bool rec_off = false;  /* PART 1: some lane walked a subtree alone (unrecorded order) */
        int lp = 0;            /* next list index (group-uniform, as are the rest) */
        bool list_open = true;
        int qh = 0, qt = 0;    /* queue head and tail (monotone; slot = value % QCAP) */
        for (;;) {
            int item = -1, nfill = 0;
            if (list_open) {
                const int i = lp + lane;
                const int e = i < a.maxu ? in[i] : -1;
                const uint64_t term = __ballot(e == -1) & gmask;
                const uint64_t first = term & (~term + 1ull);
                const bool valid = i < a.maxu && (term == 0ull || (1ull << wl) < first);
                nfill = __popcll(__ballot(valid) & gmask);
                if (valid) item = e;
                lp += nfill;
                if (term != 0ull || lp >= a.maxu) list_open = false;
            }
            const int take = min(qt - qh, GS - nfill);
            if (lane >= nfill && lane - nfill < take) item = q[(qh + lane - nfill) % QCAP];
            qh += take;
            if (nfill + take == 0 && !list_open) break;
            int nch = 0;
            bool near_leaf = false;
            int4 near_ci = make_int4(0, 0, 0, 0);
            if (item >= 0 && item != id) {
                const float4 b0 = a.t.far[3 * item], b1 = a.t.far[3 * item + 1], b2 = a.t.far[3 * item + 2];
                if (mac_ok(a, A, item, b0, b1, b2)) {
                    if (do_m2l) {
                        if constexpr (PREC) m2l_add_comp<P>(a, L, Lc, item, b0, b1, b2);
                        else m2l_add<P>(a, L, item, b0, b1, b2);
                    }
                } else {
                    const int4 ci = a.t.info[item];
                    if (ci.x != 0) {
                        near_leaf = true;
                        near_ci = ci;
                        if (PART != 1) near_p2p<PREC, NEAR_K>(a, T, nk, ci);
                    } else {
#pragma unroll
                        for (int c = 0; c < 8; ++c) nch += a.t.child[8 * item + c] >= 0 ? 1 : 0;
                    }
                }
            }
            if constexpr (PART == 1) {
                if (a.t.near_cap > 0) {
                    /* record the round's rejected leaves' particles, leaves in
                     * lane order (a group prefix sum of their counts) */
                    const int cntp = near_leaf ? near_ci.x : 0;
                    int incl = cntp;
#pragma unroll
                    for (int off = 1; off < GS; off <<= 1) {
                        const int v = __shfl_up(incl, off, GS);
                        if (lane >= off) incl += v;
                    }
                    const int total = __shfl(incl, GS - 1, GS);
                    const int pos0 = nrec + incl - cntp;
                    int32_t* rec = a.t.near_rec + (int64_t)id * a.t.near_cap;
                    const int32_t* src = a.t.octree + near_ci.y + 2;
                    for (int q = 0; q < cntp && pos0 + q < a.t.near_cap; ++q) rec[pos0 + q] = src[q];
                    nrec += total;
                }
            }
            /* group prefix sum of the child counts (0..8) from four bit ballots */
            int before = 0, total = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint64_t m = __ballot((nch >> b) & 1) & gmask;
                before += __popcll(m & ltmask) << b;
                total += __popcll(m) << b;
            }
            if (total <= QCAP - (qt - qh)) {
                if (nch > 0) {
                    int pos = qt + before;
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const int ch = a.t.child[8 * item + c];
                        if (ch >= 0) q[(pos++) % QCAP] = ch;
                    }
                }
                qt += total;
            } else if (nch > 0) {
                /* queue full: walk the branch's subtree here, in pre-order */
                rec_off = true;
                int cur = item + 1;
                const int end = a.t.sub_end[item];
                while (cur < end) {
                    if (cur == id) {
                        ++cur;
                        continue;
                    }
                    const float4 b0 = a.t.far[3 * cur], b1 = a.t.far[3 * cur + 1], b2 = a.t.far[3 * cur + 2];
                    if (mac_ok(a, A, cur, b0, b1, b2)) {
                        if (do_m2l) {
                            if constexpr (PREC) m2l_add_comp<P>(a, L, Lc, cur, b0, b1, b2);
                            else m2l_add<P>(a, L, cur, b0, b1, b2);
                        }
                        cur = a.t.sub_end[cur];
                        continue;
                    }
                    const int4 ci = a.t.info[cur];
                    ++cur;
                    if (ci.x != 0 && PART != 1) near_p2p<PREC, NEAR_K>(a, T, nk, ci);
                }
            }
            /* the queue writes of this round are read by other lanes next round */
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        if constexpr (PART == 1) {
            if (a.t.near_cap > 0) {
                const bool off = (__ballot(rec_off) & gmask) != 0ull;
                if (lane == 0) a.t.near_cnt[id] = (off || nrec > a.t.near_cap) ? -1 : nrec;
            }
        }
        }
        /* the leaf's own particles as sources, every lane taking every GS-th
         * (a target's pair with itself has r = 0 and adds nothing, as in the
         * intra-leaf P2P of helpers.h:47-81) */
        if constexpr (PART != 1)
            for (int e = lane; e < np; e += GS) near_pair<PREC, NEAR_K>(a, T, nk, a.pos[pl[e]]);
    reduce_sums:
#pragma unroll
        for (int kk = 0; kk < NEAR_K; ++kk) {
            T.gx[kk] = group_sum<GS>(T.gx[kk]);
            T.gy[kk] = group_sum<GS>(T.gy[kk]);
            T.gz[kk] = group_sum<GS>(T.gz[kk]);
            if constexpr (PREC) {
                T.gx[kk] += group_sum<GS>(T.cx[kk]);
                T.gy[kk] += group_sum<GS>(T.cy[kk]);
                T.gz[kk] += group_sum<GS>(T.cz[kk]);
                T.gp[kk] = group_sum<GS>(T.gp[kk]) + group_sum<GS>(T.cp[kk]);
            }
        }
        if (do_m2l) {
#pragma unroll
            for (int k = 0; k < NTP; ++k) {
                L.e[k] = group_sum<GS>(L.e[k]);
                if constexpr (PREC) L.e[k] += group_sum<GS>(Lc[k]);
            }
            add_incoming<P>(a, id, l0, L);
        }
        if constexpr (PART == 1) {
            if (lane == 0) store_local_total<P>(a, id, L);
            return;
        }
        if constexpr (PART == 2) load_local_total<P>(a, id, L);
        /* every lane holds every target's sums (butterfly reductions): lane kk
         * takes target kk's, so that L2P runs once with nk lanes active rather
         * than once per target with one lane active */
        float sgx = T.gx[0], sgy = T.gy[0], sgz = T.gz[0], sgp = T.gp[0];
#pragma unroll
        for (int kk = 1; kk < NEAR_K; ++kk)
            if (lane == kk) {
                sgx = T.gx[kk];
                sgy = T.gy[kk];
                sgz = T.gz[kk];
                sgp = T.gp[kk];
            }
        if (lane < nk) leaf_particle<PREC, P, false>(a, L, pl, np, k0 + lane, sgx, sgy, sgz, sgp);
    }
}

/* Down pass for GRAVITY_MAX_RECURSION == 1 with GS lanes per node.
 *
 * Lanes test GS list entries at a time against the node (MAC); accepted
 * entries are M2L'd by their lane.  The children of rejected entries are
 * then expanded GS/8 entries at a time, one lane per child: a leaf node
 * force-accepts every child (recursiveScan, helpers.h:83-123 at recursion 1);
 * a branch force-accepts leaf children, M2Ls MAC-accepted branch children and
 * pushes the rest (nonRecursiveScan, helpers.h:125-147).  Pushes are ranked
 * with a ballot in (entry, child) order -- the reference's sequential order --
 * so a push whose rank reaches max_unchecked_pointers is force-accepted
 * exactly where the reference finds its list full.  Each lane keeps a
 * partial local expansion; the group reduces them with shuffles (a fixed
 * order, so results are reproducible), adds the incoming local expansion and
 * runs L2L (lane c -> child c) or L2P + P2P (lane k -> particle k). */
template <int GS, bool NEAR, bool PREC, int P, int MODE>
__device__ __forceinline__ void down_group_body(const DownArgs& a, int s, bool valid) {
    constexpr int NTP = nterms<P>();
    constexpr int GPB = 256 / GS;
    constexpr int EPR = GS / 8;
    const int tid = threadIdx.x;
    const int lane = tid % GS;
    const int wl = tid & 63;
    const int gbase = wl - lane;
    const uint64_t gmask = GS == 64 ? ~0ull : (((1ull << GS) - 1ull) << gbase);
    const uint64_t ltmask = wl == 0 ? 0ull : (~0ull >> (64 - wl));
    int id = 0;
    bool live = valid && a.flags[0] == 0u;
    if (live) {
        id = a.level_id[s];
        live = a.t.aux[3 * id + 1].w != 0.f;
    }
    const int4 inf = live ? a.t.info[id] : make_int4(0, 0, 0, 0);
    const int np = inf.x;
    const float4 f0 = live ? a.t.far[3 * id] : make_float4(0.f, 0.f, 0.f, 0.f);
    const NodeA A{f0.x, f0.y, f0.z, f0.w, live ? a.t.aux[3 * id].w : 0.f};
    const float4 l0 = live ? a.t.loc[4 * id] : make_float4(0.f, 0.f, 0.f, 0.f);
    LocalExp L;
    L.px = l0.x;
    L.py = l0.y;
    L.pz = l0.z;
    L.lx = L.ly = L.lz = 0.f;
    if constexpr (PREC) {
        if (live) {
            const float4 cl = a.t.clo[id];
            L.lx = cl.x;
            L.ly = cl.y;
            L.lz = cl.z;
        }
    }
    float Lc[eNT3]; /* compensations (precision form) */
#pragma unroll
    for (int k = 0; k < NTP; ++k) L.e[k] = Lc[k] = 0.f;
    const int32_t* in = a.in + (live ? (int64_t)a.t.parent_slot[id] * a.maxu : 0);
    /* MODE 1: branches only; MODE 2 and 3: leaves only, their M2L part and
     * then their pair sums (separate launches, so each has the registers of
     * its own path); MODE 4: leaves only, M2L and pair sums in one walk;
     * MODE 0: every node */
    if constexpr (MODE == 1) {
        if (np != 0) return;
    }
    if constexpr (MODE == 2 || MODE == 3 || MODE == 4) {
        static_assert(NEAR, "leaf launches are for the near-field form");
        constexpr int QCAP = 512;
        __shared__ int32_t near_q[GPB * QCAP];
        if (live && np != 0)
            down_near_leaf_q<GS, PREC, P, QCAP, MODE == 4 ? 0 : MODE - 1>(a, id, inf, A, L, l0, in, lane, wl, gmask,
                                                          near_q + (tid / GS) * QCAP);
        return;
    }
    const uint32_t outbase = (uint32_t)s * (uint32_t)a.maxu;
    int running = 0;
    if (live && np == 0) {
        if (lane == 0) a.out[outbase] = id;
        running = 1;
    }
    bool more = live;
    for (int base = 0; more; base += GS) {
        const int i = base + lane;
        const int e = i < a.maxu ? in[i] : -1;
        const uint64_t term = __ballot(e == -1) & gmask;
        const uint64_t first = term & (~term + 1ull);
        const bool valid = i < a.maxu && (term == 0ull || (1ull << wl) < first);
        more = term == 0ull && base + GS < a.maxu;
        bool acc = false;
        if (valid) {
            const float4 b0 = a.t.far[3 * e], b1 = a.t.far[3 * e + 1], b2 = a.t.far[3 * e + 2];
            acc = mac_ok(a, A, e, b0, b1, b2);
            if (acc) {
                if constexpr (PREC) m2l_add_comp<P>(a, L, Lc, e, b0, b1, b2);
                else m2l_add<P>(a, L, e, b0, b1, b2);
            }
        }
        uint64_t rej = __ballot(valid && !acc) & gmask;
        while (rej != 0ull) {
            int my_src = -1;
#pragma unroll
            for (int r = 0; r < EPR; ++r) {
                if (rej != 0ull) {
                    const int b = __ffsll((long long)rej) - 1;
                    rej &= rej - 1ull;
                    if (r == (lane >> 3)) my_src = b;
                }
            }
            const int ent = __shfl(e, (my_src < 0 ? wl : my_src) - gbase, GS);
            /* near-field form: a rejected leaf entry is pushed itself (by the
             * first lane of its eight) */
            const bool leaf_ent = NEAR && my_src >= 0 && a.t.info[ent].x != 0;
            int child = -1;
            if (leaf_ent) child = (lane & 7) == 0 ? ent : -1;
            else if (my_src >= 0) child = a.t.child[8 * ent + (lane & 7)];
            bool take = false, pushc = false;
            float4 c0, c1, c2;
            if (child >= 0 && child != id) {
                c0 = a.t.far[3 * child];
                c1 = a.t.far[3 * child + 1];
                c2 = a.t.far[3 * child + 2];
                if (leaf_ent) pushc = true;
                else if (!NEAR && (np != 0 || a.t.info[child].x != 0)) take = true;
                else if (mac_ok(a, A, child, c0, c1, c2)) take = true;
                else pushc = true;
            }
            const uint64_t pm = __ballot(pushc) & gmask;
            if (pushc) {
                const int pos = running + __popcll(pm & ltmask);
                if (pos >= a.maxu) take = true;
                else a.out[outbase + (uint32_t)pos] = child;
            }
            running += __popcll(pm);
            if (take) {
                if constexpr (PREC) m2l_add_comp<P>(a, L, Lc, child, c0, c1, c2);
                else m2l_add<P>(a, L, child, c0, c1, c2);
            }
        }
    }
    if (live && np == 0 && lane == 0 && running < a.maxu) a.out[outbase + (uint32_t)running] = -1;
#pragma unroll
    for (int k = 0; k < NTP; ++k) {
        L.e[k] = group_sum<GS>(L.e[k]);
        if constexpr (PREC) L.e[k] += group_sum<GS>(Lc[k]);
    }
    if (!live) return;
    add_incoming<P>(a, id, l0, L);
    if (np == 0) {
        if (lane < 8) l2l_child<P, PREC>(a, L, a.t.child[8 * id + lane]);
    } else {
        leaf_field_lanes<PREC, P>(a, L, inf, lane, GS);
    }
}

/* One level slot per group of GS lanes.  Every slot of the level: one per
 * group, workgroups XCD-remapped.  With an active-slot list (README rule,
 * most nodes inactive): a persistent grid whose XCD x walks the x-th eighth
 * of the list, so no workgroup is launched for an inactive node.  The
 * branch-only launch (MODE 1) walks the active branches alone and the leaf
 * launches (MODE 2-4) the active leaves alone, so that no group of a wave
 * idles on a node its launch skips. */
template <int GS, bool NEAR, bool PREC, int P, int MODE>
__device__ __forceinline__ void down_group_run(const DownArgs& a) {
    constexpr int GPB = 256 / GS;
    const int g = (int)threadIdx.x / GS;
    if (!a.act_slots) {
        const int blk = xcd_block((unsigned)((a.count + GPB - 1) / GPB));
        if (blk < 0) return;
        const int sl = blk * GPB + g;
        down_group_body<GS, NEAR, PREC, P, MODE>(a, sl, sl < a.count);
        return;
    }
    constexpr int SEL = MODE == 0 ? 0 : (MODE == 1 ? 1 : 2); /* both, branches, leaves */
    const int nl = a.act_count[0], nb = a.act_count[kMaxLevels];
    const int total = SEL == 0 ? nl + nb : (SEL == 1 ? nb : nl);
    const int nlog = (total + GPB - 1) / GPB;
    const int per = (nlog + kXcds - 1) / kXcds;
    const int xcd = (int)(blockIdx.x % kXcds), nper = (int)(gridDim.x / kXcds);
    for (int k = (int)(blockIdx.x / kXcds); k < per; k += nper) {
        const int lb = xcd * per + k;
        if (lb >= nlog) break;
        const int e = lb * GPB + g;
        const bool ok = e < total;
        const bool front = SEL == 2 || (SEL == 0 && e < nl);
        const int at = front ? e : a.count - 1 - (SEL == 0 ? e - nl : e);
        down_group_body<GS, NEAR, PREC, P, MODE>(a, ok ? a.act_slots[at] : 0, ok);
    }
}

/* Occupancy asked of the compiler, measured on MI355X: the as-written
 * second-order form at 6 waves per SIMD (95 -> 80 VGPRs, 36 bytes of
 * spills, +2% steps/s at 1e6), the third-order precision branches at 4
 * (130 -> 128 VGPRs, no spills). */
template <int GS, bool NEAR, bool PREC, int P, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((!NEAR && P == 2) ? 6 : ((PREC && P == 3) ? 4 : 1)))) void k_down_group(DownArgs a) {
    down_group_run<GS, NEAR, PREC, P, MODE>(a);
}

/* The slots of a level whose node is active (the up pass's activity flag),
 * appended a wave at a time: leaves from the front of the level's list
 * (counted in nact[0]), branches from its back (nact[kMaxLevels]).  The
 * list order varies from run to run, each node's result does not. */
__global__ __launch_bounds__(256) void k_level_active(const int32_t* __restrict__ level_id, int count,
                                                      const float4* __restrict__ aux, const int4* __restrict__ info,
                                                      int32_t* __restrict__ slots, int32_t* __restrict__ nact) {
    const int sl = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int id = sl < count ? level_id[sl] : 0;
    const bool act = sl < count && aux[3 * id + 1].w != 0.f;
    const bool leaf = act && info[id].x != 0;
    const uint64_t bl = __ballot(leaf), bb = __ballot(act && !leaf);
    const int lane = (int)(threadIdx.x & 63);
    const uint64_t lt = (1ull << lane) - 1ull;
    int basel = 0, baseb = 0;
    if (lane == 0 && bl != 0ull) basel = atomicAdd(nact, (int)__popcll(bl));
    if (lane == 0 && bb != 0ull) baseb = atomicAdd(nact + kMaxLevels, (int)__popcll(bb));
    basel = __shfl(basel, 0, 64);
    baseb = __shfl(baseb, 0, 64);
    if (leaf) slots[basel + (int)__popcll(bl & lt)] = sl;
    else if (act) slots[count - 1 - (baseb + (int)__popcll(bb & lt))] = sl;
}

/* The near-field leaves of a level, M2L part.  Two waves per SIMD are asked
 * for: the P = 3 precision expansions otherwise need more registers than
 * that allows and would run one wave per SIMD. */
template <int GS, bool PREC, int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_down_leaves(DownArgs a) {
    down_group_run<GS, true, PREC, P, 2>(a);
}

/* The near-field leaves of a level, pair sums and L2P (after k_down_leaves);
 * the third-order precision form at 3 waves per SIMD (174 -> 168 VGPRs, no
 * spills). */
template <int GS, bool PREC, int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PREC && P == 3 ? 3 : 1))) void k_down_leaves_p2p(DownArgs a) {
    down_group_run<GS, true, PREC, P, 3>(a);
}

/* The near-field leaves of a level in one walk (M2L, pair sums, L2P), for
 * the second-order FP32 form: there the whole walk fits in as few registers
 * as either half (4 waves per SIMD), so splitting it would only walk each
 * subtree twice.  The other forms split it (k_down_leaves, then
 * k_down_leaves_p2p); both give the same results, bit for bit. */
template <int GS, bool PREC, int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_down_leaves_walk(DownArgs a) {
    down_group_run<GS, true, PREC, P, 4>(a);
}

/* Level 0.  zeroFirstBranches (helpers.h:15-32); the root has node id 0.
 * A root that is itself a leaf (n <= max_children_in_leaf) makes the
 * reference read particle indices as child pointers; here its particles get
 * the intra-leaf P2P with a zero local expansion instead. */
template <int P>
__global__ void k_down_root(DownArgs a) {
    if (a.flags[0] != 0u) return;
    const Tree& t = a.t;
    const int4 inf = t.info[0];
    /* the root receives no local expansion: its record is defined as zero */
    for (int q = 0; q < 4; ++q) t.loc[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (P == 3)
        for (int q = 0; q < 3; ++q) t.loc3[q] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (inf.x != 0) {
        const float4 f0 = t.far[0];
        LocalExp L;
        L.px = f0.x; L.py = f0.y; L.pz = f0.z;
        const float4 cl = a.phi ? t.clo[0] : make_float4(0.f, 0.f, 0.f, 0.f);
        L.lx = cl.x; L.ly = cl.y; L.lz = cl.z;
        for (int k = 0; k < eNT3; ++k) L.e[k] = 0.f;
        leaf_field<P>(a, L, inf);
        a.out[0] = -1;
        return;
    }
    for (int c = 0; c < 8; ++c) {
        const int cid = t.child[c];
        if (cid < 0) continue;
        const float4 f0 = t.far[3 * cid];
        t.loc[4 * cid + 0] = make_float4(f0.x, f0.y, f0.z, 0.f);
        t.loc[4 * cid + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
        t.loc[4 * cid + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
        t.loc[4 * cid + 3] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (P == 3)
            for (int q = 0; q < 3; ++q) t.loc3[3 * cid + q] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    a.out[0] = 0;
    a.out[1] = -1;
}

Tree make_tree(sph_ctx* ctx) {
    Tree t;
    t.octree = ctx->tree.octree;
    t.info = ctx->tree.node_info;
    t.child = ctx->tree.node_child;
    t.parent_slot = ctx->tree.parent_slot;
    t.sub_end = ctx->tree.subtree_end;
    t.far = ctx->tree.mp_far;
    t.aux = ctx->tree.mp_aux;
    t.loc = ctx->tree.loc;
    t.oct = ctx->tree.mp_oct;
    t.loc3 = ctx->tree.loc3;
    t.loc_tot = ctx->tree.loc_tot;
    t.near_rec = ctx->tree.near_rec;
    t.near_cnt = ctx->tree.near_cnt;
    t.near_cap = ctx->tree.near_rec ? ctx->prm.near_list_cap : 0;
    t.clo = ctx->tree.mp_clo;
    return t;
}

}  // namespace

int launch_tree_convert(sph_ctx* ctx, hipStream_t st) {
    sph_tree_dev& T = ctx->tree;
    const int threads = 256;
    const int K = T.node_count;
    k_tree_convert_init<<<(K + threads - 1) / threads, threads, 0, st>>>(T.parent_slot, T.parent_level, T.seen, K,
                                                                           ctx->grav_flags);
    SPH_HIP_TRY(hipGetLastError());
    ConvertArgs a;
    a.octree = T.octree;
    a.n_ints = T.n_ints;
    a.level_tp = T.level_tp;
    a.level_id = T.level_id;
    a.n_levels = (int)T.level_counts.size();
    a.node_count = K;
    a.n_particles = ctx->n;
    a.info = T.node_info;
    a.child = T.node_child;
    a.parent_slot = T.parent_slot;
    a.parent_level = T.parent_level;
    a.seen = T.seen;
    a.flags = ctx->grav_flags;
    /* one launch per kMaxLevels levels: one for any tree the library's builders make */
    for (int l0 = 0; l0 < a.n_levels; l0 += kMaxLevels) {
        a.first_level = l0;
        a.chunk_levels = std::min(kMaxLevels, a.n_levels - l0);
        for (int l = 0; l < a.chunk_levels; ++l) a.off[l] = T.level_offsets[l0 + l];
        a.off[a.chunk_levels] = a.off[a.chunk_levels - 1] + T.level_counts[l0 + a.chunk_levels - 1];
        const int count = a.off[a.chunk_levels] - a.off[0];
        if (count == 0) continue;
        k_tree_convert<<<(count + threads - 1) / threads, threads, 0, st>>>(a);
        SPH_HIP_TRY(hipGetLastError());
    }
    k_tree_check<<<(K + threads - 1) / threads, threads, 0, st>>>(K, T.node_info, T.parent_slot, T.parent_level,
                                                                 T.seen, ctx->grav_flags);
    SPH_HIP_TRY(hipGetLastError());
    return SPH_OK;
}

int launch_gravity_up(sph_ctx* ctx) {
    sph_tree_dev& T = ctx->tree;
    UpArgs a;
    a.t = make_tree(ctx);
    a.flags = ctx->grav_flags;
    a.pos = ctx->pos;
    a.m = ctx->m;
    a.next_active = ctx->next_active;
    a.gravnorm = ctx->gravnorm;
    a.active = ctx->active;
    a.time = ctx->time;
    a.readme_rule = ctx->prm.activity_rule == SPH_ACTIVE_README;
    a.fix_zz = (ctx->prm.fixes & SPH_FIX_M2M_ZZ) ? 1 : 0;
    a.fix_range = (ctx->prm.fixes & SPH_FIX_SYNC_RANGE) ? 1 : 0;
    a.need_preorder = (ctx->prm.fixes & SPH_FIX_NEAR_FIELD) ? 1 : 0;
    a.prec = ctx->prm.precision ? 1 : 0;
    a.order = ctx->prm.expansion_order == 3 ? 3 : 2;
    a.fix_p3 = (ctx->prm.fixes & SPH_FIX_ORDER3) ? 1 : 0;
    const int threads = 256;
    for (int l = (int)T.level_counts.size() - 1; l >= 0; --l) {
        a.count = T.level_counts[l];
        if (a.count == 0) continue;
        a.level_id = T.level_id + T.level_offsets[l];
        if (a.prec) k_up<true><<<(a.count + threads - 1) / threads, threads, 0, ctx->stream>>>(a);
        else k_up<false><<<(a.count + threads - 1) / threads, threads, 0, ctx->stream>>>(a);
        SPH_HIP_TRY(hipGetLastError());
    }
    return SPH_OK;
}

/* One level of the near-field form with GS lanes per node: branches (list
 * processing, L2L), then the leaves (balanced walk, pair sums, L2P) in one
 * or two launches; persistent over the active list, at most kDownBlocks
 * workgroups, when there is one.  The near-field form runs a wave (64
 * lanes) per node: each node's walk is a chain of dependent rounds, and
 * twice the lanes halve the rounds (measured against 32 lanes: 1e6 second
 * order 107.3 -> 113.4 steps/s, third-order precision 33.9 -> 37.1, 4e6
 * learned-h + precision with individual timesteps 224.0 -> 255.0). */
template <int GS>
static void launch_near_level(const DownArgs& a, bool compact, bool order3, hipStream_t st) {
    const int threads = 256;
    const unsigned blocks = (unsigned)(((int64_t)a.count * GS + threads - 1) / threads);
    const dim3 g(compact ? std::min(xcd_grid(blocks), (unsigned)kDownBlocks) : xcd_grid(blocks));
    if (order3) {
        if (a.phi) {
            k_down_group<GS, true, true, 3, 1><<<g, threads, 0, st>>>(a);
            k_down_leaves<GS, true, 3><<<g, threads, 0, st>>>(a);
            k_down_leaves_p2p<GS, true, 3><<<g, threads, 0, st>>>(a);
        } else {
            k_down_group<GS, true, false, 3, 1><<<g, threads, 0, st>>>(a);
            k_down_leaves<GS, false, 3><<<g, threads, 0, st>>>(a);
            k_down_leaves_p2p<GS, false, 3><<<g, threads, 0, st>>>(a);
        }
    } else {
        if (a.phi) {
            k_down_group<GS, true, true, 2, 1><<<g, threads, 0, st>>>(a);
            k_down_leaves<GS, true, 2><<<g, threads, 0, st>>>(a);
            k_down_leaves_p2p<GS, true, 2><<<g, threads, 0, st>>>(a);
        } else {
            k_down_group<GS, true, false, 2, 1><<<g, threads, 0, st>>>(a);
            k_down_leaves_walk<GS, false, 2><<<g, threads, 0, st>>>(a);
        }
    }
}

int launch_gravity_down(sph_ctx* ctx, bool bootstrap) {
    sph_tree_dev& T = ctx->tree;
    const sph_params& p = ctx->prm;
    DownArgs a;
    a.t = make_tree(ctx);
    a.flags = ctx->grav_flags;
    a.maxu = p.max_unchecked_pointers;
    a.max_rec = p.gravity_max_recursion;
    a.eta = p.gravity_eta;
    a.eps = p.gravity_softening;
    a.G = p.gravity_g;
    a.fix_soft = (p.fixes & SPH_FIX_SOFTENING) ? 1 : 0;
    a.near = (p.fixes & SPH_FIX_NEAR_FIELD) && !bootstrap ? 1 : 0;
    /* the precision form (near field required, checked at creation) also
     * writes the potential */
    a.phi = (p.precision && a.near) ? ctx->phi : nullptr;
    const bool order3 = p.expansion_order == 3 && !bootstrap;
    a.macp = (order3 && p.mac_order != 2) ? 3 : 2;
    a.fix_p3 = (p.fixes & SPH_FIX_ORDER3) ? 1 : 0;
    a.pos = ctx->pos;
    a.m = ctx->m;
    a.active = ctx->active;
    a.acc = ctx->acc;
    a.gravnorm = ctx->gravnorm;
    a.act_slots = nullptr;
    a.act_count = nullptr;
    const int threads = 256;
    /* under the README rule most nodes are inactive: the group launches run
     * over lists of the active slots of each level instead */
    const bool compact = p.activity_rule == SPH_ACTIVE_README && (a.max_rec == 1 || a.near) &&
                         T.level_counts.size() <= (size_t)kMaxLevels;
    if (compact) {
        k_fill_i32<<<1, 2 * kMaxLevels, 0, ctx->stream>>>(T.act_count, 2 * kMaxLevels, 0);
        SPH_HIP_TRY(hipGetLastError());
        for (size_t l = 1; l < T.level_counts.size(); ++l) {
            const int cnt = T.level_counts[l];
            if (cnt == 0) continue;
            k_level_active<<<(cnt + threads - 1) / threads, threads, 0, ctx->stream>>>(
                T.level_id + T.level_offsets[l], cnt, T.mp_aux, T.node_info, T.act_slots + T.level_offsets[l],
                T.act_count + l);
            SPH_HIP_TRY(hipGetLastError());
        }
    }
    for (size_t l = 0; l < T.level_counts.size(); ++l) {
        int32_t* in = (l % 2 == 0) ? ctx->list_a : ctx->list_b;
        int32_t* out = (l % 2 == 0) ? ctx->list_b : ctx->list_a;
        if (l == 0) {
            a.out = out;
            if (order3) k_down_root<3><<<1, 1, 0, ctx->stream>>>(a);
            else k_down_root<2><<<1, 1, 0, ctx->stream>>>(a);
            SPH_HIP_TRY(hipGetLastError());
            continue;
        }
        a.count = T.level_counts[l];
        if (a.count == 0) continue;
        a.level_id = T.level_id + T.level_offsets[l];
        a.in = in;
        a.out = out;
        /* the near-field form has no recursion limit (leaves walk whole subtrees) */
        if (a.max_rec == 1 || a.near) {
            if (compact) {
                a.act_slots = T.act_slots + T.level_offsets[l];
                a.act_count = T.act_count + l;
            }
            if (a.near) {
                launch_near_level<64>(a, compact, order3, ctx->stream);
            } else {
                /* as written: 32 lanes per node (64 measured slower at 1e6:
                 * 144.7 steps/s at every level, 151.0 at the levels of fewer
                 * than 8192 nodes, against 153.4) */
                constexpr int GS = 32;
                const unsigned blocks = (unsigned)(((int64_t)a.count * GS + threads - 1) / threads);
                const dim3 g(compact ? std::min(xcd_grid(blocks), (unsigned)kDownBlocks) : xcd_grid(blocks));
                if (order3) k_down_group<GS, false, false, 3, 0><<<g, threads, 0, ctx->stream>>>(a);
                else k_down_group<GS, false, false, 2, 0><<<g, threads, 0, ctx->stream>>>(a);
            }
        } else {
            a.act_slots = nullptr;
            a.act_count = nullptr;
            k_down<<<(a.count + threads - 1) / threads, threads, 0, ctx->stream>>>(a);
        }
        SPH_HIP_TRY(hipGetLastError());
    }
    return SPH_OK;
}
