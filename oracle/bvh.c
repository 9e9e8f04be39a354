/*
 * ORACLE -- test infrastructure only (never linked into alpine_amd).
 *
 * Plain-C restatement of the reference's binary BVH.  Every function cites the
 * reference lines it follows.  Float expressions keep the reference's operand
 * order; the file must be compiled with -ffp-contract=off and without
 * -ffast-math so each operation rounds once, as in the reference's default
 * x86-64 (SSE2, no FMA) build.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } f3;
typedef struct { f3 min, max; } box3;

/* std::min / std::max semantics (vector.h:224-248): min(a,b) = (b < a) ? b : a,
 * max(a,b) = (a < b) ? b : a. */
static inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_std(float a, float b) { return (a < b) ? b : a; }

static inline f3 mk3(float x, float y, float z) { f3 r = { x, y, z }; return r; }
static inline f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 muls3(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline f3 divs3(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline float comp3(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
/* vector.h:186-190 */
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* vector.h:198-206 */
static inline f3 cross3(f3 a, f3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* vector.h:192-212: v / sqrt(dot(v, v)) */
static inline f3 normalize3(f3 v) { return divs3(v, sqrtf(dot3(v, v))); }
static inline f3 min3(f3 a, f3 b) { return mk3(fmin_std(a.x, b.x), fmin_std(a.y, b.y), fmin_std(a.z, b.z)); }
static inline f3 max3(f3 a, f3 b) { return mk3(fmax_std(a.x, b.x), fmax_std(a.y, b.y), fmax_std(a.z, b.z)); }
/* vector.h:244-248 */
static inline float maxcomp3(f3 v) { return fmax_std(v.x, fmax_std(v.y, v.z)); }

/* bounding_box.h:12-13 */
static inline box3 box_empty(void)
{
    box3 b = { { FLT_MAX, FLT_MAX, FLT_MAX }, { -FLT_MAX, -FLT_MAX, -FLT_MAX } };
    return b;
}
/* bounding_box.h:59-67 */
static inline box3 box_merge(box3 a, box3 b) { box3 r = { min3(a.min, b.min), max3(a.max, b.max) }; return r; }
/* bounding_box.h:69-77 */
static inline box3 box_merge_pt(box3 a, f3 v) { box3 r = { min3(a.min, v), max3(a.max, v) }; return r; }
/* bounding_box.h:15 */
static inline f3 box_center(box3 b) { return muls3(add3(b.min, b.max), 0.5f); }
/* bounding_box.h:17-21 */
static inline float box_sa(box3 b)
{
    f3 len = sub3(b.max, b.min);
    return (len.x * len.y + len.y * len.z + len.z * len.x) * 2.0f;
}

/* util.h:8-15: MACHINE_EPSILON = 0.5 * FLT_EPSILON, gamma(n) = n*eps / (1 - n*eps) */
static const float MACHINE_EPS = 0.5f * FLT_EPSILON;
static inline float gamma_n(int n) { return ((float)n * MACHINE_EPS) / (1.0f - (float)n * MACHINE_EPS); }

/* bounding_box.h:23-56 */
static int box_intersect(const box3* bb, f3 org, float tfar, f3 inv)
{
    const float correction = 1.0f + 2.0f * gamma_n(3);

    float tMin = ((inv.x >= 0.0f ? bb->min.x : bb->max.x) - org.x) * inv.x;
    float tMax = ((inv.x >= 0.0f ? bb->max.x : bb->min.x) - org.x) * inv.x;
    tMax *= correction;

    float tyMin = ((inv.y >= 0.0f ? bb->min.y : bb->max.y) - org.y) * inv.y;
    float tyMax = ((inv.y >= 0.0f ? bb->max.y : bb->min.y) - org.y) * inv.y;
    tyMax *= correction;

    if (tMin > tyMax || tMax < tyMin) return 0;

    tMin = fmax_std(tMin, tyMin);
    tMax = fmin_std(tMax, tyMax);

    float tzMin = ((inv.z >= 0.0f ? bb->min.z : bb->max.z) - org.z) * inv.z;
    float tzMax = ((inv.z >= 0.0f ? bb->max.z : bb->min.z) - org.z) * inv.z;
    tzMax *= correction;

    if (tMin > tzMax || tMax < tzMin) return 0;

    tMin = fmax_std(tMin, tzMin);
    tMax = fmin_std(tMax, tzMax);

    return tMin < tfar && tMax > 0.0f;
}

/* Primitive (bvh_common.h:110-125) minus the Shape pointer. */
typedef struct {
    uint32_t prim_id;
    f3 v0, e0, e1, ng;
    box3 bbox;
} prim_t;

/* Primitive::updateVertices (bvh_common.cpp:105-125) */
static void prim_update(prim_t* p, const float* xyz, const uint32_t* tri)
{
    f3 v[3];
    for (int i = 0; i < 3; ++i) v[i] = mk3(xyz[3 * tri[i]], xyz[3 * tri[i] + 1], xyz[3 * tri[i] + 2]);
    p->v0 = v[0];
    p->e0 = sub3(v[1], v[0]);
    p->e1 = sub3(v[2], v[0]);
    p->ng = normalize3(cross3(p->e0, p->e1));
    p->bbox = box_empty();
    for (int i = 0; i < 3; ++i) p->bbox = box_merge_pt(p->bbox, v[i]);
}

/* Primitive::intersect (bvh_common.cpp:66-103) */
static int prim_intersect(const prim_t* p, f3 org, f3 dir, float* t, float* b0, float* b1)
{
    f3 s = sub3(org, p->v0);
    f3 s1 = cross3(dir, p->e1);
    f3 s2 = cross3(s, p->e0);
    float det = dot3(s1, p->e0);

    if (fabsf(det) < FLT_EPSILON) return 0;

    float tt = dot3(s2, p->e1) / det;
    if (tt < 0.0f) return 0;

    float bb0 = dot3(s1, s) / det;
    if (bb0 < 0.0f || bb0 > 1.0f) return 0;

    float bb1 = dot3(s2, dir) / det;
    if (bb1 < 0.0f || bb1 > 1.0f - bb0) return 0;

    *t = tt;
    *b0 = bb0;
    *b1 = bb1;
    return 1;
}

/* bvh_util.h:6-12 */
static inline uint8_t bin_index(float c, float mn, float mx, uint8_t count)
{
    uint8_t bin = (uint8_t)((float)count * (c - mn) / (mx - mn));
    uint8_t last = (uint8_t)(count - 1);
    return bin < last ? bin : last;
}

/* BuildNode (bvh_common.h:127-136) */
typedef struct bnode {
    box3 bbox;
    struct bnode* child[2];
    uint32_t offset;
    uint32_t count; /* uint16 in the reference; > 0 marks a leaf */
    uint8_t dim;
} bnode;

/* BuildPrimitive (bvh_common.cpp:131-135) */
typedef struct { box3 bbox; uint32_t index; } bprim;

enum { SPLITS_PER_DIM = 8, BIN_COUNT = SPLITS_PER_DIM + 1 };

typedef struct {
    int valid;
    uint8_t dim, split_idx, bin_count;
    float mn, mx;
} split_t;

/* bvh_util.h:100-104 */
static inline int split_is_below(const split_t* s, f3 c)
{
    return bin_index(comp3(c, s->dim), s->mn, s->mx, s->bin_count) <= s->split_idx;
}

/* findSplit (bvh_common.cpp:185-272) */
static split_t find_split(const bprim* bp, uint32_t n, float* out_cost)
{
    split_t split = { 0, 0, 0, 0, 0.0f, 0.0f };
    box3 cb = box_empty();
    for (uint32_t i = 0; i < n; ++i) {
        f3 c = box_center(bp[i].bbox);
        box3 cc = { c, c };
        cb = box_merge(cb, cc);
    }
    float diag_max = maxcomp3(sub3(cb.max, cb.min));
    if (diag_max == 0.0f) {
        *out_cost = 0.0f;
        return split;
    }

    /* The reference's std::optional<Split> holds a default Split (dim 0,
     * split 0, binCount 0) when no cost beats FLT_MAX; it is still "valid". */
    split.valid = 1;
    float min_cost = FLT_MAX;
    for (uint8_t dim = 0; dim < 3; ++dim) {
        float mn = comp3(cb.min, dim), mx = comp3(cb.max, dim);
        if (mn == mx) continue;

        box3 bin_box[BIN_COUNT];
        uint32_t bin_cnt[BIN_COUNT];
        for (int b = 0; b < BIN_COUNT; ++b) { bin_box[b] = box_empty(); bin_cnt[b] = 0; }

        for (uint32_t i = 0; i < n; ++i) {
            float c = comp3(box_center(bp[i].bbox), dim);
            uint8_t bi = bin_index(c, mn, mx, BIN_COUNT);
            if (bi > BIN_COUNT - 1) bi = BIN_COUNT - 1;
            bin_box[bi] = box_merge(bin_box[bi], bp[i].bbox);
            bin_cnt[bi]++;
        }

        float costs[SPLITS_PER_DIM];
        for (int s = 0; s < SPLITS_PER_DIM; ++s) costs[s] = 0.0f;

        box3 acc = box_empty();
        uint32_t acc_n = 0;
        for (int s = 0; s < SPLITS_PER_DIM; ++s) {
            acc = box_merge(acc, bin_box[s]);
            acc_n += bin_cnt[s];
            costs[s] += (float)acc_n * box_sa(acc);
        }
        acc = box_empty();
        acc_n = 0;
        for (int s = SPLITS_PER_DIM - 1; s >= 0; --s) {
            acc = box_merge(acc, bin_box[s + 1]);
            acc_n += bin_cnt[s + 1];
            costs[s] += (float)acc_n * box_sa(acc);
        }
        for (int s = 0; s < SPLITS_PER_DIM; ++s) {
            if (costs[s] < min_cost) {
                min_cost = costs[s];
                split.dim = dim;
                split.split_idx = (uint8_t)s;
                split.bin_count = BIN_COUNT;
                split.mn = mn;
                split.mx = mx;
            }
        }
    }
    *out_cost = min_cost;
    return split;
}

/* libstdc++ std::partition for bidirectional iterators (GCC 11
 * bits/stl_algo.h __partition, bidirectional_iterator_tag), which is what
 * std::partition over std::span iterators dispatches to. */
static uint32_t partition_bidir(bprim* a, uint32_t n, const split_t* s)
{
    uint32_t first = 0, last = n;
    for (;;) {
        for (;;) {
            if (first == last) return first;
            if (split_is_below(s, box_center(a[first].bbox))) ++first;
            else break;
        }
        --last;
        for (;;) {
            if (first == last) return first;
            if (!split_is_below(s, box_center(a[last].bbox))) --last;
            else break;
        }
        bprim tmp = a[first];
        a[first] = a[last];
        a[last] = tmp;
        ++first;
    }
}

typedef struct {
    bnode* pool;
    uint32_t pool_used, pool_cap;
    uint32_t offset;
    uint32_t leaf_threshold;
    const prim_t* prims;
    uint32_t* ordered; /* ordered index -> original primitive index */
} builder_t;

static bnode* alloc_node(builder_t* B)
{
    bnode* p = &B->pool[B->pool_used++];
    memset(p, 0, sizeof(*p));
    return p;
}

/* createLeaf (bvh_common.cpp:345-366).  Offsets come from a sequential
 * counter, i.e. the order a single-threaded build would reach the leaves;
 * the reference's fetch_add order differs between runs but the flattened
 * tree, traversal order and leaf contents do not depend on it. */
static bnode* create_leaf(builder_t* B, const bprim* bp, uint32_t n, box3 bbox)
{
    bnode* leaf = alloc_node(B);
    leaf->bbox = bbox;
    leaf->count = n;
    leaf->offset = B->offset;
    B->offset += n;
    for (uint32_t i = 0; i < n; ++i) B->ordered[leaf->offset + i] = bp[i].index;
    return leaf;
}

/* buildNode (bvh_common.cpp:274-343) */
static bnode* build_node(builder_t* B, bprim* bp, uint32_t n)
{
    box3 nb = box_empty();
    for (uint32_t i = 0; i < n; ++i) nb = box_merge(nb, bp[i].bbox);

    if (n <= B->leaf_threshold) return create_leaf(B, bp, n, nb);

    float split_cost;
    split_t s = find_split(bp, n, &split_cost);
    float leaf_cost = (float)n * box_sa(nb);
    if (!s.valid || split_cost >= leaf_cost) return create_leaf(B, bp, n, nb);

    uint32_t mid = partition_bidir(bp, n, &s);

    bnode* node = alloc_node(B);
    node->bbox = nb;
    node->dim = s.dim;
    node->child[0] = build_node(B, bp, mid);
    node->child[1] = build_node(B, bp + mid, n - mid);
    return node;
}

/* LinearNode (bvh.cpp:51-59) */
typedef struct {
    box3 bbox;
    uint32_t offset;
    uint32_t count;
    uint8_t dim;
} lnode;

struct oracle_bvh {
    uint32_t nt;
    prim_t* prims;          /* original order */
    prim_t* ordered_prims;  /* leaf order */
    uint32_t* ordered;      /* ordered -> original */
    lnode* nodes;
    uint32_t node_count;
    bnode* pool;
    bnode* root;
};

/* flatten (bvh.cpp:144-189): depth-first pre-order, left child at idx+1. */
static void flatten(oracle_bvh* b)
{
    typedef struct { const bnode* node; uint32_t* offset; } info_t;
    uint32_t cap = 1024;
    info_t* stack = (info_t*)malloc(cap * sizeof(info_t));
    uint32_t sp = 0;
    stack[sp].node = b->root;
    stack[sp].offset = NULL;
    sp++;
    b->node_count = 0;
    while (sp) {
        info_t it = stack[--sp];
        if (it.offset) *it.offset = b->node_count;
        lnode* ln = &b->nodes[b->node_count++];
        ln->bbox = it.node->bbox;
        if (it.node->count > 0) {
            ln->offset = it.node->offset;
            ln->count = it.node->count;
            ln->dim = 0;
        } else {
            ln->dim = it.node->dim;
            ln->count = 0;
            if (sp + 2 > cap) { cap *= 2; stack = (info_t*)realloc(stack, cap * sizeof(info_t)); }
            stack[sp].node = it.node->child[1];
            stack[sp].offset = &ln->offset;
            sp++;
            stack[sp].node = it.node->child[0];
            stack[sp].offset = NULL;
            sp++;
        }
    }
    free(stack);
}

oracle_bvh* oracle_bvh_create(const float* xyz, uint32_t nv, const uint32_t* idx,
                              uint32_t nt, uint32_t leaf_threshold)
{
    (void)nv;
    oracle_bvh* b = (oracle_bvh*)calloc(1, sizeof(oracle_bvh));
    b->nt = nt;
    if (nt == 0) return b;
    b->prims = (prim_t*)malloc((size_t)nt * sizeof(prim_t));
    for (uint32_t i = 0; i < nt; ++i) {
        b->prims[i].prim_id = i;
        prim_update(&b->prims[i], xyz, idx + 3 * (size_t)i);
    }
    /* buildBvh (bvh_common.cpp:369-391) */
    bprim* bp = (bprim*)malloc((size_t)nt * sizeof(bprim));
    for (uint32_t i = 0; i < nt; ++i) { bp[i].index = i; bp[i].bbox = b->prims[i].bbox; }

    builder_t B;
    B.pool_cap = 2 * nt + 1;
    B.pool = (bnode*)malloc((size_t)B.pool_cap * sizeof(bnode));
    B.pool_used = 0;
    B.offset = 0;
    B.leaf_threshold = leaf_threshold;
    B.prims = b->prims;
    b->ordered = (uint32_t*)malloc((size_t)nt * sizeof(uint32_t));
    B.ordered = b->ordered;
    b->root = build_node(&B, bp, nt);
    b->pool = B.pool;
    free(bp);

    b->ordered_prims = (prim_t*)malloc((size_t)nt * sizeof(prim_t));
    for (uint32_t i = 0; i < nt; ++i) b->ordered_prims[i] = b->prims[b->ordered[i]];

    b->nodes = (lnode*)malloc((size_t)B.pool_used * sizeof(lnode));
    flatten(b);
    return b;
}

void oracle_bvh_destroy(oracle_bvh* b)
{
    if (!b) return;
    free(b->prims);
    free(b->ordered_prims);
    free(b->ordered);
    free(b->nodes);
    free(b->pool);
    free(b);
}

/* traverse (bvh.cpp:243-319) */
static int traverse(const oracle_bvh* b, f3 org, f3 dir, float tfar, int any,
                    uint32_t* prim, float* t, float* b0, float* b1)
{
    if (b->node_count == 0) return 0;
    int found = 0;
    float tNear = tfar;
    f3 inv = mk3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    uint32_t stack[64];
    uint8_t sp = 0;
    uint32_t cur = 0;
    for (;;) {
        const lnode* ln = &b->nodes[cur];
        if (box_intersect(&ln->bbox, org, tNear, inv)) {
            if (ln->count == 0) {
                int neg = comp3(dir, ln->dim) < 0.0f;
                if (neg) { stack[sp++] = cur + 1; cur = ln->offset; }
                else { stack[sp++] = ln->offset; cur = cur + 1; }
                continue;
            }
            for (uint32_t i = 0; i < ln->count; ++i) {
                const prim_t* p = &b->ordered_prims[ln->offset + i];
                float tt, bb0, bb1;
                if (!prim_intersect(p, org, dir, &tt, &bb0, &bb1) || tt >= tNear) continue;
                *prim = p->prim_id;
                *t = tt;
                *b0 = bb0;
                *b1 = bb1;
                found = 1;
                if (any) return 1;
                tNear = tt;
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    return found;
}

typedef struct {
    const oracle_bvh* b;
    const float *org, *dir, *tfar;
    uint32_t n, begin, end;
    uint32_t* prim;
    float *t, *b0, *b1;
    uint8_t* occ;
    int any;
} job_t;

static void run_range(job_t* j)
{
    for (uint32_t i = j->begin; i < j->end; ++i) {
        f3 o = mk3(j->org[3 * i], j->org[3 * i + 1], j->org[3 * i + 2]);
        f3 d = mk3(j->dir[3 * i], j->dir[3 * i + 1], j->dir[3 * i + 2]);
        if (j->any) {
            uint32_t p; float t, b0, b1;
            j->occ[i] = (uint8_t)traverse(j->b, o, d, j->tfar ? j->tfar[i] : FLT_MAX, 1, &p, &t, &b0, &b1);
        } else {
            uint32_t p = 0xFFFFFFFFu; float t = FLT_MAX, b0 = 0.0f, b1 = 0.0f;
            traverse(j->b, o, d, FLT_MAX, 0, &p, &t, &b0, &b1);
            j->prim[i] = p; j->t[i] = t; j->b0[i] = b0; j->b1[i] = b1;
        }
    }
}

static void* run_thread(void* arg) { run_range((job_t*)arg); return NULL; }

static void run_jobs(job_t proto, uint32_t nthreads)
{
    if (nthreads <= 1 || proto.n < 1024) {
        proto.begin = 0;
        proto.end = proto.n;
        run_range(&proto);
        return;
    }
    /* 16 K-ray chunks handed out round-robin to the threads, like the
     * reference's tile tasks (alpine.cpp:276-285). */
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    job_t jobs[256];
    uint32_t chunk = (proto.n + nthreads - 1) / nthreads;
    for (uint32_t k = 0; k < nthreads; ++k) {
        jobs[k] = proto;
        jobs[k].begin = k * chunk < proto.n ? k * chunk : proto.n;
        jobs[k].end = (k + 1) * chunk < proto.n ? (k + 1) * chunk : proto.n;
        pthread_create(&th[k], NULL, run_thread, &jobs[k]);
    }
    for (uint32_t k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
}

void oracle_bvh_intersect(const oracle_bvh* b, const float* org, const float* dir,
                          uint32_t n, uint32_t* prim, float* t, float* b0, float* b1,
                          uint32_t nthreads)
{
    job_t j;
    memset(&j, 0, sizeof(j));
    j.b = b; j.org = org; j.dir = dir; j.n = n;
    j.prim = prim; j.t = t; j.b0 = b0; j.b1 = b1; j.any = 0;
    run_jobs(j, nthreads);
}

void oracle_bvh_occluded(const oracle_bvh* b, const float* org, const float* dir,
                         const float* tfar, uint32_t n, uint8_t* occluded, uint32_t nthreads)
{
    job_t j;
    memset(&j, 0, sizeof(j));
    j.b = b; j.org = org; j.dir = dir; j.tfar = tfar; j.n = n; j.occ = occluded; j.any = 1;
    run_jobs(j, nthreads);
}

void oracle_brute_intersect(const oracle_bvh* b, const float* org, const float* dir,
                            uint32_t n, uint32_t* prim, float* t, float* b0, float* b1)
{
    for (uint32_t i = 0; i < n; ++i) {
        f3 o = mk3(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
        f3 d = mk3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        uint32_t bp = 0xFFFFFFFFu;
        float bt = FLT_MAX, bb0 = 0.0f, bb1 = 0.0f;
        for (uint32_t k = 0; k < b->nt; ++k) {
            float tt, x0, x1;
            if (prim_intersect(&b->prims[k], o, d, &tt, &x0, &x1) && tt < bt) {
                bp = k; bt = tt; bb0 = x0; bb1 = x1;
            }
        }
        prim[i] = bp; t[i] = bt; b0[i] = bb0; b1[i] = bb1;
    }
}

void oracle_brute_occluded(const oracle_bvh* b, const float* org, const float* dir,
                           const float* tfar, uint32_t n, uint8_t* occluded)
{
    for (uint32_t i = 0; i < n; ++i) {
        f3 o = mk3(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
        f3 d = mk3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        float tf = tfar ? tfar[i] : FLT_MAX;
        uint8_t hit = 0;
        for (uint32_t k = 0; k < b->nt && !hit; ++k) {
            float tt, x0, x1;
            if (prim_intersect(&b->prims[k], o, d, &tt, &x0, &x1) && tt < tf) hit = 1;
        }
        occluded[i] = hit;
    }
}

int oracle_triangle_intersect(const oracle_bvh* b, uint32_t tri, const float* org,
                              const float* dir, float* t, float* b0, float* b1)
{
    return prim_intersect(&b->prims[tri], mk3(org[0], org[1], org[2]),
                          mk3(dir[0], dir[1], dir[2]), t, b0, b1);
}

static void stats_rec(const bnode* n, uint32_t depth, oracle_bvh_stats* s, double* inner,
                      double* leaf)
{
    if (depth > s->max_depth) s->max_depth = depth;
    double sa = (double)box_sa(n->bbox);
    if (n->count > 0) {
        s->leaf_count++;
        if (n->count > s->max_leaf_size) s->max_leaf_size = n->count;
        *leaf += sa * (double)n->count;
        return;
    }
    *inner += sa;
    stats_rec(n->child[0], depth + 1, s, inner, leaf);
    stats_rec(n->child[1], depth + 1, s, inner, leaf);
}

void oracle_bvh_get_stats(const oracle_bvh* b, oracle_bvh_stats* s)
{
    memset(s, 0, sizeof(*s));
    if (!b->root) return;
    s->node_count = b->node_count;
    double inner = 0.0, leaf = 0.0;
    stats_rec(b->root, 0, s, &inner, &leaf);
    double root_sa = (double)box_sa(b->root->bbox);
    s->sah_cost = (inner + leaf) / root_sa;
    s->sah_cost_leaf = leaf / root_sa;
}

void oracle_bvh_leaf_keys(const oracle_bvh* b, uint32_t* key)
{
    for (uint32_t i = 0; i < b->node_count; ++i) {
        const lnode* ln = &b->nodes[i];
        if (ln->count == 0) continue;
        uint32_t mn = 0xFFFFFFFFu;
        for (uint32_t k = 0; k < ln->count; ++k) {
            uint32_t o = b->ordered[ln->offset + k];
            if (o < mn) mn = o;
        }
        for (uint32_t k = 0; k < ln->count; ++k) key[b->ordered[ln->offset + k]] = mn;
    }
}

void oracle_bvh_ordered_prims(const oracle_bvh* b, uint32_t* order)
{
    memcpy(order, b->ordered, (size_t)b->nt * sizeof(uint32_t));
}

void oracle_bvh_nodes(const oracle_bvh* b, float* bbox6, uint32_t* offset, uint16_t* count,
                      uint8_t* dim)
{
    for (uint32_t i = 0; i < b->node_count; ++i) {
        const lnode* ln = &b->nodes[i];
        bbox6[6 * i + 0] = ln->bbox.min.x; bbox6[6 * i + 1] = ln->bbox.min.y;
        bbox6[6 * i + 2] = ln->bbox.min.z; bbox6[6 * i + 3] = ln->bbox.max.x;
        bbox6[6 * i + 4] = ln->bbox.max.y; bbox6[6 * i + 5] = ln->bbox.max.z;
        offset[i] = ln->offset;
        count[i] = (uint16_t)ln->count;
        dim[i] = ln->dim;
    }
}

int oracle_find_split(const float* boxes6, uint32_t n, uint32_t* dim, uint32_t* split_idx,
                      float* cost)
{
    bprim* bp = (bprim*)malloc((size_t)(n ? n : 1) * sizeof(bprim));
    for (uint32_t i = 0; i < n; ++i) {
        bp[i].index = i;
        bp[i].bbox.min = mk3(boxes6[6 * i], boxes6[6 * i + 1], boxes6[6 * i + 2]);
        bp[i].bbox.max = mk3(boxes6[6 * i + 3], boxes6[6 * i + 4], boxes6[6 * i + 5]);
    }
    float c = 0.0f;
    split_t s = find_split(bp, n, &c);
    free(bp);
    *dim = s.dim;
    *split_idx = s.split_idx;
    *cost = c;
    return s.valid;
}

float oracle_surface_area(const float* b6)
{
    box3 b = { mk3(b6[0], b6[1], b6[2]), mk3(b6[3], b6[4], b6[5]) };
    return box_sa(b);
}

int oracle_box_intersect(const float* b6, const float* org, const float* dir, float tfar)
{
    box3 b = { mk3(b6[0], b6[1], b6[2]), mk3(b6[3], b6[4], b6[5]) };
    f3 inv = mk3(1.0f / dir[0], 1.0f / dir[1], 1.0f / dir[2]);
    return box_intersect(&b, mk3(org[0], org[1], org[2]), tfar, inv);
}

/* ------------------------------------------------------------------------
 * WideBvh (wide_bvh.cpp) restated: the binary tree above with leaf threshold
 * 16 (wide_bvh.cpp:21), collapsed to 8-wide nodes by repeatedly opening the
 * inner child of largest surface area (collapseNode, wide_bvh.cpp:200-243),
 * flattened depth first (flattenDepthFirst, :290-356) with leaves packed into
 * 8-triangle SoA groups (createLeaf, :430-472), traversed with the ISPC
 * kernels' arithmetic: the slab test of bounding_box.ispc:15-61 (tMax scaled
 * by 1.0000004f) and the Moller-Trumbore of triangle.ispc:27-87 (|det| >=
 * 1e-6, one reciprocal multiplied in).  ISPC is not available here, so
 * whether its build contracts to FMA or approximates the reciprocal is not
 * pinned; this restatement uses IEEE division and no contraction.
 * ------------------------------------------------------------------------ */
#define WIDE_N 8
#define WIDE_LEAF 16
#define WIDE_NO_HIT FLT_MAX

typedef struct {
    float mnx[WIDE_N], mny[WIDE_N], mnz[WIDE_N], mxx[WIDE_N], mxy[WIDE_N], mxz[WIDE_N];
    uint32_t offset[WIDE_N - 1];
    uint16_t child_count;
    uint8_t is_leaf;
} wnode_t;

typedef struct {
    float vx[WIDE_N], vy[WIDE_N], vz[WIDE_N];
    float ex0[WIDE_N], ey0[WIDE_N], ez0[WIDE_N];
    float ex1[WIDE_N], ey1[WIDE_N], ez1[WIDE_N];
} trin_t;

struct oracle_wide {
    oracle_bvh* bin;
    wnode_t* nodes;
    uint32_t node_count, node_cap;
    trin_t* groups;
    uint32_t group_count, group_cap;
};

static int bnode_is_leaf(const bnode* n) { return n->count > 0; }

/* collapseNode (wide_bvh.cpp:200-243) */
static uint32_t wide_collapse(const bnode* node, const bnode** ch)
{
    uint32_t n = 0;
    ch[n++] = node->child[0];
    ch[n++] = node->child[1];
    while (n < WIDE_N) {
        uint32_t candidates = 0, max_idx = 0;
        float max_sa = 0.0f;
        for (uint32_t i = 0; i < n; ++i) {
            if (bnode_is_leaf(ch[i])) continue;
            float sa = box_sa(ch[i]->bbox);
            if (sa > max_sa) { max_sa = sa; max_idx = i; }
            candidates++;
        }
        if (candidates == 0) break;
        /* an inner child of zero area is never chosen by sa > max_sa; the reference then opens
         * slot 0, which is only defined when slot 0 is an inner node */
        if (bnode_is_leaf(ch[max_idx])) break;
        const bnode* c = ch[max_idx];
        ch[n++] = c->child[0];
        ch[max_idx] = c->child[1];
    }
    return n;
}

static void wide_leaf(struct oracle_wide* w, const bnode* node, wnode_t* ln)
{
    const oracle_bvh* b = w->bin;
    ln->offset[0] = node->offset;
    ln->child_count = (uint16_t)node->count;
    ln->is_leaf = 1;
    ln->offset[1] = w->group_count;
    uint32_t groups = (node->count - 1) / WIDE_N + 1;
    for (uint32_t g = 0; g < groups; ++g) {
        if (w->group_count == w->group_cap) {
            w->group_cap = w->group_cap ? 2 * w->group_cap : 1024;
            w->groups = (trin_t*)realloc(w->groups, (size_t)w->group_cap * sizeof(trin_t));
        }
        trin_t* t = &w->groups[w->group_count++];
        memset(t, 0, sizeof(*t));
        for (uint32_t j = 0; j < WIDE_N; ++j) {
            uint32_t idx = g * WIDE_N + j;
            if (idx >= node->count) break;
            const prim_t* p = &b->ordered_prims[node->offset + idx];
            t->vx[j] = p->v0.x; t->vy[j] = p->v0.y; t->vz[j] = p->v0.z;
            t->ex0[j] = p->e0.x; t->ey0[j] = p->e0.y; t->ez0[j] = p->e0.z;
            t->ex1[j] = p->e1.x; t->ey1[j] = p->e1.y; t->ez1[j] = p->e1.z;
        }
    }
}

/* flattenDepthFirst (wide_bvh.cpp:290-356) */
static void wide_flatten(struct oracle_wide* w)
{
    typedef struct { const bnode* node; uint32_t node_index_of_parent; int slot; } item_t;
    uint32_t cap = 1024, sp = 0;
    item_t* stack = (item_t*)malloc(cap * sizeof(item_t));
    stack[sp].node = w->bin->root; stack[sp].slot = -1; stack[sp].node_index_of_parent = 0; sp++;
    while (sp) {
        item_t it = stack[--sp];
        if (w->node_count == w->node_cap) {
            w->node_cap = w->node_cap ? 2 * w->node_cap : 1024;
            w->nodes = (wnode_t*)realloc(w->nodes, (size_t)w->node_cap * sizeof(wnode_t));
        }
        uint32_t me = w->node_count++;
        if (it.slot > 0) w->nodes[it.node_index_of_parent].offset[it.slot - 1] = me;
        wnode_t* ln = &w->nodes[me];
        for (int i = 0; i < WIDE_N; ++i) {
            ln->mnx[i] = ln->mny[i] = ln->mnz[i] = FLT_MAX;
            ln->mxx[i] = ln->mxy[i] = ln->mxz[i] = -FLT_MAX;
        }
        for (int i = 0; i < WIDE_N - 1; ++i) ln->offset[i] = 0xFFFFFFFFu;
        if (bnode_is_leaf(it.node)) { wide_leaf(w, it.node, ln); continue; }
        ln->is_leaf = 0;
        const bnode* ch[WIDE_N];
        uint32_t n = wide_collapse(it.node, ch);
        ln->child_count = (uint16_t)n;
        for (int i = (int)n - 1; i >= 0; --i) {
            box3 bb = ch[i]->bbox;
            ln->mnx[i] = bb.min.x; ln->mny[i] = bb.min.y; ln->mnz[i] = bb.min.z;
            ln->mxx[i] = bb.max.x; ln->mxy[i] = bb.max.y; ln->mxz[i] = bb.max.z;
            if (sp == cap) { cap *= 2; stack = (item_t*)realloc(stack, cap * sizeof(item_t)); }
            stack[sp].node = ch[i]; stack[sp].slot = i; stack[sp].node_index_of_parent = me; sp++;
        }
    }
    free(stack);
}

oracle_wide* oracle_wide_create(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt)
{
    oracle_wide* w = (oracle_wide*)calloc(1, sizeof(oracle_wide));
    w->bin = oracle_bvh_create(xyz, nv, idx, nt, WIDE_LEAF);
    if (nt) wide_flatten(w);
    return w;
}

void oracle_wide_destroy(oracle_wide* w)
{
    if (!w) return;
    oracle_bvh_destroy(w->bin);
    free(w->nodes);
    free(w->groups);
    free(w);
}

uint32_t oracle_wide_node_count(const oracle_wide* w) { return w->node_count; }

static inline float ispc_min(float a, float b) { return a < b ? a : b; }
static inline float ispc_max(float a, float b) { return a > b ? a : b; }

/* intersectBoundingBoxN (bounding_box.ispc:15-61) */
static void wide_boxes(const wnode_t* n, f3 org, f3 inv, float tfar, uint8_t* hit)
{
    const float corr = 1.0000004f;
    for (int i = 0; i < WIDE_N; ++i) {
        float tx0 = (n->mnx[i] - org.x) * inv.x, tx1 = (n->mxx[i] - org.x) * inv.x;
        float ty0 = (n->mny[i] - org.y) * inv.y, ty1 = (n->mxy[i] - org.y) * inv.y;
        float tz0 = (n->mnz[i] - org.z) * inv.z, tz1 = (n->mxz[i] - org.z) * inv.z;
        float tmin = ispc_max(ispc_max(ispc_min(tx0, tx1), ispc_min(ty0, ty1)), ispc_min(tz0, tz1));
        float tmax = ispc_min(ispc_min(ispc_max(tx0, tx1), ispc_max(ty0, ty1)), ispc_max(tz0, tz1));
        tmax *= corr;
        hit[i] = (uint8_t)((tmin <= tmax) & (tmax > 0.0f) & (tmin < tfar));
    }
}

/* intersectTriangleN (triangle.ispc:27-87) */
static void wide_tris(const trin_t* g, f3 o, f3 d, float* t, float* b0, float* b1)
{
    for (int i = 0; i < WIDE_N; ++i) {
        float sx = o.x - g->vx[i], sy = o.y - g->vy[i], sz = o.z - g->vz[i];
        float s1x = d.y * g->ez1[i] - d.z * g->ey1[i];
        float s1y = d.z * g->ex1[i] - d.x * g->ez1[i];
        float s1z = d.x * g->ey1[i] - d.y * g->ex1[i];
        float s2x = sy * g->ez0[i] - sz * g->ey0[i];
        float s2y = sz * g->ex0[i] - sx * g->ez0[i];
        float s2z = sx * g->ey0[i] - sy * g->ex0[i];
        float det = s1x * g->ex0[i] + s1y * g->ey0[i] + s1z * g->ez0[i];
        int hit = fabsf(det) >= 1e-6f;
        float rcp = 1.0f / det;
        float tt = (s2x * g->ex1[i] + s2y * g->ey1[i] + s2z * g->ez1[i]) * rcp;
        float u = (s1x * sx + s1y * sy + s1z * sz) * rcp;
        float v = (d.x * s2x + d.y * s2y + d.z * s2z) * rcp;
        hit &= tt >= 0.0f;
        hit &= (u >= 0.0f) & (v >= 0.0f) & (u + v <= 1.0f);
        t[i] = hit ? tt : WIDE_NO_HIT;
        b0[i] = hit ? u : WIDE_NO_HIT;
        b1[i] = hit ? v : WIDE_NO_HIT;
    }
}

/* WideBvh::Impl::traverse (wide_bvh.cpp:487-619) */
static int wide_traverse(const oracle_wide* w, f3 org, f3 dir, float tfar, int any, uint32_t* prim,
                         float* tout, float* b0out, float* b1out)
{
    if (w->node_count == 0) return 0;
    float tnear = tfar;
    f3 inv = mk3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    uint32_t stack[256];
    uint32_t sp = 0;
    int found = 0;
    stack[sp++] = 0;
    while (sp) {
        uint32_t cur = stack[--sp];
        const wnode_t* n = &w->nodes[cur];
        if (n->is_leaf) {
            uint32_t groups = (n->child_count - 1u) / WIDE_N + 1u;
            for (uint32_t g = 0; g < groups; ++g) {
                float t[WIDE_N], u[WIDE_N], v[WIDE_N];
                wide_tris(&w->groups[n->offset[1] + g], org, dir, t, u, v);
                for (uint32_t j = 0; j < WIDE_N; ++j) {
                    uint32_t idx = g * WIDE_N + j;
                    if (idx >= n->child_count) break;
                    if (t[j] == WIDE_NO_HIT || t[j] >= tnear) continue;
                    if (any) return 1;
                    found = 1;
                    tnear = t[j];
                    *prim = w->bin->ordered[n->offset[0] + idx];
                    *tout = t[j]; *b0out = u[j]; *b1out = v[j];
                }
            }
        } else {
            uint8_t hit[WIDE_N];
            wide_boxes(n, org, inv, tnear, hit);
            for (int i = (int)n->child_count - 1; i >= 0; --i) {
                if (!hit[i]) continue;
                if (sp >= 256) return found;  /* beyond the reference's stack */
                stack[sp++] = i == 0 ? cur + 1 : n->offset[i - 1];
            }
        }
    }
    return found;
}

typedef struct {
    const oracle_wide* w;
    const float *org, *dir, *tfar;
    uint32_t begin, end;
    int any;
    uint32_t* prim; float *t, *b0, *b1; uint8_t* occ;
} wjob_t;

static void* wide_run(void* arg)
{
    wjob_t* j = (wjob_t*)arg;
    for (uint32_t i = j->begin; i < j->end; ++i) {
        f3 o = mk3(j->org[3 * (size_t)i], j->org[3 * (size_t)i + 1], j->org[3 * (size_t)i + 2]);
        f3 d = mk3(j->dir[3 * (size_t)i], j->dir[3 * (size_t)i + 1], j->dir[3 * (size_t)i + 2]);
        if (j->any) {
            uint32_t p; float t, u, v;
            j->occ[i] = (uint8_t)wide_traverse(j->w, o, d, j->tfar ? j->tfar[i] : FLT_MAX, 1, &p, &t, &u, &v);
        } else {
            uint32_t p = 0xFFFFFFFFu; float t = FLT_MAX, u = 0.0f, v = 0.0f;
            if (!wide_traverse(j->w, o, d, FLT_MAX, 0, &p, &t, &u, &v)) { p = 0xFFFFFFFFu; t = FLT_MAX; u = v = 0.0f; }
            j->prim[i] = p; j->t[i] = t; j->b0[i] = u; j->b1[i] = v;
        }
    }
    return NULL;
}

static void wide_jobs(wjob_t proto, uint32_t n, uint32_t nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    wjob_t jobs[256];
    uint32_t chunk = (n + nthreads - 1) / nthreads;
    uint32_t used = 0;
    for (uint32_t k = 0; k < nthreads; ++k) {
        jobs[k] = proto;
        jobs[k].begin = k * chunk < n ? k * chunk : n;
        jobs[k].end = (k + 1) * chunk < n ? (k + 1) * chunk : n;
        if (jobs[k].begin >= jobs[k].end) continue;
        if (nthreads == 1) { wide_run(&jobs[k]); continue; }
        pthread_create(&th[used], NULL, wide_run, &jobs[k]);
        used++;
    }
    for (uint32_t k = 0; k < used; ++k) pthread_join(th[k], NULL);
}

void oracle_wide_intersect(const oracle_wide* w, const float* org, const float* dir, uint32_t n,
                           uint32_t* prim, float* t, float* b0, float* b1, uint32_t nthreads)
{
    wjob_t j = { w, org, dir, NULL, 0, 0, 0, prim, t, b0, b1, NULL };
    wide_jobs(j, n, nthreads);
}

void oracle_wide_occluded(const oracle_wide* w, const float* org, const float* dir, const float* tfar,
                          uint32_t n, uint8_t* occ, uint32_t nthreads)
{
    wjob_t j = { w, org, dir, tfar, 0, 0, 1, NULL, NULL, NULL, NULL, occ };
    wide_jobs(j, n, nthreads);
}

/* Brute force with the ISPC triangle test, for pinning the traversal. */
void oracle_wide_brute(const oracle_wide* w, const float* org, const float* dir, uint32_t n,
                       uint32_t* prim, float* tout)
{
    const oracle_bvh* b = w->bin;
    for (uint32_t i = 0; i < n; ++i) {
        f3 o = mk3(org[3 * (size_t)i], org[3 * (size_t)i + 1], org[3 * (size_t)i + 2]);
        f3 d = mk3(dir[3 * (size_t)i], dir[3 * (size_t)i + 1], dir[3 * (size_t)i + 2]);
        float best = FLT_MAX;
        uint32_t bp = 0xFFFFFFFFu;
        for (uint32_t k = 0; k < b->nt; ++k) {
            trin_t g;
            memset(&g, 0, sizeof(g));
            const prim_t* p = &b->prims[k];
            g.vx[0] = p->v0.x; g.vy[0] = p->v0.y; g.vz[0] = p->v0.z;
            g.ex0[0] = p->e0.x; g.ey0[0] = p->e0.y; g.ez0[0] = p->e0.z;
            g.ex1[0] = p->e1.x; g.ey1[0] = p->e1.y; g.ez1[0] = p->e1.z;
            float t[WIDE_N], u[WIDE_N], v[WIDE_N];
            wide_tris(&g, o, d, t, u, v);
            if (t[0] != WIDE_NO_HIT && t[0] < best) { best = t[0]; bp = k; }
        }
        prim[i] = bp;
        tout[i] = best;
    }
}
