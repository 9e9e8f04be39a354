// Host-side scene preparation, binned-SAH binary build and collapse into
// 64-byte quantised 8-wide nodes.
//
// The binary build uses the reference's split rule exactly (9 bins, 8 split
// planes, forward/backward sweeps, strict-< argmin, leaf when N <= threshold
// or split cost >= N * SA; bvh_common.cpp:185-343, bvh_util.h:6-27), so with
// the reference's leaf threshold it reproduces the reference tree's leaf sets.
// Leaves above APG_MAX_LEAF triangles are then split into balanced sub-trees
// for the wide layout; the SAH statistics and leaf keys describe the tree
// before that split.
#include "apg_internal.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <future>

namespace apg {

namespace {
constexpr uint32_t SPLITS = 8;
constexpr uint32_t BINS = SPLITS + 1;
constexpr uint32_t PARALLEL_MIN = 4096;

inline float fmin_s(float a, float b) { return (b < a) ? b : a; }
inline float fmax_s(float a, float b) { return (a < b) ? b : a; }

struct Box
{
    float lo[3] = { FLT_MAX, FLT_MAX, FLT_MAX };
    float hi[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
    void merge(const Box& b)
    {
        for (int a = 0; a < 3; ++a) { lo[a] = fmin_s(lo[a], b.lo[a]); hi[a] = fmax_s(hi[a], b.hi[a]); }
    }
    void merge_pt(const float* p)
    {
        for (int a = 0; a < 3; ++a) { lo[a] = fmin_s(lo[a], p[a]); hi[a] = fmax_s(hi[a], p[a]); }
    }
    float area() const
    {
        float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
        return (x * y + y * z + z * x) * 2.0f;
    }
    float center(int a) const { return (lo[a] + hi[a]) * 0.5f; }
};

struct BPrim
{
    Box box;
    uint32_t index;
};

inline uint8_t bin_of(float c, float mn, float mx)
{
    uint8_t b = static_cast<uint8_t>(float(BINS) * (c - mn) / (mx - mn));
    return std::min<uint8_t>(b, BINS - 1);
}

struct Split
{
    bool valid = false;
    uint8_t dim = 0, idx = 0;
    float mn = 0.0f, mx = 0.0f;
    bool below(const Box& b) const
    {
        return bin_of(b.center(dim), mn, mx) <= idx;
    }
};

Split find_split(const BPrim* p, uint32_t n, float& cost_out)
{
    Split best;
    Box cb;
    for (uint32_t i = 0; i < n; ++i)
    {
        float c[3] = { p[i].box.center(0), p[i].box.center(1), p[i].box.center(2) };
        cb.merge_pt(c);
    }
    float diag = fmax_s(cb.hi[0] - cb.lo[0], fmax_s(cb.hi[1] - cb.lo[1], cb.hi[2] - cb.lo[2]));
    if (diag == 0.0f)
    {
        cost_out = 0.0f;
        return best;
    }
    float min_cost = FLT_MAX;
    for (uint8_t d = 0; d < 3; ++d)
    {
        float mn = cb.lo[d], mx = cb.hi[d];
        if (mn == mx) continue;
        Box bb[BINS];
        uint32_t bc[BINS] = {};
        for (uint32_t i = 0; i < n; ++i)
        {
            uint8_t b = bin_of(p[i].box.center(d), mn, mx);
            bb[b].merge(p[i].box);
            bc[b]++;
        }
        float costs[SPLITS] = {};
        Box acc;
        uint32_t an = 0;
        for (uint32_t s = 0; s < SPLITS; ++s)
        {
            acc.merge(bb[s]);
            an += bc[s];
            costs[s] += float(an) * acc.area();
        }
        acc = Box();
        an = 0;
        for (int s = SPLITS - 1; s >= 0; --s)
        {
            acc.merge(bb[s + 1]);
            an += bc[s + 1];
            costs[s] += float(an) * acc.area();
        }
        for (uint32_t s = 0; s < SPLITS; ++s)
        {
            if (costs[s] < min_cost)
            {
                // valid only once a plane has won: when every cost is >= FLT_MAX (areas
                // overflowed) the reference keeps its default Split, sends every primitive
                // below and recurses without end; here the node becomes a leaf
                best.valid = true;
                min_cost = costs[s];
                best.dim = d;
                best.idx = uint8_t(s);
                best.mn = mn;
                best.mx = mx;
            }
        }
    }
    cost_out = min_cost;
    return best;
}

enum : uint32_t { F_LEAF = 1, F_SAH_LEAF = 2 };

struct Builder
{
    std::vector<BPrim>& prims;
    std::vector<BinNode>& nodes;
    std::vector<uint32_t>& flags;
    std::atomic<uint32_t> next{ 0 };
    uint32_t threshold;

    uint32_t alloc() { return next.fetch_add(1); }

    void set_box(uint32_t id, const Box& b)
    {
        BinNode& nd = nodes[id];
        for (int a = 0; a < 3; ++a) { nd.lo[a] = b.lo[a]; nd.hi[a] = b.hi[a]; }
    }

    // Split an oversized leaf range into a balanced tree of <= APG_MAX_LEAF leaves.
    void force_split(uint32_t id, uint32_t first, uint32_t count)
    {
        BinNode& nd = nodes[id];
        nd.first = first;
        nd.count = count;
        if (count <= APG_MAX_LEAF)
        {
            flags[id] |= F_LEAF;
            return;
        }
        uint32_t half = ((count + 1) / 2 + APG_MAX_LEAF - 1) / APG_MAX_LEAF * APG_MAX_LEAF;
        half = std::min(half, count - 1);
        uint32_t l = alloc(), r = alloc();
        nd.left = l;
        nd.right = r;
        Box bl, br;
        for (uint32_t i = 0; i < half; ++i) bl.merge(prims[first + i].box);
        for (uint32_t i = half; i < count; ++i) br.merge(prims[first + i].box);
        set_box(l, bl);
        set_box(r, br);
        force_split(l, first, half);
        force_split(r, first + half, count - half);
    }

    void make_leaf(uint32_t id, uint32_t first, uint32_t count)
    {
        flags[id] |= F_SAH_LEAF;
        force_split(id, first, count);
    }

    void build(uint32_t id, uint32_t first, uint32_t count)
    {
        BPrim* p = prims.data() + first;
        Box nb;
        for (uint32_t i = 0; i < count; ++i) nb.merge(p[i].box);
        set_box(id, nb);
        if (count <= threshold)
        {
            make_leaf(id, first, count);
            return;
        }
        float split_cost = 0.0f;
        Split s = find_split(p, count, split_cost);
        float leaf_cost = float(count) * nb.area();
        if (!s.valid || split_cost >= leaf_cost)
        {
            make_leaf(id, first, count);
            return;
        }
        BPrim* mid = std::partition(p, p + count, [&s](const BPrim& bp) { return s.below(bp.box); });
        uint32_t nl = uint32_t(mid - p);
        BinNode& nd = nodes[id];
        nd.first = first;
        nd.count = count;
        uint32_t l = alloc(), r = alloc();
        nd.left = l;
        nd.right = r;
        if (count >= PARALLEL_MIN)
        {
            auto fut = std::async(std::launch::async, [this, l, first, nl]() { build(l, first, nl); });
            build(r, first + nl, count - nl);
            fut.get();
        }
        else
        {
            build(l, first, nl);
            build(r, first + nl, count - nl);
        }
    }
};

inline float area_of(const BinNode& n)
{
    float x = n.hi[0] - n.lo[0], y = n.hi[1] - n.lo[1], z = n.hi[2] - n.lo[2];
    return (x * y + y * z + z * x) * 2.0f;
}
} // namespace

void build_binary_host(const std::vector<TriPrep>& tris, uint32_t leaf_threshold,
                       BinaryTree& out)
{
    const uint32_t n = uint32_t(tris.size());
    out.nodes.assign(n ? 2 * size_t(n) : 1, BinNode{});
    out.order.resize(n);
    out.leaf_keys.assign(n, 0);
    if (n == 0) return;
    std::vector<BPrim> prims(n);
    for (uint32_t i = 0; i < n; ++i)
    {
        for (int a = 0; a < 3; ++a) { prims[i].box.lo[a] = tris[i].lo[a]; prims[i].box.hi[a] = tris[i].hi[a]; }
        prims[i].index = i;
    }
    std::vector<uint32_t> flags(out.nodes.size(), 0);
    Builder b{ prims, out.nodes, flags };
    b.threshold = leaf_threshold;
    uint32_t root = b.alloc();
    b.build(root, 0, n);
    out.nodes.resize(b.next.load());
    for (uint32_t i = 0; i < n; ++i) out.order[i] = prims[i].index;

    // Statistics and leaf keys over the SAH tree (forced splits excluded).
    double inner = 0.0, leaf = 0.0;
    uint32_t sah_nodes = 0, sah_leaves = 0, maxd = 0;
    struct Item { uint32_t id, depth; };
    std::vector<Item> st{ { 0, 0 } };
    while (!st.empty())
    {
        Item it = st.back();
        st.pop_back();
        const BinNode& nd = out.nodes[it.id];
        sah_nodes++;
        maxd = std::max(maxd, it.depth);
        double sa = double(area_of(nd));
        if (flags[it.id] & F_SAH_LEAF)
        {
            sah_leaves++;
            leaf += sa * double(nd.count);
            uint32_t mn = 0xFFFFFFFFu;
            for (uint32_t k = 0; k < nd.count; ++k) mn = std::min(mn, out.order[nd.first + k]);
            for (uint32_t k = 0; k < nd.count; ++k) out.leaf_keys[out.order[nd.first + k]] = mn;
            continue;
        }
        inner += sa;
        st.push_back({ nd.right, it.depth + 1 });
        st.push_back({ nd.left, it.depth + 1 });
    }
    for (int a = 0; a < 3; ++a) { out.scene_lo[a] = out.nodes[0].lo[a]; out.scene_hi[a] = out.nodes[0].hi[a]; }
    double root_sa = double(area_of(out.nodes[0]));
    out.sah_cost = (inner + leaf) / root_sa;
    out.sah_cost_leaf = leaf / root_sa;
    out.sah_nodes = sah_nodes;
    out.sah_leaves = sah_leaves;
    out.max_depth = maxd;
    // Mark final leaves: count > 0 and no children.  Inner nodes keep count
    // for the range but left != right.
    for (size_t i = 0; i < out.nodes.size(); ++i)
        if (!(flags[i] & F_LEAF)) out.nodes[i].count = 0;
}

float quant_slack(const std::vector<TriPrep>& tris)
{
    float s = 0.0f;
    for (const auto& t : tris)
        for (int a = 0; a < 3; ++a) s = std::max(s, std::max(std::fabs(t.lo[a]), std::fabs(t.hi[a])));
    if (s == 0.0f) s = 1.0f;
    return std::ldexp(s, -17);
}

namespace {
// Quantise one axis of a wide node.  Returns the biased exponent.
uint32_t quantise_axis(float lo_f, float hi_f, float slack, const float* clo, const float* chi,
                       uint32_t nslots, const bool* valid, float& origin, uint32_t* qlo,
                       uint32_t* qhi)
{
    double lo = double(lo_f) - double(slack), hi = double(hi_f) + double(slack);
    double m = std::max(std::fabs(lo), std::fabs(hi));
    int k = 0;
    std::frexp(m > 0.0 ? m : 1.0, &k);          // m < 2^k
    int e = k - 23;                             // origin / 2^e stays below 2^24
    double ext = hi - lo;
    while (std::ldexp(62.0, e) < ext) ++e;
    e = std::max(e, -126);
    for (;;)
    {
        double step = std::ldexp(1.0, e);
        double o = std::floor(lo / step) * step;
        if (o + 63.0 * step >= hi && float(o) == o)
        {
            origin = float(o);
            for (uint32_t s = 0; s < nslots; ++s)
            {
                if (!valid[s]) { qlo[s] = APG_QMAX; qhi[s] = 0; continue; }
                double a = std::floor((double(clo[s]) - double(slack) - o) / step);
                double b = std::ceil((double(chi[s]) + double(slack) - o) / step);
                qlo[s] = uint32_t(std::clamp(a, 0.0, double(APG_QMAX)));
                qhi[s] = uint32_t(std::clamp(b, 0.0, double(APG_QMAX)));
            }
            return uint32_t(e + 127);
        }
        ++e;
    }
}
} // namespace

uint32_t collapse_to_wide(const BinaryTree& bt, const std::vector<TriPrep>& tris,
                          std::vector<ApgWideNode>& nodes, std::vector<float>& tri9,
                          std::vector<uint32_t>& tri_ids)
{
    nodes.clear();
    tri9.clear();
    tri_ids.clear();
    if (tris.empty()) return 0;
    const float slack = quant_slack(tris);
    const auto& bn = bt.nodes;
    auto is_leaf = [&](uint32_t i) { return bn[i].count > 0; };

    nodes.reserve(bn.size() / 4 + 1);
    tri9.reserve(tris.size() * 9);
    tri_ids.reserve(tris.size());

    struct Item { uint32_t bin, wide, level; };
    std::vector<Item> queue;
    nodes.push_back(ApgWideNode{});
    queue.push_back({ 0, 0, 1 });
    size_t head = 0;
    uint32_t levels = 0;
    while (head < queue.size())
    {
        Item it = queue[head++];
        levels = std::max(levels, it.level);
        // Gather up to 8 children by repeatedly opening the inner child with
        // the largest surface area (wide_bvh.cpp:200-243).
        uint32_t ch[APG_WIDE];
        uint32_t nch = 0;
        if (is_leaf(it.bin))
            ch[nch++] = it.bin;
        else
        {
            ch[nch++] = bn[it.bin].left;
            ch[nch++] = bn[it.bin].right;
            while (nch < APG_WIDE)
            {
                int best = -1;
                float best_sa = 0.0f;
                for (uint32_t i = 0; i < nch; ++i)
                {
                    if (is_leaf(ch[i])) continue;
                    float sa = area_of(bn[ch[i]]);
                    if (best < 0 || sa > best_sa) { best = int(i); best_sa = sa; }
                }
                if (best < 0) break;
                uint32_t c = ch[best];
                ch[nch++] = bn[c].left;
                ch[best] = bn[c].right;
            }
        }

        // Octant-ordered slots: slot s holds the child nearest for rays whose
        // direction sign bits equal s (bit a set = negative along axis a).
        float pc[3];
        for (int a = 0; a < 3; ++a) pc[a] = 0.5f * (bn[it.bin].lo[a] + bn[it.bin].hi[a]);
        float cost[APG_WIDE][APG_WIDE];
        for (uint32_t c = 0; c < nch; ++c)
        {
            float d[3];
            for (int a = 0; a < 3; ++a) d[a] = 0.5f * (bn[ch[c]].lo[a] + bn[ch[c]].hi[a]) - pc[a];
            for (uint32_t s = 0; s < APG_WIDE; ++s)
                cost[c][s] = ((s & 1) ? -d[0] : d[0]) + ((s & 2) ? -d[1] : d[1]) + ((s & 4) ? -d[2] : d[2]);
        }
        int slot_child[APG_WIDE];
        for (auto& v : slot_child) v = -1;
        bool child_done[APG_WIDE] = {};
        for (uint32_t round = 0; round < nch; ++round)
        {
            float bc = FLT_MAX;
            int bcc = -1, bs = -1;
            for (uint32_t c = 0; c < nch; ++c)
            {
                if (child_done[c]) continue;
                for (uint32_t s = 0; s < APG_WIDE; ++s)
                {
                    if (slot_child[s] >= 0) continue;
                    if (bcc < 0 || cost[c][s] < bc) { bc = cost[c][s]; bcc = int(c); bs = int(s); }
                }
            }
            slot_child[bs] = bcc;
            child_done[bcc] = true;
        }

        ApgWideNode nd;
        std::memset(&nd, 0, sizeof(nd));
        uint32_t imask = 0, cmask = 0;
        nd.child_base = uint32_t(nodes.size());
        nd.tri_base = uint32_t(tri_ids.size());
        float clo[3][APG_WIDE], chi[3][APG_WIDE];
        bool vmask[APG_WIDE] = {};
        float nlo[3] = { FLT_MAX, FLT_MAX, FLT_MAX }, nhi[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
        uint32_t inner_rank = 0;
        std::vector<Item> pending;
        for (uint32_t s = 0; s < APG_WIDE; ++s)
        {
            if (slot_child[s] < 0) continue;
            uint32_t c = ch[slot_child[s]];
            vmask[s] = true;
            for (int a = 0; a < 3; ++a)
            {
                clo[a][s] = bn[c].lo[a];
                chi[a][s] = bn[c].hi[a];
                nlo[a] = std::min(nlo[a], bn[c].lo[a]);
                nhi[a] = std::max(nhi[a], bn[c].hi[a]);
            }
            if (is_leaf(c))
            {
                const BinNode& lf = bn[c];
                cmask |= apg_leaf_nibble(s, lf.count);
                for (uint32_t k = 0; k < lf.count; ++k)
                {
                    uint32_t g = bt.order[lf.first + k];
                    const TriPrep& t = tris[g];
                    tri9.insert(tri9.end(), { t.v0[0], t.v0[1], t.v0[2], t.e0[0], t.e0[1], t.e0[2],
                                              t.e1[0], t.e1[1], t.e1[2] });
                    tri_ids.push_back(g);
                }
            }
            else
            {
                imask |= 1u << s;
                pending.push_back({ c, nd.child_base + inner_rank++, it.level + 1 });
            }
        }
        for (size_t i = 0; i < pending.size(); ++i) nodes.push_back(ApgWideNode{});
        uint32_t qlo[3][APG_WIDE], qhi[3][APG_WIDE];
        uint32_t eb[3];
        for (int a = 0; a < 3; ++a)
            eb[a] = quantise_axis(nlo[a], nhi[a], slack, clo[a], chi[a], APG_WIDE, vmask,
                                  nd.origin[a], qlo[a], qhi[a]);
        nd.exp_imask = eb[0] | (eb[1] << 8) | (eb[2] << 16) | (imask << 24);
        nd.cmask = cmask;
        for (uint32_t s = 0; s < APG_WIDE; ++s)
            for (int a = 0; a < 3; ++a)
            {
                apg_put_plane(nd.q, s, uint32_t(a), qlo[a][s]);
                apg_put_plane(nd.q, s, uint32_t(3 + a), qhi[a][s]);
            }
        nodes[it.wide] = nd;
        for (const Item& p : pending) queue.push_back(p);
    }
    return levels;
}

void prepare_triangles(const Context& ctx, std::vector<TriPrep>& out)
{
    out.resize(ctx.total_tris);
    for (const Shape& sh : ctx.shapes)
    {
        const uint32_t nt = uint32_t(sh.idx.size() / 3);
        for (uint32_t i = 0; i < nt; ++i)
        {
            TriPrep& t = out[sh.prim_offset + i];
            const float* v[3];
            for (int k = 0; k < 3; ++k) v[k] = &sh.xyz[3 * size_t(sh.idx[3 * size_t(i) + k])];
            for (int a = 0; a < 3; ++a)
            {
                t.v0[a] = v[0][a];
                t.e0[a] = v[1][a] - v[0][a];
                t.e1[a] = v[2][a] - v[0][a];
            }
            Box b;
            for (int k = 0; k < 3; ++k) b.merge_pt(v[k]);
            for (int a = 0; a < 3; ++a) { t.lo[a] = b.lo[a]; t.hi[a] = b.hi[a]; }
        }
    }
}

} // namespace apg
