// cwbvh_build.cpp - binned-SAH BVH2, SAH-optimal 8-wide collapse (dynamic
// programming over the BVH2, leaves of up to 3 triangles, 2 by default), octant-ordered
// child slots, and 8-bit quantisation of child bounds into the 80 B CWBVH
// node format decoded by traverse_cwbvh.cuh:51-185.
//
// Written from the decoder's semantics and from Ylitie, Karras, Laine,
// "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs"
// (HPG 2017); the reference's own builder (tinybvh, tinybvh_wrapper.cpp:53)
// is not available, so tree shape and node order are this builder's own.
#include "cwbvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <future>
#include <limits>

namespace tfg {
namespace {

struct Box {
    float mn[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                   std::numeric_limits<float>::infinity()};
    float mx[3] = {-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                   -std::numeric_limits<float>::infinity()};
    void grow(const Box& b) {
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], b.mn[a]);
            mx[a] = std::max(mx[a], b.mx[a]);
        }
    }
    void grow(const float* p) {
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], p[a]);
            mx[a] = std::max(mx[a], p[a]);
        }
    }
    double area() const {
        double dx = (double)mx[0] - mn[0], dy = (double)mx[1] - mn[1], dz = (double)mx[2] - mn[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct Node2 {
    Box box;
    int left = -1, right = -1;   // local child indices, -1 for a leaf
    int first = 0, count = 0;    // range in the index array (all nodes)
};

struct Ctx {
    const float* verts;
    int bins = 32;
    std::vector<Box> tbox;
    std::vector<float> cent;     // 3 per triangle
    std::vector<uint32_t> idx;
};

constexpr int kMaxBins = 64;
constexpr int kParallelMin = 1 << 15;

// Builds the subtree over idx[begin, end) into `nodes` (root at local 0 of
// the entry point's vector).  Leaves hold exactly one triangle.
void build_rec(Ctx& c, int begin, int end, std::vector<Node2>& nodes, int depth) {
    const int me = (int)nodes.size();
    nodes.push_back(Node2{});
    Box box, cbox;
    for (int i = begin; i < end; ++i) {
        const uint32_t t = c.idx[i];
        box.grow(c.tbox[t]);
        cbox.grow(&c.cent[3 * (size_t)t]);
    }
    nodes[me].box = box;
    nodes[me].first = begin;
    nodes[me].count = end - begin;
    const int n = end - begin;
    if (n == 1) return;

    int best_axis = -1, best_split = -1;
    double best_cost = std::numeric_limits<double>::infinity();
    for (int a = 0; a < 3; ++a) {
        const double ext = (double)cbox.mx[a] - cbox.mn[a];
        if (!(ext > 0.0)) continue;
        const int nb = c.bins;
        const double scale = nb / ext;
        Box bb[kMaxBins];
        int bc[kMaxBins] = {0};
        for (int i = begin; i < end; ++i) {
            const uint32_t t = c.idx[i];
            int b = (int)(((double)c.cent[3 * (size_t)t + a] - cbox.mn[a]) * scale);
            b = std::min(std::max(b, 0), nb - 1);
            bb[b].grow(c.tbox[t]);
            bc[b]++;
        }
        double ra[kMaxBins];
        int rc[kMaxBins];
        Box acc;
        int cnt = 0;
        for (int b = nb - 1; b > 0; --b) {
            acc.grow(bb[b]);
            cnt += bc[b];
            ra[b] = acc.area();
            rc[b] = cnt;
        }
        Box lacc;
        int lc = 0;
        for (int b = 0; b < nb - 1; ++b) {
            lacc.grow(bb[b]);
            lc += bc[b];
            const int r = rc[b + 1];
            if (lc == 0 || r == 0) continue;
            const double cost = lacc.area() * lc + ra[b + 1] * r;
            if (cost < best_cost) {
                best_cost = cost;
                best_axis = a;
                best_split = b;
            }
        }
    }

    int mid;
    if (best_axis < 0) {
        mid = begin + n / 2;   // coincident centroids: split by position
    } else {
        const double ext = (double)cbox.mx[best_axis] - cbox.mn[best_axis];
        const double scale = c.bins / ext;
        const float lo = cbox.mn[best_axis];
        auto it = std::stable_partition(c.idx.begin() + begin, c.idx.begin() + end, [&](uint32_t t) {
            int b = (int)(((double)c.cent[3 * (size_t)t + best_axis] - lo) * scale);
            b = std::min(std::max(b, 0), c.bins - 1);
            return b <= best_split;
        });
        mid = (int)(it - c.idx.begin());
        if (mid == begin || mid == end) mid = begin + n / 2;
    }

    if (n >= kParallelMin && depth < 5) {
        std::vector<Node2> ln, rn;
        auto fut = std::async(std::launch::async, [&] { build_rec(c, begin, mid, ln, depth + 1); });
        build_rec(c, mid, end, rn, depth + 1);
        fut.get();
        const int loff = (int)nodes.size();
        for (auto nd : ln) {
            if (nd.left >= 0) { nd.left += loff; nd.right += loff; }
            nodes.push_back(nd);
        }
        const int roff = (int)nodes.size();
        for (auto nd : rn) {
            if (nd.left >= 0) { nd.left += roff; nd.right += roff; }
            nodes.push_back(nd);
        }
        nodes[me].left = loff;
        nodes[me].right = roff;
    } else {
        const int l = (int)nodes.size();
        build_rec(c, begin, mid, nodes, depth + 1);
        const int r = (int)nodes.size();
        build_rec(c, mid, end, nodes, depth + 1);
        nodes[me].left = l;
        nodes[me].right = r;
    }
}

// SAH-optimal collapse of the BVH2 into 8-wide nodes (Ylitie, Karras, Laine
// 2017, section 4.1).  cost[n][i] is the least SAH cost of representing the
// subtree of BVH2 node n with at most i slots of its parent wide node:
//   i = 1: a leaf slot (at most 3 triangles: A(n) * kTriCost * count) or an
//          internal slot (A(n) * kNodeCost + best split of 8 slots between
//          n's two children);
//   i > 1: the best of i - 1 slots and every split of i slots between the
//          children.
// split[n][i] records the choice (0: i - 1 slots / the i = 1 case, k > 0:
// k slots to the left child), wide[n] the split of an internal node's 8
// slots.  Subtrees of at most max_leaf (<= 3) triangles are always leaf
// slots when given one slot, and larger ones always internal, so a slot's
// kind follows from its triangle count.  Node and triangle costs are
// BuildOptions.
struct Collapse {
    std::vector<float> cost;      // 8 per BVH2 node (slots 1..8)
    std::vector<uint8_t> split;   // 8 per BVH2 node
    std::vector<uint8_t> wide;    // left share of an internal node's 8 slots
};

void plan_collapse(const std::vector<Node2>& n2, const BuildOptions& opt, Collapse& cp) {
    const double kNodeCost = opt.node_cost, kTriCost = opt.tri_cost;
    const size_t nn = n2.size();
    cp.cost.assign(8 * nn, 0.f);
    cp.split.assign(8 * nn, 0);
    cp.wide.assign(nn, 0);
    // children always follow their parent in n2, so a reverse sweep is post-order
    for (size_t idx = nn; idx-- > 0;) {
        const Node2& n = n2[idx];
        float* C = &cp.cost[8 * idx];
        uint8_t* S = &cp.split[8 * idx];
        const double area = n.box.area();
        if (n.left < 0) {   // single triangle
            for (int i = 0; i < 8; ++i) C[i] = (float)(area * kTriCost * n.count);
            continue;
        }
        const float* CL = &cp.cost[8 * (size_t)n.left];
        const float* CR = &cp.cost[8 * (size_t)n.right];
        double dist[9];
        int dk[9];
        for (int j = 2; j <= 8; ++j) {
            dist[j] = std::numeric_limits<double>::infinity();
            dk[j] = 1;
            for (int k = 1; k < j; ++k) {
                const double c = (double)CL[k - 1] + (double)CR[j - k - 1];
                if (c < dist[j]) { dist[j] = c; dk[j] = k; }
            }
        }
        cp.wide[idx] = (uint8_t)dk[8];
        const double c1 = n.count <= opt.max_leaf ? area * kTriCost * n.count : area * kNodeCost + dist[8];
        C[0] = (float)c1;
        S[0] = 0;
        for (int i = 2; i <= 8; ++i) {
            if (dist[i] < (double)C[i - 2]) {
                C[i - 1] = (float)dist[i];
                S[i - 1] = (uint8_t)dk[i];
            } else {
                C[i - 1] = C[i - 2];
                S[i - 1] = 0;
            }
        }
    }
}

// The slots BVH2 node n takes when given at most i of them.
void expand(const std::vector<Node2>& n2, const Collapse& cp, int n, int i, int* kids, int& nk) {
    while (i > 1 && cp.split[8 * (size_t)n + i - 1] == 0) --i;
    if (i <= 1) {
        kids[nk++] = n;
        return;
    }
    const int k = cp.split[8 * (size_t)n + i - 1];
    expand(n2, cp, n2[n].left, k, kids, nk);
    expand(n2, cp, n2[n].right, i - k, kids, nk);
}

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// Power-of-two exponent e (stored as a signed byte) such that 255 * 2^e
// covers `extent`.  Clamped so that (e + 127) << 23 is a normal float.
int quant_exponent(double extent) {
    if (!(extent > 0.0)) return -100;
    int e = (int)std::ceil(std::log2(extent / 255.0));
    while (std::ldexp(255.0, e) < extent) ++e;
    return std::max(-100, std::min(e, 127));
}

void write_tri(const float* v, uint32_t prim, uint32_t format, float* out) {
    const float* a = v;
    const float* b = v + 3;
    const float* c2 = v + 6;
    if (format == TRI_WATERTIGHT) {
        out[0] = a[0]; out[1] = a[1]; out[2] = a[2]; out[3] = 0.f;
        out[4] = b[0]; out[5] = b[1]; out[6] = b[2]; out[7] = 0.f;
        out[8] = c2[0]; out[9] = c2[1]; out[10] = c2[2]; out[11] = 0.f;
        out[12] = 0.f; out[13] = 0.f; out[14] = 0.f; out[15] = u2f(prim);
        return;
    }
    // Affine world -> unit-triangle transform: rows of inverse([e1 e2 n]).
    double e1[3], e2[3], nn[3];
    for (int k = 0; k < 3; ++k) { e1[k] = (double)b[k] - a[k]; e2[k] = (double)c2[k] - a[k]; }
    nn[0] = e1[1] * e2[2] - e1[2] * e2[1];
    nn[1] = e1[2] * e2[0] - e1[0] * e2[2];
    nn[2] = e1[0] * e2[1] - e1[1] * e2[0];
    auto cross = [](const double* x, const double* y, double* r) {
        r[0] = x[1] * y[2] - x[2] * y[1];
        r[1] = x[2] * y[0] - x[0] * y[2];
        r[2] = x[0] * y[1] - x[1] * y[0];
    };
    double r0[3], r1[3];
    cross(e2, nn, r0);
    cross(nn, e1, r1);
    const double det = nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2];
    float* o = out;
    if (!(det > 0.0)) {
        for (int k = 0; k < 12; ++k) o[k] = 0.f;
        o[11] = 1.f;   // T2 = (0,0,0,1): transD = 0 -> d = -inf, rejected
    } else {
        const double* rows[3] = {r0, r1, nn};
        for (int r = 0; r < 3; ++r) {
            double w = 0.0;
            for (int k = 0; k < 3; ++k) {
                o[4 * r + k] = (float)(rows[r][k] / det);
                w -= (rows[r][k] / det) * a[k];
            }
            o[4 * r + 3] = (float)w;
        }
    }
    out[12] = a[0]; out[13] = a[1]; out[14] = a[2]; out[15] = u2f(prim);
}

}  // namespace

bool build_cwbvh(const float* verts, uint32_t n_tris, uint32_t format, CwbvhHost& out, const BuildOptions& opt) {
    out = CwbvhHost{};
    out.format = format;
    if (verts == nullptr || n_tris == 0 || n_tris >= (1u << 29)) return false;
    if (format != TRI_WATERTIGHT && format != TRI_AFFINE) return false;
    if (opt.max_leaf < 1 || opt.max_leaf > 3 || opt.bins < 4 || opt.bins > kMaxBins ||
        !(opt.node_cost > 0.0) || !(opt.tri_cost > 0.0))
        return false;

    Ctx c;
    c.verts = verts;
    c.bins = opt.bins;
    c.tbox.resize(n_tris);
    c.cent.resize(3 * (size_t)n_tris);
    c.idx.resize(n_tris);
    for (uint32_t t = 0; t < n_tris; ++t) {
        Box b;
        for (int k = 0; k < 3; ++k) {
            const float* p = verts + 9 * (size_t)t + 3 * k;
            if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) return false;
            b.grow(p);
        }
        c.tbox[t] = b;
        for (int a = 0; a < 3; ++a) c.cent[3 * (size_t)t + a] = 0.5f * (b.mn[a] + b.mx[a]);
        c.idx[t] = t;
    }

    std::vector<Node2> n2;
    n2.reserve(2 * (size_t)n_tris);
    build_rec(c, 0, (int)n_tris, n2, 0);
    Collapse cp;
    plan_collapse(n2, opt, cp);

    // ---- collapse to 8-wide, breadth-first node order --------------------
    struct Pending { uint32_t slot8; int node2; uint32_t depth; };
    std::deque<Pending> q;
    out.nodes.reserve(20 * (size_t)(n_tris / 2 + 1));
    out.tris.reserve(16 * (size_t)n_tris);
    out.nodes.resize(20);
    q.push_back({0, 0, 1});
    uint32_t tri_records = 0;
    double sah = 0.0;
    const double root_area = std::max(n2[0].box.area(), 1e-30);

    std::vector<int> stack;
    while (!q.empty()) {
        const Pending pd = q.front();
        q.pop_front();
        out.max_depth = std::max(out.max_depth, pd.depth);
        const Node2& root = n2[pd.node2];
        sah += root.box.area() / root_area;

        int kids[8];
        int nk = 0;
        if (root.left < 0) {
            kids[nk++] = pd.node2;   // single-triangle tree: one leaf child
        } else {
            const int k8 = cp.wide[pd.node2];
            expand(n2, cp, root.left, k8, kids, nk);
            expand(n2, cp, root.right, 8 - k8, kids, nk);
        }

        // Octant-ordered slot assignment: slot s is visited first by rays
        // whose direction sign octant is s, so it gets the child lying
        // furthest toward -sign(s).
        Box nb;
        for (int k = 0; k < nk; ++k) nb.grow(n2[kids[k]].box);
        double ctr[3];
        for (int a = 0; a < 3; ++a) ctr[a] = 0.5 * ((double)nb.mn[a] + nb.mx[a]);
        double cost[8][8];
        for (int k = 0; k < nk; ++k) {
            const Box& b = n2[kids[k]].box;
            double d[3];
            for (int a = 0; a < 3; ++a) d[a] = 0.5 * ((double)b.mn[a] + b.mx[a]) - ctr[a];
            for (int s = 0; s < 8; ++s) {
                const double sx = (s & 4) ? -1.0 : 1.0, sy = (s & 2) ? -1.0 : 1.0, sz = (s & 1) ? -1.0 : 1.0;
                cost[k][s] = d[0] * sx + d[1] * sy + d[2] * sz;
            }
        }
        int slot_of[8];
        int kid_in_slot[8];
        for (int s = 0; s < 8; ++s) kid_in_slot[s] = -1;
        bool kid_done[8] = {false, false, false, false, false, false, false, false};
        for (int it = 0; it < nk; ++it) {
            double bc = std::numeric_limits<double>::infinity();
            int bk = -1, bs = -1;
            for (int k = 0; k < nk; ++k) {
                if (kid_done[k]) continue;
                for (int s = 0; s < 8; ++s) {
                    if (kid_in_slot[s] >= 0) continue;
                    if (cost[k][s] < bc) { bc = cost[k][s]; bk = k; bs = s; }
                }
            }
            kid_done[bk] = true;
            kid_in_slot[bs] = bk;
            slot_of[bk] = bs;
        }
        (void)slot_of;

        // Quantisation frame.  Every child box is widened by a margin of
        // 2^-18 of the node's magnitude before rounding outward, so the
        // decoded box keeps a slack far above FP32 slab-test rounding.
        float p[3];
        int e[3];
        double scale[3], margin[3];
        for (int a = 0; a < 3; ++a) {
            const double lo = nb.mn[a], hi = nb.mx[a];
            const double mag = std::max(std::max(std::fabs(lo), std::fabs(hi)), hi - lo);
            margin[a] = std::ldexp(std::max(mag, 1e-20), -18);
            p[a] = (float)(lo - margin[a]);
            if ((double)p[a] > lo - margin[a]) p[a] = std::nextafter(p[a], -std::numeric_limits<float>::infinity());
            e[a] = quant_exponent(hi + margin[a] - (double)p[a]);
            scale[a] = std::ldexp(1.0, e[a]);
        }

        uint8_t meta[8] = {0};
        uint8_t qlo[3][8] = {{0}}, qhi[3][8] = {{0}};
        uint8_t imask = 0;
        const uint32_t child_base = (uint32_t)(out.nodes.size() / 20);
        const uint32_t tri_base = tri_records;
        uint32_t n_internal = 0;
        // internal children first, in slot order, so that child index =
        // child_base + popcount(imask below slot)
        for (int s = 0; s < 8; ++s) {
            const int k = kid_in_slot[s];
            if (k < 0) continue;
            const Node2& cn = n2[kids[k]];
            for (int a = 0; a < 3; ++a) {
                double lo = std::floor(((double)cn.box.mn[a] - margin[a] - p[a]) / scale[a]);
                double hi = std::ceil(((double)cn.box.mx[a] + margin[a] - p[a]) / scale[a]);
                lo = std::min(std::max(lo, 0.0), 255.0);
                hi = std::min(std::max(hi, 0.0), 255.0);
                qlo[a][s] = (uint8_t)lo;
                qhi[a][s] = (uint8_t)hi;
            }
            if (cn.count > opt.max_leaf) {
                imask |= (uint8_t)(1u << s);
                meta[s] = (uint8_t)((1u << 5) | (24u + s));
                out.nodes.resize(out.nodes.size() + 20);
                q.push_back({child_base + n_internal, kids[k], pd.depth + 1});
                ++n_internal;
            }
        }
        for (int s = 0; s < 8; ++s) {
            const int k = kid_in_slot[s];
            if (k < 0) continue;
            const Node2& cn = n2[kids[k]];
            if (cn.count > opt.max_leaf) continue;
            const uint32_t off = tri_records - tri_base;
            const uint32_t cnt = (uint32_t)cn.count;
            meta[s] = (uint8_t)((((1u << cnt) - 1u) << 5) | off);
            ++out.num_leaf_slots;
            for (int i = cn.first; i < cn.first + cn.count; ++i) {
                const uint32_t t = c.idx[i];
                out.tris.resize(out.tris.size() + 16);
                write_tri(verts + 9 * (size_t)t, t, format, &out.tris[out.tris.size() - 16]);
                ++tri_records;
            }
            sah += cn.box.area() / root_area * cnt;
        }

        float* nd = &out.nodes[20 * (size_t)pd.slot8];
        nd[0] = p[0];
        nd[1] = p[1];
        nd[2] = p[2];
        nd[3] = u2f((uint32_t)(uint8_t)(int8_t)e[0] | ((uint32_t)(uint8_t)(int8_t)e[1] << 8) |
                    ((uint32_t)(uint8_t)(int8_t)e[2] << 16) | ((uint32_t)imask << 24));
        nd[4] = u2f(child_base);
        nd[5] = u2f(4u * tri_base);
        auto pack4 = [](const uint8_t* b) {
            return u2f((uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24));
        };
        nd[6] = pack4(meta);
        nd[7] = pack4(meta + 4);
        nd[8] = pack4(qlo[0]);
        nd[9] = pack4(qlo[0] + 4);
        nd[10] = pack4(qlo[1]);
        nd[11] = pack4(qlo[1] + 4);
        nd[12] = pack4(qlo[2]);
        nd[13] = pack4(qlo[2] + 4);
        nd[14] = pack4(qhi[0]);
        nd[15] = pack4(qhi[0] + 4);
        nd[16] = pack4(qhi[1]);
        nd[17] = pack4(qhi[1] + 4);
        nd[18] = pack4(qhi[2]);
        nd[19] = pack4(qhi[2] + 4);
    }
    out.num_nodes = (uint32_t)(out.nodes.size() / 20);
    out.num_tris = tri_records;
    out.sah_cost = sah;
    return tri_records == n_tris;
}

}  // namespace tfg
