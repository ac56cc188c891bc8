// lp_bvh.cpp -- host-side acceleration structures for the triangle search.
//
// The reference searches all T triangles for every step
// (FindTriangleInPath, src/Simulation/MCLightPropagation.cl:1037-1093).  The
// kernel here answers the same query -- nearest face along the ray, strict
// lowest-index tie-break, refTri excluded -- through a flattened binary BVH
// built below, and skips the search entirely for steps whose length is below
// a conservative distance-to-mesh bound (the safety grid).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <numeric>
#include <thread>

#include "lp_internal.hpp"

namespace lp {
namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float* p)
    {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    void grow(const Box& b)
    {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    double area() const
    {
        double d[3];
        for (int a = 0; a < 3; ++a) d[a] = std::max(0.0, (double)hi[a] - lo[a]);
        return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
    }
};

struct TNode {
    Box box;
    int32_t child[2] = {-1, -1};
    int32_t first = 0, count = 0;  // leaf if count > 0
    int32_t depth = 0;
};

}  // namespace

Bvh build_bvh(const float* tri, const int32_t* med, int64_t ntri, int max_leaf)
{
    Bvh out;
    max_leaf = std::min(std::max(max_leaf, 1), kMaxLeaf / 2);
    std::vector<Box> tb((size_t)ntri);
    std::vector<float> cen((size_t)ntri * 3);
    Box scene;
    for (int64_t t = 0; t < ntri; ++t) {
        for (int v = 0; v < 3; ++v) tb[t].grow(tri + 9 * t + 3 * v);
        for (int a = 0; a < 3; ++a) cen[3 * t + a] = 0.5f * (tb[t].lo[a] + tb[t].hi[a]);
        scene.grow(tb[t]);
    }
    double diag = 0, maxabs = 0;
    for (int a = 0; a < 3; ++a) {
        double d = (double)scene.hi[a] - scene.lo[a];
        diag += d * d;
        maxabs = std::max({maxabs, std::fabs((double)scene.hi[a]), std::fabs((double)scene.lo[a])});
    }
    diag = std::sqrt(diag);
    out.pad = (float)std::max(1e-5 * diag, 4e-6 * maxabs) + 1e-12f;

    std::vector<int32_t> idx((size_t)ntri);
    std::iota(idx.begin(), idx.end(), 0);
    std::vector<TNode> nodes;
    nodes.reserve((size_t)std::max<int64_t>(1, 2 * ntri / std::max(1, max_leaf) + 4));
    TNode root;
    root.first = 0;
    root.count = (int32_t)ntri;
    for (int64_t t = 0; t < ntri; ++t) root.box.grow(tb[t]);
    nodes.push_back(root);

    const int kBins = 16;
    const int kMaxDepth = 56;
    std::vector<int32_t> work{0};
    while (!work.empty()) {
        int32_t ni = work.back();
        work.pop_back();
        TNode n = nodes[ni];
        if (n.count <= max_leaf) continue;
        Box cb;
        for (int32_t i = n.first; i < n.first + n.count; ++i) cb.grow(&cen[3 * (size_t)idx[i]]);
        int best_axis = -1, best_split = -1;
        double best_cost = INFINITY;
        for (int a = 0; a < 3; ++a) {
            double ext = (double)cb.hi[a] - cb.lo[a];
            if (!(ext > 0)) continue;
            Box bb[kBins];
            int cnt[kBins] = {0};
            for (int32_t i = n.first; i < n.first + n.count; ++i) {
                int b = (int)((cen[3 * (size_t)idx[i] + a] - cb.lo[a]) / ext * kBins);
                b = std::min(std::max(b, 0), kBins - 1);
                cnt[b]++;
                bb[b].grow(tb[idx[i]]);
            }
            double la[kBins], ra[kBins];
            int lc[kBins], rc[kBins];
            Box acc;
            int c = 0;
            for (int b = 0; b < kBins; ++b) {
                acc.grow(bb[b]);
                c += cnt[b];
                la[b] = acc.area();
                lc[b] = c;
            }
            acc = Box();
            c = 0;
            for (int b = kBins - 1; b >= 0; --b) {
                acc.grow(bb[b]);
                c += cnt[b];
                ra[b] = acc.area();
                rc[b] = c;
            }
            for (int b = 0; b < kBins - 1; ++b) {
                if (lc[b] == 0 || rc[b + 1] == 0) continue;
                double cost = la[b] * lc[b] + ra[b + 1] * rc[b + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = b;
                }
            }
        }
        int32_t mid;
        double leaf_cost = n.box.area() * n.count;
        bool force = n.depth >= kMaxDepth - 1;
        if (best_axis < 0 || force) {
            // median split along the longest centroid axis (ties by index)
            int a = 0;
            for (int k = 1; k < 3; ++k)
                if (cb.hi[k] - cb.lo[k] > cb.hi[a] - cb.lo[a]) a = k;
            std::sort(idx.begin() + n.first, idx.begin() + n.first + n.count, [&](int32_t x, int32_t y) {
                float cx = cen[3 * (size_t)x + a], cy = cen[3 * (size_t)y + a];
                return cx < cy || (cx == cy && x < y);
            });
            mid = n.first + n.count / 2;
        } else {
            if (n.count <= 16 && best_cost + 0.5 * n.box.area() >= leaf_cost) {
                if (n.count <= 2 * max_leaf) continue;  // keep as leaf
            }
            double ext = (double)cb.hi[best_axis] - cb.lo[best_axis];
            auto it = std::partition(idx.begin() + n.first, idx.begin() + n.first + n.count, [&](int32_t x) {
                int b = (int)((cen[3 * (size_t)x + best_axis] - cb.lo[best_axis]) / ext * kBins);
                b = std::min(std::max(b, 0), kBins - 1);
                return b <= best_split;
            });
            mid = (int32_t)(it - idx.begin());
            if (mid == n.first || mid == n.first + n.count) mid = n.first + n.count / 2;
        }
        TNode L, R;
        L.first = n.first;
        L.count = mid - n.first;
        R.first = mid;
        R.count = n.first + n.count - mid;
        L.depth = R.depth = n.depth + 1;
        for (int32_t i = L.first; i < L.first + L.count; ++i) L.box.grow(tb[idx[i]]);
        for (int32_t i = R.first; i < R.first + R.count; ++i) R.box.grow(tb[idx[i]]);
        int32_t li = (int32_t)nodes.size();
        nodes.push_back(L);
        nodes.push_back(R);
        nodes[ni].child[0] = li;
        nodes[ni].child[1] = li + 1;
        nodes[ni].count = 0;
        work.push_back(li + 1);
        work.push_back(li);
    }

    // Flatten: every tree node becomes a threaded BvhNode.  The first
    // kBfsTop nodes are numbered breadth first, so the top levels are
    // contiguous (and can be staged in LDS); each subtree below them follows
    // depth first (pre-order), so a node's first child is the next node and
    // a traversal walks forward through memory.  The order changes where
    // nodes are stored, not the threaded visiting order.
    const size_t kBfsTop = 2048;
    std::vector<int32_t> order;
    std::vector<int32_t> flat_of(nodes.size(), -1);
    std::deque<int32_t> q;
    if (ntri > 0) q.push_back(0);
    while (!q.empty() && order.size() < kBfsTop) {
        int32_t t = q.front();
        q.pop_front();
        flat_of[t] = (int32_t)order.size();
        order.push_back(t);
        if (nodes[t].count == 0)
            for (int c = 0; c < 2; ++c) q.push_back(nodes[t].child[c]);
    }
    std::vector<int32_t> stack;
    for (int32_t r : q) {
        stack.push_back(r);
        while (!stack.empty()) {
            int32_t t = stack.back();
            stack.pop_back();
            flat_of[t] = (int32_t)order.size();
            order.push_back(t);
            if (nodes[t].count == 0) {
                stack.push_back(nodes[t].child[1]);
                stack.push_back(nodes[t].child[0]);
            }
        }
    }
    std::vector<int32_t> miss_of(nodes.size(), -1);
    for (int32_t t : order) {  // parents precede children in BFS order
        if (nodes[t].count == 0) {
            miss_of[nodes[t].child[0]] = flat_of[nodes[t].child[1]];
            miss_of[nodes[t].child[1]] = miss_of[t];
        }
    }
    out.tris.resize((size_t)ntri);
    for (int64_t i = 0; i < ntri; ++i) {
        int32_t t = idx[i];
        const float* K = tri + 9 * (size_t)t;
        const float* Lv = K + 3;
        const float* M = K + 6;
        TriRec& r = out.tris[i];
        for (int a = 0; a < 3; ++a) {
            r.m[a] = M[a];
            r.e[a] = K[a] - M[a];
            r.f[a] = Lv[a] - M[a];
        }
        r.id = t;
        r.med = med ? ((med[2 * t] & 0xffff) | (med[2 * t + 1] << 16)) : 0;
        r.pad = 0;
    }
    out.nodes.resize(order.size());
    for (size_t k = 0; k < order.size(); ++k) {
        const TNode& n = nodes[order[k]];
        BvhNode& bn = out.nodes[k];
        for (int a = 0; a < 3; ++a) {
            bn.lo[a] = n.box.lo[a] - out.pad;
            bn.hi[a] = n.box.hi[a] + out.pad;
        }
        if (n.count > 0)
            bn.hit = (int32_t)(0x80000000u | ((uint32_t)n.count << 24) | (uint32_t)n.first);
        else
            bn.hit = flat_of[n.child[0]];
        bn.miss = miss_of[order[k]];
    }
    int32_t depth = 0;
    for (auto& n : nodes) depth = std::max(depth, n.depth);
    out.depth = depth;
    return out;
}

// Chebyshev distance transform over cells touched by triangle boxes.  For a
// point in cell C and a triangle touching only cells at Chebyshev distance
// >= d from C, the point-triangle distance is >= (d - 1) * min cell edge.
SafetyGrid build_safety_grid(const Bvh& bvh, const float* tri, int64_t ntri, int cells)
{
    (void)bvh;
    SafetyGrid g;
    if (ntri == 0 || cells < 4) return g;
    Box sb;
    for (int64_t t = 0; t < ntri; ++t)
        for (int v = 0; v < 3; ++v) sb.grow(tri + 9 * t + 3 * v);
    double ext[3], mx = 0;
    for (int a = 0; a < 3; ++a) {
        ext[a] = (double)sb.hi[a] - sb.lo[a];
        mx = std::max(mx, ext[a]);
    }
    double h = mx / cells;
    if (!(h > 0)) return g;
    for (int a = 0; a < 3; ++a) {
        double lo = sb.lo[a] - 2 * h;
        g.n[a] = (int32_t)std::ceil((ext[a] + 4 * h) / h);
        g.origin[a] = (float)lo;
        g.inv_cell[a] = (float)(1.0 / h);
    }
    const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int64_t N = nx * ny * nz;
    const int32_t INF = 1 << 29;
    std::vector<int32_t> d((size_t)N, INF);
    std::vector<int64_t> frontier;
    auto cell_of = [&](double x, int a) {
        int64_t c = (int64_t)std::floor((x - g.origin[a]) / h);
        return std::min<int64_t>(std::max<int64_t>(c, 0), g.n[a] - 1);
    };
    for (int64_t t = 0; t < ntri; ++t) {
        Box b;
        for (int v = 0; v < 3; ++v) b.grow(tri + 9 * t + 3 * v);
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = cell_of((double)b.lo[a] - 1e-6 * mx, a);
            hi[a] = cell_of((double)b.hi[a] + 1e-6 * mx, a);
        }
        for (int64_t z = lo[2]; z <= hi[2]; ++z)
            for (int64_t y = lo[1]; y <= hi[1]; ++y)
                for (int64_t x = lo[0]; x <= hi[0]; ++x) {
                    int64_t c = (z * ny + y) * nx + x;
                    if (d[c] != 0) {
                        d[c] = 0;
                        frontier.push_back(c);
                    }
                }
    }
    // multi-source BFS with the 26-neighbourhood = Chebyshev distance
    std::vector<int64_t> next;
    for (int32_t dist = 1; !frontier.empty(); ++dist) {
        next.clear();
        for (int64_t c : frontier) {
            int64_t x = c % nx, y = (c / nx) % ny, z = c / (nx * ny);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int64_t X = x + dx, Y = y + dy, Z = z + dz;
                        if (X < 0 || Y < 0 || Z < 0 || X >= nx || Y >= ny || Z >= nz) continue;
                        int64_t k = (Z * ny + Y) * nx + X;
                        if (d[k] > dist) {
                            d[k] = dist;
                            next.push_back(k);
                        }
                    }
        }
        frontier.swap(next);
    }
    // safety = (d - 1) * h, shrunk by 2% and a scene-relative margin that
    // covers the rounding of the float triangle test
    double margin = 1e-5 * std::sqrt(ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2]);
    g.dist.resize((size_t)N);
    for (int64_t c = 0; c < N; ++c) {
        double s = (d[c] >= INF) ? 0.0 : (d[c] - 1) * h * 0.98 - margin;
        g.dist[c] = (float)std::max(0.0, s);
    }
    return g;
}

namespace {

// Squared distance from p to the triangle (a, a + ab, a + ac), in double
// (closest-point regions of Ericson, Real-Time Collision Detection 5.1.5).
double point_tri_dist2_d(const double p[3], const double a[3], const double ab[3], const double ac[3])
{
    auto dt = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double d1 = dt(ab, ap), d2 = dt(ac, ap);
    double q[3];
    auto len2 = [&]() { return dt(q, q); };
    if (d1 <= 0 && d2 <= 0) return dt(ap, ap);
    double bp[3] = {ap[0] - ab[0], ap[1] - ab[1], ap[2] - ab[2]};
    double d3 = dt(ab, bp), d4 = dt(ac, bp);
    if (d3 >= 0 && d4 <= d3) return dt(bp, bp);
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        double v = d1 / (d1 - d3);
        for (int i = 0; i < 3; ++i) q[i] = ap[i] - v * ab[i];
        return len2();
    }
    double cp[3] = {ap[0] - ac[0], ap[1] - ac[1], ap[2] - ac[2]};
    double d5 = dt(ab, cp), d6 = dt(ac, cp);
    if (d6 >= 0 && d5 <= d6) return dt(cp, cp);
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        double w = d2 / (d2 - d6);
        for (int i = 0; i < 3; ++i) q[i] = ap[i] - w * ac[i];
        return len2();
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int i = 0; i < 3; ++i) q[i] = bp[i] - w * (ac[i] - ab[i]);
        return len2();
    }
    double den = va + vb + vc;
    if (!(den != 0)) return 0.0;  // degenerate face: treat as touching
    double v = vb / den, w = vc / den;
    for (int i = 0; i < 3; ++i) q[i] = ap[i] - ab[i] * v - ac[i] * w;
    return len2();
}

}  // namespace

namespace {

// Listing radius of a cell: r0 + the cell's half diagonal + 2 pads (a point
// the kernel places in a cell lies within half_diag + 2 pad of its centre).
double near_list_radius(float r0, float cell, float bvh_pad)
{
    return (double)r0 + 0.5 * std::sqrt(3.0) * (double)cell + 2.0 * (double)bvh_pad;
}

// Origin and cell counts of a grid reaching r_list + cell beyond [lo, hi].
int64_t fit_frame(NearGrid& g, const float lo[3], const float hi[3], double r_list, float cell)
{
    int64_t ncell = 1;
    const float grow = (float)(r_list + cell);
    for (int a = 0; a < 3; ++a) {
        g.origin[a] = lo[a] - grow;
        g.n[a] = std::max<int32_t>(1, (int32_t)std::ceil(((double)hi[a] - lo[a] + 2.0 * grow) / cell) + 1);
        g.inv_cell[a] = 1.0f / cell;
        ncell *= g.n[a];
    }
    return ncell;
}

}  // namespace

NearGrid near_grid_frame(const float lo[3], const float hi[3], float r0, float cell, float pad)
{
    NearGrid g;
    if (!(r0 > 0) || !(cell > 0)) return g;
    fit_frame(g, lo, hi, near_list_radius(r0, cell, pad), cell);
    g.r0 = r0;
    return g;
}

NearGrid build_near_grid(const Bvh& bvh, float r0, float cell, const NearGrid* frame)
{
    NearGrid g;
    const int64_t nt = (int64_t)bvh.tris.size();
    if (nt == 0 || !(r0 > 0) || !(cell > 0)) return g;
    // A point that the kernel places in cell c (by the float arithmetic of
    // near_cell) lies within the cell's box grown by the rounding pad, so it
    // is within half_diag + 2 pad of the cell centre.  Cell c lists face f
    // when dist(centre, f) <= r0 + half_diag + 2 pad: every face within r0 of
    // such a point is listed.  Cell c is "touched" by f when
    // dist(centre, f) <= half_diag + 2 pad: a face only passes through
    // touched cells.
    const double pad = 2.0 * (double)bvh.pad;
    const double half_diag = 0.5 * std::sqrt(3.0) * (double)cell;
    const double r_list = near_list_radius(r0, cell, bvh.pad);
    const double r_touch = half_diag + pad;
    Box all;
    for (int64_t i = 0; i < nt; ++i) {
        const TriRec& t = bvh.tris[i];
        for (int k = 0; k < 3; ++k) {
            float v[3];
            for (int a = 0; a < 3; ++a) v[a] = t.m[a] + (k == 1 ? t.e[a] : k == 2 ? t.f[a] : 0.0f);
            all.grow(v);
        }
    }
    int64_t ncell = 1;
    if (frame) {
        // the given frame must reach r_list + cell beyond the faces, as a
        // fitted one does
        const float grow = (float)(r_list + cell);
        for (int a = 0; a < 3; ++a) {
            g.origin[a] = frame->origin[a];
            g.n[a] = frame->n[a];
            g.inv_cell[a] = frame->inv_cell[a];
            ncell *= g.n[a];
            const double frame_hi = (double)g.origin[a] + (double)g.n[a] * cell;
            if (!(g.origin[a] <= all.lo[a] - grow && frame_hi >= (double)all.hi[a] + grow &&
                  g.inv_cell[a] == 1.0f / cell))
                return NearGrid();
        }
    } else {
        ncell = fit_frame(g, all.lo, all.hi, r_list, cell);
    }
    if (ncell > ((int64_t)1 << 31) - 1 || nt > ((int64_t)1 << 31) - 1) return NearGrid();
    g.r0 = r0;
    // cell centres follow the float indexing of near_cell: centre of cell i
    // along axis a = origin + (i + 0.5) * cell
    auto centre = [&](int a, int32_t i) { return (double)g.origin[a] + ((double)i + 0.5) * (double)cell; };

    // (cell << 32) | face | touched-bit, generated per face range in threads
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::vector<uint64_t>> pairs(nth);
    auto work = [&](unsigned w) {
        int64_t f0 = nt * w / nth, f1 = nt * (w + 1) / nth;
        std::vector<uint64_t>& out = pairs[w];
        for (int64_t i = f0; i < f1; ++i) {
            const TriRec& t = bvh.tris[i];
            double A[3], AB[3], AC[3], N[3];
            for (int a = 0; a < 3; ++a) {
                A[a] = t.m[a];
                AB[a] = t.e[a];
                AC[a] = t.f[a];
            }
            N[0] = AB[1] * AC[2] - AB[2] * AC[1];
            N[1] = AB[2] * AC[0] - AB[0] * AC[2];
            N[2] = AB[0] * AC[1] - AB[1] * AC[0];
            double nl = std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
            int32_t c0[3], c1[3];
            for (int a = 0; a < 3; ++a) {
                double lo = std::min({A[a], A[a] + AB[a], A[a] + AC[a]}) - r_list;
                double hi = std::max({A[a], A[a] + AB[a], A[a] + AC[a]}) + r_list;
                c0[a] = std::max<int32_t>(0, (int32_t)std::floor((lo - g.origin[a]) / cell) - 1);
                c1[a] = std::min<int32_t>(g.n[a] - 1, (int32_t)std::floor((hi - g.origin[a]) / cell) + 1);
            }
            for (int32_t z = c0[2]; z <= c1[2]; ++z)
                for (int32_t y = c0[1]; y <= c1[1]; ++y)
                    for (int32_t x = c0[0]; x <= c1[0]; ++x) {
                        double P[3] = {centre(0, x), centre(1, y), centre(2, z)};
                        if (nl > 0) {  // plane pre-test
                            double sd = ((P[0] - A[0]) * N[0] + (P[1] - A[1]) * N[1] + (P[2] - A[2]) * N[2]) / nl;
                            if (std::fabs(sd) > r_list) continue;
                        }
                        double d2 = point_tri_dist2_d(P, A, AB, AC);
                        if (d2 > r_list * r_list) continue;
                        uint64_t cidx = ((uint64_t)z * g.n[1] + y) * g.n[0] + x;
                        out.push_back(cidx << 32 | (uint64_t)i | (d2 <= r_touch * r_touch ? 0x80000000ull : 0));
                    }
        }
    };
    {
        std::vector<std::thread> ts;
        for (unsigned w = 1; w < nth; ++w) ts.emplace_back(work, w);
        work(0);
        for (auto& t : ts) t.join();
    }
    // counting sort by cell; faces stay in increasing order within a cell
    std::vector<uint32_t> count((size_t)ncell + 1, 0);
    std::vector<uint8_t> touched((size_t)ncell, 0);
    uint64_t total = 0;
    for (auto& v : pairs)
        for (uint64_t q : v) {
            count[q >> 32]++;
            if (q & 0x80000000ull) touched[q >> 32] = 1;
            ++total;
        }
    if (total > 0x7fffffffull) return NearGrid();  // list positions are int32 in the kernel
    g.cells.assign(2 * ((size_t)ncell + 1), 0);
    uint64_t acc = 0;
    for (int64_t c = 0; c < ncell; ++c) {
        g.cells[2 * c] = g.cells[2 * c + 1] = (uint32_t)acc;
        acc += count[c];
    }
    g.cells[2 * ncell] = (uint32_t)acc;
    std::vector<uint32_t>().swap(count);
    g.list.resize((size_t)acc);
    for (auto& v : pairs) {
        for (uint64_t q : v) g.list[g.cells[2 * (q >> 32) + 1]++] = (int32_t)(q & 0x7fffffffull);  // odd slot = fill cursor
        std::vector<uint64_t>().swap(v);
    }
    // clearance: Chebyshev cell distance d from the touched cells (exact
    // two-pass chamfer transform over the 26-neighbourhood; each pass takes
    // the 9 cells of the previous plane and 3 of the previous row as whole
    // rows, then runs along the row); a point of a cell at distance d >= 2
    // is at least (d - 1) * cell from every face
    const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
    std::vector<uint16_t> dist((size_t)ncell);
    for (int64_t c = 0; c < ncell; ++c) dist[c] = touched[c] ? 0 : 0xfffe;
    std::vector<uint8_t>().swap(touched);
    auto row = [&](int64_t y, int64_t z) { return dist.data() + (z * ny + y) * nx; };
    std::vector<uint16_t> tmp((size_t)nx);
    for (int pass = 0; pass < 2; ++pass) {
        const int64_t dir = pass == 0 ? 1 : -1;
        for (int64_t zi = 0; zi < nz; ++zi) {
            const int64_t z = pass == 0 ? zi : nz - 1 - zi;
            for (int64_t yi = 0; yi < ny; ++yi) {
                const int64_t y = pass == 0 ? yi : ny - 1 - yi;
                uint16_t* r = row(y, z);
                // previous rows: (y - dir, z) and the 3 rows of plane z - dir
                const uint16_t* prev[4];
                int np = 0;
                if (y - dir >= 0 && y - dir < ny) prev[np++] = row(y - dir, z);
                if (z - dir >= 0 && z - dir < nz)
                    for (int64_t dy = -1; dy <= 1; ++dy)
                        if (y + dy >= 0 && y + dy < ny) prev[np++] = row(y + dy, z - dir);
                if (np) {
                    for (int64_t x = 0; x < nx; ++x) {
                        uint16_t m = 0xfffe;
                        for (int k = 0; k < np; ++k) m = std::min(m, prev[k][x]);
                        tmp[x] = m;
                    }
                    // 3-wide minimum along x, + 1
                    for (int64_t x = 0; x < nx; ++x) {
                        uint16_t m = tmp[x];
                        if (x > 0) m = std::min(m, tmp[x - 1]);
                        if (x + 1 < nx) m = std::min(m, tmp[x + 1]);
                        r[x] = std::min<uint16_t>(r[x], (uint16_t)std::min(0xfffe, m + 1));
                    }
                }
                // along the row itself
                if (dir > 0)
                    for (int64_t x = 1; x < nx; ++x) r[x] = std::min<uint16_t>(r[x], (uint16_t)std::min(0xfffe, r[x - 1] + 1));
                else
                    for (int64_t x = nx - 2; x >= 0; --x)
                        r[x] = std::min<uint16_t>(r[x], (uint16_t)std::min(0xfffe, r[x + 1] + 1));
            }
        }
    }
    for (int64_t c = 0; c <= ncell; ++c) {
        double cl = (c < ncell && dist[c] >= 2 && dist[c] < 0xfffe) ? (dist[c] - 1) * (double)cell * 0.999 - pad : 0.0;
        float f = (float)std::max(0.0, cl);
        std::memcpy(&g.cells[2 * c + 1], &f, sizeof f);
    }
    return g;
}

bool build_multi_mesh(const float* tri, const int32_t* med, const std::vector<int64_t>& first, int lds_per_part,
                      float r0, float cell, MultiMesh& out, std::string& why)
{
    const int ng = (int)first.size() - 1;
    out = MultiMesh();
    std::vector<Bvh> part((size_t)ng);
    Box all;
    float pad = 0.0f;
    for (int g = 0; g < ng; ++g) {
        const int64_t f = first[g], n = first[g + 1] - first[g];
        part[g] = build_bvh(tri + 9 * f, med + 2 * f, n, 4);
        // the vertices as build_near_grid sees them (m, m + e, m + f in
        // float), so that each part's box lies in the frame fitted to this one
        for (const TriRec& t : part[g].tris)
            for (int k = 0; k < 3; ++k) {
                float v[3];
                for (int a = 0; a < 3; ++a) v[a] = t.m[a] + (k == 1 ? t.e[a] : k == 2 ? t.f[a] : 0.0f);
                all.grow(v);
            }
        pad = std::max(pad, part[g].pad);
        out.bvh.depth = std::max(out.bvh.depth, part[g].depth);
    }
    out.bvh.pad = pad;
    // node placement: the tops of all parts, then the rest of each part
    std::vector<int32_t> top((size_t)ng), top_base((size_t)ng), rest_base((size_t)ng);
    int64_t ntop = 0, nall = 0;
    for (int g = 0; g < ng; ++g) {
        const int64_t n = (int64_t)part[g].nodes.size();
        top[g] = (int32_t)std::min<int64_t>(n, std::max(1, lds_per_part));
        top_base[g] = (int32_t)ntop;
        ntop += top[g];
        nall += n;
    }
    if (nall > INT32_MAX || first[ng] >= (1 << 24)) {
        why = "merged BVH exceeds its index fields";
        return false;
    }
    int64_t rest = ntop;
    for (int g = 0; g < ng; ++g) {
        rest_base[g] = (int32_t)rest;
        rest += (int64_t)part[g].nodes.size() - top[g];
    }
    out.bvh.nodes.resize((size_t)nall);
    out.bvh.tris.resize((size_t)first[ng]);
    out.root.resize((size_t)ng);
    for (int g = 0; g < ng; ++g) {
        auto place = [&](int32_t i) { return i < 0 ? i : i < top[g] ? top_base[g] + i : rest_base[g] + (i - top[g]); };
        const Bvh& b = part[g];
        for (size_t k = 0; k < b.nodes.size(); ++k) {
            BvhNode n = b.nodes[k];
            if (n.hit >= 0) {
                n.hit = place(n.hit);
            } else {  // leaf: its range moves to the part's place in the merged leaf order
                const uint32_t lf = (uint32_t)n.hit & 0xffffffu, cnt = ((uint32_t)n.hit >> 24) & 0x7fu;
                n.hit = (int32_t)(0x80000000u | (cnt << 24) | (uint32_t)(lf + first[g]));
            }
            n.miss = place(n.miss);
            out.bvh.nodes[(size_t)place((int32_t)k)] = n;
        }
        for (size_t i = 0; i < b.tris.size(); ++i) {
            TriRec r = b.tris[i];
            r.id += (int32_t)first[g];
            out.bvh.tris[(size_t)first[g] + i] = r;
        }
        out.root[g] = top_base[g];
    }
    // near grids in one frame, part after part
    out.coff.assign((size_t)ng, 0u);
    if (!(r0 > 0) || first[ng] == 0) return true;
    const NearGrid frame = near_grid_frame(all.lo, all.hi, r0, cell, pad);
    const int64_t ncell = (int64_t)frame.n[0] * frame.n[1] * frame.n[2];
    if ((ncell + 1) * ng > (int64_t)UINT32_MAX) {
        why = "merged near grid exceeds its index fields";
        return false;
    }
    out.ng = frame;
    for (int g = 0; g < ng; ++g) {
        NearGrid q = build_near_grid(part[g], r0, cell, &frame);
        if (q.cells.size() != (size_t)(2 * (ncell + 1))) {
            why = "near grid of a geometry does not fit the shared frame";
            return false;
        }
        const uint64_t base = out.ng.list.size();
        if (base + q.list.size() > (uint64_t)UINT32_MAX) {
            why = "merged near lists exceed 2^32 entries";
            return false;
        }
        out.coff[g] = (uint32_t)(ncell + 1) * (uint32_t)g;
        for (int64_t c = 0; c <= ncell; ++c) {
            out.ng.cells.push_back(q.cells[2 * c] + (uint32_t)base);
            out.ng.cells.push_back(q.cells[2 * c + 1]);
        }
        for (int32_t e : q.list) out.ng.list.push_back(e + (int32_t)first[g]);
    }
    return true;
}

}  // namespace lp

extern "C" int lp_debug_multi_mesh(const float* tri, const int32_t* med, const int64_t* tri_count, int32_t ngeom,
                                   int32_t lds_per_part, float r0, float cell, int64_t sizes[6], uint32_t* nodes,
                                   int32_t* tri_ids, int32_t* roots, uint32_t* coffs, uint32_t* cells, int32_t* list)
{
    if (!tri || !med || !tri_count || ngeom < 1 || ngeom > lp::kMaxGeoms || !sizes) return 1;
    std::vector<int64_t> first(1, 0);
    for (int32_t g = 0; g < ngeom; ++g) {
        if (tri_count[g] < 1) return 1;
        first.push_back(first.back() + tri_count[g]);
    }
    lp::MultiMesh mm;
    std::string why;
    if (!lp::build_multi_mesh(tri, med, first, lds_per_part, r0, cell, mm, why)) return 2;
    const int64_t ncell = (int64_t)mm.ng.n[0] * mm.ng.n[1] * mm.ng.n[2];
    sizes[0] = (int64_t)mm.bvh.nodes.size();
    sizes[1] = (int64_t)mm.bvh.tris.size();
    sizes[2] = (int64_t)mm.ng.cells.size() / 2;
    sizes[3] = (int64_t)mm.ng.list.size();
    sizes[4] = mm.ng.cells.empty() ? 0 : ncell + 1;
    sizes[5] = 0;
    if (nodes && !mm.bvh.nodes.empty()) std::memcpy(nodes, mm.bvh.nodes.data(), mm.bvh.nodes.size() * sizeof(lp::BvhNode));
    if (tri_ids)
        for (size_t i = 0; i < mm.bvh.tris.size(); ++i) tri_ids[i] = mm.bvh.tris[i].id;
    if (roots) std::memcpy(roots, mm.root.data(), mm.root.size() * sizeof(int32_t));
    if (coffs) std::memcpy(coffs, mm.coff.data(), mm.coff.size() * sizeof(uint32_t));
    if (cells && !mm.ng.cells.empty()) std::memcpy(cells, mm.ng.cells.data(), mm.ng.cells.size() * sizeof(uint32_t));
    if (list && !mm.ng.list.empty()) std::memcpy(list, mm.ng.list.data(), mm.ng.list.size() * sizeof(int32_t));
    return 0;
}
