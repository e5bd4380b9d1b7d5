// Test tool: a CPU restatement, in FP32, of the reference's own cell
// construction, to count the cells where the reference's output departs from
// the exact diagram and to say why.  Not part of the library.
//
// Restated (behaviour, not code) from:
//  * voronoi_convex_cell.cuh / faster_convex_cell.cuh: the cell (64 planes,
//    64 vertices; status 1 / 2 when full), the box planes and eight corner
//    vertices of CCInit (:179-230 / :198-254), homogeneous cofactor vertices
//    (:232-254 / :256-278), the bisector plane normalised by |s - p|
//    (:263-270 / :287-294), the removal test dot3(v, n) + v.w d > 0
//    (:256-260 / :280-285), the unsigned too-far test (:367-374 / :391-408),
//    the bloom-filtered duplicate check and the intersection pre-check of
//    CCClipByPlane (:376-469 / :411-512), the boundary walk of
//    CCComputeBoundary with its 1000-step limit (status 5) (:298-364 /
//    :322-388), garbage collection (:471-511 / :517-557) and CCUpdateBounds
//    (:529-553 / :575-599);
//  * voronoi_ultra.cu:39-141 / laguerre_ultra.cu:43-174: the kNN pre-clip by
//    initial guesses, the leaf lambda (clip, update bounds, GC at 0.85 x 64),
//    the node lower bounds (d^2 / 4; the power form with its clamps), and the
//    output rule (owners of the planes the final vertices reference);
//  * voronoi_bvh.cuh:40-288 / pwr_bvh.cuh:104-356: the octant corner radii,
//    the shrinking best-first query with a 64-entry unsorted stack whose push
//    evicts the farthest entry when full.
// Stand-ins, stated because they limit what the counts mean: IEEE FP32
// arithmetic for CUDA's fast-math (approximate division and square root),
// this file's binary radix hierarchy over 30-bit Morton codes (leaf size 10)
// for cuBQL's builder, an exact kNN for cuBQL's, and the box of the exact
// point AABB for the BVH root box.  So the eviction and predicate counts are
// of this restatement; capacity and too-far outcomes depend little on them.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <vector>

#include <omp.h>

namespace {

struct F4 {
    float x, y, z, w;
};
inline float dot3(const F4 &a, const F4 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

constexpr int kMaxPlanes = 64, kMaxVerts = 64, kStack = 64, kEnd = 255;

uint32_t expand10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

struct Node {
    float lo[3], hi[3], wmax;
    int32_t child[2];  // >= 0 node, else -1 - leaf
    int32_t first, count;
};

struct Tree {
    std::vector<F4> pts;        // original order
    std::vector<int32_t> prim;  // Morton order -> original id
    std::vector<uint32_t> keys;
    std::vector<Node> nodes;    // [0] the root
    float box_lo[3], box_hi[3]; // root box -1 / +1 (voronoi_ultra.cu:237-238)
};

int32_t build(Tree &T, int32_t a, int32_t b, int leaf) {
    const int32_t me = (int32_t)T.nodes.size();
    T.nodes.push_back(Node());
    Node nd;
    for (int k = 0; k < 3; ++k) nd.lo[k] = FLT_MAX, nd.hi[k] = -FLT_MAX;
    nd.wmax = -FLT_MAX;
    for (int32_t i = a; i < b; ++i) {
        const F4 &p = T.pts[T.prim[i]];
        nd.lo[0] = std::min(nd.lo[0], p.x), nd.hi[0] = std::max(nd.hi[0], p.x);
        nd.lo[1] = std::min(nd.lo[1], p.y), nd.hi[1] = std::max(nd.hi[1], p.y);
        nd.lo[2] = std::min(nd.lo[2], p.z), nd.hi[2] = std::max(nd.hi[2], p.z);
        nd.wmax = std::max(nd.wmax, p.w);
    }
    nd.first = a, nd.count = 0;
    nd.child[0] = nd.child[1] = 0;
    if (b - a <= leaf) {
        nd.count = b - a;
        T.nodes[me] = nd;
        return me;
    }
    int32_t split;
    const uint32_t ka = T.keys[a], kb = T.keys[b - 1];
    if (ka == kb) {
        split = (a + b) / 2;
    } else {
        const int bit = 31 - __builtin_clz(ka ^ kb);
        int32_t lo = a, hi = b - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi) / 2;
            if ((T.keys[mid] >> bit) & 1u) hi = mid;
            else lo = mid + 1;
        }
        split = lo;
    }
    nd.child[0] = build(T, a, split, leaf);
    nd.child[1] = build(T, split, b, leaf);
    T.nodes[me] = nd;
    return me;
}

void build_tree(Tree &T, int64_t n, const float *p, const float *w, int leaf) {
    T.pts.resize(n);
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = 0; i < n; ++i) {
        T.pts[i] = F4{p[3 * i], p[3 * i + 1], p[3 * i + 2], w ? w[i] : 0.0f};
        for (int k = 0; k < 3; ++k) mn[k] = std::min(mn[k], p[3 * i + k]), mx[k] = std::max(mx[k], p[3 * i + k]);
    }
    for (int k = 0; k < 3; ++k) {
        volatile float a = mn[k] - 1.0f, b = mx[k] + 1.0f;
        T.box_lo[k] = a, T.box_hi[k] = b;
    }
    const float ex = std::max(std::max(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]);
    const float scale = ex > 0.0f ? 1024.0f / ex : 0.0f;
    std::vector<std::pair<uint32_t, int32_t>> kv(n);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) q[k] = (uint32_t)std::min(std::max((p[3 * i + k] - mn[k]) * scale, 0.0f), 1023.0f);
        kv[i] = {(expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]), (int32_t)i};
    }
    std::sort(kv.begin(), kv.end());
    T.prim.resize(n), T.keys.resize(n);
    for (int64_t i = 0; i < n; ++i) T.keys[i] = kv[i].first, T.prim[i] = kv[i].second;
    build(T, 0, (int32_t)n, leaf);
}

void knn(const Tree &T, int32_t q, int k, std::vector<int32_t> &out, std::vector<int32_t> &stack) {
    const F4 s = T.pts[q];
    std::vector<std::pair<float, int32_t>> best;
    auto bd = [&](const Node &nd) {
        const float dx = std::max(std::max(nd.lo[0] - s.x, s.x - nd.hi[0]), 0.0f);
        const float dy = std::max(std::max(nd.lo[1] - s.y, s.y - nd.hi[1]), 0.0f);
        const float dz = std::max(std::max(nd.lo[2] - s.z, s.z - nd.hi[2]), 0.0f);
        return dx * dx + dy * dy + dz * dz;
    };
    stack.clear();
    stack.push_back(0);
    while (!stack.empty()) {
        const Node &nd = T.nodes[stack.back()];
        stack.pop_back();
        const float worst = (int)best.size() < k ? FLT_MAX : best.back().first;
        if (bd(nd) > worst) continue;
        if (nd.count) {
            for (int32_t i = nd.first; i < nd.first + nd.count; ++i) {
                const int32_t j = T.prim[i];
                if (j == q) continue;
                const F4 &p = T.pts[j];
                const float d = (p.x - s.x) * (p.x - s.x) + (p.y - s.y) * (p.y - s.y) + (p.z - s.z) * (p.z - s.z);
                if ((int)best.size() == k && d >= best.back().first) continue;
                best.insert(std::upper_bound(best.begin(), best.end(), std::make_pair(d, j)), std::make_pair(d, j));
                if ((int)best.size() > k) best.pop_back();
            }
            continue;
        }
        const float d0 = bd(T.nodes[nd.child[0]]), d1 = bd(T.nodes[nd.child[1]]);
        const bool n0 = d0 <= d1;
        stack.push_back(n0 ? nd.child[1] : nd.child[0]);
        stack.push_back(n0 ? nd.child[0] : nd.child[1]);
    }
    out.clear();
    for (auto &b : best) out.push_back(b.second);
}

enum Flags : int32_t {
    kEvicted = 1,        // the traversal stack evicted an entry (candidates possibly lost)
    kTooFarBehind = 2,   // a point skipped as too far lies behind the seed (t < 0): it could hide the cell
    kDroppedZeroDiff = 4,  // a child with a zero prune margin was not pushed
    kPlaneFull = 8,      // status 1 was written (the last status written wins, so it may not be final)
    kVertexFull = 16,    // status 2 was written
    kWalkFailed = 32,    // status 5 was written
};

struct Cell {
    const Tree *T;
    int32_t id;
    F4 seed;
    F4 planes[kMaxPlanes];
    int32_t nbr[kMaxPlanes];
    uint8_t vtx[kMaxVerts][3];
    F4 vpos[kMaxVerts];
    uint8_t next[kMaxPlanes];
    int nv = 0, np = 0;
    float max_radius = 1e20f;
    float lb[3], ub[3];
    uint64_t bloom = 0;
    int32_t status = 0;
    int32_t flags = 0;

    static int hash_vid(int v) { return (int)(((uint32_t)v * 2654435761u) >> 26) & 63; }

    F4 vertex(const uint8_t *t) const {
        const F4 &a = planes[t[0]], &b = planes[t[1]], &c = planes[t[2]];
        const float m12 = b.x * c.y - b.y * c.x, m13 = b.x * c.z - b.z * c.x, m14 = b.x * c.w - b.w * c.x;
        const float m23 = b.y * c.z - b.z * c.y, m24 = b.y * c.w - b.w * c.y, m34 = b.z * c.w - b.w * c.z;
        return F4{-a.w * m23 - a.y * m34 + a.z * m24, a.x * m34 + a.w * m13 - a.z * m14,
                  -a.x * m24 + a.y * m14 - a.w * m12, a.x * m23 - a.y * m13 + a.z * m12};
    }
    bool beyond(int v, const F4 &e) const { return dot3(vpos[v], e) + vpos[v].w * e.w > 0.0f; }

    void init(int32_t i) {
        id = i;
        seed = T->pts[i];
        status = 0, flags = 0, bloom = 0;
        const float eps = 1e-5f;
        const float vmin[3] = {T->box_lo[0] - eps, T->box_lo[1] - eps, T->box_lo[2] - eps};
        const float vmax[3] = {T->box_hi[0] + eps, T->box_hi[1] + eps, T->box_hi[2] + eps};
        planes[0] = F4{1, 0, 0, -vmin[0]}, planes[1] = F4{-1, 0, 0, vmax[0]};
        planes[2] = F4{0, 1, 0, -vmin[1]}, planes[3] = F4{0, -1, 0, vmax[1]};
        planes[4] = F4{0, 0, 1, -vmin[2]}, planes[5] = F4{0, 0, -1, vmax[2]};
        np = 6;
        const uint8_t corners[8][3] = {{2, 5, 0}, {5, 3, 0}, {1, 5, 2}, {5, 1, 3}, {4, 2, 0}, {4, 0, 3}, {2, 4, 1}, {4, 3, 1}};
        for (int v = 0; v < 8; ++v) {
            for (int e = 0; e < 3; ++e) vtx[v][e] = corners[v][e];
            vpos[v] = vertex(vtx[v]);
        }
        nv = 8;
        for (int p = 0; p < kMaxPlanes; ++p) nbr[p] = -1;
        for (int k = 0; k < 3; ++k) {
            const float s = k == 0 ? seed.x : (k == 1 ? seed.y : seed.z);
            lb[k] = vmin[k] - s, ub[k] = vmax[k] - s;
        }
        float m2 = 0;
        for (int k = 0; k < 3; ++k) {
            const float m = std::max(std::fabs(lb[k]), std::fabs(ub[k]));
            m2 += m * m;
        }
        max_radius = std::sqrt(m2);
    }

    bool has_plane(int v) const {
        if (!(bloom & (1ull << hash_vid(v)))) return false;
        for (int i = 0; i < np; ++i)
            if (nbr[i] == v) return true;
        return false;
    }
    bool too_far(const F4 &p) {
        const float wd = p.w - seed.w;
        const float dx = seed.x - p.x, dy = seed.y - p.y, dz = seed.z - p.z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        float l = wd - d2;
        l = l * l;
        const bool far = l > 4.0f * max_radius * max_radius * d2;
        if (far && d2 + seed.w - p.w < 0.0f) flags |= kTooFarBehind;
        return far;
    }

    // The boundary of the removed triangles (nb_removed of them, at the tail)
    // as next-plane links; END when none; status 5 after 1000 steps.
    int boundary(int removed) {
        for (int i = 0; i < kMaxPlanes; ++i) next[i] = kEnd;
        int first = kEnd, iter = 0, t = nv;
        while (removed > 0) {
            if (iter++ > 1000) {
                status = 5;
                flags |= kWalkFailed;
                return first;
            }
            const uint8_t *v = vtx[t];
            bool in_border[3], opp[3];
            for (int e = 0; e < 3; ++e) in_border[e] = next[v[e]] != kEnd;
            for (int e = 0; e < 3; ++e) opp[e] = next[v[(e + 1) % 3]] == v[e];
            bool simple = true;
            for (int e = 0; e < 3; ++e)
                if (!opp[e] && !opp[(e + 1) % 3] && in_border[(e + 1) % 3]) simple = false;
            if (!opp[0] && !opp[1] && !opp[2]) {
                if (first == kEnd) {
                    for (int e = 0; e < 3; ++e) next[v[e]] = v[(e + 1) % 3];
                    first = v[0];
                } else {
                    simple = false;
                }
            }
            if (!simple) {
                if (++t == nv + removed) t = nv;
                continue;
            }
            for (int e = 0; e < 3; ++e)
                if (!opp[e]) next[v[e]] = v[(e + 1) % 3];
            for (int e = 0; e < 3; ++e)
                if (opp[e] && opp[(e + 1) % 3]) {
                    if (first == v[(e + 1) % 3]) first = next[v[(e + 1) % 3]];
                    next[v[(e + 1) % 3]] = kEnd;
                }
            for (int e = 0; e < 3; ++e) std::swap(vtx[t][e], vtx[nv + removed - 1][e]);
            t = nv;
            removed--;
        }
        return first;
    }

    bool clip(int32_t vid, const F4 &p) {
        if (status == 1) return false;
        if (too_far(p) || has_plane(vid)) return false;
        F4 dir{seed.x - p.x, seed.y - p.y, seed.z - p.z, 0};
        F4 ave{seed.x + p.x, seed.y + p.y, seed.z + p.z, 0};
        const float dt = dot3(ave, dir) - seed.w + p.w;
        const float nrm = std::sqrt(dot3(dir, dir));
        const F4 e{dir.x / nrm, dir.y / nrm, dir.z / nrm, -dt / 2 / nrm};
        bool hit = false;
        for (int k = 0; k < nv && !hit; ++k) hit = beyond(k, e);
        if (!hit) return false;
        if (np >= kMaxPlanes) {
            status = 1;
            flags |= kPlaneFull;
            return false;
        }
        const int pi = np++;
        planes[pi] = e;
        nbr[pi] = vid;
        if (vid >= 0) bloom |= 1ull << hash_vid(vid);
        int removed = 0;
        for (int i = 0; i < nv;) {
            if (beyond(i, e)) {
                nv--;
                for (int c = 0; c < 3; ++c) std::swap(vtx[i][c], vtx[nv][c]);
                std::swap(vpos[i], vpos[nv]);
                removed++;
            } else {
                i++;
            }
        }
        if (nv < 1) {
            status = 3;
            np--;
            return false;
        }
        const int first = boundary(removed);
        if (status != 0 || first == kEnd) return false;
        int cir = first, it = 0;
        bool changed = false;
        do {
            if (++it > 1000) return changed;
            const int nx = next[cir];
            if (nv >= kMaxVerts) {
                status = 2;
                flags |= kVertexFull;
                return changed;
            }
            vtx[nv][0] = (uint8_t)pi, vtx[nv][1] = (uint8_t)cir, vtx[nv][2] = (uint8_t)nx;
            vpos[nv] = vertex(vtx[nv]);
            nv++;
            cir = nx;
            changed = true;
        } while (cir != first);
        return changed;
    }

    void gc() {
        uint8_t map[kMaxPlanes];
        for (int i = 0; i < np; ++i) map[i] = kEnd;
        for (int v = 0; v < nv; ++v)
            for (int e = 0; e < 3; ++e) map[vtx[v][e]] = 0;
        int n = 0;
        for (int i = 0; i < np; ++i) {
            if (map[i] != 0) continue;
            planes[n] = planes[i], nbr[n] = nbr[i];
            map[i] = (uint8_t)n++;
        }
        for (int v = 0; v < nv; ++v)
            for (int e = 0; e < 3; ++e) vtx[v][e] = map[vtx[v][e]];
        np = n;
        bloom = 0;
        for (int i = 0; i < np; ++i)
            if (nbr[i] >= 0) bloom |= 1ull << hash_vid(nbr[i]);
    }

    void update_bounds() {
        float up[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX}, lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, m2 = 0;
        for (int v = 0; v < nv; ++v) {
            const float iw = 1.0f / vpos[v].w;
            const float d[3] = {vpos[v].x * iw - seed.x, vpos[v].y * iw - seed.y, vpos[v].z * iw - seed.z};
            m2 = std::max(m2, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            for (int k = 0; k < 3; ++k) up[k] = std::max(up[k], d[k]), lo[k] = std::min(lo[k], d[k]);
        }
        for (int k = 0; k < 3; ++k) ub[k] = up[k], lb[k] = lo[k];
        max_radius = std::sqrt(m2);
    }
};

struct Radii {
    float r[8];
    void set(const float *l, const float *u) {
        for (int c = 0; c < 8; ++c) {
            const float x = (c & 1) ? u[0] : l[0], y = (c & 2) ? u[1] : l[1], z = (c & 4) ? u[2] : l[2];
            r[c] = x * x + y * y + z * z;
        }
    }
    float safety(const F4 &s, const Node &nd) const {
        const bool i0 = nd.lo[0] <= s.x, i1 = nd.lo[1] <= s.y, i2 = nd.lo[2] <= s.z;
        const bool i3 = nd.hi[0] >= s.x, i4 = nd.hi[1] >= s.y, i5 = nd.hi[2] >= s.z;
        float m = 0.0f;
        for (int c = 0; c < 8; ++c) {
            const bool in = ((c & 1) ? i3 : i0) && ((c & 2) ? i4 : i1) && ((c & 4) ? i5 : i2);
            if (in) m = std::max(m, r[c]);
        }
        return m;
    }
};

// lower bound of the squared distance from the seed to the plane of any
// point of the node: d^2 / 4, or the power form (laguerre_ultra.cu:98-122)
float node_bound(const F4 &s, const Node &nd, bool power) {
    const float dx = std::max(std::max(nd.lo[0] - s.x, s.x - nd.hi[0]), 0.0f);
    const float dy = std::max(std::max(nd.lo[1] - s.y, s.y - nd.hi[1]), 0.0f);
    const float dz = std::max(std::max(nd.lo[2] - s.z, s.z - nd.hi[2]), 0.0f);
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (!power) return d2 * 0.25f;
    if (d2 < 1e-12f) return 0.0f;
    const float num = d2 + std::min(s.w - nd.wmax, 0.0f);
    if (num <= 0.0f) return 0.0f;
    const float t = num / (2.0f * std::sqrt(d2));
    return t * t;
}

}  // namespace

extern "C" {

// Raw rows of the reference restatement for the cells `cells` (all when
// null): counts, statuses, flags (bits above) and the rows' neighbour ids in
// plane order (unsorted; the caller sorts).  Returns the entry total, -1 when
// out_cap is too small.
int64_t refsim_run(int64_t n, const float *pts, const float *w, int32_t knn_k, int32_t leaf, const int64_t *cells,
                   int64_t m, int32_t *counts, int32_t *status, int32_t *flags, int32_t *out, int64_t out_cap,
                   int32_t nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    Tree T;
    build_tree(T, n, pts, w, leaf);
    const bool power = w != nullptr;
    std::vector<std::vector<int32_t>> rows(m);
#pragma omp parallel
    {
        Cell c;
        c.T = &T;
        std::vector<int32_t> nn, kst;
        std::vector<std::pair<float, int32_t>> stack;  // (bound, node)
#pragma omp for schedule(dynamic, 16)
        for (int64_t t = 0; t < m; ++t) {
            const int32_t i = cells ? (int32_t)cells[t] : (int32_t)t;
            c.init(i);
            Radii R;
            for (int k = 0; k < 8; ++k) R.r[k] = 1e10f;
            if (knn_k > 0 && n > 1) {  // the kNN pre-clip (initial guesses)
                knn(T, i, knn_k, nn, kst);
                for (int32_t g : nn) c.clip(g, T.pts[g]);
                c.update_bounds();
                R.set(c.lb, c.ub);
            }
            // shrinking best-first query
            stack.clear();
            const F4 s = c.seed;
            int32_t node = 0;
            bool go = true;
            while (go) {
                while (T.nodes[node].count == 0) {
                    const Node &nd = T.nodes[node];
                    const int32_t i0 = nd.child[0], i1 = nd.child[1];
                    const float r0 = R.safety(s, T.nodes[i0]), r1 = R.safety(s, T.nodes[i1]);
                    const float d0 = node_bound(s, T.nodes[i0], power), d1 = node_bound(s, T.nodes[i1], power);
                    const float f0 = d0 - r0, f1 = d1 - r1;
                    if (std::min(f0, f1) > 0.0f) {
                        node = -1;
                        break;
                    }
                    int32_t far;
                    float fd;
                    if (f0 < f1) node = i0, far = i1, fd = d1;
                    else node = i1, far = i0, fd = d0;
                    if (std::max(f0, f1) < 0.0f) {
                        if ((int)stack.size() < kStack) {
                            stack.emplace_back(fd, far);
                        } else {
                            int fp = 0;
                            for (int k = 1; k < kStack; ++k)
                                if (stack[k].first > stack[fp].first) fp = k;
                            c.flags |= kEvicted;
                            if (fd < stack[fp].first) stack[fp] = {fd, far};
                        }
                    } else if (std::max(f0, f1) == 0.0f) {
                        c.flags |= kDroppedZeroDiff;
                    }
                }
                if (node >= 0 && T.nodes[node].count) {
                    const Node &lf = T.nodes[node];
                    bool changed = false;
                    for (int32_t k = lf.first; k < lf.first + lf.count; ++k) {
                        const int32_t j = T.prim[k];
                        if (j == i) continue;
                        if (c.clip(j, T.pts[j])) {
                            changed = true;
                            c.update_bounds();
                        }
                    }
                    if (changed && c.np >= kMaxPlanes * 0.85) c.gc();
                    R.set(c.lb, c.ub);
                }
                bool found = false;
                while (!stack.empty()) {
                    size_t mp = 0;
                    for (size_t k = 1; k < stack.size(); ++k)
                        if (stack[k].first < stack[mp].first) mp = k;
                    const auto e = stack[mp];
                    stack[mp] = stack.back();
                    stack.pop_back();
                    if (e.first - R.safety(s, T.nodes[e.second]) <= 0.0f) {
                        node = e.second;
                        found = true;
                        break;
                    }
                }
                go = found;
            }
            std::vector<int32_t> &r = rows[t];
            r.clear();
            for (int v = 0; v < c.nv; ++v)
                for (int e = 0; e < 3; ++e) {
                    const int32_t o = c.nbr[c.vtx[v][e]];
                    if (o >= 0) r.push_back(o);
                }
            std::sort(r.begin(), r.end());
            r.erase(std::unique(r.begin(), r.end()), r.end());
            counts[t] = (int32_t)r.size();
            status[t] = c.status;
            flags[t] = c.flags;
        }
    }
    int64_t e = 0;
    for (int64_t t = 0; t < m; ++t) {
        if (e + (int64_t)rows[t].size() > out_cap) return -1;
        std::copy(rows[t].begin(), rows[t].end(), out + e);
        e += (int64_t)rows[t].size();
    }
    return e;
}

}  // extern "C"
