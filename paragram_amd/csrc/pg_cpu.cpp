// The diagram on host cores (C++/OpenMP): the clipping algorithm of the HIP
// kernels (pg_group_cell.hpp, pg_cell.hip), one thread per cell.  A reported
// CPU baseline and a host stand-in for the engine; declared in
// include/paragram_cpu.h.
//
// The same algorithm as the GPU path, step for step:
//  * the FP32 box of voronoi_ultra.cu:237-238 / voronoi_convex_cell.cuh:187-193;
//  * points sorted by 30-bit Morton code on the cubic 1024^3 grid, a binary
//    radix hierarchy over the codes (leaves of at most leaf_size points),
//    child boxes with the maximum weight;
//  * each cell clipped first by its 16 nearest neighbours (nearest first),
//    then a depth-first traversal, the nearer child first, that skips a box
//    when the octant corner-radius test or the per-vertex ball test proves no
//    point in it can cut the cell, and skips candidates nearer than the 16th
//    neighbour (already offered) or whose plane misses the cell's AABB;
//  * planes a . y <= H in seed-relative doubled coordinates y = 2 (x - s),
//    vertices as plane triples ordered so that D > 0, positions N / D kept in
//    FP32 with a rigorous error radius; a removal test is decided in FP32
//    when its value exceeds its error bound, else recomputed in FP64 from
//    the FP32 inputs, else decided exactly (expansion arithmetic,
//    pg_exact.hpp);
//  * rows = sorted original ids of the owners of the planes the final
//    vertices reference (voronoi_ultra.cu:125-138), every exact duplicate of
//    an owner included; status 0, 3 (empty) or 5 (inconsistent boundary);
//  * CSR offsets, and the union symmetrisation (adjacency.cuh:78-159) only
//    when some edge lacks its reverse.
// Differences are those of the machine: cells live in growable arrays (no
// capacity tiers), and divisions are IEEE instead of hardware reciprocals
// (both inside the error bounds), so the output is the same diagram.
#include "pg_exact.hpp"

#include "../../include/paragram_cpu.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

namespace {

thread_local std::string g_err;
int fail(const std::string &m) {
    g_err = m;
    return 1;
}

constexpr float kG32 = 2e-6f;
constexpr float kTiny32 = 1e-30f;
constexpr double kG64 = 1e-13;
constexpr double kTiny64 = 1e-280;
constexpr float kSlack = 1e-6f;
constexpr int kPreclip = 16;

struct P4 {
    float x, y, z, w;
};

// ------------------------------------------------------------------ index
uint32_t expand10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

struct BNode {
    float lo[3], hi[3], wmax;
    int32_t child[2];  // >= 0 node index, else leaf: -1 - leaf number
};
struct Leaf {
    int32_t first, count;
    float lo[3], hi[3], wmax;
};

struct Index {
    int64_t n = 0;
    float box_lo[3], box_hi[3];
    std::vector<P4> pts;          // Morton-sorted
    std::vector<int32_t> perm;    // sorted -> original
    std::vector<uint32_t> keys;
    std::vector<BNode> nodes;
    std::vector<Leaf> leaves;
    int32_t root = 0;             // >= 0 node, else leaf
    std::vector<int32_t> dup_next;  // next member of a duplicate group (sorted ids), -1 at the end
    std::vector<int32_t> dup_head;
    bool has_dups = false;
};

void bounds_of(const Index &I, int32_t a, int32_t b, float lo[3], float hi[3], float &wmax) {
    for (int k = 0; k < 3; ++k) lo[k] = FLT_MAX, hi[k] = -FLT_MAX;
    wmax = -FLT_MAX;
    for (int32_t i = a; i < b; ++i) {
        const P4 &p = I.pts[i];
        lo[0] = std::min(lo[0], p.x), hi[0] = std::max(hi[0], p.x);
        lo[1] = std::min(lo[1], p.y), hi[1] = std::max(hi[1], p.y);
        lo[2] = std::min(lo[2], p.z), hi[2] = std::max(hi[2], p.z);
        wmax = std::max(wmax, p.w);
    }
}

// Binary radix hierarchy over the sorted codes [a, b): split where the
// highest differing bit changes (the middle for equal codes).
int32_t build_rec(Index &I, int32_t a, int32_t b, int leaf, float lo[3], float hi[3], float &wmax) {
    if (b - a <= leaf) {
        Leaf L;
        L.first = a, L.count = b - a;
        bounds_of(I, a, b, L.lo, L.hi, L.wmax);
        for (int k = 0; k < 3; ++k) lo[k] = L.lo[k], hi[k] = L.hi[k];
        wmax = L.wmax;
        I.leaves.push_back(L);
        return -1 - (int32_t)(I.leaves.size() - 1);
    }
    int32_t split;
    const uint32_t ka = I.keys[a], kb = I.keys[b - 1];
    if (ka == kb) {
        split = (a + b) / 2;
    } else {
        const int bit = 31 - __builtin_clz(ka ^ kb);
        int32_t lo_i = a, hi_i = b - 1;  // first index with the bit set
        while (lo_i < hi_i) {
            const int32_t mid = (lo_i + hi_i) / 2;
            if ((I.keys[mid] >> bit) & 1u) hi_i = mid;
            else lo_i = mid + 1;
        }
        split = lo_i;
    }
    const int32_t me = (int32_t)I.nodes.size();
    I.nodes.push_back(BNode());
    float l0[3], h0[3], w0, l1[3], h1[3], w1;
    const int32_t c0 = build_rec(I, a, split, leaf, l0, h0, w0);
    const int32_t c1 = build_rec(I, split, b, leaf, l1, h1, w1);
    BNode &nd = I.nodes[me];
    nd.child[0] = c0, nd.child[1] = c1;
    for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(l0[k], l1[k]);
        hi[k] = std::max(h0[k], h1[k]);
    }
    wmax = std::max(w0, w1);
    for (int k = 0; k < 3; ++k) nd.lo[k] = lo[k], nd.hi[k] = hi[k];
    nd.wmax = wmax;
    return me;
}

void build_index(Index &I, int64_t n, const float *pts, const float *w, int leaf) {
    I.n = n;
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) mn[k] = std::min(mn[k], pts[3 * i + k]), mx[k] = std::max(mx[k], pts[3 * i + k]);
    for (int k = 0; k < 3; ++k) {
        // two FP32 roundings in this order (voronoi_ultra.cu:237-238, voronoi_convex_cell.cuh:187-193)
        volatile float a = mn[k] - 1.0f, b = mx[k] + 1.0f;
        I.box_lo[k] = a - 1e-5f;
        I.box_hi[k] = b + 1e-5f;
    }
    const float ex = std::max(std::max(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]);
    const float scale = ex > 0.0f ? 1024.0f / ex : 0.0f;
    std::vector<std::pair<uint32_t, int32_t>> kv((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) {
            float t = (pts[3 * i + k] - mn[k]) * scale;
            t = std::min(std::max(t, 0.0f), 1023.0f);
            q[k] = (uint32_t)t;
        }
        kv[i] = {(expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]), (int32_t)i};
    }
    std::sort(kv.begin(), kv.end());  // ties by index: a stable order
    I.pts.resize(n), I.perm.resize(n), I.keys.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t o = kv[i].second;
        I.keys[i] = kv[i].first;
        I.perm[i] = o;
        I.pts[i] = P4{pts[3 * o], pts[3 * o + 1], pts[3 * o + 2], w ? w[o] : 0.0f};
    }
    float lo[3], hi[3], wm;
    I.root = build_rec(I, 0, (int32_t)n, leaf, lo, hi, wm);
    // exact duplicates: groups of equal position and weight (sorted by value)
    std::vector<int32_t> ord(n);
    for (int64_t i = 0; i < n; ++i) ord[i] = (int32_t)i;
    auto key = [&](int32_t i, int k) { return k == 0 ? I.pts[i].x : k == 1 ? I.pts[i].y : k == 2 ? I.pts[i].z : I.pts[i].w; };
    std::sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
        for (int k = 0; k < 4; ++k)
            if (key(a, k) != key(b, k)) return key(a, k) < key(b, k);
        return a < b;
    });
    I.dup_head.assign(n, -1);
    I.dup_next.assign(n, -1);
    for (int64_t i = 0; i < n;) {
        int64_t j = i + 1;
        while (j < n && I.pts[ord[j]].x == I.pts[ord[i]].x && I.pts[ord[j]].y == I.pts[ord[i]].y &&
               I.pts[ord[j]].z == I.pts[ord[i]].z && I.pts[ord[j]].w == I.pts[ord[i]].w)
            ++j;
        for (int64_t t = i; t < j; ++t) {
            I.dup_head[ord[t]] = ord[i];
            I.dup_next[ord[t]] = t + 1 < j ? ord[t + 1] : -1;
        }
        if (j - i > 1) I.has_dups = true;
        i = j;
    }
}

// k nearest neighbours (sorted ids, itself excluded, nearest first) and the
// squared distance of the last, by a depth-first search nearer box first.
void knn_of(const Index &I, int32_t q, int k, int32_t *ids, float &kth, std::vector<int32_t> &stack) {
    const P4 s = I.pts[q];
    float best_d[kPreclip];
    int32_t best_i[kPreclip];
    int m = 0;
    auto worst = [&]() { return m < k ? FLT_MAX : best_d[m - 1]; };
    auto box_d2 = [&](const float *lo, const float *hi) {
        const float dx = std::max(std::max(lo[0] - s.x, s.x - hi[0]), 0.0f);
        const float dy = std::max(std::max(lo[1] - s.y, s.y - hi[1]), 0.0f);
        const float dz = std::max(std::max(lo[2] - s.z, s.z - hi[2]), 0.0f);
        return dx * dx + dy * dy + dz * dz;
    };
    stack.clear();
    stack.push_back(I.root);
    while (!stack.empty()) {
        const int32_t r = stack.back();
        stack.pop_back();
        if (r < 0) {
            const Leaf &L = I.leaves[-1 - r];
            if (box_d2(L.lo, L.hi) > worst()) continue;
            for (int32_t j = L.first; j < L.first + L.count; ++j) {
                if (j == q) continue;
                const float dx = I.pts[j].x - s.x, dy = I.pts[j].y - s.y, dz = I.pts[j].z - s.z;
                const float d = dx * dx + dy * dy + dz * dz;
                if (m == k && d >= best_d[m - 1]) continue;
                int t = m < k ? m++ : m - 1;
                while (t > 0 && best_d[t - 1] > d) {
                    best_d[t] = best_d[t - 1], best_i[t] = best_i[t - 1];
                    --t;
                }
                best_d[t] = d, best_i[t] = j;
            }
            continue;
        }
        const BNode &nd = I.nodes[r];
        float d0, d1;
        {
            const int32_t c0 = nd.child[0], c1 = nd.child[1];
            const float *lo0 = c0 < 0 ? I.leaves[-1 - c0].lo : I.nodes[c0].lo;
            const float *hi0 = c0 < 0 ? I.leaves[-1 - c0].hi : I.nodes[c0].hi;
            const float *lo1 = c1 < 0 ? I.leaves[-1 - c1].lo : I.nodes[c1].lo;
            const float *hi1 = c1 < 0 ? I.leaves[-1 - c1].hi : I.nodes[c1].hi;
            d0 = box_d2(lo0, hi0), d1 = box_d2(lo1, hi1);
        }
        const float wv = worst();
        const bool near0 = d0 <= d1;
        const int32_t a = near0 ? nd.child[0] : nd.child[1], b = near0 ? nd.child[1] : nd.child[0];
        const float da = near0 ? d0 : d1, db = near0 ? d1 : d0;
        if (db <= wv) stack.push_back(b);
        if (da <= wv) stack.push_back(a);
    }
    for (int t = 0; t < k; ++t) ids[t] = t < m ? best_i[t] : q;
    kth = m == k ? best_d[k - 1] : FLT_MAX;
}

// ------------------------------------------------------------------ cell
struct Plane {
    float ax, ay, az, H, pH;
    int owner;  // sorted point id, or -1 - (2 axis + side) for a box plane
};
struct Vert {
    float x, y, z, w;  // y~ and the radius (see store)
    int p[3];          // plane slots, D > 0
};

struct Stats {
    int64_t rechecks = 0, exact = 0, unresolved = 0, clips = 0, candidates = 0, nodes = 0;
};

class Cell {
  public:
    const Index *I;
    int32_t seed;
    float sx, sy, sz, sw;
    std::vector<Plane> pl;
    std::vector<Vert> vs, nv;
    std::vector<int> rem;
    std::vector<std::pair<int, int>> edges;
    std::vector<std::pair<float, Plane>> cs;  // candidates of one offer
    std::vector<int32_t> kstack;              // kNN search stack
    int status = 0;
    float L[3], U[3], EM = 0.0f;
    bool bounded = true;
    float guess_r2 = 0.0f;
    size_t gc_at = 64;
    Stats *st;

    void fp64_plane(int owner, double a[3], double &H, double pa[3], double &pH) const {
        if (owner >= 0) {
            const P4 &p = I->pts[owner];
            a[0] = (double)p.x - sx, a[1] = (double)p.y - sy, a[2] = (double)p.z - sz;
            const double dw = (double)p.w - sw;
            const double a2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
            H = a2 - dw;
            pH = a2 + std::fabs(dw);
            for (int k = 0; k < 3; ++k) pa[k] = std::fabs(a[k]);
        } else {
            const int code = -1 - owner, axis = code >> 1, side = code & 1;
            const double s = axis == 0 ? sx : (axis == 1 ? sy : sz);
            for (int k = 0; k < 3; ++k) {
                a[k] = k == axis ? (side ? 1.0 : -1.0) : 0.0;
                pa[k] = k == axis ? 1.0 : 0.0;
            }
            H = side ? 2.0 * ((double)I->box_hi[axis] - s) : 2.0 * (s - (double)I->box_lo[axis]);
            pH = std::fabs(H);
        }
    }
    void fp64_vertex(const int *t, double N[3], double &D, double pN[3], double &pD) const {
        double a[3][3], H[3], pa[3][3], pH[3];
        for (int e = 0; e < 3; ++e) fp64_plane(pl[t[e]].owner, a[e], H[e], pa[e], pH[e]);
        double c[3][3], pc[3][3];
        for (int e = 0; e < 3; ++e) {
            const double *u = a[(e + 1) % 3], *v = a[(e + 2) % 3], *pu = pa[(e + 1) % 3], *pv = pa[(e + 2) % 3];
            c[e][0] = u[1] * v[2] - u[2] * v[1];
            c[e][1] = u[2] * v[0] - u[0] * v[2];
            c[e][2] = u[0] * v[1] - u[1] * v[0];
            pc[e][0] = pu[1] * pv[2] + pu[2] * pv[1];
            pc[e][1] = pu[2] * pv[0] + pu[0] * pv[2];
            pc[e][2] = pu[0] * pv[1] + pu[1] * pv[0];
        }
        D = a[0][0] * c[0][0] + a[0][1] * c[0][1] + a[0][2] * c[0][2];
        pD = pa[0][0] * pc[0][0] + pa[0][1] * pc[0][1] + pa[0][2] * pc[0][2];
        for (int k = 0; k < 3; ++k) {
            N[k] = H[0] * c[0][k] + H[1] * c[1][k] + H[2] * c[2][k];
            pN[k] = pH[0] * pc[0][k] + pH[1] * pc[1][k] + pH[2] * pc[2][k];
        }
    }

    // w = eps (1 + s) + 2e-6 max |y~_k| (the position error and the filter's
    // own rounding term); negative: left to FP64; +inf stays +inf
    static Vert store(float x, float y, float z, float w, const int *t) {
        if (w >= 0.0f && w < FLT_MAX) w = w * (1.0f + kSlack) + kG32 * std::max(std::fabs(x), std::max(std::fabs(y), std::fabs(z)));
        Vert v;
        v.x = x, v.y = y, v.z = z, v.w = w;
        v.p[0] = t[0], v.p[1] = t[1], v.p[2] = t[2];
        return v;
    }

    Vert make_vertex(int i, int j, int k) {
        const Plane &A = pl[i], &B = pl[j], &C = pl[k];
        const float c1x = B.ay * C.az - B.az * C.ay, c1y = B.az * C.ax - B.ax * C.az, c1z = B.ax * C.ay - B.ay * C.ax;
        const float c2x = C.ay * A.az - C.az * A.ay, c2y = C.az * A.ax - C.ax * A.az, c2z = C.ax * A.ay - C.ay * A.ax;
        const float c3x = A.ay * B.az - A.az * B.ay, c3y = A.az * B.ax - A.ax * B.az, c3z = A.ax * B.ay - A.ay * B.ax;
        const float D = A.ax * c1x + A.ay * c1y + A.az * c1z;
        const float Nx = A.H * c1x + B.H * c2x + C.H * c3x;
        const float Ny = A.H * c1y + B.H * c2y + C.H * c3y;
        const float Nz = A.H * c1z + B.H * c2z + C.H * c3z;
        const float ax = std::fabs(A.ax), ay = std::fabs(A.ay), az = std::fabs(A.az);
        const float bx = std::fabs(B.ax), by = std::fabs(B.ay), bz = std::fabs(B.az);
        const float cx = std::fabs(C.ax), cy = std::fabs(C.ay), cz = std::fabs(C.az);
        const float q1x = by * cz + bz * cy, q1y = bz * cx + bx * cz, q1z = bx * cy + by * cx;
        const float q2x = cy * az + cz * ay, q2y = cz * ax + cx * az, q2z = cx * ay + cy * ax;
        const float q3x = ay * bz + az * by, q3y = az * bx + ax * bz, q3z = ax * by + ay * bx;
        const float pD = ax * q1x + ay * q1y + az * q1z;
        const float pNx = A.pH * q1x + B.pH * q2x + C.pH * q3x;
        const float pNy = A.pH * q1y + B.pH * q2y + C.pH * q3y;
        const float pNz = A.pH * q1z + B.pH * q2z + C.pH * q3z;
        const int t[3] = {i, j, k};
        const float eD = kG32 * pD;
        if (D > 2.0f * eD && pD > kTiny32) {
            const float den = (1.0f + kSlack) / (D - eD);
            const float x = Nx / D, y = Ny / D, z = Nz / D;
            const float ex = (kG32 * pNx + std::fabs(x) * eD) * den + std::fabs(x) * kSlack;
            const float ey = (kG32 * pNy + std::fabs(y) * eD) * den + std::fabs(y) * kSlack;
            const float ez = (kG32 * pNz + std::fabs(z) * eD) * den + std::fabs(z) * kSlack;
            const float w = std::max(ex, std::max(ey, ez)) + kTiny32;
            // a radius beyond 2^-12 of the position (nearly parallel or
            // coaxial planes) goes to FP64, as in the FP64-capable GPU tiers
            if (w <= 0x1p-12f * std::max(std::fabs(x), std::max(std::fabs(y), std::fabs(z))))
                return store(x, y, z, w, t);
        }
        // FP64 position; where its bound is loose (more than 2^-12 of the
        // position) the exact one; a radius of +inf when neither bounds it
        double N[3], DD, pN[3], pDD;
        fp64_vertex(t, N, DD, pN, pDD);
        const double e64 = kG64 * pDD;
        float xs[3] = {0.f, 0.f, 0.f}, eps = INFINITY;
        if (DD > 2.0 * e64 && pDD > kTiny64) {
            float ym = 0.0f;
            eps = 0.0f;
            for (int m = 0; m < 3; ++m) {
                const double X = N[m] / DD;
                const double E =
                    (kG64 * pN[m] + std::fabs(X) * e64) / (DD - e64) * (1.0 + 1e-12) + std::fabs(X) * 1e-12;
                xs[m] = (float)X;
                eps = std::max(eps, (float)E + std::fabs(xs[m]) * kSlack);
                ym = std::max(ym, std::fabs(xs[m]));
            }
            eps += kTiny32;
            if (eps <= ym * 0x1p-12f) return store(xs[0], xs[1], xs[2], eps, t);
        }
        pg::xp::PlaneIn in[3];
        for (int e = 0; e < 3; ++e) {
            const int o = pl[t[e]].owner;
            if (o >= 0) {
                const P4 &p = I->pts[o];
                in[e] = pg::xp::PlaneIn{p.x, p.y, p.z, p.w, -1};
            } else {
                in[e] = pg::xp::PlaneIn{0.f, 0.f, 0.f, 0.f, -1 - o};
            }
        }
        const float s4[4] = {sx, sy, sz, sw};
        float ye[3], we;
        if (pg::xp::exact_position(in, s4, I->box_lo, I->box_hi, ye, we) && we + kTiny32 < eps)
            return store(ye[0], ye[1], ye[2], we + kTiny32, t);
        return store(xs[0], xs[1], xs[2], eps, t);
    }

    // vertex v strictly beyond plane q?  FP32 filter, FP64 recheck, exact
    bool beyond(const Vert &v, const Plane &q, float pa1, float c2) {
        const float Pv = std::fma(q.ax, v.x, std::fma(q.ay, v.y, std::fma(q.az, v.z, -q.H)));
        const float err = std::fma(pa1, v.w, c2);
        if (std::fabs(Pv) > err) return Pv > 0.0f;
        st->rechecks++;
        double N[3], D, pN[3], pD, a[3], H, pa[3], pH;
        fp64_vertex(v.p, N, D, pN, pD);
        fp64_plane(q.owner, a, H, pa, pH);
        const double P = a[0] * N[0] + a[1] * N[1] + a[2] * N[2] - H * D;
        const double pP = pa[0] * pN[0] + pa[1] * pN[1] + pa[2] * pN[2] + pH * pD;
        const bool d_sure = std::fabs(D) > kG64 * pD && pD > kTiny64;
        if (d_sure && std::fabs(P) > kG64 * pP && pP > kTiny64) return (P > 0.0) == (D > 0.0);
        // an exact duplicate of an owner of the vertex's planes gives the same plane: P = 0
        const P4 &qp = I->pts[q.owner];
        for (int e = 0; e < 3; ++e) {
            const int o = pl[v.p[e]].owner;
            if (o < 0) continue;
            const P4 &r = I->pts[o];
            if (r.x == qp.x && r.y == qp.y && r.z == qp.z && r.w == qp.w) return false;
        }
        pg::xp::PlaneIn in[4];
        for (int e = 0; e < 4; ++e) {
            const int o = e < 3 ? pl[v.p[e]].owner : q.owner;
            if (o >= 0) {
                const P4 &p = I->pts[o];
                in[e] = pg::xp::PlaneIn{p.x, p.y, p.z, p.w, -1};
            } else {
                in[e] = pg::xp::PlaneIn{0.f, 0.f, 0.f, 0.f, -1 - o};
            }
        }
        const float s4[4] = {sx, sy, sz, sw};
        int sD = 0, sP = 0;
        if (pg::xp::exact_signs(in, s4, I->box_lo, I->box_hi, sD, sP)) {
            st->exact++;
            return sD * sP > 0;
        }
        st->unresolved++;
        return (P > 0.0) == (D >= 0.0);
    }

    void update_bounds() {
        for (int k = 0; k < 3; ++k) L[k] = FLT_MAX, U[k] = -FLT_MAX;
        EM = 0.0f;
        bounded = true;
        for (const Vert &v : vs) {
            bounded &= v.w < FLT_MAX;
            L[0] = std::min(L[0], v.x - v.w), U[0] = std::max(U[0], v.x + v.w);
            L[1] = std::min(L[1], v.y - v.w), U[1] = std::max(U[1], v.y + v.w);
            L[2] = std::min(L[2], v.z - v.w), U[2] = std::max(U[2], v.z + v.w);
            EM = std::max(EM, v.w);
        }
    }

    void init(int32_t s) {
        seed = s;
        const P4 &p = I->pts[s];
        sx = p.x, sy = p.y, sz = p.z, sw = p.w;
        status = 0;
        guess_r2 = 0.0f;
        gc_at = 64;
        pl.clear();
        vs.clear();
        for (int g = 0; g < 6; ++g) {
            const int axis = g >> 1, side = g & 1;
            const float sv = axis == 0 ? sx : (axis == 1 ? sy : sz);
            const float h = side ? 2.0f * (I->box_hi[axis] - sv) : 2.0f * (sv - I->box_lo[axis]);
            Plane q{0.f, 0.f, 0.f, h, std::fabs(h), -1 - g};
            (axis == 0 ? q.ax : axis == 1 ? q.ay : q.az) = side ? 1.0f : -1.0f;
            pl.push_back(q);
        }
        for (int c = 0; c < 8; ++c) {
            const int cx = c >> 2 & 1, cy = c >> 1 & 1, cz = c & 1;
            const int X = cx, Y = 2 + cy, Z = 4 + cz;
            const int sgn = (cx ? 1 : -1) * (cy ? 1 : -1) * (cz ? 1 : -1);
            vs.push_back(sgn > 0 ? make_vertex(X, Y, Z) : make_vertex(X, Z, Y));
        }
        update_bounds();
    }

    bool too_far(const Plane &c) const {
        if (!bounded) return false;
        const float S = std::max(c.ax * L[0], c.ax * U[0]) + std::max(c.ay * L[1], c.ay * U[1]) +
                        std::max(c.az * L[2], c.az * U[2]);
        const float M = std::fabs(c.ax) * std::max(std::fabs(L[0]), std::fabs(U[0])) +
                        std::fabs(c.ay) * std::max(std::fabs(L[1]), std::fabs(U[1])) +
                        std::fabs(c.az) * std::max(std::fabs(L[2]), std::fabs(U[2]));
        return (c.H - S) > kSlack * (c.pH + M) + kTiny32;
    }

    // the box of a node cannot hold a point that cuts the cell
    bool node_prunable(const float *lo, const float *hi, float wmax, float &key) const {
        const float alx = lo[0] - sx, aly = lo[1] - sy, alz = lo[2] - sz;
        const float ahx = hi[0] - sx, ahy = hi[1] - sy, ahz = hi[2] - sz;
        const float dx = std::max(std::max(alx, -ahx), 0.0f), dy = std::max(std::max(aly, -ahy), 0.0f),
                    dz = std::max(std::max(alz, -ahz), 0.0f);
        const float d2 = dx * dx + dy * dy + dz * dz;
        const float dw = wmax - sw;
        key = d2 - dw;
        if (!bounded || d2 == 0.0f) return false;
        const float Lq[3] = {L[0] * L[0], L[1] * L[1], L[2] * L[2]}, Uq[3] = {U[0] * U[0], U[1] * U[1], U[2] * U[2]};
        const float rho2 = (std::max(alx <= 0.0f ? Lq[0] : 0.0f, ahx >= 0.0f ? Uq[0] : 0.0f) +
                            std::max(aly <= 0.0f ? Lq[1] : 0.0f, ahy >= 0.0f ? Uq[1] : 0.0f) +
                            std::max(alz <= 0.0f ? Lq[2] : 0.0f, ahz >= 0.0f ? Uq[2] : 0.0f)) *
                           (1.0f + kSlack);
        const float d2lo = d2 * (1.0f - kSlack);
        const float wslack = kSlack * (std::fabs(wmax) + std::fabs(sw)) + kTiny32;
        const float h = d2lo - dw - wslack;
        if (4.0f * d2lo > rho2 && h > 0.0f && h * h > rho2 * d2 * (1.0f + kSlack)) return true;
        // per-vertex balls: vertex y is beyond p's plane iff |2a - y|^2 < |y|^2 + 4 dw
        const float Ax = 2.0f * alx - EM, Ay = 2.0f * aly - EM, Az = 2.0f * alz - EM;
        const float Bx = 2.0f * ahx + EM, By = 2.0f * ahy + EM, Bz = 2.0f * ahz + EM;
        const float dw4 = 4.0f * (dw + wslack);
        for (const Vert &v : vs) {
            const float rx = std::fabs(v.x) + v.w, ry = std::fabs(v.y) + v.w, rz = std::fabs(v.z) + v.w;
            const float r2 = (rx * rx + ry * ry + rz * rz) * (1.0f + 4.0f * kSlack);
            const float tx = v.x - std::min(std::max(v.x, Ax), Bx);
            const float ty = v.y - std::min(std::max(v.y, Ay), By);
            const float tz = v.z - std::min(std::max(v.z, Az), Bz);
            const float dd = tx * tx + ty * ty + tz * tz;
            if (!(dd * (1.0f - kSlack) > r2 + dw4)) return false;
        }
        return true;
    }

    void garbage_collect() {
        std::vector<int> remap(pl.size(), -1);
        for (const Vert &v : vs)
            for (int e = 0; e < 3; ++e) remap[v.p[e]] = 0;
        int m = 0;
        for (size_t p = 0; p < pl.size(); ++p) {
            if (remap[p] < 0) continue;
            remap[p] = m;
            pl[m++] = pl[p];
        }
        pl.resize(m);
        for (Vert &v : vs)
            for (int e = 0; e < 3; ++e) v.p[e] = remap[v.p[e]];
    }

    bool has_plane(int owner) const {
        for (const Plane &p : pl)
            if (p.owner == owner) return true;
        return false;
    }

    // Offers plane q; true when the cell changed
    bool clip(const Plane &q) {
        rem.clear();
        const float pa1 = std::fabs(q.ax) + std::fabs(q.ay) + std::fabs(q.az);
        const float c2 = std::fma(kG32, q.pH, kTiny32);
        for (size_t i = 0; i < vs.size(); ++i)
            if (beyond(vs[i], q, pa1, c2)) rem.push_back((int)i);
        if (rem.empty()) return false;
        if (rem.size() == vs.size()) {  // voronoi_convex_cell.cuh:430-435
            status = 3;
            vs.clear();
            return true;
        }
        if (pl.size() >= gc_at) {
            garbage_collect();
            gc_at = std::max<size_t>(64, 2 * pl.size());
        }
        const int qs = (int)pl.size();
        pl.push_back(q);
        // boundary: directed edges of removed triangles whose twin is not removed
        edges.clear();
        for (int r : rem) {
            const int *t = vs[r].p;
            for (int e = 0; e < 3; ++e) {
                const int x = t[e], y = t[(e + 1) % 3];
                bool twin = false;
                for (int u : rem) {
                    const int *o = vs[u].p;
                    if ((o[0] == y && o[1] == x) || (o[1] == y && o[2] == x) || (o[2] == y && o[0] == x)) {
                        twin = true;
                        break;
                    }
                }
                if (!twin) edges.emplace_back(x, y);
            }
        }
        // every plane of the boundary cycle starts one edge and ends one
        long sa = 0, sb = 0;
        for (auto &e : edges) sa += (long)e.first * e.first + 1, sb += (long)e.second * e.second + 1;
        if (edges.size() < 3 || sa != sb) {
            status = 5;
            return true;
        }
        nv.clear();
        size_t r = 0;
        for (size_t i = 0; i < vs.size(); ++i) {
            if (r < rem.size() && rem[r] == (int)i) {
                ++r;
                continue;
            }
            nv.push_back(vs[i]);
        }
        for (auto &e : edges) nv.push_back(make_vertex(qs, e.first, e.second));
        vs.swap(nv);
        update_bounds();
        st->clips++;
        return true;
    }

    // Candidate j: false when it cannot cut (or was offered by the pre-clip)
    bool prepare(int32_t j, Plane &c, bool &emptied, bool &maybe_guess) const {
        const P4 &p = I->pts[j];
        c.ax = p.x - sx, c.ay = p.y - sy, c.az = p.z - sz;
        const float dw = p.w - sw;
        c.owner = j;
        if (c.ax == 0.0f && c.ay == 0.0f && c.az == 0.0f) {
            emptied = dw > 0.0f;  // coincident point: empty iff w_p > w_s
            return false;
        }
        const float a2 = c.ax * c.ax + c.ay * c.ay + c.az * c.az;
        if (a2 < guess_r2) return false;
        maybe_guess = guess_r2 == 0.0f || a2 < guess_r2 * 1.00003f;
        c.H = a2 - dw;
        c.pH = a2 + std::fabs(dw);
        return !too_far(c);
    }

    // Clips by the candidates (nearest plane first)
    void offer(const int32_t *ids, int m, bool in_guess) {
        cs.clear();
        for (int t = 0; t < m; ++t) {
            const int32_t j = ids[t];
            if (j == seed || j < 0) continue;
            Plane c;
            bool emptied = false, maybe = true;
            if (prepare(j, c, emptied, maybe)) {
                if (maybe && !in_guess && has_plane(j)) continue;
                cs.emplace_back(c.H, c);
            } else if (emptied) {
                status = 3;
                vs.clear();
                return;
            }
        }
        st->candidates += (int64_t)cs.size();
        std::sort(cs.begin(), cs.end(), [](const std::pair<float, Plane> &a, const std::pair<float, Plane> &b) {
            return a.first < b.first || (a.first == b.first && a.second.owner < b.second.owner);
        });
        for (auto &c : cs) {
            if (status != 0) return;
            if (in_guess && has_plane(c.second.owner)) continue;
            clip(c.second);
        }
    }

    // Depth-first: the nearer needed child next, the farther one waits and is
    // tested again when it comes off the stack (the cell has shrunk since).
    void traverse(std::vector<std::pair<int32_t, bool>> &stack) {
        stack.clear();
        stack.emplace_back(I->root, true);
        auto box = [&](int32_t r, const float *&lo, const float *&hi, float &wm) {
            lo = r < 0 ? I->leaves[-1 - r].lo : I->nodes[r].lo;
            hi = r < 0 ? I->leaves[-1 - r].hi : I->nodes[r].hi;
            wm = r < 0 ? I->leaves[-1 - r].wmax : I->nodes[r].wmax;
        };
        while (!stack.empty() && status == 0) {
            const int32_t r = stack.back().first;
            const bool test = stack.back().second;
            stack.pop_back();
            const float *lo, *hi;
            float wm, key;
            if (test) {
                box(r, lo, hi, wm);
                if (node_prunable(lo, hi, wm, key)) continue;
            }
            if (r < 0) {
                const Leaf &L = I->leaves[-1 - r];
                int32_t ids[32];
                for (int t = 0; t < L.count; ++t) ids[t] = L.first + t;
                offer(ids, L.count, false);
                continue;
            }
            st->nodes++;
            const BNode &nd = I->nodes[r];
            const int32_t c0 = nd.child[0], c1 = nd.child[1];
            float k0, k1;
            box(c0, lo, hi, wm);
            const bool n0 = !node_prunable(lo, hi, wm, k0);
            box(c1, lo, hi, wm);
            const bool n1 = !node_prunable(lo, hi, wm, k1);
            if (n0 && n1) {
                const bool near0 = k0 <= k1;
                stack.emplace_back(near0 ? c1 : c0, true);
                stack.emplace_back(near0 ? c0 : c1, false);
            } else if (n0) {
                stack.emplace_back(c0, false);
            } else if (n1) {
                stack.emplace_back(c1, false);
            }
        }
    }

    // sorted original ids of the owners of referenced planes (with duplicates)
    void row(std::vector<int32_t> &out) const {
        out.clear();
        if (status == 3) return;
        std::vector<char> used(pl.size(), 0);
        for (const Vert &v : vs)
            for (int e = 0; e < 3; ++e) used[v.p[e]] = 1;
        for (size_t p = 0; p < pl.size(); ++p) {
            if (!used[p] || pl[p].owner < 0) continue;
            const int o = pl[p].owner;
            if (I->has_dups) {
                for (int32_t m = I->dup_head[o]; m >= 0; m = I->dup_next[m]) out.push_back(I->perm[m]);
            } else {
                out.push_back(I->perm[o]);
            }
        }
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    }
};

struct Result {
    std::vector<int32_t> counts, status, offsets, adjacency;
    Stats st;
    double seconds = 0.0, t_index = 0.0, t_cells = 0.0;
};

}  // namespace

extern "C" {

int pgc_abi_version(void) { return PGC_ABI_VERSION; }

const char *pgc_last_error(void) { return g_err.c_str(); }

int pgc_build(int64_t n, const float *points, const float *weights, int32_t leaf_size, int32_t nthreads,
              void **handle) {
    if (!handle) return fail("handle is null");
    *handle = nullptr;
    if (n < 0 || n > (int64_t)INT32_MAX / 2) return fail("number of points out of range");
    if (n > 0 && !points) return fail("points is null");
    const int leaf = leaf_size < 1 ? 32 : (leaf_size > 32 ? 32 : leaf_size);
    Result *R = new Result();
    R->counts.assign(n, 0);
    R->status.assign(n, 0);
    R->offsets.assign(n + 1, 0);
    if (n == 0) {
        *handle = R;
        return 0;
    }
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const double t0 = omp_get_wtime();
    Index I;
    build_index(I, n, points, weights, leaf);
    const double t1 = omp_get_wtime();
    R->t_index = t1 - t0;
    std::vector<std::vector<int32_t>> rows(n);
    std::vector<Stats> tst;
#pragma omp parallel
    {
#pragma omp single
        tst.resize(omp_get_num_threads());
        Stats &st = tst[omp_get_thread_num()];
        Cell c;
        c.I = &I;
        c.st = &st;
        std::vector<std::pair<int32_t, bool>> stack;
        int32_t nn[kPreclip];
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; ++i) {
            c.init((int32_t)i);
            float kth = FLT_MAX;
            const int k = (int)std::min<int64_t>(kPreclip, n - 1);
            if (k > 0) {
                knn_of(I, (int32_t)i, k, nn, kth, c.kstack);
                c.offer(nn, k, true);  // the pre-clip, nearest first (voronoi_ultra.cu:70-85)
                // every point nearer than the 16th neighbour was offered (all
                // of them when there are no more than 16 others)
                c.guess_r2 = k == kPreclip ? kth * (1.0f - 1e-5f) : FLT_MAX;
            }
            if (c.status == 0) c.traverse(stack);
            const int32_t o = I.perm[i];
            c.row(rows[o]);
            R->counts[o] = (int32_t)rows[o].size();
            R->status[o] = c.status;
        }
    }
    R->t_cells = omp_get_wtime() - t1;
    for (const Stats &s : tst) {
        R->st.rechecks += s.rechecks, R->st.exact += s.exact, R->st.unresolved += s.unresolved;
        R->st.clips += s.clips, R->st.candidates += s.candidates, R->st.nodes += s.nodes;
    }
    // symmetry check; union symmetrisation (adjacency.cuh:78-159) only when needed
    int64_t asym = 0;
#pragma omp parallel for schedule(static, 4096) reduction(+ : asym)
    for (int64_t u = 0; u < n; ++u)
        for (int32_t v : rows[u])
            asym += (rows[v].empty() || !std::binary_search(rows[v].begin(), rows[v].end(), (int32_t)u)) ? 1 : 0;
    if (asym) {
        std::vector<std::vector<int32_t>> out(n);
        for (int64_t u = 0; u < n; ++u)
            for (int32_t v : rows[u]) {
                if (rows[v].empty()) continue;
                out[u].push_back(v);
                out[v].push_back((int32_t)u);
            }
        for (int64_t u = 0; u < n; ++u) {
            std::sort(out[u].begin(), out[u].end());
            out[u].erase(std::unique(out[u].begin(), out[u].end()), out[u].end());
        }
        rows.swap(out);
    }
    int64_t e = 0;
    for (int64_t u = 0; u < n; ++u) {
        e += (int64_t)rows[u].size();
        if (e > INT32_MAX) {
            delete R;
            return fail("more adjacency entries than int32 offsets hold");
        }
        R->offsets[u + 1] = (int32_t)e;
    }
    R->adjacency.resize(e);
#pragma omp parallel for schedule(static, 4096)
    for (int64_t u = 0; u < n; ++u) std::copy(rows[u].begin(), rows[u].end(), R->adjacency.begin() + R->offsets[u]);
    R->seconds = omp_get_wtime() - t0;
    *handle = R;
    return 0;
}

int64_t pgc_result_edges(const void *handle) { return handle ? (int64_t)((const Result *)handle)->adjacency.size() : -1; }

int pgc_result_copy(const void *handle, int32_t *adjacency, int32_t *offsets, int32_t *status) {
    if (!handle) return fail("handle is null");
    const Result *R = (const Result *)handle;
    if (!R->adjacency.empty() && !adjacency) return fail("adjacency is null");
    if (!offsets || (!R->status.empty() && !status)) return fail("offsets and status must be non-null");
    std::copy(R->adjacency.begin(), R->adjacency.end(), adjacency);
    std::copy(R->offsets.begin(), R->offsets.end(), offsets);
    std::copy(R->status.begin(), R->status.end(), status);
    return 0;
}

int pgc_result_stats(const void *handle, int64_t *out) {
    if (!handle || !out) return fail("invalid argument");
    const Result *R = (const Result *)handle;
    out[0] = R->st.rechecks, out[1] = R->st.exact, out[2] = R->st.unresolved;
    out[3] = R->st.clips, out[4] = R->st.candidates, out[5] = R->st.nodes;
    out[6] = (int64_t)(R->seconds * 1e9);
    out[7] = (int64_t)(R->t_index * 1e9);
    out[8] = (int64_t)(R->t_cells * 1e9);
    return 0;
}

void pgc_free(void *handle) { delete (Result *)handle; }

}  // extern "C"
