// Per-cell construction kernel for gfx950: one lane builds one convex cell.
//
// Replaces compute_bvh_voronoi_kernel (voronoi_ultra.cu:39-141) and
// compute_bvh_power_diagram_kernel (laguerre_ultra.cu:43-174) together with
// their clipper (voronoi_convex_cell.cuh, faster_convex_cell.cuh) and
// traversal (voronoi_bvh.cuh:149-288, pwr_bvh.cuh:217-356).
//
// Geometry.  In seed-relative doubled coordinates y = 2 (x - s) a point p
// contributes the half-space a . y <= H with a = p - s and
// H = |a|^2 - (w_p - w_s) (the power bisector of faster_convex_cell.cuh:287-294
// scaled by 2|a|); box planes are a = -+e_k, H = 2 (s_k - lo_k), 2 (hi_k - s_k).
// A vertex is a triple (i, j, k) of plane slots ordered so that
// D = a_i . (a_j x a_k) > 0; its position is y = N / D with
// N = H_i (a_j x a_k) + H_j (a_k x a_i) + H_k (a_i x a_j).
// Plane q removes the vertex iff P = a_q . N - H_q D > 0 (strictly, as
// voronoi_convex_cell.cuh:256-260).
//
// Exactness.  P is evaluated in FP32 together with its permanent pP (the
// same expression over absolute values).  Every quantity is a polynomial of
// depth < 16 in the rounded seed differences, so |P~ - P| <= gamma_16 pP with
// gamma_16 < 9.6e-7 (u = 2^-24); the test is decided when |P~| > 2e-6 pP~ and
// pP~ > 1e-30 (away from underflow).  Otherwise P is recomputed in FP64 from
// the FP32 inputs with the bound 1e-13 pP (gamma_24 < 2.7e-15, u = 2^-53);
// only a test that FP64 cannot decide either is counted as unresolved.
//
// Cell state sits in a caller-owned scratch laid out [slot][lane] per wave,
// so that lanes stepping through the same slot read 64 consecutive elements.
#include "pg_common.hpp"
#include "pg_kernels.hpp"

#include <float.h>

namespace pg {

constexpr float kG32 = 2e-6f;
constexpr float kTiny32 = 1e-30f;
constexpr double kG64 = 1e-13;
constexpr double kTiny64 = 1e-280;
constexpr float kSlack = 1e-6f;  // relative margin of the conservative prune tests

struct CellParams {
    int64_t n;
    const IndexHeader *hdr;
    const float4 *pts;      // Morton-sorted points (x, y, z, w)
    const uint32_t *perm;   // sorted -> original
    const uint32_t *inv;    // original -> sorted
    const Node *nodes;
    const uint32_t *dup_head;
    const uint32_t *dup_next;
    int64_t begin, end;
    const int32_t *guesses;
    const int32_t *guess_offsets;
    int guesses_sorted;
    int seed_window;              // Morton neighbours offered first (0: none)
    int32_t *rows;                // tier-0 slab (row_stride per cell) then tier-1 rows
    int row_stride;
    int64_t *row_offsets;         // [n] sorted cell -> row start in `rows`
    int32_t *counts;
    int32_t *status;
    unsigned long long *stats;
    uint32_t *counters;           // [0] tier-0 work, [1] tier-1 queue length, [2] tier-1 work
    char *scratch;                // tier-0 per-wave scratch
    uint64_t per_wave0;
    char *scratch1;               // tier-1 per-wave scratch
    uint64_t per_wave1;
    uint32_t *big_list;           // tier-1 queue of sorted cell ids
    uint32_t big_cap;
    uint32_t *cycles;             // optional [n]: per-cell cost counter (original order)
};

// A vertex is a triple of plane slots packed kBits bits apiece.
template <int kBits>
__device__ __forceinline__ uint32_t tri_pack(uint32_t a, uint32_t b, uint32_t c) {
    return a | (b << kBits) | (c << (2 * kBits));
}
template <int kBits>
__device__ __forceinline__ uint32_t tri_get(uint32_t t, int e) {
    return (t >> (kBits * e)) & ((1u << kBits) - 1u);
}
template <int kBits>
__device__ __forceinline__ bool tri_has_edge(uint32_t t, uint32_t x, uint32_t y) {
    const uint32_t a = tri_get<kBits>(t, 0), b = tri_get<kBits>(t, 1), c = tri_get<kBits>(t, 2);
    return (a == x && b == y) || (b == x && c == y) || (c == x && a == y);
}
constexpr uint32_t kMark = 0xFFFFFFFEu;  // "plane referenced" mark in the per-plane word
__device__ __forceinline__ uint64_t bloom_bit(uint32_t id) { return 1ull << ((id * 2654435761u) >> 26); }

template <int kPCap, int kVCap, int kBits>
struct Cell {
    // scratch views (this lane)
    float4 *plane_ah;
    float2 *plane_aux;
    uint32_t *vtri;
    float4 *vnd;
    float2 *vperm;
    uint32_t *nxt;
    uint32_t *stk;
    int lane;

    const float4 *pts;
    const uint32_t *perm;
    const uint32_t *inv;
    const Node *nodes;
    uint32_t root_ref;
    float box_lo[3], box_hi[3];
    uint32_t seed;
    float sx, sy, sz, sw;
    int nv, np;
    float L[3], U[3];
    float R8[8];  // squared distance to the AABB corner of each octant (bit k: +axis k)
    bool bounded;
    int status;
    uint32_t gen;
    uint64_t bloom;
    bool in_guess;
    // counters
    uint32_t c_recheck, c_unresolved, c_inconsistent, c_overflow, c_clips, c_cands, c_nodes, c_leaves, c_skipped;

    __device__ float4 &pah(int i) { return plane_ah[i * kWave + lane]; }
    __device__ float2 &paux(int i) { return plane_aux[i * kWave + lane]; }
    __device__ uint32_t &vt(int i) { return vtri[i * kWave + lane]; }
    __device__ float4 &vn(int i) { return vnd[i * kWave + lane]; }
    __device__ float2 &vp(int i) { return vperm[i * kWave + lane]; }
    __device__ uint32_t &nx(int i) { return nxt[i * kWave + lane]; }
    __device__ uint32_t &st(int i) { return stk[i * kWave + lane]; }

    // ---------------------------------------------------------- FP64 path
    __device__ void plane64(int owner, double a[3], double &H, double pa[3], double &pH) const {
        const double s[3] = {(double)sx, (double)sy, (double)sz};
        if (owner >= 0) {
            const float4 p = pts[owner];
            a[0] = (double)p.x - s[0];
            a[1] = (double)p.y - s[1];
            a[2] = (double)p.z - s[2];
            const double dw = (double)p.w - (double)sw;
            H = a[0] * a[0] + a[1] * a[1] + a[2] * a[2] - dw;
            pH = a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + fabs(dw);
            pa[0] = fabs(a[0]);
            pa[1] = fabs(a[1]);
            pa[2] = fabs(a[2]);
        } else {
            const int code = -1 - owner, axis = code >> 1, side = code & 1;
            a[0] = a[1] = a[2] = 0.0;
            pa[0] = pa[1] = pa[2] = 0.0;
            pa[axis] = 1.0;
            if (side == 0) {
                a[axis] = -1.0;
                H = 2.0 * (s[axis] - (double)box_lo[axis]);
            } else {
                a[axis] = 1.0;
                H = 2.0 * ((double)box_hi[axis] - s[axis]);
            }
            pH = fabs(H);
        }
    }

    __device__ __noinline__ void vertex64(uint32_t t, double N[3], double &D, double pN[3], double &pD) {
        double a[3][3], H[3], pa[3][3], pH[3];
        for (int e = 0; e < 3; ++e) {
            const int owner = __float_as_int(paux(tri_get<kBits>(t, e)).y);
            plane64(owner, a[e], H[e], pa[e], pH[e]);
        }
        double c[3][3], pc[3][3];
        for (int e = 0; e < 3; ++e) {
            const double *u = a[(e + 1) % 3], *v = a[(e + 2) % 3];
            const double *pu = pa[(e + 1) % 3], *pv = pa[(e + 2) % 3];
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

    // Removal decision for vertex v against point plane q in FP64.
    __device__ __noinline__ bool recheck64(int v, int q_owner) {
        c_recheck++;
        double N[3], D, pN[3], pD;
        vertex64(vt(v), N, D, pN, pD);
        double a[3], H, pa[3], pH;
        plane64(q_owner, a, H, pa, pH);
        const double Pv = a[0] * N[0] + a[1] * N[1] + a[2] * N[2] - H * D;
        const double pP = pa[0] * pN[0] + pa[1] * pN[1] + pa[2] * pN[2] + pH * pD;
        const bool dec = fabs(Pv) > kG64 * pP && pP > kTiny64;
        if (fabs(D) > kG64 * pD && pD > kTiny64 && D < 0.0) c_inconsistent++;
        if (!dec) {
            // A point identical (position and weight) to the owner of one of
            // the vertex's planes gives the same plane: P = 0 exactly, so the
            // vertex is not strictly beyond it.
            const float4 q = pts[q_owner];
            const uint32_t t = vt(v);
            for (int e = 0; e < 3; ++e) {
                const int o = __float_as_int(paux(tri_get<kBits>(t, e)).y);
                if (o < 0) continue;
                const float4 r = pts[o];
                if (r.x == q.x && r.y == q.y && r.z == q.z && r.w == q.w) return false;
            }
            c_unresolved++;
        }
        return Pv > 0.0;
    }

    // ---------------------------------------------------------- vertices
    __device__ void make_vertex(int slot, uint32_t i, uint32_t j, uint32_t k) {
        const float4 A = pah(i), B = pah(j), C = pah(k);
        const float pHA = paux(i).x, pHB = paux(j).x, pHC = paux(k).x;
        // cross products of the normals
        const float c1x = B.y * C.z - B.z * C.y, c1y = B.z * C.x - B.x * C.z, c1z = B.x * C.y - B.y * C.x;
        const float c2x = C.y * A.z - C.z * A.y, c2y = C.z * A.x - C.x * A.z, c2z = C.x * A.y - C.y * A.x;
        const float c3x = A.y * B.z - A.z * B.y, c3y = A.z * B.x - A.x * B.z, c3z = A.x * B.y - A.y * B.x;
        const float D = A.x * c1x + A.y * c1y + A.z * c1z;
        const float Nx = A.w * c1x + B.w * c2x + C.w * c3x;
        const float Ny = A.w * c1y + B.w * c2y + C.w * c3y;
        const float Nz = A.w * c1z + B.w * c2z + C.w * c3z;
        // permanents
        const float ax = fabsf(A.x), ay = fabsf(A.y), az = fabsf(A.z);
        const float bx = fabsf(B.x), by = fabsf(B.y), bz = fabsf(B.z);
        const float cx = fabsf(C.x), cy = fabsf(C.y), cz = fabsf(C.z);
        const float q1x = by * cz + bz * cy, q1y = bz * cx + bx * cz, q1z = bx * cy + by * cx;
        const float q2x = cy * az + cz * ay, q2y = cz * ax + cx * az, q2z = cx * ay + cy * ax;
        const float q3x = ay * bz + az * by, q3y = az * bx + ax * bz, q3z = ax * by + ay * bx;
        const float pD = ax * q1x + ay * q1y + az * q1z;
        const float pNx = pHA * q1x + pHB * q2x + pHC * q3x;
        const float pNy = pHA * q1y + pHB * q2y + pHC * q3y;
        const float pNz = pHA * q1z + pHB * q2z + pHC * q3z;
        vt(slot) = tri_pack<kBits>(i, j, k);
        vn(slot) = make_float4(Nx, Ny, Nz, D);
        vp(slot) = make_float2(fmaxf(pNx, fmaxf(pNy, pNz)), pD);
        if (D < -kG32 * pD) c_inconsistent++;
    }

    // Is vertex v strictly beyond plane (a, H)?  pa1 = |a|_1, pH = permanent of H.
    __device__ __forceinline__ bool beyond(int v, float ax, float ay, float az, float H, float pa1, float pH,
                                           int q_owner) {
        const float4 nd = vn(v);
        const float2 pp = vp(v);
        const float Pv = ax * nd.x + ay * nd.y + az * nd.z - H * nd.w;
        const float pP = pa1 * pp.x + pH * pp.y;
        if (fabsf(Pv) > kG32 * pP && pP > kTiny32) return Pv > 0.0f;
        return recheck64(v, q_owner);
    }

    // ---------------------------------------------------------- bounds
    // Position and error of a vertex whose FP32 orientation is too close to
    // zero to bound its position; false when FP64 cannot bound it either.
    __device__ __noinline__ bool bounds64(int v, float &x, float &y, float &z, float &ex, float &ey, float &ez) {
        double N[3], D, pN[3], pD;
        vertex64(vt(v), N, D, pN, pD);
        const double e64 = kG64 * pD;
        if (!(D > 2.0 * e64 && pD > kTiny64)) return false;
        double X[3], E[3];
        for (int k = 0; k < 3; ++k) {
            X[k] = N[k] / D;
            E[k] = (kG64 * pN[k] + fabs(X[k]) * e64) / (D - e64) * (1.0 + 1e-12) + fabs(X[k]) * 1e-12;
        }
        x = (float)X[0];
        y = (float)X[1];
        z = (float)X[2];
        // float rounding of the centre and the error, covered by the slack
        ex = (float)E[0] + fabsf(x) * kSlack + kTiny32;
        ey = (float)E[1] + fabsf(y) * kSlack + kTiny32;
        ez = (float)E[2] + fabsf(z) * kSlack + kTiny32;
        return true;
    }

    // Rigorous AABB of the cell (doubled coordinates).  With
    // |N - N~| <= g pN and |D - D~| <= g pD:
    //     |y - N~/D~| <= (g pN + |y~| g pD) / (D~ - g pD).
    __device__ void update_bounds() {
        float lo0 = FLT_MAX, lo1 = FLT_MAX, lo2 = FLT_MAX, hi0 = -FLT_MAX, hi1 = -FLT_MAX, hi2 = -FLT_MAX;
        bool ok = true;
        for (int v = 0; v < nv; ++v) {
            const float4 nd = vn(v);
            const float2 pp = vp(v);
            const float eD = kG32 * pp.y;
            float x, y, z, ex, ey, ez;
            if (nd.w > 2.0f * eD && pp.y > kTiny32) {
                const float inv = 1.0f / nd.w;
                const float den = 1.0f / (nd.w - eD);
                const float eN = kG32 * pp.x;
                x = nd.x * inv;
                y = nd.y * inv;
                z = nd.z * inv;
                ex = (eN + fabsf(x) * eD) * den * (1.0f + kSlack) + fabsf(x) * kSlack;
                ey = (eN + fabsf(y) * eD) * den * (1.0f + kSlack) + fabsf(y) * kSlack;
                ez = (eN + fabsf(z) * eD) * den * (1.0f + kSlack) + fabsf(z) * kSlack;
            } else if (!bounds64(v, x, y, z, ex, ey, ez)) {
                ok = false;
                break;
            }
            lo0 = fminf(lo0, x - ex);
            hi0 = fmaxf(hi0, x + ex);
            lo1 = fminf(lo1, y - ey);
            hi1 = fmaxf(hi1, y + ey);
            lo2 = fminf(lo2, z - ez);
            hi2 = fmaxf(hi2, z + ez);
        }
        bounded = ok;
        L[0] = lo0, L[1] = lo1, L[2] = lo2;
        U[0] = hi0, U[1] = hi1, U[2] = hi2;
        // corner radii (voronoi_bvh.cuh:40-66), rounded up by the slack
        const float l0 = lo0 * lo0, l1 = lo1 * lo1, l2 = lo2 * lo2;
        const float u0 = hi0 * hi0, u1 = hi1 * hi1, u2 = hi2 * hi2;
#pragma unroll
        for (int o = 0; o < 8; ++o)
            R8[o] = ((o & 1 ? u0 : l0) + (o & 2 ? u1 : l1) + (o & 4 ? u2 : l2)) * (1.0f + kSlack);
    }

    // Can no point plane with normal a and offset H (permanent pH) cut the
    // cell?  max over the AABB of a . y is sum_k max(a_k L_k, a_k U_k).
    __device__ __forceinline__ bool too_far(float ax, float ay, float az, float H, float pH) const {
        if (!bounded) return false;
        const float S = fmaxf(ax * L[0], ax * U[0]) + fmaxf(ay * L[1], ay * U[1]) + fmaxf(az * L[2], az * U[2]);
        const float M = fabsf(ax) * fmaxf(fabsf(L[0]), fabsf(U[0])) + fabsf(ay) * fmaxf(fabsf(L[1]), fabsf(U[1])) +
                        fabsf(az) * fmaxf(fabsf(L[2]), fabsf(U[2]));
        return (H - S) > kSlack * (pH + M) + kTiny32;
    }

    // Node prune: H >= d_min^2 - (w_max - w_s) for every point of the box,
    // and the support of the AABB over the box is separable per axis, so its
    // maximum is the sum of per-axis maxima over the box's two faces.
    // Node prune.  Every point p of the box has a = p - s with |a| >= d and
    // H = |a|^2 - (w_p - w_s) >= |a|^2 - dw (dw = w_max - w_s).  Two bounds
    // on the cell side, either of which suffices:
    //  * octant radius (voronoi_bvh.cuh:71-107, pwr_bvh.cuh:150-185): a . y
    //    <= |a| rho with rho the largest AABB corner radius over the octants
    //    the box overlaps, so no cut when |a|^2 - rho |a| - dw >= 0 for all
    //    |a| >= d, i.e. 2d >= rho and (d^2 - dw)^2 >= rho^2 d^2, d^2 > dw;
    //  * support: a . y <= sum_k max over the box's faces of a_k L_k, a_k U_k
    //    (separable per axis), against the lower bound d^2 - dw of H.
    __device__ __forceinline__ bool node_prunable(float4 lo, float4 hi, float &key) const {
        const float alx = lo.x - sx, aly = lo.y - sy, alz = lo.z - sz;
        const float ahx = hi.x - sx, ahy = hi.y - sy, ahz = hi.z - sz;
        const float dx = fmaxf(fmaxf(alx, -ahx), 0.0f);
        const float dy = fmaxf(fmaxf(aly, -ahy), 0.0f);
        const float dz = fmaxf(fmaxf(alz, -ahz), 0.0f);
        const float d2 = dx * dx + dy * dy + dz * dz;
        const float dw = lo.w - sw;
        const float hlow = d2 - dw;
        key = hlow;
        if (!bounded) return false;
        // octant radius
        const bool nx_ = alx <= 0.0f, px_ = ahx >= 0.0f;
        const bool ny_ = aly <= 0.0f, py_ = ahy >= 0.0f;
        const bool nz_ = alz <= 0.0f, pz_ = ahz >= 0.0f;
        float rho2 = 0.0f;
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            const bool in = ((o & 1) ? px_ : nx_) && ((o & 2) ? py_ : ny_) && ((o & 4) ? pz_ : nz_);
            rho2 = in ? fmaxf(rho2, R8[o]) : rho2;
        }
        const float d2lo = d2 * (1.0f - kSlack);
        const float h = d2lo - dw - kSlack * (fabsf(lo.w) + fabsf(sw)) - kTiny32;
        if (4.0f * d2lo > rho2 && h > 0.0f && h * h > rho2 * d2 * (1.0f + kSlack)) return true;
        // support
        const float S = fmaxf(fmaxf(alx * L[0], alx * U[0]), fmaxf(ahx * L[0], ahx * U[0])) +
                        fmaxf(fmaxf(aly * L[1], aly * U[1]), fmaxf(ahy * L[1], ahy * U[1])) +
                        fmaxf(fmaxf(alz * L[2], alz * U[2]), fmaxf(ahz * L[2], ahz * U[2]));
        const float M = fmaxf(fabsf(alx), fabsf(ahx)) * fmaxf(fabsf(L[0]), fabsf(U[0])) +
                        fmaxf(fabsf(aly), fabsf(ahy)) * fmaxf(fabsf(L[1]), fabsf(U[1])) +
                        fmaxf(fabsf(alz), fabsf(ahz)) * fmaxf(fabsf(L[2]), fabsf(U[2]));
        return (hlow - S) > kSlack * (d2 + fabsf(lo.w) + fabsf(sw) + M) + kTiny32;
    }

    // ---------------------------------------------------------- planes
    // Marks the planes referenced by some vertex (kMark in their word).
    __device__ void mark_used() {
        for (int i = 0; i < np; ++i) nx(i) = 0u;
        for (int v = 0; v < nv; ++v) {
            const uint32_t t = vt(v);
            for (int e = 0; e < 3; ++e) {
                const uint32_t p = tri_get<kBits>(t, e);
                if (p < (uint32_t)np) nx(p) = kMark;
            }
        }
    }

    // Drops unreferenced planes and renumbers the vertex triples.  The
    // per-plane words then hold small remap values, whose generation field
    // (0) never matches a live boundary tag.
    __device__ void garbage_collect() {
        mark_used();
        int m = 0;
        for (int i = 0; i < np; ++i) {
            if (nx(i) != kMark) continue;
            if (m != i) {
                pah(m) = pah(i);
                paux(m) = paux(i);
            }
            nx(i) = (uint32_t)m;
            m++;
        }
        for (int v = 0; v < nv; ++v) {
            const uint32_t t = vt(v);
            const uint32_t a = tri_get<kBits>(t, 0), b = tri_get<kBits>(t, 1), c = tri_get<kBits>(t, 2);
            if (a >= (uint32_t)np || b >= (uint32_t)np || c >= (uint32_t)np) continue;
            vt(v) = tri_pack<kBits>(nx(a), nx(b), nx(c));
        }
        np = m;
    }

    __device__ bool has_plane(uint32_t owner) {
        if (!(bloom & bloom_bit(owner))) return false;
        for (int i = 0; i < np; ++i)
            if (__float_as_int(paux(i).y) == (int)owner) return true;
        return false;
    }

    // ---------------------------------------------------------- clipping
    // Returns true when the cell changed.
    __device__ bool clip(uint32_t owner, float ax, float ay, float az, float H, float pH) {
        const float pa1 = fabsf(ax) + fabsf(ay) + fabsf(az);
        // Read-only sweep first (independent loads, no calls): most offered
        // planes remove nothing and stop here.
        bool hit = false;
#pragma unroll 4
        for (int v = 0; v < nv; ++v) {
            const float4 nd = vn(v);
            const float2 pp = vp(v);
            const float Pv = ax * nd.x + ay * nd.y + az * nd.z - H * nd.w;
            const float pP = pa1 * pp.x + pH * pp.y;
            const bool dec = fabsf(Pv) > kG32 * pP && pP > kTiny32;
            hit |= !dec || Pv > 0.0f;
        }
        if (!hit) return false;
        // partition: vertices strictly beyond the plane move to the tail
        int R = 0, i = 0;
        while (i < nv - R) {
            if (beyond(i, ax, ay, az, H, pa1, pH, (int)owner)) {
                const int j = nv - 1 - R;
                if (j != i) {
                    const uint32_t t = vt(i);
                    vt(i) = vt(j);
                    vt(j) = t;
                    const float4 a = vn(i);
                    vn(i) = vn(j);
                    vn(j) = a;
                    const float2 b = vp(i);
                    vp(i) = vp(j);
                    vp(j) = b;
                }
                R++;
            } else {
                i++;
            }
        }
        if (R == 0) return false;
        if (R == nv) {  // voronoi_convex_cell.cuh:430-435
            status = kEmptyCell;
            nv = 0;
            return true;
        }
        if (np >= kPCap) garbage_collect();
        if (np >= kPCap) {
            status = kPlaneOverflow;
            c_overflow++;
            return false;
        }
        const int q = np++;
        pah(q) = make_float4(ax, ay, az, H);
        paux(q) = make_float2(pH, __int_as_float((int)owner));
        if (in_guess) bloom |= bloom_bit(owner);
        // boundary of the removed triangles: directed edges whose twin is
        // not removed; each plane starts at most one boundary edge
        // generations run 1 .. 2^(32-kBits) - 2: the all-ones generation is
        // reserved for the cleared and marked words
        gen = (gen + 1) % ((1u << (32 - kBits)) - 1u);
        if (gen == 0) gen = 1;
        const uint32_t tag = gen << kBits;
        const uint32_t tag_mask = ~((1u << kBits) - 1u);
        const int base = nv - R;
        int first = -1, nb = 0;
        for (int t = base; t < nv; ++t) {
            const uint32_t tt = vt(t);
            for (int e = 0; e < 3; ++e) {
                const uint32_t x = tri_get<kBits>(tt, e), y = tri_get<kBits>(tt, (e + 1) % 3);
                if (x >= (uint32_t)np || y >= (uint32_t)np) {
                    status = kInconsistentBoundary;
                    c_inconsistent++;
                    return false;
                }
                bool twin = false;
                for (int u = base; u < nv && !twin; ++u) twin = tri_has_edge<kBits>(vt(u), y, x);
                if (!twin) {
                    nx(x) = tag | y;
                    first = (int)x;
                    nb++;
                }
            }
        }
        if (first < 0) {
            status = kInconsistentBoundary;
            c_inconsistent++;
            return false;
        }
        // fan of new vertices (q, x, next[x]) around the boundary cycle
        int cnt = 0;
        uint32_t x = (uint32_t)first;
        do {
            const uint32_t e = nx(x);
            if ((e & tag_mask) != tag || cnt >= nb) {
                status = kInconsistentBoundary;
                c_inconsistent++;
                nv = base + cnt;
                return true;
            }
            const uint32_t y = e & ((1u << kBits) - 1u);
            if (base + cnt >= kVCap) {
                status = kVertexOverflow;
                c_overflow++;
                nv = base + cnt;
                return true;
            }
            make_vertex(base + cnt, (uint32_t)q, x, y);
            cnt++;
            x = y;
        } while (x != (uint32_t)first);
        if (cnt != nb) {
            status = kInconsistentBoundary;
            c_inconsistent++;
        }
        nv = base + cnt;
        return true;
    }

    // Offer point j (sorted index) to the cell.
    __device__ void candidate(uint32_t j) {
        const float4 p = pts[j];
        const float ax = p.x - sx, ay = p.y - sy, az = p.z - sz;
        const float dw = p.w - sw;
        if (ax == 0.0f && ay == 0.0f && az == 0.0f) {
            // coincident point: the half-space is empty iff w_p > w_s
            if (dw > 0.0f) {
                status = kEmptyCell;
                nv = 0;
            }
            return;
        }
        const float a2 = ax * ax + ay * ay + az * az;
        const float H = a2 - dw;
        const float pH = a2 + fabsf(dw);
        if (too_far(ax, ay, az, H, pH)) {
            c_skipped++;
            return;
        }
        if (bloom && has_plane(j)) return;
        c_cands++;
        if (clip(j, ax, ay, az, H, pH)) {
            c_clips++;
            if (status == kSuccess) update_bounds();
        }
    }

    __device__ void init_box() {
        const float s[3] = {sx, sy, sz};
        for (int axis = 0; axis < 3; ++axis) {
            const float hl = 2.0f * (s[axis] - box_lo[axis]);
            const float hh = 2.0f * (box_hi[axis] - s[axis]);
            float4 al = make_float4(0.f, 0.f, 0.f, hl), ah = make_float4(0.f, 0.f, 0.f, hh);
            if (axis == 0) al.x = -1.f, ah.x = 1.f;
            if (axis == 1) al.y = -1.f, ah.y = 1.f;
            if (axis == 2) al.z = -1.f, ah.z = 1.f;
            pah(2 * axis) = al;
            pah(2 * axis + 1) = ah;
            paux(2 * axis) = make_float2(fabsf(hl), __int_as_float(-1 - 2 * axis));
            paux(2 * axis + 1) = make_float2(fabsf(hh), __int_as_float(-2 - 2 * axis));
        }
        np = 6;
        int v = 0;
        for (int cx = 0; cx < 2; ++cx)
            for (int cy = 0; cy < 2; ++cy)
                for (int cz = 0; cz < 2; ++cz) {
                    const uint32_t X = cx, Y = 2 + cy, Z = 4 + cz;
                    // D(X, Y, Z) = sign_x sign_y sign_z; swap two to make it positive
                    const int sgn = (cx ? 1 : -1) * (cy ? 1 : -1) * (cz ? 1 : -1);
                    if (sgn > 0) make_vertex(v, X, Y, Z);
                    else make_vertex(v, X, Z, Y);
                    v++;
                }
        nv = 8;
        bloom = 0;
        update_bounds();
    }

    // ---------------------------------------------------------- traversal
    __device__ void traverse() {
        uint32_t ref = root_ref;
        int sp = 0;
        while (true) {
            bool have = true;
            if (is_leaf(ref)) {
                c_leaves++;
                const uint32_t f = leaf_first(ref), c = leaf_count(ref);
                for (uint32_t j = f; j < f + c && status == kSuccess; ++j)
                    if (j != seed) candidate(j);
                if (status != kSuccess) return;
                have = false;
            } else {
                c_nodes++;
                const Node nd = nodes[ref];
                float k0, k1;
                const bool p0 = node_prunable(nd.lo0, nd.hi0, k0);
                const bool p1 = node_prunable(nd.lo1, nd.hi1, k1);
                const uint32_t r0 = __float_as_uint(nd.hi0.w), r1 = __float_as_uint(nd.hi1.w);
                if (p0 && p1) {
                    have = false;
                } else if (!p0 && !p1) {
                    const bool near0 = k0 <= k1;
                    if (sp < kStackCap) st(sp++) = (ref << 1) | (near0 ? 1u : 0u);
                    else {  // cannot happen for N < 2^26 (depth < 56); never drop silently
                        status = kSecurityRadiusNotReached;
                        return;
                    }
                    ref = near0 ? r0 : r1;
                } else {
                    ref = p0 ? r1 : r0;
                }
            }
            if (!have) {
                bool found = false;
                while (sp > 0) {
                    const uint32_t e = st(--sp);
                    const uint32_t parent = e >> 1, c = e & 1u;
                    const Node nd = nodes[parent];
                    float k;
                    const float4 lo = c ? nd.lo1 : nd.lo0;
                    const float4 hi = c ? nd.hi1 : nd.hi0;
                    if (!node_prunable(lo, hi, k)) {
                        ref = __float_as_uint(hi.w);
                        found = true;
                        break;
                    }
                }
                if (!found) return;
            }
        }
    }

    // Builds the cell of sorted point `cell`.  Returns false when it
    // overflowed this tier's capacity (nothing is emitted then unless
    // `emit_partial`).
    __device__ bool run(uint32_t cell, const CellParams &prm, int32_t *row, int row_cap, bool emit_partial) {
        seed = cell;
        const float4 s = pts[cell];
        sx = s.x, sy = s.y, sz = s.z, sw = s.w;
        status = kSuccess;
        init_box();
        const uint32_t orig = perm[cell];
        in_guess = true;
        if (prm.guesses) {
            const int32_t g0 = prm.guess_offsets[orig], g1 = prm.guess_offsets[orig + 1];
            for (int32_t g = g0; g < g1 && status == kSuccess; ++g) {
                const int32_t gid = prm.guesses[g];
                if (gid < 0) continue;  // voronoi_ultra.cu:76-78
                if ((int64_t)gid >= prm.n) continue;
                const uint32_t j = prm.guesses_sorted ? (uint32_t)gid : inv[gid];
                if (j == cell) continue;
                candidate(j);
            }
        }
        in_guess = false;
        if (status == kSuccess) traverse();
        const bool overflow = status == kPlaneOverflow || status == kVertexOverflow;
        if (overflow && !emit_partial) return false;
        // emit the owners of planes referenced by surviving vertices,
        // voronoi_ultra.cu:125-138, as a sorted row of original ids
        int cnt = 0;
        bool row_full = false;
        if (status != kEmptyCell) {
            mark_used();
            const bool dups = prm.hdr->has_duplicates != 0u;
            for (int i = 0; i < np; ++i) {
                if (nx(i) != kMark) continue;
                const int owner = __float_as_int(paux(i).y);
                if (owner < 0) continue;
                // every exact duplicate of an owner carries the same face
                uint32_t m = dups ? prm.dup_head[owner] : (uint32_t)owner;
                while (m != 0xFFFFFFFFu) {
                    const int32_t id = (int32_t)perm[m];
                    int k = cnt;
                    while (k > 0 && row[k - 1] > id) k--;
                    if (!(k > 0 && row[k - 1] == id)) {
                        if (cnt < row_cap) {
                            for (int t = cnt; t > k; --t) row[t] = row[t - 1];
                            row[k] = id;
                            cnt++;
                        } else {
                            row_full = true;
                        }
                    }
                    m = dups ? prm.dup_next[m] : 0xFFFFFFFFu;
                }
            }
        }
        if (row_full && !overflow) {
            if (!emit_partial) return false;
            status = kPlaneOverflow;
            c_overflow++;
        }
        prm.counts[orig] = cnt;
        prm.status[orig] = status;
        return true;
    }
};

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <int kPCap, int kVCap, int kBits>
__device__ void setup_cell(Cell<kPCap, kVCap, kBits> &c, const CellParams &prm, char *base, int lane) {
    const CellScratchLayout SL = make_cell_scratch_layout(kPCap, kVCap);
    c.plane_ah = (float4 *)(base + SL.plane_ah);
    c.plane_aux = (float2 *)(base + SL.plane_aux);
    c.vtri = (uint32_t *)(base + SL.vert_tri);
    c.vnd = (float4 *)(base + SL.vert_nd);
    c.vperm = (float2 *)(base + SL.vert_perm);
    c.nxt = (uint32_t *)(base + SL.next);
    c.stk = (uint32_t *)(base + SL.stack);
    c.lane = lane;
    c.pts = prm.pts;
    c.perm = prm.perm;
    c.inv = prm.inv;
    c.nodes = prm.nodes;
    c.root_ref = prm.hdr->root_ref;
    for (int k = 0; k < 3; ++k) c.box_lo[k] = prm.hdr->box_lo[k], c.box_hi[k] = prm.hdr->box_hi[k];
    c.c_recheck = c.c_unresolved = c.c_inconsistent = c.c_overflow = c.c_clips = c.c_cands = 0;
    c.c_nodes = c.c_leaves = c.c_skipped = 0;
    // boundary tags grow monotonically per lane within a launch; stale words
    // from an earlier launch are cleared once
    for (int i = 0; i < kPCap; ++i) c.nx(i) = 0xFFFFFFFFu;
    c.gen = 0;
}

template <int kPCap, int kVCap, int kBits>
__device__ void flush_stats(const Cell<kPCap, kVCap, kBits> &c, const CellParams &prm, uint32_t cells, int lane) {
    const uint32_t vals[12] = {c.c_recheck, c.c_unresolved, c.c_inconsistent, c.c_overflow, c.c_clips, c.c_cands,
                               c.c_nodes, cells, 0, 0, c.c_leaves, c.c_skipped};
    for (int k = 0; k < 12; ++k) {
        const uint32_t s = wave_sum(vals[k]);
        if (lane == 0 && s) atomicAdd(&prm.stats[k], (unsigned long long)s);
    }
}

}  // namespace pg

#include "pg_group_cell.hpp"

namespace pg {

// Capacity tier 1: the queued cells, rows written after the tier-0 slab.
__global__ void __launch_bounds__(256) cell_big_kernel(const CellParams prm) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    Cell<kTier1Planes, kTier1Verts, kTier1Bits> c;
    setup_cell(c, prm, prm.scratch1 + wave * prm.per_wave1, lane);
    const uint32_t queued = min(prm.counters[1], prm.big_cap);
    while (true) {
        uint32_t c0 = 0;
        if (lane == 0) c0 = atomicAdd(&prm.counters[2], (uint32_t)kWave);
        c0 = __shfl(c0, 0);
        if (c0 >= queued) break;
        const uint32_t k = c0 + lane;
        if (k < queued) {
            const uint32_t cell = prm.big_list[k];
            const int64_t off = (int64_t)(prm.end - prm.begin) * prm.row_stride + (int64_t)k * kTier1Row;
            prm.row_offsets[cell] = off;
            const uint64_t t0 = clock64();
            if (!c.run(cell, prm, prm.rows + off, kTier1Row, true)) c.c_overflow++;
            if (prm.cycles) prm.cycles[prm.perm[cell]] = 0x80000000u | (uint32_t)min(clock64() - t0, (uint64_t)0x7fffffff);
        }
    }
    flush_stats(c, prm, 0, lane);
}

__global__ void cell_prep_kernel(uint32_t *counters) {
    if (threadIdx.x < 4) counters[threadIdx.x] = 0;
}

uint64_t cell_scratch_bytes(int waves, int waves1, int64_t big_cap) {
    (void)waves;  // tier 0 keeps its cells in LDS
    const CellScratchLayout S1 = make_cell_scratch_layout(kTier1Planes, kTier1Verts);
    return 256 + S1.per_wave * (uint64_t)waves1 + 4 * (uint64_t)big_cap;
}

hipError_t launch_cells(const CellArgs &a, hipStream_t stream) {
    const IndexLayout L = make_index_layout(a.n);
    const char *ws = (const char *)a.index_ws;
    CellParams prm;
    const IndexHeader *hdr = (const IndexHeader *)(ws + L.header);
    prm.n = a.n;
    prm.hdr = hdr;
    prm.pts = (const float4 *)(ws + L.points);
    prm.perm = (const uint32_t *)(ws + L.perm);
    prm.inv = (const uint32_t *)(ws + L.inv);
    prm.nodes = (const Node *)(ws + L.nodes);
    prm.dup_head = (const uint32_t *)(ws + L.dup_head);
    prm.dup_next = (const uint32_t *)(ws + L.dup_next);
    prm.begin = a.begin;
    prm.end = a.end;
    prm.guesses = a.guesses;
    prm.guess_offsets = a.guess_offsets;
    prm.guesses_sorted = a.guesses_sorted;
    prm.seed_window = a.seed_window;
    prm.rows = a.rows;
    prm.row_stride = a.row_stride;
    prm.row_offsets = a.row_offsets;
    prm.counts = a.counts;
    prm.status = a.status;
    prm.stats = a.stats;
    prm.cycles = a.cycles;
    const CellScratchLayout S1 = make_cell_scratch_layout(kTier1Planes, kTier1Verts);
    char *sc = (char *)a.scratch;
    prm.counters = (uint32_t *)sc;
    prm.scratch = nullptr;
    prm.per_wave0 = 0;
    prm.scratch1 = sc + 256;
    prm.per_wave1 = S1.per_wave;
    prm.big_list = (uint32_t *)(prm.scratch1 + S1.per_wave * (uint64_t)a.waves1);
    prm.big_cap = (uint32_t)a.big_cap;
    const int64_t count = a.end - a.begin;
    if (count <= 0) return hipSuccess;
    if (cell_scratch_bytes(a.waves, a.waves1, a.big_cap) > a.scratch_bytes) return hipErrorInvalidValue;
    const int64_t blocks_needed = (count + kGCellsPerBlock - 1) / kGCellsPerBlock;
    const int64_t max_blocks = (int64_t)a.waves * kWave / kGThreads;
    const int64_t blocks = max_blocks < blocks_needed ? max_blocks : blocks_needed;
    hipLaunchKernelGGL(cell_prep_kernel, dim3(1), dim3(64), 0, stream, prm.counters);
    hipLaunchKernelGGL(cell_group_kernel, dim3((unsigned)(blocks > 0 ? blocks : 1)), dim3(kGThreads), 0, stream, prm);
    hipLaunchKernelGGL(cell_big_kernel, dim3((unsigned)(a.waves1 / 4)), dim3(256), 0, stream, prm);
    return hipGetLastError();
}

}  // namespace pg
