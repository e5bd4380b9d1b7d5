// Capacity tier 0 of the cell construction: 16 lanes build one cell
// together, with the whole cell in LDS.  Included by pg_cell.hip (it uses
// CellParams, the tier constants and the FP64 helpers defined there).
//
// Per cell (2,688 bytes of LDS, plus a 1 KB wait list per wave):
//   vpos[64]  float4  vertex position y = N/D in seed-relative doubled
//                     coordinates, and a rigorous error radius (w)
//   vtri[64]  uint32  plane triple (i, j, k), ordered so that D > 0
//   pl[48]    float4  plane normal a and offset H (a . y <= H)
//   plx[48]   float2  permanent of H and the owner (sorted point id, or a
//                     negative box-plane code)
//   tmp[64]   uint32  clip and emission scratch
// Vertices live in a pool of 64 slots with a group-uniform valid mask, so a
// clip frees and refills slots without moving data.  A cell that outgrows
// the pool (64 vertices, 48 planes or 48 neighbours) is queued for tier 1.
//
// Lane parallelism: vertex sweeps (slot v on lane v mod 16), the batch of up
// to 16 candidates of a leaf (one per lane, then offered in order), the
// boundary edges of a clip and the new vertices they make, bounds
// reductions and the row sort.  All control decisions are group-uniform
// (ballots restricted to the group's 16 lanes).
#pragma once

namespace pg {

#ifndef PG_GROUP_LANES
#define PG_GROUP_LANES 16
#endif
constexpr int kGL = PG_GROUP_LANES;      // lanes per cell
constexpr uint32_t kGMask = (1u << kGL) - 1u;
constexpr int kGV = 64;                  // vertex slots (valid mask is 64 bits)
constexpr int kGP = 48;                  // plane slots
constexpr int kGStack = 64;
constexpr int kGThreads = 256;           // threads per workgroup
constexpr int kGCellsPerBlock = kGThreads / kGL;
static_assert(kGP <= 64 && kGP <= kRowCap, "plane masks are 64 bits; rows fit the tier-0 stride");

struct alignas(16) GroupLDS {
    float4 vpos[kGV];
    float4 pl[kGP];
    float2 plx[kGP];
    uint32_t vtri[kGV];
    uint32_t tmp[kGV];
};

// Best-first wait list shared by the cells of one wave (packet traversal).
constexpr int kWList = 128;
struct alignas(16) WaveList {
    uint32_t ref[kWList];     // parent << 1 | child
    float key[kWList];
};

struct Cand {
    float ax, ay, az, H, pH, pa1;
    int owner;
};

__device__ __forceinline__ int select_bit(uint64_t m, int j) {
    for (int t = 0; t < j; ++t) m &= m - 1;
    return (int)__builtin_ctzll(m);
}

struct Group {
    GroupLDS *S;
    int gl;       // lane within the group
    int gshift;   // bit offset of the group in a wave ballot

    const float4 *pts;
    const uint32_t *perm;
    const uint32_t *inv;
    const Node *nodes;
    uint32_t root_ref;
    float box_lo[3], box_hi[3];

    uint32_t seed;
    float sx, sy, sz, sw;
    uint64_t valid;
    int np;
    float L[3], U[3], L2[3], U2[3];  // AABB of the cell and its squared ends
    bool bounded;
    int status;
    bool overflow;
    bool active;    // the group holds a cell in this round
    bool dirty;     // the cell changed since the AABB was computed
    uint64_t bloom;
    bool in_guess;
    WaveList *W;
    int lane64;

    uint32_t c_recheck, c_unresolved, c_inconsistent, c_overflow, c_clips, c_cands, c_nodes, c_leaves, c_skipped;

    // ---------------------------------------------------------- group ops
    __device__ __forceinline__ uint32_t gballot(bool p) const {
        return (uint32_t)(__ballot(p) >> gshift) & kGMask;
    }
    __device__ __forceinline__ bool gany(bool p) const { return gballot(p) != 0u; }
    __device__ __forceinline__ bool live() const { return active && status == kSuccess && !overflow; }
    __device__ __forceinline__ float gbf(float x, int k) const { return __shfl(x, k, kGL); }
    __device__ __forceinline__ int gbi(int x, int k) const { return __shfl(x, k, kGL); }
    __device__ __forceinline__ uint64_t gor64(uint64_t x) const {
        uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
        for (int o = kGL / 2; o > 0; o >>= 1) {
            lo |= (uint32_t)__shfl_xor((int)lo, o, kGL);
            hi |= (uint32_t)__shfl_xor((int)hi, o, kGL);
        }
        return ((uint64_t)hi << 32) | lo;
    }
    __device__ __forceinline__ float gmin(float x) const {
        for (int o = kGL / 2; o > 0; o >>= 1) x = fminf(x, __shfl_xor(x, o, kGL));
        return x;
    }
    __device__ __forceinline__ float gmax(float x) const {
        for (int o = kGL / 2; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o, kGL));
        return x;
    }
    __device__ __forceinline__ int gsum(int x) const {
        for (int o = kGL / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kGL);
        return x;
    }
    // exclusive prefix sum over the group; total in `tot`
    __device__ __forceinline__ int gexcl(int x, int &tot) const {
        int s = x;
        for (int o = 1; o < kGL; o <<= 1) {
            const int y = __shfl_up(s, o, kGL);
            if (gl >= o) s += y;
        }
        tot = __shfl(s, kGL - 1, kGL);
        return s - x;
    }

    // ---------------------------------------------------------- FP64 path
    __device__ void plane64(int owner, double a[3], double &H, double pa[3], double &pH) const {
        const double s[3] = {(double)sx, (double)sy, (double)sz};
        if (owner >= 0) {
            const float4 p = pts[owner];
            a[0] = (double)p.x - s[0];
            a[1] = (double)p.y - s[1];
            a[2] = (double)p.z - s[2];
            const double dw = (double)p.w - (double)sw;
            const double a2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
            H = a2 - dw;
            pH = a2 + fabs(dw);
            for (int k = 0; k < 3; ++k) pa[k] = fabs(a[k]);
        } else {
            const int code = -1 - owner, axis = code >> 1, side = code & 1;
            for (int k = 0; k < 3; ++k) a[k] = 0.0, pa[k] = 0.0;
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

    __device__ void vertex64(uint32_t t, double N[3], double &D, double pN[3], double &pD) const {
        double a[3][3], H[3], pa[3][3], pH[3];
        for (int e = 0; e < 3; ++e) {
            const int owner = __float_as_int(S->plx[tri_get<8>(t, e)].y);
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

    // FP64 removal decision for vertex v against the plane of point q_owner.
    __device__ __noinline__ bool recheck64(int v, int q_owner) {
        c_recheck++;
        const uint32_t t = S->vtri[v];
        double N[3], D, pN[3], pD;
        vertex64(t, N, D, pN, pD);
        double a[3], H, pa[3], pH;
        plane64(q_owner, a, H, pa, pH);
        const double Pv = a[0] * N[0] + a[1] * N[1] + a[2] * N[2] - H * D;
        const double pP = pa[0] * pN[0] + pa[1] * pN[1] + pa[2] * pN[2] + pH * pD;
        if (fabs(D) > kG64 * pD && pD > kTiny64 && D < 0.0) c_inconsistent++;
        if (fabs(Pv) > kG64 * pP && pP > kTiny64) return Pv > 0.0;
        // A point identical (position and weight) to the owner of one of the
        // vertex's planes gives the same plane: P = 0 exactly, not beyond.
        const float4 q = pts[q_owner];
        for (int e = 0; e < 3; ++e) {
            const int o = __float_as_int(S->plx[tri_get<8>(t, e)].y);
            if (o < 0) continue;
            const float4 r = pts[o];
            if (r.x == q.x && r.y == q.y && r.z == q.z && r.w == q.w) return false;
        }
        c_unresolved++;
        return Pv > 0.0;
    }

    // Position and error radius of a vertex whose FP32 orientation is not
    // safely positive; the radius is +inf when FP64 cannot bound it either
    // (every test on that vertex then goes to FP64).
    __device__ __noinline__ float4 position64(uint32_t t) {
        double N[3], D, pN[3], pD;
        vertex64(t, N, D, pN, pD);
        const double e64 = kG64 * pD;
        if (!(D > 2.0 * e64 && pD > kTiny64)) return make_float4(0.f, 0.f, 0.f, __int_as_float(0x7f800000));
        float x[3];
        float eps = 0.0f;
        for (int k = 0; k < 3; ++k) {
            const double X = N[k] / D;
            const double E = (kG64 * pN[k] + fabs(X) * e64) / (D - e64) * (1.0 + 1e-12) + fabs(X) * 1e-12;
            x[k] = (float)X;
            eps = fmaxf(eps, (float)E + fabsf(x[k]) * kSlack);
        }
        return make_float4(x[0], x[1], x[2], eps + kTiny32);
    }

    // ---------------------------------------------------------- vertices
    // Builds vertex (i, j, k) into `slot` (one lane).  With
    // |N - N~| <= g pN and |D - D~| <= g pD (g = 2e-6), the position error is
    // |y - N~/D~| <= (g pN + |y~| g pD) / (D~ - g pD), plus division rounding.
    __device__ void make_vertex(int slot, uint32_t i, uint32_t j, uint32_t k) {
        const float4 A = S->pl[i], B = S->pl[j], C = S->pl[k];
        const float pHA = S->plx[i].x, pHB = S->plx[j].x, pHC = S->plx[k].x;
        const float c1x = B.y * C.z - B.z * C.y, c1y = B.z * C.x - B.x * C.z, c1z = B.x * C.y - B.y * C.x;
        const float c2x = C.y * A.z - C.z * A.y, c2y = C.z * A.x - C.x * A.z, c2z = C.x * A.y - C.y * A.x;
        const float c3x = A.y * B.z - A.z * B.y, c3y = A.z * B.x - A.x * B.z, c3z = A.x * B.y - A.y * B.x;
        const float D = A.x * c1x + A.y * c1y + A.z * c1z;
        const float Nx = A.w * c1x + B.w * c2x + C.w * c3x;
        const float Ny = A.w * c1y + B.w * c2y + C.w * c3y;
        const float Nz = A.w * c1z + B.w * c2z + C.w * c3z;
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
        const uint32_t tri = tri_pack<8>(i, j, k);
        const float eD = kG32 * pD;
        float4 pos;
        if (D > 2.0f * eD && pD > kTiny32) {
            const float inv = 1.0f / D;
            const float den = (1.0f + kSlack) / (D - eD);
            const float x = Nx * inv, y = Ny * inv, z = Nz * inv;
            const float ex = (kG32 * pNx + fabsf(x) * eD) * den + fabsf(x) * kSlack;
            const float ey = (kG32 * pNy + fabsf(y) * eD) * den + fabsf(y) * kSlack;
            const float ez = (kG32 * pNz + fabsf(z) * eD) * den + fabsf(z) * kSlack;
            pos = make_float4(x, y, z, fmaxf(ex, fmaxf(ey, ez)) + kTiny32);
        } else {
            if (D < -eD && pD > kTiny32) c_inconsistent++;
            pos = position64(tri);
        }
        S->vpos[slot] = pos;
        S->vtri[slot] = tri;
    }

    // Is vertex v strictly beyond the plane of q?  Filter: the error of
    // a~ . y~ - H~ against a . y - H is at most |a|_1 eps + 2e-6 (sum |a_k y_k| + pH).
    __device__ __forceinline__ bool beyond(int v, const Cand &q) {
        const float4 p = S->vpos[v];
        const float tx = q.ax * p.x, ty = q.ay * p.y, tz = q.az * p.z;
        const float P = tx + ty + tz - q.H;
        const float err = q.pa1 * p.w * (1.0f + kSlack) + kG32 * (fabsf(tx) + fabsf(ty) + fabsf(tz) + q.pH) + kTiny32;
        if (fabsf(P) > err) return P > 0.0f;
        return recheck64(v, q.owner);
    }

    // ---------------------------------------------------------- bounds
    __device__ void update_bounds() {
        float lo0 = FLT_MAX, lo1 = FLT_MAX, lo2 = FLT_MAX, hi0 = -FLT_MAX, hi1 = -FLT_MAX, hi2 = -FLT_MAX;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < kGV / kGL; ++k) {
            const int v = gl + kGL * k;
            if ((valid >> v) & 1ull) {
                const float4 p = S->vpos[v];
                ok &= p.w < FLT_MAX;
                lo0 = fminf(lo0, p.x - p.w);
                hi0 = fmaxf(hi0, p.x + p.w);
                lo1 = fminf(lo1, p.y - p.w);
                hi1 = fmaxf(hi1, p.y + p.w);
                lo2 = fminf(lo2, p.z - p.w);
                hi2 = fmaxf(hi2, p.z + p.w);
            }
        }
        bounded = !gany(!ok);
        L[0] = gmin(lo0), L[1] = gmin(lo1), L[2] = gmin(lo2);
        U[0] = gmax(hi0), U[1] = gmax(hi1), U[2] = gmax(hi2);
        for (int k = 0; k < 3; ++k) L2[k] = L[k] * L[k], U2[k] = U[k] * U[k];
    }

    __device__ __forceinline__ bool too_far(const Cand &c) const {
        if (!bounded) return false;
        const float S_ = fmaxf(c.ax * L[0], c.ax * U[0]) + fmaxf(c.ay * L[1], c.ay * U[1]) +
                         fmaxf(c.az * L[2], c.az * U[2]);
        const float M = fabsf(c.ax) * fmaxf(fabsf(L[0]), fabsf(U[0])) + fabsf(c.ay) * fmaxf(fabsf(L[1]), fabsf(U[1])) +
                        fabsf(c.az) * fmaxf(fabsf(L[2]), fabsf(U[2]));
        return (c.H - S_) > kSlack * (c.pH + M) + kTiny32;
    }

    // Same two node bounds as the tier-1 cell (octant corner radius and
    // per-axis support); see Cell::node_prunable.
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
        // the octants a box overlaps form a product set, so the largest
        // corner radius over them is a sum of per-axis maxima
        const float rho2 = (fmaxf(alx <= 0.0f ? L2[0] : 0.0f, ahx >= 0.0f ? U2[0] : 0.0f) +
                            fmaxf(aly <= 0.0f ? L2[1] : 0.0f, ahy >= 0.0f ? U2[1] : 0.0f) +
                            fmaxf(alz <= 0.0f ? L2[2] : 0.0f, ahz >= 0.0f ? U2[2] : 0.0f)) *
                           (1.0f + kSlack);
        const float d2lo = d2 * (1.0f - kSlack);
        const float h = d2lo - dw - kSlack * (fabsf(lo.w) + fabsf(sw)) - kTiny32;
        if (4.0f * d2lo > rho2 && h > 0.0f && h * h > rho2 * d2 * (1.0f + kSlack)) return true;
        const float S_ = fmaxf(fmaxf(alx * L[0], alx * U[0]), fmaxf(ahx * L[0], ahx * U[0])) +
                         fmaxf(fmaxf(aly * L[1], aly * U[1]), fmaxf(ahy * L[1], ahy * U[1])) +
                         fmaxf(fmaxf(alz * L[2], alz * U[2]), fmaxf(ahz * L[2], ahz * U[2]));
        const float M = fmaxf(fabsf(alx), fabsf(ahx)) * fmaxf(fabsf(L[0]), fabsf(U[0])) +
                        fmaxf(fabsf(aly), fabsf(ahy)) * fmaxf(fabsf(L[1]), fabsf(U[1])) +
                        fmaxf(fabsf(alz), fabsf(ahz)) * fmaxf(fabsf(L[2]), fabsf(U[2]));
        return (hlow - S_) > kSlack * (d2 + fabsf(lo.w) + fabsf(sw) + M) + kTiny32;
    }

    // ---------------------------------------------------------- planes
    __device__ uint64_t used_planes() const {
        uint64_t m = 0;
#pragma unroll
        for (int k = 0; k < kGV / kGL; ++k) {
            const int v = gl + kGL * k;
            if ((valid >> v) & 1ull) {
                const uint32_t t = S->vtri[v];
                m |= (1ull << tri_get<8>(t, 0)) | (1ull << tri_get<8>(t, 1)) | (1ull << tri_get<8>(t, 2));
            }
        }
        return gor64(m);
    }

    // Drops unreferenced planes; new index of plane p = popcount(used below p).
    __device__ void garbage_collect() {
        const uint64_t used = used_planes();
#pragma unroll
        for (int k = 0; k < (kGP + kGL - 1) / kGL; ++k) {
            const int p = gl + kGL * k;
            const bool mv = p < np && ((used >> p) & 1ull);
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
            float2 b = make_float2(0.f, 0.f);
            if (mv) {
                a = S->pl[p];
                b = S->plx[p];
            }
            // every lane has read slot p before any lane writes a slot below
            // 16 (k + 1): all later reads are at or above that
            if (mv) {
                const int dst = __popcll(used & ((1ull << p) - 1ull));
                S->pl[dst] = a;
                S->plx[dst] = b;
            }
        }
#pragma unroll
        for (int k = 0; k < kGV / kGL; ++k) {
            const int v = gl + kGL * k;
            if ((valid >> v) & 1ull) {
                const uint32_t t = S->vtri[v];
                uint32_t r[3];
                for (int e = 0; e < 3; ++e) {
                    const uint32_t p = tri_get<8>(t, e);
                    r[e] = (uint32_t)__popcll(used & ((1ull << p) - 1ull));
                }
                S->vtri[v] = tri_pack<8>(r[0], r[1], r[2]);
            }
        }
        np = __popcll(used);
    }

    __device__ bool has_plane(int owner) const {
        if (!(bloom & bloom_bit((uint32_t)owner))) return false;
        bool hit = false;
#pragma unroll
        for (int k = 0; k < (kGP + kGL - 1) / kGL; ++k) {
            const int p = gl + kGL * k;
            hit |= p < np && __float_as_int(S->plx[p].y) == owner;
        }
        return gany(hit);
    }

    // ---------------------------------------------------------- clipping
    // Offers plane q (group-uniform).  Returns true when the cell changed.
    __device__ bool clip(const Cand &q) {
        uint64_t rem = 0;
#pragma unroll
        for (int k = 0; k < kGV / kGL; ++k) {
            if (((valid >> (kGL * k)) & (uint64_t)kGMask) == 0ull) continue;
            const int v = gl + kGL * k;
            const bool r = ((valid >> v) & 1ull) && beyond(v, q);
            rem |= (uint64_t)gballot(r) << (kGL * k);
        }
        if (rem == 0ull) return false;
        const int R = __popcll(rem);
        if (R == __popcll(valid)) {  // voronoi_convex_cell.cuh:430-435
            status = kEmptyCell;
            valid = 0ull;
            return true;
        }
        if (np >= kGP) garbage_collect();
        if (np >= kGP) {
            overflow = true;
            return false;
        }
        const int qs = np++;
        if (gl == 0) {
            S->pl[qs] = make_float4(q.ax, q.ay, q.az, q.H);
            S->plx[qs] = make_float2(q.pH, __int_as_float(q.owner));
        }
        if (in_guess) bloom |= bloom_bit((uint32_t)q.owner);
        // Boundary of the removed triangles: directed edges whose twin is not
        // removed.  The R removed triangles are spread one per lane (through
        // tmp[]), and each lane checks its three edges against all of them.
        if (R > kGL) return clip_wide(q, rem, qs);
        int base = 0;
#pragma unroll
        for (int k = 0; k < kGV / kGL; ++k) {
            const uint32_t bits = (uint32_t)(rem >> (kGL * k)) & kGMask;
            if ((bits >> gl) & 1u) S->tmp[base + __popc(bits & ((1u << gl) - 1u))] = S->vtri[gl + kGL * k];
            base += __popc(bits);
        }
        const bool mine_t = gl < R;
        const uint32_t tr = mine_t ? S->tmp[gl] : 0u;
        uint32_t edge[3];
        uint32_t emask = 0;
        int sx_ = 0, sy_ = 0;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const uint32_t x = tri_get<8>(tr, e), y = tri_get<8>(tr, (e + 1) % 3);
            bool twin = false;
            for (int u = 0; u < R; ++u) twin |= tri_has_edge<8>((uint32_t)__shfl((int)tr, u, kGL), y, x);
            edge[e] = x | (y << 8);
            if (mine_t && !twin) {
                emask |= 1u << e;
                sx_ += (int)(x * x + 1u);
                sy_ += (int)(y * y + 1u);
            }
        }
        int nb;
        int j = gexcl(__popc(emask), nb);
        // every plane of the boundary cycle starts one edge and ends one
        if (nb < 3 || gsum(sx_) != gsum(sy_)) {
            status = kInconsistentBoundary;
            if (gl == 0) c_inconsistent++;
            return true;
        }
        const uint64_t freem = ~(valid & ~rem);
        if (nb > __popcll(freem)) {
            overflow = true;
            return false;
        }
        uint64_t newbits = 0;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            if (!((emask >> e) & 1u)) continue;
            const int slot = select_bit(freem, j++);
            make_vertex(slot, (uint32_t)qs, edge[e] & 255u, edge[e] >> 8);
            newbits |= 1ull << slot;
        }
        valid = (valid & ~rem) | gor64(newbits);
        dirty = true;
        return true;
    }

    // Boundary pass for clips that remove more triangles than a group has
    // lanes (early clips of the initial box): each lane walks its own rows.
    __device__ bool clip_wide(const Cand &q, uint64_t rem, int qs) {
        (void)q;
        uint32_t edge[kGV / kGL * 3];
        uint32_t emask = 0;
        int sx_ = 0, sy_ = 0;
#pragma unroll
        for (int k = 0; k < kGV / kGL; ++k) {
            const int v = gl + kGL * k;
            const bool r = (rem >> v) & 1ull;
            const uint32_t t = r ? S->vtri[v] : 0u;
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                edge[3 * k + e] = 0;
                if (!r) continue;
                const uint32_t x = tri_get<8>(t, e), y = tri_get<8>(t, (e + 1) % 3);
                bool twin = false;
                uint64_t m = rem;
                while (m && !twin) {
                    const int u = (int)__builtin_ctzll(m);
                    m &= m - 1ull;
                    twin = tri_has_edge<8>(S->vtri[u], y, x);
                }
                if (!twin) {
                    emask |= 1u << (3 * k + e);
                    edge[3 * k + e] = x | (y << 8);
                    sx_ += (int)(x * x + 1u);
                    sy_ += (int)(y * y + 1u);
                }
            }
        }
        int nb;
        int j = gexcl(__popc(emask), nb);
        if (nb < 3 || gsum(sx_) != gsum(sy_)) {
            status = kInconsistentBoundary;
            if (gl == 0) c_inconsistent++;
            return true;
        }
        const uint64_t freem = ~(valid & ~rem);
        if (nb > __popcll(freem)) {
            overflow = true;
            return false;
        }
        uint64_t newbits = 0;
#pragma unroll
        for (int t = 0; t < kGV / kGL * 3; ++t) {
            if (!((emask >> t) & 1u)) continue;
            const int slot = select_bit(freem, j++);
            make_vertex(slot, (uint32_t)qs, edge[t] & 255u, edge[t] >> 8);
            newbits |= 1ull << slot;
        }
        valid = (valid & ~rem) | gor64(newbits);
        dirty = true;
        return true;
    }

    // Prepares lane's candidate j; false when it cannot cut the cell.
    __device__ bool prepare(uint32_t j, Cand &c, bool &emptied) const {
        const float4 p = pts[j];
        c.ax = p.x - sx;
        c.ay = p.y - sy;
        c.az = p.z - sz;
        const float dw = p.w - sw;
        c.owner = (int)j;
        if (c.ax == 0.0f && c.ay == 0.0f && c.az == 0.0f) {
            emptied = dw > 0.0f;  // coincident point: half-space empty iff w_p > w_s
            return false;
        }
        const float a2 = c.ax * c.ax + c.ay * c.ay + c.az * c.az;
        c.H = a2 - dw;
        c.pH = a2 + fabsf(dw);
        c.pa1 = fabsf(c.ax) + fabsf(c.ay) + fabsf(c.az);
        return !too_far(c);
    }

    // Offers a batch of up to 16 candidates (lane-specific j when `has`).
    __device__ void offer_batch(bool has, uint32_t j) {
        Cand c;
        bool emptied = false;
        bool live = has && prepare(j, c, emptied);
        if (gany(emptied)) {
            status = kEmptyCell;
            valid = 0ull;
            return;
        }
        if (gl == 0) c_skipped += __popc(gballot(has && !live));
        uint32_t todo = gballot(live);
        while (todo) {
            const int b = __ffs(todo) - 1;
            todo &= todo - 1u;
            Cand q;
            q.ax = gbf(c.ax, b);
            q.ay = gbf(c.ay, b);
            q.az = gbf(c.az, b);
            q.H = gbf(c.H, b);
            q.pH = gbf(c.pH, b);
            q.pa1 = gbf(c.pa1, b);
            q.owner = gbi(c.owner, b);
            if (bloom && has_plane(q.owner)) continue;
            if (gl == 0) c_cands++;
            if (clip(q)) {
                if (gl == 0) c_clips++;
                if (status != kSuccess) return;
            } else if (overflow) {
                return;
            }
        }
        if (dirty) {
            update_bounds();
            dirty = false;
        }
    }

    __device__ void init_box() {
        const float s[3] = {sx, sy, sz};
        if (gl < 6) {
            const int axis = gl >> 1, side = gl & 1;
            const float h = side ? 2.0f * (box_hi[axis] - s[axis]) : 2.0f * (s[axis] - box_lo[axis]);
            float4 pl = make_float4(0.f, 0.f, 0.f, h);
            const float sg = side ? 1.0f : -1.0f;
            if (axis == 0) pl.x = sg;
            if (axis == 1) pl.y = sg;
            if (axis == 2) pl.z = sg;
            S->pl[gl] = pl;
            S->plx[gl] = make_float2(fabsf(h), __int_as_float(-1 - gl));
        }
        np = 6;
        if (gl < 8) {
            const int cx = gl >> 2 & 1, cy = gl >> 1 & 1, cz = gl & 1;
            const uint32_t X = cx, Y = 2 + cy, Z = 4 + cz;
            const int sgn = (cx ? 1 : -1) * (cy ? 1 : -1) * (cz ? 1 : -1);
            if (sgn > 0) make_vertex(gl, X, Y, Z);
            else make_vertex(gl, X, Z, Y);
        }
        valid = 0xFFull;
        bloom = 0;
        update_bounds();
        dirty = false;
    }

    // ---------------------------------------------------------- traversal
    __device__ void offer_leaf(uint32_t ref) {
        if (gl == 0) c_leaves++;
        const uint32_t f = leaf_first(ref), c = leaf_count(ref);
        for (uint32_t o = 0; o < c && status == kSuccess && !overflow; o += kGL) {
            const uint32_t j = f + o + (uint32_t)gl;
            offer_batch(o + (uint32_t)gl < c && j != seed, j);
        }
    }

    // Smallest-key entry of the wave's wait list (wave arg-min), -1 if empty.
    __device__ int list_argmin(int sz, float &kmin) const {
        float k = FLT_MAX;
        int idx = -1;
#pragma unroll
        for (int j = 0; j < kWList / kWave; ++j) {
            const int t = lane64 + kWave * j;
            if (t < sz) {
                const float kt = W->key[t];
                if (idx < 0 || kt < k) {
                    k = kt;
                    idx = t;
                }
            }
        }
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const float ok = __shfl_xor(k, o);
            const int oi = __shfl_xor(idx, o);
            if (oi >= 0 && (idx < 0 || ok < k || (ok == k && oi < idx))) {
                k = ok;
                idx = oi;
            }
        }
        kmin = k;
        return idx;
    }

    // Wave-wide: does any group still need this box?  key: smallest bound.
    __device__ bool packet_needs(float4 lo, float4 hi, float &key) const {
        float k = FLT_MAX;
        const bool need = live() && !node_prunable(lo, hi, k);
        const bool any = __ballot(gl == 0 && need) != 0ull;
        float kk = need ? k : FLT_MAX;
        for (int o = kGL; o < kWave; o <<= 1) kk = fminf(kk, __shfl_xor(kk, o));
        key = kk;
        return any;
    }

    // Best-first packet traversal: the four cells of a wave (Morton
    // neighbours) walk the hierarchy together.  A box is expanded while any
    // of them may still be cut by a point in it; its key is the smallest
    // lower bound of H over the cells that need it; leaves are offered to
    // every cell, each filtering and clipping on its own lanes.  Unpruned
    // children wait in one wave-wide keyed list and the wave always goes on
    // with the smallest key (the order voronoi_bvh.cuh:246-283 keeps with its
    // min-extraction pops), re-testing a waiting box when it comes out.  A
    // full list sends the wave's cells to the large-capacity tier instead of
    // dropping entries.
    __device__ void traverse() {
        uint32_t ref = root_ref;
        int sz = 0;
        while (true) {
            bool have = false;
            if (is_leaf(ref)) {
                if (live()) offer_leaf(ref);
            } else {
                if (lane64 == 0) c_nodes++;
                const Node nd = nodes[ref];
                float k0, k1;
                const bool n0 = packet_needs(nd.lo0, nd.hi0, k0);
                const bool n1 = packet_needs(nd.lo1, nd.hi1, k1);
                const uint32_t r0 = __float_as_uint(nd.hi0.w), r1 = __float_as_uint(nd.hi1.w);
                if (n0 || n1) {
                    const bool near0 = !n1 || (n0 && k0 <= k1);
                    const float knear = near0 ? k0 : k1;
                    if (n0 && n1) {  // the farther child waits
                        if (sz >= kWList) {
                            overflow = true;
                            return;
                        }
                        if (lane64 == 0) {
                            W->ref[sz] = (ref << 1) | (near0 ? 1u : 0u);
                            W->key[sz] = near0 ? k1 : k0;
                        }
                        sz++;
                    }
                    float kmin;
                    const int im = list_argmin(sz, kmin);
                    if (im < 0 || knear <= kmin) {
                        ref = near0 ? r0 : r1;
                        have = true;
                    } else {  // a waiting box is nearer: the near child waits too
                        if (sz >= kWList) {
                            overflow = true;
                            return;
                        }
                        if (lane64 == 0) {
                            W->ref[sz] = (ref << 1) | (near0 ? 0u : 1u);
                            W->key[sz] = knear;
                        }
                        sz++;
                    }
                }
            }
            while (!have) {
                float kmin;
                const int im = list_argmin(sz, kmin);
                if (im < 0) return;
                const uint32_t e = W->ref[im];
                if (lane64 == 0) {
                    W->ref[im] = W->ref[sz - 1];
                    W->key[im] = W->key[sz - 1];
                }
                sz--;
                const uint32_t parent = e >> 1, c = e & 1u;
                const Node nd = nodes[parent];
                float k;
                const float4 lo = c ? nd.lo1 : nd.lo0;
                const float4 hi = c ? nd.hi1 : nd.hi0;
                if (packet_needs(lo, hi, k)) {
                    ref = __float_as_uint(hi.w);
                    have = true;
                }
            }
        }
    }

    // ---------------------------------------------------------- output
    // Row = sorted original ids of the owners of referenced planes
    // (voronoi_ultra.cu:125-138).  Returns false when the row does not fit.
    __device__ bool emit(const CellParams &prm, uint32_t orig, int32_t *row) {
        int F = 0;
        if (status != kEmptyCell) {
            const uint64_t used = used_planes();
            if (prm.hdr->has_duplicates != 0u) {
                // every exact duplicate of an owner carries the same face
                bool full = false;
                if (gl == 0) {
                    for (int p = 0; p < np; ++p) {
                        if (!((used >> p) & 1ull)) continue;
                        const int owner = __float_as_int(S->plx[p].y);
                        if (owner < 0) continue;
                        uint32_t m = prm.dup_head[owner];
                        while (m != 0xFFFFFFFFu) {
                            const int32_t id = (int32_t)perm[m];
                            int k = F;
                            while (k > 0 && row[k - 1] > id) k--;
                            if (!(k > 0 && row[k - 1] == id)) {
                                if (F < kGP) {
                                    for (int t = F; t > k; --t) row[t] = row[t - 1];
                                    row[k] = id;
                                    F++;
                                } else {
                                    full = true;
                                }
                            }
                            m = prm.dup_next[m];
                        }
                    }
                }
                if (gany(full)) return false;
                F = gbi(F, 0);
            } else {
                int32_t ids[(kGP + kGL - 1) / kGL];
                int mine = 0;
#pragma unroll
                for (int k = 0; k < (kGP + kGL - 1) / kGL; ++k) {
                    const int p = gl + kGL * k;
                    ids[k] = -1;
                    if (p < np && ((used >> p) & 1ull)) {
                        const int owner = __float_as_int(S->plx[p].y);
                        if (owner >= 0) {
                            ids[k] = (int32_t)perm[owner];
                            mine++;
                        }
                    }
                }
                int at = gexcl(mine, F);
#pragma unroll
                for (int k = 0; k < (kGP + kGL - 1) / kGL; ++k)
                    if (ids[k] >= 0) S->tmp[at++] = (uint32_t)ids[k];
                // rank of each id among the row (ids are distinct)
#pragma unroll
                for (int k = 0; k < (kGP + kGL - 1) / kGL; ++k) {
                    if (ids[k] < 0) continue;
                    int rank = 0;
                    for (int t = 0; t < F; ++t) rank += (int32_t)S->tmp[t] < ids[k];
                    row[rank] = ids[k];
                }
            }
        }
        if (gl == 0) {
            prm.counts[orig] = F;
            prm.status[orig] = status;
        }
        return true;
    }

    __device__ void start(uint32_t cell, const CellParams &prm) {
        seed = cell;
        const float4 s = pts[cell];
        sx = s.x, sy = s.y, sz = s.z, sw = s.w;
        status = kSuccess;
        overflow = false;
        init_box();
        const uint32_t orig = perm[cell];
        in_guess = true;
        if (!prm.guesses && prm.seed_window > 0) {
            // Morton neighbours first: the cheap analogue of clipping by the
            // k nearest neighbours (voronoi_ultra.cu:70-85); order does not
            // change the result, only how fast the cell shrinks
            const int64_t j = (int64_t)cell - kGL / 2 + gl + (gl >= kGL / 2 ? 1 : 0);
            offer_batch(gl < prm.seed_window && j >= 0 && j < prm.n, (uint32_t)(j < 0 ? 0 : j));
        }
        if (prm.guesses) {
            const int32_t g0 = prm.guess_offsets[orig], g1 = prm.guess_offsets[orig + 1];
            for (int32_t g = g0; g < g1 && status == kSuccess && !overflow; g += kGL) {
                bool has = g + gl < g1;
                uint32_t j = 0;
                if (has) {
                    const int32_t gid = prm.guesses[g + gl];
                    has = gid >= 0 && (int64_t)gid < prm.n;  // voronoi_ultra.cu:76-78 skips negative ids
                    if (has) j = prm.guesses_sorted ? (uint32_t)gid : inv[gid];
                    has = has && j != cell;
                }
                offer_batch(has, j);
            }
        }
        in_guess = false;
    }

    __device__ void finish(uint32_t cell, const CellParams &prm) {
        const uint32_t orig = perm[cell];
        const int64_t local = (int64_t)cell - prm.begin;
        if (!overflow) {
            if (gl == 0) prm.row_offsets[cell] = local * prm.row_stride;
            if (emit(prm, orig, prm.rows + local * prm.row_stride)) return;
        }
        // queue for the large-capacity tier
        if (gl == 0) {
            c_overflow++;
            const uint32_t slot = atomicAdd(&prm.counters[1], 1u);
            if (slot < prm.big_cap) {
                prm.big_list[slot] = cell;
            } else {
                prm.row_offsets[cell] = local * prm.row_stride;
                prm.counts[orig] = 0;
                prm.status[orig] = kPlaneOverflow;
            }
        }
    }
};

// Capacity tier 0: 16 lanes per cell, a wave holds four Morton-consecutive
// cells per round (pulled from a global counter) and traverses for them
// together; cell state in LDS.
__global__ void __launch_bounds__(kGThreads) cell_group_kernel(const CellParams prm) {
    __shared__ GroupLDS lds[kGCellsPerBlock];
    __shared__ WaveList wlist[kGThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int grp = lane / kGL;
    Group g;
    g.S = &lds[(threadIdx.x / kWave) * (kWave / kGL) + grp];
    g.W = &wlist[threadIdx.x / kWave];
    g.gl = lane & (kGL - 1);
    g.lane64 = lane;
    g.gshift = grp * kGL;
    g.pts = prm.pts;
    g.perm = prm.perm;
    g.inv = prm.inv;
    g.nodes = prm.nodes;
    g.root_ref = prm.hdr->root_ref;
    for (int k = 0; k < 3; ++k) g.box_lo[k] = prm.hdr->box_lo[k], g.box_hi[k] = prm.hdr->box_hi[k];
    g.c_recheck = g.c_unresolved = g.c_inconsistent = g.c_overflow = g.c_clips = g.c_cands = 0;
    g.c_nodes = g.c_leaves = g.c_skipped = 0;
    uint32_t cells = 0;
    const int64_t count = prm.end - prm.begin;
    constexpr int kPacket = kWave / kGL;
    while (true) {
        uint32_t c0 = 0;
        if (lane == 0) c0 = atomicAdd(&prm.counters[0], (uint32_t)kPacket);
        c0 = (uint32_t)__shfl((int)c0, 0);
        if ((int64_t)c0 >= count) break;
        const uint64_t t0 = clock64();
        const int64_t c = (int64_t)c0 + grp;
        g.active = c < count;
        const uint32_t cell = (uint32_t)(prm.begin + (g.active ? c : 0));
        if (g.active) g.start(cell, prm);
        g.traverse();
        if (g.active) {
            g.finish(cell, prm);
            if (prm.cycles && g.gl == 0)
                prm.cycles[prm.perm[cell]] = (uint32_t)min(clock64() - t0, (uint64_t)0x7fffffff);
            if (g.gl == 0) cells++;
        }
    }
    const uint32_t vals[12] = {g.c_recheck, g.c_unresolved, g.c_inconsistent, g.c_overflow, g.c_clips, g.c_cands,
                               g.c_nodes, cells, 0, 0, g.c_leaves, g.c_skipped};
    for (int k = 0; k < 12; ++k) {
        const uint32_t s = wave_sum(vals[k]);
        if (lane == 0 && s) atomicAdd(&prm.stats[k], (unsigned long long)s);
    }
}

}  // namespace pg
