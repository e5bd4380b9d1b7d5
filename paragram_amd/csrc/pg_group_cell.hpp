// Cooperative cell construction: G lanes build one cell together, with the
// whole cell in LDS.  Included by pg_cell.hip (it uses CellParams, the
// precision constants and the plane-triple helpers defined there).
//
// Four instances (capacity tiers):
//   tier 0: G = 16 lanes, 64 vertex slots, 48 plane slots (about 3.3 KB of
//           LDS per cell, 4 cells per wave) -- every cell starts here;
//   tier m: G = 64 lanes, 192 vertex slots, 128 plane slots (one cell per
//           wave, 12 waves per CU) -- the few cells that outgrow tier 0
//           (cells with more than about 34 faces);
//   tier 1: G = 64 lanes, 512 vertex slots, 255 plane slots (one cell per
//           wave, 8 waves per CU) -- the cells that outgrow tier m (e.g.
//           near-range LiDAR returns with 200+ faces);
//   tier 2: G = 64 lanes, 1024 vertex slots, 1023 plane slots, 10-bit plane
//           indices (one cell per wave and block) -- the rare cells that
//           outgrow tier 1.
//
// Per cell in LDS:
//   vpos[V]  float4  vertex position y = N/D in seed-relative doubled
//                    coordinates, and a rigorous error radius (w)
//   vtri[V]  uint32  plane triple (i, j, k), ordered so that D > 0
//   pl[P]    float4  plane normal a and offset H (a . y <= H)
//   plx[P]   float2  permanent of H and the owner (sorted point id, or a
//                    negative box-plane code)
//   tmp[V]   uint32  clip and emission scratch
//   pm[P/2]  uint32  (tier 0) per plane, a 16-bit mask of the lanes whose
//                    removed triangle holds the plane; zero between clips
//   tent[8]  2 x u32 (tier 0) removal tests taken by their FP32 sign
// Vertices live in a pool of V slots with a group-uniform valid mask (slot
// v = lane + G k belongs to lane v mod G), so a clip frees and refills slots
// without moving data.
//
// Lane parallelism: vertex sweeps, the batch of up to G candidates of a leaf
// (one per lane, then offered in order), the boundary edges of a clip and
// the new vertices they make, bounds reductions and the row sort.  All
// control decisions are group-uniform (ballots restricted to the group).
// The cells of one wave (Morton neighbours) share one best-first traversal.
#pragma once

namespace pg {

template <int V, int P>
struct alignas(16) GroupLDS {
    float4 vpos[V];   // y~ and w (see store_vertex)
    float4 pl[P];
    float2 plx[P];
    uint32_t vtri[V];
    uint32_t tmp[V];
    uint32_t pm[P <= 64 ? (P + 1) / 2 : 1];
    uint32_t ctr[kNumCtr];  // diagnostics, in the order of the statistics slots
    uint32_t exact_events;  // removal tests of the current cell decided by expansion arithmetic
    // (tier 0) removal tests taken by their FP32 sign: the vertex's plane
    // triple (slots, kept and renumbered by the plane collection), and the
    // query point with the sign taken in bit 31
    uint32_t tent[P <= 64 ? kTentMax : 1][2];
    // their number (every lane of the group writes it, with the same value,
    // so each lane reads back its own store)
    uint32_t ntent;
};

// Candidates of one batch, staged for the group: (a, H) and (pH, owner).
template <int G>
struct alignas(16) CandLDS {
    float4 aH[G];
    float2 po[G];
};

struct Cand {
    float ax, ay, az, H, pH, pa1;
    float c2;  // 2e-6 pH + tiny: the plane's part of the removal filter's bound
    int owner;
};

// ------------------------------------------------------------------ masks
// A mask of W 64-bit words; slot v is bit v & 63 of word v >> 6.
template <int W>
__device__ __forceinline__ int mask_popc(const uint64_t (&m)[W]) {
    int c = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) c += __popcll(m[w]);
    return c;
}
template <int W>
__device__ __forceinline__ bool mask_bit(const uint64_t (&m)[W], int v) {
    uint64_t word = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) word = (v >> 6) == w ? m[w] : word;
    return (word >> (v & 63)) & 1ull;
}
// index of the j-th set bit (j < popcount): the word by prefix counts, then
// a binary search on popcounts within the word
template <int W>
__device__ __forceinline__ int mask_select(const uint64_t (&m)[W], int j) {
    uint64_t x = m[0];
    int base = 0;
#pragma unroll
    for (int w = 1; w < W; ++w) {
        const int c = __popcll(x);
        if (j >= c) {
            j -= c;
            x = m[w];
            base = 64 * w;
        }
    }
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const int c = __popcll(x & ((1ull << s) - 1ull));
        if (j >= c) {
            j -= c;
            x >>= s;
            base += s;
        }
    }
    return base;
}
// the lowest n set bits of m (n <= popcount)
template <int W>
__device__ __forceinline__ void mask_first(const uint64_t (&m)[W], int n, uint64_t (&out)[W]) {
    const int total = mask_popc(m);
    const int cut = n < total ? mask_select(m, n) : 64 * W;  // first bit not taken
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const int lo = 64 * w;
        const uint64_t below = cut >= lo + 64 ? ~0ull : (cut <= lo ? 0ull : ((1ull << (cut - lo)) - 1ull));
        out[w] = m[w] & below;
    }
}

// ------------------------------------------------------------------ FP64
// The FP64 recheck path, shared by all tiers.  Plane of owner o relative to
// the seed s (doubled coordinates): a = p - s, H = |a|^2 - (w_p - w_s); box
// planes a = -e_k, H = 2 (s_k - lo_k) and a = e_k, H = 2 (hi_k - s_k).  The
// p-terms are the matching permanents (the same expressions in absolute
// values), which bound the rounding error: |x~ - x| <= 1e-13 p for every
// quantity below (a few dozen roundings at 2^-53 each).
__device__ __forceinline__ void fp64_plane(const float4 *pts, const IndexHeader *hdr, int owner, double sx,
                                           double sy, double sz, double sw, double a[3], double &H, double pa[3],
                                           double &pH) {
    if (owner >= 0) {
        const float4 p = pts[owner];
        a[0] = (double)p.x - sx;
        a[1] = (double)p.y - sy;
        a[2] = (double)p.z - sz;
        const double dw = (double)p.w - sw;
        const double a2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        H = a2 - dw;
        pH = a2 + fabs(dw);
        for (int k = 0; k < 3; ++k) pa[k] = fabs(a[k]);
    } else {
        const int code = -1 - owner, axis = code >> 1, side = code & 1;
        const double s = axis == 0 ? sx : (axis == 1 ? sy : sz);
        for (int k = 0; k < 3; ++k) {
            a[k] = k == axis ? (side ? 1.0 : -1.0) : 0.0;
            pa[k] = k == axis ? 1.0 : 0.0;
        }
        H = side ? 2.0 * ((double)hdr->box_hi[axis] - s) : 2.0 * (s - (double)hdr->box_lo[axis]);
        pH = fabs(H);
    }
}

// Vertex of planes (i, j, k): N = H_i (a_j x a_k) + H_j (a_k x a_i) + H_k (a_i x a_j),
// D = a_i . (a_j x a_k), y = N / D, with permanents pN, pD.
// o0..o2: the owners of the three planes.
__device__ __forceinline__ void fp64_vertex_o(int o0, int o1, int o2, const float4 *pts, const IndexHeader *hdr,
                                              double sx, double sy, double sz, double sw, double N[3], double &D,
                                              double pN[3], double &pD) {
    double a[3][3], H[3], pa[3][3], pH[3];
    for (int e = 0; e < 3; ++e)
        fp64_plane(pts, hdr, e == 0 ? o0 : (e == 1 ? o1 : o2), sx, sy, sz, sw, a[e], H[e], pa[e], pH[e]);
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
// The same for plane triple t of a cell's plane list.
template <int kB>
__device__ __forceinline__ void fp64_vertex(const float2 *plx, const float4 *pts, const IndexHeader *hdr,
                                            uint32_t t, double sx, double sy, double sz, double sw, double N[3],
                                            double &D, double pN[3], double &pD) {
    fp64_vertex_o(__float_as_int(plx[tri_get<kB>(t, 0)].y), __float_as_int(plx[tri_get<kB>(t, 1)].y),
                  __float_as_int(plx[tri_get<kB>(t, 2)].y), pts, hdr, sx, sy, sz, sw, N, D, pN, pD);
}

// The removal test decided exactly (expansion arithmetic, pg_exact.hpp) for
// the vertex of the planes of o0..o2 against the plane of point q_owner: bit
// 0 beyond (sign(P) sign(D) > 0), bit 1 capacity exceeded (no decision).
// Its own call site keeps the expansions' private arrays off the common path.
__device__ __noinline__ int exact_removal(int o0, int o1, int o2, const float4 *pts, const IndexHeader *hdr,
                                          float sx, float sy, float sz, float sw, int q_owner) {
    xp::PlaneIn pl[4];
    for (int e = 0; e < 4; ++e) {
        const int o = e == 0 ? o0 : (e == 1 ? o1 : (e == 2 ? o2 : q_owner));
        if (o >= 0) {
            const float4 p = pts[o];
            pl[e] = xp::PlaneIn{p.x, p.y, p.z, p.w, -1};
        } else {
            pl[e] = xp::PlaneIn{0.f, 0.f, 0.f, 0.f, -1 - o};
        }
    }
    const float s4[4] = {sx, sy, sz, sw};
    int sD = 0, sP = 0;
    if (!xp::exact_signs(pl, s4, hdr->box_lo, hdr->box_hi, sD, sP)) return 2;
    return sD * sP > 0 ? 1 : 0;
}

// FP64 removal decision for the vertex of the planes of o0..o2 against the
// plane of point q_owner: beyond iff sign(P) sign(D) > 0 with
// P = a_q . N - H_q D (D > 0 for every vertex a cell keeps).  Returns bit 0:
// beyond, bit 1: undecided even by expansion arithmetic (capacity; the FP64
// signs taken), bit 2: the vertex orientation D is certainly negative,
// bit 3: decided by expansion arithmetic (the FP64 filter could not).
__device__ __forceinline__ int fp64_decide(int o0, int o1, int o2, const float4 *pts, const IndexHeader *hdr,
                                           float sx_, float sy_, float sz_, float sw_, int q_owner) {
    const double sx = sx_, sy = sy_, sz = sz_, sw = sw_;
    double N[3], D, pN[3], pD;
    fp64_vertex_o(o0, o1, o2, pts, hdr, sx, sy, sz, sw, N, D, pN, pD);
    double a[3], H, pa[3], pH;
    fp64_plane(pts, hdr, q_owner, sx, sy, sz, sw, a, H, pa, pH);
    const double Pv = a[0] * N[0] + a[1] * N[1] + a[2] * N[2] - H * D;
    const double pP = pa[0] * pN[0] + pa[1] * pN[1] + pa[2] * pN[2] + pH * pD;
    int flags = 0;
    const bool d_sure = fabs(D) > kG64 * pD && pD > kTiny64;
    if (d_sure && D < 0.0) flags |= 4;
    if (d_sure && fabs(Pv) > kG64 * pP && pP > kTiny64) return flags | ((Pv > 0.0) == (D > 0.0) ? 1 : 0);
    // A point identical (position and weight) to the owner of one of the
    // vertex's planes gives the same plane: P = 0 exactly, not beyond.
    const float4 q = pts[q_owner];
    for (int e = 0; e < 3; ++e) {
        const int o = e == 0 ? o0 : (e == 1 ? o1 : o2);
        if (o < 0) continue;
        const float4 r = pts[o];
        if (r.x == q.x && r.y == q.y && r.z == q.z && r.w == q.w) return flags;
    }
    const int x = exact_removal(o0, o1, o2, pts, hdr, sx_, sy_, sz_, sw_, q_owner);
    if (!(x & 2)) return flags | 8 | (x & 1);
    return flags | 2 | ((Pv > 0.0) == (D >= 0.0) ? 1 : 0);
}
// The same for plane triple t of a cell's plane list.
template <int kB>
__device__ __noinline__ int fp64_recheck(const float2 *plx, const float4 *pts, const IndexHeader *hdr, uint32_t t,
                                         float sx, float sy, float sz, float sw, int q_owner) {
    return fp64_decide(__float_as_int(plx[tri_get<kB>(t, 0)].y), __float_as_int(plx[tri_get<kB>(t, 1)].y),
                       __float_as_int(plx[tri_get<kB>(t, 2)].y), pts, hdr, sx, sy, sz, sw, q_owner);
}

// Position and error radius of a vertex of planes t by expansion
// arithmetic (pg_exact.hpp); false when its orientation is not positive or
// the expansions overflow.  Its own call site keeps the expansions' private
// arrays off the common path.
template <int kB>
__device__ __noinline__ bool exact_vertex(const float2 *plx, const float4 *pts, const IndexHeader *hdr, uint32_t t,
                                          float sx, float sy, float sz, float sw, float4 &out) {
    xp::PlaneIn pl[3];
    for (int e = 0; e < 3; ++e) {
        const int o = __float_as_int(plx[tri_get<kB>(t, e)].y);
        if (o >= 0) {
            const float4 p = pts[o];
            pl[e] = xp::PlaneIn{p.x, p.y, p.z, p.w, -1};
        } else {
            pl[e] = xp::PlaneIn{0.f, 0.f, 0.f, 0.f, -1 - o};
        }
    }
    const float s4[4] = {sx, sy, sz, sw};
    float y[3], w;
    if (!xp::exact_position(pl, s4, hdr->box_lo, hdr->box_hi, y, w)) return false;
    out = make_float4(y[0], y[1], y[2], w + kTiny32);
    return true;
}

// Position and error radius of a vertex whose FP32 orientation is not
// safely positive (or, in the tiers with an FP64 path, whose FP32 radius
// would be large).  FP64 first; where its bound is loose (more than 2^-12 of
// the position: D within a few thousand rounding errors of 0) the position
// is computed exactly; the radius is +inf when neither can bound it (every
// test on that vertex then goes to FP64).
template <int kB>
__device__ __noinline__ float4 fp64_position(const float2 *plx, const float4 *pts, const IndexHeader *hdr, uint32_t t,
                                             float sx, float sy, float sz, float sw) {
    double N[3], D, pN[3], pD;
    fp64_vertex<kB>(plx, pts, hdr, t, sx, sy, sz, sw, N, D, pN, pD);
    const double e64 = kG64 * pD;
    float4 r = make_float4(0.f, 0.f, 0.f, __int_as_float(0x7f800000));
    if (D > 2.0 * e64 && pD > kTiny64) {
        float x[3];
        float eps = 0.0f, ym = 0.0f;
        for (int k = 0; k < 3; ++k) {
            const double X = N[k] / D;
            const double E = (kG64 * pN[k] + fabs(X) * e64) / (D - e64) * (1.0 + 1e-12) + fabs(X) * 1e-12;
            x[k] = (float)X;
            eps = fmaxf(eps, (float)E + fabsf(x[k]) * kSlack);
            ym = fmaxf(ym, fabsf(x[k]));
        }
        r = make_float4(x[0], x[1], x[2], eps + kTiny32);
        if (eps <= ym * 0x1p-12f) return r;
    }
    float4 e;
    if (exact_vertex<kB>(plx, pts, hdr, t, sx, sy, sz, sw, e) && e.w < r.w) r = e;
    return r;
}

// ------------------------------------------------------------------ DPP
// Row (16-lane) reductions and scans on the data-parallel-primitive path of
// the VALU: no LDS round trip, unlike ds_bpermute shuffles.
constexpr int kDppXor1 = 0xB1;   // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;   // quad_perm [2, 3, 0, 1]
constexpr int kDppRor4 = 0x124;  // row_ror:4
constexpr int kDppRor8 = 0x128;  // row_ror:8
constexpr int kDppShr1 = 0x111, kDppShr2 = 0x112, kDppShr4 = 0x114, kDppShr8 = 0x118;  // row_shr:n

template <int C>
__device__ __forceinline__ int dpp(int x) {
    return __builtin_amdgcn_update_dpp(0, x, C, 0xF, 0xF, true);  // lanes shifted in from outside the row read 0
}
template <int C>
__device__ __forceinline__ float dppf(float x) {
    return __int_as_float(dpp<C>(__float_as_int(x)));
}
// Float minima and maxima over a row through an order-preserving integer
// key (sign-magnitude bits turned into two's complement): each step is one
// integer min or max with the DPP operand folded in, where fminf/fmaxf of a
// shuffled value would add two NaN-quieting instructions per step.  The
// values reduced here are never NaN (+inf is, and orders above every finite
// value; -0 orders just below +0).
__device__ __forceinline__ int f2ord(float x) {
    const int b = __float_as_int(x);
    return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float ord2f(int o) { return __int_as_float(o ^ ((o >> 31) & 0x7fffffff)); }
__device__ __forceinline__ float row_min(float x) {
    int o = f2ord(x);
    o = min(o, dpp<kDppXor1>(o));
    o = min(o, dpp<kDppXor2>(o));
    o = min(o, dpp<kDppRor4>(o));
    return ord2f(min(o, dpp<kDppRor8>(o)));
}
__device__ __forceinline__ float row_max(float x) {
    int o = f2ord(x);
    o = max(o, dpp<kDppXor1>(o));
    o = max(o, dpp<kDppXor2>(o));
    o = max(o, dpp<kDppRor4>(o));
    return ord2f(max(o, dpp<kDppRor8>(o)));
}
__device__ __forceinline__ int row_sum(int x) {
    x += dpp<kDppXor1>(x);
    x += dpp<kDppXor2>(x);
    x += dpp<kDppRor4>(x);
    return x + dpp<kDppRor8>(x);
}
__device__ __forceinline__ uint32_t row_or(uint32_t x) {
    x |= (uint32_t)dpp<kDppXor1>((int)x);
    x |= (uint32_t)dpp<kDppXor2>((int)x);
    x |= (uint32_t)dpp<kDppRor4>((int)x);
    return x | (uint32_t)dpp<kDppRor8>((int)x);
}
// inclusive prefix sum within the row
__device__ __forceinline__ int row_scan(int x) {
    x += dpp<kDppShr1>(x);
    x += dpp<kDppShr2>(x);
    x += dpp<kDppShr4>(x);
    return x + dpp<kDppShr8>(x);
}
__device__ __forceinline__ int rl(int x, int lane) { return __builtin_amdgcn_readlane(x, lane); }
__device__ __forceinline__ uint32_t rl(uint32_t x, int lane) { return (uint32_t)__builtin_amdgcn_readlane((int)x, lane); }
__device__ __forceinline__ float rl(float x, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}

// Values the whole wave agrees on, marked as such for the compiler (they
// then live in scalar registers).
__device__ __forceinline__ uint32_t uniform(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uniform(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }

// ------------------------------------------------------------------ group
// kExact: the tier carries the FP64 recheck path.  A tier without it gives
// up on a cell at the first test its FP32 filter cannot decide (`bail`), and
// the cell is rebuilt by the FP64-capable twin of the tier.
template <int G, int V, int P, int L, bool kExact>
struct Group {
    static constexpr int kRows = V / G;               // vertex slots per lane
    static constexpr int kW = V / 64;                 // words of a vertex mask
    static constexpr int kPW = (P + 63) / 64;         // words of a plane mask
    static constexpr int kPRows = (P + G - 1) / G;    // plane slots per lane
    static constexpr uint64_t kRowMask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    static constexpr int kB = P < 256 ? 8 : 10;       // bits of a plane slot in a triple
    static constexpr uint32_t kBMask = (1u << kB) - 1u;
    static_assert((G == 16 || G == 64) && L % kWave == 0 && V % 64 == 0 && P < (1 << kB) && 3 * kRows <= 64, "group geometry");

    static constexpr bool kIsExact = kExact;
    static constexpr int kLS = L / kWave;  // wait-list slots per lane
    GroupLDS<V, P> *S;
    CandLDS<G> *C;
    uint32_t lr[kLS];
    int gl;      // lane within the group
    int gshift;  // bit offset of the group in a wave ballot
    int lane64;

    const float4 *pts;
    const uint32_t *perm;
    const uint32_t *inv;
    const Node *nodes;
    const IndexHeader *hdr;
    uint32_t root_ref;

    uint32_t seed;
    float sx, sy, sz, sw;
    uint64_t valid[kW];
    int np;
    float L_[3], U_[3];  // AABB of the cell
    float Lq_[3], Uq_[3];  // their squares (the octant corner radii)
    float EM_;           // largest vertex error radius w of the cell
    bool bounded;
    int status;
    bool overflow;  // the cell left this tier (capacity, or `bail`)
    // why an FP32-only tier gave the cell up (0: it did not): 1 an undecided
    // removal test, 2 an undecided vertex orientation, 3 a predicted cube
    // that cut the cell (counted per cell after the traversal)
    uint32_t bail;
    bool cube;      // the cell started inside a predicted cube (see init_box)
    bool active;  // the group holds a cell in this round
    bool dirty;   // the cell changed since the AABB was computed
    uint64_t bloom;
    bool in_guess;
    bool bloom_all;  // every candidate whose owner hits the Bloom mask is checked for a plane already held
    // After the pre-clip: every point nearer the seed than this (squared,
    // lowered by a margin) is one of the guesses already offered.
    float guess_r2;

    // diagnostics counters live in LDS (registers are the scarce resource)
    __device__ __forceinline__ void count(int slot, uint32_t v = 1u) const { atomicAdd(&S->ctr[slot], v); }

    // ---------------------------------------------------------- group ops
    __device__ __forceinline__ uint64_t gballot(bool p) const { return (__ballot(p) >> gshift) & kRowMask; }
    __device__ __forceinline__ bool gany(bool p) const { return gballot(p) != 0ull; }
    __device__ __forceinline__ bool live() const { return active && status == kSuccess && !overflow; }
    __device__ __forceinline__ int gbi(int x, int k) const { return __shfl(x, k, G); }
    // Reductions and scans: DPP within each 16-lane row, rows combined
    // through readlane when the group is the whole wave.
    __device__ __forceinline__ uint64_t gor64(uint64_t x) const {
        uint32_t lo = row_or((uint32_t)x), hi = row_or((uint32_t)(x >> 32));
        if constexpr (G == 64) {
            lo = rl(lo, 0) | rl(lo, 16) | rl(lo, 32) | rl(lo, 48);
            hi = rl(hi, 0) | rl(hi, 16) | rl(hi, 32) | rl(hi, 48);
        }
        return ((uint64_t)hi << 32) | lo;
    }
    __device__ __forceinline__ float gmin(float x) const {
        x = row_min(x);
        if constexpr (G == 64) x = fminf(fminf(rl(x, 0), rl(x, 16)), fminf(rl(x, 32), rl(x, 48)));
        return x;
    }
    __device__ __forceinline__ float gmax(float x) const {
        x = row_max(x);
        if constexpr (G == 64) x = fmaxf(fmaxf(rl(x, 0), rl(x, 16)), fmaxf(rl(x, 32), rl(x, 48)));
        return x;
    }
    __device__ __forceinline__ int gsum(int x) const {
        x = row_sum(x);
        if constexpr (G == 64) x = rl(x, 0) + rl(x, 16) + rl(x, 32) + rl(x, 48);
        return x;
    }
    // exclusive prefix sum over the group; total in `tot`
    __device__ __forceinline__ int gexcl(int x, int &tot) const {
        const int incl = row_scan(x);
        if constexpr (G == 16) {
            tot = row_sum(x);
            return incl - x;
        } else {
            const int t0 = rl(incl, 15), t1 = rl(incl, 31), t2 = rl(incl, 47), t3 = rl(incl, 63);
            const int r = lane64 >> 4;
            tot = t0 + t1 + t2 + t3;
            return incl - x + (r > 0 ? t0 : 0) + (r > 1 ? t1 : 0) + (r > 2 ? t2 : 0);
        }
    }
    // row k of a vertex mask: the bits of slots gl + G k, lane order
    __device__ __forceinline__ uint64_t row(const uint64_t (&m)[kW], int k) const {
        return (m[(G * k) >> 6] >> ((G * k) & 63)) & kRowMask;
    }
    __device__ __forceinline__ bool own(const uint64_t (&m)[kW], int k) const {
        return (row(m, k) >> gl) & 1ull;
    }

    // ---------------------------------------------------------- FP64 path
    // (free functions below: a call that passed `this` would move the whole
    // group state from registers to scratch memory)
    __device__ __forceinline__ bool recheck64(int v, int q_owner) {
        count(kStatFp64Rechecks);
        const int r = fp64_recheck<kB>(S->plx, pts, hdr, S->vtri[v], sx, sy, sz, sw, q_owner);
        if (r & 2) count(kStatUnresolved);
        if (r & 4) count(kStatInconsistent);
        if (r & 8) atomicAdd(&S->exact_events, 1u);
        return (r & 1) != 0;
    }

    // ---------------------------------------------------------- vertices
    // Builds vertex (i, j, k) into `slot` (one lane).  With
    // |N - N~| <= g pN and |D - D~| <= g pD (g = 2e-6), the position error is
    // |y - N~/D~| <= (g pN + |y~| g pD) / (D~ - g pD), plus division rounding.
    // Returns false when the FP32 orientation is not safely positive: the
    // slot then holds radius -1 and fix_vertices() replaces it.  The tiers
    // with an FP64 path also leave to it the ill-conditioned vertices, whose
    // FP32 radius exceeds 2^-12 of the position (nearly parallel or coaxial
    // planes): one such vertex in a cell inflates its bounds until no node or
    // candidate can be pruned and every test needs FP64.  (The FP32-only
    // tier keeps them: an FP64 call site there costs it registers on every
    // cell, and a test such a vertex leaves open sends the cell to the twin.)
    __device__ __forceinline__ bool make_vertex(int slot, uint32_t i, uint32_t j, uint32_t k) {
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
        const uint32_t tri = tri_pack<kB>(i, j, k);
        const float eD = kG32 * pD;
        float4 pos = make_float4(0.f, 0.f, 0.f, -1.0f);
        bool ok = D > 2.0f * eD && pD > kTiny32;
        if (ok) {
            // hardware reciprocals (1 ulp): their error, and the products',
            // lie far inside the kSlack = 1e-6 relative terms of the bound
            const float inv = __builtin_amdgcn_rcpf(D);
            const float den = (1.0f + kSlack) * __builtin_amdgcn_rcpf(D - eD);
            const float x = Nx * inv, y = Ny * inv, z = Nz * inv;
            const float ex = (kG32 * pNx + fabsf(x) * eD) * den + fabsf(x) * kSlack;
            const float ey = (kG32 * pNy + fabsf(y) * eD) * den + fabsf(y) * kSlack;
            const float ez = (kG32 * pNz + fabsf(z) * eD) * den + fabsf(z) * kSlack;
            pos = make_float4(x, y, z, fmaxf(ex, fmaxf(ey, ez)) + kTiny32);
            if constexpr (kExact) {
                if (pos.w > 0x1p-12f * fmaxf(fabsf(x), fmaxf(fabsf(y), fabsf(z)))) {
                    pos.w = -1.0f;
                    ok = false;
                }
            }
        } else if (D < -eD && pD > kTiny32 && (kExact || S->ntent == 0u)) {
            count(kStatInconsistent);
        }
        store_vertex(slot, pos);
        S->vtri[slot] = tri;
        return ok;
    }

    // Vertex record: y~ and w = eps (1 + s) + 2e-6 max_k |y~_k|, which bounds
    // both the position error eps and the removal filter's rounding term
    // 2e-6 sum_k |a_k y~_k| <= 2e-6 |a|_1 max_k |y~_k|.  A negative w marks a
    // vertex left to the FP64 path; +inf stays +inf.
    __device__ __forceinline__ void store_vertex(int slot, float4 pos) {
        float w = pos.w;
        if (w >= 0.0f && w < FLT_MAX)
            w = w * (1.0f + kSlack) + kG32 * fmaxf(fabsf(pos.x), fmaxf(fabsf(pos.y), fabsf(pos.z)));
        S->vpos[slot] = make_float4(pos.x, pos.y, pos.z, w);
    }

    // FP64 positions for the slots of `m` that make_vertex left unresolved
    // (one call site: the rare path keeps its registers to itself).
    __device__ __forceinline__ void fix_vertices(const uint64_t (&m)[kW]) {
        uint32_t pend = 0;
#pragma unroll
        for (int k = 0; k < kRows; ++k)
            if (own(m, k) && S->vpos[gl + G * k].w < 0.0f) pend |= 1u << k;
        while (pend) {
            const int v = gl + G * (int)__builtin_ctz(pend);
            pend &= pend - 1u;
            store_vertex(v, fp64_position<kB>(S->plx, pts, hdr, S->vtri[v], sx, sy, sz, sw));
        }
    }

    // Is vertex v strictly beyond the plane of q?  Filter: the error of
    // a~ . y~ - H~ against a . y - H is at most |a|_1 eps + 2e-6 (sum |a_k y_k| + pH).
    // Returns 1 (beyond), 0 (not beyond), or when undecided (FP64 recheck)
    // -2 if the FP32 value is positive and -1 if not.
    // (with the stored w: |a|_1 w + 2e-6 pH + tiny bounds the error of the
    // FMA chain below against a . y - H)
    __device__ __forceinline__ int beyond32(int v, const Cand &q) const {
        const float4 p = S->vpos[v];
        const float Pv = fmaf(q.ax, p.x, fmaf(q.ay, p.y, fmaf(q.az, p.z, -q.H)));
        const float err = fmaf(q.pa1, p.w, q.c2);
        if (fabsf(Pv) > err) return Pv > 0.0f ? 1 : 0;
        return Pv > 0.0f ? -2 : -1;
    }

    // ---------------------------------------------------------- bounds
    __device__ __forceinline__ void update_bounds() {
        float lo0 = FLT_MAX, lo1 = FLT_MAX, lo2 = FLT_MAX, hi0 = -FLT_MAX, hi1 = -FLT_MAX, hi2 = -FLT_MAX, em = 0.0f;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < kRows; ++k) {
            if (own(valid, k)) {
                const float4 p = S->vpos[gl + G * k];
                ok &= p.w < FLT_MAX;
                lo0 = fminf(lo0, p.x - p.w);
                hi0 = fmaxf(hi0, p.x + p.w);
                lo1 = fminf(lo1, p.y - p.w);
                hi1 = fmaxf(hi1, p.y + p.w);
                lo2 = fminf(lo2, p.z - p.w);
                hi2 = fmaxf(hi2, p.z + p.w);
                em = fmaxf(em, p.w);
            }
        }
        bounded = !gany(!ok);
        L_[0] = gmin(lo0), L_[1] = gmin(lo1), L_[2] = gmin(lo2);
        U_[0] = gmax(hi0), U_[1] = gmax(hi1), U_[2] = gmax(hi2);
#pragma unroll
        for (int d = 0; d < 3; ++d) Lq_[d] = L_[d] * L_[d], Uq_[d] = U_[d] * U_[d];
        EM_ = gmax(em);
    }

    // Can the plane of candidate c not cut the cell?  max over the AABB of
    // a . y is sum_k max(a_k L_k, a_k U_k).
    __device__ __forceinline__ bool too_far(const Cand &c) const {
        if (!bounded) return false;
        const float S_ = fmaxf(c.ax * L_[0], c.ax * U_[0]) + fmaxf(c.ay * L_[1], c.ay * U_[1]) +
                         fmaxf(c.az * L_[2], c.az * U_[2]);
        const float M = fabsf(c.ax) * fmaxf(fabsf(L_[0]), fabsf(U_[0])) +
                        fabsf(c.ay) * fmaxf(fabsf(L_[1]), fabsf(U_[1])) +
                        fabsf(c.az) * fmaxf(fabsf(L_[2]), fabsf(U_[2]));
        return (c.H - S_) > kSlack * (c.pH + M) + kTiny32;
    }

    // Node prune.  Every point p of the box has a = p - s with |a| >= d and
    // H = |a|^2 - (w_p - w_s) >= |a|^2 - dw (dw = w_max - w_s).  Two tests,
    // either of which suffices:
    //  * octant radius (voronoi_bvh.cuh:71-107, pwr_bvh.cuh:150-185): a . y
    //    <= |a| rho with rho the largest AABB corner radius over the octants
    //    the box overlaps, so no cut when |a|^2 - rho |a| - dw >= 0 for all
    //    |a| >= d, i.e. 2d >= rho and (d^2 - dw)^2 >= rho^2 d^2, d^2 > dw;
    //  * the per-vertex balls (balls_clear), the exact form.
    // kSplit: the two halves of the group test two boxes at once (each lane
    // is given its half's box).
    template <bool kSplit>
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
        // a box that holds the seed meets every vertex's ball (each passes
        // through the seed), so neither test below can prune it
        if (!bounded || d2 == 0.0f) return false;
        // the octants a box overlaps form a product set, so the largest
        // corner radius over them is a sum of per-axis maxima
        const float rho2 = (fmaxf(alx <= 0.0f ? Lq_[0] : 0.0f, ahx >= 0.0f ? Uq_[0] : 0.0f) +
                            fmaxf(aly <= 0.0f ? Lq_[1] : 0.0f, ahy >= 0.0f ? Uq_[1] : 0.0f) +
                            fmaxf(alz <= 0.0f ? Lq_[2] : 0.0f, ahz >= 0.0f ? Uq_[2] : 0.0f)) *
                           (1.0f + kSlack);
        const float d2lo = d2 * (1.0f - kSlack);
        const float wslack = kSlack * (fabsf(lo.w) + fabsf(sw)) + kTiny32;
        const float h = d2lo - dw - wslack;
        if (4.0f * d2lo > rho2 && h > 0.0f && h * h > rho2 * d2 * (1.0f + kSlack)) return true;
        return balls_clear<kSplit>(alx, aly, alz, ahx, ahy, ahz, dw + wslack);
    }

    // Exact form of the prune, per vertex.  Vertex y is beyond the plane of
    // p iff a . y - |a|^2 + dw > 0, i.e. in doubled coordinates
    // |2a - y|^2 < |y|^2 + 4 dw: p lies in the ball about the vertex through
    // the seed (grown by the weight difference).  No point of the box cuts the
    // cell when the box misses every vertex's ball.  The box [2 al, 2 ah] is
    // widened by the cell's largest error radius (each coordinate of y lies
    // within w of y~), and sum_k (|y~_k| + w)^2 bounds |y|^2.  The lanes that
    // test one box (the group, or its half when kSplit) share the vertex
    // slots.
    template <bool kSplit>
    __device__ __forceinline__ bool balls_clear(float alx, float aly, float alz, float ahx, float ahy, float ahz,
                                                float dwmax) const {
        constexpr int H = kSplit ? G / 2 : G;
        const int hl = gl & (H - 1);
        const float Ax = fmaf(2.0f, alx, -EM_), Ay = fmaf(2.0f, aly, -EM_), Az = fmaf(2.0f, alz, -EM_);
        const float Bx = fmaf(2.0f, ahx, EM_), By = fmaf(2.0f, ahy, EM_), Bz = fmaf(2.0f, ahz, EM_);
        const float dw4 = 4.0f * dwmax;
        bool cut = false;
#pragma unroll
        for (int k = 0; k < V / H; ++k) {
            const int v = hl + H * k;
            if (!((valid[(H * k) >> 6] >> (((H * k) & 63) + hl)) & 1ull)) continue;
            const float4 p = S->vpos[v];
            const float rx = fabsf(p.x) + p.w, ry = fabsf(p.y) + p.w, rz = fabsf(p.z) + p.w;
            const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz)) * (1.0f + 4.0f * kSlack);
            // distance to the box per axis: y minus y clamped to [A, B]
            // (A <= B; the same difference as A - y or y - B, so the same
            // rounding)
            const float tx = p.x - __builtin_amdgcn_fmed3f(p.x, Ax, Bx);
            const float ty = p.y - __builtin_amdgcn_fmed3f(p.y, Ay, By);
            const float tz = p.z - __builtin_amdgcn_fmed3f(p.z, Az, Bz);
            const float d2 = fmaf(tx, tx, fmaf(ty, ty, tz * tz));
            cut |= !(d2 * (1.0f - kSlack) > r2 + dw4);
        }
        const uint64_t lanes = (H == 64 ? ~0ull : ((1ull << H) - 1ull)) << (gshift + (gl & ~(H - 1)));
        return (__ballot(cut) & lanes) == 0ull;
    }

    // ---------------------------------------------------------- planes
    // Planes referenced by some vertex (group-uniform mask of P bits).
    __device__ __forceinline__ void used_planes(uint64_t (&u)[kPW]) const {
#pragma unroll
        for (int w = 0; w < kPW; ++w) u[w] = 0;
#pragma unroll
        for (int k = 0; k < kRows; ++k) {
            if (own(valid, k)) {
                const uint32_t t = S->vtri[gl + G * k];
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    const uint32_t p = tri_get<kB>(t, e);
#pragma unroll
                    for (int w = 0; w < kPW; ++w) u[w] |= (int)(p >> 6) == w ? (1ull << (p & 63)) : 0ull;
                }
            }
        }
#pragma unroll
        for (int w = 0; w < kPW; ++w) u[w] = gor64(u[w]);
    }

    // new index of a kept plane: referenced planes below it
    __device__ __forceinline__ uint32_t plane_rank(const uint64_t (&u)[kPW], uint32_t p) const {
        int r = 0;
#pragma unroll
        for (int w = 0; w < kPW; ++w) {
            const uint64_t m = (int)(p >> 6) > w ? ~0ull : ((int)(p >> 6) == w ? ((1ull << (p & 63)) - 1ull) : 0ull);
            r += __popcll(u[w] & m);
        }
        return (uint32_t)r;
    }

    // Drops unreferenced planes and renumbers the vertex triples (the planes
    // of a recorded test's vertex stay, renumbered too, until the record is
    // written out with the cell).
    __device__ __forceinline__ void garbage_collect() {
        uint64_t used[kPW];
        used_planes(used);
        const uint32_t nt = kExact ? 0u : S->ntent;
        if constexpr (!kExact) {
            for (uint32_t i = 0; i < nt; ++i) {
                const uint32_t t = S->tent[i][0];
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    const uint32_t p = tri_get<kB>(t, e);
#pragma unroll
                    for (int w = 0; w < kPW; ++w) used[w] |= (int)(p >> 6) == w ? (1ull << (p & 63)) : 0ull;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kPRows; ++k) {
            const int p = gl + G * k;
            const bool mv = p < np && mask_bit(used, p);
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
            float2 b = make_float2(0.f, 0.f);
            if (mv) {
                a = S->pl[p];
                b = S->plx[p];
            }
            // every lane has read slot p before any lane writes a slot below
            // G (k + 1): all later reads are at or above that
            if (mv) {
                const uint32_t dst = plane_rank(used, (uint32_t)p);
                S->pl[dst] = a;
                S->plx[dst] = b;
            }
        }
#pragma unroll
        for (int k = 0; k < kRows; ++k) {
            if (own(valid, k)) {
                const int v = gl + G * k;
                const uint32_t t = S->vtri[v];
                S->vtri[v] = tri_pack<kB>(plane_rank(used, tri_get<kB>(t, 0)), plane_rank(used, tri_get<kB>(t, 1)),
                                         plane_rank(used, tri_get<kB>(t, 2)));
            }
        }
        if constexpr (!kExact) {
            if ((uint32_t)gl < nt) {
                const uint32_t t = S->tent[gl][0];
                S->tent[gl][0] = tri_pack<kB>(plane_rank(used, tri_get<kB>(t, 0)),
                                              plane_rank(used, tri_get<kB>(t, 1)), plane_rank(used, tri_get<kB>(t, 2)));
            }
        }
        np = mask_popc(used);
    }

    __device__ __forceinline__ bool has_plane(int owner) const {
        if (!(bloom & bloom_bit((uint32_t)owner))) return false;
        bool hit = false;
#pragma unroll
        for (int k = 0; k < kPRows; ++k) {
            const int p = gl + G * k;
            hit |= p < np && __float_as_int(S->plx[p].y) == owner;
        }
        return gany(hit);
    }

    // ---------------------------------------------------------- clipping
    // Offers plane q (group-uniform).  Returns true when the cell changed.
    __device__ __forceinline__ bool clip(const Cand &q) {
        uint64_t rem[kW];
#pragma unroll
        for (int w = 0; w < kW; ++w) rem[w] = 0;
        uint32_t mine = 0, unsure = 0, lean = 0;
#pragma unroll
        for (int k = 0; k < kRows; ++k) {
            if (!own(valid, k)) continue;
            const int b = beyond32(gl + G * k, q);
            if (b > 0) mine |= 1u << k;
            if (b < 0) unsure |= 1u << k;
            if (b == -2) lean |= 1u << k;
        }
        if constexpr (kExact) {
            while (unsure) {  // one FP64 call site
                const int k = (int)__builtin_ctz(unsure);
                unsure &= unsure - 1u;
                if (recheck64(gl + G * k, q.owner)) mine |= 1u << k;
            }
        } else if (gany(unsure != 0u)) {
            // Undecided in FP32: take the sign of the FP32 value and record
            // the test; a pass after this tier decides it in FP64 and sends
            // the cell to the FP64-capable twin if the sign was wrong.  A
            // cell with more such tests than its record holds goes there now.
            const uint32_t n0 = S->ntent;
            int tot;
            uint32_t r = n0 + (uint32_t)gexcl(__popc(unsure), tot);
            if (n0 + (uint32_t)tot > (uint32_t)kTentMax) {
                overflow = true;
                bail = 1u;
                return false;
            }
            while (unsure) {
                const int k = (int)__builtin_ctz(unsure);
                unsure &= unsure - 1u;
                S->tent[r][0] = S->vtri[gl + G * k];
                S->tent[r][1] = (uint32_t)q.owner | (((lean >> k) & 1u) << 31);
                ++r;
            }
            S->ntent = n0 + (uint32_t)tot;
            mine |= lean;
        }
        if (!gany(mine != 0u)) return false;  // no vertex beyond: the plane does not cut
#pragma unroll
        for (int k = 0; k < kRows; ++k) rem[(G * k) >> 6] |= gballot((mine >> k) & 1u) << ((G * k) & 63);
        const int R = mask_popc(rem);
        if (R == mask_popc(valid)) {  // voronoi_convex_cell.cuh:430-435
            status = kEmptyCell;
#pragma unroll
            for (int w = 0; w < kW; ++w) valid[w] = 0;
            return true;
        }
        if (np >= P) garbage_collect();
        if (np >= P) {
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
        // removed; bit 3r + e of emask marks edge e of the lane's r-th
        // removed triangle.
        uint64_t emask = 0;
        int sx_ = 0, sy_ = 0;
        uint32_t tr = 0u;
        const bool spread = R <= G;
        if (spread) {
            // one removed triangle per lane (through tmp[]), each lane checks
            // its three edges against all of them
            int base = 0;
#pragma unroll
            for (int k = 0; k < kRows; ++k) {
                const uint64_t bits = row(rem, k);
                if ((bits >> gl) & 1ull) S->tmp[base + __popcll(bits & ((1ull << gl) - 1ull))] = S->vtri[gl + G * k];
                base += __popcll(bits);
            }
            const bool has_tri = gl < R;
            tr = has_tri ? S->tmp[gl] : 0u;
            const uint32_t t0 = tri_get<kB>(tr, 0), t1 = tri_get<kB>(tr, 1), t2 = tri_get<kB>(tr, 2);
            bool tw0 = false, tw1 = false, tw2 = false;
            if constexpr (G == 16 && P <= 64) {
                // The dual of a cell is a triangulated sphere: a pair of
                // planes lies in at most two triangles, which then share it
                // as an edge.  So the twin of edge (x, y) is removed iff
                // another removed triangle holds both x and y: per plane, OR
                // in the lanes whose removed triangle holds it (LDS
                // operations of one wave complete in order), intersect, and
                // clear the words again.
                const uint32_t bit = 1u << gl;
                if (has_tri) {
                    atomicOr(&S->pm[t0 >> 1], bit << ((t0 & 1u) << 4));
                    atomicOr(&S->pm[t1 >> 1], bit << ((t1 & 1u) << 4));
                    atomicOr(&S->pm[t2 >> 1], bit << ((t2 & 1u) << 4));
                    const uint32_t m0 = S->pm[t0 >> 1] >> ((t0 & 1u) << 4);
                    const uint32_t m1 = S->pm[t1 >> 1] >> ((t1 & 1u) << 4);
                    const uint32_t m2 = S->pm[t2 >> 1] >> ((t2 & 1u) << 4);
                    const uint32_t others = 0xFFFFu & ~bit;
                    tw0 = (m0 & m1 & others) != 0u;
                    tw1 = (m1 & m2 & others) != 0u;
                    tw2 = (m2 & m0 & others) != 0u;
                    S->pm[t0 >> 1] = 0u;
                    S->pm[t1 >> 1] = 0u;
                    S->pm[t2 >> 1] = 0u;
                }
            } else {
                for (int u = 0; u < R; ++u) {  // broadcast reads of the removed triangles
                    const uint32_t o = S->tmp[u];
                    tw0 |= tri_has_edge<kB>(o, t1, t0);
                    tw1 |= tri_has_edge<kB>(o, t2, t1);
                    tw2 |= tri_has_edge<kB>(o, t0, t2);
                }
            }
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const uint32_t x = e == 0 ? t0 : (e == 1 ? t1 : t2), y = e == 0 ? t1 : (e == 1 ? t2 : t0);
                const bool twin = e == 0 ? tw0 : (e == 1 ? tw1 : tw2);
                if (has_tri && !twin) {
                    emask |= 1ull << e;
                    sx_ += (int)(x * x + 1u);
                    sy_ += (int)(y * y + 1u);
                }
            }
        } else {
            boundary_rows(rem, emask, sx_, sy_);  // early clips of the initial box
        }
        int nb;
        int j = gexcl(__popcll(emask), nb);
        // every plane of the boundary cycle starts one edge and ends one
        if (nb < 3 || gsum(sx_) != gsum(sy_)) {
            status = kInconsistentBoundary;
            // (after a test taken by its FP32 sign the cell is rebuilt instead)
            if (gl == 0 && (kExact || S->ntent == 0u)) count(kStatInconsistent);
            return true;
        }
        uint64_t freem[kW];
#pragma unroll
        for (int w = 0; w < kW; ++w) freem[w] = ~(valid[w] & ~rem[w]);
        if (nb > mask_popc(freem)) {
            overflow = true;
            return false;
        }
        // compact the boundary edges into tmp[0, nb) (after every lane has
        // read its triangle above)
        while (emask) {
            const int bit = (int)__builtin_ctzll(emask);
            emask &= emask - 1ull;
            const int k = bit / 3, e = bit - 3 * k;
            const uint32_t t = spread ? tr : S->vtri[gl + G * k];
            S->tmp[j++] = tri_get<kB>(t, e) | (tri_get<kB>(t, e == 2 ? 0 : e + 1) << kB);
        }
        // one new vertex (q, x, y) per boundary edge, in the free slots
        // (lowest first), lane-strided
        bool bad = false;
        for (int e = gl; e < nb; e += G) {
            const uint32_t ed = S->tmp[e];
            bad |= !make_vertex(mask_select(freem, e), (uint32_t)qs, ed & kBMask, ed >> kB);
        }
        uint64_t added[kW];
        mask_first(freem, nb, added);
        if (gany(bad)) {
            if constexpr (kExact) {
                fix_vertices(added);
            } else {
                overflow = true;
                bail = 2u;
                return false;
            }
        }
#pragma unroll
        for (int w = 0; w < kW; ++w) valid[w] = (valid[w] & ~rem[w]) | added[w];
        dirty = true;
        return true;
    }

    // Boundary edges of clips that remove more triangles than the group has
    // lanes: each lane walks its own removed rows.
    __device__ __forceinline__ void boundary_rows(const uint64_t (&rem)[kW], uint64_t &emask, int &sx_, int &sy_) const {
#pragma unroll
        for (int k = 0; k < kRows; ++k) {
            if (!own(rem, k)) continue;
            const uint32_t t = S->vtri[gl + G * k];
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const uint32_t x = tri_get<kB>(t, e), y = tri_get<kB>(t, e == 2 ? 0 : e + 1);
                bool twin = false;
#pragma unroll
                for (int w = 0; w < kW; ++w) {
                    uint64_t m = twin ? 0ull : rem[w];
                    while (m && !twin) {
                        const int u = 64 * w + (int)__builtin_ctzll(m);
                        m &= m - 1ull;
                        twin = tri_has_edge<kB>(S->vtri[u], y, x);
                    }
                }
                if (!twin) {
                    emask |= 1ull << (3 * k + e);
                    sx_ += (int)(x * x + 1u);
                    sy_ += (int)(y * y + 1u);
                }
            }
        }
    }

    // Prepares lane's candidate j; false when it cannot cut the cell.
    // maybe_guess: j may be one of the guesses already offered (only those
    // can already own a plane of the cell): every guess lies within the
    // guess radius, up to the few ulps that its 3e-5 widening covers.
    __device__ __forceinline__ bool prepare(uint32_t j, Cand &c, bool &emptied, bool &maybe_guess) const {
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
        if (a2 < guess_r2) return false;  // offered by the pre-clip; the cell has only shrunk since
        maybe_guess = guess_r2 == 0.0f || a2 < guess_r2 * 1.00003f || (bloom_all && (bloom & bloom_bit(j)));
        c.H = a2 - dw;
        c.pH = a2 + fabsf(dw);
        c.pa1 = fabsf(c.ax) + fabsf(c.ay) + fabsf(c.az);
        return !too_far(c);
    }

    // Offers a batch of up to G candidates (lane-specific j when `has`).
    __device__ __forceinline__ void offer_batch(bool has, uint32_t j) {
        Cand c;
        bool emptied = false, maybe_guess = true;
        const bool live_c = has && prepare(j, c, emptied, maybe_guess);
        if (gany(emptied)) {
            status = kEmptyCell;
#pragma unroll
            for (int w = 0; w < kW; ++w) valid[w] = 0;
            return;
        }
        const uint64_t skipped = gballot(has && !live_c);
        if (gl == 0 && skipped) count(kStatSkipped, (uint32_t)__popcll(skipped));
        uint64_t todo = gballot(live_c);
        if (gl == 0 && todo) count(kStatCandidates, (uint32_t)__popcll(todo));
        if (live_c) {
            C->aH[gl] = make_float4(c.ax, c.ay, c.az, c.H);
            // the owner's sign bit marks a candidate that may already own a plane
            C->po[gl] = make_float2(c.pH, __int_as_float(maybe_guess ? (c.owner | INT32_MIN) : c.owner));
        }
        float hkey = live_c ? c.H : FLT_MAX;
        while (todo) {
            // nearest plane first (smallest H = power distance of the point):
            // the cell shrinks fastest and fewer clips are undone later
            const float hmin = gmin(hkey);
            const uint64_t at = gballot(live_c && hkey == hmin) & todo;
            const int b = (int)__builtin_ctzll(at ? at : todo);
            todo &= ~(1ull << b);
            if (gl == b) hkey = FLT_MAX;
            Cand q;
            const float4 aH = C->aH[b];
            const float2 po = C->po[b];
            q.ax = aH.x;
            q.ay = aH.y;
            q.az = aH.z;
            q.H = aH.w;
            q.pH = po.x;
            q.pa1 = fabsf(q.ax) + fabsf(q.ay) + fabsf(q.az);
            q.c2 = fmaf(kG32, q.pH, kTiny32);
            const int ow = __float_as_int(po.y);
            q.owner = ow & INT32_MAX;
            if (ow < 0 && bloom && has_plane(q.owner)) continue;
            if (clip(q)) {
                if (gl == 0) count(kStatClips);
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

    // The initial cell: the box, or (cube_h > 0, a predicted half-width in
    // doubled coordinates) the box cut by the cube |y_k| <= cube_h, whose
    // planes carry owner codes -7..-12.  A cell whose final vertices touch a
    // cube plane, or that empties, is rebuilt without the cube.
    __device__ __forceinline__ void init_box(float cube_h = 0.0f) {
        bool from_cube = false;
        if (gl < 6) {
            const int axis = gl >> 1, side = gl & 1;
            // (a lane-dependent choice among the members sx, sy, sz would be
            // compiled into a choice of addresses, moving them to scratch)
            const float4 sp = pts[seed];
            const float s = axis == 0 ? sp.x : (axis == 1 ? sp.y : sp.z);
            float h = side ? 2.0f * (hdr->box_hi[axis] - s) : 2.0f * (s - hdr->box_lo[axis]);
            from_cube = cube_h > 0.0f && cube_h < h;
            if (from_cube) h = cube_h;
            float4 pl = make_float4(0.f, 0.f, 0.f, h);
            const float sg = side ? 1.0f : -1.0f;
            if (axis == 0) pl.x = sg;
            if (axis == 1) pl.y = sg;
            if (axis == 2) pl.z = sg;
            S->pl[gl] = pl;
            S->plx[gl] = make_float2(fabsf(h), __int_as_float((from_cube ? -7 : -1) - gl));
        }
        cube = gany(from_cube);
        np = 6;
        // the 8 corners go to slots 0..7 (lanes 0..7 of row 0)
        if (gl < 8) {
            const int cx = gl >> 2 & 1, cy = gl >> 1 & 1, cz = gl & 1;
            const uint32_t X = cx, Y = 2 + cy, Z = 4 + cz;
            const int sgn = (cx ? 1 : -1) * (cy ? 1 : -1) * (cz ? 1 : -1);
            // box corners: D = +-1 exactly, always resolved in FP32
            if (sgn > 0) make_vertex(gl, X, Y, Z);
            else make_vertex(gl, X, Z, Y);
        }
#pragma unroll
        for (int w = 0; w < kW; ++w) valid[w] = w == 0 ? 0xFFull : 0ull;
        bloom = 0;
        update_bounds();
        dirty = false;
    }

    // ---------------------------------------------------------- traversal
    __device__ __forceinline__ void offer_leaf(uint32_t ref) {
        if (gl == 0) count(kStatLeaves);
        const uint32_t f = leaf_first(ref), c = leaf_count(ref);
        for (uint32_t o = 0; o < c && status == kSuccess && !overflow; o += G) {
            const uint32_t j = f + o + (uint32_t)gl;
            offer_batch(o + (uint32_t)gl < c && j != seed, j);
        }
    }

    // ---------------------------------------------------------- wait list
    // Depth-first wait list (a stack) shared by the cells of one wave, in
    // registers: entry t sits in slot t / 64 of lane t % 64.  Entries are
    // parent << 1 | child.
    __device__ __forceinline__ void list_push(int t, uint32_t e) {
        const bool me = lane64 == (t & (kWave - 1));
        const int sl = t / kWave;
#pragma unroll
        for (int j = 0; j < kLS; ++j) lr[j] = (me && j == sl) ? e : lr[j];
    }
    __device__ __forceinline__ uint32_t list_pop(int &sz) {
        const int t = --sz;
        uint32_t r = 0u;
#pragma unroll
        for (int j = 0; j < kLS; ++j) r = j == t / kWave ? lr[j] : r;
        return rl(r, t & (kWave - 1));
    }

    // Wave-wide: does any cell of the wave still need this box?  key: the
    // smallest bound over the cells that need it; mine: this cell needs it.
    __device__ __forceinline__ bool packet_needs(float4 lo, float4 hi, float &key, bool &mine) const {
        float k = FLT_MAX;
        const bool need = live() && !node_prunable<false>(lo, hi, k);
        mine = need;
        const bool any = __ballot(gl == 0 && need) != 0ull;
        float kk = need ? k : FLT_MAX;
        if constexpr (G == 16) kk = fminf(fminf(rl(kk, 0), rl(kk, 16)), fminf(rl(kk, 32), rl(kk, 48)));
        key = uniform(kk);
        return any;
    }

    // Both children of a node at once: the low half of each group tests
    // child 0, the high half child 1.
    __device__ __forceinline__ void packet_needs2(const Node &nd, bool &n0, bool &n1, float &k0, float &k1, bool &m0,
                                                  bool &m1) const {
        constexpr int H = G / 2;
        const bool second = gl >= H;
        const float4 lo = second ? nd.lo1 : nd.lo0, hi = second ? nd.hi1 : nd.hi0;
        float k = FLT_MAX;
        const bool need = live() && !node_prunable<true>(lo, hi, k);
        const uint64_t b = __ballot(need);
        m0 = ((b >> gshift) & 1ull) != 0ull;
        m1 = ((b >> (gshift + H)) & 1ull) != 0ull;
        constexpr uint64_t kLead = G == 16 ? 0x0001000100010001ull : 1ull;  // lane 0 of each group
        n0 = (b & kLead) != 0ull;
        n1 = (b & (kLead << H)) != 0ull;
        const float kk = need ? k : FLT_MAX;
        if constexpr (G == 16) {
            k0 = uniform(fminf(fminf(rl(kk, 0), rl(kk, 16)), fminf(rl(kk, 32), rl(kk, 48))));
            k1 = uniform(fminf(fminf(rl(kk, H), rl(kk, 16 + H)), fminf(rl(kk, 32 + H), rl(kk, 48 + H))));
        } else {
            k0 = uniform(rl(kk, 0));
            k1 = uniform(rl(kk, H));
        }
    }

    // Depth-first packet traversal: the cells of a wave (Morton neighbours)
    // walk the hierarchy together.  A box is expanded while any of them may
    // still be cut by a point in it; of two needed children the one with the
    // smaller key (the smallest lower bound of H over the cells that need
    // it) comes first and the other waits on a wave-wide stack, re-tested
    // when it comes out.  A leaf goes only to the cells whose own test kept
    // its box (the others could not be cut by its points).  The reference
    // pops the smallest key instead (voronoi_bvh.cuh:246-283); after the kNN
    // pre-clip the order changes little of the work and the stack needs no
    // arg-min, and the diagram does not depend on it.  A full stack sends the
    // wave's cells to the next tier instead of dropping entries.
    __device__ __forceinline__ void traverse() {
        uint32_t ref = uniform(root_ref);
        int sz = 0;
        bool mine = true;  // this cell needs box ref
        while (true) {
            bool have = false;
            if (is_leaf(ref)) {
                if (live() && mine) offer_leaf(ref);
            } else {
                if (lane64 == 0) count(kStatNodes);
                const Node nd = nodes[ref];
                float k0, k1;
                bool m0, m1, n0, n1;
                if constexpr (G == 16 || V <= 512) {  // tiers 0 and 1
                    packet_needs2(nd, n0, n1, k0, k1, m0, m1);
                } else {
                    n0 = packet_needs(nd.lo0, nd.hi0, k0, m0);
                    n1 = packet_needs(nd.lo1, nd.hi1, k1, m1);
                }
                const uint32_t r0 = uniform(__float_as_uint(nd.hi0.w)), r1 = uniform(__float_as_uint(nd.hi1.w));
                if (n0 || n1) {
                    const bool near0 = !n1 || (n0 && k0 <= k1);
                    if (n0 && n1) {  // the farther child waits
                        if (sz >= L) {
                            overflow = true;
                            return;
                        }
                        list_push(sz++, (ref << 1) | (near0 ? 1u : 0u));
                    }
                    ref = near0 ? r0 : r1;
                    mine = near0 ? m0 : m1;
                    have = true;
                }
            }
            while (!have) {
                if (sz == 0) return;
                const uint32_t e = list_pop(sz);
                const uint32_t parent = e >> 1, c = e & 1u;
                const Node nd = nodes[parent];
                float k;
                const float4 lo = c ? nd.lo1 : nd.lo0;
                const float4 hi = c ? nd.hi1 : nd.hi0;
                if (packet_needs(lo, hi, k, mine)) {
                    ref = uniform(__float_as_uint(hi.w));
                    have = true;
                }
            }
        }
    }

    // ---------------------------------------------------------- output
    // Row = sorted original ids of the owners of referenced planes
    // (voronoi_ultra.cu:125-138).  stage_row() returns the row length F
    // (group-uniform; the ids are staged unsorted in tmp[0, F)), the caller
    // takes F entries of the row pool, and write_row() writes the sorted row
    // there.  With duplicate points a row also holds every exact duplicate of
    // an owner (a second copy of a plane removes no vertex, so at most one
    // plane per group of duplicates is referenced); lane 0 builds such rows.
    __device__ __forceinline__ int stage_row(const CellParams &prm) const {
        if (status == kEmptyCell) return 0;
        uint64_t used[kPW];
        used_planes(used);
        if (hdr->has_duplicates != 0u) {
            int F = 0;
            if (gl == 0) {
                for (int p = 0; p < np; ++p) {
                    if (!mask_bit(used, p)) continue;
                    const int owner = __float_as_int(S->plx[p].y);
                    if (owner < 0) continue;
                    for (uint32_t m = prm.dup_head[owner]; m != 0xFFFFFFFFu; m = prm.dup_next[m]) F++;
                }
            }
            return gbi(F, 0);
        }
        int32_t ids[kPRows];
        int mine = 0;
#pragma unroll
        for (int k = 0; k < kPRows; ++k) {
            const int p = gl + G * k;
            ids[k] = -1;
            if (p < np && mask_bit(used, p)) {
                const int owner = __float_as_int(S->plx[p].y);
                if (owner >= 0) {
                    ids[k] = (int32_t)perm[owner];
                    mine++;
                }
            }
        }
        int F;
        int at = gexcl(mine, F);
#pragma unroll
        for (int k = 0; k < kPRows; ++k)
            if (ids[k] >= 0) S->tmp[at++] = (uint32_t)ids[k];
        return F;
    }

    // Writes the staged row (sorted) to row_out and the cell's count, status
    // and extent.
    __device__ __forceinline__ void write_row(const CellParams &prm, uint32_t orig, int32_t *row_out, int F) const {
        if (F > 0) {
            if (hdr->has_duplicates != 0u) {
                // insertion into the sorted row, dropping repeats
                int len = 0;
                uint64_t used[kPW];
                used_planes(used);
                if (gl == 0) {
                    for (int p = 0; p < np; ++p) {
                        if (!mask_bit(used, p)) continue;
                        const int owner = __float_as_int(S->plx[p].y);
                        if (owner < 0) continue;
                        for (uint32_t m = prm.dup_head[owner]; m != 0xFFFFFFFFu; m = prm.dup_next[m]) {
                            const int32_t id = (int32_t)perm[m];
                            int k = len;
                            while (k > 0 && row_out[k - 1] > id) k--;
                            if (k > 0 && row_out[k - 1] == id) continue;
                            if (len >= F) break;  // (F bounds the distinct ids)
                            for (int t = len; t > k; --t) row_out[t] = row_out[t - 1];
                            row_out[k] = id;
                            len++;
                        }
                    }
                }
                F = gbi(len, 0);
            } else {
                // rank of each id among the row (ids are distinct)
                for (int t = gl; t < F; t += G) {
                    const int32_t x = (int32_t)S->tmp[t];
                    int rank = 0;
                    for (int u = 0; u < F; ++u) rank += (int32_t)S->tmp[u] < x;
                    row_out[rank] = x;
                }
            }
        }
        if (gl == 0) {
            prm.counts[orig] = F;
            prm.status[orig] = status;
            if (prm.extent)
                prm.extent[orig] = status == kEmptyCell ? 0.0f
                                                       : 0.5f * fmaxf(fmaxf(fmaxf(-L_[0], U_[0]), fmaxf(-L_[1], U_[1])),
                                                                      fmaxf(-L_[2], U_[2]));
        }
    }

    // Does a final vertex lie on a plane of the predicted cube (group op)?
    __device__ __forceinline__ bool cube_touched() const {
        uint64_t used[kPW];
        used_planes(used);
        bool hit = false;
#pragma unroll
        for (int k = 0; k < kPRows; ++k) {
            const int p = gl + G * k;
            if (p < np && mask_bit(used, p)) hit |= __float_as_int(S->plx[p].y) <= -7;
        }
        return gany(hit);
    }

    __device__ __forceinline__ void start(uint32_t cell, const CellParams &prm) {
        seed = cell;
        if constexpr (kExact) {
            if (gl == 0) S->exact_events = 0u;
        } else {
            S->ntent = 0u;
        }
        const float4 s = pts[cell];
        sx = s.x, sy = s.y, sz = s.z, sw = s.w;
        status = kSuccess;
        overflow = false;
        bail = 0u;
        const uint32_t orig = perm[cell];
        init_box((!kExact && prm.init_h) ? 2.0f * prm.init_h[orig] : 0.0f);
        in_guess = true;
        bloom_all = false;
        guess_r2 = 0.0f;
        if (!prm.guesses && prm.seed_window > 0) {
            // Morton neighbours first: the cheap analogue of clipping by the
            // k nearest neighbours (voronoi_ultra.cu:70-85); order does not
            // change the result, only how fast the cell shrinks
            for (int o = 0; o < prm.seed_window && status == kSuccess && !overflow; o += G) {
                const int i = o + gl;
                const int64_t j = (int64_t)cell - prm.seed_window / 2 + i + (i >= prm.seed_window / 2 ? 1 : 0);
                offer_batch(i < prm.seed_window && j >= 0 && j < prm.n, (uint32_t)(j < 0 ? 0 : j));
            }
        }
        if (prm.guesses) {
            // rows by original id (CSR), or Morton-ordered rows of guess_stride
            // sorted ids (the library's own kNN pre-clip)
            const bool strided = prm.guess_stride > 0;
            const int64_t g0 = strided ? (int64_t)cell * prm.guess_stride : (int64_t)prm.guess_offsets[orig];
            const int64_t g1 = strided ? g0 + prm.guess_stride : (int64_t)prm.guess_offsets[orig + 1];
            for (int64_t g = g0; g < g1 && status == kSuccess && !overflow; g += G) {
                bool has = g + gl < g1;
                uint32_t j = 0;
                if (has) {
                    const int32_t gid = prm.guesses[g + gl];
                    has = gid >= 0 && (int64_t)gid < prm.n;  // voronoi_ultra.cu:76-78 skips negative ids
                    if (has) j = (strided || prm.guesses_sorted) ? (uint32_t)gid : inv[gid];
                    has = has && j != cell;
                }
                offer_batch(has, j);
            }
            // The row holds the guess_stride nearest points (kNN order, ties
            // and all); a point off the row has squared distance >= the
            // row's last, up to the few ulps of the two FP32 evaluations of
            // |p - s|^2, which the 1e-5 relative margin covers.
            if (strided && prm.guess_d2) guess_r2 = prm.guess_d2[cell] * (1.0f - 1e-5f);
        }
        in_guess = false;
    }

    // The owners of the planes held, in plane order (point ids in the index's
    // order, negative for the box planes, -1 past them), for the large tier
    // (group op).
    __device__ __forceinline__ void store_owners(int32_t *out) const {
#pragma unroll
        for (int k = 0; k < kPRows; ++k) {
            const int p = gl + G * k;
            if (p < kPromoPlanes) out[p] = p < np ? __float_as_int(S->plx[p].y) : -1;
        }
    }

    // Offers the planes a smaller tier held when it gave the cell up (point
    // ids, negative for none): the cell starts from nearly all of its faces.  They
    // join the Bloom mask, and every later candidate that hits the mask is
    // checked against the planes held, so none is clipped twice.
    __device__ __forceinline__ void offer_owners(const int32_t *ids, int m) {
        if (status != kSuccess || overflow) return;
        in_guess = true;
        bloom_all = true;
        for (int o = 0; o < m && status == kSuccess && !overflow; o += G) {
            const int32_t id = o + gl < m ? ids[o + gl] : -1;
            offer_batch(id >= 0 && (uint32_t)id != seed, (uint32_t)(id < 0 ? 0 : id));
        }
        in_guess = false;
    }

    __device__ __forceinline__ void setup(const CellParams &prm, GroupLDS<V, P> *s, CandLDS<G> *c) {
        const int lane = threadIdx.x & (kWave - 1);
        const int grp = lane / G;
        S = s;
        C = c;
#pragma unroll
        for (int j = 0; j < kLS; ++j) lr[j] = 0u;
        gl = lane & (G - 1);
        lane64 = lane;
        gshift = grp * G;
        pts = prm.pts;
        perm = prm.perm;
        inv = prm.inv;
        nodes = prm.nodes;
        set_cloud(prm.hdr);
        if (gl < kNumCtr) S->ctr[gl] = 0u;
        for (int w = gl; w < (int)(sizeof(S->pm) / 4); w += G) S->pm[w] = 0u;
    }

    // The cloud the next cells belong to (its clipping box and subtree root).
    __device__ __forceinline__ void set_cloud(const IndexHeader *h) {
        hdr = h;
        root_ref = h->root_ref;
    }

    __device__ __forceinline__ void flush(const CellParams &prm, uint32_t cells) const {
        uint32_t vals[kNumCtr];
#pragma unroll
        for (int k = 0; k < kNumCtr; ++k) vals[k] = gl == 0 ? S->ctr[k] : 0u;
        vals[kStatCells] = cells;
#pragma unroll
        for (int k = 0; k < kNumCtr; ++k) {
            const uint32_t s = wave_sum(vals[k]);
            if (lane64 == 0 && s) atomicAdd(&prm.stats[ctr_stat_slot(k)], (unsigned long long)s);
        }
    }
};

// ------------------------------------------------------------------ batches
// Largest c < n_clouds with off[c] <= x (off non-decreasing): the cloud of a
// sorted cell (off = cloud cell offsets) or of a work quad (off = quad
// offsets; empty clouds are skipped because the search takes the last of
// equal offsets).
__device__ __forceinline__ int batch_find(const int64_t *off, int n_clouds, int64_t x) {
    int lo = 0, hi = n_clouds - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------ tiers
constexpr int kT0Lanes = 16, kT0Verts = 64, kT0Planes = 48, kT0List = 128, kT0Threads = 256;
// (a depth-first stack holds at most one entry per level of the hierarchy,
// fewer than 64 levels: 128 entries never fill)
constexpr int kTmLanes = 64, kTmVerts = 192, kTmPlanes = 128, kTmList = 128, kTmThreads = 128;
constexpr int kT1Lanes = 64, kT1Verts = 512, kT1Planes = 255, kT1List = 128, kT1Threads = 128;
constexpr int kT2Lanes = 64, kT2Verts = 1024, kT2Planes = 1023, kT2List = 128, kT2Threads = 64;
using Tier0 = Group<kT0Lanes, kT0Verts, kT0Planes, kT0List, false>;
using Tier0x = Group<kT0Lanes, kT0Verts, kT0Planes, kT0List, true>;
using TierM = Group<kTmLanes, kTmVerts, kTmPlanes, kTmList, true>;
using Tier1 = Group<kT1Lanes, kT1Verts, kT1Planes, kT1List, true>;
using Tier2 = Group<kT2Lanes, kT2Verts, kT2Planes, kT2List, true>;

// Tier 0: a wave holds four cells per round and traverses for them
// together.  The FP32-only instance takes Morton-consecutive cells of the
// range from a global counter; a cell it gives up on (`bail`) is flagged for
// the FP64-capable instance (kExact), whose waves scan the flags in windows
// of 64 Morton-consecutive cells and take the flagged cells four at a time
// (neighbours again, so the shared traversal stays coherent).  A cell that
// outgrows the tier is queued for tier m.
// The address of the block's copy of the launch's parameters (in LDS: one
// copy per block, where a kernel argument whose address is taken would be
// copied per lane to scratch), opaque to the optimiser: fields read through
// it are loaded where they are used, not held in registers across the
// traversal.
__device__ __forceinline__ const CellParams *reread(const CellParams *p) {
    asm volatile("" : "+s"(p));
    return p;
}

// A row that does not fit the row pool: provisional status 1 and an empty
// row, counted, with the pool size the launch needs (the host runs the
// launch again with a pool that large).
__device__ __forceinline__ void pool_short(const CellParams &P, uint32_t cell, uint32_t orig, uint64_t need) {
    P.row_offsets[cell] = 0;
    P.counts[orig] = 0;
    P.status[orig] = kPlaneOverflow;
    atomicAdd(&P.stats[kStatPoolShort], 1ull);
    atomicMax(&P.stats[kStatPoolDemand], (unsigned long long)need);
}
// A cell beyond the largest tier: the reference's capacity status with an
// empty row (counted, never silent).
__device__ __forceinline__ void truncate_cell(const CellParams &P, uint32_t orig) {
    P.counts[orig] = 0;
    P.status[orig] = kPlaneOverflow;
    atomicAdd(&P.stats[kStatTruncated], 1ull);
}
// Counts the removal tests of a cell that expansion arithmetic decided, and
// records the cell's original id in the first free log slot.
__device__ __forceinline__ void log_exact(const CellParams &P, uint32_t orig, uint32_t events) {
    const unsigned long long k = atomicAdd(&P.stats[kStatExact], (unsigned long long)events);
    if (k < (unsigned long long)kExactLogSlots) P.stats[kStatExactLog + k] = (unsigned long long)orig + 1ull;
}
// Takes n entries of the row pool (lane 0 asks; every lane gets the base).
__device__ __forceinline__ uint64_t pool_take(const CellParams &P, uint32_t n) {
    unsigned long long b = 0;
    if ((threadIdx.x & (kWave - 1)) == 0 && n > 0u) b = atomicAdd(P.pool_top, (unsigned long long)n);
    return ((uint64_t)rl((uint32_t)(b >> 32), 0) << 32) | (uint64_t)rl((uint32_t)b, 0);
}

// kFlagged: the instance that takes the cells flagged by the first one
// (else the cells of the range, or the batch's quads).
template <typename T, bool kFlagged>
__device__ __forceinline__ void tier0_body(const CellParams &prm) {
    constexpr bool kExactTier = kFlagged;
    constexpr int kCellsPerWave = kWave / kT0Lanes;
    __shared__ GroupLDS<kT0Verts, kT0Planes> lds[kT0Threads / kT0Lanes];
    __shared__ CandLDS<kT0Lanes> cands[kT0Threads / kT0Lanes];
    __shared__ CellParams sprm;
    // each wave's unused part of the row pool: next entry and entries left
    __shared__ unsigned long long chunk_at[kT0Threads / kWave];
    __shared__ uint32_t chunk_left[kT0Threads / kWave];
    if (threadIdx.x == 0) sprm = prm;
    if (threadIdx.x < kT0Threads / kWave) {
        chunk_at[threadIdx.x] = 0ull;
        chunk_left[threadIdx.x] = 0u;
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int grp = lane / kT0Lanes;
    if (kExactTier && prm.counters[5] == 0u) return;  // no cell was flagged
    if (kExactTier && prm.counters[5] <= kFewFlagged) {
        // a few flagged cells: queued for tier m, with no plane owners (the
        // flags read in a grid-stride loop: no work counter to contend for)
        const int64_t count = prm.end - prm.begin;
        const int64_t step = (int64_t)gridDim.x * kT0Threads;
        for (int64_t c = (int64_t)blockIdx.x * kT0Threads + threadIdx.x; c < count; c += step) {
            if (prm.exact_list[c] != 0u) {
                const uint32_t slot = atomicAdd(&prm.counters[1], 1u);
                const uint32_t cell = (uint32_t)(prm.begin + c);
                if (slot < prm.queue_cap) prm.mid_list[slot] = cell;
                else truncate_cell(prm, prm.perm[cell]);
                if (slot < prm.promo_cap)
                    for (int p = 0; p < kPromoPlanes; ++p) prm.promo[(size_t)slot * kPromoPlanes + p] = -1;
            }
        }
        return;
    }
    T g;
    g.setup(prm, &lds[(threadIdx.x / kWave) * kCellsPerWave + grp], &cands[(threadIdx.x / kWave) * kCellsPerWave + grp]);
    uint32_t cells = 0;
    const int64_t count = prm.end - prm.begin;
    // one round: cell begin + c for an active group
    auto round = [&](bool active, int64_t c) {
        const uint64_t t0 = clock64();
        g.active = active;
        const uint32_t cell = (uint32_t)(prm.begin + (active ? c : 0));
        if (active) g.start(cell, *reread(&sprm));
        g.traverse();
        // the launch's fields again from the argument segment (scalar loads)
        // rather than kept in registers across the traversal
        const CellParams &P = *reread(&sprm);
        const uint32_t orig = P.perm[cell];
        if (kExactTier && active && g.gl == 0 && g.S->exact_events) log_exact(P, orig, g.S->exact_events);
        // a cell built inside a predicted cube stands only if the cube cut nothing
        if (active && !kExactTier && !g.overflow && g.cube && (g.status != kSuccess || g.cube_touched()))
            g.overflow = true, g.bail = 3u;
        if constexpr (!kExactTier) {
            // the removal tests taken by their FP32 sign go to the record for
            // the FP64 check after this tier, with the cell; a cell that ended
            // inconsistent after such a test, or that the record cannot hold,
            // goes to the FP64 twin instead
            const uint32_t nt = active && !g.overflow ? g.S->ntent : 0u;
            if (nt > 0u && g.status == kInconsistentBoundary) {
                g.overflow = true, g.bail = 1u;
            } else if (nt > 0u) {
                // (the base comes from the group's lane 0 through a shuffle:
                // a store and load of one LDS word by different lanes are not
                // ordered for the compiler)
                uint32_t base = 0u;
                if (g.gl == 0) base = atomicAdd(&P.counters[7], nt);
                base = (uint32_t)g.gbi((int)base, 0);
                const bool fits = (uint64_t)base + nt <= (uint64_t)P.tent_cap;
                if ((uint32_t)g.gl < nt && (uint64_t)base + g.gl < (uint64_t)P.tent_cap) {
                    uint32_t *rec = P.tent + 5 * ((uint64_t)base + g.gl);
                    const uint32_t t = g.S->tent[g.gl][0];
                    rec[0] = fits ? cell : kNoCell;
                    rec[1] = __float_as_uint(g.S->plx[tri_get<T::kB>(t, 0)].y);
                    rec[2] = __float_as_uint(g.S->plx[tri_get<T::kB>(t, 1)].y);
                    rec[3] = __float_as_uint(g.S->plx[tri_get<T::kB>(t, 2)].y);
                    rec[4] = g.S->tent[g.gl][1];
                }
                if (!fits) g.overflow = true, g.bail = 1u;
            }
        }
        const bool emit = active && !g.overflow;
        const int F = emit ? g.stage_row(P) : 0;
        // the wave's four rows, back to back in the wave's current pool chunk
        const int f0 = rl(F, 0), f1 = rl(F, 16), f2 = rl(F, 32), f3 = rl(F, 48);
        // (the chunk's position lives in LDS, not in registers across the traversal)
        const uint32_t need = (uint32_t)(f0 + f1 + f2 + f3);
        const int wv = threadIdx.x / kWave;
        uint32_t left = uniform(chunk_left[wv]);
        uint64_t at;
        if (need > left) {
            left = need > P.pool_chunk ? need : P.pool_chunk;
            at = pool_take(P, left);
        } else {
            const unsigned long long a = chunk_at[wv];
            at = ((uint64_t)uniform((uint32_t)(a >> 32)) << 32) | (uint64_t)uniform((uint32_t)a);
        }
        const uint64_t base = at + (uint64_t)(grp > 0 ? f0 : 0) + (uint64_t)(grp > 1 ? f1 : 0) +
                              (uint64_t)(grp > 2 ? f2 : 0);
        if (lane == 0) {
            chunk_at[wv] = at + need;
            chunk_left[wv] = left - need;
        }
        if (emit) {
            if (base + (uint64_t)F <= P.rows_elems) {
                if (g.gl == 0) P.row_offsets[cell] = (int64_t)base;
                g.write_row(P, orig, P.rows + base, F);
            } else if (g.gl == 0) {
                pool_short(P, cell, orig, base + (uint64_t)F);
            }
        }
        uint32_t slot = 0u;
        if (!emit && active && g.gl == 0) {
            if (!kExactTier && g.bail) {  // flag for the FP64-capable instance
                if (g.bail < 3u) g.count(g.bail == 1u ? kCtrBailTest : kCtrBailVertex);
                P.exact_list[c] = 1u;
                atomicAdd(&P.counters[5], 1u);
            } else {  // queue for the large-capacity tier
                g.count(kStatPromoted);
                slot = atomicAdd(&P.counters[1], 1u);
                if (slot < P.queue_cap) P.mid_list[slot] = cell;
                else truncate_cell(P, orig);
            }
        }
        // with the promotion, the owners of the planes held (the large tier
        // starts from them); the slot comes from the group's lane 0
        if (active && !emit && (kExactTier || !g.bail)) {
            slot = (uint32_t)g.gbi((int)slot, 0);
            if (slot < P.promo_cap) g.store_owners(P.promo + (size_t)slot * kPromoPlanes);
        }
        if (active && P.cycles && g.gl == 0) P.cycles[orig] = (uint32_t)min(clock64() - t0, (uint64_t)0x7fffffff);
        if (active && g.gl == 0 && !g.bail) cells++;
    };
    if constexpr (!kExactTier) {
        if (prm.clouds == nullptr) {
            while (true) {
                uint32_t c0 = 0;
                if (lane == 0) c0 = atomicAdd(&prm.counters[0], (uint32_t)kCellsPerWave);
                c0 = uniform(c0);
                if ((int64_t)c0 >= count) break;
                round((int64_t)c0 + grp < count, (int64_t)c0 + grp);
            }
        } else {
            // a batch: one quad (four consecutive cells of one cloud) per round,
            // so the cells of a wave share their cloud's subtree
            const int64_t nq = prm.quads[prm.n_clouds];
            while (true) {
                uint32_t q0 = 0;
                if (lane == 0) q0 = atomicAdd(&prm.counters[0], 1u);
                const int64_t q = (int64_t)uniform(q0);
                if (q >= nq) break;
                const int cl = uniform(batch_find(prm.quads, prm.n_clouds, q));
                g.set_cloud(&prm.clouds[cl]);
                const int64_t lo = max(prm.cloud_cells[cl], prm.begin), hi = min(prm.cloud_cells[cl + 1], prm.end);
                const int64_t cell = lo + kCellsPerWave * (q - prm.quads[cl]) + grp;
                round(cell < hi, cell - prm.begin);
            }
        }
    } else {
        while (true) {
            uint32_t w = 0;
            if (lane == 0) w = atomicAdd(&prm.counters[6], 1u);
            const int64_t base = (int64_t)uniform(w) * kWave;
            if (base >= count) break;
            uint64_t bits = __ballot(base + lane < count && prm.exact_list[base + lane] != 0u);
            // in a batch the cells of a round must share a cloud
            const int lc = (prm.clouds != nullptr && base + lane < count)
                               ? batch_find(prm.cloud_cells, prm.n_clouds, prm.begin + base + lane) : 0;
            while (bits) {
                uint64_t take = bits;
                if (prm.clouds != nullptr) {
                    const int cl = uniform(rl(lc, (int)__builtin_ctzll(bits)));
                    g.set_cloud(&prm.clouds[cl]);
                    take = bits & __ballot(lc == cl);
                }
                uint64_t b = take;  // the grp-th flagged cell of the round
                for (int t = 0; t < grp; ++t) b &= b - 1ull;
                uint64_t used = take;
                for (int t = 0; t < kCellsPerWave; ++t) used &= used - 1ull;
                bits &= ~(take & ~used);
                round(b != 0ull, base + (b ? (int64_t)__builtin_ctzll(b) : 0));
            }
        }
    }
    g.flush(prm, cells);
}

__global__ void __launch_bounds__(kT0Threads) __attribute__((amdgpu_waves_per_eu(3, 3))) cell_group_kernel(const CellParams prm) {
    tier0_body<Tier0, false>(prm);
}
__global__ void __launch_bounds__(kT0Threads) cell_group_exact_kernel(const CellParams prm) {
    tier0_body<Tier0x, true>(prm);
}
// The same at 3 waves per SIMD, as tier 0, for batches of clouds, where many
// cells come here (a fifth of a LiDAR batch): the third wave gains more than
// the 76 registers it spills cost.  A single cloud sends few cells here, and
// every persistent wave writes its spills once, so it keeps the uncapped one.
__global__ void __launch_bounds__(kT0Threads) __attribute__((amdgpu_waves_per_eu(3, 3)))
cell_group_exact_batch_kernel(const CellParams prm) {
    tier0_body<Tier0x, true>(prm);
}

// Decides in FP64 (where FP64 cannot, exactly) the removal tests tier 0 took
// by their FP32 sign, and flags the cell of every test whose sign was wrong
// for the FP64-capable twin, which rebuilds the cell (its row, count and
// status replace tier 0's).  A test whose vertex lies on a plane of a
// predicted cube (owner codes -7..-12, not a plane of the diagram) is not
// decided: its cell is flagged as well, and the twin builds it without the
// cube.  One thread per recorded test.
__global__ void __launch_bounds__(256) cell_tentative_check_kernel(const CellParams prm) {
    const uint32_t m = min(prm.counters[7], prm.tent_cap);
    uint32_t tests = 0, decided = 0, unresolved = 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x) {
        const uint32_t *rec = prm.tent + 5 * (uint64_t)r;
        const uint32_t cell = rec[0];
        if (cell == kNoCell) continue;
        const IndexHeader *hdr =
            prm.clouds != nullptr ? &prm.clouds[batch_find(prm.cloud_cells, prm.n_clouds, cell)] : prm.hdr;
        const float4 s = prm.pts[cell];
        const uint32_t qv = rec[4];
        const int o0 = (int)rec[1], o1 = (int)rec[2], o2 = (int)rec[3];
        bool wrong = true;
        ++tests;
        if (o0 >= -6 && o1 >= -6 && o2 >= -6) {
            const int res = fp64_decide(o0, o1, o2, prm.pts, hdr, s.x, s.y, s.z, s.w, (int)(qv & 0x7FFFFFFFu));
            ++decided;
            if (res & 2) ++unresolved;
            if (res & 8) log_exact(prm, prm.perm[cell], 1u);
            wrong = (uint32_t)(res & 1) != (qv >> 31);
        }
        if (wrong) {
            prm.exact_list[cell - prm.begin] = 1u;
            atomicAdd(&prm.counters[5], 1u);
            const unsigned long long k = atomicAdd(&prm.stats[kStatTentativeWrong], 1ull);
            if (k < (unsigned long long)kTentativeLogSlots)
                prm.stats[kStatTentativeLog + k] = (unsigned long long)prm.perm[cell] + 1ull;
        }
    }
    tests = wave_sum(tests);
    decided = wave_sum(decided);
    unresolved = wave_sum(unresolved);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (tests) {
            atomicAdd(&prm.stats[kStatTentative], (unsigned long long)tests);
            if (decided) atomicAdd(&prm.stats[kStatFp64Rechecks], (unsigned long long)decided);
        }
        if (unresolved) atomicAdd(&prm.stats[kStatUnresolved], (unsigned long long)unresolved);
    }
}

// Tiers m and 1: one cell per wave, the cells queued in `in` (count in
// counters[kIn], work counter counters[kIn + 1]); a cell that outgrows the
// tier goes to `out` (count in counters[kOut]).  kOwners: the queue slots
// below promo_cap carry the plane owners tier 0 held (offered first).  An
// empty queue ends the launch before any set-up.
template <typename T, int V, int P, int kThreads, int kIn, int kOut, bool kOwners>
__device__ __forceinline__ void queued_tier_body(const CellParams &prm, const uint32_t *in, uint32_t *out) {
    __shared__ GroupLDS<V, P> lds[kThreads / kWave];
    __shared__ CandLDS<kWave> cands[kThreads / kWave];
    __shared__ CellParams sprm;
    const uint32_t queued = min(prm.counters[kIn], prm.queue_cap);
    if (queued == 0u) return;
    if (threadIdx.x == 0) sprm = prm;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    // the wave's first cell before its set-up: a wave without one ends here
    // (after the block's only barrier), before it writes any spilled register
    uint32_t k = 0;
    if (lane == 0) k = atomicAdd(&prm.counters[kIn + 1], 1u);
    k = (uint32_t)__shfl((int)k, 0);
    if (k >= queued) return;
    T g;
    g.setup(prm, &lds[threadIdx.x / kWave], &cands[threadIdx.x / kWave]);
    while (true) {
        const uint64_t t0 = clock64();
        const uint32_t cell = in[k];
        if (prm.clouds != nullptr) g.set_cloud(&prm.clouds[uniform(batch_find(prm.cloud_cells, prm.n_clouds, cell))]);
        g.active = true;
        g.start(cell, *reread(&sprm));
        if constexpr (kOwners) {
            if (k < prm.promo_cap) g.offer_owners(prm.promo + (size_t)k * kPromoPlanes, kPromoPlanes);
        }
        g.traverse();
        const CellParams &Q = *reread(&sprm);
        const uint32_t orig = Q.perm[cell];
        if (lane == 0 && g.S->exact_events) log_exact(Q, orig, g.S->exact_events);
        if (!g.overflow) {
            const int F = g.stage_row(Q);
            const uint64_t base = pool_take(Q, (uint32_t)F);
            if (base + (uint64_t)F <= Q.rows_elems) {
                if (lane == 0) Q.row_offsets[cell] = (int64_t)base;
                g.write_row(Q, orig, Q.rows + base, F);
            } else if (lane == 0) {
                pool_short(Q, cell, orig, base + (uint64_t)F);
            }
        } else if (lane == 0) {  // queue for the next tier
            g.count(kStatPromoted);
            const uint32_t slot = atomicAdd(&Q.counters[kOut], 1u);
            if (slot < Q.queue_cap) out[slot] = cell;
            else truncate_cell(Q, orig);
        }
        if (Q.cycles && lane == 0)
            Q.cycles[orig] = 0x80000000u | (uint32_t)min(clock64() - t0, (uint64_t)0x7fffffff);
        if (lane == 0) k = atomicAdd(&prm.counters[kIn + 1], 1u);
        k = (uint32_t)__shfl((int)k, 0);
        if (k >= queued) break;
    }
    g.flush(prm, 0);
}

// Tier m: 192 vertices / 128 planes at three waves per SIMD, the cells tier 0
// (or its twin) gave up on capacity.
__global__ void __launch_bounds__(kTmThreads) __attribute__((amdgpu_waves_per_eu(3, 3)))
cell_group_mid_kernel(const CellParams prm) {
    queued_tier_body<TierM, kTmVerts, kTmPlanes, kTmThreads, 1, 8, true>(prm, prm.mid_list, prm.big_list);
}

// Tier 1: 512 vertices / 255 planes, the cells that outgrow tier m.
__global__ void __launch_bounds__(kT1Threads) cell_group_big_kernel(const CellParams prm) {
    queued_tier_body<Tier1, kT1Verts, kT1Planes, kT1Threads, 8, 3, false>(prm, prm.big_list, prm.huge_list);
}

// Tier 2: one cell per wave, one wave per block (57 KB of LDS), the cells
// that outgrow tier 1.
__global__ void __launch_bounds__(kT2Threads) cell_group_huge_kernel(const CellParams prm) {
    __shared__ GroupLDS<kT2Verts, kT2Planes> lds;
    __shared__ CandLDS<kT2Lanes> cands;
    __shared__ CellParams sprm;
    const uint32_t queued = min(prm.counters[3], prm.queue_cap);
    if (queued == 0u) return;
    if (threadIdx.x == 0) sprm = prm;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t k = 0;  // the wave's first cell before its set-up, as in the tiers above
    if (lane == 0) k = atomicAdd(&prm.counters[4], 1u);
    k = (uint32_t)__shfl((int)k, 0);
    if (k >= queued) return;
    Tier2 g;
    g.setup(prm, &lds, &cands);
    while (true) {
        const uint64_t t0 = clock64();
        const uint32_t cell = prm.huge_list[k];
        if (prm.clouds != nullptr) g.set_cloud(&prm.clouds[uniform(batch_find(prm.cloud_cells, prm.n_clouds, cell))]);
        g.active = true;
        g.start(cell, *reread(&sprm));
        g.traverse();
        const CellParams &P = *reread(&sprm);
        const uint32_t orig = P.perm[cell];
        if (lane == 0 && g.S->exact_events) log_exact(P, orig, g.S->exact_events);
        if (!g.overflow) {
            const int F = g.stage_row(P);
            const uint64_t base = pool_take(P, (uint32_t)F);
            if (base + (uint64_t)F <= P.rows_elems) {
                if (lane == 0) P.row_offsets[cell] = (int64_t)base;
                g.write_row(P, orig, P.rows + base, F);
            } else if (lane == 0) {
                pool_short(P, cell, orig, base + (uint64_t)F);
            }
        } else if (lane == 0) {
            truncate_cell(P, orig);  // beyond every tier
        }
        if (P.cycles && lane == 0)
            P.cycles[orig] = 0x80000000u | (uint32_t)min(clock64() - t0, (uint64_t)0x7fffffff);
        if (lane == 0) k = atomicAdd(&prm.counters[4], 1u);
        k = (uint32_t)__shfl((int)k, 0);
        if (k >= queued) break;
    }
    g.flush(prm, 0);
}

}  // namespace pg
