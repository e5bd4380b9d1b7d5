// Exact signs of the clipping predicates by floating-point expansion
// arithmetic (J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic
// and Fast Robust Geometric Predicates", Discrete & Computational Geometry
// 18, 1997: Two-Sum, Fast-Two-Sum, Two-Product with a fused multiply-add,
// Grow-Expansion, Scale-Expansion, Compress).
//
// Used where the FP64 filter of a removal test cannot decide its sign (the
// test a vertex of planes i, j, k against the plane of q, as in
// voronoi_convex_cell.cuh:256-260 / faster_convex_cell.cuh:280-285, but
// decided exactly instead of in fast-math FP32).  All inputs are FP32 values
// (coordinates, weights, the box), so every quantity below is a polynomial in
// exactly representable differences and the expansions hold it exactly.
//
// An expansion is a sum of nonoverlapping doubles in increasing magnitude
// with no zero components; its sign is the sign of its largest component.
// Capacities are fixed (no allocation on the device): a result that would
// need more than kCap components after compression is marked invalid
// (n < 0), and the caller falls back to the FP64 sign and counts it.  With
// kCap = 12 that needs inputs whose exponents span hundreds of bits.
//
// Compiled for the device (pg_cell.hip), for the host by hipcc (the
// library's check of this code against rational arithmetic) and by a plain
// C++ compiler (the host-core build, pg_cpu.cpp), with contraction of
// multiplies and adds into FMAs switched off: the error-free transforms rely
// on every product and sum being rounded on its own (plain C++ builds pass
// -ffp-contract=off).
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
#define __host__
#define __device__
#endif
#include <math.h>

namespace pg {
namespace xp {

constexpr int kCap = 12;

struct Ex {
    int n;  // components in use; < 0: capacity exceeded
    double c[kCap];
};

// a value held exactly as hi + lo (a difference of two FP32 numbers)
struct E2 {
    double hi, lo;
};

__host__ __device__ inline void two_sum(double a, double b, double &x, double &y) {
#pragma clang fp contract(off)
    x = a + b;
    const double bv = x - a;
    const double av = x - bv;
    y = (a - av) + (b - bv);
}
// |a| >= |b| (or a == 0)
__host__ __device__ inline void fast_two_sum(double a, double b, double &x, double &y) {
#pragma clang fp contract(off)
    x = a + b;
    y = b - (x - a);
}
__host__ __device__ inline void two_prod(double a, double b, double &x, double &y) {
#pragma clang fp contract(off)
    x = a * b;
    y = fma(a, b, -x);
}

__host__ __device__ inline void set_zero(Ex &e) { e.n = 0; }
__host__ __device__ inline void set_one(Ex &e, double v) {
    e.n = 0;
    if (v != 0.0) e.c[e.n++] = v;
}
__host__ __device__ inline void from2(Ex &e, const E2 &v) {
    e.n = 0;
    if (v.lo != 0.0) e.c[e.n++] = v.lo;
    if (v.hi != 0.0) e.c[e.n++] = v.hi;
}
// a - b exactly
__host__ __device__ inline E2 diff2(double a, double b) {
    E2 r;
    two_sum(a, -b, r.hi, r.lo);
    if (r.hi == 0.0) r.hi = r.lo, r.lo = 0.0;
    return r;
}

// Compress (Shewchuk, section 2.8) of e[0..m) in place; returns the new
// length.  The result is nonoverlapping, in increasing magnitude, with no
// zero components.
__host__ __device__ inline int compress(double *e, int m) {
#pragma clang fp contract(off)
    if (m <= 0) return 0;
    int bottom = m - 1;
    double q = e[m - 1];
    for (int i = m - 2; i >= 0; --i) {
        double x, y;
        fast_two_sum(q, e[i], x, y);
        if (y != 0.0) {
            e[bottom--] = x;
            q = y;
        } else {
            q = x;
        }
    }
    e[bottom] = q;
    int top = 0;
    for (int i = bottom + 1; i < m; ++i) {
        double x, y;
        fast_two_sum(e[i], q, x, y);
        if (y != 0.0) e[top++] = y;
        q = x;
    }
    if (q != 0.0) e[top++] = q;
    return top;
}

// Grow-Expansion in place: e[0..m) + b, room for m + 1 components; returns
// the new length (zero components dropped).
__host__ __device__ inline int grow(double *e, int m, double b) {
#pragma clang fp contract(off)
    double q = b;
    int out = 0;
    for (int i = 0; i < m; ++i) {
        double x, y;
        two_sum(q, e[i], x, y);
        if (y != 0.0) e[out++] = y;
        q = x;
    }
    if (q != 0.0) e[out++] = q;
    return out;
}

// h = a + s b (s = +1 or -1); h may be a or b (both are read before h is written)
__host__ __device__ inline void add(Ex &h, const Ex &a, const Ex &b, double s = 1.0) {
    if (a.n < 0 || b.n < 0) {
        h.n = -1;
        return;
    }
    double t[2 * kCap + 1];
    int m = a.n;
    for (int i = 0; i < m; ++i) t[i] = a.c[i];
    for (int j = 0; j < b.n; ++j) m = grow(t, m, s * b.c[j]);
    m = compress(t, m);
    if (m > kCap) {
        h.n = -1;
        return;
    }
    h.n = m;
    for (int i = 0; i < m; ++i) h.c[i] = t[i];
}

// Scale-Expansion: out[0..) = e b, room for 2 e.n components; returns the
// length (zero components dropped).
__host__ __device__ inline int scale(double *out, const Ex &e, double b) {
#pragma clang fp contract(off)
    if (e.n <= 0 || b == 0.0) return 0;
    int m = 0;
    double q, t0;
    two_prod(e.c[0], b, q, t0);
    if (t0 != 0.0) out[m++] = t0;
    for (int i = 1; i < e.n; ++i) {
        double T, t, x, y;
        two_prod(e.c[i], b, T, t);
        two_sum(q, t, x, y);
        if (y != 0.0) out[m++] = y;
        fast_two_sum(T, x, q, y);
        if (y != 0.0) out[m++] = y;
    }
    if (q != 0.0) out[m++] = q;
    return m;
}

// h = a b; h may be a or b (both are read before h is written)
__host__ __device__ inline void mul(Ex &h, const Ex &a, const Ex &b) {
    if (a.n < 0 || b.n < 0) {
        h.n = -1;
        return;
    }
    double acc[3 * kCap + 1];
    double t[2 * kCap];
    int m = 0;
    for (int j = 0; j < b.n; ++j) {
        const int k = scale(t, a, b.c[j]);
        for (int i = 0; i < k; ++i) m = grow(acc, m, t[i]);
        m = compress(acc, m);
        if (m > kCap) {
            h.n = -1;
            return;
        }
    }
    h.n = m;
    for (int i = 0; i < m; ++i) h.c[i] = acc[i];
}

__host__ __device__ inline int sign(const Ex &e) {
    if (e.n <= 0) return 0;
    const double top = e.c[e.n - 1];
    return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

// One plane of a cell in seed-relative doubled coordinates y = 2 (x - s):
// a . y <= H with, for a point p, a = p - s and H = |a|^2 - (w_p - w_s), and
// for a box plane (code = 2 axis + side) a = -e_axis, H = 2 (s - lo) (side 0)
// or a = e_axis, H = 2 (hi - s) (side 1).
struct PlaneIn {
    float x, y, z, w;  // the point (ignored for a box plane)
    int box;           // -1: point plane, else 2 axis + side
};

__host__ __device__ inline void plane_a(E2 a[3], const PlaneIn &p, const float s[4]) {
    if (p.box < 0) {
        a[0] = diff2((double)p.x, (double)s[0]);
        a[1] = diff2((double)p.y, (double)s[1]);
        a[2] = diff2((double)p.z, (double)s[2]);
    } else {
        const int axis = p.box >> 1;
        for (int k = 0; k < 3; ++k) a[k] = E2{k == axis ? ((p.box & 1) ? 1.0 : -1.0) : 0.0, 0.0};
    }
}
__host__ __device__ inline void plane_H(Ex &H, const E2 a[3], const PlaneIn &p, const float s[4], const float lo[3],
                                        const float hi[3]) {
    Ex u, v, t;
    if (p.box >= 0) {
        const int axis = p.box >> 1;
        const E2 d = (p.box & 1) ? diff2((double)hi[axis], (double)s[axis]) : diff2((double)s[axis], (double)lo[axis]);
        from2(H, E2{2.0 * d.hi, 2.0 * d.lo});  // scaling by 2 is exact
        return;
    }
    set_zero(H);
    for (int k = 0; k < 3; ++k) {
        from2(u, a[k]);
        mul(t, u, u);
        add(v, H, t);
        H = v;
    }
    from2(u, diff2((double)p.w, (double)s[3]));
    add(v, H, u, -1.0);
    H = v;
}

// det[u; v; w] = u . (v x w)
__host__ __device__ inline void det3(Ex &out, const E2 *u, const E2 *v, const E2 *w) {
    Ex acc, x, y, p, q, r;
    set_zero(acc);
    for (int k = 0; k < 3; ++k) {
        const int i = (k + 1) % 3, j = (k + 2) % 3;
        from2(x, v[i]);
        from2(y, w[j]);
        mul(p, x, y);
        from2(x, v[j]);
        from2(y, w[i]);
        mul(q, x, y);
        add(r, p, q, -1.0);  // (v x w)_k
        from2(x, u[k]);
        mul(p, x, r);
        add(q, acc, p);
        acc = q;
    }
    out = acc;
}

// Exact signs for the vertex of planes pl[0..2] against the plane pl[3] of
// the cell of seed s (x, y, z, w) with box [lo, hi]:
//   D = det[a_0; a_1; a_2], the vertex's orientation, and
//   P = a_q . N - H_q D with N = H_0 (a_1 x a_2) + H_1 (a_2 x a_0) + H_2 (a_0 x a_1),
// so that the vertex N / D lies strictly beyond plane q iff sign(P) sign(D) > 0.
// Returns false (signs unset) when an expansion exceeds its capacity.
// P is evaluated as the expansion of the 4 x 4 determinant along its H
// column: P = H_0 det(a_1, a_2, a_q) + H_1 det(a_2, a_0, a_q)
//           + H_2 det(a_0, a_1, a_q) - H_q det(a_0, a_1, a_2).
__host__ __device__ inline bool exact_signs(const PlaneIn pl[4], const float s[4], const float lo[3],
                                            const float hi[3], int &sD, int &sP) {
    E2 a[4][3];
    for (int t = 0; t < 4; ++t) plane_a(a[t], pl[t], s);
    Ex D, P, M, H, T, U;
    det3(D, a[0], a[1], a[2]);
    plane_H(H, a[3], pl[3], s, lo, hi);
    mul(T, H, D);
    set_zero(P);
    add(P, P, T, -1.0);
    for (int t = 0; t < 3; ++t) {
        const int i = (t + 1) % 3, j = (t + 2) % 3;
        det3(M, a[i], a[j], a[3]);
        plane_H(H, a[t], pl[t], s, lo, hi);
        mul(T, H, M);
        add(U, P, T);
        P = U;
    }
    if (D.n < 0 || P.n < 0) return false;
    sD = sign(D);
    sP = sign(P);
    return true;
}

// The value of an expansion to within a few units in the last place of a
// double: its components summed smallest first (they do not overlap, so each
// partial sum at least doubles the last and the rounding errors stay below
// one unit in the last place of the result).
__host__ __device__ inline double approx(const Ex &e) {
#pragma clang fp contract(off)
    double v = 0.0;
    for (int i = 0; i < e.n; ++i) v += e.c[i];
    return v;
}

// Position of the vertex of planes pl[0..2] of the cell of seed s with box
// [lo, hi]: y = N / D, N and D as in exact_signs and held exactly, so that
// y is known to a few units in the last place of a double however small D
// is (three nearly coaxial planes, where the FP64 bound on N / D fails).
// out[k] = y_k rounded to FP32 and w >= max_k |y_k - out[k]|.  Returns false
// when D is not positive or an expansion exceeds its capacity.
__host__ __device__ inline bool exact_position(const PlaneIn pl[3], const float s[4], const float lo[3],
                                               const float hi[3], float out[3], float &w) {
    E2 a[3][3];
    for (int t = 0; t < 3; ++t) plane_a(a[t], pl[t], s);
    Ex D;
    det3(D, a[0], a[1], a[2]);
    if (D.n < 0 || sign(D) <= 0) return false;
    Ex H[3];
    for (int t = 0; t < 3; ++t) plane_H(H[t], a[t], pl[t], s, lo, hi);
    const double d = approx(D);
    double err = 0.0;
    for (int k = 0; k < 3; ++k) {
        const int i = (k + 1) % 3, j = (k + 2) % 3;
        Ex acc, x, y, p, q, r;
        set_zero(acc);
        for (int t = 0; t < 3; ++t) {
            const E2 *u = a[(t + 1) % 3], *v = a[(t + 2) % 3];
            from2(x, u[i]);
            from2(y, v[j]);
            mul(p, x, y);
            from2(x, u[j]);
            from2(y, v[i]);
            mul(q, x, y);
            add(r, p, q, -1.0);  // (u x v)_k
            mul(p, H[t], r);
            add(q, acc, p);
            acc = q;
        }
        if (acc.n < 0) return false;
        // N_k / D from two values each within 2^-52 relative, one division:
        // within 1e-15 relative; 1e-13 of |y_k| covers it with room
        const double yk = approx(acc) / d;
        out[k] = (float)yk;
        const double e = fabs(yk - (double)out[k]) + fabs(yk) * 1e-13;
        err = e > err ? e : err;
    }
    w = (float)(err * (1.0 + 1e-6)) + 1e-30f;
    return true;
}

}  // namespace xp
}  // namespace pg
