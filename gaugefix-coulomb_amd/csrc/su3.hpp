// su3.hpp -- register-resident 3x3 complex algebra for the CDNA4 kernels.
//
// All functions are templated on the real type T (float or double) and fully
// unrolled so that every matrix lives in VGPRs (no dynamic register indexing).
// The SU(3) operations restate the reference's group algebra
// (src/SU3_ops.c, src/SU2_ops.c of jesuel-marques/GaugeFix-Coulomb).
#pragma once

#include <hip/hip_runtime.h>

namespace gfc {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T re, im;
};

template <typename T>
struct mat3 {
    cplx<T> e[9];
};

template <typename T>
__device__ __forceinline__ cplx<T> c_make(T re, T im) {
    cplx<T> r;
    r.re = re;
    r.im = im;
    return r;
}

// acc += a * b
template <typename T>
__device__ __forceinline__ void c_mac(cplx<T> &acc, const cplx<T> &a, const cplx<T> &b) {
    acc.re += a.re * b.re;
    acc.re -= a.im * b.im;
    acc.im += a.re * b.im;
    acc.im += a.im * b.re;
}

// acc += a * conj(b)
template <typename T>
__device__ __forceinline__ void c_mac_cb(cplx<T> &acc, const cplx<T> &a, const cplx<T> &b) {
    acc.re += a.re * b.re;
    acc.re += a.im * b.im;
    acc.im += a.im * b.re;
    acc.im -= a.re * b.im;
}

// acc += conj(a) * conj(b)
template <typename T>
__device__ __forceinline__ void c_mac_cc(cplx<T> &acc, const cplx<T> &a, const cplx<T> &b) {
    acc.re += a.re * b.re;
    acc.re -= a.im * b.im;
    acc.im -= a.re * b.im;
    acc.im -= a.im * b.re;
}

template <typename T>
__device__ __forceinline__ void m_zero(mat3<T> &m) {
#pragma unroll
    for (int k = 0; k < 9; k++) m.e[k] = c_make<T>(0, 0);
}

template <typename T>
__device__ __forceinline__ void m_identity(mat3<T> &m) {
#pragma unroll
    for (int k = 0; k < 9; k++) m.e[k] = c_make<T>((k % 4 == 0) ? T(1) : T(0), 0);
}

// c = a * b
template <typename T>
__device__ __forceinline__ void m_mul(const mat3<T> &a, const mat3<T> &b, mat3<T> &c) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<T> s = c_make<T>(0, 0);
#pragma unroll
            for (int k = 0; k < 3; k++) c_mac(s, a.e[3 * i + k], b.e[3 * k + j]);
            c.e[3 * i + j] = s;
        }
}

// acc += a * b^dagger
template <typename T>
__device__ __forceinline__ void m_mac_nd(mat3<T> &acc, const mat3<T> &a, const mat3<T> &b) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) c_mac_cb(acc.e[3 * i + j], a.e[3 * i + k], b.e[3 * j + k]);
}

// acc += a^dagger * b^dagger  (= (b a)^dagger)
template <typename T>
__device__ __forceinline__ void m_mac_dd(mat3<T> &acc, const mat3<T> &a, const mat3<T> &b) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) c_mac_cc(acc.e[3 * i + j], a.e[3 * k + i], b.e[3 * j + k]);
}

// c = a * b^dagger
template <typename T>
__device__ __forceinline__ void m_mul_nd(const mat3<T> &a, const mat3<T> &b, mat3<T> &c) {
    m_zero(c);
    m_mac_nd(c, a, b);
}

// Third row of an SU(3) matrix from its first two: r2 = conj(r0 x r1).
template <typename T>
__device__ __forceinline__ void m_fill_row2(mat3<T> &m) {
    const cplx<T> *r0 = m.e, *r1 = m.e + 3;
    cplx<T> c0 = c_make<T>(0, 0), c1 = c_make<T>(0, 0), c2 = c_make<T>(0, 0);
    c_mac_cc(c0, r0[1], r1[2]);
    c_mac_cc(c1, r0[2], r1[0]);
    c_mac_cc(c2, r0[0], r1[1]);
    cplx<T> d0 = c_make<T>(0, 0), d1 = c_make<T>(0, 0), d2 = c_make<T>(0, 0);
    c_mac_cc(d0, r0[2], r1[1]);
    c_mac_cc(d1, r0[0], r1[2]);
    c_mac_cc(d2, r0[1], r1[0]);
    m.e[6] = c_make<T>(c0.re - d0.re, c0.im - d0.im);
    m.e[7] = c_make<T>(c1.re - d1.re, c1.im - d1.im);
    m.e[8] = c_make<T>(c2.re - d2.re, c2.im - d2.im);
}

template <typename T>
__device__ __forceinline__ T inv_sqrt(T x);
// FP64: the hardware estimate (v_rsq_f64) and two Newton steps
// y <- y + y (1 - x y^2) / 2, within an ulp or two of 1.0 / sqrt(x) for the
// positive normal x it is used on (norms of the update's SU(2) vectors and
// rows), in a third of the instructions of a square root and a division.
template <>
__device__ __forceinline__ double inv_sqrt<double>(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double t = x * y, e = __builtin_fma(-t, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
    t = x * y;
    e = __builtin_fma(-t, y, 1.0);
    return __builtin_fma(0.5 * y, e, y);
}

// 1 / d for d > 0 in FP64: the hardware estimate (v_rcp_f64) and two Newton
// steps y <- y + y (1 - d y), within an ulp of 1.0 / d.
__device__ __forceinline__ double rcp_nr(double d) {
    double y = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-d, y, 1.0);
    return __builtin_fma(y, e, y);
}
template <>
__device__ __forceinline__ float inv_sqrt<float>(float x) {
    return rsqrtf(x);
}

// acc += a b^dagger for SU(3) a and b, where only rows 0 and 1 of a are set:
// rows 0 and 1 of the product from those, row 2 = conj(r0 x r1) (the product
// is in SU(3)): 18 complex products and a 6-product cross product instead of
// 27 and the 6-product rebuild of a's third row.
template <typename T>
__device__ __forceinline__ void m_mac_nd_su3(mat3<T> &acc, const mat3<T> &a, const mat3<T> &b) {
    mat3<T> P;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<T> s = c_make<T>(0, 0);
#pragma unroll
            for (int k = 0; k < 3; k++) c_mac_cb(s, a.e[3 * i + k], b.e[3 * j + k]);
            P.e[3 * i + j] = s;
        }
    m_fill_row2(P);
#pragma unroll
    for (int k = 0; k < 9; k++) acc.e[k] = c_make<T>(acc.e[k].re + P.e[k].re, acc.e[k].im + P.e[k].im);
}

// acc += (a b)^dagger for SU(3) a and b, where only rows 0 and 1 of a are
// set (rows 0 and 1 of a b from those, row 2 by closure).
template <typename T>
__device__ __forceinline__ void m_mac_dagger_su3(mat3<T> &acc, const mat3<T> &a,
                                                 const mat3<T> &b) {
    mat3<T> Q;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            cplx<T> s = c_make<T>(0, 0);
#pragma unroll
            for (int k = 0; k < 3; k++) c_mac(s, a.e[3 * i + k], b.e[3 * k + j]);
            Q.e[3 * i + j] = s;
        }
    m_fill_row2(Q);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            acc.e[3 * i + j] = c_make<T>(acc.e[3 * i + j].re + Q.e[3 * j + i].re,
                                         acc.e[3 * i + j].im - Q.e[3 * j + i].im);
}

// projectSU3 (src/SU3_ops.c:994-1049): normalise row 0; row 2 = conj(r0 x r1)
// normalised; row 1 = conj(r2 x r0).  Returns false when row 0 or the cross
// product is null; the rows are then written unnormalised, as in the
// reference.
template <typename T>
__device__ __forceinline__ bool m_project_su3(mat3<T> &m) {
    bool ok = true;
    T n0 = 0;
#pragma unroll
    for (int b = 0; b < 3; b++) n0 += m.e[b].re * m.e[b].re + m.e[b].im * m.e[b].im;
    if (n0 != T(0)) {
        T r = inv_sqrt(n0);
#pragma unroll
        for (int b = 0; b < 3; b++) m.e[b] = c_make<T>(m.e[b].re * r, m.e[b].im * r);
    } else {
        ok = false;
    }
    m_fill_row2(m);
    T n2 = 0;
#pragma unroll
    for (int b = 6; b < 9; b++) n2 += m.e[b].re * m.e[b].re + m.e[b].im * m.e[b].im;
    if (n2 != T(0)) {
        T r = inv_sqrt(n2);
#pragma unroll
        for (int b = 6; b < 9; b++) m.e[b] = c_make<T>(m.e[b].re * r, m.e[b].im * r);
    } else {
        ok = false;
    }
    // row 1 = conj(r2 x r0)
    const cplx<T> *r0 = m.e, *r2 = m.e + 6;
    cplx<T> c0 = c_make<T>(0, 0), c1 = c_make<T>(0, 0), c2 = c_make<T>(0, 0);
    c_mac_cc(c0, r2[1], r0[2]);
    c_mac_cc(c1, r2[2], r0[0]);
    c_mac_cc(c2, r2[0], r0[1]);
    cplx<T> d0 = c_make<T>(0, 0), d1 = c_make<T>(0, 0), d2 = c_make<T>(0, 0);
    c_mac_cc(d0, r2[2], r0[1]);
    c_mac_cc(d1, r2[0], r0[2]);
    c_mac_cc(d2, r2[1], r0[0]);
    m.e[3] = c_make<T>(c0.re - d0.re, c0.im - d0.im);
    m.e[4] = c_make<T>(c1.re - d1.re, c1.im - d1.im);
    m.e[5] = c_make<T>(c2.re - d2.re, c2.im - d2.im);
    return ok;
}

// determinant3x3 (src/SU3_ops.c:251-281), Leibniz expansion along row 0.
template <typename T>
__device__ __forceinline__ cplx<T> m_det(const mat3<T> &m) {
    const cplx<T> *e = m.e;
    cplx<T> c0 = c_make<T>(0, 0), c1 = c_make<T>(0, 0), c2 = c_make<T>(0, 0);
    c_mac(c0, e[4], e[8]);
    c_mac(c1, e[5], e[6]);
    c_mac(c2, e[3], e[7]);
    cplx<T> d0 = c_make<T>(0, 0), d1 = c_make<T>(0, 0), d2 = c_make<T>(0, 0);
    c_mac(d0, e[5], e[7]);
    c_mac(d1, e[3], e[8]);
    c_mac(d2, e[4], e[6]);
    cplx<T> m0 = c_make<T>(c0.re - d0.re, c0.im - d0.im);
    cplx<T> m1 = c_make<T>(c1.re - d1.re, c1.im - d1.im);
    cplx<T> m2 = c_make<T>(c2.re - d2.re, c2.im - d2.im);
    cplx<T> r = c_make<T>(0, 0);
    c_mac(r, e[0], m0);
    c_mac(r, e[1], m1);
    c_mac(r, e[2], m2);
    return r;
}

// The SU(2) element of a Cabibbo-Marinari step on rows (A, B) of w
// (updateSubLosAlamos, src/SU3_ops.c:776-840 with projectSU2,
// src/SU2_ops.c:369-402): x = a/|a| with a = (Re(w_AA + w_BB),
// -Im(w_AB + w_BA), -Re(w_AB - w_BA), -Im(w_AA - w_BB)) maximises
// Re Tr(x w_sub).  False for a null a (the step is then skipped).
template <typename T, int A, int B>
__device__ __forceinline__ bool cm_su2(const mat3<T> &w, T &a0, T &a1, T &a2, T &a3) {
    const cplx<T> waa = w.e[3 * A + A], wbb = w.e[3 * B + B];
    const cplx<T> wab = w.e[3 * A + B], wba = w.e[3 * B + A];
    a0 = waa.re + wbb.re;
    a1 = -(wab.im + wba.im);
    a2 = -(wab.re - wba.re);
    a3 = -(waa.im - wbb.im);
    const T s = a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3;
    if (s == T(0)) return false;
    const T r = inv_sqrt(s);
    a0 *= r;
    a1 *= r;
    a2 *= r;
    a3 *= r;
    return true;
}

// Rows A, B of m left-multiplied by the SU(2) element
// [[x0 + i x3, x2 + i x1], [-x2 + i x1, x0 - i x3]].
template <typename T, int A, int B>
__device__ __forceinline__ void cm_rows(mat3<T> &m, T a0, T a1, T a2, T a3) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const cplx<T> p = m.e[3 * A + c], q = m.e[3 * B + c];
        m.e[3 * A + c] = c_make<T>(a0 * p.re - a3 * p.im + a2 * q.re - a1 * q.im,
                                   a0 * p.im + a3 * p.re + a2 * q.im + a1 * q.re);
        m.e[3 * B + c] = c_make<T>(-a2 * p.re - a1 * p.im + a0 * q.re + a3 * q.im,
                                   -a2 * p.im + a1 * p.re + a0 * q.im - a3 * q.re);
    }
}

// One Cabibbo-Marinari step on rows (A, B): rows A, B of w and of the
// accumulated update g are left-multiplied by the step's SU(2) element.
template <typename T, int A, int B>
__device__ __forceinline__ void cm_step(mat3<T> &w, mat3<T> &g) {
    T a0, a1, a2, a3;
    if (!cm_su2<T, A, B>(w, a0, a1, a2, a3)) return;
    cm_rows<T, A, B>(w, a0, a1, a2, a3);
    cm_rows<T, A, B>(g, a0, a1, a2, a3);
}

// The first two steps of the first hit, from g = 1 (what cm_step<0, 1> then
// cm_step<0, 2> compute there, without the products with the identity's zeros
// and ones): step (0, 1) sets rows 0, 1 of g to the SU(2) element x, step
// (0, 2) mixes row 0 = (x00, x01, 0) with row 2 = (0, 0, 1).  The values are
// those of the general steps (a product with an exact 0 or 1 is exact).
template <typename T>
__device__ __forceinline__ void cm_first_two(mat3<T> &w, mat3<T> &g) {
    m_identity(g);
    T a0, a1, a2, a3;
    if (cm_su2<T, 0, 1>(w, a0, a1, a2, a3)) {
        cm_rows<T, 0, 1>(w, a0, a1, a2, a3);
        g.e[0] = c_make<T>(a0, a3);
        g.e[1] = c_make<T>(a2, a1);
        g.e[3] = c_make<T>(-a2, a1);
        g.e[4] = c_make<T>(a0, -a3);
    }
    if (!cm_su2<T, 0, 2>(w, a0, a1, a2, a3)) return;
    cm_rows<T, 0, 2>(w, a0, a1, a2, a3);
    // g rows 0 and 2: columns 0, 1 from row 0 alone (row 2 is 0 there),
    // column 2 from row 2 alone (row 0 is 0 there, row 2 is 1)
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const cplx<T> p = g.e[c];
        g.e[c] = c_make<T>(a0 * p.re - a3 * p.im, a0 * p.im + a3 * p.re);
        g.e[6 + c] = c_make<T>(-a2 * p.re - a1 * p.im, -a2 * p.im + a1 * p.re);
    }
    g.e[2] = c_make<T>(a2, a1);
    g.e[8] = c_make<T>(a0, -a3);
}

// Local Los-Alamos + overrelaxation update from w(n)
// (gaugefixOverrelaxationLocal / updateLosAlamos / power3x3Binomial,
// src/gauge_fixing.c:537-708, src/SU3_ops.c:734-771).
//
// The reference forms g_LA = w_final * w^-1; here the same SU(2) factors are
// accumulated directly into g_LA (g_LA = x_k ... x_1, which equals
// w_final * w^-1 exactly in real arithmetic) -- no complex division and a
// smaller rounding error.  The reference runs the hits only when the real
// part of det(w) is non-zero (inverse3x3 returns the determinant as a double);
// otherwise g_LA = w w^-1 = 1, or g_LA = 0 when det(w) = 0 (the zero inverse).
// Then g_OR = Proj_SU3(1 + omega (g_LA - 1)), or the identity if the
// projection meets a null row.
template <typename T>
__device__ __forceinline__ void la_or_update(mat3<T> &w, int hits, T omega, mat3<T> &g) {
    cplx<T> d = m_det(w);
    if (d.re == T(0)) {
        if (d.im == T(0))
            m_zero(g);
        else
            m_identity(g);
    } else if (hits > 0) {
        cm_first_two(w, g);
        cm_step<T, 1, 2>(w, g);
        for (int h = 1; h < hits - 1; h++) {
            cm_step<T, 0, 1>(w, g);
            cm_step<T, 0, 2>(w, g);
            cm_step<T, 1, 2>(w, g);
        }
        // the last hit outside the loop: the entries of w that no later step
        // reads (row 1's first and row 0's last two columns after (0, 1), rows 0
        // and 2 but for w21, w22 after (0, 2), everything after (1, 2)) are then
        // not computed at all
        if (hits > 1) {
            cm_step<T, 0, 1>(w, g);
            cm_step<T, 0, 2>(w, g);
            cm_step<T, 1, 2>(w, g);
        }
    } else {
        m_identity(g);
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        T re = g.e[k].re, im = g.e[k].im;
        if (k % 4 == 0)
            g.e[k] = c_make<T>(T(1) + omega * (re - T(1)), omega * im);
        else
            g.e[k] = c_make<T>(omega * re, omega * im);
    }
    if (!m_project_su3(g)) m_identity(g);
}

}  // namespace gfc
