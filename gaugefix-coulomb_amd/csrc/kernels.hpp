// kernels.hpp -- HIP kernels of the Coulomb gauge-fixing engine (gfx950).
//
// Device state of one lattice (V = ns^3 nt sites, Vh = V/2 per parity):
//   cfg : AoS FP64, the reference's U image (4V matrices, site-major, mu=T,Z,Y,X)
//   u0  : SoA spatial links of the configuration being fixed
//   g   : SoA gauge transformation
// u0 and g hold SU(3) matrices as their first two rows (6 complex numbers; an
// FP64 u0 link as five, see Stored below); row 2 = conj(r0 x r1) is rebuilt in
// registers on every load.  One tiled field
// (see field_size below) per parity p and, for u0, direction i = Z,Y,X:
//   u0 field p*3 + i,   g field p.
// Within a parity the index is h = lex >> 1 with lex = ((t ns + z) ns + y) ns + x
// (ns even), so the y/z neighbours of a site sit at a constant offset +-ns/2,
// +-ns^2/2 in the other parity's field and the x neighbours at h or h+-1; a
// timeslice t is the range [t ns^3/2, (t+1) ns^3/2) of h.
//
// The sweep keeps the reference's state in factored form: the current links are
// U_i(n) = G(n) u0_i(n) G(n+i)^dagger, so a half-sweep only rewrites G on the
// active parity while u0 stays read-only.  w(n) = sum_i U_i(n) + U_i(n-i)^dagger
// (src/gauge_fixing.c:493-531) becomes w(n) = G(n) W(n) with
// W(n) = sum_i [u0_i(n) G(n+i)^dagger + u0_i(n-i)^dagger G(n-i)^dagger].
#pragma once

#include "su3.hpp"

namespace gfc {

constexpr int kBlock = 256;
constexpr int kPlanes = 6;  // planes of a tile: complex numbers of a two-row matrix

// Kernels and device helpers have internal linkage: each translation unit
// that launches them (engine.hip, server.hip) carries its own code object.
namespace {

// Loads that stream (read once per sweep of a lattice larger than the
// Infinity Cache) carry the non-temporal hint so that the L2 keeps the
// neighbour G matrices, which six sites read.
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ cplx<double> ld_stream(const cplx<double> *p) {
    dvec2 v = __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(p));
    return c_make<double>(v.x, v.y);
}
__device__ __forceinline__ cplx<float> ld_stream(const cplx<float> *p) {
    fvec2 v = __builtin_nontemporal_load(reinterpret_cast<const fvec2 *>(p));
    return c_make<float>(v.x, v.y);
}
__device__ __forceinline__ void st_stream(cplx<double> *p, const cplx<double> &c) {
    dvec2 v = {c.re, c.im};
    __builtin_nontemporal_store(v, reinterpret_cast<dvec2 *>(p));
}
__device__ __forceinline__ void st_stream(cplx<float> *p, const cplx<float> &c) {
    fvec2 v = {c.re, c.im};
    __builtin_nontemporal_store(v, reinterpret_cast<fvec2 *>(p));
}

// Tiled layout of a matrix field over n parity sites: the six stored complex
// numbers of 32 consecutive sites form one 6 x 32 tile, element k of site h at
// tile_elem(h) + 32 k.  One 32-bit offset (on a uniform base) plus immediate
// steps of 32 elements addresses a whole matrix, and 32 lanes with
// consecutive h read 32 consecutive numbers.  A field holds field_size(n)
// complex numbers (n rounded up to whole tiles) and must stay below 4 GiB.
constexpr int kTile = 32;

__host__ __device__ constexpr int64_t field_size(int64_t n) {
    return (n + kTile - 1) / kTile * kTile * kPlanes;
}

__device__ __forceinline__ uint32_t tile_elem(int h) {
    return (uint32_t)(h >> 5) * (uint32_t)(kTile * kPlanes) + (uint32_t)(h & (kTile - 1));
}

template <typename T>
__device__ __forceinline__ const cplx<T> *elem_ptr(const cplx<T> *field, uint32_t byte_off,
                                                   int k) {
    return reinterpret_cast<const cplx<T> *>(reinterpret_cast<const char *>(field) +
                                             (uint64_t)byte_off + k * kTile * sizeof(cplx<T>));
}
template <typename T>
__device__ __forceinline__ cplx<T> *elem_ptr(cplx<T> *field, uint32_t byte_off, int k) {
    return reinterpret_cast<cplx<T> *>(reinterpret_cast<char *>(field) + (uint64_t)byte_off +
                                       k * kTile * sizeof(cplx<T>));
}

template <typename T>
__device__ __forceinline__ uint32_t tile_bytes(int h) {
    return tile_elem(h) * (uint32_t)sizeof(cplx<T>);
}

// u0 fields.  The sweeps read all six u0 links of a site per update, two
// thirds of their bytes, and they are bound by bytes; so an FP64 u0 matrix
// keeps five complex numbers: row 0, and the two elements of row 1 that do not
// sit under row 0's element of largest modulus (its pivot).  The pivot's row-1
// element follows from the orthogonality of the rows (u0_unpack), to a few
// ulps of the projected input.  FP32 u0 keeps both rows (six numbers): the
// FP32 sweep is also at its VALU limit.  G keeps both rows too: stored in five
// numbers it made the 48^3 x 96 FP64 fix 6% slower (the unpacking sits on the
// neighbour loads' path).  Tiles hold U0Planes<T>::n planes of
// 32 sites; the allocation keeps six planes per field (it also stages FP32
// input on upload).
template <typename T>
struct U0Planes {
    static constexpr int n = kPlanes;
};
template <>
struct U0Planes<double> {
    static constexpr int n = 5;
};

template <typename T>
__host__ __device__ constexpr int64_t u0_field_size(int64_t n) {
    return (n + kTile - 1) / kTile * kTile * U0Planes<T>::n;
}

template <typename T>
__device__ __forceinline__ uint32_t u0_tile_bytes(int h) {
    return ((uint32_t)(h >> 5) * (uint32_t)(kTile * U0Planes<T>::n) +
            (uint32_t)(h & (kTile - 1))) * (uint32_t)sizeof(cplx<T>);
}

// Field of the spatial links u0_i on parity p; field of G on parity p.
template <typename T>
__device__ __forceinline__ const cplx<T> *u0_field(const cplx<T> *u0, int Vh, int p, int i) {
    return u0 + (p * 3 + i) * u0_field_size<T>(Vh);
}
template <typename T>
__device__ __forceinline__ cplx<T> *u0_field(cplx<T> *u0, int Vh, int p, int i) {
    return u0 + (p * 3 + i) * u0_field_size<T>(Vh);
}
template <typename T>
__device__ __forceinline__ const cplx<T> *g_field(const cplx<T> *g, int Vh, int p) {
    return g + p * field_size(Vh);
}
template <typename T>
__device__ __forceinline__ cplx<T> *g_field(cplx<T> *g, int Vh, int p) {
    return g + p * field_size(Vh);
}

// SU(3) matrix from its two stored rows.
template <bool STREAM, typename T>
__device__ __forceinline__ void load_su3(mat3<T> &m, const cplx<T> *__restrict__ field, int h) {
    const uint32_t off = tile_bytes<T>(h);
#pragma unroll
    for (int k = 0; k < kPlanes; k++)
        m.e[k] = STREAM ? ld_stream(elem_ptr(field, off, k)) : *elem_ptr(field, off, k);
    m_fill_row2(m);
}

// The two stored rows only (row 2 left unset).
template <bool STREAM, typename T>
__device__ __forceinline__ void load_rows01(mat3<T> &m, const cplx<T> *__restrict__ field, int h) {
    const uint32_t off = tile_bytes<T>(h);
#pragma unroll
    for (int k = 0; k < kPlanes; k++)
        m.e[k] = STREAM ? ld_stream(elem_ptr(field, off, k)) : *elem_ptr(field, off, k);
}

// Index of row 0's element of largest modulus (ties to the lower index); the
// products are explicit FMAs, so a builder and a reader compute the same bits.
__device__ __forceinline__ int u0_pivot(const cplx<double> *r0) {
    const double n0 = __builtin_fma(r0[0].re, r0[0].re, r0[0].im * r0[0].im);
    const double n1 = __builtin_fma(r0[1].re, r0[1].re, r0[1].im * r0[1].im);
    const double n2 = __builtin_fma(r0[2].re, r0[2].re, r0[2].im * r0[2].im);
    return (n0 >= n1 && n0 >= n2) ? 0 : (n1 >= n2 ? 1 : 2);
}

// Rows 0 and 1 of an FP64 u0 matrix from its five stored numbers e[0..4]:
// row 0 = e[0..2]; row 1 = e[3], e[4] in the two non-pivot columns, and in the
// pivot column k the element x with sum_j r0_j conj(r1_j) = 0, i.e.
// x = -conj(S) r0_k / |r0_k|^2, S = sum_{j != k} r0_j conj(r1_j).  |r0_k|^2 is
// at least 1/3, so the division (a reciprocal, rcp_nr) is well conditioned.
__device__ __forceinline__ cplx<double> csel(bool c, cplx<double> x, cplx<double> y) {
    return c_make<double>(c ? x.re : y.re, c ? x.im : y.im);
}

__device__ __forceinline__ void u0_unpack(mat3<double> &m, const cplx<double> *e) {
    const int k = u0_pivot(e);
    const cplx<double> e0 = e[0], e1 = e[1], e2 = e[2], s0 = e[3], s1 = e[4];
    const cplx<double> r = csel(k == 0, e0, csel(k == 1, e1, e2));
    const cplx<double> a = csel(k == 0, e1, e0);  // row 0 over the first stored element
    const cplx<double> b = csel(k == 2, e1, e2);  // ... and over the second
    // S = a conj(s0) + b conj(s1)
    const double sre = __builtin_fma(a.re, s0.re, __builtin_fma(a.im, s0.im,
                       __builtin_fma(b.re, s1.re, b.im * s1.im)));
    const double sim = __builtin_fma(a.im, s0.re, -__builtin_fma(a.re, s0.im,
                       -__builtin_fma(b.im, s1.re, -(b.re * s1.im))));
    const double nr = __builtin_fma(r.re, r.re, r.im * r.im);
    const double inv = rcp_nr(nr);
    const cplx<double> x = c_make<double>(-__builtin_fma(sre, r.re, sim * r.im) * inv,
                                          -__builtin_fma(sre, r.im, -(sim * r.re)) * inv);
    m.e[0] = e0;
    m.e[1] = e1;
    m.e[2] = e2;
    m.e[3] = csel(k == 0, x, s0);
    m.e[4] = csel(k == 0, s0, csel(k == 1, x, s1));
    m.e[5] = csel(k == 2, x, s1);
}

// The N stored numbers of a matrix at byte offset off of a field, as loaded.
template <bool STREAM, typename T, int N>
__device__ __forceinline__ void load_raw(cplx<T> (&e)[N], const cplx<T> *__restrict__ field,
                                         uint32_t off) {
#pragma unroll
    for (int k = 0; k < N; k++)
        e[k] = STREAM ? ld_stream(elem_ptr(field, off, k)) : *elem_ptr(field, off, k);
}

// Rows 0 and 1 (row 2 left unset) from the stored numbers of a G matrix ...
template <typename T>
__device__ __forceinline__ void rows01(mat3<T> &m, const cplx<T> (&e)[kPlanes]) {
#pragma unroll
    for (int k = 0; k < kPlanes; k++) m.e[k] = e[k];
}

// ... and of a u0 matrix.
template <typename T>
__device__ __forceinline__ void u0_rows01(mat3<T> &m, const cplx<T> (&e)[U0Planes<T>::n]) {
    if constexpr (U0Planes<T>::n == kPlanes)
        rows01(m, e);
    else
        u0_unpack(m, e);
}

// Rows 0 and 1 of a u0 matrix (row 2 left unset).
template <bool STREAM, typename T>
__device__ __forceinline__ void load_u0_rows01(mat3<T> &m, const cplx<T> *__restrict__ field,
                                               int h) {
    cplx<T> e[U0Planes<T>::n];
    load_raw<STREAM>(e, field, u0_tile_bytes<T>(h));
    u0_rows01(m, e);
}

// A u0 matrix in FP64 (rows 0 and 1 as stored, row 2 rebuilt).
template <typename T>
__device__ __forceinline__ void load_u0_d(mat3<double> &m, const cplx<T> *__restrict__ field,
                                          int h) {
    mat3<T> t;
    load_u0_rows01<false>(t, field, h);
#pragma unroll
    for (int k = 0; k < kPlanes; k++) m.e[k] = c_make<double>((double)t.e[k].re, (double)t.e[k].im);
    m_fill_row2(m);
}

template <bool STREAM, typename T>
__device__ __forceinline__ void store_su3(const mat3<T> &m, cplx<T> *__restrict__ field, int h) {
    const uint32_t off = tile_bytes<T>(h);
#pragma unroll
    for (int k = 0; k < kPlanes; k++) {
        if (STREAM)
            st_stream(elem_ptr(field, off, k), m.e[k]);
        else
            *elem_ptr(field, off, k) = m.e[k];
    }
}

// The reunitarization of G (reunitarizeField, src/fields.c:144-187) always
// projects in FP64: an FP32 G (the FP32 phase of the mixed mode) is promoted,
// projected and rounded back.  The projection reads rows 0 and 1 only.
__device__ __forceinline__ void project_g(mat3<double> &m) { m_project_su3(m); }
__device__ __forceinline__ void project_g(mat3<float> &m) {
    mat3<double> d;
#pragma unroll
    for (int k = 0; k < kPlanes; k++) d.e[k] = c_make<double>(m.e[k].re, m.e[k].im);
    m_project_su3(d);
#pragma unroll
    for (int k = 0; k < kPlanes; k++) m.e[k] = c_make<float>((float)d.e[k].re, (float)d.e[k].im);
}

// Site geometry of parity index h on parity p.
struct Site {
    int x, y, z, t, xh, row;
};

__device__ __forceinline__ Site site_of(int h, int p, int ns) {
    const int nh = ns >> 1;
    Site s;
    s.xh = h % nh;
    int r = h / nh;
    s.y = r % ns;
    r /= ns;
    s.z = r % ns;
    s.t = r / ns;
    s.x = 2 * s.xh + ((s.y + s.z + s.t + p) & 1);
    s.row = h - s.xh;
    return s;
}

// Parity indices (in the other parity) of the six spatial neighbours, in the
// direction order i = 0 (Z), 1 (Y), 2 (X).
struct Nbr {
    int fwd[3], bwd[3];
};

__device__ __forceinline__ Nbr neighbours(const Site &s, int h, int ns) {
    const int nh = ns >> 1, sz = ns * nh;
    Nbr n;
    n.fwd[0] = h + (s.z + 1 == ns ? -(ns - 1) * sz : sz);
    n.bwd[0] = h + (s.z == 0 ? (ns - 1) * sz : -sz);
    n.fwd[1] = h + (s.y + 1 == ns ? -(ns - 1) * nh : nh);
    n.bwd[1] = h + (s.y == 0 ? (ns - 1) * nh : -nh);
    n.fwd[2] = s.row + ((s.x + 1 == ns ? 0 : s.x + 1) >> 1);
    n.bwd[2] = s.row + ((s.x == 0 ? ns - 1 : s.x - 1) >> 1);
    return n;
}

// ---------------------------------------------------------------------------
// The site update of parity index h on parity p: reads 6 u0 links, 6
// neighbour G and G(n); writes G(n).  Shared by the single-lattice and the
// batched (serving) kernels, so a site computes the same bits in both.
// POUT projects the new G(n) onto SU(3) (in FP64, project_g) before storing
// it: the same bits as a reunitarization of this parity's G right after the
// sweep, since the projection reads rows 0 and 1 only.  PIN projects G(n)
// right after loading it and rebuilds its third row from the projected rows,
// as a store and reload would: the same bits as a reunitarization of this
// parity's G before the sweep (nothing else reads it in between).  STREAM
// reads u0 and G(n) with non-temporal loads (data read once per sweep);
// without it they are cached, for sweeps whose data stays in the Infinity
// Cache from one half-sweep to the next.
template <typename T, bool POUT = false, bool STREAM = true, bool PIN = false>
__device__ __forceinline__ void sweep_site(const cplx<T> *__restrict__ u0,
                                           cplx<T> *__restrict__ g, int ns, int Vh, int p,
                                           T omega, int hits, int h) {
    const Site s = site_of(h, p, ns);
    const Nbr nb = neighbours(s, h, ns);
    const cplx<T> *gq = g_field(static_cast<const cplx<T> *>(g), Vh, 1 - p);  // other parity
    cplx<T> *gp = g_field(g, Vh, p);
    // with PIN, G(n) is loaded and projected first and its rows 0 and 1 kept
    // through the W products (loaded later, next to its use, the projection
    // makes the compiler hold twice the registers)
    mat3<T> gn;
    if (PIN) {
        load_rows01<STREAM>(gn, static_cast<const cplx<T> *>(gp), h);
        project_g(gn);
        __builtin_amdgcn_sched_barrier(0);
    }
    // the six products are SU(3): rows 0, 1 from the stored rows of the left
    // factor, row 2 by closure (m_mac_nd_su3, m_mac_dagger_su3).  The four
    // matrices of a direction are loaded together, one memory round trip per
    // direction rather than the compiler's two (48^3 x 96: FP64 fix 1%, mixed
    // 2% faster).
    mat3<T> W;
    m_zero(W);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        constexpr int NU = U0Planes<T>::n;
        cplx<T> ruf[NU], rub[NU], rgf[kPlanes], rgb[kPlanes];
        load_raw<STREAM>(ruf, u0_field(u0, Vh, p, i), u0_tile_bytes<T>(h));
        load_raw<false>(rgf, gq, tile_bytes<T>(nb.fwd[i]));
        load_raw<STREAM>(rub, u0_field(u0, Vh, 1 - p, i), u0_tile_bytes<T>(nb.bwd[i]));
        load_raw<false>(rgb, gq, tile_bytes<T>(nb.bwd[i]));
        __builtin_amdgcn_sched_barrier(0);
        mat3<T> uf, gf, ub, gb;
        u0_rows01(uf, ruf);
        rows01(gf, rgf);
        m_fill_row2(gf);
        m_mac_nd_su3(W, uf, gf);  // u0_i(n) G(n+i)^dagger
        u0_rows01(ub, rub);
        m_fill_row2(ub);
        rows01(gb, rgb);
        m_mac_dagger_su3(W, gb, ub);  // (G(n-i) u0_i(n-i))^dagger
    }
    if (PIN)
        m_fill_row2(gn);  // the third row a store and reload would give
    else
        load_su3<STREAM>(gn, static_cast<const cplx<T> *>(gp), h);
    mat3<T> w, upd, out;
    m_mul(gn, W, w);
    la_or_update(w, hits, omega, upd);
    m_mul(upd, gn, out);
    if (POUT) project_g(out);
    store_su3<true>(out, gp, h);
}

// Workgroup order.  The hardware hands workgroup i of a launch to XCD i % 8,
// and each XCD has its own L2.  xcd_order renumbers the workgroups so that,
// in rounds of 8 x 64 workgroups, each XCD takes one contiguous run of the
// round (a launch of at most 512 workgroups: one contiguous eighth of it).
// The z-neighbour G of a site, ns^2 / 2 sites away (4.5 workgroups at 48^3),
// are then mostly in the same L2.  A bijection of [0, n).  The sweeps with
// cached loads use it: measured at 48^3 x 96 with cache-resident waves of
// single-timeslice launches (216 workgroups), an FP64 half-sweep is 11%
// faster and an FP32 one 10%.  The streaming sweeps keep the plain order,
// under which all XCDs stream one stretch of sites together (with xcd_order
// they were 1-2% slower; an XCD per contiguous eighth of a whole-lattice
// launch, 5% slower).
__device__ __forceinline__ unsigned xcd_order(unsigned i, unsigned n) {
    constexpr unsigned kXcd = 8, kRound = kXcd * 64;
    const unsigned base = i / kRound * kRound;
    const unsigned m = n - base < kRound ? n - base : kRound;  // workgroups in i's round
    const unsigned j = i - base, q = m / kXcd, r = m % kXcd, x = j % kXcd;
    return base + x * q + (x < r ? x : r) + j / kXcd;
}

// One red/black half-sweep on parity p over parity indices [h0, h0 + count).
template <typename T, bool POUT = false, bool STREAM = true, bool PIN = false>
__global__ __launch_bounds__(kBlock) void k_sweep(const cplx<T> *__restrict__ u0,
                                                  cplx<T> *__restrict__ g, int ns, int Vh,
                                                  int p, T omega, int hits, int h0,
                                                  int count) {
    const unsigned b = STREAM ? blockIdx.x : xcd_order(blockIdx.x, gridDim.x);
    const int tid = (int)(b * kBlock + threadIdx.x);
    if (tid >= count) return;
    sweep_site<T, POUT, STREAM, PIN>(u0, g, ns, Vh, p, omega, hits, h0 + tid);
}

// ---------------------------------------------------------------------------
// Per-site G kernels: one thread per site of the parities [p0, p0 + np),
// idx in [0, np Vh), parity p = p0 + idx / Vh.

// G <- identity (setFieldToIdentity, src/fields.c:64-100).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_g_identity(cplx<T> *__restrict__ g, int Vh) {
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= 2 * Vh) return;
    const int p = idx >= Vh, h = idx - p * Vh;
    mat3<T> m;
    m_identity(m);
    store_su3<false>(m, g_field(g, Vh, p), h);
}

// Projects G onto SU(3) at parity index h of parity p.
template <typename T>
__device__ __forceinline__ void reunit_site(cplx<T> *__restrict__ g, int Vh, int p, int h) {
    cplx<T> *base = g_field(g, Vh, p);
    const uint32_t off = tile_bytes<T>(h);
    mat3<T> m;
#pragma unroll
    for (int k = 0; k < kPlanes; k++) m.e[k] = *elem_ptr(base, off, k);
    project_g(m);
    store_su3<false>(m, base, h);
}

// Parity indices [h0, h0 + count) of parities [p0, p0 + np).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_g_reunit(cplx<T> *__restrict__ g, int Vh, int p0,
                                                     int np, int h0, int count) {
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= np * count) return;
    reunit_site<T>(g, Vh, p0 + idx / count, h0 + idx % count);
}

// FP32 -> FP64 hand-off of G: promote rows 0, 1 then project in FP64.
__device__ __forceinline__ void promote_site(const cplx<float> *__restrict__ gf,
                                             cplx<double> *__restrict__ g, int Vh, int p,
                                             int h) {
    const cplx<float> *src = g_field(gf, Vh, p);
    const uint32_t off = tile_bytes<float>(h);
    mat3<double> m;
#pragma unroll
    for (int k = 0; k < kPlanes; k++) {
        cplx<float> v = *elem_ptr(src, off, k);
        m.e[k] = c_make<double>(v.re, v.im);
    }
    m_project_su3(m);
    store_su3<false>(m, g_field(g, Vh, p), h);
}

__global__ __launch_bounds__(kBlock) void k_g_promote(const cplx<float> *__restrict__ gf,
                                                      cplx<double> *__restrict__ g, int Vh) {
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= 2 * Vh) return;
    const int p = idx >= Vh;
    promote_site(gf, g, Vh, p, idx - p * Vh);
}

// cfg (AoS FP64) spatial links -> u0 planes of type T (rows 0 and 1).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_build_u0(const cplx<double> *__restrict__ cfg,
                                                     cplx<T> *__restrict__ u0, int ns,
                                                     int Vh) {
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= 2 * Vh) return;
    const int p = idx >= Vh, h = idx - p * Vh;
    const Site s = site_of(h, p, ns);
    const int64_t lex = (int64_t)2 * s.row + s.x;
    const uint32_t off = u0_tile_bytes<T>(h);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const cplx<double> *src = cfg + (lex * 4 + 1 + i) * 9;
        cplx<T> *dst = u0_field(u0, Vh, p, i);
        cplx<double> e[kPlanes];
#pragma unroll
        for (int k = 0; k < kPlanes; k++) e[k] = src[k];
        if constexpr (U0Planes<T>::n == 5) {
            // row 1 without the element under row 0's pivot (u0_unpack)
            const int piv = u0_pivot(e);
            const cplx<double> s0 = csel(piv == 0, e[4], e[3]);
            const cplx<double> s1 = csel(piv == 2, e[4], e[5]);
            e[3] = s0;
            e[4] = s1;
        }
#pragma unroll
        for (int k = 0; k < U0Planes<T>::n; k++)
            *elem_ptr(dst, off, k) = c_make<T>((T)e[k].re, (T)e[k].im);
    }
}

// cfg <- G(n) cfg G(n+mu)^dagger for all four directions
// (applyGaugeTransformationU, src/fields.c:375-435); with PROJECT every link
// is then projected onto SU(3), the reunitarization of U that ends a
// converged gaugefixOverrelaxation (src/gauge_fixing.c:822).  One thread per
// link, 64 sites (256 links, 36 KiB) per workgroup: the links are read from
// and written back to the AoS array through LDS in consecutive 16-byte
// pieces, so that every wave's loads and stores cover whole cache lines (one
// thread per site reading its 576 bytes directly left the accesses of a wave
// 576 bytes apart).
template <bool PROJECT>
__global__ __launch_bounds__(kBlock) void k_materialize(cplx<double> *__restrict__ cfg,
                                                        const cplx<double> *__restrict__ g,
                                                        int ns, int nt, int Vh) {
    __shared__ double2 sh[kBlock * 9];
    const int64_t link0 = (int64_t)blockIdx.x * kBlock;
    const int nl = (int)min<int64_t>(kBlock, 8 * (int64_t)Vh - link0);  // links here
    double2 *base = reinterpret_cast<double2 *>(cfg) + link0 * 9;
    for (int k = threadIdx.x; k < nl * 9; k += kBlock) sh[k] = base[k];
    __syncthreads();
    if ((int)threadIdx.x < nl) {
        const int64_t link = link0 + threadIdx.x;
        const int64_t lex = link >> 2;
        const int mu = (int)(link & 3);
        const int x = (int)(lex % ns);
        int64_t r = lex / ns;
        const int y = (int)(r % ns);
        r /= ns;
        const int z = (int)(r % ns);
        const int t = (int)(r / ns);
        const int p = (x + y + z + t) & 1;
        int c[4] = {t, z, y, x};
        c[mu] += 1;
        if (c[mu] == (mu == 0 ? nt : ns)) c[mu] = 0;
        const int64_t ln = (((int64_t)c[0] * ns + c[1]) * ns + c[2]) * ns + c[3];
        mat3<double> gn, gm, u, tmp, out;
        load_su3<false>(gn, g_field(g, Vh, p), (int)(lex >> 1));
        load_su3<false>(gm, g_field(g, Vh, 1 - p), (int)(ln >> 1));
        double2 *mine = sh + threadIdx.x * 9;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            u.e[k].re = mine[k].x;
            u.e[k].im = mine[k].y;
        }
        m_mul(gn, u, tmp);
        m_mul_nd(tmp, gm, out);
        if (PROJECT) m_project_su3(out);
#pragma unroll
        for (int k = 0; k < 9; k++) mine[k] = double2{out.e[k].re, out.e[k].im};
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nl * 9; k += kBlock) base[k] = sh[k];
}

// Projects n AoS FP64 matrices onto SU(3); sets *flag when a row was null
// (reunitarizeField's -1).
__global__ __launch_bounds__(kBlock) void k_reunit_aos(cplx<double> *__restrict__ m_,
                                                       int64_t n, int *flag) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    cplx<double> *ptr = m_ + i * 9;
    mat3<double> m;
#pragma unroll
    for (int k = 0; k < 9; k++) m.e[k] = ptr[k];
    if (!m_project_su3(m)) atomicOr(flag, 1);
#pragma unroll
    for (int k = 0; k < 9; k++) ptr[k] = m.e[k];
}

__global__ __launch_bounds__(kBlock) void k_f32_to_f64(const float *__restrict__ in,
                                                       double *__restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = (double)in[i];
}

// SoA G (FP64) -> AoS image (site-major, full 3x3).
__global__ __launch_bounds__(kBlock) void k_g_to_aos(const cplx<double> *__restrict__ g,
                                                     cplx<double> *__restrict__ out, int ns,
                                                     int Vh) {
    const int64_t lex = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (lex >= 2 * (int64_t)Vh) return;
    const int x = (int)(lex % ns);
    int64_t r = lex / ns;
    const int y = (int)(r % ns);
    r /= ns;
    const int z = (int)(r % ns);
    const int t = (int)(r / ns);
    const int p = (x + y + z + t) & 1;
    mat3<double> m;
    load_su3<false>(m, g_field(g, Vh, p), (int)(lex >> 1));
#pragma unroll
    for (int k = 0; k < 9; k++) out[lex * 9 + k] = m.e[k];
}

// ---------------------------------------------------------------------------
// Observables.  Per site: D = sum_i [(U_i(n) - U_i(n)^dag) - (U_i(n-i) - U_i(n-i)^dag)]
// made traceless, div A = -(i/2) D (divergenceA/calculateA,
// src/four_potential.c:35-113); e2 density = sum_a c_a^2 with the Gell-Mann
// components of decomposeAlgebraSU3 (src/SU3_ops.c:1055-1107); theta density
// = Re Tr(divA divA^dag) (calculateTheta, src/gauge_fixing.c:341-384); and
// Re Tr sum_i U_i(n) for F (calculateF, :289-337).  All arithmetic in FP64.

struct Obs {
    double e2, theta, f;
};

__device__ __forceinline__ void accum_link(mat3<double> &D, const mat3<double> &u, double sgn) {
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const cplx<double> uab = u.e[3 * a + b], uba = u.e[3 * b + a];
            D.e[3 * a + b].re += sgn * (uab.re - uba.re);
            D.e[3 * a + b].im += sgn * (uab.im + uba.im);
        }
}

__device__ __forceinline__ Obs site_obs(mat3<double> &D, double f) {
    const double tr = (D.e[0].im + D.e[4].im + D.e[8].im) / 3.0;
    D.e[0].im -= tr;
    D.e[4].im -= tr;
    D.e[8].im -= tr;
    // A = -(i/2) D : A.re = D.im / 2, A.im = -D.re / 2
    mat3<double> A;
#pragma unroll
    for (int k = 0; k < 9; k++) A.e[k] = c_make<double>(0.5 * D.e[k].im, -0.5 * D.e[k].re);
    double c[8];
    c[0] = A.e[1].re + A.e[3].re;
    c[1] = -A.e[1].im + A.e[3].im;
    c[2] = A.e[0].re - A.e[4].re;
    c[3] = A.e[2].re + A.e[6].re;
    c[4] = -A.e[2].im + A.e[6].im;
    c[5] = A.e[5].re + A.e[7].re;
    c[6] = -A.e[5].im + A.e[7].im;
    c[7] = (A.e[0].re + A.e[4].re - 2.0 * A.e[8].re) * 0.57735026918962573;
    Obs o;
    o.e2 = 0;
#pragma unroll
    for (int a = 0; a < 8; a++) o.e2 += c[a] * c[a];
    o.theta = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) o.theta += A.e[k].re * A.e[k].re + A.e[k].im * A.e[k].im;
    o.f = f;
    return o;
}

template <typename T>
__device__ __forceinline__ void load_su3_d(mat3<double> &m, const cplx<T> *__restrict__ field,
                                           int h) {
    const uint32_t off = tile_bytes<T>(h);
#pragma unroll
    for (int k = 0; k < kPlanes; k++) {
        const cplx<T> v = *elem_ptr(field, off, k);
        m.e[k] = c_make<double>((double)v.re, (double)v.im);
    }
    m_fill_row2(m);
}

// Deterministic block reduction of three doubles; writes part[3*block + j].
__device__ __forceinline__ void block_reduce3(Obs o, double *__restrict__ part, unsigned blk) {
    __shared__ double sh[3][kBlock / 64];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        o.e2 += __shfl_xor(o.e2, off, 64);
        o.theta += __shfl_xor(o.theta, off, 64);
        o.f += __shfl_xor(o.f, off, 64);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][wv] = o.e2;
        sh[1][wv] = o.theta;
        sh[2][wv] = o.f;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double s = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; w++) s += sh[threadIdx.x][w];
        part[3 * blk + threadIdx.x] = s;
    }
}

// Observables of the factored state at site index idx in [0, 2 Vh) (parity
// idx >= Vh); u0 of type TU, G of type TG, arithmetic in FP64.  They come
// from the site's w = G(n) W(n), formed as the sweep forms it:
// w = sum_i [U_i(n) + U_i(n-i)^dagger], so w - w^dagger is the divergence
// sum_i [(U - U^dagger)_i(n) - (U - U^dagger)_i(n-i)] that e2 and theta
// measure (calculateA, divergenceA: src/four_potential.c), and
// Re Tr w / 2 summed over all sites is the sum over links of Re Tr U that F
// measures (calculateF, src/gauge_fixing.c).  Seven products instead of the
// twelve of forming every link.
template <typename TU, typename TG>
__device__ __forceinline__ Obs obs_site(const cplx<TU> *__restrict__ u0,
                                        const cplx<TG> *__restrict__ g, int ns, int Vh,
                                        int idx) {
    Obs o = {0, 0, 0};
    if (idx < 2 * Vh) {
        const int p = idx >= Vh, h = idx - p * Vh;
        const Site s = site_of(h, p, ns);
        const Nbr nb = neighbours(s, h, ns);
        const cplx<TG> *gq = g_field(g, Vh, 1 - p);
        mat3<double> W;
        m_zero(W);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            mat3<double> uf, gf, ub, gb;
            load_u0_d(uf, u0_field(u0, Vh, p, i), h);
            load_su3_d(gf, gq, nb.fwd[i]);
            load_u0_d(ub, u0_field(u0, Vh, 1 - p, i), nb.bwd[i]);
            load_su3_d(gb, gq, nb.bwd[i]);
            m_mac_nd_su3(W, uf, gf);      // u0_i(n) G(n+i)^dagger
            m_mac_dagger_su3(W, gb, ub);  // (G(n-i) u0_i(n-i))^dagger
            // one direction's loads at a time (hoisting all of them took 2
            // waves/SIMD)
            __builtin_amdgcn_sched_barrier(0);
        }
        mat3<double> gn, w, D;
        load_su3_d(gn, g_field(g, Vh, p), h);
        m_mul(gn, W, w);
        m_zero(D);
        accum_link(D, w, 1.0);
        o = site_obs(D, 0.5 * (w.e[0].re + w.e[4].re + w.e[8].re));
    }
    return o;
}

// One thread per site, both parities; per-block partial sums in part, in
// site order.  The workgroups run in XCD-contiguous order (xcd_order), so that
// the neighbour G and backward u0 links a site reads were mostly brought into
// its XCD's L2 by the sites next to it.
template <typename TU, typename TG>
__global__ __launch_bounds__(kBlock) void k_obs_factored(const cplx<TU> *__restrict__ u0,
                                                         const cplx<TG> *__restrict__ g,
                                                         int ns, int Vh,
                                                         double *__restrict__ part) {
    const unsigned b = xcd_order(blockIdx.x, gridDim.x);
    const int idx = (int)(b * kBlock + threadIdx.x);
    block_reduce3(obs_site<TU, TG>(u0, g, ns, Vh, idx), part, b);
}

// Observables of an AoS FP64 configuration; one thread per site.
__global__ __launch_bounds__(kBlock) void k_obs_aos(const cplx<double> *__restrict__ cfg,
                                                    int ns, int64_t V,
                                                    double *__restrict__ part) {
    const int64_t lex = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    Obs o = {0, 0, 0};
    if (lex < V) {
        const int x = (int)(lex % ns);
        int64_t r = lex / ns;
        const int y = (int)(r % ns);
        r /= ns;
        const int z = (int)(r % ns);
        const int t = (int)(r / ns);
        mat3<double> D;
        m_zero(D);
        double f = 0;
#pragma unroll
        for (int mu = 1; mu <= 3; mu++) {
            int c[4] = {t, z, y, x};
            c[mu] = c[mu] == 0 ? ns - 1 : c[mu] - 1;
            const int64_t lb = (((int64_t)c[0] * ns + c[1]) * ns + c[2]) * ns + c[3];
            mat3<double> u;
            const cplx<double> *pf = cfg + (lex * 4 + mu) * 9;
#pragma unroll
            for (int k = 0; k < 9; k++) u.e[k] = pf[k];
            accum_link(D, u, 1.0);
            f += u.e[0].re + u.e[4].re + u.e[8].re;
            const cplx<double> *pb = cfg + (lb * 4 + mu) * 9;
#pragma unroll
            for (int k = 0; k < 9; k++) u.e[k] = pb[k];
            accum_link(D, u, -1.0);
        }
        o = site_obs(D, f);
    }
    block_reduce3(o, part, blockIdx.x);
}

// Plaquettes of an AoS FP64 configuration (TrPlaquette, src/measurement.c:38-84):
// Re Tr[U_mu(n) U_nu(n+mu) U_mu(n+nu)^dag U_nu(n)^dag] / 3, spatial planes
// (XY, XZ, YZ) into slot 0 and temporal planes (TZ, TY, TX) into slot 1.
__global__ __launch_bounds__(kBlock) void k_plaquette_aos(const cplx<double> *__restrict__ cfg,
                                                          int ns, int nt, int64_t V,
                                                          double *__restrict__ part) {
    const int64_t lex = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    Obs o = {0, 0, 0};
    if (lex < V) {
        const int x = (int)(lex % ns);
        int64_t r = lex / ns;
        const int y = (int)(r % ns);
        r /= ns;
        const int z = (int)(r % ns);
        const int t = (int)(r / ns);
        const int planes[6][2] = {{3, 2}, {3, 1}, {2, 1}, {0, 1}, {0, 2}, {0, 3}};
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int mu = planes[q][0], nu = planes[q][1];
            int cm[4] = {t, z, y, x}, cn[4] = {t, z, y, x};
            cm[mu] += 1;
            if (cm[mu] == (mu == 0 ? nt : ns)) cm[mu] = 0;
            cn[nu] += 1;
            if (cn[nu] == (nu == 0 ? nt : ns)) cn[nu] = 0;
            const int64_t lm = (((int64_t)cm[0] * ns + cm[1]) * ns + cm[2]) * ns + cm[3];
            const int64_t ln = (((int64_t)cn[0] * ns + cn[1]) * ns + cn[2]) * ns + cn[3];
            mat3<double> a, b, c, d, ab, abc;
            const cplx<double> *pa = cfg + (lex * 4 + mu) * 9, *pb = cfg + (lm * 4 + nu) * 9;
            const cplx<double> *pc = cfg + (ln * 4 + mu) * 9, *pd = cfg + (lex * 4 + nu) * 9;
#pragma unroll
            for (int k = 0; k < 9; k++) {
                a.e[k] = pa[k];
                b.e[k] = pb[k];
                c.e[k] = pc[k];
                d.e[k] = pd[k];
            }
            m_mul(a, b, ab);
            m_mul_nd(ab, c, abc);
            // Re Tr(abc d^dag) = sum_ij Re(abc_ij conj(d_ij))
            double tr = 0;
#pragma unroll
            for (int k = 0; k < 9; k++) tr += abc.e[k].re * d.e[k].re + abc.e[k].im * d.e[k].im;
            if (q < 3)
                o.e2 += tr / 3.0;
            else
                o.theta += tr / 3.0;
        }
    }
    block_reduce3(o, part, blockIdx.x);
}

// Streaming copy of n 16-byte elements (bandwidth probe / counter calibration).
__global__ __launch_bounds__(kBlock) void k_copy16(const double2 *__restrict__ in,
                                                   double2 *__restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// Read-rate probe: `passes` reads of gridDim.x contiguous chunks of `chunk`
// 16-byte elements, eight loads in flight per lane.  In pass p workgroup b
// reads chunk (b + p) mod gridDim.x, so no workgroup reads the same bytes twice
// in a row and the next reader of a chunk runs on another XCD: every pass
// streams the buffer from where it resides (the Infinity Cache up to 256 MiB,
// HBM above), not from an L2.  The sum is stored only if it takes a value a
// zeroed buffer cannot give.
__global__ __launch_bounds__(kBlock) void k_read16(const double2 *__restrict__ in, int64_t chunk,
                                                   int passes, double2 *__restrict__ sink) {
    constexpr int kInFlight = 8;
    double2 acc = {0.0, 0.0};
    for (int p = 0; p < passes; p++) {
        asm volatile("" ::: "memory");  // every pass loads again
        const double2 *c = in + (int64_t)((blockIdx.x + (unsigned)p) % gridDim.x) * chunk;
        int64_t j = threadIdx.x;
        for (; j + (kInFlight - 1) * kBlock < chunk; j += kInFlight * kBlock) {
            double2 v[kInFlight];
#pragma unroll
            for (int u = 0; u < kInFlight; u++) v[u] = c[j + u * kBlock];
#pragma unroll
            for (int u = 0; u < kInFlight; u++) acc += v[u];
        }
        for (; j < chunk; j += kBlock) acc += c[j];
    }
    if (acc.x == 1.2345e-300) sink[(int64_t)blockIdx.x * kBlock + threadIdx.x] = acc;
}

// Sums part[3*b + j] over b in fixed order with one block; out[j].
__device__ __forceinline__ void reduce_parts(const double *__restrict__ part, int nblocks,
                                             double *__restrict__ out) {
    Obs o = {0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += kBlock) {
        o.e2 += part[3 * b];
        o.theta += part[3 * b + 1];
        o.f += part[3 * b + 2];
    }
    block_reduce3(o, out, 0);
}

__global__ __launch_bounds__(kBlock) void k_reduce_final(const double *__restrict__ part,
                                                         int nblocks, double *__restrict__ out) {
    reduce_parts(part, nblocks, out);
}

// ---------------------------------------------------------------------------
// Batched kernels for the serving scheduler: many lattices per launch.  A
// work item is one 256-site chunk of one lattice; the per-site code is the
// same device function as in the single-lattice kernels, and reductions use
// the same block decomposition and order, so every lattice gets the same
// bits alone or packed with others.

struct BatchJob {
    const void *u0;   // cplx<T> planes of the lattice's u0 (T = sweep launch type)
    void *g;          // cplx<T> planes of its G (T = sweep launch type)
    const void *u0d;  // FP64 u0 (observables always read FP64 links)
    const void *gd;   // FP64 G, for FP64-state observables
    const void *gf;   // FP32 G (MIXED only, else null), for FP32-state observables
    double omega;
    int ns, Vh, hits;
    int part_off;     // first partial-sum slot of this lattice (observables)
};

// One workgroup per item = (job index, first parity index, parity, unused).
// (Looping a workgroup over several items measured 17% slower.)  STREAM as in
// sweep_site.
template <typename T, bool STREAM = true>
__global__ __launch_bounds__(kBlock) void k_sweep_batch(const BatchJob *__restrict__ jobs,
                                                        const int4 *__restrict__ items,
                                                        int nitems) {
    const int i = (int)(STREAM ? blockIdx.x : xcd_order(blockIdx.x, gridDim.x));
    if (i >= nitems) return;
    const int4 it = items[i];
    const BatchJob J = jobs[it.x];
    const int h = it.y + (int)threadIdx.x;
    if (h < J.Vh)
        sweep_site<T, false, STREAM>(static_cast<const cplx<T> *>(J.u0),
                                     static_cast<cplx<T> *>(J.g), J.ns, J.Vh, it.z, (T)J.omega,
                                     J.hits, h);
}

// item = (job index, block index within the lattice's [0, 2 Vh) range, fp32 G?, _)
__global__ __launch_bounds__(kBlock) void k_obs_batch(const BatchJob *__restrict__ jobs,
                                                      const int4 *__restrict__ items,
                                                      double *__restrict__ part) {
    const int4 it = items[blockIdx.x];
    const BatchJob J = jobs[it.x];
    const int idx = it.y * kBlock + (int)threadIdx.x;
    const cplx<double> *u0 = static_cast<const cplx<double> *>(J.u0d);
    Obs o = it.z ? obs_site<double, float>(u0, static_cast<const cplx<float> *>(J.gf), J.ns,
                                           J.Vh, idx)
                 : obs_site<double, double>(u0, static_cast<const cplx<double> *>(J.gd), J.ns,
                                            J.Vh, idx);
    block_reduce3(o, part + 3 * (int64_t)J.part_off, it.y);
}

// One block per evaluated lattice: jobs[sel[k]] partials -> out[3k + j].
__global__ __launch_bounds__(kBlock) void k_reduce_batch(const BatchJob *__restrict__ jobs,
                                                         const int *__restrict__ sel,
                                                         const double *__restrict__ part,
                                                         double *__restrict__ out) {
    const BatchJob J = jobs[sel[blockIdx.x]];
    const int nb = (2 * J.Vh + kBlock - 1) / kBlock;
    reduce_parts(part + 3 * (int64_t)J.part_off, nb, out + 3 * blockIdx.x);
}

// item = (job index, block index within [0, 2 Vh), _, _): project G
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reunit_batch(const BatchJob *__restrict__ jobs,
                                                         const int4 *__restrict__ items) {
    const int4 it = items[blockIdx.x];
    const BatchJob J = jobs[it.x];
    const int idx = it.y * kBlock + (int)threadIdx.x;
    if (idx >= 2 * J.Vh) return;
    const int p = idx >= J.Vh;
    reunit_site<T>(static_cast<cplx<T> *>(J.g), J.Vh, p, idx - p * J.Vh);
}

// ---------------------------------------------------------------------------
// Device-side synthetic configuration: the same construction as the host
// generator (host_gen.c: splitmix64 draws, Box-Muller normals, Gram-Schmidt
// Haar SU(3), exp(i eps H) by scaling and squaring of a 16-term Taylor
// polynomial, U_mu(n) = g(n) exp(i eps H_mu(n)) g(n+mu)^dagger).  The device
// libm and FMA contraction make it agree with the host one to rounding, not
// bit for bit; the serving scheduler uses it to synthesise request inputs.

__device__ __forceinline__ uint64_t gen_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ void gen_normals(uint64_t base, uint64_t site, unsigned k,
                                            double &a, double &b) {
    const uint64_t gold = 0x9E3779B97F4A7C15ULL;
    const uint64_t z1 = gen_mix64(base + (site * 64 + k + 1) * gold);
    const uint64_t z2 = gen_mix64(base + (site * 64 + k + 2) * gold);
    const double u1 = (double)((z1 >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)((z2 >> 11) + 1) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincos(6.283185307179586 * u2, &sn, &cs);
    a = r * cs;
    b = r * sn;
}

__device__ void gen_haar(uint64_t base, uint64_t site, mat3<double> &g) {
    double v[18];
#pragma unroll
    for (unsigned k = 0; k < 18; k += 2) gen_normals(base, site, k, v[k], v[k + 1]);
#pragma unroll
    for (int k = 0; k < 9; k++) g.e[k] = c_make<double>(v[2 * k], v[2 * k + 1]);
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int q = 0; q < i; q++) {
            cplx<double> ip = c_make<double>(0, 0);
#pragma unroll
            for (int j = 0; j < 3; j++) c_mac_cb(ip, g.e[3 * i + j], g.e[3 * q + j]);
            // ip = <r_q, r_i> = sum conj(r_q) r_i
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const cplx<double> rq = g.e[3 * q + j];
                g.e[3 * i + j].re -= ip.re * rq.re - ip.im * rq.im;
                g.e[3 * i + j].im -= ip.re * rq.im + ip.im * rq.re;
            }
        }
        double n = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) n += g.e[3 * i + j].re * g.e[3 * i + j].re +
                                         g.e[3 * i + j].im * g.e[3 * i + j].im;
        n = 1.0 / sqrt(n);
#pragma unroll
        for (int j = 0; j < 3; j++) g.e[3 * i + j] = c_make(g.e[3 * i + j].re * n, g.e[3 * i + j].im * n);
    }
    const cplx<double> d = m_det(g);
    double sn, cs;
    sincos(-atan2(d.im, d.re) / 3.0, &sn, &cs);
#pragma unroll
    for (int k = 0; k < 9; k++)
        g.e[k] = c_make(g.e[k].re * cs - g.e[k].im * sn, g.e[k].re * sn + g.e[k].im * cs);
}

__device__ void gen_expi(uint64_t base, uint64_t site, int mu, double eps, mat3<double> &E) {
    double d[10];
    const unsigned k0 = 18 + 10 * (unsigned)mu;
#pragma unroll
    for (unsigned k = 0; k < 10; k += 2) gen_normals(base, site, k0 + k, d[k], d[k + 1]);
    const double s2 = 0.7071067811865476;
    const double mean = (d[0] + d[1] + d[2]) / 3.0;
    mat3<double> H;
    H.e[0] = c_make(d[0] - mean, 0.0);
    H.e[4] = c_make(d[1] - mean, 0.0);
    H.e[8] = c_make(d[2] - mean, 0.0);
    H.e[1] = c_make(d[3] * s2, d[4] * s2);
    H.e[2] = c_make(d[5] * s2, d[6] * s2);
    H.e[5] = c_make(d[7] * s2, d[8] * s2);
    H.e[3] = c_make(H.e[1].re, -H.e[1].im);
    H.e[6] = c_make(H.e[2].re, -H.e[2].im);
    H.e[7] = c_make(H.e[5].re, -H.e[5].im);
    double nrm = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double rs = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) rs += hypot(H.e[3 * i + j].re, H.e[3 * i + j].im);
        nrm = rs > nrm ? rs : nrm;
    }
    nrm *= fabs(eps);
    int sq = 0;
    while (nrm > 0.5 && sq < 64) {
        nrm *= 0.5;
        sq++;
    }
    const double scale = ldexp(eps, -sq);
    mat3<double> A, T;
    // A = i scale H
#pragma unroll
    for (int k = 0; k < 9; k++) A.e[k] = c_make(-scale * H.e[k].im, scale * H.e[k].re);
    m_identity(E);
    for (int n = 16; n >= 1; n--) {
        m_mul(A, E, T);
#pragma unroll
        for (int k = 0; k < 9; k++)
            E.e[k] = c_make(T.e[k].re / n + ((k % 4 == 0) ? 1.0 : 0.0), T.e[k].im / n);
    }
    for (int s = 0; s < sq; s++) {
        m_mul(E, E, T);
        E = T;
    }
    m_project_su3(E);
}

// g(n) for every site into tmp (AoS, 9 complex per site).
__global__ __launch_bounds__(kBlock) void k_gen_haar(uint64_t base, int64_t V,
                                                     cplx<double> *__restrict__ tmp) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= V) return;
    mat3<double> g;
    gen_haar(base, (uint64_t)s, g);
#pragma unroll
    for (int k = 0; k < 9; k++) tmp[s * 9 + k] = g.e[k];
}

__global__ __launch_bounds__(kBlock) void k_gen_links(uint64_t base, double eps, int ns, int nt,
                                                      const cplx<double> *__restrict__ tmp,
                                                      cplx<double> *__restrict__ cfg) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t V = (int64_t)ns * ns * ns * nt;
    if (s >= V) return;
    const int x = (int)(s % ns), y = (int)((s / ns) % ns), z = (int)((s / ns / ns) % ns);
    const int t = (int)(s / ((int64_t)ns * ns * ns));
    mat3<double> gs;
#pragma unroll
    for (int k = 0; k < 9; k++) gs.e[k] = tmp[s * 9 + k];
    for (int mu = 0; mu < 4; mu++) {
        int c[4] = {t, z, y, x};
        c[mu] += 1;
        if (c[mu] == (mu == 0 ? nt : ns)) c[mu] = 0;
        const int64_t sn = (((int64_t)c[0] * ns + c[1]) * ns + c[2]) * ns + c[3];
        mat3<double> X, T1, gn, out;
        gen_expi(base, (uint64_t)s, mu, eps, X);
#pragma unroll
        for (int k = 0; k < 9; k++) gn.e[k] = tmp[sn * 9 + k];
        m_mul(gs, X, T1);
        m_mul_nd(T1, gn, out);
#pragma unroll
        for (int k = 0; k < 9; k++) cfg[(s * 4 + mu) * 9 + k] = out.e[k];
    }
}

}  // namespace
}  // namespace gfc
