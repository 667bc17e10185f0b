// Level-0 (fine lattice, 2 spin components) kernels.
//
// Vectors are [nrhs][ns][2] complex; site n = x*nt + t.  One thread per site,
// 256 threads per workgroup; the right-hand side is blockIdx.y.
#pragma once
#include "common.hpp"
#include "layout.hpp"

namespace amgs {


// Workgroup -> tile map that keeps consecutive tiles on one XCD: the hardware
// deals workgroups round-robin over the 8 XCDs, so tile = (wg % 8) * per + wg / 8
// gives each XCD a contiguous run of lattice rows (its x-neighbours stay in
// the same L2).  Identity when the grid is not a multiple of 8.
__device__ __forceinline__ int xcd_tile(int wg, int nwg) {
    if (nwg & 7) return wg;
    return (wg & 7) * (nwg >> 3) + (wg >> 3);
}

// ---------------------------------------------------------------------------
// Wilson-Dirac stencil at one site (PDF eq.6; src/dirac_operator.cpp:3-54)
//   D:  out = (m0+2) psi - 1/2 sum_mu [(1-s_mu) U_mu(n) psi(n+mu) + (1+s_mu) U_mu(n-mu)^* psi(n-mu)]
//   D^dagger swaps the projectors.  s_0 = sigma_x (time), s_1 = sigma_y (space),
//   antiperiodic in time.  With spin projection every hop costs one complex
//   product:  A = f0 (p0-p1)(r0), B = f1 (p0+ip1)(r1), C = b0 (p0+p1)(l0),
//   E = b1 (p0-ip1)(l1);  out0 = mm p0 - (A+B+C+E)/2,  out1 = mm p1 - (-A-iB+C+iE)/2
//   (for D^dagger the same with the forward/backward hops exchanged).
// MASK(n') says whether neighbour n' contributes (zero Dirichlet for SAP blocks).
// ---------------------------------------------------------------------------
// Site readers: the two spin components of site n of one vector.  Site
// indices are non-negative and below 2^28, so element offsets are formed as
// unsigned 32-bit values: with a uniform base the loads then take the
// scalar-base + 32-bit-offset form instead of 64-bit address arithmetic.
template <typename R> struct SpinPtr {  // complex FP32/FP64 storage [ns][2]
    const cplx<R>* p;
    __device__ __forceinline__ void operator()(int n, cplx<R>& a, cplx<R>& b) const {
        const unsigned u = 2u * (unsigned)n;
        a = p[u];
        b = p[u + 1u];
    }
};
// FP16 storage: (re0, im0, re1, im1) of a site in 8 bytes; values are read as
// stored (the power-of-two scale lives in the Krylov scale factors)
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
struct SpinH16 {
    const f16x4* p;
    __device__ __forceinline__ f16x4 raw(int n) const { return p[(unsigned)n]; }
    __device__ __forceinline__ void operator()(int n, c32& a, c32& b) const {
        const f16x4 v = p[(unsigned)n];
        a = c32{float(v.x), float(v.y)};
        b = c32{float(v.z), float(v.w)};
    }
};
__device__ __forceinline__ f16x4 to_h16(c32 a, c32 b) {
    return f16x4{(_Float16)a.re, (_Float16)a.im, (_Float16)b.re, (_Float16)b.im};
}

// Spin projections of FP16 site entries (x, y) = p0, (z, w) = p1 straight
// from the halves: b + s a as one mixed-precision FMA, the exact sum rounded
// once, as in FP32.  The +-1 pass through an empty asm: folded to an add, the
// FMA would become two conversions and an add.
struct H16Proj {
    float p1, m1;
    __device__ __forceinline__ H16Proj() : p1(1.f), m1(-1.f) { __asm__("" : "+s"(p1), "+s"(m1)); }
    __device__ __forceinline__ static float pm(_Float16 a, float sg, _Float16 b) { return fmaf((float)a, sg, (float)b); }
    __device__ __forceinline__ c32 sum(f16x4 v, bool neg) const {  // p0 +- p1
        return c32{pm(v.z, neg ? m1 : p1, v.x), pm(v.w, neg ? m1 : p1, v.y)};
    }
    __device__ __forceinline__ c32 sumi(f16x4 v, bool neg) const {  // p0 +- i p1
        return c32{pm(v.w, neg ? p1 : m1, v.x), pm(v.z, neg ? m1 : p1, v.y)};
    }
};

// The hop terms of a site from its neighbours' spin projections (right t,
// right x, left t, left x: D takes p0 - p1, p0 + i p1, p0 + p1, p0 - i p1,
// D^dagger the opposite signs), its links (with the antiperiodic signs
// applied) and the neighbour masks:
//   out0 = mm c0 - (A+B+C+E)/2,  out1 = mm c1 - (-A-iB+C+iE)/2.
template <typename R, bool DAG>
__device__ __forceinline__ void wilson_combine(cplx<R> c0, cplx<R> c1, cplx<R> dr0, cplx<R> dr1, cplx<R> dl0,
                                               cplx<R> dl1, cplx<R> f0, cplx<R> f1, cplx<R> b0, cplx<R> b1, bool mr0,
                                               bool mr1, bool ml0, bool ml1, R mm, cplx<R>& o0, cplx<R>& o1) {
    const cplx<R> z = czero<R>();
    cplx<R> A, B, C, E;
    if (!DAG) {
        A = mr0 ? f0 * dr0 : z;
        B = mr1 ? f1 * dr1 : z;
        C = ml0 ? b0 * dl0 : z;
        E = ml1 ? b1 * dl1 : z;
    } else {
        A = ml0 ? b0 * dl0 : z;
        B = ml1 ? b1 * dl1 : z;
        C = mr0 ? f0 * dr0 : z;
        E = mr1 ? f1 * dr1 : z;
    }
    const R h = R(0.5);
    if constexpr (std::is_same<R, float>::value) {
        const cplx<R> s0 = (A + B) + (C + E), s1 = (C - A) + mul_i(E - B);
        o0 = cplx<R>{fmaf(-h, s0.re, mm * c0.re), fmaf(-h, s0.im, mm * c0.im)};
        o1 = cplx<R>{fmaf(-h, s1.re, mm * c1.re), fmaf(-h, s1.im, mm * c1.im)};
    } else {
        o0 = mm * c0 - h * (A + B + C + E);
        o1 = mm * c1 - h * ((C - A) + mul_i(E - B));
    }
}

// Stencil core on given neighbours r0 = n+t, r1 = n+x, l0 = n-t, l1 = n-x,
// their masks and the antiperiodic wraps of the time hops; c0, c1 return the
// site's own input values.
template <typename R, bool DAG, typename In>
__device__ __forceinline__ void wilson_core(const cplx<R>* __restrict__ U, In in, int n, int r0, int r1, int l0, int l1,
                                            bool mr0, bool mr1, bool ml0, bool ml1, bool wr, bool wl, R mm,
                                            cplx<R>& o0, cplx<R>& o1, cplx<R>& c0, cplx<R>& c1) {
    in(n, c0, c1);
    const unsigned un = 2u * (unsigned)n;
    cplx<R> f0 = U[un], f1 = U[un + 1u];
    cplx<R> b0 = conj(U[2u * (unsigned)l0]), b1 = conj(U[2u * (unsigned)l1 + 1u]);
    if (wr) f0 = -f0;
    if (wl) b0 = -b0;
    // neighbour reads without branches: a masked-out neighbour reads the
    // site's own (cached) entries and its term is replaced by zero, so the four
    // loads are in flight together
    // spin projections of the four neighbours (right t, right x, left t,
    // left x): D takes p0 - p1, p0 + i p1, p0 + p1, p0 - i p1; D^dagger the
    // opposite signs
    cplx<R> dr0, dr1, dl0, dl1;
    if constexpr (std::is_same<In, SpinH16>::value) {
        const f16x4 h0 = in.raw(mr0 ? r0 : n), h1 = in.raw(mr1 ? r1 : n);
        const f16x4 h2 = in.raw(ml0 ? l0 : n), h3 = in.raw(ml1 ? l1 : n);
        const H16Proj P;
        dr0 = P.sum(h0, !DAG);
        dr1 = P.sumi(h1, DAG);
        dl0 = P.sum(h2, DAG);
        dl1 = P.sumi(h3, !DAG);
    } else {
        cplx<R> q00, q01, q10, q11, q20, q21, q30, q31;
        in(mr0 ? r0 : n, q00, q01);
        in(mr1 ? r1 : n, q10, q11);
        in(ml0 ? l0 : n, q20, q21);
        in(ml1 ? l1 : n, q30, q31);
        if (!DAG) {
            dr0 = q00 - q01;
            dr1 = q10 + mul_i(q11);
            dl0 = q20 + q21;
            dl1 = q30 - mul_i(q31);
        } else {
            dr0 = q00 + q01;
            dr1 = q10 - mul_i(q11);
            dl0 = q20 - q21;
            dl1 = q30 + mul_i(q31);
        }
    }
    wilson_combine<R, DAG>(c0, c1, dr0, dr1, dl0, dl1, f0, f1, b0, b1, mr0, mr1, ml0, ml1, mm, o0, o1);
}

template <typename R, bool DAG, typename Mask, typename In>
__device__ __forceinline__ void wilson_site_in(const cplx<R>* __restrict__ U, In in, int n, int nx, int nt, R mm,
                                               Mask mask, cplx<R>& o0, cplx<R>& o1) {
    const int r0 = nb_right(n, 0, nx, nt), r1 = nb_right(n, 1, nx, nt);
    const int l0 = nb_left(n, 0, nx, nt), l1 = nb_left(n, 1, nx, nt);
    cplx<R> c0, c1;
    wilson_core<R, DAG>(U, in, n, r0, r1, l0, l1, mask(r0), mask(r1), mask(l0), mask(l1), wraps_right(n, 0, nt),
                        wraps_left(n, 0, nt), mm, o0, o1, c0, c1);
}

template <typename R, bool DAG, typename Mask>
__device__ __forceinline__ void wilson_site(const cplx<R>* __restrict__ U, const cplx<R>* __restrict__ in,
                                            int n, int nx, int nt, R mm, Mask mask, cplx<R>& o0, cplx<R>& o1) {
    wilson_site_in<R, DAG>(U, SpinPtr<R>{in}, n, nx, nt, mm, mask, o0, o1);
}

struct AllSites {
    __device__ bool operator()(int) const { return true; }
};
struct InBlock {
    Geo g;
    int blk;
    __device__ bool operator()(int n) const { return sap_block_of(g, n) == blk; }
};
// Block-vector reads of the Schwarz-block stencil on non-square lattices: the
// reference indexes the block vector by the local index m(n') that its block
// test computes for a neighbour n' (getMandBlock, include/level.h:48-59, with
// x = n / Nx; SAP_fine_level::D_B, src/sap.cpp:196-210), so an in-block
// neighbour reads the entry of the block's site with local index m(n'),
// sap_site(blk, m(n')), which differs from n' when Nx != Nt.  The centre site
// (and a masked-out neighbour, which reads the centre) reads its own entry.
template <typename R, typename In> struct BlockLocalIn {
    In in;
    Geo g;
    int blk, n;
    __device__ __forceinline__ void operator()(int i, cplx<R>& a, cplx<R>& b) const {
        in(i == n ? n : sap_site(g, blk, sap_local_of(g, i)), a, b);
    }
};

// out = D in  (or D^dagger), all sites, all right-hand sides
// Several gauge configurations in one batch: right-hand side r uses the links
// at U + r * ucs and the mass term mmv[r] (ucs = 0 and mmv = null for one).
template <typename R>
__device__ __forceinline__ R cfg_mm(R mm, const R* __restrict__ mmv, int rhs) { return mmv ? mmv[rhs] : mm; }

template <typename R, bool DAG>
__global__ void __launch_bounds__(FT) k_wilson(Geo g, R mm, const cplx<R>* __restrict__ U,
                                               const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out,
                                               size_t ucs, const R* __restrict__ mmv) {
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int n = tile * FT + threadIdx.x;
    if (n >= g.ns) return;
    const size_t off = (size_t)blockIdx.y * g.ns * 2;
    cplx<R> o0, o1;
    wilson_site<R, DAG>(U + blockIdx.y * ucs, in + off, n, g.nx, g.nt, cfg_mm(mm, mmv, blockIdx.y), AllSites{}, o0, o1);
    out[off + 2 * n] = o0;
    out[off + 2 * n + 1] = o1;
}

// r = v - D x; optional per-workgroup partial sums of |r|^2 (for SAP's
// convergence check) into part[rhs][tile].  skip[rhs] != 0 leaves r unchanged;
// with want (one flag per Schwarz block) only the flagged blocks' sites are
// written, the others keep their values.
template <typename R>
__global__ void __launch_bounds__(FT) k_resid(Geo g, R mm, const cplx<R>* __restrict__ U,
                                              const cplx<R>* __restrict__ v, const cplx<R>* __restrict__ x,
                                              cplx<R>* __restrict__ r, double* __restrict__ part,
                                              const int* __restrict__ skip,
                                              const unsigned char* __restrict__ want, size_t ucs = 0,
                                              const R* __restrict__ mmv = nullptr) {
    __shared__ c64 red[4];
    const int rhs = blockIdx.y;
    if (skip && skip[rhs]) return;
    U += rhs * ucs;
    mm = cfg_mm(mm, mmv, rhs);
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int n = tile * FT + threadIdx.x;
    const size_t off = (size_t)rhs * g.ns * 2;
    double nn = 0;
    bool on = n < g.ns;
    if (on && want) {
        // only the sites of the Schwarz blocks flagged in want (true coordinates)
        const int x = n / g.nt, t = n - x * g.nt;
        on = want[(x / g.sxel) * g.sbt + t / g.stel] != 0;
    }
    if (on) {
        cplx<R> o0, o1;
        wilson_site<R, false>(U, x + off, n, g.nx, g.nt, mm, AllSites{}, o0, o1);
        cplx<R> a = v[off + 2 * n] - o0, b = v[off + 2 * n + 1] - o1;
        r[off + 2 * n] = a;
        r[off + 2 * n + 1] = b;
        nn = double(norm2(a)) + double(norm2(b));
    }
    if (part) {
        c64 q[1] = {{nn, 0.0}};
        block_sum<1>(q, red);
        if (threadIdx.x == 0) part[(size_t)rhs * gridDim.x + tile] = q[0].re;
    }
}

// ---------------------------------------------------------------------------
// SAP on the fine lattice (src/sap.cpp:44-174, include/sap.h:27-47).
//
// All Schwarz blocks of one colour and all right-hand sides are solved at once:
// an item is (block of the colour, rhs); each item runs its own restarted
// GMRES(m) with the reference's control flow (x0 = b_B, |g_{j+1}| < tol ||b_B||
// breaks a cycle, true residual after each cycle, maxIt carried across cycles).
// Vectors keep the full-lattice layout: the colour's blocks are disjoint, so
// block-local vectors of all items of a colour share one lattice-sized buffer.
//
// An item covers nchunk workgroups of FT sites each.  Scalars (Hessenberg
// column, Givens rotations, residual estimate) are recomputed identically by
// every workgroup of an item from the previous kernel's per-workgroup partial
// sums (fixed summation order), and stored by chunk 0 for later kernels; a
// kernel never reads state that the same kernel writes.
//
// Modified Gram-Schmidt is evaluated in one pass: with G_ki = <V_k, V_i>,
//   H_ij = <w, V_i> - sum_{k<i} H_kj G_ki
// equals the sequential MGS coefficient <w - sum_{k<i} H_kj V_k, V_i> exactly
// in exact arithmetic.  Basis vectors are stored unnormalised (Vt_j) with
// V_j = s_j Vt_j, s_j = 1/H_{j,j-1}, so no separate normalisation pass runs.
// ---------------------------------------------------------------------------
template <typename R> struct SapArgs {
    Geo g;
    R mm;
    const cplx<R>* U;
    // several gauge configurations in one batch: rhs r uses the links at
    // U + r * ucs and the mass term mmv[r] (ucs = 0, mmv = null: one configuration)
    size_t ucs;
    const R* mmv;
    const int* blocks;  // block ids of this colour [ncol]
    int ncol, nchunk, cs, nrhs, m;  // cs: sites per chunk (a multiple of FT)
    double tol;
    const cplx<R>* rglob;  // b of each item = global residual restricted to the block
    cplx<R>* Vt;           // (m+1) slots of [nrhs][ns][2]; with FP16 storage only slot 0 (V_0) and slot 1 (w)
    f16x4* Vh;             // FP16 storage: slots 0..m of [nrhs][ns] (both spins of a site in 8 bytes)
    const double* rsig;    // FP16 storage: power-of-two scale of Vt_0 at the start [rhs]
    float sig;             // FP16 storage: power-of-two scale of the stored slots (1 otherwise)
    int pairs;             // step A may take sites in pairs along t (block_stencil_pair's conditions)
    float rcp_stel;        // 1 / stel (block-local index -> coordinates)
    cplx<R>* xl;           // [nrhs][ns][2] block solutions
    const int* skip;       // per-rhs: SAP already converged
    SapState st;
};

struct ItemIdx {
    int item, chunk, rhs, blk, m_lo, m_hi;
    int x0, t0;  // lattice coordinates of the block's first site
    double mm;   // m0 + 2 of the item's configuration
};

template <typename R> __device__ __forceinline__ ItemIdx item_of(const SapArgs<R>& a) {
    ItemIdx t;
    const int i = blockIdx.x / a.nchunk;
    t.chunk = blockIdx.x - i * a.nchunk;
    t.rhs = blockIdx.y;
    t.item = t.rhs * a.ncol + i;
    t.blk = a.blocks[i];
    const int X = t.blk / a.g.sbt;
    t.x0 = X * a.g.sxel;
    t.t0 = (t.blk - X * a.g.sbt) * a.g.stel;
    t.m_lo = t.chunk * a.cs;
    t.m_hi = min(a.g.sspb, t.m_lo + a.cs);
    t.mm = a.mmv ? (double)a.mmv[t.rhs] : (double)a.mm;
    return t;
}

// Site m (x-major local index, src/sap.cpp:3-24) of the item's Schwarz block:
// global site and coordinates.  m / stel through the FP32 reciprocal, exact for
// blocks of up to 2^20 sites.
struct BSite {
    int n, x, t, mx, mt;
};
template <typename R> __device__ __forceinline__ BSite bsite(const SapArgs<R>& a, const ItemIdx& t, int m) {
    BSite b;
    b.mx = (int)((float(m) + 0.5f) * a.rcp_stel);
    b.mt = m - b.mx * a.g.stel;
    b.x = t.x0 + b.mx;
    b.t = t.t0 + b.mt;
    b.n = b.x * a.g.nt + b.t;
    return b;
}

// D_B at a site of the item's block, zero Dirichlet outside the block
// (src/level.cpp:304-368).  On square lattices a neighbour's membership
// follows from the local coordinates; otherwise it is the reference's
// getMandBlock test with its x = n / Nx (include/level.h:48-59).
// c0, c1 return the site's own input values.
template <bool SQ, typename R, typename In>
__device__ __forceinline__ void block_stencil(const SapArgs<R>& a, const ItemIdx& t, const BSite& b, In in, cplx<R>& o0,
                                              cplx<R>& o1, cplx<R>& c0, cplx<R>& c1) {
    const Geo& g = a.g;
    const cplx<R>* U = a.U + (size_t)t.rhs * a.ucs;
    const R mm = R(t.mm);
    if constexpr (SQ) {
        const bool tr = b.t + 1 == g.nt, tl = b.t == 0;
        const int r0 = tr ? b.n - b.t : b.n + 1;
        const int l0 = tl ? b.n + g.nt - 1 : b.n - 1;
        const int r1 = b.x + 1 == g.nx ? b.t : b.n + g.nt;
        const int l1 = b.x == 0 ? b.n + (g.nx - 1) * g.nt : b.n - g.nt;
        const bool wt = g.stel == g.nt, wx = g.sxel == g.nx;  // block spans the extent
        wilson_core<R, false>(U, in, b.n, r0, r1, l0, l1, wt || b.mt + 1 < g.stel, wx || b.mx + 1 < g.sxel,
                              wt || b.mt > 0, wx || b.mx > 0, tr, tl, mm, o0, o1, c0, c1);
    } else {
        wilson_site_in<R, false>(U, BlockLocalIn<R, In>{in, g, t.blk, b.n}, b.n, g.nx, g.nt, mm,
                                 InBlock{g, t.blk}, o0, o1);
        in(b.n, c0, c1);
    }
}
template <bool SQ, typename R, typename In>
__device__ __forceinline__ void block_stencil(const SapArgs<R>& a, const ItemIdx& t, const BSite& b, In in, cplx<R>& o0,
                                              cplx<R>& o1) {
    cplx<R> c0, c1;
    block_stencil<SQ>(a, t, b, in, o0, o1, c0, c1);
}

// D_B at the two sites n, n+1 of a pair along t (the block-local t of site n
// even; square lattice, FP16 input, blocks of even t extent that do not span
// the lattice).  Each site is the other's t-neighbour, so the pair reads its
// own entries, its x-neighbour pairs and its links with 16-byte loads and only
// the two outer t-neighbours singly: about half the loads and index arithmetic
// of two single-site stencils, with the same arithmetic per site.
// o[s][spin] are the outputs, c[s][spin] the pair's own input values.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ void block_stencil_pair(const SapArgs<float>& a, const ItemIdx& t, const BSite& b,
                                                   const f16x4* __restrict__ in, c32 (&o)[2][2], c32 (&c)[2][2]) {
    const Geo& g = a.g;
    const c32* Ut = a.U + (size_t)t.rhs * a.ucs;
    const float mm = float(t.mm);
    const int n = b.n;
    const bool tl = b.t == 0, tr1 = b.t + 2 == g.nt;  // wraps: site n to the left, site n+1 to the right
    const bool ml0 = b.mt > 0, mr1t = b.mt + 2 < g.stel;         // outer t-neighbours in the block
    const bool mx_r = b.mx + 1 < g.sxel, mx_l = b.mx > 0;        // x-neighbours in the block
    const int l0 = tl ? n + g.nt - 1 : n - 1;
    const int r0b = tr1 ? n + 1 - (b.t + 1) : n + 2;
    const int xr = b.x + 1 == g.nx ? b.t : n + g.nt;
    const int xl = b.x == 0 ? n + (g.nx - 1) * g.nt : n - g.nt;
    // loads: a masked-out neighbour reads the pair's own entries (its term is zero)
    const f16x8 cp = *reinterpret_cast<const f16x8*>(in + (unsigned)n);
    const f16x8 rp = *reinterpret_cast<const f16x8*>(in + (unsigned)(mx_r ? xr : n));
    const f16x8 lp = *reinterpret_cast<const f16x8*>(in + (unsigned)(mx_l ? xl : n));
    const f16x4 tL = in[(unsigned)(ml0 ? l0 : n)], tR = in[(unsigned)(mr1t ? r0b : n)];
    using F4 = float __attribute__((ext_vector_type(4)));
    const F4* U4 = reinterpret_cast<const F4*>(Ut);
    const F4 ua = U4[(unsigned)n], ub = U4[(unsigned)n + 1u];  // f0, f1 of n; f0, f1 of n+1
    const c32 ul0 = Ut[2u * (unsigned)l0];
    const c32 ux0 = Ut[2u * (unsigned)xl + 1u], ux1 = Ut[2u * (unsigned)xl + 3u];
    const f16x4 h0 = cp.lo, h1 = cp.hi, r0v = rp.lo, r1v = rp.hi, l0v = lp.lo, l1v = lp.hi;
    c[0][0] = c32{float(h0.x), float(h0.y)};
    c[0][1] = c32{float(h0.z), float(h0.w)};
    c[1][0] = c32{float(h1.x), float(h1.y)};
    c[1][1] = c32{float(h1.z), float(h1.w)};
    c32 f00{ua.x, ua.y}, f01{ua.z, ua.w}, f10{ub.x, ub.y}, f11{ub.z, ub.w};
    c32 b00 = conj(ul0), b10 = conj(f00);  // left t links of n and n+1
    if (tl) b00 = -b00;
    if (tr1) f10 = -f10;
    const c32 b01 = conj(ux0), b11 = conj(ux1);  // left x links
    const H16Proj P;
    // site n: right t = n+1, left t = tL; site n+1: right t = tR, left t = n
    wilson_combine<float, false>(c[0][0], c[0][1], P.sum(h1, true), P.sumi(r0v, false), P.sum(tL, false),
                                 P.sumi(l0v, true), f00, f01, b00, b01, true, mx_r, ml0, mx_l, mm, o[0][0], o[0][1]);
    wilson_combine<float, false>(c[1][0], c[1][1], P.sum(tR, true), P.sumi(r1v, false), P.sum(h0, false),
                                 P.sumi(l1v, true), f10, f11, b10, b11, mr1t, mx_r, true, mx_l, mm, o[1][0], o[1][1]);
}

template <typename R> __device__ __forceinline__ cplx<R>* slot(const SapArgs<R>& a, int j, int rhs) {
    return a.Vt + ((size_t)j * a.nrhs + rhs) * a.g.ns * 2;
}
// FP16 slot j
template <typename R> __device__ __forceinline__ f16x4* hslot(const SapArgs<R>& a, int j, int rhs) {
    return a.Vh + ((size_t)j * a.nrhs + rhs) * a.g.ns;
}
// both spins of Krylov slot i at site n: FP16 slots when H, else complex R
template <typename R, bool H>
__device__ __forceinline__ void ld_slot(const SapArgs<R>& a, int i, int rhs, int n, cplx<R>& b0, cplx<R>& b1) {
    if constexpr (H) {
        SpinH16{hslot(a, i, rhs)}(n, b0, b1);
        return;
    }
    SpinPtr<R>{slot(a, i, rhs)}(n, b0, b1);
}
// Vt_0 = r at site n; returns |stored|^2.  With FP16 storage the stored
// values are sg r rounded (sg a power of two that brings the entries to order 1).
template <typename R, bool H>
__device__ __forceinline__ double st_slot0(const SapArgs<R>& a, int rhs, int n, cplx<R> r0, cplx<R> r1, double sg) {
    if constexpr (H) {
        const float f = float(sg);
        const f16x4 h = to_h16(c32{f * r0.re, f * r0.im}, c32{f * r1.re, f * r1.im});
        hslot(a, 0, rhs)[(unsigned)n] = h;
        return double(h.x) * double(h.x) + double(h.y) * double(h.y) + double(h.z) * double(h.z) +
               double(h.w) * double(h.w);
    } else {
        cplx<R>* v0 = slot(a, 0, rhs);
        v0[2u * (unsigned)n] = r0;
        v0[2u * (unsigned)n + 1u] = r1;
        return double(norm2(r0)) + double(norm2(r1));
    }
}

// Per-item scalar state held in LDS during a kernel (matrices with row
// stride M, as in global memory).
struct ItemLds {
    int active, maxIt, broke, done, jbreak;
    double normb;
    double s[SAP_MAXM + 1];
    c64 Hraw[(SAP_MAXM + 1) * SAP_MAXM];
    c64 R[(SAP_MAXM + 1) * SAP_MAXM];
    c64 G[(SAP_MAXM + 1) * (SAP_MAXM + 1)];
    c64 gpre[SAP_MAXM + 1], gpost[SAP_MAXM + 1];
    c64 cn[SAP_MAXM], sn[SAP_MAXM];
    c64 Hcol[SAP_MAXM + 1];
    c64 eta[SAP_MAXM];
    c64 dq[2 * SAP_MAXM];
    double dr[2];
    c64 tmp[FT];  // reduction scratch
};

// All threads: copy the item's state into LDS (one parallel pass of loads).
__device__ __forceinline__ void load_item(ItemLds& L, const SapState& st, int item, int M) {
    const int t = threadIdx.x;
    const int nH = (M + 1) * M, nG = (M + 1) * (M + 1);
    for (int i = t; i < nH; i += FT) {
        L.Hraw[i] = st.Hraw[(size_t)item * nH + i];
        L.R[i] = st.Hrot[(size_t)item * nH + i];
    }
    for (int i = t; i < nG; i += FT) L.G[i] = st.G[(size_t)item * nG + i];
    if (t < M + 1) {
        L.s[t] = st.s[(size_t)item * (M + 1) + t];
        L.gpre[t] = st.gpre[(size_t)item * (M + 1) + t];
        L.gpost[t] = st.gpost[(size_t)item * (M + 1) + t];
    }
    if (t >= 64 && t < 64 + M) {
        L.cn[t - 64] = st.cn[(size_t)item * M + t - 64];
        L.sn[t - 64] = st.sn[(size_t)item * M + t - 64];
    }
    if (t == 128) {
        L.done = st.done[item];
        L.jbreak = st.jbreak[item];
        L.maxIt = st.maxIt[item];
        L.normb = st.normb[item];
    }
}

// Sum of per-chunk partials base[c*stride + q] over chunks for q < nq, in a
// fixed order (16 threads per quantity, then the 16 partials in order).
template <typename T>
__device__ __forceinline__ void reduce_chunks(const T* base, int stride, int nq, int nchunk, c64* tmp, T* out) {
    const int q = threadIdx.x >> 4, sub = threadIdx.x & 15;
    T sacc{};
    if (q < nq)
        for (int c = sub; c < nchunk; c += 16) sacc += base[(size_t)c * stride + q];
    reinterpret_cast<T*>(tmp)[threadIdx.x] = sacc;
    __syncthreads();
    if (threadIdx.x < nq) {
        T r = reinterpret_cast<T*>(tmp)[threadIdx.x * 16];
        for (int i = 1; i < 16; i++) r += reinterpret_cast<T*>(tmp)[threadIdx.x * 16 + i];
        out[threadIdx.x] = r;
    }
    __syncthreads();
}

// Finalise Arnoldi step j (src/fgmres.cpp:35-49): H_{j+1,j} = |w|, Givens
// rotation of column j, residual estimate, break test.  Thread 0 on LDS
// state; chunk 0 also stores the results.
// hn2 is |sig w|^2 of the stored vector: H_{j+1,j} = |w| = sqrt(hn2) / sig and the
// scale of the stored vector is s_{j+1} = 1 / sqrt(hn2).
__device__ void sap_finalize_step(const SapState& st, int item, double hn2, int M, int j, double tol, bool store,
                                  ItemLds& L, double sig) {
    const double hs = sqrt(hn2), hn = hs / sig;
    c64* col = L.Hcol;
    for (int i = 0; i <= j; i++) col[i] = L.Hraw[i * M + j];
    col[j + 1] = c64{hn, 0.0};
    for (int i = 0; i < j; i++) {
        const c64 a = col[i], b = col[i + 1];
        col[i + 1] = -(L.sn[i] * a) + L.cn[i] * b;
        col[i] = conj(L.cn[i]) * a + conj(L.sn[i]) * b;
    }
    const double den = sqrt(norm2(col[j]) + norm2(col[j + 1]));
    const c64 snj = {col[j + 1].re / den, col[j + 1].im / den};
    const c64 cnj = {col[j].re / den, col[j].im / den};
    col[j] = conj(cnj) * col[j] + conj(snj) * col[j + 1];
    const c64 gj = L.gpre[j];
    const c64 gn = -(snj * gj);
    const c64 gp = conj(cnj) * gj;
    const bool broke = sqrt(norm2(gn)) < tol * L.normb;
    for (int i = 0; i <= j; i++) L.R[i * M + j] = col[i];
    L.cn[j] = cnj;
    L.sn[j] = snj;
    L.gpost[j] = gp;
    L.gpre[j + 1] = gn;
    L.s[j + 1] = hs > 0 ? 1.0 / hs : 0.0;
    L.broke = broke;
    if (store) {
        c64* Ro = st.Hrot + (size_t)item * (M + 1) * M;
        for (int i = 0; i <= j; i++) Ro[i * M + j] = col[i];
        st.cn[(size_t)item * M + j] = cnj;
        st.sn[(size_t)item * M + j] = snj;
        st.gpost[(size_t)item * (M + 1) + j] = gp;
        st.gpre[(size_t)item * (M + 1) + j + 1] = gn;
        st.s[(size_t)item * (M + 1) + j + 1] = L.s[j + 1];
        if (broke) {
            st.jbreak[item] = j;
            st.maxIt[item] = j + 1;
        }
    }
}

// Hessenberg column j of a block GMRES by modified Gram-Schmidt algebra on
// the step's inner products of the stored vectors, L.dq: d[i] = <W, Vt_i>
// (i <= j, W = D_B Vt_j unscaled), d[j+1+k] = <Vt_j, Vt_k> (k < j).  Sets the
// Gram row j (G_jk = s_j s_k <Vt_j, Vt_k>), Hcol (H_ij = <w, V_i> - sum_{k<i}
// H_kj G_ik with w = s_j W) and the coefficients eta_i = s_i H_ij on the Vt_i.
__device__ __forceinline__ void sap_hessenberg_column(ItemLds& L, int M, int j) {
    const c64* d = L.dq;
    for (int kk = 0; kk < j; kk++) L.G[j * (M + 1) + kk] = (L.s[j] * L.s[kk]) * d[j + 1 + kk];
    for (int i = 0; i <= j; i++) {
        c64 h = (L.s[i] * L.s[j]) * d[i];
        for (int kk = 0; kk < i; kk++) h -= L.Hcol[kk] * conj(L.G[i * (M + 1) + kk]);
        L.Hcol[i] = h;
    }
    for (int i = 0; i <= j; i++) L.eta[i] = L.s[i] * L.Hcol[i];
}

// End of a cycle: y from the rotated triangular system (src/fgmres.cpp:52-59,
// 107-115) over maxIt columns, as coefficients eta_i = s_i y_i on the Vt_i.
__device__ __forceinline__ void sap_backsolve(ItemLds& L, int M, int maxIt) {
    for (int i = maxIt - 1; i >= 0; i--) {
        c64 e = L.gpost[i];
        for (int q = i + 1; q < maxIt; q++) e -= L.R[i * M + q] * L.eta[q];
        L.eta[i] = cdiv(e, L.R[i * M + i]);
    }
    for (int i = 0; i < maxIt; i++) L.eta[i] = L.s[i] * L.eta[i];
}

// FP16 storage: power-of-two scale of the next cycle's stored Vt_0 from the
// cycle's final residual estimate spread over the block's 2 sspb entries
// (entries of order 1); 0 when the estimate gives none.
__device__ __forceinline__ double sap_next_sig0(const ItemLds& L, int maxIt, int sspb) {
    const double est = sqrt(norm2(L.gpre[maxIt]));
    const double want = sqrt(2.0 * sspb) / est;
    return est > 0 && want < 1e30 ? exp2(rint(log2(want))) : 0.0;
}

// Wave-reduce v for quantity q and park it in tmp[wave][q] (no private arrays
// with run-time indices).
__device__ __forceinline__ void park(c64 v, int q, c64* tmp, int nqmax) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) tmp[(threadIdx.x >> 6) * nqmax + q] = v;
}

// init: x_B = b_B (x0 = b, src/sap.cpp:61), Vt_0 = b_B - D_B b_B; partials |r|^2, |b|^2
template <typename R, bool H, bool SQ> __global__ void __launch_bounds__(FT) k_sap_init(SapArgs<R> a) {
    __shared__ c64 red[8];
    const ItemIdx t = item_of(a);
    if (a.skip && a.skip[t.rhs]) return;
    const size_t off = (size_t)t.rhs * a.g.ns * 2;
    const cplx<R>* rg = a.rglob + off;
    cplx<R>* xl = a.xl + off;
    const double sg = H ? a.rsig[t.rhs] : 1.0;
    c64 q[2] = {{0, 0}, {0, 0}};
    for (int m = t.m_lo + threadIdx.x; m < t.m_hi; m += FT) {
        const BSite b = bsite(a, t, m);
        const int n = b.n;
        cplx<R> o0, o1;
        block_stencil<SQ>(a, t, b, SpinPtr<R>{a.rglob + off}, o0, o1);
        const unsigned u = 2u * (unsigned)n;
        const cplx<R> b0 = rg[u], b1 = rg[u + 1u];
        xl[u] = b0;
        xl[u + 1u] = b1;
        const cplx<R> r0 = b0 - o0, r1 = b1 - o1;
        q[0].re += st_slot0<R, H>(a, t.rhs, n, r0, r1, sg);
        q[1].re += double(norm2(b0)) + double(norm2(b1));
    }
    block_sum<2>(q, red);
    if (threadIdx.x == 0) {
        a.st.qr[((size_t)t.item * a.nchunk + t.chunk) * 2] = q[0].re;
        a.st.qr[((size_t)t.item * a.nchunk + t.chunk) * 2 + 1] = q[1].re;
        if (t.chunk == 0) {
            a.st.maxIt[t.item] = a.m;
            a.st.done[t.item] = 0;
            a.st.jbreak[t.item] = -1;
            if (H) a.st.sig0[t.item] = sg;
        }
    }
}

// Arnoldi step J, part A: W = D_B Vt_J (unscaled; step B applies s_J) into
// Vt_{J+1} (slot 1 with FP16 storage H), partials <W, Vt_i> (i <= J) and
// <Vt_J, Vt_i> (i < J).  The stream does not depend on the item's scalar
// state: only chunk 0 of an item first finalises step J-1 (or starts cycle k)
// and stores the result, which every later kernel reads; the other chunks
// start streaming at once, skipping only items whose end an earlier kernel
// recorded.  An item that this kernel's chunk 0 finds finished has streamed
// once for nothing; step B and the update skip it.
template <typename R, int J, bool H, bool SQ>
__global__ void __launch_bounds__(FT, (H && SQ && J <= 3 ? 8 : 1)) k_sap_arnoldi_a(SapArgs<R> a, int k, int* act) {
    __shared__ ItemLds L;
    const ItemIdx t = item_of(a);
    const int M = a.m;
    if (a.skip && a.skip[t.rhs]) return;
    if (t.chunk == 0) {
        load_item(L, a.st, t.item, M);
        if (J == 0) {
            reduce_chunks<double>(a.st.qr + (size_t)t.item * a.nchunk * 2, 2, 2, a.nchunk, L.tmp, L.dr);
        } else {
            reduce_chunks<double>(a.st.qb + (size_t)t.item * a.nchunk, 1, 1, a.nchunk, L.tmp, L.dr);
        }
        if (threadIdx.x == 0) {
            const SapState& st = a.st;
            int active = !L.done;
            if (active && J == 0) {
                // cycle start (src/fgmres.cpp:8-17, 61-65); Vt_0 is stored
                // scaled by sig0 (1 without FP16 storage)
                const double est = sqrt(L.dr[0]), err = H ? est / st.sig0[t.item] : est;
                const double nb = k == 0 ? sqrt(L.dr[1]) : L.normb;
                const bool done = (k > 0 && err < a.tol * nb) || nb == 0.0;
                if (done) {
                    active = 0;
                    st.done[t.item] = 1;
                } else {
                    if (k == 0) st.normb[t.item] = nb;
                    st.s[(size_t)t.item * (M + 1)] = 1.0 / est;
                    st.gpre[(size_t)t.item * (M + 1)] = c64{err, 0.0};
                    st.jbreak[t.item] = -1;
                }
            } else if (active) {
                if (L.jbreak >= 0) active = 0;
                else {
                    sap_finalize_step(st, t.item, L.dr[0], M, J - 1, a.tol, true, L, H ? a.sig : 1.0);
                    if (L.broke) active = 0;
                }
            }
            if (act && active) atomicAdd(act, 1);
        }
        __syncthreads();  // L.tmp is reused below
    } else if (a.st.done[t.item] || (J > 0 && a.st.jbreak[t.item] >= 0)) {
        // finished in an earlier cycle, or broke down at step J-2 or before
        return;
    }
    cplx<R>* w = slot(a, H ? 1 : J + 1, t.rhs);
    // per-thread partial sums in the working precision, combined in FP64
    cplx<R> pw[J + 1], pv[J + 1];
#pragma unroll
    for (int i = 0; i <= J; i++) pw[i] = pv[i] = czero<R>();
    if constexpr (H && SQ) {
        if (a.pairs) {
            // two sites along t per thread (the host checked the geometry)
            const f16x4* vin = hslot(a, J, t.rhs);
#pragma unroll 1
            for (int m = t.m_lo + 2 * threadIdx.x; m < t.m_hi; m += 2 * FT) {
                const BSite b = bsite(a, t, m);
                const unsigned n = (unsigned)b.n;
                c32 o[2][2], cc[2][2];
                block_stencil_pair(a, t, b, vin, o, cc);
                const f16x4 w0 = to_h16(o[0][0], o[0][1]), w1 = to_h16(o[1][0], o[1][1]);
                *reinterpret_cast<f16x8*>(reinterpret_cast<f16x4*>(w) + n) =
                    f16x8{w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
                const c32 s00{float(w0.x), float(w0.y)}, s01{float(w0.z), float(w0.w)};
                const c32 s10{float(w1.x), float(w1.y)}, s11{float(w1.z), float(w1.w)};
#pragma unroll
                for (int i = 0; i <= J; i++) {
                    c32 v00 = cc[0][0], v01 = cc[0][1], v10 = cc[1][0], v11 = cc[1][1];
                    if (i < J) {
                        const f16x8 v = *reinterpret_cast<const f16x8*>(hslot(a, i, t.rhs) + n);
                        v00 = c32{float(v.s0), float(v.s1)};
                        v01 = c32{float(v.s2), float(v.s3)};
                        v10 = c32{float(v.s4), float(v.s5)};
                        v11 = c32{float(v.s6), float(v.s7)};
                    }
                    cmadc(pw[i], s00, v00);
                    cmadc(pw[i], s01, v01);
                    cmadc(pw[i], s10, v10);
                    cmadc(pw[i], s11, v11);
                }
            }
        }
    }
#pragma unroll (J >= 4 ? 1 : 2)
    for (int m = t.m_lo + threadIdx.x; m < t.m_hi && !(H && SQ && a.pairs); m += FT) {
        const BSite b = bsite(a, t, m);
        const int n = b.n;
        // a0, a1: Vt_J at the site, read once by the stencil
        cplx<R> o0, o1, a0, a1;
        if constexpr (H)
            block_stencil<SQ>(a, t, b, SpinH16{hslot(a, J, t.rhs)}, o0, o1, a0, a1);
        else
            block_stencil<SQ>(a, t, b, SpinPtr<R>{slot(a, J, t.rhs)}, o0, o1, a0, a1);
        if constexpr (H) {
            // FP16 storage: W in FP16 too (in the first half of slot 1); the
            // dot products use the stored values, so that the Hessenberg
            // column belongs to the vector that step B reads
            const f16x4 hw = to_h16(o0, o1);
            reinterpret_cast<f16x4*>(w)[(unsigned)n] = hw;
            o0 = cplx<R>{R(hw.x), R(hw.y)};
            o1 = cplx<R>{R(hw.z), R(hw.w)};
        } else {
            w[2u * (unsigned)n] = o0;
            w[2u * (unsigned)n + 1u] = o1;
        }
#pragma unroll
        for (int i = 0; i <= J; i++) {
            cplx<R> b0 = a0, b1 = a1;
            if (i < J) ld_slot<R, H>(a, i, t.rhs, n, b0, b1);
            cmadc(pw[i], o0, b0);
            cmadc(pw[i], o1, b1);
            if (i < J) {
                cmadc(pv[i], a0, b0);
                cmadc(pv[i], a1, b1);
            }
        }
    }
#pragma unroll
    for (int i = 0; i <= J; i++) {
        park(cvt<double>(pw[i]), i, L.tmp, 2 * SAP_MAXM);
        if (i < J) park(cvt<double>(pv[i]), J + 1 + i, L.tmp, 2 * SAP_MAXM);
    }
    __syncthreads();
    if (threadIdx.x < 2 * J + 1) {
        c64 r = L.tmp[threadIdx.x];
        for (int wv = 1; wv < FT / 64; wv++) r += L.tmp[wv * 2 * SAP_MAXM + threadIdx.x];
        a.st.qa[((size_t)t.item * a.nchunk + t.chunk) * 2 * M + threadIdx.x] = r;
    }
}

// Arnoldi step j, part B: prologue builds Hessenberg column j (one-pass MGS),
// body w -= sum_i H_ij V_i (in place in Vt_{j+1}; with FP16 storage H from
// slot 1 into FP16 slot j+1 as sig w); partial |w|^2 of the stored values.
template <typename R, bool H> __global__ void __launch_bounds__(FT) k_sap_arnoldi_b(SapArgs<R> a, int j, int* act) {
    __shared__ ItemLds L;
    const ItemIdx t = item_of(a);
    const int M = a.m;
    if (a.skip && a.skip[t.rhs]) return;
    load_item(L, a.st, t.item, M);
    reduce_chunks<c64>(a.st.qa + (size_t)t.item * a.nchunk * 2 * M, 2 * M, 2 * j + 1, a.nchunk, L.tmp, L.dq);
    if (threadIdx.x == 0) {
        const SapState& st = a.st;
        const int active = !L.done && L.jbreak < 0;
        if (active) {
            sap_hessenberg_column(L, M, j);
            if (t.chunk == 0) {
                c64* Hr = st.Hraw + (size_t)t.item * (M + 1) * M;
                for (int i = 0; i <= j; i++) Hr[i * M + j] = L.Hcol[i];
                c64* Gm = st.G + (size_t)t.item * (M + 1) * (M + 1);
                for (int kk = 0; kk < j; kk++) Gm[j * (M + 1) + kk] = L.G[j * (M + 1) + kk];
            }
        }
        L.active = active;
        if (act && active && t.chunk == 0) atomicAdd(act, 1);
    }
    __syncthreads();
    if (!L.active) return;
    cplx<R>* w = slot(a, H ? 1 : j + 1, t.rhs);
    // the update runs in the engine's precision (the Gram coefficients were
    // formed in FP64 in the prologue)
    cplx<R> eta[SAP_MAXM + 1];
#pragma unroll
    for (int i = 0; i <= SAP_MAXM; i++) eta[i] = i <= j ? cvt<R>(L.eta[i]) : czero<R>();
    const R sj = R(L.s[j]);
    double nn = 0;
    if constexpr (H) {
        if (a.pairs) {
            // two sites along t per thread, 16-byte loads and stores
            const float sg = a.sig;
            for (int m = t.m_lo + 2 * threadIdx.x; m < t.m_hi; m += 2 * FT) {
                const unsigned n = (unsigned)bsite(a, t, m).n;
                const f16x8 hw = *reinterpret_cast<const f16x8*>(reinterpret_cast<const f16x4*>(w) + n);
                c32 u[4] = {sj * c32{float(hw.s0), float(hw.s1)}, sj * c32{float(hw.s2), float(hw.s3)},
                            sj * c32{float(hw.s4), float(hw.s5)}, sj * c32{float(hw.s6), float(hw.s7)}};
#pragma unroll
                for (int i = 0; i <= SAP_MAXM; i++) {
                    if (i <= j) {
                        const f16x8 v = *reinterpret_cast<const f16x8*>(hslot(a, i, t.rhs) + n);
                        const c32 e = cvt<float>(eta[i]);
                        cmsb(u[0], e, c32{float(v.s0), float(v.s1)});
                        cmsb(u[1], e, c32{float(v.s2), float(v.s3)});
                        cmsb(u[2], e, c32{float(v.s4), float(v.s5)});
                        cmsb(u[3], e, c32{float(v.s6), float(v.s7)});
                    }
                }
                const f16x4 h0 = to_h16(c32{sg * u[0].re, sg * u[0].im}, c32{sg * u[1].re, sg * u[1].im});
                const f16x4 h1 = to_h16(c32{sg * u[2].re, sg * u[2].im}, c32{sg * u[3].re, sg * u[3].im});
                *reinterpret_cast<f16x8*>(hslot(a, j + 1, t.rhs) + n) =
                    f16x8{h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
                nn += double(h0.x) * double(h0.x) + double(h0.y) * double(h0.y) + double(h0.z) * double(h0.z) +
                      double(h0.w) * double(h0.w);
                nn += double(h1.x) * double(h1.x) + double(h1.y) * double(h1.y) + double(h1.z) * double(h1.z) +
                      double(h1.w) * double(h1.w);
            }
        }
    }
    for (int m = t.m_lo + threadIdx.x; m < t.m_hi && !(H && a.pairs); m += FT) {
        const int n = bsite(a, t, m).n;
        cplx<R> w0, w1;
        if constexpr (H) {
            const f16x4 hw = reinterpret_cast<const f16x4*>(w)[(unsigned)n];
            w0 = sj * cplx<R>{R(hw.x), R(hw.y)};
            w1 = sj * cplx<R>{R(hw.z), R(hw.w)};
        } else {
            w0 = sj * w[2u * (unsigned)n];
            w1 = sj * w[2u * (unsigned)n + 1u];
        }
#pragma unroll
        for (int i = 0; i <= SAP_MAXM; i++) {
            if (i <= j) {
                cplx<R> v0, v1;
                ld_slot<R, H>(a, i, t.rhs, n, v0, v1);
                cmsb(w0, eta[i], v0);
                cmsb(w1, eta[i], v1);
            }
        }
        if constexpr (H) {
            const float sg = a.sig;
            const f16x4 h = to_h16(c32{sg * w0.re, sg * w0.im}, c32{sg * w1.re, sg * w1.im});
            hslot(a, j + 1, t.rhs)[(unsigned)n] = h;
            nn += double(h.x) * double(h.x) + double(h.y) * double(h.y) + double(h.z) * double(h.z) +
                  double(h.w) * double(h.w);
        } else {
            w[2u * (unsigned)n] = w0;
            w[2u * (unsigned)n + 1u] = w1;
            nn += double(norm2(w0)) + double(norm2(w1));
        }
    }
    c64 q[1] = {{nn, 0.0}};
    block_sum<1>(q, L.tmp);
    if (threadIdx.x == 0) a.st.qb[(size_t)t.item * a.nchunk + t.chunk] = q[0].re;
}

// End of cycle: finalise the last step if the cycle ran to m, back-substitute
// (src/fgmres.cpp:52-59, 107-115) and update x_B += sum_j eta_j V_j.  With
// xacc set (last cycle) the block solution is added into the global x
// instead (src/sap.cpp:117-125), also for items that converged earlier.
template <typename R, bool H> __global__ void __launch_bounds__(FT, (H ? 8 : 1)) k_sap_update(SapArgs<R> a, cplx<R>* xacc) {
    __shared__ ItemLds L;
    const ItemIdx t = item_of(a);
    const int M = a.m;
    if (a.skip && a.skip[t.rhs]) return;
    load_item(L, a.st, t.item, M);
    reduce_chunks<double>(a.st.qb + (size_t)t.item * a.nchunk, 1, 1, a.nchunk, L.tmp, L.dr);
    if (threadIdx.x == 0) {
        const int active = !L.done;
        int maxIt = 0;
        if (active) {
            maxIt = L.maxIt;
            if (L.jbreak < 0) {
                sap_finalize_step(a.st, t.item, L.dr[0], M, M - 1, a.tol, t.chunk == 0, L, H ? a.sig : 1.0);
                if (L.broke) maxIt = M;
            }
            sap_backsolve(L, M, maxIt);
            if (H && t.chunk == 0) {
                const double sg = sap_next_sig0(L, maxIt, a.g.sspb);
                if (sg > 0) a.st.sig0[t.item] = sg;
            }
        }
        L.maxIt = maxIt;
        L.active = active;
    }
    __syncthreads();
    if (!L.active && !xacc) return;
    const size_t off = (size_t)t.rhs * a.g.ns * 2;
    cplx<R>* xl = a.xl + off;
    const int mi = L.maxIt;
    // x_B += sum_j eta_j Vt_j in the engine's precision
    if constexpr (H) {
        if (a.pairs) {
            // two sites along t per thread, 16-byte loads and stores
            using F4 = float __attribute__((ext_vector_type(4)));
            for (int m = t.m_lo + 2 * threadIdx.x; m < t.m_hi; m += 2 * FT) {
                const unsigned n = (unsigned)bsite(a, t, m).n;
                F4* xp = reinterpret_cast<F4*>(xl) + n;  // sites n, n+1: 2 x 16 bytes
                F4 xa = xp[0], xb = xp[1];
                c32 x[4] = {c32{xa.x, xa.y}, c32{xa.z, xa.w}, c32{xb.x, xb.y}, c32{xb.z, xb.w}};
                for (int jj = 0; jj < mi; jj++) {
                    const f16x8 v = *reinterpret_cast<const f16x8*>(hslot(a, jj, t.rhs) + n);
                    const c32 c = cvt<float>(L.eta[jj]);
                    cmad(x[0], c, c32{float(v.s0), float(v.s1)});
                    cmad(x[1], c, c32{float(v.s2), float(v.s3)});
                    cmad(x[2], c, c32{float(v.s4), float(v.s5)});
                    cmad(x[3], c, c32{float(v.s6), float(v.s7)});
                }
                if (xacc) {
                    F4* gp = reinterpret_cast<F4*>(xacc + off) + n;
                    F4 ga = gp[0], gb = gp[1];
                    gp[0] = F4{ga.x + x[0].re, ga.y + x[0].im, ga.z + x[1].re, ga.w + x[1].im};
                    gp[1] = F4{gb.x + x[2].re, gb.y + x[2].im, gb.z + x[3].re, gb.w + x[3].im};
                } else {
                    xp[0] = F4{x[0].re, x[0].im, x[1].re, x[1].im};
                    xp[1] = F4{x[2].re, x[2].im, x[3].re, x[3].im};
                }
            }
            return;
        }
    }
    for (int m = t.m_lo + threadIdx.x; m < t.m_hi; m += FT) {
        const int n = bsite(a, t, m).n;
        const unsigned u = 2u * (unsigned)n;
        cplx<R> x0 = xl[u], x1 = xl[u + 1u];
        for (int jj = 0; jj < mi; jj++) {
            cplx<R> v0, v1;
            ld_slot<R, H>(a, jj, t.rhs, n, v0, v1);
            const cplx<R> c = cvt<R>(L.eta[jj]);
            cmad(x0, c, v0);
            cmad(x1, c, v1);
        }
        if (xacc) {
            xacc[off + u] += x0;
            xacc[off + u + 1u] += x1;
        } else {
            xl[u] = x0;
            xl[u + 1u] = x1;
        }
    }
}

// True block residual after a cycle: Vt_0 = b_B - D_B x_B, partial |stored|^2.
template <typename R, bool H, bool SQ> __global__ void __launch_bounds__(FT) k_sap_resid(SapArgs<R> a) {
    __shared__ c64 red[8];
    __shared__ int act;
    __shared__ double sg;
    const ItemIdx t = item_of(a);
    if (a.skip && a.skip[t.rhs]) return;
    if (threadIdx.x == 0) {
        act = !a.st.done[t.item];
        sg = H ? a.st.sig0[t.item] : 1.0;
    }
    __syncthreads();
    if (!act) return;
    const size_t off = (size_t)t.rhs * a.g.ns * 2;
    const cplx<R>* rg = a.rglob + off;
    c64 q[1] = {{0, 0}};
    for (int m = t.m_lo + threadIdx.x; m < t.m_hi; m += FT) {
        const BSite b = bsite(a, t, m);
        const int n = b.n;
        cplx<R> o0, o1;
        block_stencil<SQ>(a, t, b, SpinPtr<R>{a.xl + off}, o0, o1);
        const unsigned u = 2u * (unsigned)n;
        const cplx<R> r0 = rg[u] - o0, r1 = rg[u + 1u] - o1;
        q[0].re += st_slot0<R, H>(a, t.rhs, n, r0, r1, sg);
    }
    block_sum<1>(q, red);
    if (threadIdx.x == 0) a.st.qr[((size_t)t.item * a.nchunk + t.chunk) * 2] = q[0].re;
}

// per-rhs SAP convergence |r| < tol |v| after a sweep (src/sap.cpp:159-168);
// one workgroup per rhs
__global__ void __launch_bounds__(FT) k_sap_check(const double* part_r, const double* part_v, int ntile, double tol,
                                                  int* skip) {
    __shared__ c64 red[8];
    const int rhs = blockIdx.x;
    if (skip[rhs]) return;
    c64 q[2] = {{0, 0}, {0, 0}};
    for (int i = threadIdx.x; i < ntile; i += FT) {
        q[0].re += part_r[(size_t)rhs * ntile + i];
        q[1].re += part_v[(size_t)rhs * ntile + i];
    }
    block_sum<2>(q, red);
    if (threadIdx.x == 0 && sqrt(q[0].re) < tol * sqrt(q[1].re)) skip[rhs] = 1;
}

// FP16 storage: power-of-two scale rsig[rhs] of the first cycle's stored Vt_0,
// from |v|^2 partials of the smoother's input: its nent entries brought to
// order 1 (the residuals the blocks start from are no larger than v)
__global__ void __launch_bounds__(FT) k_sap_rscale(const double* part_v, int ntile, double nent, double* rsig) {
    __shared__ c64 red[8];
    const int rhs = blockIdx.x;
    c64 q[1] = {{0, 0}};
    for (int i = threadIdx.x; i < ntile; i += FT) q[0].re += part_v[(size_t)rhs * ntile + i];
    block_sum<1>(q, red);
    if (threadIdx.x == 0) {
        const double want = q[0].re > 0 ? sqrt(nent / q[0].re) : 1.0;
        rsig[rhs] = want < 1e30 ? exp2(rint(log2(want))) : 1.0;
    }
}

// |v|^2 per tile (for SAP's ||v||)
template <typename R>
__global__ void __launch_bounds__(FT) k_norm_part(int ns, const cplx<R>* __restrict__ v, double* part) {
    __shared__ c64 red[4];
    const int tile = blockIdx.x, n = tile * FT + threadIdx.x;
    const size_t off = (size_t)blockIdx.y * ns * 2;
    double nn = 0;
    if (n < ns) nn = double(norm2(v[off + 2 * n])) + double(norm2(v[off + 2 * n + 1]));
    c64 q[1] = {{nn, 0.0}};
    block_sum<1>(q, red);
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + tile] = q[0].re;
}

// ---------------------------------------------------------------------------
// Interpolation between level 0 and level 1 (src/level.cpp:182-237).
// P columns: w[cc][n][s] (test vector cc), coarse dof 2*cc + s, coarse site =
// aggregation block of n.
// ---------------------------------------------------------------------------

// out[X][2cc+s] = sum_{n in X} conj(w_cc[n][s]) v[n][s]; one workgroup per (X, rhs)
template <typename R, int NV>
__global__ void __launch_bounds__(FT) k_restrict0(Geo g, int nv, const cplx<R>* __restrict__ w,
                                                  const cplx<R>* __restrict__ v, cplx<R>* __restrict__ out,
                                                  int dofc, int nsc, size_t wcs = 0, const int* rmask = nullptr) {
    __shared__ cplx<R> red[4 * 2 * NV];
    const int X = blockIdx.x, rhs = blockIdx.y;
    if (rmask && !rmask[rhs]) return;
    w += rhs * wcs;
    const size_t off = (size_t)rhs * g.ns * 2, nvec = (size_t)g.ns * 2;
    cplx<R> acc[2 * NV];
#pragma unroll
    for (int q = 0; q < 2 * NV; q++) acc[q] = czero<R>();
    for (int m = threadIdx.x; m < g.spb; m += FT) {
        const int n = agg_site(g, X, m);
        const cplx<R> v0 = v[off + 2 * n], v1 = v[off + 2 * n + 1];
#pragma unroll
        for (int cc = 0; cc < NV; cc++) {
            if (cc < nv) {
                acc[2 * cc] += mulc(v0, w[cc * nvec + 2 * n]);
                acc[2 * cc + 1] += mulc(v1, w[cc * nvec + 2 * n + 1]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 2 * NV; q++) {
        if (q < 2 * nv) {
            const cplx<R> v = wave_sum(acc[q]);
            if ((threadIdx.x & 63) == 0) red[(threadIdx.x >> 6) * 2 * NV + q] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * nv) {
        cplx<R> r = red[threadIdx.x];
        for (int wv = 1; wv < FT / 64; wv++) r += red[wv * 2 * NV + threadIdx.x];
        out[(size_t)rhs * nsc * dofc + (size_t)X * dofc + threadIdx.x] = r;
    }
}

// out[n][s] (+)= sum_cc w_cc[n][s] vc[X(n)][2cc+s]
template <typename R>
__global__ void __launch_bounds__(FT) k_prolong0(Geo g, int nv, const cplx<R>* __restrict__ w,
                                                 const cplx<R>* __restrict__ vc, cplx<R>* __restrict__ out,
                                                 int dofc, int nsc, int accumulate, size_t wcs = 0,
                                                 const int* rmask = nullptr) {
    const int n = blockIdx.x * FT + threadIdx.x;
    if (n >= g.ns) return;
    const int rhs = blockIdx.y;
    if (rmask && !rmask[rhs]) return;
    w += rhs * wcs;
    const size_t off = (size_t)rhs * g.ns * 2, nvec = (size_t)g.ns * 2;
    const cplx<R>* c = vc + (size_t)rhs * nsc * dofc + (size_t)agg_block_of(g, n) * dofc;
    cplx<R> o0 = czero<R>(), o1 = czero<R>();
    for (int cc = 0; cc < nv; cc++) {
        o0 += w[cc * nvec + 2 * n] * c[2 * cc];
        o1 += w[cc * nvec + 2 * n + 1] * c[2 * cc + 1];
    }
    if (accumulate) {
        out[off + 2 * n] += o0;
        out[off + 2 * n + 1] += o1;
    } else {
        out[off + 2 * n] = o0;
        out[off + 2 * n + 1] = o1;
    }
}

// The FP32 engine's interpolation with the test vectors read once per group
// of right-hand sides (the kernels above read them once per rhs: 32 passes
// over them per launch at 32 rhs).  Same sums in the same order, with fused
// complex multiply-adds.
// Restriction, two rhs per workgroup (grid: aggregates x rhs pairs).
template <int NV>
__global__ void __launch_bounds__(FT) k_restrict0_g(Geo g, int nv, const c32* __restrict__ w,
                                                    const c32* __restrict__ v, c32* __restrict__ out, int dofc,
                                                    int nsc, int nrhs, const int* __restrict__ rmask) {
    __shared__ c32 red[4 * 2 * 2 * NV];
    const int X = blockIdx.x, rA = 2 * blockIdx.y, rB = rA + 1 < nrhs ? rA + 1 : rA;
    if (rmask && !rmask[rA] && !rmask[rB]) return;
    const size_t nvec = (size_t)g.ns * 2;
    c32 acc[2][2 * NV];
#pragma unroll
    for (int q = 0; q < 2 * NV; q++) acc[0][q] = acc[1][q] = czero<float>();
    for (int m = threadIdx.x; m < g.spb; m += FT) {
        const unsigned n = 2u * (unsigned)agg_site(g, X, m);
        const c32 a0 = v[rA * nvec + n], a1 = v[rA * nvec + n + 1u];
        const c32 b0 = v[rB * nvec + n], b1 = v[rB * nvec + n + 1u];
#pragma unroll
        for (int cc = 0; cc < NV; cc++) {
            if (cc < nv) {
                const c32 w0 = w[cc * nvec + n], w1 = w[cc * nvec + n + 1u];
                cmadc(acc[0][2 * cc], a0, w0);
                cmadc(acc[0][2 * cc + 1], a1, w1);
                cmadc(acc[1][2 * cc], b0, w0);
                cmadc(acc[1][2 * cc + 1], b1, w1);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int q = 0; q < 2 * NV; q++) {
            if (q < 2 * nv) {
                const c32 x = wave_sum(acc[r][q]);
                if ((threadIdx.x & 63) == 0) red[((threadIdx.x >> 6) * 2 + r) * 2 * NV + q] = x;
            }
        }
    __syncthreads();
    const int r = threadIdx.x / (2 * NV), q = threadIdx.x - r * 2 * NV;
    if (r < 2 && q < 2 * nv && (r == 0 || rB != rA)) {
        c32 x = red[r * 2 * NV + q];
        for (int wv = 1; wv < FT / 64; wv++) x += red[(wv * 2 + r) * 2 * NV + q];
        out[(size_t)(r ? rB : rA) * nsc * dofc + (size_t)X * dofc + q] = x;
    }
}

// Prolongation: a site's test-vector entries held in registers across up to
// 8 rhs (grid: site tiles x rhs groups of 8).
template <int NV>
__global__ void __launch_bounds__(FT) k_prolong0_g(Geo g, int nv, const c32* __restrict__ w,
                                                   const c32* __restrict__ vc, c32* __restrict__ out, int dofc,
                                                   int nsc, int nrhs, int accumulate, const int* __restrict__ rmask) {
    const int n = blockIdx.x * FT + threadIdx.x;
    if (n >= g.ns) return;
    const size_t nvec = (size_t)g.ns * 2;
    const unsigned u = 2u * (unsigned)n;
    c32 w0[NV], w1[NV];
#pragma unroll
    for (int cc = 0; cc < NV; cc++) {
        w0[cc] = cc < nv ? w[cc * nvec + u] : czero<float>();
        w1[cc] = cc < nv ? w[cc * nvec + u + 1u] : czero<float>();
    }
    const size_t X = (size_t)agg_block_of(g, n);
    const int r1 = min(nrhs, 8 * (int)blockIdx.y + 8);
    for (int rhs = 8 * blockIdx.y; rhs < r1; rhs++) {
        if (rmask && !rmask[rhs]) continue;
        const c32* c = vc + (size_t)rhs * nsc * dofc + X * dofc;
        c32 o0 = czero<float>(), o1 = czero<float>();
#pragma unroll
        for (int cc = 0; cc < NV; cc++) {
            if (cc < nv) {
                cmad(o0, w0[cc], c[2 * cc]);
                cmad(o1, w1[cc], c[2 * cc + 1]);
            }
        }
        c32* op = out + (size_t)rhs * nvec + u;
        if (accumulate) {
            op[0] += o0;
            op[1] += o1;
        } else {
            op[0] = o0;
            op[1] = o1;
        }
    }
}

// precision conversion / scaling of whole vectors
template <typename To, typename From>
__global__ void k_convert(size_t n, const cplx<From>* __restrict__ in, cplx<To>* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cvt<To>(in[i]);
}
// out = sigma_3 in: the spin-1 components change sign (D^dagger = s3 D s3)
__global__ void k_sigma3(size_t nsites_total, const c64* in, c64* out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsites_total) return;
    out[2 * i] = in[2 * i];
    out[2 * i + 1] = -in[2 * i + 1];
}

// Fermion force of S = phi^dagger (D D^dagger)^-1 phi with Y = D^-1 phi and
// X = D^dagger^-1 Y: dS/dtheta_mu(n) = -2 Re X^dagger (dD/dtheta_mu(n)) Y with
// dU/dtheta = iU, i.e. Re[i s U_mu(n) X(n)^+(1-s_mu)Y(n+mu) - i s U_mu(n)^* X(n+mu)^+(1+s_mu)Y(n)]
// (s = -1 on the antiperiodic time hop).  F: [nrhs][ns][2] real.
__global__ void __launch_bounds__(FT) k_force(Geo g, const c64* __restrict__ U, const c64* __restrict__ X,
                                              const c64* __restrict__ Y, double* __restrict__ F, size_t ucs) {
    const int n = blockIdx.x * FT + threadIdx.x;
    if (n >= g.ns) return;
    const size_t off = (size_t)blockIdx.y * g.ns * 2;
    U += blockIdx.y * ucs;
    const c64* x = X + off;
    const c64* y = Y + off;
    const c64 xn0 = x[2 * n], xn1 = x[2 * n + 1], yn0 = y[2 * n], yn1 = y[2 * n + 1];
    for (int mu = 0; mu < 2; mu++) {
        const int r = nb_right(n, mu, g.nx, g.nt);
        const double s = wraps_right(n, mu, g.nt) ? -1.0 : 1.0;
        const c64 u = U[2 * n + mu];
        const c64 xr0 = x[2 * r], xr1 = x[2 * r + 1], yr0 = y[2 * r], yr1 = y[2 * r + 1];
        c64 t1, t2;
        if (mu == 0) {  // sigma_x: (1-s)v = (v0-v1)(1,-1), (1+s)v = (v0+v1)(1,1)
            t1 = mulc(yr0 - yr1, xn0 - xn1);
            t2 = mulc(yn0 + yn1, xr0 + xr1);
        } else {  // sigma_y: (1-s)v = (v0+iv1)(1,-i), (1+s)v = (v0-iv1)(1,i)
            t1 = mulc(yr0 + mul_i(yr1), xn0 + mul_i(xn1));
            t2 = mulc(yn0 - mul_i(yn1), xr0 - mul_i(xr1));
        }
        const c64 z = u * t1 - conj(u) * t2;  // F = Re(i s z) = -s Im z
        F[(size_t)blockIdx.y * g.ns * 2 + 2 * n + mu] = -s * z.im;
    }
}

// out = a x + b y (real a, b; out may alias x or y)
template <typename R>
__global__ void k_lin(size_t n, double a, const cplx<R>* x, double b, const cplx<R>* y, cplx<R>* out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = R(a) * x[i] + R(b) * y[i];
}
template <typename R> __global__ void k_zero(size_t n, cplx<R>* out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = czero<R>();
}
// Per-rhs SAP skip flags at the start of a smoother application: 0, or 1
// where the caller's mask marks the rhs as finished (a kernel rather than a
// memset node, so that a captured K-cycle replays correctly)
__global__ void k_skip_init(int n, const int* __restrict__ mask, int* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = mask && mask[i] == 0 ? 1 : 0;
}

}  // namespace amgs
