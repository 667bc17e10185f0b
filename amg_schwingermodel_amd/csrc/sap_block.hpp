// Level-0 SAP with one workgroup per item (Schwarz block of a colour x rhs):
// the item's whole restarted block GMRES of one colour sweep runs inside the
// workgroup (SAP_C::I_D_B_1_It / GMRES_D_B, src/sap.cpp:44-174, include/sap.h:27-47,
// with the control flow of src/fgmres.cpp:5-88 on D_B).
//
// The block vector that the stencil reads (the new Krylov vector of an
// Arnoldi step) sits in an LDS tile of the whole block, so D_B never reads a
// neighbour from memory; the rest of the basis streams through HBM; W = D_B
// Vt_j stays in registers from the step that forms it to the next step's
// orthogonalisation, so it never goes to memory; every inner product is a
// workgroup reduction in a fixed order, and the item's scalars (Hessenberg
// column, Givens rotations, break test, back-substitution) are computed by
// one thread from LDS.  One launch per colour replaces the five kernels per
// Arnoldi step and restart of the chunked form (k_sap_init / arnoldi_a /
// arnoldi_b / update / resid), with the same arithmetic per item: the
// one-pass modified Gram-Schmidt algebra, the FP16 slots with power-of-two
// scales, the reference's restart and stale-maxIt rules.
//
// Conditions (checked by the host): square lattice, sites in pairs along t
// (even block extent in t, blocks that span neither direction), the block's
// tile within the LDS budget, and sspb / 2 pairs spread evenly as PP pairs
// per thread over a multiple of 64 threads (at most 1024).
#pragma once
#include "fine.hpp"

namespace amgs {

constexpr int SBT = 1024;                     // threads per workgroup, at most
constexpr int SBQ = 2 * SAP_MAXM + 2;         // reduced quantities per step, at most
constexpr size_t SB_TILE_MAX = 128 * 1024;    // LDS bytes of the block tile, at most

// both spins of a site: FP16 storage as one f16x4, else two complex values
template <typename R> struct Spin2 {
    cplx<R> a, b;
};
template <typename R, bool H> using sb_tile_t = typename std::conditional<H, f16x4, Spin2<R>>::type;
template <typename R, bool H> constexpr size_t sb_site_bytes() { return sizeof(sb_tile_t<R, H>); }

// a pair (sites n, n+1; spins 0, 1) in the arithmetic precision
template <typename R> struct Pair4 {
    cplx<R> v[4];
};

template <typename R, bool H> struct SbIO {
    // Krylov slot i of the item at global site n (pair n, n+1)
    static __device__ __forceinline__ Pair4<R> load(const SapArgs<R>& a, int i, int rhs, unsigned n) {
        Pair4<R> p;
        if constexpr (H) {
            const f16x8 h = *reinterpret_cast<const f16x8*>(hslot(a, i, rhs) + n);
            p.v[0] = c32{float(h.s0), float(h.s1)};
            p.v[1] = c32{float(h.s2), float(h.s3)};
            p.v[2] = c32{float(h.s4), float(h.s5)};
            p.v[3] = c32{float(h.s6), float(h.s7)};
        } else {
            const cplx<R>* s = slot(a, i, rhs) + 2u * n;
            p.v[0] = s[0];
            p.v[1] = s[1];
            p.v[2] = s[2];
            p.v[3] = s[3];
        }
        return p;
    }
    // the tile entries of local sites m, m+1
    static __device__ __forceinline__ Pair4<R> tile_pair(const sb_tile_t<R, H>* tile, int m) {
        Pair4<R> p;
        if constexpr (H) {
            const f16x8 h = *reinterpret_cast<const f16x8*>(tile + m);
            p.v[0] = c32{float(h.s0), float(h.s1)};
            p.v[1] = c32{float(h.s2), float(h.s3)};
            p.v[2] = c32{float(h.s4), float(h.s5)};
            p.v[3] = c32{float(h.s6), float(h.s7)};
        } else {
            p.v[0] = tile[m].a;
            p.v[1] = tile[m].b;
            p.v[2] = tile[m + 1].a;
            p.v[3] = tile[m + 1].b;
        }
        return p;
    }
    // store u as Krylov slot i (FP16: sg u rounded) and into the tile; the
    // stored values come back in u, their squared norm is returned
    static __device__ __forceinline__ double store(const SapArgs<R>& a, int i, int rhs, unsigned n, sb_tile_t<R, H>* tile,
                                                   int m, Pair4<R>& u, float sg) {
        if constexpr (H) {
            const f16x4 h0 = to_h16(c32{sg * u.v[0].re, sg * u.v[0].im}, c32{sg * u.v[1].re, sg * u.v[1].im});
            const f16x4 h1 = to_h16(c32{sg * u.v[2].re, sg * u.v[2].im}, c32{sg * u.v[3].re, sg * u.v[3].im});
            const f16x8 h{h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
            *reinterpret_cast<f16x8*>(hslot(a, i, rhs) + n) = h;
            *reinterpret_cast<f16x8*>(tile + m) = h;
            u.v[0] = c32{float(h.s0), float(h.s1)};
            u.v[1] = c32{float(h.s2), float(h.s3)};
            u.v[2] = c32{float(h.s4), float(h.s5)};
            u.v[3] = c32{float(h.s6), float(h.s7)};
            double q = 0;
#pragma unroll
            for (int e = 0; e < 4; e++) q += double(u.v[e].re) * double(u.v[e].re) + double(u.v[e].im) * double(u.v[e].im);
            return q;
        } else {
            cplx<R>* s = slot(a, i, rhs) + 2u * n;
            s[0] = u.v[0];
            s[1] = u.v[1];
            s[2] = u.v[2];
            s[3] = u.v[3];
            tile[m] = Spin2<R>{u.v[0], u.v[1]};
            tile[m + 1] = Spin2<R>{u.v[2], u.v[3]};
            return double(norm2(u.v[0])) + double(norm2(u.v[1])) + double(norm2(u.v[2])) + double(norm2(u.v[3]));
        }
    }
};

// spin projections of a tile entry: p0 +- p1 and p0 +- i p1 (neg: the minus sign)
template <typename R> __device__ __forceinline__ cplx<R> sb_sum(const Spin2<R>& v, bool neg) {
    return neg ? v.a - v.b : v.a + v.b;
}
template <typename R> __device__ __forceinline__ cplx<R> sb_sumi(const Spin2<R>& v, bool neg) {
    return neg ? v.a - mul_i(v.b) : v.a + mul_i(v.b);
}

// D_B at the pair of local sites m, m+1 (b: the coordinates of site m, its
// local t even) from the block tile; the links come from memory.  Within a
// block that spans neither direction no in-block hop wraps around the
// lattice, so the antiperiodic signs never meet an unmasked term.  A
// masked-out neighbour reads the pair's own entry; its term is zero.
template <typename R, bool H>
__device__ __forceinline__ void sb_pair_stencil(const SapArgs<R>& a, const cplx<R>* __restrict__ Ut, R mm, const BSite& b,
                                                int m, const sb_tile_t<R, H>* tile, Pair4<R>& o, Pair4<R>& c) {
    const Geo& g = a.g;
    const unsigned n = (unsigned)b.n;
    const bool ml0 = b.mt > 0, mr1t = b.mt + 2 < g.stel, mx_r = b.mx + 1 < g.sxel, mx_l = b.mx > 0;
    const int st = g.stel;
    const unsigned l0 = ml0 ? n - 1u : n, xl = mx_l ? n - (unsigned)g.nt : n;
    using F4 = typename std::conditional<std::is_same<R, float>::value, float __attribute__((ext_vector_type(4))),
                                         double __attribute__((ext_vector_type(4)))>::type;
    const F4* U4 = reinterpret_cast<const F4*>(Ut);
    const F4 ua = U4[n], ub = U4[n + 1u];  // t and x links of n, of n+1
    const cplx<R> ul0 = Ut[2u * l0], ux0 = Ut[2u * xl + 1u], ux1 = Ut[2u * xl + 3u];
    const cplx<R> f00{ua.x, ua.y}, f01{ua.z, ua.w}, f10{ub.x, ub.y}, f11{ub.z, ub.w};
    const cplx<R> b00 = conj(ul0), b10 = conj(f00), b01 = conj(ux0), b11 = conj(ux1);
    const int mr = mx_r ? m + st : m, ml = mx_l ? m - st : m;
    if constexpr (H) {
        const f16x8 cp = *reinterpret_cast<const f16x8*>(tile + m);
        const f16x8 rp = *reinterpret_cast<const f16x8*>(tile + mr);
        const f16x8 lp = *reinterpret_cast<const f16x8*>(tile + ml);
        const f16x4 tL = tile[ml0 ? m - 1 : m], tR = tile[mr1t ? m + 2 : m + 1];
        const f16x4 h0 = cp.lo, h1 = cp.hi;
        c.v[0] = c32{float(h0.x), float(h0.y)};
        c.v[1] = c32{float(h0.z), float(h0.w)};
        c.v[2] = c32{float(h1.x), float(h1.y)};
        c.v[3] = c32{float(h1.z), float(h1.w)};
        const H16Proj P;
        wilson_combine<float, false>(c.v[0], c.v[1], P.sum(h1, true), P.sumi(rp.lo, false), P.sum(tL, false),
                                     P.sumi(lp.lo, true), f00, f01, b00, b01, true, mx_r, ml0, mx_l, mm, o.v[0], o.v[1]);
        wilson_combine<float, false>(c.v[2], c.v[3], P.sum(tR, true), P.sumi(rp.hi, false), P.sum(h0, false),
                                     P.sumi(lp.hi, true), f10, f11, b10, b11, mr1t, mx_r, true, mx_l, mm, o.v[2], o.v[3]);
    } else {
        const Spin2<R> s0 = tile[m], s1 = tile[m + 1], r0 = tile[mr], r1 = tile[mr + 1], q0 = tile[ml],
                       q1 = tile[ml + 1];
        const Spin2<R> tL = tile[ml0 ? m - 1 : m], tR = tile[mr1t ? m + 2 : m + 1];
        c.v[0] = s0.a;
        c.v[1] = s0.b;
        c.v[2] = s1.a;
        c.v[3] = s1.b;
        wilson_combine<R, false>(s0.a, s0.b, sb_sum(s1, true), sb_sumi(r0, false), sb_sum(tL, false), sb_sumi(q0, true),
                                 f00, f01, b00, b01, true, mx_r, ml0, mx_l, mm, o.v[0], o.v[1]);
        wilson_combine<R, false>(s1.a, s1.b, sb_sum(tR, true), sb_sumi(r1, false), sb_sum(s0, false), sb_sumi(q1, true),
                                 f10, f11, b10, b11, mr1t, mx_r, true, mx_l, mm, o.v[2], o.v[3]);
    }
}

// Block-local coordinates of local site m (global site n known): the
// reciprocal-based split of bsite() without recomputing n.
struct BSiteMap {
    int stel, x0, t0;
    float rcp_stel;
    __device__ __forceinline__ BSite operator()(int m, unsigned n) const {
        BSite b;
        b.mx = (int)((float(m) + 0.5f) * rcp_stel);
        b.mt = m - b.mx * stel;
        b.x = x0 + b.mx;
        b.t = t0 + b.mt;
        b.n = (int)n;
        return b;
    }
};

// W of a pair kept in registers between steps (FP16: the stored rounding)
template <typename R, bool H> struct SbW;
template <> struct SbW<float, true> {
    f16x8 h;
    __device__ __forceinline__ void set(Pair4<float>& o) {
        const f16x4 a = to_h16(o.v[0], o.v[1]), b = to_h16(o.v[2], o.v[3]);
        h = f16x8{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        o.v[0] = c32{float(h.s0), float(h.s1)};
        o.v[1] = c32{float(h.s2), float(h.s3)};
        o.v[2] = c32{float(h.s4), float(h.s5)};
        o.v[3] = c32{float(h.s6), float(h.s7)};
    }
    __device__ __forceinline__ Pair4<float> get() const {
        Pair4<float> p;
        p.v[0] = c32{float(h.s0), float(h.s1)};
        p.v[1] = c32{float(h.s2), float(h.s3)};
        p.v[2] = c32{float(h.s4), float(h.s5)};
        p.v[3] = c32{float(h.s6), float(h.s7)};
        return p;
    }
};
template <typename R> struct SbW<R, false> {
    Pair4<R> w;
    __device__ __forceinline__ void set(Pair4<R>& o) { w = o; }
    __device__ __forceinline__ Pair4<R> get() const { return w; }
};

// W[k] for a run-time pair index k of a loop that is not unrolled: each case
// names its element, so W stays in registers while the loop body (the heavy
// part) exists once and the compiler cannot lift work of all pairs out of it.
template <int PP, typename Wt, typename F> __device__ __forceinline__ void sb_with(Wt (&W)[PP], int k, F f) {
    switch (k) {
#define AMGS_SB_CASE(i)                \
    case i:                            \
        if constexpr (i < PP) f(W[i]); \
        break;
        AMGS_SB_CASE(0) AMGS_SB_CASE(1) AMGS_SB_CASE(2) AMGS_SB_CASE(3)
        AMGS_SB_CASE(4) AMGS_SB_CASE(5) AMGS_SB_CASE(6) AMGS_SB_CASE(7)
#undef AMGS_SB_CASE
    }
}

// Park per-thread partial q of the step (wave sum in FP64) in LDS.
__device__ __forceinline__ void sb_park(c64 v, int q, c64 (*park)[SBQ]) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) park[threadIdx.x >> 6][q] = v;
}
// Thread 0: the workgroup total of quantity q (waves in order).
__device__ __forceinline__ c64 sb_total(c64 (*park)[SBQ], int q) {
    const int nw = blockDim.x >> 6;
    c64 r = park[0][q];
    for (int w = 1; w < nw; w++) r += park[w][q];
    return r;
}

// The thread's pairs: pair k has block-local first site m0 + k mstep and
// global site n0 + k nstep (mstep a multiple of the block's t extent, so pair
// k lies mstep / stel rows further along x).
struct SbPairs {
    unsigned n0, nstep;
    int m0, mstep;
    __device__ __forceinline__ unsigned n(int k) const { return n0 + (unsigned)k * nstep; }
    __device__ __forceinline__ int m(int k) const { return m0 + k * mstep; }
};

// Everything an Arnoldi step of the item needs besides W.
template <typename R, bool H> struct SbItem {
    const SapArgs<R>& a;
    int rhs;
    SbPairs P;
    sb_tile_t<R, H>* tile;
    ItemLds& L;
    c64 (*park)[SBQ];
    const cplx<R>* Ut;
    R mm;
    float sg;
};

// u = s_{J-1} W - sum_{i<J} eta_i Vt_i for every pair (one-pass MGS, the
// update of step B): with STORE the result is Vt_J (memory and tile), else
// only its squared norm as it would be stored.  Two pairs per iteration with
// their J loads each in flight together.  Returns the thread's |stored|^2.
template <typename R, bool H, int PP, int J, bool STORE>
__device__ __forceinline__ double sb_orthogonalise(const SbItem<R, H>& it, SbW<R, H> (&W)[PP]) {
    using IO = SbIO<R, H>;
    const R sj = R(it.L.s[J - 1]);
    cplx<R> eta[J];
#pragma unroll
    for (int i = 0; i < J; i++) eta[i] = cvt<R>(it.L.eta[i]);
    constexpr int K2 = PP >= 2 ? 2 : 1;
    double nn = 0;
#pragma unroll 1
    for (int kb = 0; kb < PP; kb += K2) {
        Pair4<R> u[K2];
#pragma unroll
        for (int q = 0; q < K2; q++) sb_with(W, kb + q, [&](const SbW<R, H>& w) { u[q] = w.get(); });
#pragma unroll
        for (int q = 0; q < K2; q++)
#pragma unroll
            for (int e = 0; e < 4; e++) u[q].v[e] = sj * u[q].v[e];
#pragma unroll
        for (int i = 0; i < J; i++)
#pragma unroll
            for (int q = 0; q < K2; q++) {
                const Pair4<R> v = IO::load(it.a, i, it.rhs, it.P.n(kb + q));
#pragma unroll
                for (int e = 0; e < 4; e++) cmsb(u[q].v[e], eta[i], v.v[e]);
            }
#pragma unroll
        for (int q = 0; q < K2; q++) {
            if constexpr (STORE) {
                nn += IO::store(it.a, J, it.rhs, it.P.n(kb + q), it.tile, it.P.m(kb + q), u[q], it.sg);
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    if constexpr (H) {
                        const _Float16 re = (_Float16)(it.sg * u[q].v[e].re), im = (_Float16)(it.sg * u[q].v[e].im);
                        nn += double(re) * double(re) + double(im) * double(im);
                    } else {
                        nn += double(norm2(u[q].v[e]));
                    }
                }
            }
        }
    }
    return nn;
}

// Arnoldi step J of a cycle (src/fgmres.cpp:20-49): Vt_J (J > 0), W = D_B Vt_J
// into registers, the inner products <W, Vt_i> (i <= J) and <Vt_J, Vt_i>
// (i < J), then thread 0 finalises step J-1 and forms column J.  Returns
// false when the cycle ends here (breakdown or the residual estimate met).
template <typename R, bool H, int PP, int J>
__device__ __forceinline__ bool sb_step(const SbItem<R, H>& it, SbW<R, H> (&W)[PP], const BSiteMap& bm, int M,
                                        double tol, double sigd, int& stop, double& bytes) {
    using IO = SbIO<R, H>;
    constexpr double S = (double)sizeof(sb_tile_t<R, H>);
    double nn = 0;
    if constexpr (J > 0) {
        nn = sb_orthogonalise<R, H, PP, J, true>(it, W);
        bytes += S * (J + 1);  // Vt_0..Vt_{J-1} read, Vt_J written
    }
    __syncthreads();  // the tile holds Vt_J
    cplx<R> pw[J + 1], pv[J > 0 ? J : 1];
#pragma unroll
    for (int i = 0; i <= J; i++) pw[i] = czero<R>();
#pragma unroll
    for (int i = 0; i < (J > 0 ? J : 1); i++) pv[i] = czero<R>();
#pragma unroll 1
    for (int k = 0; k < PP; k++) {
        const int m = it.P.m(k);
        const unsigned n = it.P.n(k);
        Pair4<R> o, c;
        sb_pair_stencil<R, H>(it.a, it.Ut, it.mm, bm(m, n), m, it.tile, o, c);
        sb_with(W, k, [&](SbW<R, H>& w) { w.set(o); });
#pragma unroll
        for (int e = 0; e < 4; e++) cmadc(pw[J], o.v[e], c.v[e]);
#pragma unroll
        for (int i = 0; i < J; i++) {
            const Pair4<R> v = IO::load(it.a, i, it.rhs, n);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                cmadc(pw[i], o.v[e], v.v[e]);
                cmadc(pv[i], c.v[e], v.v[e]);
            }
        }
    }
    bytes += S * J;  // Vt_0..Vt_{J-1} read again
#pragma unroll
    for (int i = 0; i <= J; i++) sb_park(cvt<double>(pw[i]), i, it.park);
#pragma unroll
    for (int i = 0; i < J; i++) sb_park(cvt<double>(pv[i]), J + 1 + i, it.park);
    sb_park(c64{nn, 0.0}, SBQ - 1, it.park);
    __syncthreads();
    if (threadIdx.x == 0) {
        ItemLds& L = it.L;
        // finalise step J-1 with H_{J,J-1} = |Vt_J| / sig, then column J
        if (J > 0) {
            sap_finalize_step(it.a.st, 0, sb_total(it.park, SBQ - 1).re, M, J - 1, tol, false, L, sigd);
            if (L.broke) {
                L.jbreak = J - 1;
                L.maxIt = J;
                stop = 1;
            }
        }
        if (!stop) {
            for (int q = 0; q <= J; q++) L.dq[q] = sb_total(it.park, q);
            for (int q = 0; q < J; q++) L.dq[J + 1 + q] = sb_total(it.park, J + 1 + q);
            sap_hessenberg_column(L, M, J);
            for (int i = 0; i <= J; i++) L.Hraw[i * M + J] = L.Hcol[i];
        }
    }
    __syncthreads();
    return !stop;
}

// |Vt_M|^2 of the vector step B would store after the last step of a cycle
// (M run-time, the inner loops compiled for each possible M).
template <typename R, bool H, int PP>
__device__ __forceinline__ double sb_last_norm(const SbItem<R, H>& it, SbW<R, H> (&W)[PP], int M) {
    switch (M) {
#define AMGS_SB_NORM(j) \
    case j: return sb_orthogonalise<R, H, PP, j, false>(it, W);
        AMGS_SB_NORM(1) AMGS_SB_NORM(2) AMGS_SB_NORM(3) AMGS_SB_NORM(4)
        AMGS_SB_NORM(5) AMGS_SB_NORM(6) AMGS_SB_NORM(7) AMGS_SB_NORM(8)
#undef AMGS_SB_NORM
    }
    return 0.0;
}

// One colour of one SAP sweep for every item: x_B = b_B, then `restarts`
// cycles of GMRES(m) on D_B x_B = b_B; after the last cycle x_B is added to
// xacc (src/sap.cpp:117-125).  Grid (colour blocks, rhs); block sspb / (2 PP)
// threads; dynamic LDS: the block tile.  act (may be null): the algorithmic
// bytes per block site moved by each item, summed over the items.
template <typename R, bool H, int PP>
__global__ void __launch_bounds__(SBT) k_sap_block(SapArgs<R> a, cplx<R>* __restrict__ xacc, int restarts, int* act) {
    using T = sb_tile_t<R, H>;
    using IO = SbIO<R, H>;
    extern __shared__ __align__(16) char sb_dyn[];
    T* tile = reinterpret_cast<T*>(sb_dyn);
    __shared__ ItemLds L;
    __shared__ c64 park[SBT / 64][SBQ];
    __shared__ double sig0;  // FP16 storage: scale of the stored Vt_0
    __shared__ int stop;     // the current cycle broke down
    const int rhs = blockIdx.y;
    if (a.skip && a.skip[rhs]) return;
    ItemIdx t;
    t.chunk = 0;
    t.rhs = rhs;
    t.item = rhs * a.ncol + blockIdx.x;
    t.blk = a.blocks[blockIdx.x];
    {
        const int X = t.blk / a.g.sbt;
        t.x0 = X * a.g.sxel;
        t.t0 = (t.blk - X * a.g.sbt) * a.g.stel;
    }
    t.m_lo = 0;
    t.m_hi = a.g.sspb;
    t.mm = a.mmv ? (double)a.mmv[rhs] : (double)a.mm;
    const int M = a.m, tid = threadIdx.x, nth = blockDim.x;
    const size_t off = (size_t)rhs * a.g.ns * 2;
    const cplx<R>* __restrict__ bg = a.rglob + off;
    cplx<R>* __restrict__ xl = a.xl + off;
    const double sigd = H ? (double)a.sig : 1.0;
    constexpr double S = (double)sizeof(T);  // bytes per site of a Krylov slot
    double bytes = 0;                        // algorithmic bytes per block site (accounting)
    SbPairs P;
    P.m0 = 2 * tid;
    P.mstep = 2 * nth;
    P.n0 = (unsigned)bsite(a, t, P.m0).n;
    P.nstep = (unsigned)(P.mstep / a.g.stel * a.g.nt);
    const BSiteMap bm{a.g.stel, t.x0, t.t0, a.rcp_stel};
    const SbItem<R, H> it{a, rhs, P, tile, L, park, a.U + (size_t)rhs * a.ucs, R(t.mm), H ? a.sig : 1.0f};
    SbW<R, H> W[PP];
    // pairs whose accumulators are live at once in the update loop
    constexpr int KB = PP < 2 ? PP : 2;

    // ---- x_B = b_B, Vt_0 = b_B - D_B b_B (src/sap.cpp:61, src/fgmres.cpp:8-17) ----
    {
        const double sgi = H ? a.rsig[rhs] : 1.0;
        double qr = 0, qb = 0;
#pragma unroll 1
        for (int k = 0; k < PP; k++) {
            const int m = P.m(k);
            const BSite b0 = bm(m, P.n(k)), b1 = bm(m + 1, P.n(k) + 1u);
            Pair4<R> o, c;
            block_stencil<true>(a, t, b0, SpinPtr<R>{bg}, o.v[0], o.v[1], c.v[0], c.v[1]);
            block_stencil<true>(a, t, b1, SpinPtr<R>{bg}, o.v[2], o.v[3], c.v[2], c.v[3]);
            const unsigned n = P.n(k);
            Pair4<R> r;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                xl[2u * n + e] = c.v[e];
                r.v[e] = c.v[e] - o.v[e];
                qb += double(norm2(c.v[e]));
            }
            qr += IO::store(a, 0, rhs, n, tile, m, r, float(sgi));
        }
        sb_park(c64{qr, 0.0}, 0, park);
        sb_park(c64{qb, 0.0}, 1, park);
        __syncthreads();
        if (tid == 0) {
            const double est = sqrt(sb_total(park, 0).re), nb = sqrt(sb_total(park, 1).re);
            L.maxIt = M;
            L.jbreak = -1;
            L.done = nb == 0.0;
            L.normb = nb;
            L.s[0] = 1.0 / est;
            L.gpre[0] = c64{H ? est / sgi : est, 0.0};
            sig0 = sgi;
        }
        __syncthreads();
        bytes += 2 * sizeof(cplx<R>) * 2 + S;  // b read, x_B written, Vt_0 written
    }

    for (int kr = 0; kr < restarts; kr++) {
        const bool last = kr + 1 == restarts;
        if (!L.done) {
            if (tid == 0) stop = 0;
            // ---- the Arnoldi steps of the cycle (compiled per step index) ----
            bool go = true;
#define AMGS_SB_STEP(j) \
    if (go && j < M) go = sb_step<R, H, PP, j>(it, W, bm, M, a.tol, sigd, stop, bytes);
            AMGS_SB_STEP(0) AMGS_SB_STEP(1) AMGS_SB_STEP(2) AMGS_SB_STEP(3)
            AMGS_SB_STEP(4) AMGS_SB_STEP(5) AMGS_SB_STEP(6) AMGS_SB_STEP(7)
#undef AMGS_SB_STEP
            // ---- end of the cycle: H_{m,m-1}, back-substitution, x_B update ----
            if (go) {
                const double nn = sb_last_norm<R, H, PP>(it, W, M);
                bytes += S * M;
                sb_park(c64{nn, 0.0}, 0, park);
                __syncthreads();
                if (tid == 0) L.dr[0] = sb_total(park, 0).re;
            }
            if (tid == 0) {
                // src/fgmres.cpp:52-59 with the stale maxIt of earlier cycles (a
                // cycle that ran to m without breaking keeps it)
                int maxIt = L.maxIt;
                if (go) {
                    sap_finalize_step(a.st, t.item, L.dr[0], M, M - 1, a.tol, false, L, sigd);
                    if (L.broke) {
                        L.jbreak = M - 1;
                        L.maxIt = M;
                        maxIt = M;
                    }
                }
                sap_backsolve(L, M, maxIt);
                if (H) {
                    const double s0 = sap_next_sig0(L, maxIt, a.g.sspb);
                    if (s0 > 0) sig0 = s0;
                }
                L.active = maxIt;  // the coefficients of this update
            }
            __syncthreads();
            const int mi = L.active;
#pragma unroll 1
            for (int kb = 0; kb < PP; kb += KB) {
                Pair4<R> x[KB];
#pragma unroll
                for (int q = 0; q < KB; q++)
#pragma unroll
                    for (int e = 0; e < 4; e++) x[q].v[e] = xl[2u * P.n(kb + q) + e];
#pragma unroll 1
                for (int i = 0; i < mi; i++) {
                    const cplx<R> c = cvt<R>(L.eta[i]);
#pragma unroll
                    for (int q = 0; q < KB; q++) {
                        const Pair4<R> v = IO::load(a, i, rhs, P.n(kb + q));
#pragma unroll
                        for (int e = 0; e < 4; e++) cmad(x[q].v[e], c, v.v[e]);
                    }
                }
#pragma unroll
                for (int q = 0; q < KB; q++)
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        if (last) xacc[off + 2u * P.n(kb + q) + e] += x[q].v[e];
                        else xl[2u * P.n(kb + q) + e] = x[q].v[e];
                    }
            }
            bytes += S * mi + (last ? 2 * sizeof(cplx<R>) * 3 : 2 * sizeof(cplx<R>) * 2);
        } else if (last) {
            // converged in an earlier cycle: x_B goes into x unchanged
#pragma unroll 1
            for (int k = 0; k < PP; k++) {
                const unsigned n = P.n(k);
#pragma unroll
                for (int e = 0; e < 4; e++) xacc[off + 2u * n + e] += xl[2u * n + e];
            }
            bytes += 2 * sizeof(cplx<R>) * 3;
        }
        if (last || L.done) continue;
        // ---- next cycle: Vt_0 = b_B - D_B x_B, its start (src/fgmres.cpp:61-65, 8-17) ----
        __syncthreads();  // x_B of the whole block written
        {
            const float s0 = float(sig0);
            double qr = 0;
#pragma unroll 1
            for (int k = 0; k < PP; k++) {
                const int m = P.m(k);
                const unsigned n = P.n(k);
                const BSite b0 = bm(m, n), b1 = bm(m + 1, n + 1u);
                Pair4<R> o;
                block_stencil<true>(a, t, b0, SpinPtr<R>{xl}, o.v[0], o.v[1]);
                block_stencil<true>(a, t, b1, SpinPtr<R>{xl}, o.v[2], o.v[3]);
                Pair4<R> r;
#pragma unroll
                for (int e = 0; e < 4; e++) r.v[e] = bg[2u * n + e] - o.v[e];
                qr += IO::store(a, 0, rhs, n, tile, m, r, H ? s0 : 1.0f);
            }
            sb_park(c64{qr, 0.0}, 0, park);
            __syncthreads();
            if (tid == 0) {
                const double est = sqrt(sb_total(park, 0).re), err = H ? est / sig0 : est;
                if (err < a.tol * L.normb || L.normb == 0.0) {
                    L.done = 1;
                } else {
                    L.s[0] = 1.0 / est;
                    L.gpre[0] = c64{err, 0.0};
                    L.jbreak = -1;
                }
            }
            __syncthreads();
            bytes += 2 * sizeof(cplx<R>) * 2 + S;  // x_B, b read, Vt_0 written
        }
    }
    if (act && tid == 0) atomicAdd(act, (int)bytes);
}

}  // namespace amgs
