// Level-0 SAP with one workgroup per item (Schwarz block of a colour x rhs),
// streaming form for the FP16 Krylov basis.  The item runs the same
// restarted block GMRES as k_sap_block (SAP_C::I_D_B_1_It / GMRES_D_B,
// src/sap.cpp:44-174, with the control flow of src/fgmres.cpp:5-88 on D_B):
// the one-pass modified Gram-Schmidt algebra, FP16 slots with power-of-two
// scales, the reference's restart and stale-maxIt rules.  It moves fewer
// bytes per restart cycle:
//
// * Each thread owns one pair of sites along t (pair l of the block's
//   TL = stel / 2) in RW consecutive block rows, its run (block-local
//   x = rho RW + k).  A row's x-neighbours are the thread's own previous and
//   next rows and its t-neighbours sit in the adjacent lanes (lane shuffles);
//   only the first and last row of every run pass through LDS.
// * An Arnoldi step is one pass over the rows: the new vector
//   Vt_J = s_{J-1} W - sum_i eta_i Vt_i of row k+1, the stencil W = D_B Vt_J
//   of row k and the inner products of row k share one load of each stored
//   slot Vt_0..Vt_{J-1} (k_sap_block loads them twice per step).
// * After the last step of a cycle the norm of the vector the next step would
//   store, |s W - sum eta_i Vt_i|, follows from that step's inner products,
//   the Gram entries and |W|^2 instead of another pass over the basis.
// * The x_B update and the next cycle's residual b_B - D_B x_B are one pass.
//
// Inner products take the FP16 operands as stored, two v_dot2_f32_f16 per
// complex product.  A workgroup has at most 512 threads so that a thread's W
// rows and the streaming state fit its registers (256 per lane).
//
// Conditions (checked by the host): square lattice, blocks that span neither
// direction, stel even with TL dividing 64, sxel a multiple of RW,
// TL * sxel / RW threads, a multiple of 64 and at most SRT, and a block GMRES
// restart length of at most SR_MAXM.
#pragma once
#include "sap_block.hpp"

namespace amgs {

typedef int sr_i32x2 __attribute__((ext_vector_type(2)));
typedef int sr_i32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

constexpr int SRT = 512;  // threads per workgroup, at most

// dynamic LDS bytes per thread: the two boundary rows of its run in FP16
// (Krylov vectors) and in FP32 (x_B at a restart)
constexpr size_t SR_LDS_PER_THREAD = 2 * sizeof(f16x8) + 2 * sizeof(Pair4<float>);
// restart length of the block GMRES, at most (the reference's default; the
// registers of longer cycles do not fit the 1024-thread workgroup)
constexpr int SR_MAXM = 5;

// Global accesses as a wave-uniform base plus a 32-bit byte offset: the
// loads and stores then take the scalar-base + vector-offset form and no lane
// holds a 64-bit address (every offset here is below 2^31).
template <typename T> __device__ __forceinline__ T sr_ld(const void* base, unsigned off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <typename T> __device__ __forceinline__ void sr_st(void* base, unsigned off, const T& v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v;
}
typedef float sr_f4 __attribute__((ext_vector_type(4)));
// both spins of the sites n, n+1 (n even) of an FP32 vector [ns][2]
__device__ __forceinline__ Pair4<float> sr_ld_pair(const c32* base, unsigned n) {
    const sr_f4 a = sr_ld<sr_f4>(base, 16u * n), b = sr_ld<sr_f4>(base, 16u * n + 16u);
    Pair4<float> p;
    p.v[0] = c32{a.x, a.y};
    p.v[1] = c32{a.z, a.w};
    p.v[2] = c32{b.x, b.y};
    p.v[3] = c32{b.z, b.w};
    return p;
}
__device__ __forceinline__ void sr_st_pair(c32* base, unsigned n, const Pair4<float>& p) {
    sr_st<sr_f4>(base, 16u * n, sr_f4{p.v[0].re, p.v[0].im, p.v[1].re, p.v[1].im});
    sr_st<sr_f4>(base, 16u * n + 16u, sr_f4{p.v[2].re, p.v[2].im, p.v[3].re, p.v[3].im});
}

__device__ __forceinline__ Pair4<float> sr_unpack(const f16x8 h) {
    Pair4<float> p;
    p.v[0] = c32{float(h.s0), float(h.s1)};
    p.v[1] = c32{float(h.s2), float(h.s3)};
    p.v[2] = c32{float(h.s4), float(h.s5)};
    p.v[3] = c32{float(h.s6), float(h.s7)};
    return p;
}
// sg u rounded to FP16
__device__ __forceinline__ f16x8 sr_pack(const Pair4<float>& u, float sg) {
    const f16x4 a = to_h16(c32{sg * u.v[0].re, sg * u.v[0].im}, c32{sg * u.v[1].re, sg * u.v[1].im});
    const f16x4 b = to_h16(c32{sg * u.v[2].re, sg * u.v[2].im}, c32{sg * u.v[3].re, sg * u.v[3].im});
    return f16x8{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}
// |h|^2 of a stored pair (8 halves, exact products, FP32 sums)
__device__ __forceinline__ float sr_norm2(const f16x8 h) {
    float q = __builtin_amdgcn_fdot2(h.s01, h.s01, 0.f, false);
    q = __builtin_amdgcn_fdot2(h.s23, h.s23, q, false);
    q = __builtin_amdgcn_fdot2(h.s45, h.s45, q, false);
    return __builtin_amdgcn_fdot2(h.s67, h.s67, q, false);
}
// The four complex entries of a stored pair and their companions
// (im, -re): acc += a conj(b) is re += a.b, im += (im a, -re a).b
struct SrC4 {
    h2 v[4], s[4];
    __device__ __forceinline__ explicit SrC4(const f16x8 h) {
        v[0] = h.s01;
        v[1] = h.s23;
        v[2] = h.s45;
        v[3] = h.s67;
#pragma unroll
        for (int e = 0; e < 4; e++) s[e] = h2{v[e].y, -v[e].x};
    }
};
// acc += sum_e a_e conj(b_e) over the pair's four entries
__device__ __forceinline__ void sr_dotc(c32& acc, const SrC4& a, const f16x8 b) {
    const h2 bv[4] = {b.s01, b.s23, b.s45, b.s67};
#pragma unroll
    for (int e = 0; e < 4; e++) {
        acc.re = __builtin_amdgcn_fdot2(a.v[e], bv[e], acc.re, false);
        acc.im = __builtin_amdgcn_fdot2(a.s[e], bv[e], acc.im, false);
    }
}
// the site before the pair (lane - 1's second site) and after it (lane + 1's
// first site); lanes at a run's t-edges receive a value they do not use.  The
// pair is reinterpreted whole (words 0-1: first site, 2-3: second): a
// bit_cast of a .lo / .hi swizzle of it has been compiled as the .lo half
// for both.
__device__ __forceinline__ void sr_tnb(const f16x8 h, f16x4& tl, f16x4& tr) {
    const sr_i32x4 w = __builtin_bit_cast(sr_i32x4, h);
    const sr_i32x2 l{__shfl_up(w.z, 1u, 64), __shfl_up(w.w, 1u, 64)};
    const sr_i32x2 r{__shfl_down(w.x, 1u, 64), __shfl_down(w.y, 1u, 64)};
    tl = __builtin_bit_cast(f16x4, l);
    tr = __builtin_bit_cast(f16x4, r);
}
__device__ __forceinline__ c32 sr_shfl_up(c32 v) { return c32{__shfl_up(v.re, 1u, 64), __shfl_up(v.im, 1u, 64)}; }
__device__ __forceinline__ c32 sr_shfl_down(c32 v) {
    return c32{__shfl_down(v.re, 1u, 64), __shfl_down(v.im, 1u, 64)};
}

// The links a pair's stencil reads: both links of n and n+1 (the two loads
// of a row, ua and ub), the t link of n-1 and the x links of the pair one row
// down.  The t link of n-1 is the left lane's second site's (a lane shuffle)
// and the x links one row down are the thread's previous row's own (carried
// in SrXCarry); lanes and rows whose hop is masked out get a value they do
// not use.
struct SrLinks {
    c32 f00, f01, f10, f11, b00, b10, b01, b11;
    __device__ __forceinline__ SrLinks(const sr_f4 ua, const sr_f4 ub, c32 ux0, c32 ux1) {
        f00 = c32{ua.x, ua.y};
        f01 = c32{ua.z, ua.w};
        f10 = c32{ub.x, ub.y};
        f11 = c32{ub.z, ub.w};
        b00 = conj(sr_shfl_up(f10));
        b10 = conj(f00);
        b01 = conj(ux0);
        b11 = conj(ux1);
    }
};
// x links (direction 1) of the pair one row down, for the next row: loaded
// for the run's first row (its own pair's where the row below lies outside
// the block), then taken from each row's own links.
struct SrXCarry {
    c32 x0, x1;
    __device__ __forceinline__ SrXCarry(const c32* __restrict__ Ut, unsigned n0, unsigned nt, bool below) {
        const unsigned xl = below ? n0 - nt : n0;
        x0 = sr_ld<c32>(Ut, 16u * xl + 8u);
        x1 = sr_ld<c32>(Ut, 16u * xl + 24u);
    }
    __device__ __forceinline__ void next(const sr_f4 ua, const sr_f4 ub) {
        x0 = c32{ua.z, ua.w};
        x1 = c32{ub.z, ub.w};
    }
};

// D_B at the pair (sites n, n+1) from FP16 entries as stored: cp the pair,
// rp / lp the pairs one row up / down in x, tl / tr the sites n-1 / n+2.
// Masks: mx_r / mx_l the x-rows in the block, ml0 site n-1, mr1t site n+2.
// The block spans neither direction, so no in-block hop meets the
// antiperiodic sign.  c returns the pair's own values.
__device__ __forceinline__ void sr_stencil16(const SrLinks& K, float mm, const f16x8 cp, const f16x8 rp, const f16x8 lp,
                                             const f16x4 tl, const f16x4 tr, bool mx_r, bool mx_l, bool ml0, bool mr1t,
                                             Pair4<float>& o, Pair4<float>& c) {
    const f16x4 h0 = cp.lo, h1 = cp.hi;
    c = sr_unpack(cp);
    const H16Proj P;
    wilson_combine<float, false>(c.v[0], c.v[1], P.sum(h1, true), P.sumi(rp.lo, false), P.sum(tl, false),
                                 P.sumi(lp.lo, true), K.f00, K.f01, K.b00, K.b01, true, mx_r, ml0, mx_l, mm, o.v[0],
                                 o.v[1]);
    wilson_combine<float, false>(c.v[2], c.v[3], P.sum(tr, true), P.sumi(rp.hi, false), P.sum(h0, false),
                                 P.sumi(lp.hi, true), K.f10, K.f11, K.b10, K.b11, mr1t, mx_r, true, mx_l, mm, o.v[2],
                                 o.v[3]);
}
// The same on FP32 entries (x_B at a restart); tl / tr: both spins of n-1 / n+2.
__device__ __forceinline__ void sr_stencil32(const SrLinks& K, float mm, const Pair4<float>& cp,
                                             const Pair4<float>& rp, const Pair4<float>& lp, const Spin2<float>& tl,
                                             const Spin2<float>& tr, bool mx_r, bool mx_l, bool ml0, bool mr1t,
                                             Pair4<float>& o) {
    const Spin2<float> s0{cp.v[0], cp.v[1]}, s1{cp.v[2], cp.v[3]};
    const Spin2<float> r0{rp.v[0], rp.v[1]}, r1{rp.v[2], rp.v[3]}, q0{lp.v[0], lp.v[1]}, q1{lp.v[2], lp.v[3]};
    wilson_combine<float, false>(s0.a, s0.b, sb_sum(s1, true), sb_sumi(r0, false), sb_sum(tl, false),
                                 sb_sumi(q0, true), K.f00, K.f01, K.b00, K.b01, true, mx_r, ml0, mx_l, mm, o.v[0],
                                 o.v[1]);
    wilson_combine<float, false>(s1.a, s1.b, sb_sum(tr, true), sb_sumi(r1, false), sb_sum(s0, false),
                                 sb_sumi(q1, true), K.f10, K.f11, K.b10, K.b11, mr1t, mx_r, true, mx_l, mm, o.v[2],
                                 o.v[3]);
}

// W rows of the thread, FP16 as stored: the first WR in registers (row k by
// a run-time index through a switch, so the array stays in registers), the
// others in LDS (row k at wl[(k - WR) * nth]).  The 16-row form keeps six in
// registers: with more, its row loops hold spill reloads, and a reload waits
// (vector-memory counters retire in order) for every load issued before it,
// the slot prefetch included.
template <int RW> constexpr int sr_wreg() { return RW < 8 ? RW : RW == 8 ? 8 : 6; }
template <int RW> constexpr size_t sr_lds_per_thread() {
    return SR_LDS_PER_THREAD + (size_t)(RW - sr_wreg<RW>()) * sizeof(f16x8);
}
template <int RW> struct SrW {
    static_assert(RW >= 2 && RW <= 16, "rows per thread");
    static constexpr int WR = sr_wreg<RW>();
    f16x8 h[WR];
    f16x8* wl;  // this thread's LDS rows
    unsigned nth;
    __device__ __forceinline__ f16x8 get(int k) const {
        if (WR < RW && k >= WR) return wl[(unsigned)(k - WR) * nth];
        f16x8 r = h[0];
        switch (k) {
#define AMGS_SR_GET(i) \
    case i:            \
        if constexpr (i < WR) r = h[i]; \
        break;
            AMGS_SR_GET(1) AMGS_SR_GET(2) AMGS_SR_GET(3) AMGS_SR_GET(4) AMGS_SR_GET(5) AMGS_SR_GET(6) AMGS_SR_GET(7)
#undef AMGS_SR_GET
        }
        return r;
    }
    __device__ __forceinline__ void set(int k, const f16x8 v) {
        if (WR < RW && k >= WR) {
            wl[(unsigned)(k - WR) * nth] = v;
            return;
        }
        switch (k) {
#define AMGS_SR_SET(i)                   \
    case i:                              \
        if constexpr (i < WR) h[i] = v; \
        break;
            AMGS_SR_SET(0) AMGS_SR_SET(1) AMGS_SR_SET(2) AMGS_SR_SET(3)
            AMGS_SR_SET(4) AMGS_SR_SET(5) AMGS_SR_SET(6) AMGS_SR_SET(7)
#undef AMGS_SR_SET
        }
    }
};

// A thread's place in the item and the item's shared state (plain values:
// a reference into the kernel arguments would put a copy of them in scratch).
struct SrItem {
    f16x4* Vh0;       // FP16 slot 0 of this rhs; slot i at Vh0 + i * sst
    size_t sst;
    int rho, l, TL, nruns;
    unsigned n0, nt;  // global site of the pair in row 0 of the run; row stride
    const c32* Ut;
    float mm, sg;
    f16x8* b16;          // [nruns][2][TL] FP16 boundary rows (first, last) of every run
    Pair4<float>* b32;   // [nruns][2][TL] FP32 boundary rows
    ItemLds& L;
    c64 (*park)[SBQ];
    c64* tot;            // [SBQ] workgroup totals of the parked quantities
    const SapState& st;  // unused state (the scalar helpers store nothing here)
    __device__ __forceinline__ unsigned n(int k) const { return n0 + (unsigned)k * nt; }
    __device__ __forceinline__ int bi(int run, int side) const { return (run * 2 + side) * TL + l; }
    __device__ __forceinline__ f16x8 slot_ld(int i, unsigned n) const { return sr_ld<f16x8>(Vh0 + i * sst, 8u * n); }
    __device__ __forceinline__ void slot_st(int i, unsigned n, const f16x8 h) const {
        sr_st<f16x8>(Vh0 + i * sst, 8u * n, h);
    }
    // the x-neighbour rows of row k lie in the block
    __device__ __forceinline__ bool xlo(int k) const { return rho > 0 || k > 0; }
    __device__ __forceinline__ bool xhi(int k, int RW) const { return rho + 1 < nruns || k + 1 < RW; }
};

// Workgroup totals of quantities 0..nq-1 and of q2 (parked per wave): wave 0
// sums them in fixed order into tot; the caller's thread 0 reads them after
// this wave's own LDS stores have completed.
__device__ __forceinline__ void sr_totals(const SrItem& I, int nq, int q2a, int q2b) {
    if (threadIdx.x < 64) {
        const int q = threadIdx.x;
        const bool want = q < nq || q == q2a || q == q2b;
        if (want) {
            const int nw = blockDim.x >> 6;
            c64 r = I.park[0][q];
            for (int w = 1; w < nw; w++) r += I.park[w][q];
            I.tot[q] = r;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

constexpr int SRQ_NN = SBQ - 1;  // |stored Vt_J|^2
constexpr int SRQ_WW = SBQ - 2;  // |W|^2

// Arnoldi step J of a cycle (src/fgmres.cpp:20-49) as one pass over the rows:
// Vt_J (J > 0), W = D_B Vt_J into W's registers, the inner products
// <W, Vt_i> (i <= J), <Vt_J, Vt_i> (i < J) and |W|^2; then thread 0
// finalises step J-1 and forms column J.  For J = 0, W holds the stored Vt_0
// rows on entry and the LDS boundary rows hold its first and last rows.
// The slots of row k+2 are loaded during row k (each row's slots once: for
// the new vector of that row and its inner products).
// Returns false when the cycle ends here (breakdown or the estimate met).
template <int RW, int J>
__device__ __forceinline__ bool sr_step(const SrItem& I, SrW<RW>& W, int M, double tol, double sigd, int& stop,
                                        double& bytes) {
    constexpr int JV = J > 0 ? J : 1;
    double nn = 0;
    float sj = 0.f;
    c32 eta[JV];
    // u = s_{J-1} W_row - sum_i eta_i Vt_i(row), stored as slot J unless
    // `to` names another slot
    auto form = [&](const f16x8 w, const f16x8 (&v)[JV], unsigned n, int to = J) {
        Pair4<float> u = sr_unpack(w);
#pragma unroll
        for (int e = 0; e < 4; e++) u.v[e] = sj * u.v[e];
#pragma unroll
        for (int i = 0; i < J; i++) {
            const Pair4<float> vi = sr_unpack(v[i]);
#pragma unroll
            for (int e = 0; e < 4; e++) cmsb(u.v[e], eta[i], vi.v[e]);
        }
        const f16x8 h = sr_pack(u, I.sg);
        I.slot_st(to, n, h);
        nn += (double)sr_norm2(h);
        return h;
    };
    auto load = [&](f16x8 (&v)[JV], int k) {
        if constexpr (J > 0) {
            const unsigned n = I.n(k < RW ? k : RW - 1);
#pragma unroll
            for (int i = 0; i < J; i++) v[i] = I.slot_ld(i, n);
        }
    };
    if constexpr (J > 0) {
        sj = float(I.L.s[J - 1]);
#pragma unroll
        for (int i = 0; i < J; i++) eta[i] = cvt<float>(I.L.eta[i]);
        // the first and last row of the run first: their neighbours read them
        f16x8 v0[JV], v1[JV];
        load(v0, 0);
        load(v1, RW - 1);
        I.b16[I.bi(I.rho, 0)] = form(W.h[0], v0, I.n(0));
        I.b16[I.bi(I.rho, 1)] = form(W.get(RW - 1), v1, I.n(RW - 1));
        __syncthreads();
        bytes += 8.0 * (J + 1);  // Vt_0..Vt_{J-1} read once, Vt_J written
    }
    f16x8 uc = J == 0 ? W.h[0] : I.b16[I.bi(I.rho, 0)];
    f16x8 up = I.rho > 0 ? I.b16[I.bi(I.rho - 1, 1)] : uc;
    c32 pw[J + 1], pv[JV];
#pragma unroll
    for (int i = 0; i <= J; i++) pw[i] = czero<float>();
#pragma unroll
    for (int i = 0; i < JV; i++) pv[i] = czero<float>();
    double ww = 0;
    const bool ml0 = I.l > 0, mr1t = I.l + 1 < I.TL;
    // row k with the slots of rows k (vc) and k+1 (vn) at hand; the slots of
    // row k+2 go to vp.  Three register sets take the three roles in turn (the
    // row loop runs three rows per trip), so no copy of a set waits for its
    // loads.
    SrXCarry X(I.Ut, I.n(0), I.nt, I.xlo(0));
    auto row = [&](int k, const f16x8 (&vc)[JV], const f16x8 (&vn)[JV], f16x8 (&vp)[JV]) {
        const unsigned n = I.n(k);
        const bool mx_l = I.xlo(k), mx_r = I.xhi(k, RW);
        // the row's links first: vector-memory counters retire in order, so
        // the stencil's wait for them then leaves the prefetch below in flight
        const sr_f4 ua = sr_ld<sr_f4>(I.Ut, 16u * n), ub = sr_ld<sr_f4>(I.Ut, 16u * n + 16u);
        // The same memory operations on every path (a clamped prefetch row; a
        // store of the value at hand into slot J+1, written later before any
        // read or never read, where no row is formed): the compiler's waits
        // after a join then need not drain the prefetch.
        load(vp, k + 2);
        f16x8 un;
        if (k + 2 < RW) {  // row k+1 lies inside the run: form it now
            un = J == 0 ? W.get(k + 1) : form(W.get(k + 1), vn, I.n(k + 1));
        } else {
            if (k + 2 == RW)  // the run's last row, formed before the pass
                un = J == 0 ? W.get(RW - 1) : I.b16[I.bi(I.rho, 1)];
            else  // the next run's first row
                un = I.rho + 1 < I.nruns ? I.b16[I.bi(I.rho + 1, 0)] : uc;
            if constexpr (J > 0) I.slot_st(J + 1, I.n(RW - 1), un);
        }
        f16x4 tl, tr;
        sr_tnb(uc, tl, tr);
        const SrLinks K(ua, ub, X.x0, X.x1);
        X.next(ua, ub);
        Pair4<float> o, c;
        sr_stencil16(K, I.mm, uc, un, up, tl, tr, mx_r, mx_l, ml0, mr1t, o, c);
        // W as stored in FP16 (the Hessenberg column belongs to these values)
        const f16x8 wh = sr_pack(o, 1.0f);
        W.set(k, wh);
        ww += (double)sr_norm2(wh);
        const SrC4 wc(wh);
        sr_dotc(pw[J], wc, uc);
        if constexpr (J > 0) {
            const SrC4 cc(uc);
#pragma unroll
            for (int i = 0; i < J; i++) {
                sr_dotc(pw[i], wc, vc[i]);
                sr_dotc(pv[i], cc, vc[i]);
            }
        }
        up = uc;
        uc = un;
    };
    f16x8 sa[JV], sb[JV], sc[JV];
    load(sa, 0);
    load(sb, 1);
#pragma unroll 1
    for (int k = 0; k < RW; k += 3) {
        row(k, sa, sb, sc);
        if (k + 1 < RW) row(k + 1, sb, sc, sa);
        if (k + 2 < RW) row(k + 2, sc, sa, sb);
    }
#pragma unroll
    for (int i = 0; i <= J; i++) sb_park(cvt<double>(pw[i]), i, I.park);
#pragma unroll
    for (int i = 0; i < J; i++) sb_park(cvt<double>(pv[i]), J + 1 + i, I.park);
    sb_park(c64{nn, 0.0}, SRQ_NN, I.park);
    sb_park(c64{ww, 0.0}, SRQ_WW, I.park);
    __syncthreads();
    sr_totals(I, 2 * J + 1, SRQ_NN, SRQ_WW);
    if (threadIdx.x == 0) {
        ItemLds& L = I.L;
        if (J > 0) {
            sap_finalize_step(I.st, 0, I.tot[SRQ_NN].re, M, J - 1, tol, false, L, sigd);
            if (L.broke) {
                L.jbreak = J - 1;
                L.maxIt = J;
                stop = 1;
            }
        }
        if (!stop) {
            for (int q = 0; q <= J; q++) L.dq[q] = I.tot[q];
            for (int q = 0; q < J; q++) L.dq[J + 1 + q] = I.tot[J + 1 + q];
            sap_hessenberg_column(L, M, J);
            for (int i = 0; i <= J; i++) L.Hraw[i * M + J] = L.Hcol[i];
            L.dr[1] = I.tot[SRQ_WW].re;
        }
    }
    __syncthreads();
    return !stop;
}

// |u|^2 of u = w - sum_{i<=j} H_ij V_i (w = s_j W, V_i = s_i Vt_i) from the
// quantities of step j: <V_i, w> = s_i s_j dq[i], G[i][k] = <V_k, V_i> (k < i),
// <V_i, V_i> = 1 and |W|^2 (dr[1]); thread 0, before step j is finalised.
//   |u|^2 = |w|^2 - 2 Re sum_i conj(H_i) <V_i, w> + sum_i |H_i|^2
//           + 2 Re sum_{k<i} conj(H_i) H_k conj(G[i][k])
__device__ __forceinline__ double sr_last_norm2(const ItemLds& L, int M, int j) {
    double u2 = (L.s[j] * L.s[j]) * L.dr[1];
    for (int i = 0; i <= j; i++) {
        const c64 h = L.Hcol[i];
        const c64 ai = (L.s[i] * L.s[j]) * L.dq[i];
        u2 += norm2(h) - 2.0 * (conj(h) * ai).re;
        for (int k = 0; k < i; k++) u2 += 2.0 * (conj(h) * L.Hcol[k] * conj(L.G[i * (M + 1) + k])).re;
    }
    return u2 > 0 ? u2 : 0.0;
}

// One colour of one SAP sweep for every item: x_B = b_B, then `restarts`
// cycles of GMRES(m) on D_B x_B = b_B; after the last cycle x_B is added to
// xacc (src/sap.cpp:117-125).  Grid (colour blocks, rhs); block TL sxel / RW
// threads; dynamic LDS sr_lds_per_thread<RW>() bytes per thread.  act (may be null):
// the algorithmic bytes per block site moved by each item, summed over items.
template <int RW>
__global__ void __launch_bounds__(SRT) k_sap_run(SapArgs<float> a, c32* __restrict__ xacc, int restarts, int* act) {
    extern __shared__ __align__(16) char sr_dyn[];
    __shared__ ItemLds L;
    __shared__ c64 park[SRT / 64][SBQ];
    __shared__ c64 tot[SBQ];
    __shared__ double sig0;  // scale of the stored Vt_0
    __shared__ int stop;     // the current cycle broke down
    const int rhs = blockIdx.y;
    if (a.skip && a.skip[rhs]) return;
    const Geo g = a.g;
    const int blk = a.blocks[blockIdx.x];
    const int X = blk / g.sbt;
    const int x0 = X * g.sxel, t0 = (blk - X * g.sbt) * g.stel;
    const int tid = threadIdx.x, TL = g.stel >> 1, rho = tid / TL, l = tid - rho * TL;
    const float mm = a.mmv ? a.mmv[rhs] : a.mm;
    __shared__ SapState st_unused;
    const SrItem I{a.Vh + (size_t)rhs * g.ns,
                   (size_t)a.nrhs * g.ns,
                   rho,
                   l,
                   TL,
                   g.sxel / RW,
                   (unsigned)((x0 + rho * RW) * g.nt + t0 + 2 * l),
                   (unsigned)g.nt,
                   a.U + (size_t)rhs * a.ucs,
                   mm,
                   a.sig,
                   reinterpret_cast<f16x8*>(sr_dyn),
                   reinterpret_cast<Pair4<float>*>(sr_dyn + blockDim.x * 2 * sizeof(f16x8)),
                   L,
                   park,
                   tot,
                   st_unused};
    const int M = a.m;
    const size_t off = (size_t)rhs * g.ns * 2;
    const c32* __restrict__ bg = a.rglob + off;
    c32* __restrict__ xl = a.xl + off;
    c32* __restrict__ xa0 = xacc + off;
    const double sigd = (double)a.sig;
    double bytes = 0;  // algorithmic bytes per block site (accounting)
    const bool ml0 = l > 0, mr1t = l + 1 < TL;
    SrW<RW> W;
    W.wl = reinterpret_cast<f16x8*>(sr_dyn + blockDim.x * SR_LDS_PER_THREAD) + tid;
    W.nth = blockDim.x;
    // update coefficients of the cycle's x_B update (L.eta after the
    // back-substitution), as FP32
    c32 cf[SR_MAXM];
    int mi = 0;
    // The residual pass reads per row x_B (or b in the first cycle, x0 = b)
    // and the slots of all M restart steps (those at mi <= i < M, from
    // earlier cycles or never written, are not used; the basis holds M + 1
    // slots, so i < M keeps every load inside it), loaded one row ahead.
    struct SrX {
        Pair4<float> x;
        f16x8 v[SR_MAXM];
    };
    auto xload = [&](SrX& r, int k, auto first) {
        const unsigned n = I.n(k < RW ? k : RW - 1);
        r.x = sr_ld_pair(first.value ? bg : xl, n);
        if constexpr (!decltype(first)::value) {
#pragma unroll
            for (int i = 0; i < SR_MAXM; i++)
                if (i < M) r.v[i] = I.slot_ld(i, n);
        }
    };
    // x_B (first cycle: b; else x_B + sum_{i<mi} eta_i Vt_i) of a loaded row
    auto xfin = [&](const SrX& r, auto first) {
        Pair4<float> x = r.x;
        if constexpr (!decltype(first)::value) {
#pragma unroll
            for (int i = 0; i < SR_MAXM; i++) {
                if (i < mi) {
                    const Pair4<float> vi = sr_unpack(r.v[i]);
#pragma unroll
                    for (int e = 0; e < 4; e++) cmad(x.v[e], cf[i], vi.v[e]);
                }
            }
        }
        return x;
    };
    auto xstore = [&](int k, const Pair4<float>& x) { sr_st_pair(xl, I.n(k), x); };
    // x_B stored, then Vt_0 = sg (b_B - D_B x_B) stored (slot 0, W rows, the
    // run's boundary rows in LDS) in one pass over the rows, the x_B boundary
    // rows exchanged through LDS first.  Returns the thread's |stored Vt_0|^2
    // (and |b|^2 in qb on the first cycle).
    auto resid_pass = [&](auto first, float sg, double& qb) {
        constexpr bool FIRST = decltype(first)::value;
        auto xst = [&](int k) {
            SrX r;
            xload(r, k, first);
            const Pair4<float> x = xfin(r, first);
            xstore(k, x);
            return x;
        };
        I.b32[I.bi(rho, 0)] = xst(0);
        I.b32[I.bi(rho, 1)] = xst(RW - 1);
        __syncthreads();
        Pair4<float> xc = I.b32[I.bi(rho, 0)];
        Pair4<float> xp = rho > 0 ? I.b32[I.bi(rho - 1, 1)] : xc;
        double qr = 0;
        SrXCarry X(I.Ut, I.n(0), I.nt, I.xlo(0));
#pragma unroll 1
        for (int k = 0; k < RW; k++) {
            const unsigned n = I.n(k);
            const bool mx_l = I.xlo(k), mx_r = I.xhi(k, RW);
            const sr_f4 ua = sr_ld<sr_f4>(I.Ut, 16u * n), ub = sr_ld<sr_f4>(I.Ut, 16u * n + 16u);
            // b of the row (x_B itself in the first cycle)
            const Pair4<float> bk = FIRST ? xc : sr_ld_pair(bg, n);
            Pair4<float> xn;
            if (k + 2 < RW) xn = xst(k + 1);
            else if (k + 2 == RW) xn = I.b32[I.bi(rho, 1)];
            else xn = rho + 1 < I.nruns ? I.b32[I.bi(rho + 1, 0)] : xc;
            const Spin2<float> tl{sr_shfl_up(xc.v[2]), sr_shfl_up(xc.v[3])};
            const Spin2<float> tr{sr_shfl_down(xc.v[0]), sr_shfl_down(xc.v[1])};
            const SrLinks K(ua, ub, X.x0, X.x1);
            X.next(ua, ub);
            Pair4<float> o;
            sr_stencil32(K, mm, xc, xn, xp, tl, tr, mx_r, mx_l, ml0, mr1t, o);
            Pair4<float> r;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                r.v[e] = bk.v[e] - o.v[e];
                if constexpr (FIRST) qb += double(norm2(bk.v[e]));
            }
            const f16x8 h = sr_pack(r, sg);
            I.slot_st(0, n, h);
            qr += (double)sr_norm2(h);
            W.set(k, h);
            if (k == 0) I.b16[I.bi(rho, 0)] = h;
            if (k == RW - 1) I.b16[I.bi(rho, 1)] = h;
            xp = xc;
            xc = xn;
        }
        return qr;
    };
    using First = std::integral_constant<bool, true>;
    using Restart = std::integral_constant<bool, false>;

    // ---- x_B = b_B, Vt_0 = b_B - D_B b_B (src/sap.cpp:61, src/fgmres.cpp:8-17) ----
    {
        const float sgi = (float)a.rsig[rhs];
        double qb = 0;
        const double qr = resid_pass(First{}, sgi, qb);
        sb_park(c64{qr, 0.0}, 0, park);
        sb_park(c64{qb, 0.0}, 1, park);
        __syncthreads();
        sr_totals(I, 2, -1, -1);
        if (tid == 0) {
            const double est = sqrt(tot[0].re), nb = sqrt(tot[1].re);
            L.maxIt = M;
            L.jbreak = -1;
            L.done = nb == 0.0;
            L.normb = nb;
            L.s[0] = 1.0 / est;
            L.gpre[0] = c64{est / sgi, 0.0};
            sig0 = sgi;
        }
        __syncthreads();
        bytes += 16.0 + 16.0 + 8.0;  // b read, x_B written, Vt_0 written
    }

    for (int kr = 0; kr < restarts; kr++) {
        const bool last = kr + 1 == restarts;
        if (!L.done) {
            if (tid == 0) stop = 0;
            // ---- the Arnoldi steps of the cycle (compiled per step index) ----
            bool go = true;
#define AMGS_SR_STEP(j) \
    if (go && j < M) go = sr_step<RW, j>(I, W, M, a.tol, sigd, stop, bytes);
            AMGS_SR_STEP(0) AMGS_SR_STEP(1) AMGS_SR_STEP(2) AMGS_SR_STEP(3) AMGS_SR_STEP(4)
            static_assert(SR_MAXM == 5, "steps compiled up to SR_MAXM - 1");
#undef AMGS_SR_STEP
            // ---- end of the cycle: H_{m,m-1}, back-substitution ----
            if (tid == 0) {
                // src/fgmres.cpp:52-59 with the stale maxIt of earlier cycles (a
                // cycle that ran to m without breaking keeps it)
                int maxIt = L.maxIt;
                if (go) {
                    const double hn2 = sr_last_norm2(L, M, M - 1) * sigd * sigd;
                    sap_finalize_step(st_unused, 0, hn2, M, M - 1, a.tol, false, L, sigd);
                    if (L.broke) {
                        L.jbreak = M - 1;
                        L.maxIt = M;
                        maxIt = M;
                    }
                }
                sap_backsolve(L, M, maxIt);
                const double s0 = sap_next_sig0(L, maxIt, g.sspb);
                if (s0 > 0) sig0 = s0;
                L.active = maxIt;  // the coefficients of this update
            }
            __syncthreads();
            mi = L.active;
#pragma unroll
            for (int i = 0; i < SR_MAXM; i++) cf[i] = cvt<float>(L.eta[i]);
            if (last) {
                // x += x_B + sum_i eta_i Vt_i, two rows' loads in flight
#pragma unroll 1
                for (int k = 0; k < RW; k += 2) {
                    SrX A, B;
                    xload(A, k, Restart{});
                    xload(B, k + 1, Restart{});
                    const Pair4<float> xa = xfin(A, Restart{}), xb = xfin(B, Restart{});
                    const unsigned n = I.n(k);
                    Pair4<float> ya = sr_ld_pair(xa0, n), yb = sr_ld_pair(xa0, n + I.nt);
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        ya.v[e] += xa.v[e];
                        yb.v[e] += xb.v[e];
                    }
                    sr_st_pair(xa0, n, ya);
                    sr_st_pair(xa0, n + I.nt, yb);
                }
                bytes += 8.0 * mi + 16.0 * 3;
            } else {
                // ---- x_B += sum_i eta_i Vt_i and the next cycle's Vt_0 =
                //      b_B - D_B x_B in one pass (src/fgmres.cpp:61-65, 8-17) ----
                double unused = 0;
                const double qr = resid_pass(Restart{}, float(sig0), unused);
                sb_park(c64{qr, 0.0}, 0, park);
                __syncthreads();
                sr_totals(I, 1, -1, -1);
                if (tid == 0) {
                    const double est = sqrt(tot[0].re), err = est / sig0;
                    if (err < a.tol * L.normb || L.normb == 0.0) {
                        L.done = 1;
                    } else {
                        L.s[0] = 1.0 / est;
                        L.gpre[0] = c64{err, 0.0};
                        L.jbreak = -1;
                    }
                }
                __syncthreads();
                // x_B read and written, Vt_0..Vt_{mi-1} read, b read, Vt_0 written
                bytes += 16.0 * 3 + 8.0 * mi + 8.0;
            }
        } else if (last) {
            // converged in an earlier cycle: x_B goes into x unchanged
#pragma unroll 1
            for (int k = 0; k < RW; k++) {
                const unsigned n = I.n(k);
                Pair4<float> y = sr_ld_pair(xa0, n);
                const Pair4<float> xb = sr_ld_pair(xl, n);
#pragma unroll
                for (int e = 0; e < 4; e++) y.v[e] += xb.v[e];
                sr_st_pair(xa0, n, y);
            }
            bytes += 16.0 * 3;
        }
    }
    if (act && tid == 0) atomicAdd(act, (int)bytes);
}

}  // namespace amgs
