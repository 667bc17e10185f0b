// Shared types and device helpers for libamgs (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <type_traits>

#define AMGS_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------------------
// complex arithmetic on interleaved (re, im) pairs
// ---------------------------------------------------------------------------
template <typename R> struct alignas(2 * sizeof(R)) cplx {
    R re, im;
};
using c64 = cplx<double>;
using c32 = cplx<float>;

template <typename R> AMGS_HD cplx<R> mk(R re, R im) { return cplx<R>{re, im}; }
template <typename R> AMGS_HD cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> AMGS_HD cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> AMGS_HD cplx<R> operator-(cplx<R> a) { return {-a.re, -a.im}; }
template <typename R> AMGS_HD cplx<R> operator*(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> AMGS_HD cplx<R> operator*(R s, cplx<R> a) { return {s * a.re, s * a.im}; }
template <typename R> AMGS_HD cplx<R>& operator+=(cplx<R>& a, cplx<R> b) { a.re += b.re; a.im += b.im; return a; }
template <typename R> AMGS_HD cplx<R>& operator-=(cplx<R>& a, cplx<R> b) { a.re -= b.re; a.im -= b.im; return a; }
template <typename R> AMGS_HD cplx<R> conj(cplx<R> a) { return {a.re, -a.im}; }
// a * conj(b): the summand of the reference dot product (include/lin_alg_op.h:32-40)
template <typename R> AMGS_HD cplx<R> mulc(cplx<R> a, cplx<R> b) {
    return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
// acc += a b, acc += a conj(b) and acc -= a b.  In FP32 as four fused
// multiply-adds (contraction does not reach across the separate statements of
// an accumulation); FP64 keeps the plain expressions and their rounding.
template <typename R> AMGS_HD void cmad(cplx<R>& acc, cplx<R> a, cplx<R> b) {
    if constexpr (std::is_same<R, float>::value) {
        acc.re = fmaf(a.re, b.re, fmaf(-a.im, b.im, acc.re));
        acc.im = fmaf(a.re, b.im, fmaf(a.im, b.re, acc.im));
    } else {
        acc += a * b;
    }
}
template <typename R> AMGS_HD void cmadc(cplx<R>& acc, cplx<R> a, cplx<R> b) {
    if constexpr (std::is_same<R, float>::value) {
        acc.re = fmaf(a.re, b.re, fmaf(a.im, b.im, acc.re));
        acc.im = fmaf(a.im, b.re, fmaf(-a.re, b.im, acc.im));
    } else {
        acc += mulc(a, b);
    }
}
template <typename R> AMGS_HD void cmsb(cplx<R>& acc, cplx<R> a, cplx<R> b) {
    if constexpr (std::is_same<R, float>::value) {
        acc.re = fmaf(-a.re, b.re, fmaf(a.im, b.im, acc.re));
        acc.im = fmaf(-a.re, b.im, fmaf(-a.im, b.re, acc.im));
    } else {
        acc -= a * b;
    }
}
// i * a and -i * a
template <typename R> AMGS_HD cplx<R> mul_i(cplx<R> a) { return {-a.im, a.re}; }
template <typename R> AMGS_HD cplx<R> mul_mi(cplx<R> a) { return {a.im, -a.re}; }
template <typename R> AMGS_HD R norm2(cplx<R> a) { return a.re * a.re + a.im * a.im; }
template <typename R> AMGS_HD cplx<R> czero() { return {R(0), R(0)}; }

template <typename To, typename From> AMGS_HD cplx<To> cvt(cplx<From> a) { return {To(a.re), To(a.im)}; }

// complex division as a / b with Smith-free scaling by |b|^2 (FP64 scalars only)
AMGS_HD c64 cdiv(c64 a, c64 b) {
    double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// ---------------------------------------------------------------------------
// lattice geometry of one level (host + device POD)
// ---------------------------------------------------------------------------
struct Geo {
    int nx, nt, ns, dof, colors;
    // aggregation towards level l+1 (valid for l < levels-1)
    int bx, bt, ntest, nagg, xel, tel, spb;
    // Schwarz (SAP) blocks
    int sbx, sbt, sxel, stel, sspb, nsb, ncol;
};

AMGS_HD int nb_right(int n, int mu, int nx, int nt) {
    int x = n / nt, t = n - x * nt;
    return mu == 0 ? x * nt + (t + 1 == nt ? 0 : t + 1) : (x + 1 == nx ? 0 : x + 1) * nt + t;
}
AMGS_HD int nb_left(int n, int mu, int nx, int nt) {
    int x = n / nt, t = n - x * nt;
    return mu == 0 ? x * nt + (t == 0 ? nt - 1 : t - 1) : (x == 0 ? nx - 1 : x - 1) * nt + t;
}
// antiperiodic time boundary (include/boundary.h:39-40)
AMGS_HD bool wraps_right(int n, int mu, int nt) { return mu == 0 && (n % nt) == nt - 1; }
AMGS_HD bool wraps_left(int n, int mu, int nt) { return mu == 0 && (n % nt) == 0; }

// Schwarz block of a site as the reference computes it (getMandBlock,
// include/level.h:48-59: x = n / Nx, t = n % Nt; equal to the true
// coordinates on square lattices).
AMGS_HD int sap_block_of(const Geo& g, int n) {
    int x = n / g.nx, t = n % g.nt;
    return (x / g.sxel) * g.sbt + t / g.stel;
}
AMGS_HD int sap_local_of(const Geo& g, int n) {
    int x = n / g.nx, t = n % g.nt;
    return (x % g.sxel) * g.stel + t % g.stel;
}
// global site of local index m in Schwarz block blk (src/sap.cpp:3-24)
AMGS_HD int sap_site(const Geo& g, int blk, int m) {
    int X = blk / g.sbt, T = blk - X * g.sbt;
    int mx = m / g.stel, mt = m - mx * g.stel;
    return (X * g.sxel + mx) * g.nt + T * g.stel + mt;
}
// aggregation block of a site (Level::getLatticeBlock, include/level.h:267-278)
AMGS_HD int agg_block_of(const Geo& g, int n) {
    int x = n / g.nt, t = n % g.nt;
    return (x / g.xel) * g.bt + t / g.tel;
}
// site of entry m (x-major within the block) of aggregation block X
AMGS_HD int agg_site(const Geo& g, int X, int m) {
    int BX = X / g.bt, BT = X - BX * g.bt;
    int mx = m / g.tel, mt = m - mx * g.tel;
    return (BX * g.xel + mx) * g.nt + BT * g.tel + mt;
}

// ---------------------------------------------------------------------------
// deterministic reductions (fixed order: lane butterfly, then waves in order)
// ---------------------------------------------------------------------------
// Lane exchanges for a butterfly reduction: DPP row operations for partners
// 1, 2 (quad permutes), 4 (half-row mirror) and 8 (row mirror); after each
// step all lanes of the group hold the same partial sum, so a mirror pairs
// groups exactly like an xor.  Partners 16 and 32 go through ds_bpermute.
template <int CTRL> __device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ __forceinline__ double dpp_d(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);  // row_half_mirror
    v += dpp_d<0x140>(v);  // row_mirror
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_f<0xB1>(v);
    v += dpp_f<0x4E>(v);
    v += dpp_f<0x141>(v);
    v += dpp_f<0x140>(v);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}
template <typename R> __device__ __forceinline__ cplx<R> wave_sum(cplx<R> v) {
    return {wave_sum(v.re), wave_sum(v.im)};
}

// lane l's value of v (l wave-uniform), for every lane
__device__ __forceinline__ double readlane_d(double v, int l) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Block-wide sum of NQ complex values per thread; every thread gets the
// totals in out[].  scratch: >= NQ * nwaves complex entries of LDS.
template <int NQ, typename R>
__device__ __forceinline__ void block_sum(cplx<R>* vals, cplx<R>* scratch, int nq = NQ) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++)
        if (q < nq) vals[q] = wave_sum(vals[q]);
    if (lane == 0)
        for (int q = 0; q < nq; q++) scratch[q * nw + wv] = vals[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        if (q < nq) {
            cplx<R> s = scratch[q * nw];
            for (int w = 1; w < nw; w++) s += scratch[q * nw + w];
            vals[q] = s;
        }
    }
    __syncthreads();
}

#define HIP_CHECK(x)                                                              \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) throw HipError(e_, #x, __FILE__, __LINE__);         \
    } while (0)
