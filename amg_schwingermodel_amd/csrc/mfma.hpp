// The coarse operator over a block of right-hand sides as matrix-core
// block-GEMMs (gfx950 MFMA): Level::D_operator at l >= 1
// (src/level.cpp:161-180) applied to nrhs vectors at once,
//   out_r[x] = (m0+2) v_r[x] - sum_k f_k(x) L_k[x] v_r[nbr_k(x)],
// k = A, B_0, B_1, C_0, C_1 (f = -1 on the antiperiodic time hops that wrap,
// include/boundary.h:39-40).  For one coarse site the five D x D link blocks
// side by side form a D x 5D complex matrix and the five neighbour vectors of
// all right-hand sides a 5D x nrhs one: OUT^T (nrhs x D) = V^T (nrhs x 5D) x
// L^T (5D x D), a complex GEMM done as four real MFMAs per K step,
//   re += Vr Lr - Vi Li,   im += Vr Li + Vi Lr,
// with 16x16x4 tiles (M = 16 right-hand sides, N = 16 rows, K = 4 columns).
//
// Operand layouts of v_mfma_{f32,f64}_16x16x4 (wave64, one block), checked on
// the device: A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k; the result
// D[i][j] sits in lane j + 16 (i / 4), register i % 4 (f32) and in lane
// j + 16 (i % 4), register i / 4 (f64).
//
// Launch shape (the tile autotuner's choice): a workgroup of 64 * WPG threads
// handles SPW coarse sites; each wavefront owns one (rhs tile, row tile) pair
// of one site and walks the whole K range.
#pragma once
#include "common.hpp"

namespace amgs {

template <typename R> struct MfmaT;
template <> struct MfmaT<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // (row i, column j) of the 16x16 result held by register v of lane l
    static __device__ __forceinline__ int row(int l, int v) { return 4 * (l >> 4) + v; }
};
template <> struct MfmaT<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int v) { return 4 * v + (l >> 4); }
};

struct CdopShape {
    int wpg;  // wavefronts per workgroup
};

// links: [5][ns][D][D] column-major per site (entry (c, r) at c D + r);
// in, out: [nrhs][ns][D].  Grid: (ceil(ns * tiles / wpg)), tiles = rhs tiles x row tiles per site.
template <typename R>
__global__ void __launch_bounds__(512) k_cdop_mfma(Geo g, R mm, const cplx<R>* __restrict__ links,
                                                   const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out,
                                                   int nrhs) {
    using T = MfmaT<R>;
    using v4 = typename T::v4;
    const int D = g.dof, ns = g.ns, lane = threadIdx.x & 63;
    const int mt = (nrhs + 15) >> 4, nt = (D + 15) >> 4, tiles = mt * nt;
    const int task = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (task >= ns * tiles) return;
    const int x = task / tiles, t = task - x * tiles, rt = t / nt, ct = t - rt * nt;
    const int li = lane & 15, lk = lane >> 4;
    const int rhs = 16 * rt + li, row = 16 * ct + li;  // A row (rhs) and B column (row) of this lane
    const bool rhs_ok = rhs < nrhs, row_ok = row < D;
    const size_t vstride = (size_t)ns * D, msz = (size_t)D * D, ksz = (size_t)ns * msz;
    int nbr[5];
    R f[5];
    nbr[0] = x;
    nbr[1] = nb_right(x, 0, g.nx, g.nt);
    nbr[2] = nb_right(x, 1, g.nx, g.nt);
    nbr[3] = nb_left(x, 0, g.nx, g.nt);
    nbr[4] = nb_left(x, 1, g.nx, g.nt);
    f[0] = R(1);
    f[1] = wraps_right(x, 0, g.nt) ? R(-1) : R(1);
    f[2] = R(1);
    f[3] = wraps_left(x, 0, g.nt) ? R(-1) : R(1);
    f[4] = R(1);
    const cplx<R>* vr = in + (size_t)(rhs_ok ? rhs : 0) * vstride;
    v4 are = {0, 0, 0, 0}, aim = {0, 0, 0, 0};
    const int K = 5 * D;
    for (int k0 = 0; k0 < K; k0 += 4) {
        const int kk = k0 + lk;
        cplx<R> a = czero<R>(), b = czero<R>();
        if (kk < K) {
            const int blk = kk / D, c = kk - blk * D;
            if (rhs_ok) a = f[blk] * vr[(size_t)nbr[blk] * D + c];
            if (row_ok) b = links[blk * ksz + x * msz + (size_t)c * D + row];
        }
        are = T::mma(a.re, b.re, are);
        are = T::mma(-a.im, b.im, are);
        aim = T::mma(a.re, b.im, aim);
        aim = T::mma(a.im, b.re, aim);
    }
    // lane holds result column j = li (a row of D), rows i = rhs offsets
    if (!row_ok) return;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int r = 16 * rt + T::row(lane, v);
        if (r < nrhs) {
            const size_t o = (size_t)r * vstride + (size_t)x * D + row;
            const cplx<R> vin = in[o];
            out[o] = cplx<R>{mm * vin.re - are[v], mm * vin.im - aim[v]};
        }
    }
}

}  // namespace amgs
