// Outer FGMRES on the fine lattice (FP64, src/fgmres.cpp:5-115 with the
// FGMRES_AMG operator and preconditioner, include/amg.h:99-145).  An item is
// one right-hand side; heavy kernels cover the lattice, scalar kernels run one
// workgroup per right-hand side.  Basis vectors are normalised explicitly.
#pragma once
#include "fine.hpp"

namespace amgs {

// r = b - D x with partials |r|^2, |b|^2 (src/fgmres.cpp:10-11, 60-62)
__global__ void __launch_bounds__(FT) k_outer_resid(Geo g, double mm, const c64* __restrict__ U,
                                                    const c64* __restrict__ b, const c64* __restrict__ x,
                                                    c64* __restrict__ r, OuterState st) {
    __shared__ c64 red[8];
    const int rhs = blockIdx.y;
    if (st.done[rhs]) return;
    U += rhs * st.ucs;
    mm = cfg_mm(mm, st.mmv, rhs);
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int n = tile * FT + threadIdx.x;
    const size_t off = (size_t)rhs * g.ns * 2;
    c64 q[2] = {{0, 0}, {0, 0}};
    if (n < g.ns) {
        c64 o0, o1;
        wilson_site<double, false>(U, x + off, n, g.nx, g.nt, mm, AllSites{}, o0, o1);
        const c64 b0 = b[off + 2 * n], b1 = b[off + 2 * n + 1];
        const c64 r0 = b0 - o0, r1 = b1 - o1;
        r[off + 2 * n] = r0;
        r[off + 2 * n + 1] = r1;
        q[0].re = norm2(r0) + norm2(r1);
        q[1].re = norm2(b0) + norm2(b1);
    }
    block_sum<2>(q, red);
    if (threadIdx.x == 0) {
        st.qr[((size_t)rhs * st.ntile + tile) * 2] = q[0].re;
        st.qr[((size_t)rhs * st.ntile + tile) * 2 + 1] = q[1].re;
    }
}

// Cycle start k: error estimate, convergence test of the previous cycle,
// V_0 scale (src/fgmres.cpp:61-67, 16-18).  One workgroup per rhs.
__global__ void __launch_bounds__(FT) k_outer_start(OuterState st, int k, double tol, int m) {
    __shared__ c64 red[8];
    const int rhs = blockIdx.x;
    if (st.done[rhs]) {
        if (threadIdx.x == 0) st.active[rhs] = 0;
        return;
    }
    c64 q[2] = {{0, 0}, {0, 0}};
    for (int t = threadIdx.x; t < st.ntile; t += FT) {
        q[0].re += st.qr[((size_t)rhs * st.ntile + t) * 2];
        q[1].re += st.qr[((size_t)rhs * st.ntile + t) * 2 + 1];
    }
    block_sum<2>(q, red);
    if (threadIdx.x != 0) return;
    const double err = sqrt(q[0].re);
    if (k == 0) st.normb[rhs] = sqrt(q[1].re);
    const double nb = st.normb[rhs];
    if (nb == 0.0) {  // zero right-hand side: x = 0 is exact
        st.done[rhs] = 1;
        st.active[rhs] = 0;
        st.iters[rhs] = 0;
        return;
    }
    if (k > 0 && err < tol * nb) {
        st.done[rhs] = 1;
        st.active[rhs] = 0;
        st.iters[rhs] = (k - 1) * m + st.maxIt[rhs];
        return;
    }
    if (k == 0) st.maxIt[rhs] = m;
    st.s[(size_t)rhs * (st.M + 1)] = 1.0 / err;
    st.g[(size_t)rhs * (st.M + 1)] = c64{err, 0.0};
    st.jbreak[rhs] = -1;
    st.active[rhs] = 1;
}

// V_j = s_j * src (V_0 from r, V_{j+1} from w)
// dst = s_j src; with dst32 also its FP32 copy (the mixed-precision
// preconditioner's input, so that no separate conversion pass reads dst)
__global__ void __launch_bounds__(FT) k_outer_scale(int ns, const c64* __restrict__ src, c64* __restrict__ dst,
                                                    OuterState st, int j, c32* __restrict__ dst32) {
    const int rhs = blockIdx.y;
    if (!st.active[rhs]) return;
    const int n = blockIdx.x * FT + threadIdx.x;
    if (n >= ns) return;
    const double s = st.s[(size_t)rhs * (st.M + 1) + j];
    const size_t off = (size_t)rhs * ns * 2;
    const c64 a = s * src[off + 2 * n], b = s * src[off + 2 * n + 1];
    dst[off + 2 * n] = a;
    dst[off + 2 * n + 1] = b;
    if (dst32) {
        dst32[off + 2 * n] = cvt<float>(a);
        dst32[off + 2 * n + 1] = cvt<float>(b);
    }
}

// w = D Z_j; partials <w, V_i> (i <= j), <V_j, V_k> (k < j)
// Z entries read in FP64: stored FP64, or FP32 (the mixed modes' preconditioner
// outputs) converted exactly
template <typename ZT> struct ZIn {
    const cplx<ZT>* p;
    __device__ __forceinline__ void operator()(int n, c64& a, c64& b) const {
        const unsigned u = 2u * (unsigned)n;
        a = cvt<double>(p[u]);
        b = cvt<double>(p[u + 1u]);
    }
};

template <typename ZT>
__global__ void __launch_bounds__(FT) k_outer_a(Geo g, double mm, const c64* __restrict__ U,
                                                const cplx<ZT>* __restrict__ Zj, const c64* __restrict__ V,
                                                c64* __restrict__ w, OuterState st, int j, size_t vstride) {
    constexpr int NQ = 2 * OUT_MAXM + 1;
    __shared__ c64 red[(FT / 64) * NQ];
    const int rhs = blockIdx.y;
    if (!st.active[rhs]) return;
    U += rhs * st.ucs;
    mm = cfg_mm(mm, st.mmv, rhs);
    // OUT_SPT consecutive tiles of FT sites per workgroup: the 2j+1 block
    // reductions are shared by OUT_SPT sites per thread
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const size_t off = (size_t)rhs * g.ns * 2;
    const int nq = 2 * j + 1;
    c64 o0[OUT_SPT], o1[OUT_SPT], a0[OUT_SPT], a1[OUT_SPT];
    bool live[OUT_SPT];
    const c64* vj = V + (size_t)j * vstride + off;
#pragma unroll
    for (int s = 0; s < OUT_SPT; s++) {
        const int n = (tile * OUT_SPT + s) * FT + threadIdx.x;
        live[s] = n < g.ns;
        o0[s] = o1[s] = a0[s] = a1[s] = c64{0, 0};
        if (live[s]) {
            wilson_site_in<double, false>(U, ZIn<ZT>{Zj + off}, n, g.nx, g.nt, mm, AllSites{}, o0[s], o1[s]);
            w[off + 2 * n] = o0[s];
            w[off + 2 * n + 1] = o1[s];
            a0[s] = vj[2 * n];
            a1[s] = vj[2 * n + 1];
        }
    }
    for (int i = 0; i < j; i++) {
        c64 pw = {0, 0}, pv = {0, 0};
        const c64* vi = V + (size_t)i * vstride + off;
#pragma unroll
        for (int s = 0; s < OUT_SPT; s++) {
            if (live[s]) {
                const int n = (tile * OUT_SPT + s) * FT + threadIdx.x;
                const c64 b0 = vi[2 * n], b1 = vi[2 * n + 1];
                pw += mulc(o0[s], b0) + mulc(o1[s], b1);
                pv += mulc(a0[s], b0) + mulc(a1[s], b1);
            }
        }
        park(pw, i, red, NQ);
        park(pv, j + 1 + i, red, NQ);
    }
    {  // i = j: V_j is at hand in a0, a1
        c64 pw = {0, 0};
#pragma unroll
        for (int s = 0; s < OUT_SPT; s++)
            if (live[s]) pw += mulc(o0[s], a0[s]) + mulc(o1[s], a1[s]);
        park(pw, j, red, NQ);
    }
    __syncthreads();
    if (threadIdx.x < nq) {
        c64 r = red[threadIdx.x];
        for (int wv = 1; wv < FT / 64; wv++) r += red[wv * NQ + threadIdx.x];
        st.qa[((size_t)rhs * (2 * st.M + 1) + threadIdx.x) * st.ntile_a + tile] = r;
    }
}

// Hessenberg column j by one-pass MGS: H_ij = <w,V_i> - sum_{k<i} H_kj <V_k,V_i>.
// Stores the unrotated column in R[.][j] temporarily (rotated in k_outer_fin).
static_assert(OUT_MAXM < 64, "one lane per Hessenberg row in k_outer_hcol");
__global__ void __launch_bounds__(FT) k_outer_hcol(OuterState st, int j) {
    __shared__ c64 d[2 * OUT_MAXM + 1];
    __shared__ c64 Gs[OUT_MAXM][OUT_MAXM];  // Gram rows 1..j-1 of the cycle
    const int rhs = blockIdx.x;
    if (!st.active[rhs]) return;
    const int nq = 2 * j + 1, M = st.M;
    // the tile partials of quantity q lie contiguous: one wavefront per
    // quantity, lane sums over tiles in order, then the wave butterfly
    const c64* base = st.qa + (size_t)rhs * (2 * M + 1) * st.ntile_a;
    const int lane = threadIdx.x & 63;
    for (int q = threadIdx.x >> 6; q < nq; q += FT / 64) {
        c64 v = {0, 0};
        for (int t = lane; t < st.ntile_a; t += 64) v += base[(size_t)q * st.ntile_a + t];
        v = wave_sum(v);
        if (lane == 0) d[q] = v;
    }
    c64* G = st.G + (size_t)rhs * (M + 1) * (M + 1);
    c64* Rm = st.R + (size_t)rhs * M * M;
    for (int e = threadIdx.x; e < j * j; e += FT) {
        const int i = e / j, k = e - i * j;
        if (k < i) Gs[i][k] = G[i * (M + 1) + k];
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    // lane i forms row i; H_kj is final once the lanes below k are (j < 64)
    const int l = threadIdx.x;
    c64 h = {0.0, 0.0};
    if (l <= j) h = d[l];
    for (int k = 0; k <= j; k++) {
        const c64 hk = {readlane_d(h.re, k), readlane_d(h.im, k)};
        if (l > k && l <= j) h -= hk * conj(l == j ? d[j + 1 + k] : Gs[l][k]);
    }
    if (l <= j) Rm[l * M + j] = h;
    if (l < j) G[j * (M + 1) + l] = d[j + 1 + l];
}

// w -= sum_i H_ij V_i; partial |w|^2
__global__ void __launch_bounds__(FT) k_outer_b(int ns, const c64* __restrict__ V, c64* __restrict__ w, OuterState st,
                                                int j, size_t vstride) {
    __shared__ c64 red[4];
    __shared__ c64 H[OUT_MAXM + 1];
    const int rhs = blockIdx.y;
    if (!st.active[rhs]) return;
    const int M = st.M;
    if (threadIdx.x <= j) H[threadIdx.x] = st.R[(size_t)rhs * M * M + threadIdx.x * M + j];
    __syncthreads();
    const int tile = blockIdx.x, n = tile * FT + threadIdx.x;
    const size_t off = (size_t)rhs * ns * 2;
    double nn = 0;
    if (n < ns) {
        c64 w0 = w[off + 2 * n], w1 = w[off + 2 * n + 1];
        for (int i = 0; i <= j; i++) {
            const c64* vi = V + (size_t)i * vstride + off;
            w0 -= H[i] * vi[2 * n];
            w1 -= H[i] * vi[2 * n + 1];
        }
        w[off + 2 * n] = w0;
        w[off + 2 * n + 1] = w1;
        nn = norm2(w0) + norm2(w1);
    }
    c64 q[1] = {{nn, 0.0}};
    block_sum<1>(q, red);
    if (threadIdx.x == 0) st.qb[(size_t)rhs * st.ntile + tile] = q[0].re;
}

// H_{j+1,j} = |w|, Givens rotation, residual estimate, break test (src/fgmres.cpp:35-49)
__global__ void __launch_bounds__(FT) k_outer_fin(OuterState st, int j, double tol) {
    __shared__ c64 red[4];
    const int rhs = blockIdx.x;
    if (!st.active[rhs]) return;
    const int M = st.M;
    c64 q[1] = {{0, 0}};
    for (int t = threadIdx.x; t < st.ntile; t += FT) q[0].re += st.qb[(size_t)rhs * st.ntile + t];
    block_sum<1>(q, red);
    if (threadIdx.x != 0) return;
    const double hn = sqrt(q[0].re);
    c64* Rm = st.R + (size_t)rhs * M * M;
    c64* cn = st.cn + (size_t)rhs * M;
    c64* sn = st.sn + (size_t)rhs * M;
    c64* g = st.g + (size_t)rhs * (M + 1);
    __shared__ c64 col[OUT_MAXM + 1];
    for (int i = 0; i <= j; i++) col[i] = Rm[i * M + j];
    col[j + 1] = c64{hn, 0.0};
    for (int i = 0; i < j; i++) {
        c64 t = conj(cn[i]) * col[i] + conj(sn[i]) * col[i + 1];
        col[i + 1] = -(sn[i] * col[i]) + cn[i] * col[i + 1];
        col[i] = t;
    }
    const double den = sqrt(norm2(col[j]) + norm2(col[j + 1]));
    const c64 snj = {col[j + 1].re / den, col[j + 1].im / den};
    const c64 cnj = {col[j].re / den, col[j].im / den};
    col[j] = conj(cnj) * col[j] + conj(snj) * col[j + 1];
    for (int i = 0; i <= j; i++) Rm[i * M + j] = col[i];
    cn[j] = cnj;
    sn[j] = snj;
    const c64 gj = g[j];
    g[j + 1] = -(snj * gj);
    g[j] = conj(cnj) * gj;
    st.s[(size_t)rhs * (M + 1) + j + 1] = hn > 0 ? 1.0 / hn : 0.0;
    const double res = sqrt(norm2(g[j + 1]));
    if (st.nhist[rhs] < st.hcap) st.hist[(size_t)rhs * st.hcap + st.nhist[rhs]] = res;
    st.nhist[rhs]++;
    if (res < tol * st.normb[rhs]) {
        st.maxIt[rhs] = j + 1;
        st.jbreak[rhs] = j;
        st.active[rhs] = 0;
    }
}

// back substitution with maxIt (src/fgmres.cpp:52-53, 107-115); re-activates
// every rhs that is not done for the update/residual kernels
__global__ void k_outer_backsolve(OuterState st, int nrhs) {
    const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
    if (rhs >= nrhs || st.done[rhs]) return;
    const int M = st.M, mi = st.maxIt[rhs];
    const c64* Rm = st.R + (size_t)rhs * M * M;
    const c64* g = st.g + (size_t)rhs * (M + 1);
    c64* eta = st.eta + (size_t)rhs * M;
    for (int i = mi - 1; i >= 0; i--) {
        c64 e = g[i];
        for (int q = i + 1; q < mi; q++) e -= Rm[i * M + q] * eta[q];
        eta[i] = cdiv(e, Rm[i * M + i]);
    }
    st.active[rhs] = 1;
}

// x += sum_{j < maxIt} eta_j Z_j (src/fgmres.cpp:55-59)
template <typename ZT>
__global__ void __launch_bounds__(FT) k_outer_update(int ns, const cplx<ZT>* __restrict__ Z, c64* __restrict__ x,
                                                     OuterState st, size_t zstride) {
    const int rhs = blockIdx.y;
    if (st.done[rhs]) return;
    const int n = blockIdx.x * FT + threadIdx.x;
    if (n >= ns) return;
    const int M = st.M, mi = st.maxIt[rhs];
    const size_t off = (size_t)rhs * ns * 2;
    c64 x0 = x[off + 2 * n], x1 = x[off + 2 * n + 1];
    for (int j = 0; j < mi; j++) {
        const c64 e = st.eta[(size_t)rhs * M + j];
        const cplx<ZT>* zj = Z + (size_t)j * zstride + off;
        x0 = x0 + e * cvt<double>(zj[2 * n]);
        x1 = x1 + e * cvt<double>(zj[2 * n + 1]);
    }
    x[off + 2 * n] = x0;
    x[off + 2 * n + 1] = x1;
}

// after the last cycle: non-converged count restarts*m (src/fgmres.cpp:84-87)
__global__ void k_outer_final(OuterState st, int nrhs, int restarts, int m) {
    const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
    if (rhs >= nrhs) return;
    if (!st.done[rhs]) st.iters[rhs] = restarts * m;
}

}  // namespace amgs
