// Setup-phase kernels (FP64): aggregate-wise orthonormalisation of the test
// vectors and the Galerkin coarse links (src/level.cpp:242-301, 373-417).
#pragma once
#include "common.hpp"

namespace amgs {

constexpr int ST = 256;

// Modified Gram-Schmidt inside every aggregate a = 2*X + s over the test
// vectors in order, then normalisation (src/level.cpp:373-417).
// w: [ntest][ns][dof]; one workgroup per aggregate.
__global__ void __launch_bounds__(ST) k_orthonormalize(Geo g, c64* w) {
    __shared__ c64 red[8];
    const int a = blockIdx.x, X = a >> 1, s = a & 1;
    const int E = g.spb * g.colors;
    const size_t nvec = (size_t)g.ns * g.dof;
    auto idx = [&](int e) {
        const int m = e / g.colors, c = e - m * g.colors;
        return (size_t)agg_site(g, X, m) * g.dof + 2 * c + s;
    };
    for (int nt = 0; nt < g.ntest; nt++) {
        c64* wn = w + nt * nvec;
        for (int ntt = 0; ntt < nt; ntt++) {
            const c64* wm = w + ntt * nvec;
            c64 q[1] = {{0, 0}};
            for (int e = threadIdx.x; e < E; e += ST) {
                const size_t i = idx(e);
                q[0] += mulc(wn[i], wm[i]);
            }
            block_sum<1>(q, red);
            for (int e = threadIdx.x; e < E; e += ST) {
                const size_t i = idx(e);
                wn[i] -= q[0] * wm[i];
            }
            __syncthreads();
        }
        c64 q[1] = {{0, 0}};
        for (int e = threadIdx.x; e < E; e += ST) q[0].re += norm2(wn[idx(e)]);
        block_sum<1>(q, red);
        const double inv = 1.0 / sqrt(q[0].re);
        for (int e = threadIdx.x; e < E; e += ST) {
            const size_t i = idx(e);
            wn[i] = inv * wn[i];
        }
        __syncthreads();
    }
}

// Level-l stencil coefficient (row (c, alpha), column (b, beta)) of link kind
// k (0 = on-site A, 1+mu = forward hop, 3+mu = backward hop), without the
// boundary sign.  Level 0: A = 0, forward 1/2 (1 - sigma_mu) U_mu(n),
// backward 1/2 (1 + sigma_mu) U_mu(n - mu)^* (src/level.cpp:87-125).
struct LinkSrc {
    Geo g;
    int level0;
    const c64* U;      // level 0
    const c64* links;  // level >= 1: [5][ns][col][row]
    __device__ c64 operator()(int k, int n, int row, int col) const {
        if (!level0) return links[(((size_t)k * g.ns + n) * g.dof + col) * g.dof + row];
        if (k == 0) return c64{0, 0};
        const int mu = (k - 1) & 1;
        const int a = row, b = col;
        c64 p;  // projector entry
        if (k <= 2) {  // 1 - sigma_mu
            if (mu == 0) p = (a == b) ? c64{1, 0} : c64{-1, 0};
            else p = (a == b) ? c64{1, 0} : (a == 0 ? c64{0, 1} : c64{0, -1});
            return 0.5 * (p * U[2 * n + mu]);
        } else {       // 1 + sigma_mu
            if (mu == 0) p = c64{1, 0};
            else p = (a == b) ? c64{1, 0} : (a == 0 ? c64{0, -1} : c64{0, 1});
            return 0.5 * (p * conj(U[2 * nb_left(n, mu, g.nx, g.nt) + mu]));
        }
    }
};

// Galerkin links of level l+1 from level l (src/level.cpp:242-301): for the
// coarse site X and spins (alpha, beta), one workgroup computes
//   A[(p,alpha),(s,beta)]    = sum_{n in X} w_p^* (G1 w_s(n) + in-block hops)
//   B_mu[(p,alpha),(s,beta)] = sum over hops n -> n+mu landing in block X+mu
//   C_mu[(p,alpha),(s,beta)] = sum over hops n -> n-mu landing in block X-mu
// out: links of level l+1, [5][nsc][col][row].
__global__ void __launch_bounds__(ST) k_galerkin(LinkSrc src, const c64* __restrict__ w, int nsc, Geo gc,
                                                 c64* __restrict__ out) {
    const Geo& g = src.g;
    const int X = blockIdx.x, al = blockIdx.y >> 1, be = blockIdx.y & 1;
    const int Nv = g.ntest, C = g.colors, D = g.dof, Dc = 2 * Nv;
    const size_t nvec = (size_t)g.ns * D;
    int XR[2], XL[2];
    for (int mu = 0; mu < 2; mu++) {
        XR[mu] = nb_right(X, mu, gc.nx, gc.nt);
        XL[mu] = nb_left(X, mu, gc.nx, gc.nt);
    }
    for (int o = threadIdx.x; o < 5 * Nv * Nv; o += ST) {
        const int k = o / (Nv * Nv), p = (o / Nv) % Nv, s = o % Nv;
        const c64* wp = w + p * nvec;
        const c64* ws = w + s * nvec;
        c64 acc = {0, 0};
        for (int m = 0; m < g.spb; m++) {
            const int n = agg_site(g, X, m);
            for (int c = 0; c < C; c++) {
                const int fr = 2 * c + al;
                const c64 wpc = conj(wp[(size_t)n * D + fr]);
                for (int b = 0; b < C; b++) {
                    const int fc = 2 * b + be;
                    if (k == 0) {
                        acc += wpc * src(0, n, fr, fc) * ws[(size_t)n * D + fc];
                        for (int mu = 0; mu < 2; mu++) {
                            const int r = nb_right(n, mu, g.nx, g.nt), lf = nb_left(n, mu, g.nx, g.nt);
                            if (agg_block_of(g, r) == X) acc += wpc * src(1 + mu, n, fr, fc) * ws[(size_t)r * D + fc];
                            if (agg_block_of(g, lf) == X) acc += wpc * src(3 + mu, n, fr, fc) * ws[(size_t)lf * D + fc];
                        }
                    } else if (k <= 2) {
                        const int mu = k - 1, r = nb_right(n, mu, g.nx, g.nt), br = agg_block_of(g, r);
                        if (br != X && br == XR[mu]) acc += wpc * src(k, n, fr, fc) * ws[(size_t)r * D + fc];
                    } else {
                        const int mu = k - 3, lf = nb_left(n, mu, g.nx, g.nt), bl = agg_block_of(g, lf);
                        if (bl != X && bl == XL[mu]) acc += wpc * src(k, n, fr, fc) * ws[(size_t)lf * D + fc];
                    }
                }
            }
        }
        const int row = 2 * p + al, col = 2 * s + be;
        out[(((size_t)k * nsc + X) * Dc + col) * Dc + row] = acc;
    }
}

}  // namespace amgs
