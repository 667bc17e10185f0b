// Fine-lattice comparison solvers in FP64, batched over right-hand sides:
// Bi-CGstab on D (src/bi_cgstab.cpp:5-87) and CG on D D^dagger
// (src/conjugate_gradient.cpp:5-65).  Vector work is fused into few passes
// that also produce the per-tile partial sums of the next scalar; the host
// reduces nothing: k_sum_parts folds the partials per right-hand side and the
// few scalars per iteration travel to the host, which evaluates the
// recurrences in the reference's order and passes the coefficients back as
// kernel arguments.
#pragma once
#include "fine.hpp"

namespace amgs {

constexpr int KMAXB = 32;  // right-hand sides per batch (one bit each in `act`)

struct KCoef {
    c64 a[KMAXB], b[KMAXB], c[KMAXB];
    unsigned act;  // right-hand sides still iterating
};

__device__ __forceinline__ bool k_active(const KCoef& k, int rhs) { return (k.act >> rhs) & 1u; }

// out = D in (DAG: D^dagger) with partial sums per tile in part[rhs][tile][2]:
//   mode 1: q0 = sum out . conj(y)
//   mode 2: q0 = sum y . conj(out), q1 = sum |out|^2
template <bool DAG>
__global__ void __launch_bounds__(FT) k_stencil_dots(Geo g, double mm, const c64* __restrict__ U,
                                                     const c64* __restrict__ in, c64* __restrict__ out,
                                                     const c64* __restrict__ y, c64* __restrict__ part, int mode,
                                                     unsigned act) {
    __shared__ c64 red[8];
    const int rhs = blockIdx.y;
    if (!((act >> rhs) & 1u)) return;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int n = tile * FT + threadIdx.x;
    const size_t off = (size_t)rhs * g.ns * 2;
    c64 q[2] = {{0, 0}, {0, 0}};
    if (n < g.ns) {
        c64 o0, o1;
        wilson_site<double, DAG>(U, in + off, n, g.nx, g.nt, mm, AllSites{}, o0, o1);
        out[off + 2 * n] = o0;
        out[off + 2 * n + 1] = o1;
        if (mode != 0) {
            const c64 y0 = y[off + 2 * n], y1 = y[off + 2 * n + 1];
            if (mode == 1) {
                q[0] = mulc(o0, y0) + mulc(o1, y1);
            } else {
                q[0] = mulc(y0, o0) + mulc(y1, o1);
                q[1].re = norm2(o0) + norm2(o1);
            }
        }
    }
    if (mode == 0) return;
    block_sum<2>(q, red);
    if (threadIdx.x == 0) {
        part[((size_t)rhs * gridDim.x + tile) * 2] = q[0];
        part[((size_t)rhs * gridDim.x + tile) * 2 + 1] = q[1];
    }
}

enum KOp {
    KOP_DIR = 0,   // d = r + b (d - c Ad)                       (Bi-CGstab direction)
    KOP_S = 1,     // s = r - a Ad;            q0 = |s|^2
    KOP_RX = 2,    // r = s - c t; x = x + a d + c s;  q0 = <r, rt>
    KOP_XD = 3,    // x = x + a d                                (Bi-CGstab converged)
    KOP_CGXR = 4,  // x += a d; r -= a Ad;     q0 = |r|^2        (CG)
    KOP_CGD = 5,   // d = b d + r                                (CG, b real)
    KOP_DOT = 6,   // q0 = <v0, v1>
};

// One fused vector pass per op over [rhs][ns][2]; vectors v0..v5 as the op
// names them: DIR (d, r, Ad), S (s, r, Ad), RX (r, s, t, x, d, rt), XD (x, d),
// CGXR (x, d, r, Ad), CGD (d, r), DOT (a, b).
__global__ void __launch_bounds__(FT) k_vec(int ns, int op, KCoef k, c64* v0, c64* v1, c64* v2, c64* v3, c64* v4,
                                            c64* v5, c64* part) {
    __shared__ c64 red[8];
    const int rhs = blockIdx.y;
    if (!k_active(k, rhs)) return;
    const int tile = blockIdx.x;
    const size_t off = (size_t)rhs * ns * 2;
    const c64 a = k.a[rhs], b = k.b[rhs], c = k.c[rhs];
    c64 q[1] = {{0, 0}};
    for (int e = tile * 2 * FT + threadIdx.x; e < min(2 * ns, (tile + 1) * 2 * FT); e += FT) {
        const size_t i = off + e;
        switch (op) {
            case KOP_DIR: v0[i] = v1[i] + b * (v0[i] - c * v2[i]); break;
            case KOP_S: {
                const c64 s = v1[i] - a * v2[i];
                v0[i] = s;
                q[0].re += norm2(s);
                break;
            }
            case KOP_RX: {
                const c64 s = v1[i];
                const c64 r = s - c * v2[i];
                v0[i] = r;
                v3[i] = v3[i] + a * v4[i] + c * s;
                q[0] += mulc(r, v5[i]);
                break;
            }
            case KOP_XD: v0[i] = v0[i] + a * v1[i]; break;
            case KOP_CGXR: {
                v0[i] += a * v1[i];
                const c64 r = v2[i] - a * v3[i];
                v2[i] = r;
                q[0].re += norm2(r);
                break;
            }
            case KOP_CGD: {
                c64 d = b * v0[i];
                d += v1[i];
                v0[i] = d;
                break;
            }
            case KOP_DOT: q[0] += mulc(v0[i], v1[i]); break;
        }
    }
    if (!part) return;
    block_sum<1>(q, red);
    if (threadIdx.x == 0) {
        part[((size_t)rhs * gridDim.x + tile) * 2] = q[0];
        part[((size_t)rhs * gridDim.x + tile) * 2 + 1] = c64{0, 0};
    }
}

// out[rhs][q] = sum over tiles of part[rhs][tile][q], q = 0, 1 (fixed order)
__global__ void __launch_bounds__(FT) k_sum_parts(const c64* __restrict__ part, int ntile, c64* __restrict__ out) {
    __shared__ c64 red[8];
    const int rhs = blockIdx.x;
    c64 q[2] = {{0, 0}, {0, 0}};
    for (int t = threadIdx.x; t < ntile; t += FT) {
        q[0] += part[((size_t)rhs * ntile + t) * 2];
        q[1] += part[((size_t)rhs * ntile + t) * 2 + 1];
    }
    block_sum<2>(q, red);
    if (threadIdx.x == 0) {
        out[rhs * 2] = q[0];
        out[rhs * 2 + 1] = q[1];
    }
}

}  // namespace amgs
