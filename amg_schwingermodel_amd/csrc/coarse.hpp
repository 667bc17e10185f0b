// Coarse levels (l >= 1): one persistent workgroup per right-hand side runs the
// whole nested solve of the reference on the device (src/amg.cpp:135-178,
// include/amg.h:19-40, src/fgmres.cpp:5-115, src/sap.cpp:64-174,
// src/level.cpp:161-237,304-368): K-cycle FGMRES, K-cycle recursion, SAP with
// per-block GMRES, coarsest GMRES.  Control flow never returns to the host.
//
// Layout: a level-l vector is [ns][dof]; coarse links are [5][ns][col][row]
// (k = A, B_0, B_1, C_0, C_1) so that the lanes of one site (consecutive rows)
// read consecutive addresses.  Test vectors are [ntest][ns*dof].
//
// Workgroup-level reductions are deterministic (fixed lane butterfly, waves in
// order).  SAP blocks of one colour run concurrently, one wavefront per block,
// each with its own control flow and wave-level reductions (no block barrier).
#pragma once
#include <type_traits>

#include "common.hpp"

namespace amgs {

constexpr int CT = 512;        // threads per coarse workgroup
constexpr int CW = CT / 64;    // wavefronts per coarse workgroup
constexpr int MAXL = 5;        // maximum number of levels
constexpr int KC_MAXM = 8;     // max restart length of K-cycle FGMRES and SAP GMRES
constexpr int CRS_MAXM = 24;   // max restart length of the coarsest GMRES
static_assert(KC_MAXM < 64 && CRS_MAXM < 64, "one lane per Hessenberg row in the column algebra");
constexpr int GRP_MAX = 8;     // max workgroups per right-hand side

enum CMode { CM_KC_SOLVE = 0, CM_SAP = 1, CM_DOP = 2, CM_RESTRICT = 3, CM_PROLONG = 4, CM_CYCLE = 5 };

template <typename R> struct CLevel {
    Geo g;
    const cplx<R>* links;  // [5][ns][dof][dof], column-major per site
    const cplx<R>* links2; // FP32 engine, SAP levels: the same with column pairs interleaved per row, or null
    const cplx<R>* w;      // test vectors [ntest][ns*dof]  (levels < L-1)
    const int* red;        // SAP colour lists [ncol]
    const int* black;
    // per-rhs workspace offsets (complex elements)
    size_t kcV, kcZ, kcR, kcW, kcDx;   // FGMRES of the K-cycle at this level
    size_t crV, crR, crW, crDx;        // coarsest GMRES (l == L-1)
    size_t kDpsi, kR, kEta1, kPsi1;    // k_cycle temporaries (l < L-1)
    size_t sR, sDx;                    // SAP residual / D x
    size_t blk;                        // CW per-wave block workspaces
    size_t blkStride;                  // (sap_m + 3) * sspb * dof
};

template <typename R> struct CArgs {
    int L;
    R mm;
    CLevel<R> lv[MAXL];
    int kc_m, kc_restarts;
    double kc_tol;
    int crs_m, crs_restarts;
    double crs_tol;
    int sap_m, sap_restarts;
    double sap_gtol, sap_tol;
    int nu1, nu2;
    int vcycle;  // 1: V-cycle (src/amg.cpp:90-132), 0: K-cycle (src/amg.cpp:135-178)
    cplx<R>* ws;
    size_t ws_per_rhs;  // workspace entries of one workgroup
    // Workgroups per right-hand side (grp >= 1, modes CM_KC_SOLVE and
    // CM_CYCLE): the grp members of a right-hand side run the whole nested
    // solve redundantly, each in its own workspace, and split the Schwarz
    // blocks of every SAP colour between them; the blocks' results meet in the
    // exchange vector xchg[rhs] behind a barrier of the members (arrival
    // counter bar[rhs], zero at launch).  x_top: workspace offset of the
    // member's private copy of the solve's output.
    int grp;
    // several gauge configurations (mmv non-null): configuration r's links and
    // test vectors at lcs[l], wcs[l] entries past the first's, m0 + 2 = mmv[r]
    size_t lcs[MAXL], wcs[MAXL];
    const R* mmv;
    unsigned* bar;
    cplx<R>* xchg;
    size_t xchg_n, x_top;
    int* err;  // set when a member barrier waits beyond its bound
    const int* rmask;  // per-rhs flags of the calling outer solve (0: skip that rhs); null: none
    // algorithmic stencil flops of this launch (member 0 of each rhs adds
    // them), or null: 8 flops per complex multiply-add of every level
    // operator and Schwarz-block stencil applied
    unsigned long long* cnt;
    int crs_lds;  // coarsest GMRES vectors in dynamic LDS ((crs_m + 4) * n entries)
    int sap_lds[MAXL];  // SAP block vectors of level l in dynamic LDS (CW * blkStride entries)
    // level l's Schwarz blocks solved one per member by the whole workgroup
    // with the block's links staged in dynamic LDS (g_block_gmres_2x2)
    int sap_wg[MAXL];
    unsigned long long* prof;  // optional phase timers [rhs][32] (wall clock ticks)
    int mode, level, nu;
    const cplx<R>* in;
    cplx<R>* out;
    size_t in_stride, out_stride;
};

template <int M> struct FgScal {
    c64 R[M][M];        // rotated Hessenberg (upper triangle)
    c64 G[M + 1][M + 1];  // G[i][k] = <V_i, V_k>, k < i
    c64 g[M + 1], cn[M], sn[M], eta[M];
    c64 col[M + 1], dq[2 * M + 1];
    int brk;
};
struct WaveScal {
    c64 R[KC_MAXM][KC_MAXM];
    c64 G[KC_MAXM][KC_MAXM];  // Gram rows of the cycle's basis: G[i][k] = <V_i, V_k>, k < i
    c64 g[KC_MAXM + 1], cn[KC_MAXM], sn[KC_MAXM], eta[KC_MAXM];
    c64 col[KC_MAXM + 1];
    c64 rcol[KC_MAXM + 1];    // the column as rotated (w_gs_rotate)
    c64 dq[2 * KC_MAXM + 2];  // quantities of a step's inner-product pass
    double bsc[2];  // |b_B|, |r_B| broadcast to the workgroup (g_block_gmres_2x2)
    double hn;      // |w - sum h_i V_i| of the step, broadcast (g_block_gmres_2x2)
    int brk;
};
constexpr int VIN_BYTES = 16384;  // LDS staging of stencil inputs
// A suspended solve of the nested K-cycle: FGMRES at a level (in = b, out = x),
// k_cycle at a level (in = eta, out = psi) or the coarsest GMRES.
struct CFrame {
    const void* in;
    void* out;
    int pc, k, j, maxIt;
    double nb, err;
};
struct CLds {
    c64 red[CT];
    alignas(16) char vin[VIN_BYTES];
    FgScal<KC_MAXM> kc[MAXL];
    FgScal<CRS_MAXM> crs;
    WaveScal wv[CW];
    CFrame fg[MAXL], kcf[MAXL], crsf;  // frames per level and kind
    int stk[2 * MAXL + 2], sp;         // frame stack: kind * 16 + level
    unsigned bar_epi;                  // member barriers passed so far
};

// ---------------------------------------------------------------------------
// workgroup-level vector primitives (all threads call them uniformly)
// ---------------------------------------------------------------------------
template <typename R> struct CK {
    const CArgs<R>& A;
    cplx<R>* W;    // this workgroup's workspace (global memory)
    CLds& S;
    cplx<R>* dyn;  // dynamic LDS region
    unsigned long long* prof;
    int rhs, member;  // right-hand side and member index (0 .. grp-1)
};

// Barrier of the grp members of one right-hand side.  Every thread makes its
// Member barrier.  Every wavefront first waits until its own global stores
// (the published blocks) have reached its XCD's L2; after the workgroup
// barrier thread 0 alone releases at agent scope (one L2 write-back: the
// members may sit on different XCDs, whose L2 caches are not coherent with
// each other), counts the arrival, waits for the episode's total and
// acquires at agent scope (one invalidate of the CU's L1 and the XCD's L2),
// and the closing workgroup barrier orders every wavefront's later reads
// after it.  The wait is bounded: past the bound the error flag is raised and
// the member goes on (the host reports the error), so a launch always drains.
constexpr unsigned MEMBER_SPIN_LIMIT = 1u << 22;
template <typename R> __device__ __forceinline__ void member_sync(CK<R>& K) {
    const CArgs<R>& A = K.A;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned target = (unsigned)A.grp * ++K.S.bar_epi;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(A.bar + K.rhs, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (__hip_atomic_load(A.bar + K.rhs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > MEMBER_SPIN_LIMIT) {
                __hip_atomic_store(A.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// zero the member-barrier arrival counters before a grouped launch (a kernel,
// not a memset, so that the launch sequence can be captured in a HIP graph)
static __global__ void k_member_reset(int n, unsigned* bar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bar[i] = 0u;
}

// phase timers (thread 0, device wall clock), only when a buffer is given
struct PhaseTimer {
    unsigned long long* p;
    unsigned long long t;
    __device__ __forceinline__ PhaseTimer(unsigned long long* prof) : p(prof), t(0) {
        if (p && threadIdx.x == 0) t = wall_clock64();
    }
    __device__ __forceinline__ void lap(int id) {
        if (p && threadIdx.x == 0) {
            const unsigned long long n = wall_clock64();
            p[id] += n - t;
            t = n;
        }
    }
    __device__ __forceinline__ void count(int id, unsigned long long v = 1) {
        if (p && threadIdx.x == 0) p[id] += v;
    }
};

// per-wavefront phase timers (lane 0 of each wavefront), summed over the
// wavefronts of the workgroup, only when a buffer is given
struct WaveTimer {
    unsigned long long* p;
    unsigned long long t;
    __device__ __forceinline__ WaveTimer(unsigned long long* prof) : p(prof), t(0) {
        if (p && (threadIdx.x & 63) == 0) t = wall_clock64();
    }
    __device__ __forceinline__ void lap(int id) {
        if (p && (threadIdx.x & 63) == 0) {
            const unsigned long long n = wall_clock64();
            atomicAdd(p + id, n - t);
            t = n;
        }
    }
    __device__ __forceinline__ void count(int id) {
        if (p && (threadIdx.x & 63) == 0) atomicAdd(p + id, 1ull);
    }
};

template <typename R> __device__ __forceinline__ c64 c_dot(CLds& S, const cplx<R>* a, const cplx<R>* b, int n) {
    c64 q[1] = {{0, 0}};
    for (int e = threadIdx.x; e < n; e += CT) q[0] += mulc(cvt<double>(a[e]), cvt<double>(b[e]));
    __syncthreads();
    block_sum<1>(q, S.red);
    return q[0];
}

// out = D_l in (src/level.cpp:161-180): (m0+2) v - A v - sum_mu (B_mu s+ v(x+mu) + C_mu s- v(x-mu)).
// One thread per output row; the five link blocks are walked column by column
// with independent accumulators (many loads in flight).  The input vector is
// staged in LDS when it is not there already (all rows of a site read it).
// Global-memory view of a pointer: loads through it are global (not flat)
// instructions, so they do not serialise against LDS traffic.
template <typename T> using gptr = const __attribute__((address_space(1))) T*;
template <typename T> __device__ __forceinline__ gptr<T> as_global(const T* p) { return (gptr<T>)p; }
template <typename R> __device__ __forceinline__ cplx<R> ld(gptr<cplx<R>> p) {
    using V2 = R __attribute__((ext_vector_type(2)));
    const V2 v = *(const __attribute__((address_space(1))) V2*)p;
    return {v.x, v.y};
}
template <typename R> __device__ __forceinline__ cplx<R> ld(const cplx<R>* p) { return *p; }

// Row r of site x: the five link blocks (column stride D) against the on-site
// and the four neighbour vectors, five independent accumulators.
template <typename R, typename LP, typename VP>
__device__ __forceinline__ void stencil_row_p(LP L0, LP L1, LP L2, LP L3, LP L4, int D, VP v0, VP vr0, VP vr1, VP vl0,
                                              VP vl1, R f1, R f2, R f3, R f4, cplx<R>& acc) {
    cplx<R> a0 = czero<R>(), a1 = czero<R>(), a2 = czero<R>(), a3 = czero<R>(), a4 = czero<R>();
    // U columns of the five blocks in flight per round trip (fewer in FP64)
    constexpr int U = sizeof(R) == 4 ? 8 : 4;
    for (int c0 = 0; c0 < D; c0 += U) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = c0 + u;
            if (c < D) {
                const int o = c * D;
                a0 += ld<R>(L0 + o) * ld<R>(v0 + c);
                a1 += ld<R>(L1 + o) * ld<R>(vr0 + c);
                a2 += ld<R>(L2 + o) * ld<R>(vr1 + c);
                a3 += ld<R>(L3 + o) * ld<R>(vl0 + c);
                a4 += ld<R>(L4 + o) * ld<R>(vl1 + c);
            }
        }
    }
    acc -= a0;
    acc -= f1 * a1 + f3 * a3;
    acc -= f2 * a2 + f4 * a4;
}

template <typename R, typename LP, typename VP>
__device__ __forceinline__ void stencil_row(LP links, int ksz, int msz, int D, int x, int r, VP v0, VP vr0, VP vr1,
                                            VP vl0, VP vl1, R f1, R f2, R f3, R f4, cplx<R>& acc) {
    LP L0 = links + x * msz + r;
    stencil_row_p<R>(L0, L0 + ksz, L0 + 2 * ksz, L0 + 3 * ksz, L0 + 4 * ksz, D, v0, vr0, vr1, vl0, vl1, f1, f2, f3, f4,
                     acc);
}

template <typename R, typename LP, typename VP>
__device__ __forceinline__ cplx<R> dop_row(const CArgs<R>& A, const Geo& g, LP links, VP src, int e) {
    const int D = g.dof;
    const int msz = D * D, ksz = g.ns * msz;
    const int x = e / D, r = e - x * D;
    const int xr0 = nb_right(x, 0, g.nx, g.nt), xr1 = nb_right(x, 1, g.nx, g.nt);
    const int xl0 = nb_left(x, 0, g.nx, g.nt), xl1 = nb_left(x, 1, g.nx, g.nt);
    const R f1 = wraps_right(x, 0, g.nt) ? R(-1) : R(1);
    const R f3 = wraps_left(x, 0, g.nt) ? R(-1) : R(1);
    cplx<R> acc = A.mm * ld<R>(src + e);
    stencil_row<R>(links, ksz, msz, D, x, r, src + x * D, src + xr0 * D, src + xr1 * D, src + xl0 * D, src + xl1 * D,
                   f1, R(1), f3, R(1), acc);
    return acc;
}
template <typename R, typename LP, typename VP>
__device__ __forceinline__ void dop_rows_l(const CArgs<R>& A, int l, LP links, VP src, cplx<R>* out) {
    const Geo& g = A.lv[l].g;
    const int n = g.ns * g.dof;
    for (int e = threadIdx.x; e < n; e += CT) out[e] = dop_row<R>(A, g, links, src, e);
}
template <typename R, typename VP>
__device__ __forceinline__ void dop_rows(const CArgs<R>& A, int l, VP src, cplx<R>* out) {
    dop_rows_l<R>(A, l, as_global(A.lv[l].links), src, out);
}

// out = D_l in (src/level.cpp:161-180): (m0+2) v - A v - sum_mu (B_mu s+ v(x+mu) + C_mu s- v(x-mu)).
// One thread per output row; the five link blocks are walked column by column
// with independent accumulators.  The input is first copied to the LDS
// staging buffer when it fits (every row of a site reads it).
template <typename R>
__device__ __forceinline__ void c_dop(const CArgs<R>& A, int l, const cplx<R>* in, cplx<R>* out, cplx<R>* stage = nullptr,
                                      int stage_cap = 0) {
    const int n = A.lv[l].g.ns * A.lv[l].g.dof;
    if (A.cnt && threadIdx.x == 0 && blockIdx.x % A.grp == 0)
        atomicAdd(A.cnt, 8ull * 5 * A.lv[l].g.dof * (unsigned long long)n);
    __syncthreads();
    if (stage && n <= stage_cap) {
        for (int e = threadIdx.x; e < n; e += CT) stage[e] = in[e];
        __syncthreads();
        dop_rows<R>(A, l, (const cplx<R>*)stage, out);
    } else {
        dop_rows<R>(A, l, as_global(in), out);
    }
    __syncthreads();
}

// r = v - D_l x on the rows of Schwarz block blk only (the rows of every
// other block keep what they hold), with the bits of c_dop and the
// subtraction after it: D_l x is stored (Dx) and subtracted after a barrier,
// so no contraction joins the two.
template <typename R>
__device__ __forceinline__ void c_resid_block(const CArgs<R>& A, int l, int blk, const cplx<R>* v, const cplx<R>* x,
                                              cplx<R>* Dx, cplx<R>* r, cplx<R>* stage, int stage_cap) {
    const Geo& g = A.lv[l].g;
    const int D = g.dof, nb = g.sspb * D, n = g.ns * D;
    if (A.cnt && threadIdx.x == 0 && blockIdx.x % A.grp == 0)
        atomicAdd(A.cnt, 8ull * 5 * D * (unsigned long long)nb);
    __syncthreads();
    // x read as c_dop reads it (the staged copy where it fits), so that the
    // rows come from the same code
    auto rows = [&](auto src) {
        for (int q = threadIdx.x; q < nb; q += CT) {
            const int m = q / D, d = q - m * D;
            const int e = sap_site(g, blk, m) * D + d;
            Dx[e] = dop_row<R>(A, g, as_global(A.lv[l].links), src, e);
        }
    };
    if (stage && n <= stage_cap) {
        for (int e = threadIdx.x; e < n; e += CT) stage[e] = x[e];
        __syncthreads();
        rows((const cplx<R>*)stage);
    } else {
        rows(as_global(x));
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nb; q += CT) {
        const int m = q / D, d = q - m * D;
        const int e = sap_site(g, blk, m) * D + d;
        r[e] = v[e] - Dx[e];
    }
    __syncthreads();
}

// out = P^dagger_l vf (src/level.cpp:212-237); one thread per coarse entry,
// summed over the aggregate in the reference order (x, t, colour).
template <typename R> __device__ __forceinline__ void c_restrict(const CArgs<R>& A, int l, const cplx<R>* vf, cplx<R>* out) {
    const CLevel<R>& L = A.lv[l];
    const Geo& g = L.g;
    const int Dc = A.lv[l + 1].g.dof, nc = A.lv[l + 1].g.ns * Dc, D = g.dof;
    const size_t nvec = (size_t)g.ns * D;
    __syncthreads();
    for (int e = threadIdx.x; e < nc; e += CT) {
        const int X = e / Dc, q = e - X * Dc, cc = q >> 1, s = q & 1;
        c64 acc = {0, 0};
        const cplx<R>* w = L.w + cc * nvec;
        for (int m = 0; m < g.spb; m++) {
            const int n = agg_site(g, X, m);
            for (int c = 0; c < g.colors; c++) {
                const size_t i = (size_t)n * D + 2 * c + s;
                acc += mulc(cvt<double>(vf[i]), cvt<double>(w[i]));
            }
        }
        out[e] = cvt<R>(acc);
    }
    __syncthreads();
}

// out (+)= P_l vc (src/level.cpp:182-209)
template <typename R>
__device__ __forceinline__ void c_prolong(const CArgs<R>& A, int l, const cplx<R>* vc, cplx<R>* out, bool accumulate) {
    const CLevel<R>& L = A.lv[l];
    const Geo& g = L.g;
    const int D = g.dof, n = g.ns * D, Dc = A.lv[l + 1].g.dof;
    const size_t nvec = (size_t)n;
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += CT) {
        const int x = e / D, d = e - x * D, s = d & 1;
        const cplx<R>* c = vc + (size_t)agg_block_of(g, x) * Dc;
        cplx<R> o = czero<R>();
        for (int cc = 0; cc < g.ntest; cc++) o += L.w[cc * nvec + e] * c[2 * cc + s];
        out[e] = accumulate ? out[e] + o : o;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// wavefront-level pieces for one Schwarz block (src/sap.cpp:44-62,
// src/level.cpp:304-368).  Block vectors are [sspb][dof], local index
// m = mx * stel + mt.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <typename R> __device__ __forceinline__ c64 w_dot(const cplx<R>* a, const cplx<R>* b, int n) {
    const int lane = threadIdx.x & 63;
    c64 s = {0, 0};
    for (int e = lane; e < n; e += 64) s += mulc(cvt<double>(a[e]), cvt<double>(b[e]));
    return wave_sum(s);
}

// block-local stencil, neighbours outside block blk read as zero: multiplier
// 0, and both the vector and the link reads are redirected to the site's own
// (already loaded) entries, so no link block outside the block is fetched
template <typename R, typename LP>
__device__ __forceinline__ void w_dlocal_l(const CArgs<R>& A, int l, int blk, LP links, const cplx<R>* in, cplx<R>* out) {
    const CLevel<R>& L = A.lv[l];
    const Geo& g = L.g;
    const int D = g.dof, nb = g.sspb * D, lane = threadIdx.x & 63;
    const int msz = D * D, ksz = g.ns * msz;
    for (int e = lane; e < nb; e += 64) {
        const int m = e / D, r = e - m * D;
        const int n = sap_site(g, blk, m);
        int nbr[4] = {nb_right(n, 0, g.nx, g.nt), nb_right(n, 1, g.nx, g.nt), nb_left(n, 0, g.nx, g.nt),
                      nb_left(n, 1, g.nx, g.nt)};
        R f[4];
        const cplx<R>* vp[4];
        LP lk[5];
        lk[0] = links + n * msz + r;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const bool inb = sap_block_of(g, nbr[q]) == blk;
            const bool wrap = (q == 0 && wraps_right(n, 0, g.nt)) || (q == 2 && wraps_left(n, 0, g.nt));
            f[q] = inb ? (wrap ? R(-1) : R(1)) : R(0);
            vp[q] = in + (size_t)(inb ? sap_local_of(g, nbr[q]) : m) * D;
            lk[q + 1] = inb ? lk[0] + (q + 1) * ksz : lk[0];
        }
        cplx<R> acc = A.mm * in[e];
        stencil_row_p<R>(lk[0], lk[1], lk[2], lk[3], lk[4], D, in + m * D, vp[0], vp[1], vp[2], vp[3], f[0], f[1], f[2],
                         f[3], acc);
        out[e] = acc;
    }
    wave_sync();
}
// 2x2-site Schwarz blocks on a square level larger than 2x2 (the reference's
// blocking of the coarse grids): every site has exactly one in-block
// neighbour in t and one in x and no block wraps the time boundary, so a row
// walks three link blocks (A, the t hop, the x hop) instead of five; the two
// dropped hops are the exact zeros of the general form.  Local index
// m = 2 mx + mt; the t neighbour is m ^ 1, the x neighbour m ^ 2.
template <typename R, typename LP>
__device__ __forceinline__ void w_dlocal_2x2(const CArgs<R>& A, int l, int blk, LP links, const cplx<R>* in,
                                             cplx<R>* out) {
    const Geo& g = A.lv[l].g;
    const int D = g.dof, nb = 4 * D, lane = threadIdx.x & 63;
    const int msz = D * D, ksz = g.ns * msz;
    for (int e = lane; e < nb; e += 64) {
        const int m = e / D, r = e - m * D;
        const int n = sap_site(g, blk, m);
        const LP L0 = links + n * msz + r;
        const LP Lt = L0 + ((m & 1) ? 3 : 1) * ksz;  // C_0 (backward) from mt = 1, B_0 (forward) from mt = 0
        const LP Lx = L0 + ((m & 2) ? 4 : 2) * ksz;  // C_1 from mx = 1, B_1 from mx = 0
        const cplx<R>* v0 = in + m * D;
        const cplx<R>* vt = in + (m ^ 1) * D;
        const cplx<R>* vx = in + (m ^ 2) * D;
        cplx<R> a0 = czero<R>(), at = czero<R>(), ax = czero<R>();
        // U columns of the three blocks in flight per round trip (fewer in FP64,
        // whose complex values take twice the registers)
        constexpr int U = sizeof(R) == 4 ? 10 : 5;
        for (int c0 = 0; c0 < D; c0 += U) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int c = c0 + u;
                if (c < D) {
                    const int o = c * D;
                    a0 += ld<R>(L0 + o) * ld<R>(v0 + c);
                    at += ld<R>(Lt + o) * ld<R>(vt + c);
                    ax += ld<R>(Lx + o) * ld<R>(vx + c);
                }
            }
        }
        cplx<R> acc = A.mm * in[e];
        acc -= a0;
        acc -= at;
        acc -= ax;
        out[e] = acc;
    }
    wave_sync();
}

// The same 2x2-block stencil with its 3 x 4D (row, link block) partial sums
// spread evenly over the lanes (at most 4 per lane, all their columns in
// flight together: two round trips instead of four for D = 20), combined per
// row through a wave-private LDS area in the same order as w_dlocal_2x2.
template <typename R, typename LP>
__device__ __forceinline__ void w_dlocal_2x2s(const CArgs<R>& A, int l, int blk, LP links, const cplx<R>* in,
                                              cplx<R>* out, cplx<R>* part) {
    const Geo& g = A.lv[l].g;
    const int D = g.dof, nb = 4 * D, nu = 3 * nb, lane = threadIdx.x & 63;
    const int msz = D * D, ksz = g.ns * msz;
    constexpr int S = 4;  // units per lane (nu <= 4 * 64)
    LP Lp[S];
    const cplx<R>* vp[S];
    cplx<R> acc[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int u0 = lane + 64 * s, u = u0 < nu ? u0 : 0;
        const int q = u / nb, e = u - q * nb;  // link block (0 A, 1 t hop, 2 x hop) and row
        const int m = e / D, r = e - m * D;
        const int n = sap_site(g, blk, m);
        const int k = q == 0 ? 0 : q == 1 ? ((m & 1) ? 3 : 1) : ((m & 2) ? 4 : 2);
        Lp[s] = links + k * ksz + n * msz + r;
        vp[s] = in + (q == 0 ? m : q == 1 ? (m ^ 1) : (m ^ 2)) * D;
        acc[s] = czero<R>();
    }
    constexpr int U = 10;
    for (int c0 = 0; c0 < D; c0 += U) {
#pragma unroll
        for (int uu = 0; uu < U; uu++) {
            const int c = c0 + uu;
            if (c < D) {
#pragma unroll
                for (int s = 0; s < S; s++) acc[s] += ld<R>(Lp[s] + c * D) * ld<R>(vp[s] + c);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; s++)
        if (lane + 64 * s < nu) part[lane + 64 * s] = acc[s];
    wave_sync();
    for (int e = lane; e < nb; e += 64) {
        cplx<R> o = A.mm * in[e];
        o -= part[e];
        o -= part[nb + e];
        o -= part[2 * nb + e];
        out[e] = o;
    }
    wave_sync();
}

// w_dlocal_2x2s with the links in column-pair order [5][ns][dof/2][dof][2]:
// each lane reads columns c, c+1 of its row, and of its input, with one
// 16-byte load each, half the vector-memory instructions of the column-major
// form, which bound the stencil (a workgroup's eight wavefronts share the
// CU's memory pipeline).  Same products in the same order, so the result is
// that of w_dlocal_2x2s.  Needs an even dof and 16-byte aligned block vectors.
template <typename R>
__device__ __forceinline__ void w_dlocal_2x2p(const CArgs<R>& A, int l, int blk, gptr<cplx<R>> links2,
                                              const cplx<R>* in, cplx<R>* out, cplx<R>* part) {
    static_assert(sizeof(R) == 4, "FP32 engine only");
    using V4 = float __attribute__((ext_vector_type(4)));
    const Geo& g = A.lv[l].g;
    const int D = g.dof, nb = 4 * D, nu = 3 * nb, lane = threadIdx.x & 63;
    const int msz = D * D, ksz = g.ns * msz, np = D / 2;
    constexpr int S = 4;  // units per lane (nu <= 4 * 64)
    const __attribute__((address_space(1))) V4* Lp[S];
    const V4* vp[S];
    cplx<R> acc[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int u0 = lane + 64 * s, u = u0 < nu ? u0 : 0;
        const int q = u / nb, e = u - q * nb;  // link block (0 A, 1 t hop, 2 x hop) and row
        const int m = e / D, r = e - m * D;
        const int n = sap_site(g, blk, m);
        const int k = q == 0 ? 0 : q == 1 ? ((m & 1) ? 3 : 1) : ((m & 2) ? 4 : 2);
        Lp[s] = (const __attribute__((address_space(1))) V4*)(links2 + k * ksz + n * msz + 2 * r);
        vp[s] = (const V4*)(in + (q == 0 ? m : q == 1 ? (m ^ 1) : (m ^ 2)) * D);
        acc[s] = czero<R>();
    }
    constexpr int U = 5;  // column pairs per round
    for (int p0 = 0; p0 < np; p0 += U) {
#pragma unroll
        for (int uu = 0; uu < U; uu++) {
            const int p = p0 + uu;
            if (p < np) {
#pragma unroll
                for (int s = 0; s < S; s++) {
                    const V4 lv = Lp[s][p * D], v = vp[s][p];
                    acc[s] += cplx<R>{lv.x, lv.y} * cplx<R>{v.x, v.y};
                    acc[s] += cplx<R>{lv.z, lv.w} * cplx<R>{v.z, v.w};
                }
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; s++)
        if (lane + 64 * s < nu) part[lane + 64 * s] = acc[s];
    wave_sync();
    for (int e = lane; e < nb; e += 64) {
        cplx<R> o = A.mm * in[e];
        o -= part[e];
        o -= part[nb + e];
        o -= part[2 * nb + e];
        out[e] = o;
    }
    wave_sync();
}

// part: wave-private LDS scratch of >= 12 D complex entries (FP32 engine), or null
template <typename R>
__device__ __forceinline__ void w_dlocal(const CArgs<R>& A, int l, int blk, const cplx<R>* in, cplx<R>* out,
                                         cplx<R>* part = nullptr) {
    const Geo& g = A.lv[l].g;
    // stencil flops (member 0 of the rhs): rows x in-block hops x dof
    if (A.cnt && (threadIdx.x & 63) == 0 && blockIdx.x % A.grp == 0) {
        const bool b22 = g.sxel == 2 && g.stel == 2 && g.nx == g.nt && g.nx > 2;
        atomicAdd(A.cnt, 8ull * (b22 ? 3 : 5) * g.dof * (unsigned long long)(g.sspb * g.dof));
    }
    if (g.sxel == 2 && g.stel == 2 && g.nx == g.nt && g.nx > 2) {
        if constexpr (std::is_same<R, float>::value) {
            if (part && 12 * g.dof <= 4 * 64) {
                if (A.lv[l].links2) w_dlocal_2x2p<R>(A, l, blk, as_global(A.lv[l].links2), in, out, part);
                else w_dlocal_2x2s<R>(A, l, blk, as_global(A.lv[l].links), in, out, part);
                return;
            }
        }
        w_dlocal_2x2<R>(A, l, blk, as_global(A.lv[l].links), in, out);
    } else {
        w_dlocal_l<R>(A, l, blk, as_global(A.lv[l].links), in, out);
    }
}

// Modified Gram-Schmidt of block GMRES step j (src/fgmres.cpp:27-34) from
// one pass of inner products instead of j + 1 dependent ones.  For w = D_B V_j
// and the stored unit vectors V_0..V_j, MGS projects one direction at a time,
// h_i = <w - sum_{k<i} h_k V_k, V_i> = d_i - sum_{k<i} h_k conj(G_ik) with
// d_i = <w, V_i> and G_ik = <V_i, V_k>, and each projection takes |h_i|^2
// off the squared norm: |w - sum_i h_i V_i|^2 = |w|^2 - sum_i |h_i|^2.
// Pass quantity q (one wavefront: lane partial sums over e = lane + 64 t in
// order, then the wave butterfly): q <= j: d_q; j < q <= 2j: G_{j,q-j-1};
// q = 2j + 1: |w|^2.
template <typename R, typename VF>
__device__ __forceinline__ c64 w_gs_quantity(int q, int j, const cplx<R>* w, VF V, int nb) {
    const cplx<R>* a = q <= j || q == 2 * j + 1 ? w : V(j);
    const cplx<R>* b = q <= j ? V(q) : q == 2 * j + 1 ? w : V(q - j - 1);
    return w_dot(a, b, nb);
}
// One wavefront, after the pass quantities are in S.dq: the column h_0..h_j
// into S.col (lane i forms h_i; h_k is final once the lanes below it are),
// the Gram row j into S.G; returns |w - sum_i h_i V_i|^2 (wave-uniform).
__device__ __forceinline__ double w_gs_column(WaveScal& S, int j) {
    const int lane = threadIdx.x & 63;
    c64 d = {0.0, 0.0};
    if (lane <= j) d = S.dq[lane];
    double u2 = S.dq[2 * j + 1].re;
    for (int k = 0; k <= j; k++) {
        const c64 hk = {readlane_d(d.re, k), readlane_d(d.im, k)};
        u2 -= norm2(hk);
        if (lane > k && lane <= j) {
            const c64 gik = lane == j ? S.dq[j + 1 + k] : S.G[lane][k];
            d -= hk * conj(gik);
        }
    }
    if (lane <= j) S.col[lane] = d;
    if (lane < j) S.G[j][lane] = S.dq[j + 1 + lane];
    return u2 > 0 ? u2 : 0.0;
}
// Element e of V_{j+1} = (w - sum_i h_i V_i) / |.| (w is stored in V_{j+1});
// hn = |.| from the algebra (0: the difference is stored unscaled).
template <typename R, typename VF>
__device__ __forceinline__ void w_gs_form(const WaveScal& S, VF V, int j, double hn, int e) {
    cplx<R>* w = V(j + 1);
    c64 u = cvt<double>(w[e]);
    for (int i = 0; i <= j; i++) u -= S.col[i] * cvt<double>(V(i)[e]);
    w[e] = cvt<R>(hn > 0 ? (1.0 / hn) * u : u);
}

// Givens rotation of column col[0..j+1] with stored cn/sn (src/fgmres.cpp:90-104);
// col, cn, sn live in LDS; one thread executes it.  col[j] becomes the
// diagonal of R; returns cn_j, sn_j.
__device__ __forceinline__ void givens_col(c64* col, const c64* cn, const c64* sn, int j, c64& cnj, c64& snj) {
    for (int i = 0; i < j; i++) {
        const c64 a = col[i], b = col[i + 1];
        col[i + 1] = -(sn[i] * a) + cn[i] * b;
        col[i] = conj(cn[i]) * a + conj(sn[i]) * b;
    }
    const c64 a = col[j], b = col[j + 1];
    const double den = sqrt(norm2(a) + norm2(b));
    snj = c64{b.re / den, b.im / den};
    cnj = c64{a.re / den, a.im / den};
    col[j] = conj(cnj) * a + conj(snj) * b;
}

// Step j's Givens rotation and residual estimate (src/fgmres.cpp:35-49) on a
// copy of the column (col stays unrotated for w_gs_form); one lane.
__device__ __forceinline__ void w_gs_rotate(WaveScal& S, int j, double hn, double gtol) {
    for (int i = 0; i <= j; i++) S.rcol[i] = S.col[i];
    S.rcol[j + 1] = c64{hn, 0.0};
    c64 cnj, snj;
    givens_col(S.rcol, S.cn, S.sn, j, cnj, snj);
    for (int i = 0; i <= j; i++) S.R[i][j] = S.rcol[i];
    const c64 gj = S.g[j];
    const c64 gn = -(snj * gj);
    S.cn[j] = cnj;
    S.sn[j] = snj;
    S.g[j] = conj(cnj) * gj;
    S.g[j + 1] = gn;
    S.brk = sqrt(norm2(gn)) < gtol;
}

// GMRES(m, restarts, tol) on D_B for one Schwarz block, x0 = b_B; then
// x[B] += x_B.  Executed by one wavefront; modified Gram-Schmidt from one
// pass of inner products per step (w_gs_column).
template <typename R>
__device__ __forceinline__ void w_block_gmres(const CArgs<R>& A, CK<R>& K, int l, int blk, const cplx<R>* rglob, cplx<R>* x) {
    const CLevel<R>& L = A.lv[l];
    const Geo& g = L.g;
    const int D = g.dof, nb = g.sspb * D, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = A.sap_m;
    WaveScal& S = K.S.wv[wv];
    // block Krylov vectors in the dynamic LDS region when they fit (it is
    // otherwise used only by the coarsest GMRES, never live during SAP)
    cplx<R>* base = (A.sap_lds[l] ? K.dyn : K.W + L.blk) + (size_t)wv * L.blkStride;
    auto V = [&](int j) { return base + (size_t)j * nb; };
    cplx<R>* xb = base + (size_t)(m + 1) * nb;
    cplx<R>* bb = base + (size_t)(m + 2) * nb;
    // partial sums of the block stencil: this wave's slice of the stencil
    // staging buffer (idle while the colour's blocks are solved)
    constexpr int PART = VIN_BYTES / CW / (int)sizeof(cplx<R>);
    cplx<R>* part = 12 * D <= PART ? (cplx<R>*)K.S.vin + wv * PART : nullptr;
    // gather b_B; x_B = b_B
    for (int e = lane; e < nb; e += 64) {
        const int mm = e / D, d = e - mm * D;
        const cplx<R> v = rglob[(size_t)sap_site(g, blk, mm) * D + d];
        bb[e] = v;
        xb[e] = v;
    }
    wave_sync();
    w_dlocal(A, l, blk, xb, V(1), part);
    for (int e = lane; e < nb; e += 64) V(0)[e] = bb[e] - V(1)[e];
    wave_sync();
    WaveTimer T(K.prof);
    const double normb = sqrt(w_dot(bb, bb, nb).re);
    double err = sqrt(w_dot(V(0), V(0), nb).re);
    T.lap(24);
    if (normb == 0.0) return;  // zero block right-hand side: x_B = 0
    int maxIt = m;
    for (int k = 0; k < A.sap_restarts; k++) {
        {
            const R ib = R(1.0 / err);
            for (int e = lane; e < nb; e += 64) V(0)[e] = ib * V(0)[e];
        }
        c64 g0 = {err, 0.0};
        if (lane == 0) S.g[0] = g0;
        wave_sync();
        for (int j = 0; j < m; j++) {
            w_dlocal(A, l, blk, V(j), V(j + 1), part);
            T.lap(25);
            T.count(19);
            cplx<R>* w = V(j + 1);
            for (int q = 0; q <= 2 * j + 1; q++) {
                const c64 v = w_gs_quantity<R>(q, j, w, V, nb);
                if (lane == 0) S.dq[q] = v;
            }
            wave_sync();
            const double hn = sqrt(w_gs_column(S, j));
            wave_sync();
            T.lap(26);
            for (int e = lane; e < nb; e += 64) w_gs_form<R>(S, V, j, hn, e);
            wave_sync();
            T.lap(27);
            if (lane == 0) w_gs_rotate(S, j, hn, A.sap_gtol * normb);
            wave_sync();
            T.lap(28);
            if (S.brk) {
                maxIt = j + 1;
                break;
            }
        }
        if (lane == 0) {
            for (int i = maxIt - 1; i >= 0; i--) {
                c64 e = S.g[i];
                for (int q = i + 1; q < maxIt; q++) e -= S.R[i][q] * S.eta[q];
                S.eta[i] = cdiv(e, S.R[i][i]);
            }
        }
        wave_sync();
        T.lap(29);
        for (int e = lane; e < nb; e += 64) {
            c64 s = cvt<double>(xb[e]);
            for (int j = 0; j < maxIt; j++) s += S.eta[j] * cvt<double>(V(j)[e]);
            xb[e] = cvt<R>(s);
        }
        wave_sync();
        if (k + 1 == A.sap_restarts) break;  // the last cycle's residual decides nothing
        w_dlocal(A, l, blk, xb, V(1), part);
        for (int e = lane; e < nb; e += 64) V(0)[e] = bb[e] - V(1)[e];
        wave_sync();
        err = sqrt(w_dot(V(0), V(0), nb).re);
        T.lap(30);
        if (err < A.sap_gtol * normb) break;
    }
    for (int e = lane; e < nb; e += 64) {
        const int mm = e / D, d = e - mm * D;
        x[(size_t)sap_site(g, blk, mm) * D + d] += xb[e];
    }
}

// Dynamic LDS of g_block_gmres_2x2 (bytes): the block's links in column-pair
// order, one partial sum per (row, link block) and the block's m + 3 vectors.
__host__ __device__ inline size_t wg_block_lds_bytes(int dof, int m) {
    const size_t nb = 4 * (size_t)dof, nu = 3 * nb, np = dof / 2;
    return np * nu * 16 + nu * 8 + (size_t)(m + 3) * nb * 8;
}

// The block stencil of g_block_gmres_2x2: out = D_B in on a 2x2-site block,
// in and out in LDS.  Thread u < 12 D owns one (row, link block) partial sum
// over the row's D columns, reading the staged links and the input in
// 16-byte pairs; the row sums are formed in the order of w_dlocal_2x2p.
template <typename R>
__device__ __forceinline__ void wg_dlocal_2x2(R mm, int D, const float __attribute__((ext_vector_type(4)))* Lk,
                                              cplx<R>* part, const cplx<R>* in, cplx<R>* out,
                                              unsigned long long* cnt = nullptr, bool synced = false) {
    using V4 = float __attribute__((ext_vector_type(4)));
    const int nb = 4 * D, nu = 3 * nb, np = D / 2, u = threadIdx.x;
    if (cnt && u == 0) atomicAdd(cnt, 8ull * 3 * D * (unsigned long long)nb);
    if (!synced) __syncthreads();  // in is complete (synced: the caller's barrier saw to it)
    if (u < nu) {
        const int q = u / nb, e = u - q * nb, m = e / D;
        const V4* vp = (const V4*)(in + (q == 0 ? m : q == 1 ? (m ^ 1) : (m ^ 2)) * D);
        cplx<R> acc = czero<R>();
        for (int p = 0; p < np; p++) {
            const V4 lv = Lk[p * nu + u], v = vp[p];
            acc += cplx<R>{lv.x, lv.y} * cplx<R>{v.x, v.y};
            acc += cplx<R>{lv.z, lv.w} * cplx<R>{v.z, v.w};
        }
        part[u] = acc;
    }
    __syncthreads();
    if (u < nb) {
        cplx<R> o = mm * in[u];
        o -= part[u];
        o -= part[nb + u];
        o -= part[2 * nb + u];
        out[u] = o;
    }
    __syncthreads();
}

// w_block_gmres for one 2x2-site Schwarz block run by the whole workgroup (a
// member of a grouped launch that owns one block of the colour): the block's
// links are staged in LDS once and serve every stencil of the block's
// GMRES(m, restarts), which the workgroup's threads share one (row, link
// block) partial sum each; a step's inner products are spread over the
// wavefronts, each computed as w_block_gmres computes it, and the column,
// the new vector and the scalars stay on wavefront 0 in the order of
// w_block_gmres, so both give the same bits.  Needs the FP32 engine's
// column-pair links and sap_wg[l].
template <typename R>
__device__ __noinline__ void g_block_gmres_2x2(const CArgs<R>& A, CK<R>& K, int l, int blk, const cplx<R>* rglob,
                                               cplx<R>* x) {
    static_assert(sizeof(R) == 4, "FP32 engine only");
    using V4 = float __attribute__((ext_vector_type(4)));
    const CLevel<R>& L = A.lv[l];
    const Geo& g = L.g;
    const int D = g.dof, nb = 4 * D, nu = 3 * nb, np = D / 2;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int m = A.sap_m;
    WaveScal& S = K.S.wv[0];
    unsigned long long* cnt = K.member == 0 ? A.cnt : nullptr;
    V4* Lk = (V4*)K.dyn;                      // [np][nu]
    cplx<R>* part = (cplx<R>*)(Lk + np * nu);  // [nu]
    cplx<R>* base = part + nu;                 // V_0 .. V_m, x_B, b_B
    auto V = [&](int j) { return base + (size_t)j * nb; };
    cplx<R>* xb = base + (size_t)(m + 1) * nb;
    cplx<R>* bb = base + (size_t)(m + 2) * nb;
    PhaseTimer T(K.prof);
    // stage the links; gather b_B; x_B = b_B
    {
        const int msz = D * D, ksz = g.ns * msz;
        const gptr<cplx<R>> links2 = as_global(L.links2);
        for (int i = tid; i < np * nu; i += CT) {
            const int p = i / nu, u = i - p * nu;
            const int q = u / nb, e = u - q * nb, mm = e / D, r = e - mm * D;
            const int n = sap_site(g, blk, mm);
            const int k = q == 0 ? 0 : q == 1 ? ((mm & 1) ? 3 : 1) : ((mm & 2) ? 4 : 2);
            Lk[p * nu + u] = ((const __attribute__((address_space(1))) V4*)(links2 + k * ksz + n * msz + 2 * r))[p * D];
        }
        for (int e = tid; e < nb; e += CT) {
            const int mm = e / D, d = e - mm * D;
            const cplx<R> v = rglob[(size_t)sap_site(g, blk, mm) * D + d];
            bb[e] = v;
            xb[e] = v;
        }
    }
    wg_dlocal_2x2<R>(A.mm, D, Lk, part, xb, V(1), cnt);
    if (wv == 0) {
        for (int e = lane; e < nb; e += 64) V(0)[e] = bb[e] - V(1)[e];
        wave_sync();
        const double nrm = sqrt(w_dot(bb, bb, nb).re);
        const double er = sqrt(w_dot(V(0), V(0), nb).re);
        if (lane == 0) {
            S.bsc[0] = nrm;
            S.bsc[1] = er;
        }
    }
    __syncthreads();
    T.lap(18);
    const double normb = S.bsc[0];
    double err = S.bsc[1];
    if (normb == 0.0) return;  // zero block right-hand side: x_B = 0
    int maxIt = m;
    for (int k = 0; k < A.sap_restarts; k++) {
        if (wv == 0) {
            const R ib = R(1.0 / err);
            for (int e = lane; e < nb; e += 64) V(0)[e] = ib * V(0)[e];
            if (lane == 0) S.g[0] = c64{err, 0.0};
        }
        for (int j = 0; j < m; j++) {
            T.lap(17);
            // after step j-1's closing barrier V(j) is complete
            wg_dlocal_2x2<R>(A.mm, D, Lk, part, V(j), V(j + 1), cnt, j > 0);
            T.lap(5);
            T.count(16);
            // the pass quantities over the wavefronts, then the column on wave 0
            for (int q = wv; q <= 2 * j + 1; q += CW) {
                const c64 v = w_gs_quantity<R>(q, j, V(j + 1), V, nb);
                if (lane == 0) S.dq[q] = v;
            }
            __syncthreads();
            T.lap(24);
            if (wv == 0) {
                const double hn = sqrt(w_gs_column(S, j));
                if (lane == 0) S.hn = hn;
            }
            __syncthreads();
            T.lap(25);
            // V_{j+1} from the column on wavefront 0 while wavefront 1 rotates a copy
            if (wv == 0) {
                const double hn = S.hn;
                for (int e = lane; e < nb; e += 64) w_gs_form<R>(S, V, j, hn, e);
            } else if (wv == 1 && lane == 0) {
                w_gs_rotate(S, j, S.hn, A.sap_gtol * normb);
            }
            __syncthreads();
            T.lap(6);
            if (S.brk) {
                maxIt = j + 1;
                break;
            }
        }
        if (wv == 0) {
            if (lane == 0) {
                for (int i = maxIt - 1; i >= 0; i--) {
                    c64 e = S.g[i];
                    for (int q = i + 1; q < maxIt; q++) e -= S.R[i][q] * S.eta[q];
                    S.eta[i] = cdiv(e, S.R[i][i]);
                }
            }
            wave_sync();
            for (int e = lane; e < nb; e += 64) {
                c64 s = cvt<double>(xb[e]);
                for (int j = 0; j < maxIt; j++) s += S.eta[j] * cvt<double>(V(j)[e]);
                xb[e] = cvt<R>(s);
            }
        }
        if (k + 1 == A.sap_restarts) break;  // the last cycle's residual decides nothing
        wg_dlocal_2x2<R>(A.mm, D, Lk, part, xb, V(1), cnt);
        if (wv == 0) {
            for (int e = lane; e < nb; e += 64) V(0)[e] = bb[e] - V(1)[e];
            wave_sync();
            const double er = sqrt(w_dot(V(0), V(0), nb).re);
            if (lane == 0) S.bsc[1] = er;
        }
        __syncthreads();
        err = S.bsc[1];
        if (err < A.sap_gtol * normb) break;
    }
    if (wv == 0) {
        wave_sync();
        for (int e = lane; e < nb; e += 64) {
            const int mm = e / D, d = e - mm * D;
            x[(size_t)sap_site(g, blk, mm) * D + d] += xb[e];
        }
    }
}

// SAP_C::SAP (src/sap.cpp:64-174): red then black blocks, residual after each colour.
template <typename R> __device__ __forceinline__ void c_sap(const CArgs<R>& A, CK<R>& K, int l, const cplx<R>* v, cplx<R>* x, int nu) {
    const CLevel<R>& L = A.lv[l];
    const int n = L.g.ns * L.g.dof;
    cplx<R>* r = K.W + L.sR;
    cplx<R>* Dx = K.W + L.sDx;
    const double vn = sqrt(c_dot(K.S, v, v, n).re);
    const int grp = A.grp, mem = K.member;
    // With one block per member (the workgroup form) a member reads the
    // residual only on the block it solves next: the residual before the
    // first red colour and those before black colours are formed on that
    // block alone; the one after a black colour stays whole for the sweep's
    // tolerance check.
    bool own = false;
    if constexpr (sizeof(R) == 4) own = A.sap_wg[l] != 0;
    auto resid = [&](const int* next) {
        if (own) {
            if (mem < L.g.ncol)
                c_resid_block(A, l, next[mem], v, x, Dx, r, (cplx<R>*)K.S.vin, VIN_BYTES / (int)sizeof(cplx<R>));
            else __syncthreads();
        } else {
            c_dop(A, l, x, Dx, (cplx<R>*)K.S.vin, VIN_BYTES / (int)sizeof(cplx<R>));
            for (int e = threadIdx.x; e < n; e += CT) r[e] = v[e] - Dx[e];
        }
    };
    resid(L.red);
    cplx<R>* xs = grp > 1 ? A.xchg + (size_t)K.rhs * A.xchg_n : nullptr;
    const int D = L.g.dof, nbs = L.g.sspb * D;
    for (int it = 0; it < nu; it++) {
        for (int colour = 0; colour < 2; colour++) {
            const int* list = colour == 0 ? L.red : L.black;
            __syncthreads();
            PhaseTimer T(K.prof);
            bool wg = false;
            if constexpr (sizeof(R) == 4) {
                // one block per member, solved by the whole workgroup
                if (A.sap_wg[l]) {
                    wg = true;
                    if (mem < L.g.ncol) {
                        g_block_gmres_2x2<R>(A, K, l, list[mem], r, x);
                        T.lap(22);
                        if (xs && threadIdx.x < 64) {
                            for (int e = threadIdx.x; e < nbs; e += 64) {
                                const int m = e / D, d = e - m * D;
                                const size_t i = (size_t)sap_site(L.g, list[mem], m) * D + d;
                                xs[i] = x[i];
                            }
                        }
                    }
                }
            }
            // block bi of the colour: member bi % grp, wavefront (bi / grp) % CW
            for (int bi = mem + grp * (threadIdx.x >> 6); !wg && bi < L.g.ncol; bi += grp * CW) {
                w_block_gmres(A, K, l, list[bi], r, x);
                if (xs) {  // publish the block's new x for the other members
                    for (int e = threadIdx.x & 63; e < nbs; e += 64) {
                        const int m = e / D, d = e - m * D;
                        const size_t i = (size_t)sap_site(L.g, list[bi], m) * D + d;
                        xs[i] = x[i];
                    }
                }
            }
            if (xs) {
                member_sync(K);
                // the blocks that the other members solved
                for (int e = threadIdx.x; e < L.g.ncol * nbs; e += CT) {
                    const int bi = e / nbs;
                    if (bi % grp == mem) continue;
                    const int q = e - bi * nbs, m = q / D, d = q - m * D;
                    const size_t i = (size_t)sap_site(L.g, list[bi], m) * D + d;
                    x[i] = xs[i];
                }
            }
            __syncthreads();
            T.lap(wg ? 23 : 20);
            // the residual after the last colour of the last sweep and its
            // tolerance check decide nothing: x is final
            if (colour == 1 && it + 1 == nu) return;
            if (colour == 0) {
                resid(L.black);
            } else {  // whole: the sweep's tolerance check reads it
                c_dop(A, l, x, Dx, (cplx<R>*)K.S.vin, VIN_BYTES / (int)sizeof(cplx<R>));
                for (int e = threadIdx.x; e < n; e += CT) r[e] = v[e] - Dx[e];
            }
            T.lap(21);
        }
        const double err = sqrt(c_dot(K.S, r, r, n).re);
        if (err < A.sap_tol * vn) return;
    }
}

// ---------------------------------------------------------------------------
// workgroup FGMRES (src/fgmres.cpp:5-88) and the K-cycle recursion, run as a
// state machine over per-level frames in LDS: FGMRES_k_cycle[l].fgmres ->
// k_cycle(l) -> FGMRES_k_cycle[l+1].fgmres -> ... -> the coarsest GMRES.  Each
// level body exists once (run-time level), so the kernel stays compact; a
// frame that needs a nested solve saves its loop state, pushes the child and
// is resumed when the child pops.  The order of operations within each solve
// is that of the recursive formulation.
// ---------------------------------------------------------------------------
enum { FR_FG = 0, FR_KC = 1, FR_CRS = 2 };

// thread 0 only; the run loop's barrier publishes the change
__device__ __forceinline__ void frame_push(CLds& S, int kind, int l) { S.stk[++S.sp] = kind * 16 + l; }
__device__ __forceinline__ void frame_pop(CLds& S) { S.sp--; }

// Restarted FGMRES from x0 = b (frame F, level l).  CRS: the coarsest GMRES
// (identity preconditioner, GMRES_level_l); otherwise the K-cycle FGMRES whose
// preconditioner is k_cycle(l) started from the previous cycle's Z_j
// (include/amg.h:37-39).  One-pass MGS with the Gram correction:
//   H_ij = <w, V_i> - sum_{k<i} H_kj <V_k, V_i>.
// Scalars (Hessenberg column, Givens, residual estimate) are computed by
// thread 0 in FP64 and broadcast through LDS.  Loop state that survives a
// nested solve (k, j, maxIt, |b|, |r|) lives in the frame.
template <typename R, int M, bool CRS>
__device__ __forceinline__ void fg_run(const CArgs<R>& A, CK<R>& K, int l, FgScal<M>& S, CFrame& F) {
    constexpr bool crs = CRS;
    const CLevel<R>& L = A.lv[l];
    const int n = L.g.ns * L.g.dof;
    const int m = crs ? A.crs_m : A.kc_m;
    const int restarts = crs ? A.crs_restarts : A.kc_restarts;
    const double tol = crs ? A.crs_tol : A.kc_tol;
    // the coarsest GMRES keeps its vectors in LDS when they fit
    const bool lds = crs && A.crs_lds;
    cplx<R>* Vb = lds ? K.dyn : K.W + (crs ? L.crV : L.kcV);
    cplx<R>* Zb = crs ? Vb : K.W + L.kcZ;
    cplx<R>* r = lds ? K.dyn + (size_t)(m + 1) * n : K.W + (crs ? L.crR : L.kcR);
    cplx<R>* w = lds ? K.dyn + (size_t)(m + 2) * n : K.W + (crs ? L.crW : L.kcW);
    cplx<R>* Dx = lds ? K.dyn + (size_t)(m + 3) * n : K.W + (crs ? L.crDx : L.kcDx);
    auto V = [&](int j) { return Vb + (size_t)j * n; };
    auto Z = [&](int j) { return Zb + (size_t)j * n; };
    const cplx<R>* b = (const cplx<R>*)F.in;
    cplx<R>* x = (cplx<R>*)F.out;
    int pc = F.pc, k = F.k, j = F.j, maxIt = F.maxIt;
    double nb = F.nb, err = F.err;
    __syncthreads();  // every thread has read the frame before thread 0 rewrites it
    for (;;) {
        if (pc == 0) {
            // setZeros (include/fgmres.h:93-120): Z_j start from zero in the first cycle
            if (!crs)
                for (int e = threadIdx.x; e < m * n; e += CT) Zb[e] = czero<R>();
            for (int e = threadIdx.x; e < n; e += CT) x[e] = b[e];
            c_dop(A, l, x, Dx, (cplx<R>*)K.S.vin, VIN_BYTES / (int)sizeof(cplx<R>));
            for (int e = threadIdx.x; e < n; e += CT) r[e] = b[e] - Dx[e];
            nb = sqrt(c_dot(K.S, b, b, n).re);
            err = sqrt(c_dot(K.S, r, r, n).re);
            if (err == 0.0) break;
            maxIt = m;
            k = 0;
            pc = 1;
        }
        if (pc == 1) {  // cycle k starts
            const R ib = R(1.0 / err);
            for (int e = threadIdx.x; e < n; e += CT) V(0)[e] = ib * r[e];
            if (threadIdx.x == 0) S.g[0] = c64{err, 0.0};
            j = 0;
            pc = 2;
        }
        if (pc == 2) {  // step j: the preconditioner, a nested solve
            if (!crs) {
                if (threadIdx.x == 0) {
                    F.pc = 3;
                    F.k = k;
                    F.j = j;
                    F.maxIt = maxIt;
                    F.nb = nb;
                    F.err = err;
                    CFrame& C = K.S.kcf[l];
                    C.in = V(j);
                    C.out = Z(j);
                    C.pc = 0;
                    frame_push(K.S, FR_KC, l);
                }
                return;
            }
            pc = 3;
        }
        if (pc == 3) {  // step j after the preconditioner
            cplx<R>* zj = Z(j);
            const int pb = crs ? 0 : 8 + 4 * (l - 1 > 2 ? 2 : l - 1);  // timer slots
            PhaseTimer T(K.prof);
            T.count(crs ? 7 : pb + 3);
            c_dop(A, l, zj, w, (cplx<R>*)K.S.vin, VIN_BYTES / (int)sizeof(cplx<R>));
            T.lap(pb + 0);
            // one pass: <w, V_i> (i <= j) and <V_j, V_k> (k < j); nq dots,
            // one wavefront per dot (lane partial sums in element order, then
            // the wave butterfly)
            const int nq = 2 * j + 1;
            for (int q = threadIdx.x >> 6; q < nq; q += CW) {
                const cplx<R>* a = q <= j ? w : V(j);
                const cplx<R>* bv = q <= j ? V(q) : V(q - j - 1);
                const c64 v = w_dot(a, bv, n);
                if ((threadIdx.x & 63) == 0) S.dq[q] = v;
            }
            __syncthreads();
            T.lap(pb + 1);
            if (threadIdx.x < 64) {
                // H_ij = <w,V_i> - sum_{k<i} H_kj conj(G_ik): lane i forms row i,
                // H_kj is final once the lanes below k are (m <= 64)
                const int lane = threadIdx.x;
                c64 d = {0.0, 0.0};
                if (lane <= j) d = S.dq[lane];
                for (int k = 0; k <= j; k++) {
                    const c64 hk = {readlane_d(d.re, k), readlane_d(d.im, k)};
                    if (lane > k && lane <= j) {
                        const c64 gik = lane == j ? S.dq[j + 1 + k] : S.G[lane][k];
                        d -= hk * conj(gik);
                    }
                }
                if (lane <= j) S.col[lane] = d;
                if (lane < j) S.G[j][lane] = S.dq[j + 1 + lane];
            }
            __syncthreads();
            T.lap(pb + 2);
            c64 nq1[1] = {{0, 0}};
            for (int e = threadIdx.x; e < n; e += CT) {
                c64 we = cvt<double>(w[e]);
                for (int i = 0; i <= j; i++) we -= S.col[i] * cvt<double>(V(i)[e]);
                const cplx<R> wr = cvt<R>(we);
                w[e] = wr;
                nq1[0] += mulc(cvt<double>(wr), cvt<double>(wr));
            }
            block_sum<1>(nq1, K.S.red);
            const double hn = sqrt(nq1[0].re);
            T.lap(pb + 3);
            if (hn > 0) {
                const R ih = R(1.0 / hn);
                for (int e = threadIdx.x; e < n; e += CT) V(j + 1)[e] = ih * w[e];
            }
            if (threadIdx.x == 0) {
                S.col[j + 1] = c64{hn, 0.0};
                c64 cnj, snj;
                givens_col(S.col, S.cn, S.sn, j, cnj, snj);
                for (int i = 0; i <= j; i++) S.R[i][j] = S.col[i];
                const c64 gj = S.g[j];
                const c64 gn = -(snj * gj);
                S.cn[j] = cnj;
                S.sn[j] = snj;
                S.g[j] = conj(cnj) * gj;
                S.g[j + 1] = gn;
                S.brk = sqrt(norm2(gn)) < tol * nb;
            }
            __syncthreads();
            T.lap(crs ? 4 : pb + 2);
            if (S.brk) {
                maxIt = j + 1;
                pc = 4;
            } else if (++j < m) {
                pc = 2;
                continue;
            } else {
                pc = 4;
            }
        }
        if (pc == 4) {  // end of cycle k: x += Z eta, true residual
            if (threadIdx.x == 0) {
                for (int i = maxIt - 1; i >= 0; i--) {
                    c64 e = S.g[i];
                    for (int q = i + 1; q < maxIt; q++) e -= S.R[i][q] * S.eta[q];
                    S.eta[i] = cdiv(e, S.R[i][i]);
                }
            }
            __syncthreads();
            for (int e = threadIdx.x; e < n; e += CT) {
                c64 s = cvt<double>(x[e]);
                for (int q = 0; q < maxIt; q++) s += S.eta[q] * cvt<double>(Z(q)[e]);
                x[e] = cvt<R>(s);
            }
            // after the last cycle the true residual decides nothing (only x
            // leaves the solve)
            if (k + 1 == restarts) break;
            c_dop(A, l, x, Dx, (cplx<R>*)K.S.vin, VIN_BYTES / (int)sizeof(cplx<R>));
            for (int e = threadIdx.x; e < n; e += CT) r[e] = b[e] - Dx[e];
            err = sqrt(c_dot(K.S, r, r, n).re);
            if (err < tol * nb) break;
            ++k;
            pc = 1;
        }
    }
    if (threadIdx.x == 0) frame_pop(K.S);
}

// k_cycle(l, eta, psi) (src/amg.cpp:135-178), or v_cycle(l, eta, psi)
// (src/amg.cpp:90-132) when A.vcycle: the coarse correction is then a
// recursive cycle from a zero guess instead of the level's K-cycle FGMRES.
// At the coarsest level the frame becomes the coarsest GMRES from x0 = eta.
template <typename R> __device__ __forceinline__ void kc_run(const CArgs<R>& A, CK<R>& K, int l) {
    CFrame& F = K.S.kcf[l];
    const cplx<R>* eta = (const cplx<R>*)F.in;
    cplx<R>* psi = (cplx<R>*)F.out;
    const int pc = F.pc;
    __syncthreads();  // every thread has read the frame before thread 0 rewrites it
    if (l == A.L - 1) {
        if (threadIdx.x == 0) {
            K.S.crsf.in = eta;
            K.S.crsf.out = psi;
            K.S.crsf.pc = 0;
            K.S.stk[K.S.sp] = FR_CRS * 16 + l;  // tail call
        }
        return;
    }
    const CLevel<R>& L = A.lv[l];
    const int n = L.g.ns * L.g.dof;
    cplx<R>* Dpsi = K.W + L.kDpsi;
    cplx<R>* rr = K.W + L.kR;
    cplx<R>* eta1 = K.W + L.kEta1;
    cplx<R>* psi1 = K.W + L.kPsi1;
    if (pc == 0) {
        if (A.nu1 > 0) c_sap(A, K, l, eta, psi, A.nu1);  // pre-smoothing
        // coarse-grid correction
        c_dop(A, l, psi, Dpsi, (cplx<R>*)K.S.vin, VIN_BYTES / (int)sizeof(cplx<R>));
        for (int e = threadIdx.x; e < n; e += CT) rr[e] = eta[e] - Dpsi[e];
        c_restrict(A, l, rr, eta1);
        if (A.vcycle) {
            const int nc = A.lv[l + 1].g.ns * A.lv[l + 1].g.dof;
            for (int e = threadIdx.x; e < nc; e += CT) psi1[e] = czero<R>();
        }
        if (threadIdx.x == 0) {
            F.pc = 1;
            CFrame& C = A.vcycle ? K.S.kcf[l + 1] : K.S.fg[l + 1];
            C.in = eta1;
            C.out = psi1;
            C.pc = 0;
            frame_push(K.S, A.vcycle ? FR_KC : FR_FG, l + 1);
        }
        return;
    }
    c_prolong(A, l, psi1, psi, true);
    if (A.nu2 > 0) c_sap(A, K, l, eta, psi, A.nu2);  // post-smoothing
    if (threadIdx.x == 0) frame_pop(K.S);
}

// Runs the frame stack until it is empty.
template <typename R> __device__ __forceinline__ void c_run(const CArgs<R>& A, CK<R>& K) {
    for (;;) {
        __syncthreads();
        const int sp = K.S.sp;
        if (sp < 0) break;
        const int top = K.S.stk[sp], kind = top >> 4, l = top & 15;
        if (kind == FR_FG) fg_run<R, KC_MAXM, false>(A, K, l, K.S.kc[l], K.S.fg[l]);
        else if (kind == FR_CRS) fg_run<R, CRS_MAXM, true>(A, K, l, K.S.crs, K.S.crsf);
        else kc_run<R>(A, K, l);
    }
}

// The arguments travel by value in the kernel-argument segment (no per-launch
// copy, so the launch can be captured in a HIP graph); each workgroup copies
// them to LDS, where they are read on demand (read in place through the
// segment pointer, the kernel kept many of them in scalar registers and
// spilled several hundred of those: 10 % slower per launch).
template <typename R> __device__ __forceinline__ void coarse_main(const CArgs<R>& A, CLds& S, char* dyn_lds) {
    const int rhs = blockIdx.x / A.grp, member = blockIdx.x - rhs * A.grp;
    if (A.rmask && !A.rmask[rhs]) return;  // the whole group skips a finished rhs
    CK<R> K{A, A.ws + (size_t)blockIdx.x * A.ws_per_rhs, S, reinterpret_cast<cplx<R>*>(dyn_lds),
            A.prof && member == 0 ? A.prof + (size_t)rhs * 32 : nullptr, rhs, member};
    PhaseTimer Ttot(K.prof);
    const cplx<R>* in = A.in + (size_t)rhs * A.in_stride;
    cplx<R>* out = A.out + (size_t)rhs * A.out_stride;
    // with several members each iterates on a private copy of the output (a
    // member ahead of another must not change what the other still reads);
    // member 0 stores the result
    cplx<R>* xo = A.grp > 1 ? K.W + A.x_top : out;
    if (threadIdx.x == 0) S.bar_epi = 0;
    const int l = A.level;
    switch (A.mode) {
        case CM_KC_SOLVE:  // FGMRES_k_cycle[1].fgmres(eta, eta, psi) (src/amg.cpp:161)
            if (threadIdx.x == 0) {
                S.fg[1].in = in;
                S.fg[1].out = xo;
                S.fg[1].pc = 0;
                S.sp = -1;
                frame_push(S, FR_FG, 1);
            }
            c_run(A, K);
            break;
        case CM_SAP: c_sap(A, K, l, in, out, A.nu); break;
        case CM_DOP: c_dop(A, l, in, out); break;
        case CM_RESTRICT: c_restrict(A, l, in, out); break;
        case CM_PROLONG: c_prolong(A, l, in, out, false); break;
        case CM_CYCLE: {  // the cycle at level l from psi = 0
            const int n = A.lv[l].g.ns * A.lv[l].g.dof;
            for (int e = threadIdx.x; e < n; e += CT) xo[e] = czero<R>();
            if (threadIdx.x == 0) {
                S.kcf[l].in = in;
                S.kcf[l].out = xo;
                S.kcf[l].pc = 0;
                S.sp = -1;
                frame_push(S, FR_KC, l);
            }
            c_run(A, K);
            break;
        }
    }
    if (xo != out && member == 0 && (A.mode == CM_KC_SOLVE || A.mode == CM_CYCLE)) {
        const int n = A.lv[l].g.ns * A.lv[l].g.dof;
        __syncthreads();
        for (int e = threadIdx.x; e < n; e += CT) out[e] = xo[e];
    }
    Ttot.lap(31);
}

// Several gauge configurations in one launch (right-hand side r belongs to
// configuration r, mmv non-null): the copy of the arguments gets the
// workgroup's configuration's coarse links, test vectors and mass term.
template <typename R> __global__ void __launch_bounds__(CT) k_coarse(const CArgs<R> Aval) {
    (void)Aval;
    const CArgs<R>& Ak = *(const CArgs<R>*)__builtin_amdgcn_kernarg_segment_ptr();
    __shared__ CLds S;
    __shared__ CArgs<R> Ac;
    extern __shared__ __align__(16) char dyn_lds[];
    static_assert(sizeof(CArgs<R>) % 4 == 0, "argument copy by words");
    for (int i = threadIdx.x; i < (int)(sizeof(CArgs<R>) / 4); i += CT)
        reinterpret_cast<int*>(&Ac)[i] = reinterpret_cast<const int*>(&Ak)[i];
    __syncthreads();
    if (threadIdx.x == 0 && Ak.mmv) {
        const int cfg = blockIdx.x / Ak.grp;
        for (int l = 1; l < MAXL; l++) {
            if (Ac.lv[l].links) Ac.lv[l].links += cfg * Ak.lcs[l];
            if (Ac.lv[l].links2) Ac.lv[l].links2 += cfg * Ak.lcs[l];
            if (Ac.lv[l].w) Ac.lv[l].w += cfg * Ak.wcs[l];
        }
        Ac.mm = Ak.mmv[cfg];
    }
    __syncthreads();
    coarse_main(Ac, S, dyn_lds);
}

}  // namespace amgs
