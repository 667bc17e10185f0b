// libamgs: context, device memory, orchestration and the C ABI (include/amgs.h).
//
// Execution model (one HIP device, one stream per context):
//   - outer FGMRES (FP64) on the fine lattice: kernels per Arnoldi step, scalar
//     work in small per-rhs kernels, one host synchronisation per step for the
//     break test (src/fgmres.cpp:5-88);
//   - preconditioner K-cycle(0): restriction, one persistent coarse kernel
//     (levels >= 1, whole nested recursion on the device), prolongation, SAP on
//     the fine lattice as a fixed sequence of predicated kernels;
//   - every right-hand side of a batch follows its own control flow exactly as
//     a single solve would; finished right-hand sides are skipped.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/amgs.h"
#include "common.hpp"
#include "engine.hpp"
#include "fine.hpp"
#include "host_util.hpp"
#include "krylov.hpp"
#include "mfma.hpp"
#include "outer.hpp"
#include "sap_block.hpp"
#include "sap_run.hpp"
#include "setup.hpp"

namespace amgs {

// ---------------------------------------------------------------------------
// geometry (include/params.h:30-65, include/level.h:147-155, include/sap.h:64-78)
// ---------------------------------------------------------------------------
static void build_geometry(Ctx& c) {
    const amgs_params& p = c.P;
    if (p.levels < 2 || p.levels > MAXL) throw ApiError(AMGS_ERR_GEOMETRY, "levels must be in [2, 5]");
    if (p.nx < 2 || p.nt < 2) throw ApiError(AMGS_ERR_GEOMETRY, "lattice too small");
    c.L = p.levels;
    for (int l = 0; l < c.L; l++) {
        Geo& g = c.geo[l];
        g.nx = l == 0 ? p.nx : p.bx[l - 1];
        g.nt = l == 0 ? p.nt : p.bt[l - 1];
        g.ns = g.nx * g.nt;
        g.colors = l == 0 ? 1 : p.ntest[l - 1];
        g.dof = 2 * g.colors;
        if (l < c.L - 1) {
            g.bx = p.bx[l];
            g.bt = p.bt[l];
            g.ntest = p.ntest[l];
            g.sbx = p.sapx[l];
            g.sbt = p.sapt[l];
            if (g.ntest < 1 || g.ntest > 24) throw ApiError(AMGS_ERR_GEOMETRY, "ntest must be in [1, 24]");
        } else {
            g.bx = g.bt = g.ntest = g.sbx = g.sbt = 1;
        }
        if (g.nx < 1 || g.nt < 1 || g.bx < 1 || g.bt < 1 || g.sbx < 1 || g.sbt < 1 || g.nx % g.bx || g.nt % g.bt ||
            g.nx % g.sbx || g.nt % g.sbt)
            throw ApiError(AMGS_ERR_GEOMETRY, "level " + std::to_string(l) + ": block counts must divide the lattice");
        g.nagg = 2 * g.bx * g.bt;
        g.xel = g.nx / g.bx;
        g.tel = g.nt / g.bt;
        g.spb = g.xel * g.tel;
        g.sxel = g.nx / g.sbx;
        g.stel = g.nt / g.sbt;
        g.sspb = g.sxel * g.stel;
        g.nsb = g.sbx * g.sbt;
        if (l == 0 && g.sspb > (1 << 20))
            throw ApiError(AMGS_ERR_GEOMETRY, "level-0 Schwarz blocks are limited to 2^20 sites");
        g.ncol = g.nsb != 1 ? g.nsb / 2 : 1;
        // red/black lists (src/sap.cpp:26-36): red iff (X + T) even; block b
        // goes to slot b / 2; a single block serves as both colours
        c.red[l].assign(g.ncol, 0);
        c.black[l].assign(g.ncol, 0);
        for (int X = 0; X < g.sbx; X++)
            for (int T = 0; T < g.sbt; T++) {
                const int blk = X * g.sbt + T;
                bool red = (g.sbt % 2 == 0) ? ((X % 2 == 0) ? blk % 2 == 0 : blk % 2 != 0) : blk % 2 == 0;
                if (blk / 2 < g.ncol) (red ? c.red[l] : c.black[l])[blk / 2] = blk;
            }
    }
    if (p.sap_m < 1 || p.sap_m > SAP_MAXM || p.kc_m < 1 || p.kc_m > KC_MAXM || p.crs_m < 1 || p.crs_m > CRS_MAXM ||
        p.out_m < 1 || p.out_m > OUT_MAXM)
        throw ApiError(AMGS_ERR_ARG, "restart lengths exceed the compiled limits (sap 8, kc 8, coarsest 24, outer 32)");
    if (p.sap_restarts < 1 || p.kc_restarts < 1 || p.crs_restarts < 1 || p.out_restarts < 1)
        throw ApiError(AMGS_ERR_ARG, "restart counts must be >= 1");
}

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
template <typename R> void Engine<R>::init(Ctx* ctx, int capacity) {
    c = ctx;
    cap = capacity;
    const Geo& g0 = c->geo[0];
    const size_t n0 = c->nvec(0);
    const int m = c->P.sap_m;
    const bool h16 = std::is_same<R, float>::value && c->h16();
    Vt.alloc(sizeof(cplx<R>) * (h16 ? 2 : m + 1) * cap * n0);
    if (h16) Vh.alloc(sizeof(f16x4) * (m + 1) * cap * g0.ns);
    xl.alloc(sizeof(cplx<R>) * cap * n0);
    r.alloc(sizeof(cplx<R>) * cap * n0);
    res0.alloc(sizeof(cplx<R>) * cap * n0);
    size_t nmax = n0;
    for (int l = 1; l < c->L; l++) nmax = std::max(nmax, c->nvec(l));
    zero.alloc(sizeof(cplx<R>) * cap * nmax);
    HIP_CHECK(hipMemsetAsync(zero.p, 0, zero.bytes, c->stream));
    nchunk_max = (int)cdiv_u(g0.sspb, FT);
    const size_t items = (size_t)g0.ncol * cap, M = m;
    // state layout
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    const size_t o_maxIt = take(4 * items), o_jb = take(4 * items), o_done = take(4 * items);
    const size_t o_nb = take(8 * items), o_s = take(8 * items * (M + 1));
    const size_t o_Hr = take(16 * items * (M + 1) * M), o_Ho = take(16 * items * (M + 1) * M);
    const size_t o_G = take(16 * items * (M + 1) * (M + 1));
    const size_t o_gp = take(16 * items * (M + 1)), o_gq = take(16 * items * (M + 1));
    const size_t o_cn = take(16 * items * M), o_sn = take(16 * items * M);
    const size_t o_qa = take(16 * items * nchunk_max * 2 * M), o_qb = take(8 * items * nchunk_max);
    const size_t o_qr = take(16 * items * nchunk_max);
    const size_t o_sig0 = take(8 * items);
    st_mem.alloc(off);
    char* b = st_mem.as<char>();
    st.maxIt = (int*)(b + o_maxIt);
    st.jbreak = (int*)(b + o_jb);
    st.done = (int*)(b + o_done);
    st.normb = (double*)(b + o_nb);
    st.s = (double*)(b + o_s);
    st.Hraw = (c64*)(b + o_Hr);
    st.Hrot = (c64*)(b + o_Ho);
    st.G = (c64*)(b + o_G);
    st.gpre = (c64*)(b + o_gp);
    st.gpost = (c64*)(b + o_gq);
    st.cn = (c64*)(b + o_cn);
    st.sn = (c64*)(b + o_sn);
    st.qa = (c64*)(b + o_qa);
    st.qb = (double*)(b + o_qb);
    st.qr = (double*)(b + o_qr);
    st.sig0 = (double*)(b + o_sig0);
    skip.alloc(sizeof(int) * cap);
    const unsigned ntile = cdiv_u(g0.ns, FT);
    part_r.alloc(sizeof(double) * cap * ntile);
    part_v.alloc(sizeof(double) * cap * ntile);
    rsig.alloc(sizeof(double) * cap);
    // K-cycle(0) and coarse workspace
    const size_t n1 = c->nvec(1);
    eta1.alloc(sizeof(cplx<R>) * cap * n1);
    psi1.alloc(sizeof(cplx<R>) * cap * n1);
    ca = CArgs<R>{};
    ca.L = c->L;
    ca.kc_m = c->P.kc_m;
    ca.kc_restarts = c->P.kc_restarts;
    ca.kc_tol = c->P.kc_tol;
    ca.crs_m = c->P.crs_m;
    ca.crs_restarts = c->P.crs_restarts;
    ca.crs_tol = c->P.crs_tol;
    ca.sap_m = c->P.sap_m;
    ca.sap_restarts = c->P.sap_restarts;
    ca.sap_gtol = c->P.sap_gmres_tol;
    ca.sap_tol = c->P.sap_tol;
    ca.nu1 = c->P.nu1;
    ca.nu2 = c->P.nu2;
    ca.vcycle = c->P.cycle == 0;
    size_t wo = 0;
    for (int l = 1; l < c->L; l++) {
        CLevel<R>& lv = ca.lv[l];
        const Geo& g = c->geo[l];
        const size_t n = c->nvec(l);
        lv.g = g;
        lv.red = c->redl[l].as<int>();
        lv.black = c->blackl[l].as<int>();
        lv.kcV = wo; wo += (ca.kc_m + 1) * n;
        lv.kcZ = wo; wo += ca.kc_m * n;
        lv.kcR = wo; wo += n;
        lv.kcW = wo; wo += n;
        lv.kcDx = wo; wo += n;
        if (l == c->L - 1) {
            lv.crV = wo; wo += (ca.crs_m + 1) * n;
            lv.crR = wo; wo += n;
            lv.crW = wo; wo += n;
            lv.crDx = wo; wo += n;
        } else {
            const size_t nc = c->nvec(l + 1);
            lv.kDpsi = wo; wo += n;
            lv.kR = wo; wo += n;
            lv.kEta1 = wo; wo += nc;
            lv.kPsi1 = wo; wo += nc;
            lv.sR = wo; wo += n;
            lv.sDx = wo; wo += n;
            lv.blkStride = (size_t)(ca.sap_m + 3) * g.sspb * g.dof;
            lv.blk = wo; wo += CW * lv.blkStride;
        }
    }
    // private copy of a grouped launch's output (any coarse level)
    size_t nco = 0;
    for (int l = 1; l < c->L; l++) nco = std::max(nco, c->nvec(l));
    ca.x_top = wo;
    wo += nco;
    ca.ws_per_rhs = (wo + 63) & ~size_t(63);
    const char* pe = std::getenv("AMGS_PROFILE");
    if (pe && pe[0] && pe[0] != '0') {
        prof.alloc(sizeof(unsigned long long) * 32 * cap);
        HIP_CHECK(hipMemsetAsync(prof.p, 0, prof.bytes, c->stream));
    }
    // grouped coarse launches: up to one member per Schwarz block of a colour
    // on the coarse levels with SAP smoothing (at most GRP_MAX)
    grp_max = 1;
    size_t nsap = 0;
    for (int l = 1; l < c->L - 1; l++) {
        grp_max = std::max(grp_max, c->geo[l].ncol);
        nsap = std::max(nsap, c->nvec(l));
    }
    grp_max = std::min(grp_max, GRP_MAX);
    cws.alloc(sizeof(cplx<R>) * ca.ws_per_rhs * cap * grp_max);
    ca.ws = cws.as<cplx<R>>();
    ca.grp = 1;
    ca.err = c->cerr.as<int>();
    if (grp_max > 1) {
        bar.alloc(sizeof(unsigned) * cap);
        HIP_CHECK(hipMemsetAsync(bar.p, 0, bar.bytes, c->stream));
        xchg.alloc(sizeof(cplx<R>) * cap * nsap);
        ca.bar = bar.as<unsigned>();
        ca.xchg = xchg.as<cplx<R>>();
        ca.xchg_n = nsap;
    }
}

template <typename R> void Engine<R>::restrict0(int nrhs, const cplx<R>* w, const cplx<R>* v, cplx<R>* out) {
    const Geo& g = c->geo[0];
    const Geo& g1 = c->geo[1];
    const dim3 grid(g1.ns, nrhs);
    const int nv = g.ntest;
    const size_t wcs = tvcs[0];
    if constexpr (std::is_same<R, float>::value) {
        if (nv <= 12 && wcs == 0) {
            const dim3 g2(g1.ns, cdiv_u(nrhs, 2));
            if (nv <= 8)
                hipLaunchKernelGGL(k_restrict0_g<8>, g2, dim3(FT), 0, c->stream, g, nv, w, v, out, g1.dof, g1.ns, nrhs,
                                   rmask);
            else
                hipLaunchKernelGGL(k_restrict0_g<12>, g2, dim3(FT), 0, c->stream, g, nv, w, v, out, g1.dof, g1.ns, nrhs,
                                   rmask);
            HIP_CHECK(hipGetLastError());
            return;
        }
    }
    if (nv <= 8)
        hipLaunchKernelGGL((k_restrict0<R, 8>), grid, dim3(FT), 0, c->stream, g, nv, w, v, out, g1.dof, g1.ns, wcs,
                           rmask);
    else if (nv <= 16)
        hipLaunchKernelGGL((k_restrict0<R, 16>), grid, dim3(FT), 0, c->stream, g, nv, w, v, out, g1.dof, g1.ns, wcs,
                           rmask);
    else
        hipLaunchKernelGGL((k_restrict0<R, 24>), grid, dim3(FT), 0, c->stream, g, nv, w, v, out, g1.dof, g1.ns, wcs,
                           rmask);
    HIP_CHECK(hipGetLastError());
}

// Chunks per SAP item (block x rhs): enough workgroups to fill the chip
// (about AMGS_SAP_WGS, default 2048), each chunk a multiple of FT sites.
static int sap_default_wgs() {
    static const int target = [] {
        const char* e = std::getenv("AMGS_SAP_WGS");
        const int v = e ? std::atoi(e) : 0;
        return v > 0 ? v : 2048;
    }();
    return target;
}
static int sap_chunks(const Geo& g, int nrhs, int target) {
    const int items = g.ncol * nrhs, maxc = (int)cdiv_u(g.sspb, FT);
    return std::max(1, std::min(maxc, (int)cdiv_u(target, items)));
}
// FP16 Krylov storage exists for the FP32 engine only
template <typename R> constexpr bool kH16 = std::is_same<R, float>::value;

// The level-0 SAP with one workgroup per item (k_sap_block): pairs per thread
// for this context's geometry and precision, 0 where the form does not apply
// (non-square lattice, blocks without site pairs along t or spanning a
// direction, a block tile beyond the LDS budget, fewer than 64 or an uneven
// split of the pairs over the threads).  AMGS_SAP_BLOCK=0 switches it off.
template <typename R> static int sap_block_pp(const Ctx& c) {
    static const bool off = [] {
        const char* e = std::getenv("AMGS_SAP_BLOCK");
        return e && e[0] == '0';
    }();
    const Geo& g = c.geo[0];
    if (off || g.nx != g.nt || g.stel % 2 || g.stel >= g.nt || g.sxel >= g.nx || g.sspb % 2 || g.nt % 2 ||
        c.P.sap_m > SAP_MAXM)
        return 0;
    const bool h16 = kH16<R> && c.h16();
    const size_t site = h16 ? sb_site_bytes<float, true>() : sb_site_bytes<R, false>();
    if ((size_t)g.sspb * site > SB_TILE_MAX) return 0;
    const int pairs = g.sspb / 2, ppmax = h16 ? 8 : (std::is_same<R, float>::value ? 4 : 2);
    // a thread's pairs lie whole rows apart: 2 pairs / pp sites a multiple of stel
    for (int pp = 1; pp <= ppmax; pp *= 2)
        if (pairs % pp == 0 && pairs / pp <= SBT && (pairs / pp) % 64 == 0 && (2 * pairs / pp) % g.stel == 0) return pp;
    return 0;
}

// The streaming level-0 SAP with one workgroup per item (k_sap_run): rows per
// thread (2, 4, 8 or 16, the fewest that keep a workgroup within SRT threads)
// for this context's geometry, 0 where the form does not apply (FP64 or FP32
// basis, non-square lattice, blocks spanning a direction, odd t extent, more
// than 64 pairs per block row or a row width not dividing 64, fewer than 64
// threads, a block GMRES restart length above SR_MAXM).  AMGS_SAP_RUN=0
// switches it off.
template <typename R> static int sap_run_rw(const Ctx& c) {
    static const bool off = [] {
        const char* e = std::getenv("AMGS_SAP_RUN");
        return e && e[0] == '0';
    }();
    const Geo& g = c.geo[0];
    if (off || !std::is_same<R, float>::value || !c.h16() || g.nx != g.nt || g.stel % 2 || g.stel >= g.nt ||
        g.sxel >= g.nx || c.P.sap_m > SR_MAXM)
        return 0;
    const int TL = g.stel / 2;
    if (TL > 64 || 64 % TL) return 0;
    for (int rw = 2; rw <= 16; rw *= 2) {
        if (g.sxel % rw) continue;
        const long long nth = (long long)TL * (g.sxel / rw);
        if (nth <= SRT && nth % 64 == 0) return rw;
    }
    return 0;
}

// workgroup target of the level-0 SAP launches for this batch size: the
// autotuned (amgs_autotune) or pinned (amgs_set_tuning) value if there is
// one, else the default.  Targets 1 and 2 are one workgroup per item: 1 the
// block-tile kernel (k_sap_block), 2 the streaming kernel (k_sap_run), each
// where it applies; larger targets split items into chunks.
template <typename R> int Engine<R>::sap_wgs(int nrhs) const {
    for (const auto& t : sap_tuned)
        if (t.first == nrhs) return t.second;
    // default: a one-workgroup-per-item form when its workgroups fill the chip
    // (at least 64 Ki threads) or when the chunked form could not split a
    // block either; the streaming form first
    const Geo& g = c->geo[0];
    const int rw = sap_run_rw<R>(*c);
    if (rw > 0) {
        const long long threads = (long long)g.ncol * nrhs * (g.stel / 2) * (g.sxel / rw);
        if (threads >= 65536 || g.sspb <= FT) return 2;
    }
    const int pp = sap_block_pp<R>(*c);
    if (pp > 0) {
        const long long threads = (long long)g.ncol * nrhs * (g.sspb / (2 * pp));
        if (threads >= 65536 || g.sspb <= FT) return 1;
    }
    return sap_default_wgs();
}

template <int RW>
static void sap_run_launch(const Ctx& c, int nrhs, const SapArgs<float>& a, c32* x, int* act, KTimer& kt) {
    const Geo& g = c.geo[0];
    const int nth = (g.stel / 2) * (g.sxel / RW);
    const size_t dyn = (size_t)nth * sr_lds_per_thread<RW>();
    static bool attr = false;
    if (!attr) {
        HIP_CHECK(hipFuncSetAttribute((const void*)k_sap_run<RW>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)(SRT * sr_lds_per_thread<RW>())));
        attr = true;
    }
    const dim3 grid(g.ncol, nrhs), block(nth);
    if (act)
        hipExtLaunchKernelGGL((k_sap_run<RW>), grid, block, dyn, c.stream, kt.start(), kt.stop(), 0, a, x,
                              c.P.sap_restarts, act);
    else
        hipLaunchKernelGGL((k_sap_run<RW>), grid, block, dyn, c.stream, a, x, c.P.sap_restarts, act);
}

// one colour of one sweep by k_sap_run (rw: sap_run_rw)
template <typename R>
static void sap_run_colour(const Ctx& c, int nrhs, int rw, const SapArgs<R>& a, cplx<R>* x, int* act, KTimer& kt) {
    if constexpr (std::is_same<R, float>::value) {
        switch (rw) {
            case 2: sap_run_launch<2>(c, nrhs, a, x, act, kt); return;
            case 4: sap_run_launch<4>(c, nrhs, a, x, act, kt); return;
            case 8: sap_run_launch<8>(c, nrhs, a, x, act, kt); return;
            case 16: sap_run_launch<16>(c, nrhs, a, x, act, kt); return;
        }
    }
    throw ApiError(AMGS_ERR_ARG, "SAP streaming kernel: no instance for this shape");
}

template <typename R, bool H, int PP>
static void sap_block_launch(const Ctx& c, int nrhs, const SapArgs<R>& a, cplx<R>* x, int* act, KTimer& kt) {
    const Geo& g = c.geo[0];
    const size_t dyn = (size_t)g.sspb * sb_site_bytes<R, H>();
    static bool attr = false;
    if (!attr) {
        HIP_CHECK(hipFuncSetAttribute((const void*)k_sap_block<R, H, PP>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)SB_TILE_MAX));
        attr = true;
    }
    const dim3 grid(g.ncol, nrhs), block(g.sspb / (2 * PP));
    if (act)
        hipExtLaunchKernelGGL((k_sap_block<R, H, PP>), grid, block, dyn, c.stream, kt.start(), kt.stop(), 0, a, x,
                              c.P.sap_restarts, act);
    else
        hipLaunchKernelGGL((k_sap_block<R, H, PP>), grid, block, dyn, c.stream, a, x, c.P.sap_restarts, act);
}

// one colour of one sweep by k_sap_block (pp: sap_block_pp)
template <typename R>
static void sap_block_colour(const Ctx& c, int nrhs, int pp, bool h16, const SapArgs<R>& a, cplx<R>* x, int* act,
                             KTimer& kt) {
    if constexpr (std::is_same<R, float>::value) {
        if (h16) {
            switch (pp) {
                case 1: sap_block_launch<float, true, 1>(c, nrhs, a, x, act, kt); return;
                case 2: sap_block_launch<float, true, 2>(c, nrhs, a, x, act, kt); return;
                case 4: sap_block_launch<float, true, 4>(c, nrhs, a, x, act, kt); return;
                case 8: sap_block_launch<float, true, 8>(c, nrhs, a, x, act, kt); return;
            }
        } else {
            switch (pp) {
                case 1: sap_block_launch<float, false, 1>(c, nrhs, a, x, act, kt); return;
                case 2: sap_block_launch<float, false, 2>(c, nrhs, a, x, act, kt); return;
                case 4: sap_block_launch<float, false, 4>(c, nrhs, a, x, act, kt); return;
            }
        }
    } else {
        switch (pp) {
            case 1: sap_block_launch<double, false, 1>(c, nrhs, a, x, act, kt); return;
            case 2: sap_block_launch<double, false, 2>(c, nrhs, a, x, act, kt); return;
        }
    }
    throw ApiError(AMGS_ERR_ARG, "SAP block kernel: no instance for this shape");
}

template <typename R, bool H, bool SQ>
static void sap_arnoldi_a(dim3 grid, hipStream_t s, const SapArgs<R>& a, int j, int k, int* act, KTimer& kt) {
    switch (j) {
#define AMGS_SAP_A(J)                                                                                          \
    case J:                                                                                                    \
        if (act)                                                                                               \
            hipExtLaunchKernelGGL((k_sap_arnoldi_a<R, J, H, SQ>), grid, dim3(FT), 0, s, kt.start(), kt.stop(), 0, \
                                  a, k, act);                                                                  \
        else                                                                                                   \
            hipLaunchKernelGGL((k_sap_arnoldi_a<R, J, H, SQ>), grid, dim3(FT), 0, s, a, k, act);               \
        break;
        AMGS_SAP_A(0) AMGS_SAP_A(1) AMGS_SAP_A(2) AMGS_SAP_A(3)
        AMGS_SAP_A(4) AMGS_SAP_A(5) AMGS_SAP_A(6) AMGS_SAP_A(7)
#undef AMGS_SAP_A
        default: throw ApiError(AMGS_ERR_ARG, "SAP restart length exceeds 8");
    }
}

// SAP_C::SAP at level 0 (src/sap.cpp:64-174) for nrhs right-hand sides.
template <typename R> void Engine<R>::sap0(int nrhs, const cplx<R>* v, cplx<R>* x, int nu) {
    if (nrhs > cap) throw ApiError(AMGS_ERR_ARG, "batch exceeds engine capacity");
    const Geo& g = c->geo[0];
    const R mm = R(c->P.m0 + 2.0);
    const unsigned ntile = cdiv_u(g.ns, FT);
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_skip_init, dim3(cdiv_u(nrhs, 64)), dim3(64), 0, s, nrhs, rmask, skip.as<int>());
    if (nu <= 0) return;
    hipLaunchKernelGGL(k_norm_part<R>, dim3(ntile, nrhs), dim3(FT), 0, s, g.ns, v, part_v.as<double>());
    // r is read only on the blocks of the colour about to be smoothed: the
    // residual before the red blocks is formed on the red blocks only
    const unsigned char* mred = c->colmask.as<unsigned char>();
    const unsigned char* mblack = mred + g.nsb;
    hipLaunchKernelGGL(k_resid<R>, dim3(ntile, nrhs), dim3(FT), 0, s, g, mm, U, v, x, r.as<cplx<R>>(),
                       (double*)nullptr, skip.as<int>(), mred, ucs, mmv);
    const int target = sap_wgs(nrhs), bpp = target == 1 ? sap_block_pp<R>(*c) : 0;
    const int rrw = target == 2 ? sap_run_rw<R>(*c) : 0;
    const int nchunk = sap_chunks(g, nrhs, target), cs = (int)cdiv_u(cdiv_u(g.sspb, nchunk), FT) * FT;
    SapArgs<R> a{};
    a.g = g;
    a.mm = mm;
    a.U = U;
    a.ucs = ucs;
    a.mmv = mmv;
    a.ncol = g.ncol;
    a.nchunk = (int)cdiv_u(g.sspb, cs);
    a.cs = cs;
    a.nrhs = nrhs;
    a.m = c->P.sap_m;
    a.tol = c->P.sap_gmres_tol;
    a.rglob = r.as<cplx<R>>();
    a.Vt = Vt.as<cplx<R>>();
    a.xl = xl.as<cplx<R>>();
    a.skip = skip.as<int>();
    a.st = st;
    const bool h16 = std::is_same<R, float>::value && c->h16();
    a.Vh = h16 ? Vh.as<f16x4>() : nullptr;
    a.rsig = rsig.as<double>();
    if (h16)
        hipLaunchKernelGGL(k_sap_rscale, dim3(nrhs), dim3(FT), 0, s, (const double*)part_v.as<double>(), (int)ntile,
                           2.0 * g.ns, rsig.as<double>());
    // power-of-two scale of the FP16 slots: a unit-norm block vector has
    // entries of order 1 / sqrt(2 sspb), stored as entries of order 1
    a.sig = h16 ? std::ldexp(1.0f, (int)std::lround(0.5 * std::log2(2.0 * g.sspb))) : 1.0f;
    a.rcp_stel = 1.0f / (float)g.stel;
    // sites in pairs along t in step A: FP16 storage, square lattice, blocks of
    // even t extent that span neither direction, chunks of even length
    a.pairs = h16 && g.nx == g.nt && g.stel % 2 == 0 && g.stel < g.nt && g.sxel < g.nx && a.cs % 2 == 0 &&
              g.sspb % 2 == 0 && g.nt % 2 == 0;
    // square lattice: block membership of a neighbour from local coordinates
    const bool sq = g.nx == g.nt;
    const dim3 grid(g.ncol * a.nchunk, nrhs);
    for (int it = 0; it < nu; it++) {
        for (int colour = 0; colour < 2; colour++) {
            a.blocks = (colour == 0 ? c->redl[0] : c->blackl[0]).as<int>();
            if (rrw > 0) {
                // the colour's whole block solve, one workgroup per item,
                // streaming over block rows; bytes counted by the kernel
                KTimer& kb = c->kt[KT_SAP_BLOCK];
                int* act = kb.begin(s, (double)g.ncol * g.sspb * 2 * sizeof(cplx<R>), (double)g.sspb);
                sap_run_colour<R>(*c, nrhs, rrw, a, x, act, kb);
            } else if (bpp > 0) {
                // the colour's whole block solve, one workgroup per item;
                // algorithmic bytes: links once, per item the bytes per block
                // site that the kernel counts
                KTimer& kb = c->kt[KT_SAP_BLOCK];
                int* act = kb.begin(s, (double)g.ncol * g.sspb * 2 * sizeof(cplx<R>), (double)g.sspb);
                sap_block_colour<R>(*c, nrhs, bpp, h16, a, x, act, kb);
            } else {
                if (h16) {
                    if (sq) hipLaunchKernelGGL((k_sap_init<R, kH16<R>, true>), grid, dim3(FT), 0, s, a);
                    else hipLaunchKernelGGL((k_sap_init<R, kH16<R>, false>), grid, dim3(FT), 0, s, a);
                } else {
                    if (sq) hipLaunchKernelGGL((k_sap_init<R, false, true>), grid, dim3(FT), 0, s, a);
                    else hipLaunchKernelGGL((k_sap_init<R, false, false>), grid, dim3(FT), 0, s, a);
                }
                for (int k = 0; k < c->P.sap_restarts; k++) {
                    for (int j = 0; j < a.m; j++) {
                        // algorithmic bytes: links of the colour's sites once; per active
                        // item Vt_j read, w written, Vt_i (i < j) read for the dots.
                        // With FP16 slots: w and every slot FP16.
                        const double S = sizeof(cplx<R>), sites = (double)g.ncol * g.sspb;
                        const double pa = h16 ? 16.0 + 8.0 * j : 2 * S * (2 + j);
                        KTimer& ka = c->kt[KT_SAP_A];
                        int* act = ka.begin(s, sites * 2 * S, g.sspb * pa, j);
                        if (h16) {
                            if (sq) sap_arnoldi_a<R, kH16<R>, true>(grid, s, a, j, k, act, ka);
                            else sap_arnoldi_a<R, kH16<R>, false>(grid, s, a, j, k, act, ka);
                        } else {
                            if (sq) sap_arnoldi_a<R, false, true>(grid, s, a, j, k, act, ka);
                            else sap_arnoldi_a<R, false, false>(grid, s, a, j, k, act, ka);
                        }
                        // w and Vt_i (i <= j) read, w written
                        const double pb = h16 ? 24.0 + 8.0 * j : 2 * S * (3 + j);
                        KTimer& kb = c->kt[KT_SAP_B];
                        act = kb.begin(s, 0.0, g.sspb * pb);
                        if (act) {
                            if (h16)
                                hipExtLaunchKernelGGL((k_sap_arnoldi_b<R, kH16<R>>), grid, dim3(FT), 0, s, kb.start(),
                                                      kb.stop(), 0, a, j, act);
                            else
                                hipExtLaunchKernelGGL((k_sap_arnoldi_b<R, false>), grid, dim3(FT), 0, s, kb.start(),
                                                      kb.stop(), 0, a, j, act);
                        } else if (h16) {
                            hipLaunchKernelGGL((k_sap_arnoldi_b<R, kH16<R>>), grid, dim3(FT), 0, s, a, j, act);
                        } else {
                            hipLaunchKernelGGL((k_sap_arnoldi_b<R, false>), grid, dim3(FT), 0, s, a, j, act);
                        }
                    }
                    const bool last = k + 1 == c->P.sap_restarts;
                    cplx<R>* xacc = last ? x : (cplx<R>*)nullptr;
                    if (h16) hipLaunchKernelGGL((k_sap_update<R, kH16<R>>), grid, dim3(FT), 0, s, a, xacc);
                    else hipLaunchKernelGGL((k_sap_update<R, false>), grid, dim3(FT), 0, s, a, xacc);
                    if (!last) {
                        if (h16) {
                            if (sq) hipLaunchKernelGGL((k_sap_resid<R, kH16<R>, true>), grid, dim3(FT), 0, s, a);
                            else hipLaunchKernelGGL((k_sap_resid<R, kH16<R>, false>), grid, dim3(FT), 0, s, a);
                        } else {
                            if (sq) hipLaunchKernelGGL((k_sap_resid<R, false, true>), grid, dim3(FT), 0, s, a);
                            else hipLaunchKernelGGL((k_sap_resid<R, false, false>), grid, dim3(FT), 0, s, a);
                        }
                    }
                }
            }
            // after red: r on the black blocks; after black: all of r and its
            // norm for the tolerance check, which decides nothing after the
            // last sweep and is left out there
            if (colour == 0)
                hipLaunchKernelGGL(k_resid<R>, dim3(ntile, nrhs), dim3(FT), 0, s, g, mm, U, v, x, r.as<cplx<R>>(),
                                   (double*)nullptr, skip.as<int>(), mblack, ucs, mmv);
            else if (it + 1 < nu)
                hipLaunchKernelGGL(k_resid<R>, dim3(ntile, nrhs), dim3(FT), 0, s, g, mm, U, v, x, r.as<cplx<R>>(),
                                   part_r.as<double>(), skip.as<int>(), (const unsigned char*)nullptr, ucs, mmv);
        }
        if (it + 1 < nu)
            hipLaunchKernelGGL(k_sap_check, dim3(nrhs), dim3(FT), 0, s, (const double*)part_r.as<double>(),
                               (const double*)part_v.as<double>(), (int)ntile, c->P.sap_tol, skip.as<int>());
    }
    HIP_CHECK(hipGetLastError());
}

// One application psi = K(eta) at level 0 from psi = 0 (include/amg.h:134-144,
// src/amg.cpp:135-178); with P.cycle == 0 the V-cycle (src/amg.cpp:90-132).
template <typename R> void Engine<R>::kcycle0(int nrhs, const cplx<R>* eta, cplx<R>* psi, bool from_zero) {
    const Geo& g = c->geo[0];
    const size_t n0 = c->nvec(0), n1 = c->nvec(1);
    hipStream_t s = c->stream;
    const unsigned ntile = cdiv_u(g.ns, FT);
    // from zero without pre-smoothing nothing reads psi before the
    // prolongation, which then writes it instead of adding to zero
    const bool direct = from_zero && c->P.nu1 <= 0;
    const int acc = direct ? 0 : 1;
    if (from_zero && !direct)
        hipLaunchKernelGGL(k_zero<R>, dim3(cdiv_u(nrhs * n0, 256)), dim3(256), 0, s, nrhs * n0, psi);
    const cplx<R>* rr = eta;  // r = eta - D psi = eta for psi = 0
    if (c->P.nu1 > 0 || !from_zero) {
        if (c->P.nu1 > 0) sap0(nrhs, eta, psi, c->P.nu1);
        hipLaunchKernelGGL(k_resid<R>, dim3(ntile, nrhs), dim3(FT), 0, s, g, R(c->P.m0 + 2.0), U, eta, psi,
                           res0.as<cplx<R>>(), (double*)nullptr, (const int*)nullptr, nullptr, ucs, mmv);
        rr = res0.as<cplx<R>>();
    }
    restrict0(nrhs, tv[0], rr, eta1.as<cplx<R>>());
    // K-cycle: FGMRES_k_cycle[1] (src/amg.cpp:161); V-cycle: v_cycle(1) from zero (src/amg.cpp:110)
    coarse(c->P.cycle == 0 ? CM_CYCLE : CM_KC_SOLVE, 1, nrhs, eta1.as<cplx<R>>(), n1, psi1.as<cplx<R>>(), n1);
    if (tvcs[0] != 0) {  // one test-vector set per configuration
        hipLaunchKernelGGL(k_prolong0<R>, dim3(ntile, nrhs), dim3(FT), 0, s, g, g.ntest, tv[0],
                           (const cplx<R>*)psi1.as<cplx<R>>(), psi, c->geo[1].dof, c->geo[1].ns, acc, tvcs[0], rmask);
    } else if constexpr (std::is_same<R, float>::value) {
        const dim3 gp(ntile, cdiv_u(nrhs, 8));
        const int nv = g.ntest, dofc = c->geo[1].dof, nsc = c->geo[1].ns;
        const c32* pc = (const c32*)psi1.as<c32>();
        if (nv <= 8)
            hipLaunchKernelGGL(k_prolong0_g<8>, gp, dim3(FT), 0, s, g, nv, tv[0], pc, psi, dofc, nsc, nrhs, acc, rmask);
        else if (nv <= 12)
            hipLaunchKernelGGL(k_prolong0_g<12>, gp, dim3(FT), 0, s, g, nv, tv[0], pc, psi, dofc, nsc, nrhs, acc, rmask);
        else if (nv <= 16)
            hipLaunchKernelGGL(k_prolong0_g<16>, gp, dim3(FT), 0, s, g, nv, tv[0], pc, psi, dofc, nsc, nrhs, acc, rmask);
        else
            hipLaunchKernelGGL(k_prolong0<R>, dim3(ntile, nrhs), dim3(FT), 0, s, g, g.ntest, tv[0],
                               (const cplx<R>*)psi1.as<cplx<R>>(), psi, dofc, nsc, acc, (size_t)0, rmask);
    } else {
        hipLaunchKernelGGL(k_prolong0<R>, dim3(ntile, nrhs), dim3(FT), 0, s, g, g.ntest, tv[0],
                           (const cplx<R>*)psi1.as<cplx<R>>(), psi, c->geo[1].dof, c->geo[1].ns, acc, (size_t)0, rmask);
    }
    if (c->P.nu2 > 0) sap0(nrhs, eta, psi, c->P.nu2);
    HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// hierarchy set-up (src/amg.cpp:3-46)
// ---------------------------------------------------------------------------
// multi: the configurations side by side (rhs r = configuration r); otherwise
// one configuration at the buffers' start
template <typename R>
static void bind_engine(Ctx& c, Engine<R>& e, DevBuf& U, DevBuf* tv, DevBuf* links, bool multi = false) {
    e.U = U.as<cplx<R>>();
    for (int l = 0; l < c.L; l++) {
        e.tv[l] = tv[l].as<cplx<R>>();
        e.links[l] = links[l].as<cplx<R>>();
        e.tvcs[l] = multi && l < c.L - 1 ? (size_t)c.geo[l].ntest * c.nvec(l) : 0;
        e.lkcs[l] = multi && l >= 1 ? (size_t)5 * c.geo[l].ns * c.geo[l].dof * c.geo[l].dof : 0;
    }
    e.ucs = multi ? c.nvec(0) : 0;
    if constexpr (std::is_same<R, double>::value) e.mmv = multi ? c.mm64.as<double>() : nullptr;
    else e.mmv = multi ? c.mm32.as<float>() : nullptr;
}

// The FP64 hierarchy buffers of configuration i (of a multi-configuration
// context) seen as a one-configuration context's for the set-up: the
// buffers' pointers and the mass are switched for the object's life.
struct CfgView {
    Ctx& c;
    void* U;
    void* tv[MAXL];
    void* lk[MAXL];
    double m0;
    CfgView(Ctx& cc, int i) : c(cc), U(cc.U64.p), m0(cc.P.m0) {
        c.U64.p = (char*)U + sizeof(c64) * c.nvec(0) * i;
        for (int l = 0; l < c.L; l++) {
            tv[l] = c.tv64[l].p;
            lk[l] = c.links64[l].p;
            if (l < c.L - 1) c.tv64[l].p = (char*)tv[l] + sizeof(c64) * c.geo[l].ntest * c.nvec(l) * i;
            if (l >= 1)
                c.links64[l].p = (char*)lk[l] + sizeof(c64) * 5 * c.geo[l].ns * c.geo[l].dof * c.geo[l].dof * i;
        }
        c.P.m0 = c.m0cfg[i];
    }
    ~CfgView() {
        c.U64.p = U;
        for (int l = 0; l < c.L; l++) {
            c.tv64[l].p = tv[l];
            c.links64[l].p = lk[l];
        }
        c.P.m0 = m0;
    }
};

// column-major links [5][ns][D][D] (entry (c, r) at c*D + r) to column-pair
// order [5][ns][D/2][D][2] (entry (c, r) at (c/2)*2D + 2r + c%2), D even
__global__ void k_links_pairs(size_t n, int D, const c32* in, c32* out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t msz = (size_t)D * D, site = i / msz;
    const int w = (int)(i - site * msz), c = w / D, r = w - c * D;
    out[site * msz + (size_t)(c / 2) * 2 * D + 2 * r + (c & 1)] = in[i];
}

static void convert64to32(Ctx& c, DevBuf& src, DevBuf& dst, size_t ncomplex) {
    dst.alloc(sizeof(c32) * ncomplex);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(cdiv_u(ncomplex, 256)), dim3(256), 0, c.stream, ncomplex,
                       src.as<c64>(), dst.as<c32>());
    HIP_CHECK(hipGetLastError());
}

static void smooth_level(Ctx& c, int l) {
    Engine<double>& e = c.e64;
    const Geo& g = c.geo[l];
    const size_t n = c.nvec(l);
    if (c.P.tv_sap_iters <= 0) return;
    if (l == 0) {
        e.sap0(g.ntest, e.zero.as<c64>(), c.tv64[0].as<c64>(), c.P.tv_sap_iters);
    } else {
        e.coarse(CM_SAP, l, g.ntest, e.zero.as<c64>(), n, c.tv64[l].as<c64>(), n, c.P.tv_sap_iters);
    }
}

// orthonormalise level l's test vectors and build level l+1's links
// (src/level.cpp:242-301, 373-417)
static void ortho_links(Ctx& c, int l) {
    const Geo& g = c.geo[l];
    hipLaunchKernelGGL(k_orthonormalize, dim3(g.nagg), dim3(ST), 0, c.stream, g, c.tv64[l].as<c64>());
    const Geo& gc = c.geo[l + 1];
    LinkSrc src{g, l == 0, c.U64.as<c64>(), l == 0 ? nullptr : c.links64[l].as<c64>()};
    hipLaunchKernelGGL(k_galerkin, dim3(gc.ns, 4), dim3(ST), 0, c.stream, src, (const c64*)c.tv64[l].as<c64>(), gc.ns,
                       gc, c.links64[l + 1].as<c64>());
    HIP_CHECK(hipGetLastError());
}

// ---- the coarse operator over a block of right-hand sides (MFMA) ----
// Level::D_operator (src/level.cpp:161-180) at level l >= 1 on nrhs vectors as
// matrix-core block-GEMMs (k_cdop_mfma) in the engine's precision.
template <typename R>
static void cdop_mfma(Ctx& c, Engine<R>& e, int l, int nrhs, const cplx<R>* in, cplx<R>* out, int wpg = 0) {
    const Geo& g = c.geo[l];
    if (wpg <= 0) wpg = e.cdop_wpg(l, nrhs);
    const size_t tasks = (size_t)g.ns * cdiv_u(nrhs, 16) * cdiv_u(g.dof, 16);
    hipLaunchKernelGGL(k_cdop_mfma<R>, dim3(cdiv_u(tasks, wpg)), dim3(64 * wpg), 0, c.stream, g, R(c.P.m0 + 2.0),
                       e.links[l], in, out, nrhs);
    HIP_CHECK(hipGetLastError());
}

// Bootstrap (src/amg.cpp:52-84) in FP64: per level, every test vector v
// becomes v + cycle_l(v - D_l v) from a zero guess with the current
// hierarchy (all test vectors of the level as one batch), then the level's
// interpolator and the next level's links are rebuilt.
static void bootstrap(Ctx& c) {
    Engine<double>& e = c.e64;
    DevBuf rhs, tvn;
    for (int it = 0; it < c.P.nit; it++)
        for (int l = 0; l < c.L - 1; l++) {
            const Geo& g = c.geo[l];
            const int nv = g.ntest;
            const size_t n = c.nvec(l), tot = (size_t)nv * n;
            rhs.alloc(sizeof(c64) * tot);
            tvn.alloc(sizeof(c64) * tot);
            c64* w = c.tv64[l].as<c64>();
            const unsigned nb = cdiv_u(tot, 256);
            if (l == 0) {
                hipLaunchKernelGGL(k_resid<double>, dim3(cdiv_u(g.ns, FT), nv), dim3(FT), 0, c.stream, g, c.P.m0 + 2.0,
                                   (const c64*)c.U64.as<c64>(), (const c64*)w, (const c64*)w, rhs.as<c64>(),
                                   (double*)nullptr, (const int*)nullptr, nullptr);
                e.kcycle0(nv, rhs.as<c64>(), tvn.as<c64>());
            } else {
                cdop_mfma(c, e, l, nv, (const c64*)w, rhs.as<c64>());
                hipLaunchKernelGGL(k_lin<double>, dim3(nb), dim3(256), 0, c.stream, tot, 1.0, (const c64*)w, -1.0,
                                   (const c64*)rhs.as<c64>(), rhs.as<c64>());
                e.coarse(CM_CYCLE, l, nv, rhs.as<c64>(), n, tvn.as<c64>(), n);
            }
            hipLaunchKernelGGL(k_lin<double>, dim3(nb), dim3(256), 0, c.stream, tot, 1.0, (const c64*)tvn.as<c64>(),
                               1.0, (const c64*)w, w);
            ortho_links(c, l);
        }
    c.sync();
}

// the FP64 part of the set-up of one configuration (after the raw test vectors)
static void setup_fp64(Ctx& c, bool smooth) {
    bind_engine(c, c.e64, c.U64, c.tv64, c.links64);
    for (int l = 0; l < c.L - 1; l++) {
        if (smooth) smooth_level(c, l);
        ortho_links(c, l);
    }
    if (c.P.nit > 0) bootstrap(c);
}

// FP32 copies of every configuration's hierarchy, engine bindings
static void setup_finish(Ctx& c) {
    const bool multi = c.ncfg > 1;
    const size_t nc = (size_t)c.ncfg;
    bind_engine(c, c.e64, c.U64, c.tv64, c.links64, multi);
    if (c.mixed()) {
        convert64to32(c, c.U64, c.U32, nc * c.nvec(0));
        for (int l = 0; l < c.L; l++) {
            if (l < c.L - 1) convert64to32(c, c.tv64[l], c.tv32[l], nc * c.geo[l].ntest * c.nvec(l));
            if (l >= 1)
                convert64to32(c, c.links64[l], c.links32[l], nc * 5 * c.geo[l].ns * c.geo[l].dof * c.geo[l].dof);
        }
        bind_engine(c, c.e32, c.U32, c.tv32, c.links32, multi);
        // levels with SAP smoothing: column-pair copies for the block solves
        for (int l = 1; l < c.L - 1; l++) {
            const Geo& g = c.geo[l];
            if (g.dof % 2) continue;
            const size_t nl = nc * 5 * g.ns * g.dof * g.dof;
            c.links2[l].alloc(sizeof(c32) * nl);
            hipLaunchKernelGGL(k_links_pairs, dim3(cdiv_u(nl, 256)), dim3(256), 0, c.stream, nl, g.dof,
                               (const c32*)c.links32[l].as<c32>(), c.links2[l].as<c32>());
            HIP_CHECK(hipGetLastError());
            c.e32.links2[l] = c.links2[l].as<c32>();
        }
    }
    c.sync();
    c.setup_done = true;
}

static void setup_common(Ctx& c, bool smooth) {
    c.clear_graphs();
    setup_fp64(c, smooth);
    setup_finish(c);
}

static void setup_random(Ctx& c, bool finish = true) {
    TvStream rng(c.P.tv_seed);
    std::vector<double> host;
    const Geo& g0 = c.geo[0];
    host.resize((size_t)2 * g0.ntest * c.nvec(0));
    rng.fill(host.data(), (size_t)g0.ntest * c.nvec(0));
    HIP_CHECK(hipMemcpyAsync(c.tv64[0].p, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    bind_engine(c, c.e64, c.U64, c.tv64, c.links64);
    // levels 1..L-2: P^dagger of the raw level-(l-1) vectors, then fresh draws
    for (int l = 1; l < c.L - 1; l++) {
        const Geo& g = c.geo[l];
        const Geo& gp = c.geo[l - 1];
        const size_t n = c.nvec(l), np = c.nvec(l - 1);
        const int nr = std::min(g.ntest, gp.ntest);
        if (l == 1)
            c.e64.restrict0(nr, c.tv64[0].as<c64>(), c.tv64[0].as<c64>(), c.tv64[1].as<c64>());
        else
            c.e64.coarse(CM_RESTRICT, l - 1, nr, c.tv64[l - 1].as<c64>(), np, c.tv64[l].as<c64>(), n);
        if (g.ntest > gp.ntest) {
            host.assign((size_t)2 * (g.ntest - gp.ntest) * n, 0.0);
            rng.fill(host.data(), (size_t)(g.ntest - gp.ntest) * n);
            HIP_CHECK(hipMemcpyAsync(c.tv64[l].as<c64>() + (size_t)gp.ntest * n, host.data(),
                                     host.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
        }
        c.sync();  // host buffer reuse
    }
    c.clear_graphs();
    setup_fp64(c, true);
    if (finish) setup_finish(c);
}

// every configuration of a multi-configuration context: each one's set-up is
// that of a context of its own (the same test-vector stream, its own field
// and mass), then the FP32 copies of all of them
static void setup_random_multi(Ctx& c) {
    for (int i = 0; i < c.ncfg; i++) {
        CfgView view(c, i);
        setup_random(c, false);
        c.sync();
    }
    setup_finish(c);
}

// ---------------------------------------------------------------------------
// outer FGMRES (FP64)
// ---------------------------------------------------------------------------
static void outer_alloc(Ctx& c, int cap) {
    const size_t n0 = c.nvec(0);
    const int M = c.P.out_m;
    c.ocap = cap;
    c.oV.alloc(sizeof(c64) * (M + 1) * cap * n0);
    // mixed modes: the multigrid preconditioner's FP32 outputs are kept as
    // they are (FP32 Z_j, exactly representable), not widened to FP64; the
    // FP64 basis is then allocated only by the solvers that use it
    // (solve_with: one SAP sweep or no preconditioner)
    if (c.mixed()) c.oZ32.alloc(sizeof(c32) * M * cap * n0);
    else c.oZ.alloc(sizeof(c64) * M * cap * n0);
    c.ox.alloc(sizeof(c64) * cap * n0);
    c.ob.alloc(sizeof(c64) * cap * n0);
    c.or_.alloc(sizeof(c64) * cap * n0);
    c.ow.alloc(sizeof(c64) * cap * n0);
    if (c.mixed()) {
        c.oeta32.alloc(sizeof(c32) * cap * n0);
        c.opsi32.alloc(sizeof(c32) * cap * n0);
    }
    const int ntile = (int)cdiv_u(c.geo[0].ns, FT);
    // per-tile partial sums of the final true residual check
    c.rel_part.alloc(sizeof(double) * 2 * (size_t)cap * ntile);
    const int hcap = M * c.P.out_restarts;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    const size_t o_i = take(4 * cap * 6), o_nb = take(8 * cap), o_s = take(8 * cap * (M + 1)),
                 o_h = take(8 * (size_t)cap * hcap), o_R = take(16 * (size_t)cap * M * M),
                 o_G = take(16 * (size_t)cap * (M + 1) * (M + 1)), o_g = take(16 * cap * (M + 1)),
                 o_cn = take(16 * cap * M), o_sn = take(16 * cap * M), o_eta = take(16 * cap * M),
                 o_qa = take(16 * (size_t)cap * ntile * (2 * M + 1)), o_qb = take(8 * (size_t)cap * ntile),
                 o_qr = take(16 * (size_t)cap * ntile);
    c.ost_mem.alloc(off);
    char* b = c.ost_mem.as<char>();
    OuterState& s = c.ost;
    int* ib = (int*)(b + o_i);
    s.maxIt = ib;
    s.jbreak = ib + cap;
    s.done = ib + 2 * cap;
    s.active = ib + 3 * cap;
    s.iters = ib + 4 * cap;
    s.nhist = ib + 5 * cap;
    s.normb = (double*)(b + o_nb);
    s.s = (double*)(b + o_s);
    s.hist = (double*)(b + o_h);
    s.R = (c64*)(b + o_R);
    s.G = (c64*)(b + o_G);
    s.g = (c64*)(b + o_g);
    s.cn = (c64*)(b + o_cn);
    s.sn = (c64*)(b + o_sn);
    s.eta = (c64*)(b + o_eta);
    s.qa = (c64*)(b + o_qa);
    s.qb = (double*)(b + o_qb);
    s.qr = (double*)(b + o_qr);
    s.M = M;
    s.ntile = ntile;
    s.ntile_a = (int)cdiv_u(ntile, OUT_SPT);
    s.hcap = hcap;
}

// Z_j = K(V_j) for all rhs of the batch.  eta32_ready: the FP32 staging
// input already holds V_j (written by the kernel that produced it); zj32: in
// the mixed modes, the FP32 Z_j that receives the result instead of zj.
static void outer_precond(Ctx& c, int nrhs, const c64* vj, c64* zj, bool eta32_ready, c32* zj32) {
    const size_t n0 = c.nvec(0), tot = nrhs * n0;
    if (c.outer_prec == 1) {
        // FGMRES_SAP::preconditioner (include/sap.h:212-224): zero, one SAP sweep
        hipLaunchKernelGGL(k_zero<double>, dim3(cdiv_u(tot, 256)), dim3(256), 0, c.stream, tot, zj);
        c.e64.sap0(nrhs, vj, zj, 1);
        return;
    }
    if (c.outer_prec == 2) {  // FGMRES_fine_level::preconditioner: identity (include/fgmres.h:144-147)
        HIP_CHECK(hipMemcpyAsync(zj, vj, sizeof(c64) * tot, hipMemcpyDeviceToDevice, c.stream));
        return;
    }
    if (c.mixed()) {
        if (!eta32_ready)
            hipLaunchKernelGGL((k_convert<float, double>), dim3(cdiv_u(tot, 256)), dim3(256), 0, c.stream, tot, vj,
                               c.oeta32.as<c32>());
        c.e32.kcycle0(nrhs, c.oeta32.as<c32>(), zj32 ? zj32 : c.opsi32.as<c32>());
        if (!zj32)
            hipLaunchKernelGGL((k_convert<double, float>), dim3(cdiv_u(tot, 256)), dim3(256), 0, c.stream, tot,
                               (const c32*)c.opsi32.as<c32>(), zj);
    } else {
        c.e64.kcycle0(nrhs, vj, zj);
    }
}

// Executable graph of one outer restart cycle for this batch size and outer
// preconditioner, captured from `body` on first use (every buffer of the
// outer solve is fixed for the context's life, and no kernel of the cycle
// takes the restart index).
template <typename F> static hipGraphExec_t cycle_graph(Ctx& c, int nrhs, F&& body) {
    for (auto& g : c.pgraphs)
        if (g.nrhs == nrhs && g.outer_prec == c.outer_prec) return g.exec;
    hipGraph_t graph = nullptr;
    HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    try {
        body();
    } catch (...) {
        if (hipStreamEndCapture(c.stream, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
        throw;
    }
    HIP_CHECK(hipStreamEndCapture(c.stream, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_CHECK(e);
    c.pgraphs.push_back({nrhs, c.outer_prec, exec});
    return exec;
}

// The outer solve's per-rhs active flags gate the K-cycle's work (SAP, the
// coarse kernel, restriction and prolongation skip a converged rhs) for the
// lifetime of the guard.
struct RhsMask {
    Ctx& c;
    RhsMask(Ctx& ctx, const int* m) : c(ctx) { c.e64.rmask = c.e32.rmask = m; }
    ~RhsMask() { c.e64.rmask = c.e32.rmask = nullptr; }
};

// The outer solve's per-rhs active flags on the host.  The copy goes to
// pinned memory and the host spins on an event behind it: this wait happens
// once per outer Arnoldi step, and a blocking wait (or a copy to pageable
// memory) adds the host's wake-up latency to every step.
static void read_flags(Ctx& c, const int* dflags, int nrhs, std::vector<int>& out) {
    if (!c.evf) HIP_CHECK(hipEventCreateWithFlags(&c.evf, hipEventDisableTiming));
    if (nrhs > c.hflags_cap) {
        if (c.hflags) HIP_CHECK(hipHostFree(c.hflags));
        c.hflags = nullptr;
        HIP_CHECK(hipHostMalloc((void**)&c.hflags, sizeof(int) * nrhs));
        c.hflags_cap = nrhs;
    }
    HIP_CHECK(hipMemcpyAsync(c.hflags, dflags, sizeof(int) * nrhs, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipEventRecord(c.evf, c.stream));
    hipError_t e;
    while ((e = hipEventQuery(c.evf)) == hipErrorNotReady) {
    }
    HIP_CHECK(e);
    std::copy(c.hflags, c.hflags + nrhs, out.begin());
}

// FGMRES_AMG::fgmres(b, 0, x) on device buffers ob -> ox (src/fgmres.cpp:5-88)
static void outer_solve(Ctx& c, int nrhs) {
    const Geo& g = c.geo[0];
    const size_t n0 = c.nvec(0), vstride = (size_t)c.ocap * n0;
    const int M = c.P.out_m;
    const double mm = c.P.m0 + 2.0, tol = c.P.out_tol;
    const unsigned ntile = cdiv_u(g.ns, FT);
    hipStream_t s = c.stream;
    OuterState& st = c.ost;
    // several configurations: rhs r's links and mass
    st.ucs = c.ncfg > 1 ? n0 : 0;
    st.mmv = c.ncfg > 1 ? c.mm64.as<double>() : nullptr;
    c64* V = c.oV.as<c64>();
    c64* Z = c.oZ.as<c64>();
    c64* x = c.ox.as<c64>();
    c64* b = c.ob.as<c64>();
    c64* r = c.or_.as<c64>();
    c64* w = c.ow.as<c64>();
    // the mixed multigrid preconditioner reads an FP32 copy of V_j: the
    // scaling kernel that writes V_j writes it too
    const bool eta32 = c.mixed() && c.outer_prec == 0;
    c32* oeta32 = eta32 ? c.oeta32.as<c32>() : nullptr;
    c32* Z32 = eta32 ? c.oZ32.as<c32>() : nullptr;
    HIP_CHECK(hipMemsetAsync(x, 0, sizeof(c64) * nrhs * n0, s));
    HIP_CHECK(hipMemsetAsync(st.done, 0, sizeof(int) * nrhs, s));
    HIP_CHECK(hipMemsetAsync(st.nhist, 0, sizeof(int) * nrhs, s));
    HIP_CHECK(hipMemsetAsync(st.iters, 0, sizeof(int) * nrhs, s));
    const dim3 g2(ntile, nrhs);
    std::vector<int> flags(nrhs);
    const char* ve = std::getenv("AMGS_VERBOSE");
    const bool verbose = ve && ve[0] && ve[0] != '0';
    // one line per restart cycle (long solves report their progress)
    const char* pe = std::getenv("AMGS_PROGRESS");
    const bool progress = pe && pe[0] && pe[0] != '0';
    const auto tstart = std::chrono::steady_clock::now();
    int step = 0;
    const RhsMask mask(c, st.active);
    // Arnoldi step j (src/fgmres.cpp:20-49): Z_j = K(V_j), w = D Z_j, its
    // Hessenberg column and Givens rotation, the break test, V_{j+1}.  Every
    // kernel skips an rhs whose flag the break test cleared.
    auto arnoldi_step = [&](int j) {
        outer_precond(c, nrhs, V + (size_t)j * vstride, Z + (size_t)j * vstride, eta32,
                      Z32 ? Z32 + (size_t)j * vstride : nullptr);
        if (Z32)
            hipLaunchKernelGGL(k_outer_a<float>, dim3(st.ntile_a, nrhs), dim3(FT), 0, s, g, mm,
                               (const c64*)c.U64.as<c64>(), (const c32*)(Z32 + (size_t)j * vstride), (const c64*)V, w,
                               st, j, vstride);
        else
            hipLaunchKernelGGL(k_outer_a<double>, dim3(st.ntile_a, nrhs), dim3(FT), 0, s, g, mm,
                               (const c64*)c.U64.as<c64>(), (const c64*)(Z + (size_t)j * vstride), (const c64*)V, w,
                               st, j, vstride);
        hipLaunchKernelGGL(k_outer_hcol, dim3(nrhs), dim3(FT), 0, s, st, j);
        hipLaunchKernelGGL(k_outer_b, g2, dim3(FT), 0, s, g.ns, (const c64*)V, w, st, j, vstride);
        hipLaunchKernelGGL(k_outer_fin, dim3(nrhs), dim3(FT), 0, s, st, j, tol);
        if (j + 1 < M)
            hipLaunchKernelGGL(k_outer_scale, g2, dim3(FT), 0, s, g.ns, (const c64*)w, V + (size_t)(j + 1) * vstride,
                               st, j + 1, eta32 ? oeta32 : nullptr);
    };
    auto cycle_start = [&]() {
        hipLaunchKernelGGL(k_outer_scale, g2, dim3(FT), 0, s, g.ns, (const c64*)r, V, st, 0, eta32 ? oeta32 : nullptr);
    };
    // end of a cycle: y from the triangular system, x += Z y, r = b - D x
    // (src/fgmres.cpp:52-59, 61-65)
    auto cycle_end = [&]() {
        hipLaunchKernelGGL(k_outer_backsolve, dim3(cdiv_u(nrhs, 64)), dim3(64), 0, s, st, nrhs);
        if (Z32) hipLaunchKernelGGL(k_outer_update<float>, g2, dim3(FT), 0, s, g.ns, (const c32*)Z32, x, st, vstride);
        else hipLaunchKernelGGL(k_outer_update<double>, g2, dim3(FT), 0, s, g.ns, (const c64*)Z, x, st, vstride);
        hipLaunchKernelGGL(k_outer_resid, g2, dim3(FT), 0, s, g, mm, (const c64*)c.U64.as<c64>(), (const c64*)b,
                           (const c64*)x, r, st);
    };
    // graph mode: a whole restart cycle is one graph launch and the host
    // reads the flags once per cycle; a batch that converges inside a cycle
    // runs its remaining steps as launches that skip every rhs.  Eager
    // launches (and the per-launch kernel timers, which resolve after every
    // step) read them after every step and leave the cycle early.
    const bool timers = std::any_of(std::begin(c.kt), std::end(c.kt), [](const KTimer& kt) { return kt.on; });
    const bool whole_cycle = c.graph && !timers && !verbose;
    hipLaunchKernelGGL(k_outer_resid, g2, dim3(FT), 0, s, g, mm, (const c64*)c.U64.as<c64>(), (const c64*)b,
                       (const c64*)x, r, st);
    for (int k = 0; k < c.P.out_restarts; k++) {
        hipLaunchKernelGGL(k_outer_start, dim3(nrhs), dim3(FT), 0, s, st, k, tol, M);
        read_flags(c, st.active, nrhs, flags);
        if (progress && k > 0) {
            int act = 0;
            for (int f : flags) act += f != 0;
            const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tstart).count();
            std::fprintf(stderr, "[amgs] %dx%d: restart cycle %d done, %d of %d rhs still active, %.0f ms\n", g.nx, g.nt, k,
                         act, nrhs, el);
            std::fflush(stderr);
        }
        if (std::none_of(flags.begin(), flags.end(), [](int f) { return f != 0; })) break;
        if (whole_cycle) {
            HIP_CHECK(hipGraphLaunch(cycle_graph(c, nrhs,
                                                 [&]() {
                                                     cycle_start();
                                                     for (int j = 0; j < M; j++) arnoldi_step(j);
                                                     cycle_end();
                                                 }),
                                     s));
            step += M;
        } else {
            cycle_start();
            for (int j = 0; j < M; j++) {
                arnoldi_step(j);
                read_flags(c, st.active, nrhs, flags);
                for (KTimer& kt : c.kt) kt.resolve(s);
                step++;
                if (verbose) {
                    int nh = 0;
                    double h = 0, nb = 0;
                    HIP_CHECK(hipMemcpy(&nh, st.nhist, sizeof(int), hipMemcpyDeviceToHost));
                    if (nh > 0 && nh <= st.hcap)
                        HIP_CHECK(hipMemcpy(&h, st.hist + nh - 1, sizeof(double), hipMemcpyDeviceToHost));
                    HIP_CHECK(hipMemcpy(&nb, st.normb, sizeof(double), hipMemcpyDeviceToHost));
                    const double el =
                        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tstart).count();
                    int act = 0;
                    for (int f : flags) act += f != 0;
                    std::fprintf(stderr, "[amgs] cycle %d step %d (total %d): |g|/|b| rhs0 = %.3e, active %d/%d, %.1f ms\n",
                                 k, j, step, nb > 0 ? h / nb : 0.0, act, nrhs, el);
                    std::fflush(stderr);
                }
                if (std::none_of(flags.begin(), flags.end(), [](int f) { return f != 0; })) break;
            }
            cycle_end();
        }
        if (k + 1 == c.P.out_restarts)
            hipLaunchKernelGGL(k_outer_start, dim3(nrhs), dim3(FT), 0, s, st, k + 1, tol, M);
    }
    hipLaunchKernelGGL(k_outer_final, dim3(cdiv_u(nrhs, 64)), dim3(64), 0, s, st, nrhs, c.P.out_restarts, M);
    HIP_CHECK(hipGetLastError());
    auto dump = [&](DevBuf& pb, const char* tag) {
        if (!pb.p) return;
        std::vector<unsigned long long> h(32);
        c.sync();
        HIP_CHECK(hipMemcpy(h.data(), pb.p, 32 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        std::fprintf(stderr, "[amgs profile %s, rhs 0, ms at 100 MHz]", tag);
        for (int i = 0; i < 32; i++)
            if (h[i]) std::fprintf(stderr, " %d:%.3f", i, (i == 7 || i == 11 || i == 15 || i == 16 || i == 19) ? (double)h[i] : h[i] * 1e-5);
        std::fprintf(stderr, "\n");
    };
    dump(c.e64.prof, "fp64");
    dump(c.e32.prof, "fp32");
}

// true relative residual |b - D x| / |b| in FP64
static void true_relres(Ctx& c, int nrhs, double* out) {
    const Geo& g = c.geo[0];
    const size_t n0 = c.nvec(0);
    const unsigned ntile = cdiv_u(g.ns, FT);
    double* pr = c.rel_part.as<double>();
    double* pb = pr + (size_t)nrhs * ntile;
    hipLaunchKernelGGL(k_resid<double>, dim3(ntile, nrhs), dim3(FT), 0, c.stream, g, c.P.m0 + 2.0,
                       (const c64*)c.U64.as<c64>(), (const c64*)c.ob.as<c64>(), (const c64*)c.ox.as<c64>(),
                       c.ow.as<c64>(), pr, (const int*)nullptr, nullptr, c.ncfg > 1 ? c.nvec(0) : 0,
                       c.ncfg > 1 ? c.mm64.as<double>() : nullptr);
    hipLaunchKernelGGL(k_norm_part<double>, dim3(ntile, nrhs), dim3(FT), 0, c.stream, g.ns,
                       (const c64*)c.ob.as<c64>(), pb);
    std::vector<double> hr((size_t)nrhs * ntile), hb((size_t)nrhs * ntile);
    HIP_CHECK(hipMemcpyAsync(hr.data(), pr, hr.size() * 8, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(hb.data(), pb, hb.size() * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    (void)n0;
    for (int r = 0; r < nrhs; r++) {
        double a = 0, bb = 0;
        for (unsigned t = 0; t < ntile; t++) {
            a += hr[(size_t)r * ntile + t];
            bb += hb[(size_t)r * ntile + t];
        }
        out[r] = bb > 0 ? std::sqrt(a / bb) : 0.0;
    }
}

}  // namespace amgs

// ===========================================================================
// C ABI
// ===========================================================================
using namespace amgs;

// ---- static workspace of amgs_sap_dev_noalloc ----
// A file-scope device array that every GPU the library runs on holds until the
// process ends.  Host code never passes the array's name to a kernel (on the
// host it is only a stand-in symbol): amgs_create looks up its device address
// once per device with hipGetSymbolAddress and the entry point uses that
// address.  Every call on one GPU shares the array; calls that overlap on one
// GPU race.
static __device__ __align__(256) char scratch[AMGS_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // per device, null until looked up

static void scratch_setup(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = ptr;
}

static void* scratch_ptr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

struct amgs_ctx {
    Ctx c;
};

#define API_BEGIN try {
#define API_END                                             \
    }                                                       \
    catch (const ApiError& e) {                             \
        ctx->c.err = e.what();                              \
        return e.code;                                      \
    }                                                       \
    catch (const HipError& e) {                             \
        ctx->c.err = e.what();                              \
        return AMGS_ERR_HIP;                                \
    }                                                       \
    catch (const std::exception& e) {                       \
        ctx->c.err = e.what();                              \
        return AMGS_ERR_STATE;                              \
    }

static void need(bool cond, int code, const char* msg) {
    if (!cond) throw ApiError(code, msg);
}

// entry points that serve one gauge configuration per context
static void single_config(const Ctx& c) {
    need(c.ncfg == 1, AMGS_ERR_STATE, "not available on a multi-configuration context");
}

static void h2d(Ctx& c, DevBuf& dst, const double* src, size_t ncomplex) {
    dst.alloc(ncomplex * sizeof(c64));
    HIP_CHECK(hipMemcpyAsync(dst.p, src, ncomplex * sizeof(c64), hipMemcpyHostToDevice, c.stream));
}

static int create_impl(const amgs_params* p, int ncfg, const double* m0, amgs_ctx** out) {
    if (!p || !out || ncfg < 1 || (ncfg > 1 && !m0)) return AMGS_ERR_ARG;
    *out = nullptr;
    auto* ctx = new amgs_ctx();
    try {
        Ctx& c = ctx->c;
        c.P = *p;
        if (c.P.max_rhs < 1) c.P.max_rhs = 1;
        c.ncfg = ncfg;
        if (ncfg > 1) {
            c.P.max_rhs = ncfg;  // one right-hand side per configuration
            c.m0cfg.assign(m0, m0 + ncfg);
            for (double v : c.m0cfg)
                if (!std::isfinite(v)) throw ApiError(AMGS_ERR_NAN, "bare mass holds NaN/Inf");
        }
        if (c.P.precision != AMGS_PREC_FP64 && c.P.precision != AMGS_PREC_MIXED && c.P.precision != AMGS_PREC_MIXED16)
            throw ApiError(AMGS_ERR_ARG, "precision must be AMGS_PREC_FP64, AMGS_PREC_MIXED or AMGS_PREC_MIXED16");
        build_geometry(c);
        HIP_CHECK(hipGetDevice(&c.device));
        scratch_setup(c.device);  // with c.device current
        HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&c.ev0));
        HIP_CHECK(hipEventCreate(&c.ev1));
        const size_t n0 = c.nvec(0), nc = (size_t)ncfg;
        c.U64.alloc(sizeof(c64) * n0 * nc);
        if (ncfg > 1) {
            std::vector<double> m64(ncfg);
            std::vector<float> m32(ncfg);
            for (int i = 0; i < ncfg; i++) {
                m64[i] = c.m0cfg[i] + 2.0;
                m32[i] = float(c.m0cfg[i] + 2.0);
            }
            c.mm64.alloc(sizeof(double) * ncfg);
            c.mm32.alloc(sizeof(float) * ncfg);
            HIP_CHECK(hipMemcpy(c.mm64.p, m64.data(), sizeof(double) * ncfg, hipMemcpyHostToDevice));
            HIP_CHECK(hipMemcpy(c.mm32.p, m32.data(), sizeof(float) * ncfg, hipMemcpyHostToDevice));
        }
        c.cerr.alloc(sizeof(int));
        HIP_CHECK(hipMemsetAsync(c.cerr.p, 0, sizeof(int), c.stream));
        {
            // level-0 Schwarz block flags: [0][blk] red, [1][blk] black
            const Geo& g0 = c.geo[0];
            std::vector<unsigned char> m(2 * (size_t)g0.nsb, 0);
            for (int b : c.red[0]) m[b] = 1;
            for (int b : c.black[0]) m[g0.nsb + b] = 1;
            c.colmask.alloc(m.size());
            HIP_CHECK(hipMemcpy(c.colmask.p, m.data(), m.size(), hipMemcpyHostToDevice));
        }
        for (int l = 0; l < c.L; l++) {
            c.redl[l].alloc(sizeof(int) * c.red[l].size());
            c.blackl[l].alloc(sizeof(int) * c.black[l].size());
            HIP_CHECK(hipMemcpyAsync(c.redl[l].p, c.red[l].data(), sizeof(int) * c.red[l].size(), hipMemcpyHostToDevice, c.stream));
            HIP_CHECK(hipMemcpyAsync(c.blackl[l].p, c.black[l].data(), sizeof(int) * c.black[l].size(),
                                     hipMemcpyHostToDevice, c.stream));
            if (l < c.L - 1) c.tv64[l].alloc(sizeof(c64) * c.geo[l].ntest * c.nvec(l) * nc);
            if (l >= 1) c.links64[l].alloc(sizeof(c64) * 5 * c.geo[l].ns * c.geo[l].dof * c.geo[l].dof * nc);
        }
        int ntest_max = 1;
        for (int l = 0; l < c.L - 1; l++) ntest_max = std::max(ntest_max, c.geo[l].ntest);
        const bool mixed = c.mixed();
        c.e64.init(&c, mixed ? ntest_max : std::max(ntest_max, c.P.max_rhs));
        if (mixed) c.e32.init(&c, c.P.max_rhs);
        outer_alloc(c, c.P.max_rhs);
        c.sync();
    } catch (const ApiError& e) {
        std::fprintf(stderr, "amgs_create: %s\n", e.what());
        int code = e.code;
        delete ctx;
        return code;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "amgs_create: %s\n", e.what());
        delete ctx;
        return AMGS_ERR_HIP;
    }
    *out = ctx;
    return AMGS_OK;
}

extern "C" int amgs_create(const amgs_params* p, amgs_ctx** out) { return create_impl(p, 1, nullptr, out); }

extern "C" int amgs_create_multi(const amgs_params* p, int ncfg, const double* m0, amgs_ctx** out) {
    return create_impl(p, ncfg, m0, out);
}

extern "C" int amgs_num_configs(const amgs_ctx* ctx) { return ctx ? ctx->c.ncfg : 0; }

extern "C" void amgs_destroy(amgs_ctx* ctx) {
    if (!ctx) return;
    if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
    Ctx& c = ctx->c;
    hipStream_t s = c.stream;
    hipEvent_t a = c.ev0, b = c.ev1, f = c.evf;
    int* hf = c.hflags;
    for (int k = 0; k < 2; k++) {
        if (c.stage_ev[k]) (void)hipEventDestroy(c.stage_ev[k]);
        if (c.stage[k]) (void)hipHostFree(c.stage[k]);
    }
    c.clear_graphs();
    delete ctx;  // frees device buffers
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (f) (void)hipEventDestroy(f);
    if (hf) (void)hipHostFree(hf);
    if (s) (void)hipStreamDestroy(s);
}

extern "C" const char* amgs_last_error(const amgs_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

extern "C" int amgs_set_gauge(amgs_ctx* ctx, const double* U) {
    if (!ctx || !U) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    const size_t n = c.nvec(0) * c.ncfg;  // [ncfg][ns][2]
    for (size_t i = 0; i < 2 * n; i++) need(std::isfinite(U[i]), AMGS_ERR_NAN, "gauge field holds NaN/Inf");
    h2d(c, c.U64, U, n);
    c.sync();
    c.gauge_set = true;
    c.setup_done = false;
    return AMGS_OK;
    API_END
}

extern "C" int amgs_setup(amgs_ctx* ctx) {
    if (!ctx) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    need(c.gauge_set, AMGS_ERR_STATE, "amgs_set_gauge first");
    auto t0 = std::chrono::steady_clock::now();
    if (c.ncfg > 1) setup_random_multi(c);
    else setup_random(c);
    c.last_ms[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return AMGS_OK;
    API_END
}

extern "C" int amgs_setup_from_tv(amgs_ctx* ctx, const double* const* tv, int smooth) {
    if (!ctx || !tv) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    need(c.gauge_set, AMGS_ERR_STATE, "amgs_set_gauge first");
    single_config(c);
    for (int l = 0; l < c.L - 1; l++) {
        need(tv[l] != nullptr, AMGS_ERR_ARG, "missing test vectors");
        HIP_CHECK(hipMemcpyAsync(c.tv64[l].p, tv[l], sizeof(c64) * c.geo[l].ntest * c.nvec(l), hipMemcpyHostToDevice,
                                 c.stream));
    }
    setup_common(c, smooth != 0);
    return AMGS_OK;
    API_END
}

// A member barrier of a grouped coarse launch that waited beyond its bound
// (the members of a right-hand side did not all run): the results are void.
static void check_coarse_error(Ctx& c) {
    int e = 0;
    HIP_CHECK(hipMemcpy(&e, c.cerr.p, sizeof(int), hipMemcpyDeviceToHost));
    if (e) {
        HIP_CHECK(hipMemset(c.cerr.p, 0, sizeof(int)));
        throw ApiError(AMGS_ERR_HIP, "coarse-level solve: a member barrier timed out (results discarded)");
    }
}

static int solve_impl(Ctx& c, int nrhs, int* iters, double* relres, double* hist, int hist_cap, int* nhist) {
    need(c.setup_done || c.outer_prec != 0, AMGS_ERR_STATE, "amgs_setup first");
    need(c.gauge_set, AMGS_ERR_STATE, "amgs_set_gauge first");
    need(nrhs >= 1 && nrhs <= c.P.max_rhs, AMGS_ERR_ARG, "nrhs must be in [1, max_rhs]");
    need(c.ncfg == 1 || nrhs == c.ncfg, AMGS_ERR_ARG, "a multi-configuration context solves one rhs per configuration");
    auto t0 = std::chrono::steady_clock::now();
    outer_solve(c, nrhs);
    c.sync();
    check_coarse_error(c);
    c.last_ms[2] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c.last_ms[0] = c.last_ms[1] + c.last_ms[2];
    std::vector<int> it(nrhs), nh(nrhs);
    HIP_CHECK(hipMemcpy(it.data(), c.ost.iters, sizeof(int) * nrhs, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(nh.data(), c.ost.nhist, sizeof(int) * nrhs, hipMemcpyDeviceToHost));
    if (iters) std::copy(it.begin(), it.end(), iters);
    if (hist && hist_cap > 0) {
        std::vector<double> h((size_t)nrhs * c.ost.hcap);
        HIP_CHECK(hipMemcpy(h.data(), c.ost.hist, h.size() * 8, hipMemcpyDeviceToHost));
        for (int r = 0; r < nrhs; r++)
            for (int i = 0; i < std::min(hist_cap, std::min(nh[r], c.ost.hcap)); i++)
                hist[(size_t)r * hist_cap + i] = h[(size_t)r * c.ost.hcap + i];
    }
    if (nhist)
        for (int r = 0; r < nrhs; r++) nhist[r] = std::min(nh[r], hist_cap);
    std::vector<double> rr(nrhs);
    true_relres(c, nrhs, rr.data());
    if (relres) std::copy(rr.begin(), rr.end(), relres);
    // convergence is judged on the true residual (an iteration count of
    // restarts*m is also what a solve converging in the last step returns)
    bool all = true;
    for (int r = 0; r < nrhs; r++) all = all && rr[r] < c.P.out_tol * 1.0000001;
    return all ? AMGS_OK : AMGS_NOT_CONVERGED;
}

// Host buffers of a solve pass through two pinned chunks of STAGE_BYTES: the
// host's threads copy one chunk while the DMA of the other runs.
constexpr size_t STAGE_BYTES = size_t(32) << 20;
static void stage_init(Ctx& c) {
    if (c.stage[0]) return;
    for (int k = 0; k < 2; k++) {
        HIP_CHECK(hipHostMalloc(&c.stage[k], STAGE_BYTES));
        HIP_CHECK(hipEventCreateWithFlags(&c.stage_ev[k], hipEventDisableTiming));
        HIP_CHECK(hipEventRecord(c.stage_ev[k], c.stream));
    }
}
static void h2d_staged(Ctx& c, void* dst, const void* src, size_t bytes) {
    stage_init(c);
    int k = 0;
    for (size_t off = 0; off < bytes; off += STAGE_BYTES, k ^= 1) {
        const size_t n = std::min(STAGE_BYTES, bytes - off);
        HIP_CHECK(hipEventSynchronize(c.stage_ev[k]));  // the chunk's previous DMA has read it
        par_copy(c.stage[k], static_cast<const char*>(src) + off, n);
        HIP_CHECK(hipMemcpyAsync(static_cast<char*>(dst) + off, c.stage[k], n, hipMemcpyHostToDevice, c.stream));
        HIP_CHECK(hipEventRecord(c.stage_ev[k], c.stream));
    }
}
static void d2h_staged(Ctx& c, void* dst, const void* src, size_t bytes) {
    stage_init(c);
    // chunk i's DMA is queued before chunk i-1 is copied out of its buffer
    size_t prev_off = 0, prev_n = 0;
    int k = 0, prev_k = -1;
    auto drain = [&]() {
        if (prev_k < 0) return;
        HIP_CHECK(hipEventSynchronize(c.stage_ev[prev_k]));
        par_copy(static_cast<char*>(dst) + prev_off, c.stage[prev_k], prev_n);
    };
    for (size_t off = 0; off < bytes; off += STAGE_BYTES, k ^= 1) {
        const size_t n = std::min(STAGE_BYTES, bytes - off);
        HIP_CHECK(hipMemcpyAsync(c.stage[k], static_cast<const char*>(src) + off, n, hipMemcpyDeviceToHost, c.stream));
        HIP_CHECK(hipEventRecord(c.stage_ev[k], c.stream));
        drain();
        prev_off = off;
        prev_n = n;
        prev_k = k;
    }
    drain();
}

extern "C" int amgs_solve(amgs_ctx* ctx, int nrhs, const double* b, double* x, int* iters, double* relres, double* hist,
                          int hist_cap, int* nhist) {
    if (!ctx || !b || !x) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    need(nrhs >= 1 && nrhs <= c.P.max_rhs, AMGS_ERR_ARG, "nrhs must be in [1, max_rhs]");
    const size_t n0 = c.nvec(0), bytes = sizeof(c64) * n0 * nrhs;
    need(all_finite(b, 2 * n0 * nrhs), AMGS_ERR_NAN, "right-hand side holds NaN/Inf");
    h2d_staged(c, c.ob.p, b, bytes);
    int rc = solve_impl(c, nrhs, iters, relres, hist, hist_cap, nhist);
    d2h_staged(c, x, c.ox.p, bytes);
    return rc;
    API_END
}

extern "C" int amgs_solve_dev(amgs_ctx* ctx, int nrhs, const void* b_dev, void* x_dev, int* iters, double* relres) {
    if (!ctx || !b_dev || !x_dev) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    need(nrhs >= 1 && nrhs <= c.P.max_rhs, AMGS_ERR_ARG, "nrhs must be in [1, max_rhs]");
    const size_t n0 = c.nvec(0);
    HIP_CHECK(hipMemcpyAsync(c.ob.p, b_dev, sizeof(c64) * n0 * nrhs, hipMemcpyDeviceToDevice, c.stream));
    int rc = solve_impl(c, nrhs, iters, relres, nullptr, 0, nullptr);
    HIP_CHECK(hipMemcpyAsync(x_dev, c.ox.p, sizeof(c64) * n0 * nrhs, hipMemcpyDeviceToDevice, c.stream));
    c.sync();
    return rc;
    API_END
}

// ---- autotuner of the level-0 SAP launch shape ----
// Times one SAP sweep (all kernels of both colours) for a set of workgroup
// targets, i.e. chunks per (block, rhs) item, and keeps the fastest for this
// batch size.  Deterministic afterwards: the choice only fixes the chunking.
__global__ void k_fill_pattern(size_t n, c32* out32, c64* out64) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = 0.5 - (double)((i * 2654435761u) % 1000003u) / 1000003.0;
    const double b = 0.5 - (double)((i * 40503u + 7u) % 999983u) / 999983.0;
    if (out32) out32[i] = c32{(float)a, (float)b};
    if (out64) out64[i] = c64{a, b};
}

template <typename R> static int autotune_sap(Ctx& c, Engine<R>& e, int nrhs, double* best_ms) {
    const Geo& g = c.geo[0];
    const size_t tot = (size_t)nrhs * c.nvec(0);
    DevBuf v, x;
    v.alloc(sizeof(cplx<R>) * tot);
    x.alloc(sizeof(cplx<R>) * tot);
    const bool f32 = std::is_same<R, float>::value;
    hipLaunchKernelGGL(k_fill_pattern, dim3(cdiv_u(tot, 256)), dim3(256), 0, c.stream, tot,
                       f32 ? (c32*)v.p : nullptr, f32 ? nullptr : (c64*)v.p);
    const int maxc = (int)cdiv_u(g.sspb, FT);
    std::vector<int> cand;
    if (sap_run_rw<R>(c) > 0) cand.push_back(2);    // one workgroup per item, streaming (k_sap_run)
    if (sap_block_pp<R>(c) > 0) cand.push_back(1);  // one workgroup per item (k_sap_block)
    for (int t : {512, 1024, 1536, 2048, 2560, 3072, 4096}) {
        const int nch = sap_chunks(g, nrhs, t);
        bool dup = false;
        for (int u : cand) dup = dup || sap_chunks(g, nrhs, u) == nch;
        if (!dup && nch <= maxc) cand.push_back(t);
    }
    int best = sap_default_wgs();
    double bms = 1e30;
    for (int t : cand) {
        e.sap_tuned.assign(1, {nrhs, t});
        float ms = 0, acc = 0;
        for (int rep = 0; rep < 3; rep++) {
            hipLaunchKernelGGL(k_zero<R>, dim3(cdiv_u(tot, 256)), dim3(256), 0, c.stream, tot, x.as<cplx<R>>());
            HIP_CHECK(hipEventRecord(c.ev0, c.stream));
            e.sap0(nrhs, v.as<cplx<R>>(), x.as<cplx<R>>(), 1);
            HIP_CHECK(hipEventRecord(c.ev1, c.stream));
            HIP_CHECK(hipEventSynchronize(c.ev1));
            HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
            if (rep > 0) acc += ms;  // the first run warms up
        }
        if (acc < bms) {
            bms = acc;
            best = t;
        }
    }
    e.sap_tuned.assign(1, {nrhs, best});
    if (best_ms) *best_ms = bms / 2;
    return best;
}

// Workgroups per right-hand side of the coarse K-cycle launch: times the
// coarse solve of one K-cycle (FGMRES at level 1 and everything below it) on a
// fixed input for each candidate and keeps the fastest for this batch size.
// The members split only the Schwarz blocks, so every candidate gives the same
// result bit for bit.
template <typename R> static int autotune_coarse(Ctx& c, Engine<R>& e, int nrhs, double* best_ms) {
    if (e.grp_max <= 1) {
        if (best_ms) *best_ms = 0;
        return 1;
    }
    const size_t n1 = c.nvec(1), tot = (size_t)nrhs * n1;
    DevBuf v, x;
    v.alloc(sizeof(cplx<R>) * tot);
    x.alloc(sizeof(cplx<R>) * tot);
    const bool f32 = std::is_same<R, float>::value;
    hipLaunchKernelGGL(k_fill_pattern, dim3(cdiv_u(tot, 256)), dim3(256), 0, c.stream, tot,
                       f32 ? (c32*)v.p : nullptr, f32 ? nullptr : (c64*)v.p);
    const int mode = c.P.cycle == 0 ? CM_CYCLE : CM_KC_SOLVE;
    int best = 1;
    double bms = 1e30;
    for (int g = 1; g <= e.grp_max; g *= 2) {
        e.coarse_tuned.assign(1, {nrhs, g});
        float ms = 0, acc = 0;
        for (int rep = 0; rep < 3; rep++) {
            HIP_CHECK(hipEventRecord(c.ev0, c.stream));
            e.coarse(mode, 1, nrhs, v.as<cplx<R>>(), n1, x.as<cplx<R>>(), n1);
            HIP_CHECK(hipEventRecord(c.ev1, c.stream));
            HIP_CHECK(hipEventSynchronize(c.ev1));
            HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
            if (rep > 0) acc += ms;  // the first run warms up
        }
        check_coarse_error(c);
        if (acc < bms) {
            bms = acc;
            best = g;
        }
    }
    e.coarse_tuned.assign(1, {nrhs, best});
    if (best_ms) *best_ms = bms / 2;
    return best;
}

extern "C" int amgs_autotune(amgs_ctx* ctx, int nrhs, int* best_wgs, double* sweep_ms) {
    if (!ctx) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    need(c.setup_done, AMGS_ERR_STATE, "amgs_setup first");
    need(nrhs >= 1 && nrhs <= c.P.max_rhs, AMGS_ERR_ARG, "nrhs must be in [1, max_rhs]");
    c.clear_graphs();
    const int best = c.mixed() ? autotune_sap(c, c.e32, nrhs, sweep_ms) : autotune_sap(c, c.e64, nrhs, sweep_ms);
    if (c.mixed()) autotune_coarse(c, c.e32, nrhs, nullptr);
    else autotune_coarse(c, c.e64, nrhs, nullptr);
    if (best_wgs) *best_wgs = best;
    return AMGS_OK;
    API_END
}

extern "C" int amgs_get_tuning(amgs_ctx* ctx, int nrhs, int* sap_wgs, int* coarse_grp) {
    if (!ctx || nrhs < 1) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    if (c.mixed()) {
        if (sap_wgs) *sap_wgs = c.e32.sap_wgs(nrhs);
        if (coarse_grp) *coarse_grp = c.e32.coarse_grp(nrhs);
    } else {
        if (sap_wgs) *sap_wgs = c.e64.sap_wgs(nrhs);
        if (coarse_grp) *coarse_grp = c.e64.coarse_grp(nrhs);
    }
    return AMGS_OK;
    API_END
}

template <typename R> static void set_tuning(Engine<R>& e, int nrhs, int sap_wgs, int coarse_grp) {
    auto put = [nrhs](std::vector<std::pair<int, int>>& v, int val) {
        for (auto it = v.begin(); it != v.end();)
            it = it->first == nrhs ? v.erase(it) : it + 1;
        if (val > 0) v.push_back({nrhs, val});
    };
    if (sap_wgs >= 0) put(e.sap_tuned, sap_wgs);
    if (coarse_grp >= 0) put(e.coarse_tuned, coarse_grp);
}

extern "C" int amgs_set_tuning(amgs_ctx* ctx, int nrhs, int sap_wgs, int coarse_grp) {
    if (!ctx || nrhs < 1 || coarse_grp > GRP_MAX) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    c.clear_graphs();
    if (c.mixed()) set_tuning(c.e32, nrhs, sap_wgs, coarse_grp);
    set_tuning(c.e64, nrhs, sap_wgs, coarse_grp);
    return AMGS_OK;
    API_END
}

extern "C" int amgs_set_graph(amgs_ctx* ctx, int enable) {
    if (!ctx) return AMGS_ERR_ARG;
    ctx->c.graph = enable != 0;
    return AMGS_OK;
}

// ---- uneven multi-configuration batch ----
// Independent contexts (each its own lattice size, hierarchy, mass and HIP
// stream) solve concurrently: one host thread per context drives its solve, so
// the kernels of all configurations overlap on the device.  A thread may wait
// a given time before its first launch, to offset the phases of sub-batches
// that would otherwise run in lockstep.
extern "C" int amgs_solve_multi_staggered(int nctx, amgs_ctx* const* ctxs, const int* nrhs, const void* const* b_dev,
                                          void* const* x_dev, const double* start_ms, int* iters, double* relres,
                                          int* status) {
    if (nctx < 1 || !ctxs || !nrhs || !b_dev || !x_dev || !iters || !relres || !status) return AMGS_ERR_ARG;
    std::vector<size_t> off(nctx + 1, 0);
    for (int i = 0; i < nctx; i++) {
        if (!ctxs[i] || nrhs[i] < 1) return AMGS_ERR_ARG;
        if (start_ms && !(start_ms[i] >= 0.0 && start_ms[i] <= 1000.0)) return AMGS_ERR_ARG;
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return AMGS_ERR_ARG;  // one solve per context at a time
        off[i + 1] = off[i] + nrhs[i];
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    th.reserve(nctx);
    for (int i = 0; i < nctx; i++)
        th.emplace_back([&, i]() {
            if (hipSetDevice(ctxs[i]->c.device) != hipSuccess) {
                status[i] = AMGS_ERR_HIP;
                return;
            }
            if (start_ms && start_ms[i] > 0.0)
                std::this_thread::sleep_until(t0 + std::chrono::duration<double, std::milli>(start_ms[i]));
            status[i] = amgs_solve_dev(ctxs[i], nrhs[i], b_dev[i], x_dev[i], iters + off[i], relres + off[i]);
        });
    for (std::thread& t : th) t.join();
    int rc = AMGS_OK;
    for (int i = 0; i < nctx; i++) {
        if (status[i] < 0) return status[i];
        if (status[i] == AMGS_NOT_CONVERGED) rc = AMGS_NOT_CONVERGED;
    }
    return rc;
}

extern "C" int amgs_solve_multi(int nctx, amgs_ctx* const* ctxs, const int* nrhs, const void* const* b_dev,
                                void* const* x_dev, int* iters, double* relres, int* status) {
    return amgs_solve_multi_staggered(nctx, ctxs, nrhs, b_dev, x_dev, nullptr, iters, relres, status);
}

// ---- comparison solvers on the fine lattice (FP64) ----
using zc = std::complex<double>;

struct KryRun {
    Ctx& c;
    int nrhs;
    unsigned ntile;
    KCoef k{};
    std::vector<c64> sums;
    KryRun(Ctx& cc, int n) : c(cc), nrhs(n), ntile(cdiv_u(cc.geo[0].ns, FT)), sums(2 * n) {
        c.kpart.alloc(sizeof(c64) * 2 * ntile * nrhs);
        c.ksum.alloc(sizeof(c64) * 2 * nrhs);
        k.act = nrhs >= 32 ? 0xffffffffu : ((1u << nrhs) - 1u);
    }
    c64* part() { return c.kpart.as<c64>(); }
    void vec(int op, c64* v0, c64* v1 = nullptr, c64* v2 = nullptr, c64* v3 = nullptr, c64* v4 = nullptr,
             c64* v5 = nullptr, bool dots = false) {
        hipLaunchKernelGGL(k_vec, dim3(ntile, nrhs), dim3(FT), 0, c.stream, c.geo[0].ns, op, k, v0, v1, v2, v3, v4, v5,
                           dots ? part() : (c64*)nullptr);
    }
    template <bool DAG> void stencil(const c64* in, c64* out, const c64* y, int mode) {
        hipLaunchKernelGGL(k_stencil_dots<DAG>, dim3(ntile, nrhs), dim3(FT), 0, c.stream, c.geo[0], c.P.m0 + 2.0,
                           (const c64*)c.U64.as<c64>(), in, out, y, part(), mode, k.act);
    }
    // per-rhs sums of the last partials (host synchronises)
    void fold() {
        hipLaunchKernelGGL(k_sum_parts, dim3(nrhs), dim3(FT), 0, c.stream, (const c64*)part(), (int)ntile,
                           c.ksum.as<c64>());
        HIP_CHECK(hipMemcpyAsync(sums.data(), c.ksum.p, sizeof(c64) * 2 * nrhs, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    }
    zc q(int r, int i) const { return zc(sums[2 * r + i].re, sums[2 * r + i].im); }
    bool on(int r) const { return (k.act >> r) & 1u; }
};

static c64 to_c(zc z) { return c64{z.real(), z.imag()}; }

// Bi-CGstab (src/bi_cgstab.cpp:5-87) on D_phi from x0 = 0, b in c.ob, x in
// c.ox; iters[r] = k+1 at convergence or max_iter; hist gets |s| per iteration.
static void bicgstab_impl(Ctx& c, int nrhs, int max_iter, double tol, int* iters, double* hist, int hcap, int* nhist) {
    const size_t tot = c.nvec(0) * nrhs;
    c.kbuf.alloc(sizeof(c64) * tot * 6);
    c64* base = c.kbuf.as<c64>();
    c64 *r = base, *rt = base + tot, *d = base + 2 * tot, *s = base + 3 * tot, *t = base + 4 * tot, *Ad = base + 5 * tot;
    c64* x = c.ox.as<c64>();
    c64* b = c.ob.as<c64>();
    KryRun K(c, nrhs);
    HIP_CHECK(hipMemsetAsync(x, 0, sizeof(c64) * tot, c.stream));
    HIP_CHECK(hipMemcpyAsync(r, b, sizeof(c64) * tot, hipMemcpyDeviceToDevice, c.stream));  // r = b - D 0
    HIP_CHECK(hipMemcpyAsync(rt, b, sizeof(c64) * tot, hipMemcpyDeviceToDevice, c.stream));
    K.vec(KOP_DOT, b, b, nullptr, nullptr, nullptr, nullptr, true);
    K.fold();
    std::vector<double> nb(nrhs);
    std::vector<zc> alpha(nrhs), omega(nrhs), rho(nrhs), rho_prev(nrhs);
    for (int q = 0; q < nrhs; q++) {
        nb[q] = std::sqrt(K.q(q, 0).real());
        iters[q] = max_iter;
        if (nhist) nhist[q] = 0;
    }
    K.vec(KOP_DOT, r, rt, nullptr, nullptr, nullptr, nullptr, true);  // rho_0 = <r, rt>
    K.fold();
    for (int q = 0; q < nrhs; q++) rho[q] = K.q(q, 0);
    for (int k = 0; k < max_iter && K.k.act; k++) {
        if (k == 0) {
            HIP_CHECK(hipMemcpyAsync(d, r, sizeof(c64) * tot, hipMemcpyDeviceToDevice, c.stream));
        } else {
            for (int q = 0; q < nrhs; q++)
                if (K.on(q)) {
                    K.k.b[q] = to_c(alpha[q] * rho[q] / (omega[q] * rho_prev[q]));
                    K.k.c[q] = to_c(omega[q]);
                }
            K.vec(KOP_DIR, d, r, Ad);
        }
        K.stencil<false>(d, Ad, rt, 1);  // Ad = D d, <Ad, rt>
        K.fold();
        for (int q = 0; q < nrhs; q++)
            if (K.on(q)) {
                alpha[q] = rho[q] / K.q(q, 0);
                K.k.a[q] = to_c(alpha[q]);
            }
        K.vec(KOP_S, s, r, Ad, nullptr, nullptr, nullptr, true);  // s = r - alpha Ad, |s|^2
        K.fold();
        unsigned conv = 0;
        for (int q = 0; q < nrhs; q++)
            if (K.on(q)) {
                const double err = std::sqrt(K.q(q, 0).real());
                if (hist && nhist && nhist[q] < hcap) hist[(size_t)q * hcap + nhist[q]++] = err;
                if (err < tol * nb[q]) {
                    conv |= 1u << q;
                    iters[q] = k + 1;
                }
            }
        if (conv) {  // x = x + alpha d for the right-hand sides that converged now
            const unsigned keep = K.k.act;
            K.k.act = conv;
            K.vec(KOP_XD, x, d);
            K.k.act = keep & ~conv;
        }
        if (!K.k.act) break;
        K.stencil<false>(s, t, s, 2);  // t = D s, <s, t>, <t, t>
        K.fold();
        for (int q = 0; q < nrhs; q++)
            if (K.on(q)) {
                omega[q] = K.q(q, 0) / K.q(q, 1);
                K.k.c[q] = to_c(omega[q]);
            }
        K.vec(KOP_RX, r, s, t, x, d, rt, true);  // r = s - omega t; x += alpha d + omega s; <r, rt>
        K.fold();
        for (int q = 0; q < nrhs; q++)
            if (K.on(q)) {
                rho_prev[q] = rho[q];
                rho[q] = K.q(q, 0);
            }
    }
    c.sync();
}

// CG on D D^dagger (src/conjugate_gradient.cpp:5-65) from x0 = phi (c.ob);
// iters[r] = loop passes at convergence (the reference prints one less), or
// -max_iter.
static void cg_impl(Ctx& c, int nrhs, int max_iter, double tol, int* iters) {
    const size_t tot = c.nvec(0) * nrhs;
    c.kbuf.alloc(sizeof(c64) * tot * 4);
    c64* base = c.kbuf.as<c64>();
    c64 *r = base, *d = base + tot, *Ad = base + 2 * tot, *tmp = base + 3 * tot;
    c64* x = c.ox.as<c64>();
    c64* phi = c.ob.as<c64>();
    KryRun K(c, nrhs);
    HIP_CHECK(hipMemcpyAsync(x, phi, sizeof(c64) * tot, hipMemcpyDeviceToDevice, c.stream));
    K.stencil<true>(x, tmp, nullptr, 0);
    K.stencil<false>(tmp, Ad, nullptr, 0);
    for (int q = 0; q < nrhs; q++) K.k.a[q] = c64{1.0, 0.0};
    HIP_CHECK(hipMemcpyAsync(r, phi, sizeof(c64) * tot, hipMemcpyDeviceToDevice, c.stream));
    // r = phi - Ad (as x += 0*d; r -= 1*Ad) and |r|^2
    K.vec(KOP_CGXR, tmp, tmp, r, Ad, nullptr, nullptr, true);
    K.fold();
    HIP_CHECK(hipMemcpyAsync(d, r, sizeof(c64) * tot, hipMemcpyDeviceToDevice, c.stream));
    std::vector<zc> rn(nrhs);
    for (int q = 0; q < nrhs; q++) rn[q] = K.q(q, 0);
    K.vec(KOP_DOT, phi, phi, nullptr, nullptr, nullptr, nullptr, true);
    K.fold();
    std::vector<double> np(nrhs);
    for (int q = 0; q < nrhs; q++) {
        np[q] = std::sqrt(K.q(q, 0).real());
        iters[q] = -max_iter;
    }
    for (int k = 0; k < max_iter && K.k.act; k++) {
        K.stencil<true>(d, tmp, nullptr, 0);
        K.stencil<false>(tmp, Ad, d, 2);  // Ad = D D^dagger d, <d, Ad>
        K.fold();
        for (int q = 0; q < nrhs; q++)
            if (K.on(q)) K.k.a[q] = to_c(rn[q] / K.q(q, 0));
        K.vec(KOP_CGXR, x, d, r, Ad, nullptr, nullptr, true);  // x += a d; r -= a Ad; |r|^2
        K.fold();
        unsigned conv = 0;
        for (int q = 0; q < nrhs; q++)
            if (K.on(q)) {
                const double e2 = K.q(q, 0).real();
                if (std::sqrt(e2) < tol * np[q]) {
                    conv |= 1u << q;
                    iters[q] = k + 1;
                } else {
                    K.k.b[q] = to_c(zc(e2, 0.0) / rn[q]);
                    rn[q] = zc(e2, 0.0);
                }
            }
        K.k.act &= ~conv;
        if (!K.k.act) break;
        K.vec(KOP_CGD, d, r);  // d = beta d + r
    }
    c.sync();
}

static void krylov_prologue(Ctx& c, int nrhs) {
    single_config(c);
    need(c.gauge_set, AMGS_ERR_STATE, "amgs_set_gauge first");
    need(nrhs >= 1 && nrhs <= std::min(c.P.max_rhs, KMAXB), AMGS_ERR_ARG, "nrhs must be in [1, min(max_rhs, 32)]");
}

// relative residual |b - A x| / |b| with A = D (op 0) or D D^dagger (op 1)
static void krylov_relres(Ctx& c, int nrhs, int op, double* relres) {
    const size_t tot = c.nvec(0) * nrhs;
    c.kbuf.alloc(sizeof(c64) * tot * 3);
    c64* t1 = c.kbuf.as<c64>();
    c64* t2 = t1 + tot;
    c64* t3 = t2 + tot;
    KryRun K(c, nrhs);
    if (op == 1) {
        K.stencil<true>(c.ox.as<c64>(), t1, nullptr, 0);
        K.stencil<false>(t1, t2, nullptr, 0);
    } else {
        K.stencil<false>(c.ox.as<c64>(), t2, nullptr, 0);
    }
    for (int q = 0; q < nrhs; q++) K.k.a[q] = c64{1.0, 0.0};
    HIP_CHECK(hipMemcpyAsync(t1, c.ob.p, sizeof(c64) * tot, hipMemcpyDeviceToDevice, c.stream));
    K.vec(KOP_CGXR, t3, t3, t1, t2, nullptr, nullptr, true);  // t1 = b - A x, |t1|^2 (t3 is scratch)
    K.fold();
    std::vector<double> e(nrhs);
    for (int q = 0; q < nrhs; q++) e[q] = K.q(q, 0).real();
    K.vec(KOP_DOT, c.ob.as<c64>(), c.ob.as<c64>(), nullptr, nullptr, nullptr, nullptr, true);
    K.fold();
    for (int q = 0; q < nrhs; q++) {
        const double nb = K.q(q, 0).real();
        relres[q] = nb > 0 ? std::sqrt(e[q] / nb) : 0.0;
    }
}

extern "C" int amgs_bicgstab(amgs_ctx* ctx, int nrhs, const double* b, double* x, int max_iter, double tol,
                             int* iters, double* relres, double* hist, int hist_cap, int* nhist) {
    if (!ctx || !b || !x || !iters || !relres || max_iter < 1) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    krylov_prologue(c, nrhs);
    const size_t tot = c.nvec(0) * nrhs;
    HIP_CHECK(hipMemcpyAsync(c.ob.p, b, sizeof(c64) * tot, hipMemcpyHostToDevice, c.stream));
    bicgstab_impl(c, nrhs, max_iter, tol, iters, hist_cap > 0 ? hist : nullptr, hist_cap, nhist);
    krylov_relres(c, nrhs, 0, relres);
    HIP_CHECK(hipMemcpy(x, c.ox.p, sizeof(c64) * tot, hipMemcpyDeviceToHost));
    for (int q = 0; q < nrhs; q++)
        if (!(relres[q] < tol * 1.0000001)) return AMGS_NOT_CONVERGED;
    return AMGS_OK;
    API_END
}

extern "C" int amgs_cg_normal(amgs_ctx* ctx, int nrhs, const double* phi, double* x, int max_iter, double tol,
                              int* iters, double* relres) {
    if (!ctx || !phi || !x || !iters || !relres || max_iter < 1) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    krylov_prologue(c, nrhs);
    const size_t tot = c.nvec(0) * nrhs;
    HIP_CHECK(hipMemcpyAsync(c.ob.p, phi, sizeof(c64) * tot, hipMemcpyHostToDevice, c.stream));
    cg_impl(c, nrhs, max_iter, tol, iters);
    krylov_relres(c, nrhs, 1, relres);
    HIP_CHECK(hipMemcpy(x, c.ox.p, sizeof(c64) * tot, hipMemcpyDeviceToHost));
    for (int q = 0; q < nrhs; q++)
        if (iters[q] < 0) return AMGS_NOT_CONVERGED;
    return AMGS_OK;
    API_END
}

// outer FGMRES with another preconditioner (1: one SAP sweep, 2: none)
static int solve_with(amgs_ctx* ctx, int prec, int nrhs, const double* b, double* x, int* iters, double* relres,
                      double* hist, int hist_cap, int* nhist) {
    Ctx& c = ctx->c;
    single_config(c);
    need(c.gauge_set, AMGS_ERR_STATE, "amgs_set_gauge first");
    need(nrhs >= 1 && nrhs <= c.P.max_rhs && (prec != 1 || nrhs <= c.e64.cap), AMGS_ERR_ARG, "bad nrhs");
    const size_t n0 = c.nvec(0);
    bind_engine(c, c.e64, c.U64, c.tv64, c.links64);
    // the FP64 preconditioned basis (mixed-precision contexts allocate it here)
    c.oZ.alloc(sizeof(c64) * c.P.out_m * c.ocap * n0);
    HIP_CHECK(hipMemcpyAsync(c.ob.p, b, sizeof(c64) * n0 * nrhs, hipMemcpyHostToDevice, c.stream));
    c.outer_prec = prec;
    int rc;
    try {
        rc = solve_impl(c, nrhs, iters, relres, hist, hist_cap, nhist);
    } catch (...) {
        c.outer_prec = 0;
        throw;
    }
    c.outer_prec = 0;
    HIP_CHECK(hipMemcpy(x, c.ox.p, sizeof(c64) * n0 * nrhs, hipMemcpyDeviceToHost));
    return rc;
}

extern "C" int amgs_fgmres_sap(amgs_ctx* ctx, int nrhs, const double* b, double* x, int* iters, double* relres,
                               double* hist, int hist_cap, int* nhist) {
    if (!ctx || !b || !x) return AMGS_ERR_ARG;
    API_BEGIN
    return solve_with(ctx, 1, nrhs, b, x, iters, relres, hist, hist_cap, nhist);
    API_END
}

extern "C" int amgs_gmres(amgs_ctx* ctx, int nrhs, const double* b, double* x, int* iters, double* relres,
                          double* hist, int hist_cap, int* nhist) {
    if (!ctx || !b || !x) return AMGS_ERR_ARG;
    API_BEGIN
    return solve_with(ctx, 2, nrhs, b, x, iters, relres, hist, hist_cap, nhist);
    API_END
}

// ---- fermion force (new: adjoint of the solve) ----
// S = phi^dagger (D D^dagger)^-1 phi = |Y|^2 with Y = D^-1 phi; X = D^dagger^-1 Y
// = s3 D^-1 s3 Y (D^dagger = s3 D s3); dS/dtheta from k_force.  The form
// phi^dagger (D^dagger D)^-1 phi is the same at s3 phi.  Two solves of the
// configured kind (plain or mixed), both to out_tol.
extern "C" int amgs_force(amgs_ctx* ctx, int nrhs, int form, const double* phi, double* dsdtheta, double* action) {
    if (!ctx || !phi || !dsdtheta || !action) return AMGS_ERR_ARG;
    if (form != AMGS_FORCE_DDAG_D && form != AMGS_FORCE_D_DDAG) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    need(nrhs >= 1 && nrhs <= std::min(c.P.max_rhs, KMAXB), AMGS_ERR_ARG, "nrhs must be in [1, min(max_rhs, 32)]");
    need(c.ncfg == 1 || nrhs == c.ncfg, AMGS_ERR_ARG, "a multi-configuration context takes one phi per configuration");
    const Geo& g = c.geo[0];
    const size_t tot = c.nvec(0) * nrhs, nsr = (size_t)g.ns * nrhs;
    DevBuf Yb, Fb;
    Yb.alloc(sizeof(c64) * tot);
    Fb.alloc(sizeof(double) * 2 * nsr);
    HIP_CHECK(hipMemcpyAsync(c.ob.p, phi, sizeof(c64) * tot, hipMemcpyHostToDevice, c.stream));
    if (form == AMGS_FORCE_DDAG_D)
        hipLaunchKernelGGL(k_sigma3, dim3(cdiv_u(nsr, 256)), dim3(256), 0, c.stream, nsr, (const c64*)c.ob.as<c64>(),
                           c.ob.as<c64>());
    std::vector<int> it(nrhs);
    std::vector<double> rr(nrhs);
    int rc1 = solve_impl(c, nrhs, it.data(), rr.data(), nullptr, 0, nullptr);  // Y = D^-1 phi
    HIP_CHECK(hipMemcpyAsync(Yb.p, c.ox.p, sizeof(c64) * tot, hipMemcpyDeviceToDevice, c.stream));
    hipLaunchKernelGGL(k_sigma3, dim3(cdiv_u(nsr, 256)), dim3(256), 0, c.stream, nsr, (const c64*)Yb.as<c64>(),
                       c.ob.as<c64>());
    int rc2 = solve_impl(c, nrhs, it.data(), rr.data(), nullptr, 0, nullptr);  // s3 X = D^-1 s3 Y
    hipLaunchKernelGGL(k_sigma3, dim3(cdiv_u(nsr, 256)), dim3(256), 0, c.stream, nsr, (const c64*)c.ox.as<c64>(),
                       c.ox.as<c64>());
    hipLaunchKernelGGL(k_force, dim3(cdiv_u(g.ns, FT), nrhs), dim3(FT), 0, c.stream, g, (const c64*)c.U64.as<c64>(),
                       (const c64*)c.ox.as<c64>(), (const c64*)Yb.as<c64>(), Fb.as<double>(),
                       c.ncfg > 1 ? c.nvec(0) : 0);
    KryRun K(c, nrhs);
    K.vec(KOP_DOT, Yb.as<c64>(), Yb.as<c64>(), nullptr, nullptr, nullptr, nullptr, true);
    K.fold();
    for (int q = 0; q < nrhs; q++) action[q] = K.q(q, 0).real();
    HIP_CHECK(hipMemcpy(dsdtheta, Fb.p, sizeof(double) * 2 * nsr, hipMemcpyDeviceToHost));
    return (rc1 == AMGS_OK && rc2 == AMGS_OK) ? AMGS_OK : AMGS_NOT_CONVERGED;
    API_END
}

// Tile autotuner of the batched coarse operator: wavefronts per workgroup
// (each owns one 16-rhs x 16-row tile of a site) from 1 to 8, timed on a
// fixed input; the fastest is kept per (level, batch size).  Every candidate
// computes the same tiles in the same order, so results do not depend on it.
template <typename R> static int autotune_cdop(Ctx& c, Engine<R>& e, int l, int nrhs, double* best_us) {
    const size_t tot = (size_t)nrhs * c.nvec(l);
    DevBuf v, x;
    v.alloc(sizeof(cplx<R>) * tot);
    x.alloc(sizeof(cplx<R>) * tot);
    const bool f32 = std::is_same<R, float>::value;
    hipLaunchKernelGGL(k_fill_pattern, dim3(cdiv_u(tot, 256)), dim3(256), 0, c.stream, tot,
                       f32 ? (c32*)v.p : nullptr, f32 ? nullptr : (c64*)v.p);
    int best = 4;
    double bus = 1e30;
    const int reps = 20;
    for (int wpg : {1, 2, 4, 8}) {
        cdop_mfma(c, e, l, nrhs, v.as<cplx<R>>(), x.as<cplx<R>>(), wpg);  // warm-up
        HIP_CHECK(hipEventRecord(c.ev0, c.stream));
        for (int r = 0; r < reps; r++) cdop_mfma(c, e, l, nrhs, v.as<cplx<R>>(), x.as<cplx<R>>(), wpg);
        HIP_CHECK(hipEventRecord(c.ev1, c.stream));
        HIP_CHECK(hipEventSynchronize(c.ev1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        if (ms * 1e3 / reps < bus) {
            bus = ms * 1e3 / reps;
            best = wpg;
        }
    }
    const long long key = (long long)l << 32 | (unsigned)nrhs;
    for (auto it = e.cdop_tuned.begin(); it != e.cdop_tuned.end();)
        it = it->first == key ? e.cdop_tuned.erase(it) : it + 1;
    e.cdop_tuned.push_back({key, best});
    if (best_us) *best_us = bus;
    return best;
}

extern "C" int amgs_apply_D_batched_dev(amgs_ctx* ctx, int level, int nrhs, int precision, const void* in, void* out) {
    if (!ctx || !in || !out) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.setup_done, AMGS_ERR_STATE, "coarse operators need amgs_setup");
    need(level >= 1 && level < c.L, AMGS_ERR_ARG, "level must be a coarse level (1 .. levels-1)");
    need(nrhs >= 1, AMGS_ERR_ARG, "nrhs must be >= 1");
    if (precision == AMGS_PREC_FP64) {
        cdop_mfma(c, c.e64, level, nrhs, (const c64*)in, (c64*)out);
    } else {
        need(c.mixed(), AMGS_ERR_STATE, "the FP32 hierarchy exists in mixed-precision contexts only");
        cdop_mfma(c, c.e32, level, nrhs, (const c32*)in, (c32*)out);
    }
    c.sync();
    return AMGS_OK;
    API_END
}

extern "C" int amgs_autotune_batched_D(amgs_ctx* ctx, int level, int nrhs, int precision, int* best_wpg,
                                       double* best_us) {
    if (!ctx) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.setup_done, AMGS_ERR_STATE, "amgs_setup first");
    need(level >= 1 && level < c.L && nrhs >= 1, AMGS_ERR_ARG, "bad level / nrhs");
    need(precision == AMGS_PREC_FP64 || c.mixed(), AMGS_ERR_STATE, "FP32 hierarchy needs a mixed-precision context");
    const int w = precision == AMGS_PREC_FP64 ? autotune_cdop(c, c.e64, level, nrhs, best_us)
                                              : autotune_cdop(c, c.e32, level, nrhs, best_us);
    if (best_wpg) *best_wpg = w;
    return AMGS_OK;
    API_END
}

// ---- component entry points (FP64 engine) ----
static size_t level_n(Ctx& c, int l) { return c.nvec(l); }

extern "C" int amgs_apply_D_dev(amgs_ctx* ctx, int level, int dagger, int nrhs, const void* in, void* out) {
    if (!ctx || !in || !out) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    need(level >= 0 && level < c.L, AMGS_ERR_ARG, "level out of range");
    need(nrhs >= 1, AMGS_ERR_ARG, "nrhs must be >= 1");
    need(c.gauge_set, AMGS_ERR_STATE, "amgs_set_gauge first");
    const Geo& g = c.geo[level];
    if (level == 0) {
        const dim3 grid(cdiv_u(g.ns, FT), nrhs);
        const size_t ucs = c.ncfg > 1 ? c.nvec(0) : 0;
        const double* mmv = c.ncfg > 1 ? c.mm64.as<double>() : nullptr;
        need(c.ncfg == 1 || nrhs == c.ncfg, AMGS_ERR_ARG, "one right-hand side per configuration");
        if (dagger)
            hipLaunchKernelGGL((k_wilson<double, true>), grid, dim3(FT), 0, c.stream, g, c.P.m0 + 2.0,
                               (const c64*)c.U64.as<c64>(), (const c64*)in, (c64*)out, ucs, mmv);
        else
            hipLaunchKernelGGL((k_wilson<double, false>), grid, dim3(FT), 0, c.stream, g, c.P.m0 + 2.0,
                               (const c64*)c.U64.as<c64>(), (const c64*)in, (c64*)out, ucs, mmv);
        HIP_CHECK(hipGetLastError());
    } else {
        need(!dagger, AMGS_ERR_ARG, "D^dagger only on the fine level");
        need(c.setup_done, AMGS_ERR_STATE, "coarse operators need amgs_setup");
        single_config(c);
        // all right-hand sides at once as matrix-core block-GEMMs
        cdop_mfma(c, c.e64, level, nrhs, (const c64*)in, (c64*)out);
    }
    c.sync();
    return AMGS_OK;
    API_END
}

// host-pointer wrapper: copy in, run, copy out
template <typename F>
static int host_wrap(amgs_ctx* ctx, size_t nin, const double* in, size_t nout, double* out, F&& f, double* inout = nullptr) {
    Ctx& c = ctx->c;
    c.io_a.alloc(sizeof(c64) * std::max<size_t>(nin, 1));
    c.io_b.alloc(sizeof(c64) * std::max<size_t>(nout, 1));
    if (in) HIP_CHECK(hipMemcpyAsync(c.io_a.p, in, sizeof(c64) * nin, hipMemcpyHostToDevice, c.stream));
    if (inout) HIP_CHECK(hipMemcpyAsync(c.io_b.p, inout, sizeof(c64) * nout, hipMemcpyHostToDevice, c.stream));
    f(c.io_a.as<c64>(), c.io_b.as<c64>());
    HIP_CHECK(hipMemcpyAsync(out ? out : inout, c.io_b.p, sizeof(c64) * nout, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return AMGS_OK;
}

extern "C" int amgs_apply_D(amgs_ctx* ctx, int level, int dagger, int nrhs, const double* in, double* out) {
    if (!ctx || !in || !out) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    need(level >= 0 && level < c.L && nrhs >= 1, AMGS_ERR_ARG, "bad level / nrhs");
    const size_t n = level_n(c, level) * nrhs;
    int rc = AMGS_OK;
    host_wrap(ctx, n, in, n, out, [&](c64* a, c64* b) { rc = amgs_apply_D_dev(ctx, level, dagger, nrhs, a, b); });
    return rc;
    API_END
}

extern "C" int amgs_apply_P(amgs_ctx* ctx, int level, int nrhs, const double* coarse, double* fine) {
    if (!ctx || !coarse || !fine) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.setup_done, AMGS_ERR_STATE, "amgs_setup first");
    need(level >= 0 && level < c.L - 1 && nrhs >= 1 && nrhs <= c.e64.cap, AMGS_ERR_ARG, "bad level / nrhs");
    const size_t nf = c.nvec(level), nc = c.nvec(level + 1);
    return host_wrap(ctx, nc * nrhs, coarse, nf * nrhs, fine, [&](c64* a, c64* b) {
        if (level == 0) {
            const Geo& g = c.geo[0];
            hipLaunchKernelGGL(k_prolong0<double>, dim3(cdiv_u(g.ns, FT), nrhs), dim3(FT), 0, c.stream, g, g.ntest,
                               (const c64*)c.tv64[0].as<c64>(), (const c64*)a, b, c.geo[1].dof, c.geo[1].ns, 0);
            HIP_CHECK(hipGetLastError());
        } else {
            c.e64.coarse(CM_PROLONG, level, nrhs, a, nc, b, nf);
        }
    });
    API_END
}

extern "C" int amgs_apply_Pt(amgs_ctx* ctx, int level, int nrhs, const double* fine, double* coarse) {
    if (!ctx || !coarse || !fine) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.setup_done, AMGS_ERR_STATE, "amgs_setup first");
    need(level >= 0 && level < c.L - 1 && nrhs >= 1 && nrhs <= c.e64.cap, AMGS_ERR_ARG, "bad level / nrhs");
    const size_t nf = c.nvec(level), nc = c.nvec(level + 1);
    return host_wrap(ctx, nf * nrhs, fine, nc * nrhs, coarse, [&](c64* a, c64* b) {
        if (level == 0) c.e64.restrict0(nrhs, c.tv64[0].as<c64>(), a, b);
        else c.e64.coarse(CM_RESTRICT, level, nrhs, a, nf, b, nc);
    });
    API_END
}

extern "C" int amgs_sap(amgs_ctx* ctx, int level, int nrhs, const double* v, double* x, int nu) {
    if (!ctx || !v || !x) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.gauge_set, AMGS_ERR_STATE, "amgs_set_gauge first");
    need(level >= 0 && level < c.L - 1 && nrhs >= 1 && nrhs <= c.e64.cap, AMGS_ERR_ARG, "bad level / nrhs");
    need(level == 0 || c.setup_done, AMGS_ERR_STATE, "coarse SAP needs amgs_setup");
    bind_engine(c, c.e64, c.U64, c.tv64, c.links64);
    const size_t n = c.nvec(level) * nrhs;
    return host_wrap(ctx, n, v, n, nullptr, [&](c64* a, c64* b) {
        if (level == 0) c.e64.sap0(nrhs, a, b, nu);
        else c.e64.coarse(CM_SAP, level, nrhs, a, c.nvec(level), b, c.nvec(level), nu);
    }, x);
    API_END
}

// amgs_sap at level 0 in the inner precision of a mixed context: the FP32
// engine's smoother, with the FP16 Krylov basis under AMGS_PREC_MIXED16 (the
// kernels the mixed solves run, in the form the tuning selects), on host
// buffers converted to FP32 and back.
extern "C" int amgs_sap_inner(amgs_ctx* ctx, int nrhs, const double* v, double* x, int nu) {
    if (!ctx || !v || !x) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.mixed(), AMGS_ERR_STATE, "amgs_sap_inner needs a mixed-precision context");
    need(c.setup_done, AMGS_ERR_STATE, "amgs_setup first");
    need(nrhs >= 1 && nrhs <= c.e32.cap, AMGS_ERR_ARG, "bad nrhs");
    bind_engine(c, c.e32, c.U32, c.tv32, c.links32);
    const size_t n = c.nvec(0) * nrhs;
    DevBuf v32, x32;
    v32.alloc(sizeof(c32) * n);
    x32.alloc(sizeof(c32) * n);
    return host_wrap(ctx, n, v, n, nullptr, [&](c64* a, c64* b) {
        const dim3 grid(cdiv_u(n, 256));
        hipLaunchKernelGGL((k_convert<float, double>), grid, dim3(256), 0, c.stream, n, (const c64*)a, v32.as<c32>());
        hipLaunchKernelGGL((k_convert<float, double>), grid, dim3(256), 0, c.stream, n, (const c64*)b, x32.as<c32>());
        c.e32.sap0(nrhs, v32.as<c32>(), x32.as<c32>(), nu);
        hipLaunchKernelGGL((k_convert<double, float>), grid, dim3(256), 0, c.stream, n, (const c32*)x32.as<c32>(), b);
        HIP_CHECK(hipGetLastError());
    }, x);
    API_END
}

// amgs_sap on a coarse level (l >= 1) on device buffers with the per-rhs
// workspace of the block solves (residual, D x, the Schwarz-block Krylov
// vectors of the workgroup's wavefronts) in the static array above instead of
// the context's: nothing is allocated.  Every workspace entry is written before
// it is read, so the array needs no clearing between calls.
extern "C" int amgs_sap_dev_noalloc(amgs_ctx* ctx, int level, int nrhs, const void* v_dev, void* x_dev, int nu) {
    if (!ctx || !v_dev || !x_dev) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.setup_done, AMGS_ERR_STATE, "coarse SAP needs amgs_setup");
    need(level >= 1 && level < c.L - 1, AMGS_ERR_ARG, "level must be a coarse level with SAP smoothing (1 .. levels-2)");
    need(nrhs >= 1, AMGS_ERR_ARG, "nrhs must be >= 1");
    void* ws = scratch_ptr(c.device);
    need(ws != nullptr, AMGS_ERR_STATE, "static workspace not set up on this device");
    const size_t n = c.nvec(level);
    CArgs<double> a = c.e64.ca;
    CLevel<double>& lv = a.lv[level];
    lv.sR = 0;
    lv.sDx = n;
    lv.blk = 2 * n;
    a.ws_per_rhs = (2 * n + CW * lv.blkStride + 15) & ~size_t(15);  // 256-byte aligned per rhs
    a.ws = (c64*)ws;
    const size_t need_bytes = sizeof(c64) * a.ws_per_rhs * (size_t)nrhs;
    if (need_bytes > AMGS_NOALLOC_SCRATCH_BYTES)
        throw ApiError(AMGS_ERR_ARG, "amgs_sap_dev_noalloc: " + std::to_string(nrhs) + " right-hand sides need " +
                                         std::to_string(need_bytes) + " bytes of workspace, above the static " +
                                         std::to_string((size_t)AMGS_NOALLOC_SCRATCH_BYTES) + "-byte limit");
    bind_engine(c, c.e64, c.U64, c.tv64, c.links64);
    c.e64.coarse(CM_SAP, level, nrhs, (const c64*)v_dev, n, (c64*)x_dev, n, nu, &a);
    c.sync();
    return AMGS_OK;
    API_END
}

extern "C" int amgs_kcycle(amgs_ctx* ctx, int nrhs, const double* eta, double* psi) {
    if (!ctx || !eta || !psi) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.setup_done, AMGS_ERR_STATE, "amgs_setup first");
    need(nrhs >= 1 && nrhs <= c.e64.cap, AMGS_ERR_ARG, "bad nrhs");
    const size_t n = c.nvec(0) * nrhs;
    return host_wrap(ctx, n, eta, n, psi, [&](c64* a, c64* b) { c.e64.kcycle0(nrhs, a, b); });
    API_END
}

// AlgebraicMG::applyMultilevel (src/amg.cpp:181-231): x <- cycle(b, x) until
// |b - D x| < tol |b| or max_cycles, in FP64; x is the initial guess on entry.
// Each right-hand side keeps the iterate of the cycle at which it converged.
extern "C" int amgs_multigrid(amgs_ctx* ctx, int nrhs, const double* b, double* x, int max_cycles, double tol,
                              int* cycles, double* relres) {
    if (!ctx || !b || !x || !cycles || !relres || max_cycles < 1) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.setup_done, AMGS_ERR_STATE, "amgs_setup first");
    need(nrhs >= 1 && nrhs <= c.e64.cap && nrhs <= c.ocap, AMGS_ERR_ARG, "bad nrhs");
    const Geo& g = c.geo[0];
    const size_t n = c.nvec(0), tot = n * nrhs;
    const unsigned ntile = cdiv_u(g.ns, FT);
    c64* db = c.ob.as<c64>();
    c64* dx = c.ox.as<c64>();
    c64* dr = c.or_.as<c64>();
    c64* keep = c.ow.as<c64>();
    HIP_CHECK(hipMemcpyAsync(db, b, sizeof(c64) * tot, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(dx, x, sizeof(c64) * tot, hipMemcpyHostToDevice, c.stream));
    std::vector<double> pr((size_t)nrhs * ntile), pv((size_t)nrhs * ntile);
    DevBuf part;
    part.alloc(sizeof(double) * 2 * nrhs * ntile);
    double* dpr = part.as<double>();
    double* dpv = dpr + (size_t)nrhs * ntile;
    hipLaunchKernelGGL(k_norm_part<double>, dim3(ntile, nrhs), dim3(FT), 0, c.stream, g.ns, (const c64*)db, dpv);
    HIP_CHECK(hipMemcpyAsync(pv.data(), dpv, sizeof(double) * nrhs * ntile, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    std::vector<double> nb(nrhs, 0.0);
    for (int r = 0; r < nrhs; r++)
        for (unsigned t = 0; t < ntile; t++) nb[r] += pv[(size_t)r * ntile + t];
    for (int r = 0; r < nrhs; r++) {
        nb[r] = std::sqrt(nb[r]);
        cycles[r] = 0;
        relres[r] = 1.0;
    }
    int left = nrhs;
    for (int it = 0; it < max_cycles && left > 0; it++) {
        c.e64.kcycle0(nrhs, db, dx, false);
        hipLaunchKernelGGL(k_resid<double>, dim3(ntile, nrhs), dim3(FT), 0, c.stream, g, c.P.m0 + 2.0,
                           (const c64*)c.U64.as<c64>(), (const c64*)db, (const c64*)dx, dr, dpr, (const int*)nullptr, nullptr);
        HIP_CHECK(hipMemcpyAsync(pr.data(), dpr, sizeof(double) * nrhs * ntile, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        for (int r = 0; r < nrhs; r++) {
            if (cycles[r] > 0) continue;
            double e = 0;
            for (unsigned t = 0; t < ntile; t++) e += pr[(size_t)r * ntile + t];
            e = std::sqrt(e);
            relres[r] = nb[r] > 0 ? e / nb[r] : 0.0;
            if (e < tol * nb[r] || it + 1 == max_cycles) {
                cycles[r] = e < tol * nb[r] ? it + 1 : -max_cycles;
                HIP_CHECK(hipMemcpyAsync(keep + (size_t)r * n, dx + (size_t)r * n, sizeof(c64) * n,
                                         hipMemcpyDeviceToDevice, c.stream));
                left--;
            }
        }
    }
    HIP_CHECK(hipMemcpyAsync(x, keep, sizeof(c64) * tot, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    for (int r = 0; r < nrhs; r++)
        if (cycles[r] < 0) return AMGS_NOT_CONVERGED;
    return AMGS_OK;
    API_END
}

extern "C" int amgs_get_tv(amgs_ctx* ctx, int level, int i, double* out) {
    if (!ctx || !out) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.setup_done, AMGS_ERR_STATE, "amgs_setup first");
    need(level >= 0 && level < c.L - 1 && i >= 0 && i < c.geo[level].ntest, AMGS_ERR_ARG, "bad level / index");
    const size_t n = c.nvec(level);
    HIP_CHECK(hipMemcpy(out, c.tv64[level].as<c64>() + (size_t)i * n, sizeof(c64) * n, hipMemcpyDeviceToHost));
    return AMGS_OK;
    API_END
}

extern "C" int amgs_get_links(amgs_ctx* ctx, int level, int which, int mu, double* out) {
    if (!ctx || !out) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.setup_done, AMGS_ERR_STATE, "amgs_setup first");
    need(level >= 1 && level < c.L && which >= 0 && which <= 2 && mu >= 0 && mu <= 1, AMGS_ERR_ARG, "bad arguments");
    const Geo& g = c.geo[level];
    const int D = g.dof;
    const int k = which == 0 ? 0 : (which == 1 ? 1 + mu : 3 + mu);
    std::vector<c64> h((size_t)g.ns * D * D);
    HIP_CHECK(hipMemcpy(h.data(), c.links64[level].as<c64>() + (size_t)k * g.ns * D * D, h.size() * sizeof(c64),
                        hipMemcpyDeviceToHost));
    // device layout [x][col][row] -> [x][row][col]
    for (int x = 0; x < g.ns; x++)
        for (int r = 0; r < D; r++)
            for (int col = 0; col < D; col++) {
                const c64 v = h[((size_t)x * D + col) * D + r];
                out[2 * (((size_t)x * D + r) * D + col)] = v.re;
                out[2 * (((size_t)x * D + r) * D + col) + 1] = v.im;
            }
    return AMGS_OK;
    API_END
}

extern "C" int amgs_time_kernel(amgs_ctx* ctx, int kernel, int nrhs, int reps, double* avg_ms) {
    if (!ctx || !avg_ms || reps < 1) return AMGS_ERR_ARG;
    API_BEGIN
    Ctx& c = ctx->c;
    single_config(c);
    need(c.gauge_set, AMGS_ERR_STATE, "amgs_set_gauge first");
    need(nrhs >= 1 && nrhs <= c.ocap, AMGS_ERR_ARG, "nrhs exceeds capacity");
    const Geo& g = c.geo[0];
    const dim3 grid(cdiv_u(g.ns, FT), nrhs);
    auto launch = [&]() {
        if (kernel == 0)
            hipLaunchKernelGGL((k_wilson<double, false>), grid, dim3(FT), 0, c.stream, g, c.P.m0 + 2.0,
                               (const c64*)c.U64.as<c64>(), (const c64*)c.ob.as<c64>(), c.ow.as<c64>(), (size_t)0,
                               (const double*)nullptr);
        else if (kernel == 1) {
            need(c.mixed() && c.setup_done, AMGS_ERR_STATE, "FP32 operator needs mixed setup");
            hipLaunchKernelGGL((k_wilson<float, false>), grid, dim3(FT), 0, c.stream, g, float(c.P.m0 + 2.0),
                               (const c32*)c.U32.as<c32>(), (const c32*)c.oeta32.as<c32>(), c.opsi32.as<c32>(),
                               (size_t)0, (const float*)nullptr);
        } else
            throw ApiError(AMGS_ERR_ARG, "unknown kernel id");
    };
    launch();
    HIP_CHECK(hipEventRecord(c.ev0, c.stream));
    for (int i = 0; i < reps; i++) launch();
    HIP_CHECK(hipEventRecord(c.ev1, c.stream));
    HIP_CHECK(hipEventSynchronize(c.ev1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    *avg_ms = ms / reps;
    return AMGS_OK;
    API_END
}

extern "C" int amgs_kernel_timing(amgs_ctx* ctx, int kclass, int enable) {
    if (!ctx || kclass < 0 || kclass >= KT_CLASSES) return AMGS_ERR_ARG;
    API_BEGIN
    KTimer& k = ctx->c.kt[kclass];
    k.resolve(ctx->c.stream);
    k.reset();
    k.on = enable != 0;
    return AMGS_OK;
    API_END
}

extern "C" int amgs_kernel_stats_ex(amgs_ctx* ctx, int kclass, int tag, long long* launches, double* total_ms,
                                    double* bytes, double* flops) {
    if (!ctx || kclass < 0 || kclass >= KT_CLASSES || tag < -1 || tag >= KTimer::TAGS || !launches || !total_ms ||
        !bytes || !flops)
        return AMGS_ERR_ARG;
    API_BEGIN
    KTimer& k = ctx->c.kt[kclass];
    k.resolve(ctx->c.stream);
    if (tag < 0) {
        *launches = k.launches;
        *total_ms = k.ms;
        *bytes = k.bytes;
        *flops = k.flops;
    } else {
        *launches = k.t_launches[tag];
        *total_ms = k.t_ms[tag];
        *bytes = k.t_bytes[tag];
        *flops = 0;
    }
    return AMGS_OK;
    API_END
}

extern "C" int amgs_kernel_stats(amgs_ctx* ctx, int kclass, long long* launches, double* total_ms, double* bytes) {
    if (!ctx || kclass < 0 || kclass >= KT_CLASSES || !launches || !total_ms || !bytes) return AMGS_ERR_ARG;
    API_BEGIN
    KTimer& k = ctx->c.kt[kclass];
    k.resolve(ctx->c.stream);
    *launches = k.launches;
    *total_ms = k.ms;
    *bytes = k.bytes;
    return AMGS_OK;
    API_END
}

extern "C" int amgs_last_timings(amgs_ctx* ctx, double* ms3) {
    if (!ctx || !ms3) return AMGS_ERR_ARG;
    for (int i = 0; i < 3; i++) ms3[i] = ctx->c.last_ms[i];
    return AMGS_OK;
}
