// Host launcher of the persistent coarse-level kernel (levels >= 1), kept in
// translation units of its own (one per precision) so that the large kernel
// compiles in parallel with the rest of the library and only when it changes.
#pragma once
#include <hip/hip_ext.h>

#include "engine.hpp"

namespace amgs {

template <typename R>
void Engine<R>::coarse(int mode, int level, int nrhs, const cplx<R>* in, size_t in_stride, cplx<R>* out,
                       size_t out_stride, int nu, const CArgs<R>* base) {
    if (!base && nrhs > cap) throw ApiError(AMGS_ERR_ARG, "batch exceeds engine capacity");
    static_assert(sizeof(CArgs<R>) <= 4096, "coarse-kernel arguments exceed the kernel-argument segment");
    CArgs<R> a = base ? *base : ca;
    a.mm = R(c->P.m0 + 2.0);
    for (int l = 1; l < c->L; l++) {
        a.lv[l].links = links[l];
        a.lv[l].links2 = links2[l];
        a.lv[l].w = l < c->L - 1 ? tv[l] : nullptr;
        a.lcs[l] = lkcs[l];
        a.wcs[l] = tvcs[l];
    }
    a.mmv = mmv;
    a.rmask = rmask;
    a.mode = mode;
    a.level = level;
    a.nu = nu;
    // coarsest GMRES vectors in dynamic LDS if they fit beside the static part
    const size_t ncrs = c->nvec(c->L - 1);
    const size_t dyn = (size_t)(a.crs_m + 4) * ncrs * sizeof(cplx<R>);
    // (the kernel's static LDS: CLds and its copy of the arguments)
    const size_t lds_cap = 160 * 1024 - sizeof(CLds) - 1024 - sizeof(CArgs<R>);
    a.crs_lds = dyn <= lds_cap ? 1 : 0;
    size_t dyn_bytes = a.crs_lds ? dyn : 0;
    // SAP block vectors of a level share that region (never live at the same time)
    for (int l = 0; l < MAXL; l++) {
        const size_t need = l >= 1 && l < c->L - 1 ? CW * a.lv[l].blkStride * sizeof(cplx<R>) : 0;
        a.sap_lds[l] = need > 0 && need <= lds_cap ? 1 : 0;
        if (a.sap_lds[l]) dyn_bytes = std::max(dyn_bytes, need);
    }
    a.prof = prof.p ? prof.as<unsigned long long>() : nullptr;
    a.in = in;
    a.out = out;
    a.in_stride = in_stride;
    a.out_stride = out_stride;
    // the K-cycle / cycle solves split their SAP blocks over grp workgroups per rhs
    a.grp = (!base && (mode == CM_KC_SOLVE || mode == CM_CYCLE)) ? coarse_grp(nrhs) : 1;
    if (a.grp > 1) {
        hipLaunchKernelGGL(k_member_reset, dim3(cdiv_u(nrhs, 64)), dim3(64), 0, c->stream, nrhs, a.bar);
        HIP_CHECK(hipGetLastError());
    }
    // FP32, one Schwarz block of a colour per member, 2x2-site blocks on a
    // square level: the member's whole workgroup solves its block with the
    // block's links staged in LDS
    for (int l = 0; l < MAXL; l++) {
        a.sap_wg[l] = 0;
        if (!std::is_same<R, float>::value || l < 1 || l >= c->L - 1 || !links2[l]) continue;
        const Geo& g = c->geo[l];
        const size_t need = wg_block_lds_bytes(g.dof, a.sap_m);
        if (a.grp >= g.ncol && g.sxel == 2 && g.stel == 2 && g.nx == g.nt && g.nx > 2 && g.dof % 2 == 0 &&
            12 * g.dof <= CT && need <= lds_cap) {
            a.sap_wg[l] = 1;
            dyn_bytes = std::max(dyn_bytes, need);
        }
    }
    if (dyn_bytes > 0) {
        static bool attr_set[2] = {false, false};
        bool& done = attr_set[sizeof(R) == 8];
        if (!done) {
            HIP_CHECK(hipFuncSetAttribute((const void*)k_coarse<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)(160 * 1024 - sizeof(CLds) - sizeof(CArgs<R>) - 1024)));
            done = true;
        }
    }
    // per-launch timing (kernel class KT_COARSE): events of the launch's own
    // dispatch, and the stencil flops the kernel counts (member 0 of each rhs)
    KTimer& kt = c->kt[KT_COARSE];
    a.cnt = nullptr;
    if (kt.begin(c->stream, 0.0, 0.0)) {
        a.cnt = kt.flop_counter();
        hipExtLaunchKernelGGL(k_coarse<R>, dim3(nrhs * a.grp), dim3(CT), dyn_bytes, c->stream, kt.start(), kt.stop(), 0,
                              a);
    } else {
        hipLaunchKernelGGL(k_coarse<R>, dim3(nrhs * a.grp), dim3(CT), dyn_bytes, c->stream, a);
    }
    HIP_CHECK(hipGetLastError());
}

}  // namespace amgs
