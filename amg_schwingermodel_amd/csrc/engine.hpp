// Host-side state of libamgs shared by its translation units: error types,
// device buffers, per-launch kernel timers, the per-precision engine (fine
// SAP, K-cycle at level 0, coarse-kernel arguments) and the context.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/amgs.h"
#include "coarse.hpp"
#include "common.hpp"
#include "layout.hpp"

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t e, const char* what, const char* file, int line)
        : std::runtime_error(std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")"),
          code(e) {}
};
struct ApiError : std::runtime_error {
    int code;
    ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

namespace amgs {

// ---------------------------------------------------------------------------
// device buffer
// ---------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t b) {
        if (b <= bytes && p) return;
        release();
        HIP_CHECK(hipMalloc(&p, std::max<size_t>(b, 256)));
        bytes = std::max<size_t>(b, 256);
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

static inline unsigned cdiv_u(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

struct Ctx;

// ---------------------------------------------------------------------------
// Per-launch timing of one kernel class inside solves: a HIP event pair on the
// library stream recorded by every launch's own dispatch (hipExtLaunchKernel:
// the pair brackets the kernel, not the gaps around it), and a device counter
// of the items that did work in that launch (for the algorithmic byte count).
// ---------------------------------------------------------------------------
struct KTimer {
    static constexpr int CAP = 8192;  // launches recorded between two resolves
    static constexpr int TAGS = 16;   // per-tag totals (the SAP step's Arnoldi index J)
    bool on = false;
    std::vector<hipEvent_t> ev;
    std::vector<double> fixed, per_item;  // bytes = fixed + per_item * active items
    std::vector<int> tag;
    DevBuf actlog, flog;  // per launch: active items; algorithmic flops (kernels that count them)
    long long launches = 0;
    double ms = 0, bytes = 0, flops = 0;
    long long t_launches[TAGS] = {};
    double t_ms[TAGS] = {}, t_bytes[TAGS] = {};

    ~KTimer() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    void reset() {
        launches = 0;
        ms = bytes = flops = 0;
        for (int t = 0; t < TAGS; t++) {
            t_launches[t] = 0;
            t_ms[t] = t_bytes[t] = 0;
        }
        fixed.clear();
        per_item.clear();
        tag.clear();
    }
    // returns the counter for this launch, or nullptr when not recording
    int* begin(hipStream_t s, double fx, double pi, int tg = 0) {
        if (!on || (int)fixed.size() >= CAP) return nullptr;
        if (!actlog.p) {
            actlog.alloc(sizeof(int) * CAP);
            HIP_CHECK(hipMemset(actlog.p, 0, sizeof(int) * CAP));
            flog.alloc(sizeof(unsigned long long) * CAP);
            HIP_CHECK(hipMemset(flog.p, 0, sizeof(unsigned long long) * CAP));
        }
        const size_t i = fixed.size();
        while (ev.size() < 2 * i + 2) {
            hipEvent_t e;
            HIP_CHECK(hipEventCreate(&e));
            ev.push_back(e);
        }
        fixed.push_back(fx);
        per_item.push_back(pi);
        tag.push_back(tg < 0 || tg >= TAGS ? 0 : tg);
        return actlog.as<int>() + i;
    }
    // the flop counter of the launch that the last begin() opened
    unsigned long long* flop_counter() const { return flog.as<unsigned long long>() + (fixed.size() - 1); }
    // the start and stop events of the launch that the last begin() opened
    hipEvent_t start() const { return ev[2 * (fixed.size() - 1)]; }
    hipEvent_t stop() const { return ev[2 * (fixed.size() - 1) + 1]; }
    // fold the pending records into the totals (synchronises the stream)
    void resolve(hipStream_t s) {
        if (fixed.empty()) return;
        HIP_CHECK(hipStreamSynchronize(s));
        std::vector<int> act(fixed.size());
        std::vector<unsigned long long> fl(fixed.size());
        HIP_CHECK(hipMemcpy(act.data(), actlog.p, sizeof(int) * act.size(), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(fl.data(), flog.p, sizeof(unsigned long long) * fl.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < fixed.size(); i++) {
            float t = 0;
            HIP_CHECK(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
            const double b = act[i] > 0 ? fixed[i] + per_item[i] * act[i] : 0.0;
            ms += t;
            launches++;
            bytes += b;
            flops += (double)fl[i];
            t_launches[tag[i]]++;
            t_ms[tag[i]] += t;
            t_bytes[tag[i]] += b;
        }
        HIP_CHECK(hipMemset(actlog.p, 0, sizeof(int) * CAP));
        HIP_CHECK(hipMemset(flog.p, 0, sizeof(unsigned long long) * CAP));
        fixed.clear();
        per_item.clear();
        tag.clear();
    }
};
enum { KT_SAP_A = 0, KT_SAP_B = 1, KT_COARSE = 2, KT_SAP_BLOCK = 3, KT_CLASSES = 4 };

// ---------------------------------------------------------------------------
// per-precision engine: fine SAP, K-cycle(0), coarse kernel arguments
// ---------------------------------------------------------------------------
template <typename R> struct Engine {
    Ctx* c = nullptr;
    int cap = 0;
    // hierarchy in precision R
    const cplx<R>* U = nullptr;
    const cplx<R>* tv[MAXL] = {};
    const cplx<R>* links[MAXL] = {};
    const cplx<R>* links2[MAXL] = {};  // column-pair copies for the coarse SAP block stencil (FP32 engine)
    // several gauge configurations (rhs r = configuration r): entries between
    // consecutive configurations' links U, test vectors tv[l] and coarse links
    // links[l] (0 with one configuration), and m0 + 2 per configuration
    size_t ucs = 0, tvcs[MAXL] = {}, lkcs[MAXL] = {};
    const R* mmv = nullptr;
    // per-rhs flags of the outer solve that calls the K-cycle (0: that rhs has
    // converged, its preconditioner work is skipped); null outside it
    const int* rmask = nullptr;
    // fine SAP workspace
    DevBuf Vt, Vh, xl, r, st_mem, skip, part_r, part_v, zero, rsig;
    SapState st{};
    int nchunk_max = 0;
    // K-cycle(0) temporaries
    DevBuf eta1, psi1, res0;
    // coarse workspace; the kernel arguments (ca + per-launch fields) travel by value
    DevBuf cws, prof, bar, xchg;
    CArgs<R> ca{};
    // workgroups per right-hand side of the coarse K-cycle launches: at most
    // grp_max; the autotuned value per batch size, else the default
    int grp_max = 1;
    std::vector<std::pair<int, int>> coarse_tuned;
    int coarse_grp(int nrhs) const {
        for (const auto& t : coarse_tuned)
            if (t.first == nrhs) return std::min(t.second, grp_max);
        const char* e = std::getenv("AMGS_COARSE_GRP");
        const int v = e ? std::atoi(e) : 0;
        if (v > 0) return std::min(v, grp_max);
        // default: fill the chip's 256 compute units, one workgroup each
        return std::max(1, std::min(grp_max, 256 / std::max(nrhs, 1)));
    }

    void init(Ctx* ctx, int capacity);
    void sap0(int nrhs, const cplx<R>* v, cplx<R>* x, int nu);
    void kcycle0(int nrhs, const cplx<R>* eta, cplx<R>* psi, bool from_zero = true);
    // base: kernel arguments to start from instead of ca (a caller-provided
    // workspace layout); the batch is then limited by that workspace, not cap
    void coarse(int mode, int level, int nrhs, const cplx<R>* in, size_t in_stride, cplx<R>* out, size_t out_stride,
                int nu = 0, const CArgs<R>* base = nullptr);
    void restrict0(int nrhs, const cplx<R>* w, const cplx<R>* v, cplx<R>* out);
    // level-0 SAP launch shape chosen by the autotuner, per batch size
    std::vector<std::pair<int, int>> sap_tuned;
    // wavefronts per workgroup of the batched coarse operator (k_cdop_mfma),
    // per (level, batch size): the autotuned value, else 4
    std::vector<std::pair<long long, int>> cdop_tuned;
    int cdop_wpg(int level, int nrhs) const {
        const long long key = (long long)level << 32 | (unsigned)nrhs;
        for (const auto& t : cdop_tuned)
            if (t.first == key) return t.second;
        return 4;
    }
    int sap_wgs(int nrhs) const;
};

struct Ctx {
    amgs_params P{};
    int L = 0;
    Geo geo[MAXL]{};
    std::vector<int> red[MAXL], black[MAXL];
    hipStream_t stream = nullptr;
    int device = 0;
    std::string err;
    bool gauge_set = false, setup_done = false;
    // FP64 hierarchy (always), FP32 copy in mixed mode
    DevBuf U64, U32, tv64[MAXL], tv32[MAXL], links64[MAXL], links32[MAXL], links2[MAXL], redl[MAXL], blackl[MAXL];
    DevBuf colmask;  // level-0 Schwarz block flags [2][nsb]: red, black
    Engine<double> e64;
    Engine<float> e32;
    // outer FGMRES
    DevBuf oV, oZ, oZ32, ox, ob, or_, ow, ost_mem, oeta32, opsi32, rel_part;
    DevBuf cerr;  // raised by a coarse-kernel member barrier that timed out
    // several gauge configurations of one lattice shape (amgs_create_multi):
    // ncfg hierarchies side by side, right-hand side r solved with configuration r
    int ncfg = 1;
    std::vector<double> m0cfg;
    DevBuf mm64, mm32;  // m0 + 2 per configuration
    OuterState ost{};
    int ocap = 0;
    // io staging
    DevBuf io_a, io_b;
    // comparison solvers: Krylov vectors, partial sums
    DevBuf kbuf, kpart, ksum;
    int outer_prec = 0;  // outer FGMRES preconditioner: 0 multigrid cycle, 1 one SAP sweep, 2 none
    double last_ms[3] = {0, 0, 0};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // pinned copy of the outer solve's per-rhs active flags and the event the
    // host spins on for it (read once per outer Arnoldi step)
    int* hflags = nullptr;
    int hflags_cap = 0;
    hipEvent_t evf = nullptr;
    // two pinned chunks that stage host-buffer solves (allocated on first use)
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    KTimer kt[KT_CLASSES];
    // HIP-graph replay of the outer FGMRES restart cycle (m Arnoldi steps with
    // their preconditioner applications, the update and the new residual):
    // one executable graph per batch size
    bool graph = false;
    struct PGraph {
        int nrhs, outer_prec;
        hipGraphExec_t exec;
    };
    std::vector<PGraph> pgraphs;
    void clear_graphs() {
        for (auto& g : pgraphs) (void)hipGraphExecDestroy(g.exec);
        pgraphs.clear();
    }

    size_t nvec(int l) const { return (size_t)geo[l].ns * geo[l].dof; }
    // FP32 inner levels (AMGS_PREC_MIXED, AMGS_PREC_MIXED16)
    bool mixed() const { return P.precision != AMGS_PREC_FP64; }
    // FP16 Krylov basis in the level-0 SAP block GMRES (AMGS_PREC_MIXED16)
    bool h16() const { return P.precision == AMGS_PREC_MIXED16; }
    void sync() { HIP_CHECK(hipStreamSynchronize(stream)); }
};

}  // namespace amgs
