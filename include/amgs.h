/*
 * amgs.h - C ABI of libamgs, the MI355X-native DD-alpha-AMG solver for the
 * 2-D Schwinger-model Wilson-Dirac operator.
 *
 * This is the drop-in boundary for the reference's solve path
 * (Fabian2598/AMG_SchwingerModel): FGMRES with a K-cycle multigrid
 * preconditioner (src/fgmres.cpp:5-88, include/amg.h:99-145, src/amg.cpp:3-178).
 * The reference has no FFI; its in-process API is the C++ class contract
 * FGMRES_AMG(dim1, dim2, m, restarts, tol, GaugeConf, m0) + fgmres(b, x0, x)
 * and its file surface (parameters.dat, x t mu re im gauge/RHS files,
 * Residuals_FGMRES_*.res).  Each entry point below names what it replaces.
 *
 * Conventions
 *   - complex numbers are interleaved (re, im) doubles;
 *   - a level-l vector is [nsites_l][dof_l], site n = x*nt_l + t,
 *     dof = 2*colour + spin (src/level.cpp:166); the fine level has dof 2;
 *   - gauge links U[n][mu], mu = 0 time, mu = 1 space (include/boundary.h:18-21);
 *   - batched vectors are [nrhs][nsites][dof] (right-hand side outermost);
 *   - host pointers unless the name ends in _dev (device pointers on the
 *     context's HIP device, visible to the context's stream);
 *   - return value: AMGS_OK (0), AMGS_NOT_CONVERGED (1, results still
 *     written) or a negative error code; amgs_last_error() describes it.
 *   - one context = one hierarchy, one HIP stream, no global state; calls on
 *     one context must not overlap.
 */
#ifndef AMGS_H
#define AMGS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMGS_OK 0
#define AMGS_NOT_CONVERGED 1
#define AMGS_ERR_ARG (-1)
#define AMGS_ERR_GEOMETRY (-2)
#define AMGS_ERR_HIP (-3)
#define AMGS_ERR_STATE (-4)
#define AMGS_ERR_IO (-5)
#define AMGS_ERR_NAN (-6)

#define AMGS_MAX_LEVELS 6

/* Arithmetic of the inner levels.  The outer FGMRES, its Hessenberg
 * least-squares problem and the true-residual check are FP64 in every mode. */
#define AMGS_PREC_FP64 0   /* plain FP64 everywhere: the drop-in */
#define AMGS_PREC_MIXED 1  /* SAP, P/P^dagger and coarse levels in FP32 */
#define AMGS_PREC_MIXED16 2 /* as MIXED, and the level-0 SAP block GMRES keeps its
                             * Krylov basis in FP16 with a power-of-two scale
                             * (FP32 arithmetic, FP64 scalars) */

typedef struct {
    /* lattice and hierarchy: CMakeLists.txt:22-27 (NS, NT, LEVELS) and the
     * parameters.dat rows "level bx bt ntest sapx sapt" (block COUNTS,
     * include/params.h:30-46) for levels 0..levels-2 */
    int nx, nt, levels;
    int bx[AMGS_MAX_LEVELS], bt[AMGS_MAX_LEVELS], ntest[AMGS_MAX_LEVELS];
    int sapx[AMGS_MAX_LEVELS], sapt[AMGS_MAX_LEVELS];
    double m0;                                       /* mass::m0, src/main.cpp:64 */
    int nu1, nu2;                                    /* src/variables.cpp:63-64   */
    int sap_m, sap_restarts; double sap_gmres_tol;   /* src/variables.cpp:47-49   */
    double sap_tol;                                  /* src/variables.cpp:50      */
    int kc_m, kc_restarts; double kc_tol;            /* src/variables.cpp:70-72   */
    int crs_m, crs_restarts; double crs_tol;         /* include/params.h:63-65    */
    int out_m, out_restarts; double out_tol;         /* src/variables.cpp:79-81   */
    int tv_sap_iters;                                /* src/main.cpp:63           */
    uint32_t tv_seed;                                /* src/amg.cpp:8 (50)        */
    int precision;                                   /* AMGS_PREC_*               */
    int max_rhs;                                     /* batch capacity (>= 1)     */
    int cycle;                                       /* 1 K-cycle, 0 V-cycle (src/main.cpp:58) */
    int nit;                                         /* bootstrap iterations (src/amg.cpp:52-84) */
} amgs_params;

typedef struct amgs_ctx amgs_ctx;

/* Reference defaults (src/variables.cpp:46-87) with main()'s overrides
 * (src/main.cpp:61-64,129: K-cycle, 4 SAP sweeps on test vectors, Nit = 0). */
int amgs_default_params(amgs_params* p);

/* Reads the 6-column parameters.dat (include/params.h:17-68) into p->rows and
 * sets p->levels = rows + 1.  nx, nt are not touched. */
int amgs_read_parameters(const char* path, amgs_params* p);

/* Replaces the FGMRES_AMG constructor's allocation part (include/amg.h:101-104,
 * include/level.h:91-134).  Validates the blocking; allocates device memory. */
int amgs_create(const amgs_params* p, amgs_ctx** out);
/* Several gauge configurations of one lattice shape and blocking in one
 * context (new; the reference solves one configuration per process,
 * src/main.cpp:82-89): ncfg configurations, each with its own bare mass m0[i],
 * gauge field and multigrid hierarchy (set up as a context of its own would
 * be), share every kernel launch.  Right-hand side i of a batch belongs to
 * configuration i: solves and forces take exactly ncfg right-hand sides
 * (p->max_rhs is set to ncfg), amgs_set_gauge takes U [ncfg][nx*nt][2], and
 * amgs_apply_D at level 0 applies configuration i's D to vector i.  The
 * component entry points of single hierarchies (coarse levels, P, SAP, the
 * comparison solvers) return AMGS_ERR_STATE on such a context. */
int amgs_create_multi(const amgs_params* p, int ncfg, const double* m0, amgs_ctx** out);
int amgs_num_configs(const amgs_ctx* c);
void amgs_destroy(amgs_ctx* c);
const char* amgs_last_error(const amgs_ctx* c);

/* GaugeConf::read_conf / setGconf (src/gauge_conf.cpp:18-32): U [nx*nt][2]
 * ([ncfg][nx*nt][2] on a multi-configuration context). */
int amgs_set_gauge(amgs_ctx* c, const double* U);

/* AlgebraicMG::setUpPhase(Nit = 0) (src/amg.cpp:3-46): test vectors from
 * mt19937(tv_seed), SAP smoothing, aggregate orthonormalisation, Galerkin
 * coarse links, on the GPU. */
int amgs_setup(amgs_ctx* c);

/* Setup from caller-given raw test vectors tv[l] = [ntest_l][nsites_l][dof_l]
 * (levels 0..levels-2), bypassing the generator; smooth = 0 skips SAP. */
int amgs_setup_from_tv(amgs_ctx* c, const double* const* tv, int smooth);

/* FGMRES_AMG::fgmres(b, x0 = 0, x) (src/fgmres.cpp:5-88) for nrhs right-hand
 * sides at once; every right-hand side follows its own iteration exactly as a
 * single solve would.  iters[nrhs] receives the reference's count
 * (k*m + maxIt, or restarts*m); relres[nrhs] the FP64 true relative residual;
 * hist (may be NULL) [nrhs][hist_cap] the |g_{j+1}| history of each solve,
 * nhist[nrhs] its length (the Residuals_FGMRES_*.res content,
 * src/fgmres.cpp:65-79).  Returns AMGS_NOT_CONVERGED if any did not converge. */
int amgs_solve(amgs_ctx* c, int nrhs, const double* b, double* x, int* iters, double* relres,
               double* hist, int hist_cap, int* nhist);
int amgs_solve_dev(amgs_ctx* c, int nrhs, const void* b_dev, void* x_dev, int* iters, double* relres);

/* HIP-graph replay of the outer restart cycle (new): with enable != 0, later
 * solves capture one whole FGMRES restart cycle - its m Arnoldi steps with
 * their multigrid preconditioner applications (restriction, the persistent
 * coarse-level kernel, prolongation, SAP; several hundred launches each), the
 * update and the new residual - once per batch size into a HIP graph and
 * replay it in every cycle; the host waits once per cycle, never inside one.
 * The break test stays on the device: a right-hand side that converges inside
 * a cycle is skipped by every later kernel.  Results are identical to the
 * eager launches (which read the flags after every Arnoldi step and leave the
 * cycle early).  Graphs are rebuilt after a new set-up; solves with kernel
 * timing enabled (amgs_kernel_timing) launch eagerly. */
int amgs_set_graph(amgs_ctx* c, int enable);

/* Autotuner (new): times one level-0 SAP sweep for a set of launch shapes
 * (workgroups per launch, i.e. lattice chunks per Schwarz block and rhs) and
 * keeps the fastest for batches of nrhs right-hand sides in this context;
 * then times the coarse-level solve of one K-cycle for 1, 2, 4, 8 workgroups
 * per right-hand side (the coarse levels' Schwarz blocks of a colour split
 * over them) and keeps the fastest.  Results stay deterministic: the SAP
 * choice fixes the chunking (and with it the order of the per-chunk partial
 * sums) for all later solves of that size, and the coarse choice does not
 * change any arithmetic.  best_wgs (may be NULL) receives the workgroup
 * target, sweep_ms (may be NULL) the sweep time it measured.  Captured
 * graphs are dropped. */
int amgs_autotune(amgs_ctx* c, int nrhs, int* best_wgs, double* sweep_ms);
/* The launch shapes in force for batches of nrhs (new): level-0 SAP
 * workgroup target and coarse workgroups per right-hand side.  set: a value
 * > 0 fixes it, 0 restores the default, < 0 leaves it (drops captured graphs). */
int amgs_get_tuning(amgs_ctx* c, int nrhs, int* sap_wgs, int* coarse_grp);
int amgs_set_tuning(amgs_ctx* c, int nrhs, int sap_wgs, int coarse_grp);

/* Uneven multi-configuration batch (new; the reference solves one gauge
 * configuration per process): the solves of nctx independent contexts - each
 * with its own lattice size, blocking, mass and gauge field, set up
 * beforehand - run concurrently, one host thread per context, their kernels
 * overlapping on the device.  Context i solves nrhs[i] right-hand sides from
 * device buffers b_dev[i] into x_dev[i] exactly as amgs_solve_dev would;
 * iters and relres are concatenated over the contexts in order, status[i] is
 * that context's return value.  A context may appear only once.  Returns the
 * first negative status, else AMGS_NOT_CONVERGED if any did not converge,
 * else AMGS_OK. */
int amgs_solve_multi(int nctx, amgs_ctx* const* ctxs, const int* nrhs, const void* const* b_dev, void* const* x_dev,
                     int* iters, double* relres, int* status);

/* amgs_solve_multi with staggered starts (new, a scheduling form with no
 * reference counterpart): context i's host thread waits start_ms[i]
 * milliseconds (0 to 1000; a null start_ms means no wait) before its first
 * launch.  Equal-sized sub-batches of one workload otherwise run their
 * phases in lockstep, so that their bandwidth-bound passes (the outer Arnoldi
 * steps) contend with each other while the latency-bound coarse kernels leave
 * the memory idle; an offset of a few milliseconds interleaves them.  The
 * results do not depend on the offsets (the contexts share no data). */
int amgs_solve_multi_staggered(int nctx, amgs_ctx* const* ctxs, const int* nrhs, const void* const* b_dev,
                               void* const* x_dev, const double* start_ms, int* iters, double* relres, int* status);

/* ---- component entry points (the reference's building blocks, used by the
 *      parity tests and by callers that assemble their own solvers) ---- */

/* D_phi / D_dagger_phi (src/dirac_operator.cpp:3-54) at level 0 and
 * Level::D_operator (src/level.cpp:161-180) at level >= 1 (dagger must be 0). */
int amgs_apply_D(amgs_ctx* c, int level, int dagger, int nrhs, const double* in, double* out);
int amgs_apply_D_dev(amgs_ctx* c, int level, int dagger, int nrhs, const void* in_dev, void* out_dev);
/* Level::D_operator (src/level.cpp:161-180) at a coarse level (1 .. levels-1)
 * on nrhs device vectors [nrhs][nsites][dof] at once, as matrix-core
 * block-GEMMs (new): per coarse site the five dof x dof link blocks against
 * the five neighbour vectors of all right-hand sides, in 16x16x4 MFMA tiles
 * (16 rhs x 16 rows x 4 columns; FP64 MFMA for precision AMGS_PREC_FP64, FP32
 * MFMA on the FP32 hierarchy of a mixed-precision context otherwise).
 * amgs_apply_D at coarse levels runs the FP64 form. */
int amgs_apply_D_batched_dev(amgs_ctx* c, int level, int nrhs, int precision, const void* in_dev, void* out_dev);
/* Tile autotuner of amgs_apply_D_batched_dev (new): times 1, 2, 4, 8
 * wavefronts (one tile each) per workgroup for this level and batch size and
 * keeps the fastest; results do not depend on the choice.  best_wpg, best_us
 * (may be NULL): the choice and its time per apply in microseconds. */
int amgs_autotune_batched_D(amgs_ctx* c, int level, int nrhs, int precision, int* best_wpg, double* best_us);
/* Level::P_v and Level::Pt_v (src/level.cpp:182-237), level = the finer one. */
int amgs_apply_P(amgs_ctx* c, int level, int nrhs, const double* coarse, double* fine);
int amgs_apply_Pt(amgs_ctx* c, int level, int nrhs, const double* fine, double* coarse);
/* SAP_C::SAP(v, x, nu) (src/sap.cpp:64-174); x is updated in place. */
int amgs_sap(amgs_ctx* c, int level, int nrhs, const double* v, double* x, int nu);
/* SAP_C::SAP(v, x, nu) at level 0 (src/sap.cpp:64-174) in the inner precision
 * of a mixed-precision context (FP32 arithmetic; FP16 Krylov basis under
 * AMGS_PREC_MIXED16), the smoother its solves run, in the launch form the
 * tuning selects; host buffers, x updated in place.  AMGS_ERR_STATE for an
 * FP64 context or before amgs_setup. */
int amgs_sap_inner(amgs_ctx* c, int nrhs, const double* v, double* x, int nu);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): amgs_sap at a
 * coarse level (1 .. levels-2) on device buffers, with the block solves'
 * workspace in a static device array of AMGS_NOALLOC_SCRATCH_BYTES per GPU
 * instead of allocated memory; the call allocates nothing.  The workspace
 * holds 16 * (2 n + 8 * (sap_m + 3) * sspb * dof) bytes per right-hand side
 * (n = nsites * dof of the level, sspb = sites per Schwarz block); larger
 * batches return AMGS_ERR_ARG.  Calls that overlap on one GPU share the array
 * and race.  Results equal amgs_sap's. */
#define AMGS_NOALLOC_SCRATCH_BYTES (8 << 20)
int amgs_sap_dev_noalloc(amgs_ctx* c, int level, int nrhs, const void* v_dev, void* x_dev, int nu);
/* One preconditioner application psi = K(eta) from psi = 0 (include/amg.h:134-144):
 * the K-cycle (src/amg.cpp:135-178), or the V-cycle (src/amg.cpp:90-132) when
 * params.cycle == 0. */
int amgs_kcycle(amgs_ctx* c, int nrhs, const double* eta, double* psi);
/* AlgebraicMG::applyMultilevel(it, rhs, out, tol) (src/amg.cpp:181-231): the
 * stand-alone multigrid x <- cycle(b, x) (K or V per params.cycle, FP64) until
 * |b - D x| < tol |b| or max_cycles.  x holds the initial guess on entry.
 * cycles[r] = cycles used, or -max_cycles if right-hand side r did not
 * converge (then AMGS_NOT_CONVERGED); relres[r] = its relative residual. */
int amgs_multigrid(amgs_ctx* c, int nrhs, const double* b, double* x, int max_cycles, double tol, int* cycles,
                   double* relres);
/* ---- comparison solvers on the fine lattice (FP64; nrhs <= min(max_rhs, 32);
 *      they need amgs_set_gauge only, no set-up) ---- */
/* bi_cgstab(D_phi, ..., x0 = 0, max_iter, tol) (src/bi_cgstab.cpp:5-87):
 * iters[r] = k+1 at convergence or max_iter; hist [nrhs][hist_cap] gets |s| of
 * every iteration (the Residuals_BiCG file); relres = |b - D x| / |b|. */
int amgs_bicgstab(amgs_ctx* c, int nrhs, const double* b, double* x, int max_iter, double tol, int* iters,
                  double* relres, double* hist, int hist_cap, int* nhist);
/* conjugate_gradient(U, phi, x, m0) on D D^dagger from x0 = phi
 * (src/conjugate_gradient.cpp:5-65): iters[r] = loop passes at convergence
 * (the reference prints this minus one) or -max_iter; relres = |phi - D D^dagger x| / |phi|. */
int amgs_cg_normal(amgs_ctx* c, int nrhs, const double* phi, double* x, int max_iter, double tol, int* iters,
                   double* relres);
/* FGMRES_SAP::fgmres(b, 0, x) (include/sap.h:194-225): outer FGMRES(out_m,
 * out_restarts, out_tol) with one SAP sweep (level-0 Schwarz blocks) as
 * preconditioner.  Same outputs as amgs_solve. */
int amgs_fgmres_sap(amgs_ctx* c, int nrhs, const double* b, double* x, int* iters, double* relres, double* hist,
                    int hist_cap, int* nhist);
/* FGMRES_fine_level::fgmres(b, 0, x) (include/fgmres.h:126-148): GMRES on D
 * without preconditioning (out_m, out_restarts, out_tol). */
int amgs_gmres(amgs_ctx* c, int nrhs, const double* b, double* x, int* iters, double* relres, double* hist,
               int hist_cap, int* nhist);

/* ---- fermion force (new; no counterpart in the reference, which stops at
 *      the solve; the D^dagger it needs is D_dagger_phi, src/dirac_operator.cpp:30-54) ----
 * The pseudofermion action and its derivative dS/dtheta_mu(n) for
 * U_mu(n) = exp(i theta_mu(n)), by the adjoint method (two multigrid solves):
 *   form AMGS_FORCE_DDAG_D: S = phi^dagger (D^dagger D)^-1 phi,
 *   form AMGS_FORCE_D_DDAG: S = phi^dagger (D D^dagger)^-1 phi
 *     = |Y|^2 with Y = D^-1 phi, X = (D^dagger)^-1 Y, dS/dtheta = -2 Re X^dagger (dD/dtheta) Y.
 * Since D^dagger = s3 D s3, (D^dagger D)^-1 = s3 (D D^dagger)^-1 s3: the first
 * form is the second at s3 phi.  phi [nrhs][nx*nt][2] complex; dsdtheta
 * [nrhs][nx*nt][2] real (mu = 0 time, 1 space); action[nrhs].
 * Needs amgs_setup; nrhs <= min(max_rhs, 32). */
#define AMGS_FORCE_DDAG_D 0
#define AMGS_FORCE_D_DDAG 1
int amgs_force(amgs_ctx* c, int nrhs, int form, const double* phi, double* dsdtheta, double* action);

/* Level::interpolator_columns and the coarse links G1/G2/G3 (include/level.h:143-164).
 * tv: [nsites_l][dof_l] of test vector i;  which 0/1/2 = A/B_mu/C_mu as
 * [nsites_l][dof_l][dof_l] (row = 2c+alpha, col = 2b+beta). */
int amgs_get_tv(amgs_ctx* c, int level, int i, double* out);
int amgs_get_links(amgs_ctx* c, int level, int which, int mu, double* out);

/* ---- workload generators (new: the reference only reads files) ---- */
/* Quenched U(1) plaquette-action Metropolis from a cold start; deterministic
 * in (nx, nt, beta, sweeps, seed); writes U [nx*nt][2]; *plaq = <Re U_p>. */
int amgs_gauge_metropolis(int nx, int nt, double beta, int sweeps, uint64_t seed, double* U, double* plaq);
/* Gaussian complex vector: re, im ~ N(0, 1), deterministic in (n, seed). */
int amgs_gaussian(int64_t ncomplex, uint64_t seed, double* out);

/* ---- file formats (src/gauge_conf.cpp:25-28, src/variables.cpp:117-155, 101-115) ---- */
int amgs_read_conf(const char* path, int nx, int nt, double* U);
int amgs_write_conf(const char* path, int nx, int nt, const double* U);
int amgs_read_rhs(const char* path, int nx, int nt, double* b);
int amgs_write_rhs(const char* path, int nx, int nt, const double* b);
int amgs_write_residuals(const char* path, const double* hist, int n);

/* ---- measurement ---- */
/* Average device time (ms) of `reps` back-to-back launches of one kernel on
 * the context's stream, bracketed by HIP events on that stream.
 * kernel: 0 = fine D (FP64), 1 = fine D (FP32, needs a mixed-precision
 * setup); other ids return AMGS_ERR_ARG (the level-0 SAP sweep is timed by
 * amgs_autotune).  Inputs are whatever the context holds. */
int amgs_time_kernel(amgs_ctx* c, int kernel, int nrhs, int reps, double* avg_ms);
/* Phase timers of the last solve (ms): [0] total, [1] setup, [2] solve. */
int amgs_last_timings(amgs_ctx* c, double* ms3);
/* Per-launch timing of one kernel class during later solves: a HIP event
 * pair on the context's stream around every launch of the class, plus the
 * algorithmic bytes of the items that did work in that launch.
 * kclass: 0 = SAP Arnoldi step A (block stencil + dot products),
 *         1 = SAP Arnoldi step B (orthogonalisation + norm),
 *         2 = the persistent coarse-level kernel (its algorithmic stencil
 *             flops: 8 per complex multiply-add of every level operator and
 *             Schwarz-block stencil it applies, counted once per rhs).
 * enable != 0 starts recording (and clears the totals), 0 stops it.
 * amgs_kernel_stats returns the totals: launches, summed device time (ms)
 * and summed algorithmic bytes; amgs_kernel_stats_ex also the flops, and
 * with tag >= 0 the totals of the launches with that tag (step A: the
 * Arnoldi index J). */
int amgs_kernel_timing(amgs_ctx* c, int kclass, int enable);
int amgs_kernel_stats(amgs_ctx* c, int kclass, long long* launches, double* total_ms, double* bytes);
int amgs_kernel_stats_ex(amgs_ctx* c, int kclass, int tag, long long* launches, double* total_ms, double* bytes,
                         double* flops);

#ifdef __cplusplus
}
#endif
#endif
