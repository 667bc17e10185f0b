/*
 * CPU oracle for the DD-alpha-AMG Wilson-Dirac solve (2-D Schwinger model).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C11 restatement of the reference
 * algorithm (Fabian2598/AMG_SchwingerModel) in FP64 complex arithmetic, used by
 * tests/ and by bench.py's CPU-baseline leg as the checker.  The product
 * library (libamgs) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Pinning:
 *   - all functions are compared with golden vectors produced by the compiled
 *     reference (tests/golden/, 256x256, 4 levels: the shipped config.h);
 *   - the Dirac operator additionally against closed forms (free-field
 *     spectrum, gamma5-hermiticity, gauge covariance) and a dense brute force;
 *   - P/P^dagger, orthonormalisation and coarse links against P^H P = 1 and the
 *     Galerkin identity P^H D_l P = D_{l+1}.
 *
 * Storage: a level-l vector is complex128[nsites_l][dof_l], site n = x*nt_l + t,
 * dof = 2*colour + spin.  Gauge links U[n][mu], mu = 0 (time), 1 (space).
 */
#ifndef AMGS_ORACLE_H
#define AMGS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXL 8

typedef struct {
    int nx, nt, levels;               /* levels includes finest and coarsest       */
    /* parameters.dat rows for l < levels-1 (block COUNTS, include/params.h:30-46) */
    int bx[OR_MAXL], bt[OR_MAXL], ntest[OR_MAXL], sapx[OR_MAXL], sapt[OR_MAXL];
    double m0;
    int nu1, nu2;                                   /* src/variables.cpp:63-64   */
    int sap_m, sap_restarts; double sap_gmres_tol;  /* src/variables.cpp:47-49   */
    double sap_tol;                                 /* src/variables.cpp:50      */
    int kc_m, kc_restarts; double kc_tol;           /* src/variables.cpp:70-72   */
    int crs_m, crs_restarts; double crs_tol;        /* include/params.h:63-65    */
    int out_m, out_restarts; double out_tol;        /* src/variables.cpp:79-81   */
    int tv_sap_iters;                               /* src/main.cpp:63           */
    unsigned seed;                                  /* src/amg.cpp:8             */
    int threads;                                    /* OpenMP threads (0: default) */
    int cycle;                                      /* 1 K-cycle, 0 V-cycle (src/main.cpp:58) */
    int nit;                                        /* bootstrap iterations (src/amg.cpp:52-84) */
} or_params;

typedef struct or_ctx or_ctx;

/* Fills p with the reference defaults and main()'s overrides. */
void or_default_params(or_params* p);

/* Returns NULL on inconsistent geometry (indivisible blocking, odd counts). */
or_ctx* or_create(const or_params* p, const double* U /* complex [nx*nt][2] */);
void    or_destroy(or_ctx* c);

int  or_nsites(const or_ctx* c, int l);
int  or_dof(const or_ctx* c, int l);

/* Fine operators in the D_phi form (src/dirac_operator.cpp:3-54). */
void or_dphi(or_ctx* c, const double* in, double* out);
void or_ddag(or_ctx* c, const double* in, double* out);

/* Level operators (src/level.cpp:161-237). */
void or_dop(or_ctx* c, int l, const double* in, double* out);
void or_prolong(or_ctx* c, int l, const double* coarse, double* fine);
void or_restrict(or_ctx* c, int l, const double* fine, double* coarse);

/* SAP smoother (src/sap.cpp:64-174); returns 1 if converged. */
int  or_sap(or_ctx* c, int l, const double* v, double* x, int nu);

/* Setup (src/amg.cpp:3-46, Nit = 0).  draw_order 0: real part drawn first. */
void or_setup(or_ctx* c, int draw_order);
/* Setup from given level-l interpolator columns (bypasses the RNG); smooth=0
 * skips the SAP smoothing so the given columns are used as they are. */
void or_setup_from_tv(or_ctx* c, const double* const* tv, int smooth);

/* Read access to the hierarchy. */
void or_get_tv(or_ctx* c, int l, int i, double* out);
/* which: 0 = G1 (A), 1 = G2 (B), 2 = G3 (C); mu ignored for G1.
 * out: complex [nsites_l][dof_l][dof_l], row = 2c+alpha, col = 2b+beta. */
void or_get_links(or_ctx* c, int l, int which, int mu, double* out);

/* One K-cycle application psi = K(eta) at level 0 from psi = 0 (include/amg.h:134-144). */
void or_kcycle0(or_ctx* c, const double* eta, double* psi);
/* Stand-alone multigrid (src/amg.cpp:181-231) from the x given; returns the
 * cycles used or -it. */
int or_multigrid(or_ctx* c, const double* b, double* x, int it, double tol, double* relres);
/* Comparison solvers on the fine lattice (FP64): Bi-CGstab on D from 0
 * (src/bi_cgstab.cpp:5-87), CG on D D^dagger from x0 = phi
 * (src/conjugate_gradient.cpp:5-65), FGMRES with one SAP sweep as
 * preconditioner (include/sap.h:194-225; out_m/out_restarts/out_tol). */
int or_bicgstab(or_ctx* c, const double* b, double* x, int max_iter, double tol, double* hist, int hist_cap,
                int* nhist);
int or_cg(or_ctx* c, const double* phi, double* x, int max_iter, double tol);
int or_fgmres_sap(or_ctx* c, const double* b, double* x, double* hist, int hist_cap, int* nhist);
/* Fermion force of S = phi^+ (D D^+)^-1 phi: F [ns][2] real = dS/dtheta_mu(n);
 * returns S.  Two multigrid solves (out_tol). */
double or_force(or_ctx* c, const double* phi, double* F);
/* V-cycle (src/amg.cpp:90-132) at level l from psi as given. */
void or_vcycle(or_ctx* c, int l, const double* eta, double* psi);

/* Outer FGMRES with the K-cycle (src/fgmres.cpp:5-88, include/amg.h:99-145).
 * Returns the reference's iteration count (k*m + maxIt, or restarts*m).
 * hist receives |g_{j+1}| for every Arnoldi step (up to hist_cap entries). */
int  or_solve(or_ctx* c, const double* b, double* x, double* hist, int hist_cap, int* nhist,
              double* true_relres);

/* Solver statistics since creation, per level l:
 * out[6l..6l+5] = K-cycle FGMRES calls, iterations; coarsest GMRES calls,
 * iterations; SAP block GMRES calls, iterations. */
void or_stats(or_ctx* c, long* out);

/* Plain restarted FGMRES on the fine D with identity preconditioner
 * (FGMRES_fine_level, include/fgmres.h:126-148). */
int  or_gmres_fine(or_ctx* c, const double* b, double* x, int m, int restarts, double tol);

#ifdef __cplusplus
}
#endif
#endif
