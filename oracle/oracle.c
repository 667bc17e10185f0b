/*
 * CPU oracle: C11 restatement of the reference DD-alpha-AMG solve.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Reference paths are relative to
 * the Fabian2598/AMG_SchwingerModel tree.
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cx;

/* ------------------------------------------------------------------ */
/* BLAS-1 with the reference conventions (include/lin_alg_op.h:32-111) */
/* ------------------------------------------------------------------ */

/* dot(x, y) = sum_i x_i conj(y_i), summed in index order. */
static cx vdot(const cx* x, const cx* y, int n) {
    cx z = 0;
    for (int i = 0; i < n; i++) z += x[i] * conj(y[i]);
    return z;
}
static double vnorm(const cx* x, int n) { return sqrt(creal(vdot(x, x, n))); }

/* ------------------------------------------------------------------ */
/* mt19937 and libstdc++'s uniform_real_distribution<double>(-1, 1)    */
/* (src/amg.cpp:8-16): generate_canonical<double,53> takes two 32-bit  */
/* words, sum = w0 + w1 * 2^32, u = sum / 2^64 (clamped below 1).      */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t s[624]; int i; } mt_t;

static void mt_seed(mt_t* m, uint32_t seed) {
    m->s[0] = seed;
    for (int k = 1; k < 624; k++)
        m->s[k] = 1812433253u * (m->s[k - 1] ^ (m->s[k - 1] >> 30)) + (uint32_t)k;
    m->i = 624;
}
static uint32_t mt_next(mt_t* m) {
    if (m->i >= 624) {
        for (int k = 0; k < 624; k++) {
            uint32_t y = (m->s[k] & 0x80000000u) | (m->s[(k + 1) % 624] & 0x7fffffffu);
            m->s[k] = m->s[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        m->i = 0;
    }
    uint32_t y = m->s[m->i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}
static double mt_uniform_pm1(mt_t* m) {
    double sum = (double)mt_next(m);
    sum += (double)mt_next(m) * 4294967296.0;
    double u = sum / 18446744073709551616.0;
    if (u >= 1.0) u = nextafter(1.0, 0.0);
    return u * 2.0 + (-1.0);
}
static cx mt_complex(mt_t* m, int draw_order) {
    double a = mt_uniform_pm1(m), b = mt_uniform_pm1(m);
    return draw_order == 0 ? CMPLX(a, b) : CMPLX(b, a);
}

/* ------------------------------------------------------------------ */
/* Context                                                            */
/* ------------------------------------------------------------------ */
typedef struct {
    int n, m, restarts;
    double tol;
    cx *V, *Z, *H, *g, *sn, *cn, *eta, *r, *w, *Dx;
    long calls, iters;  /* statistics: fgmres() calls and summed iteration counts */
} fg_ws;

typedef struct {
    int nx, nt, ns, dof, colors;
    /* aggregation towards level l+1 (include/level.h:147-155) */
    int bx, bt, ntest, nagg, xel, tel, spb;
    int* agg;                   /* [nagg][colors*spb] flat entry n*dof + 2c + s */
    /* Schwarz blocks (include/sap.h:53-79) */
    int sbx, sbt, sxel, stel, sspb, nsb, ncol;
    int* sap_sites;             /* [nsb][sspb] */
    int *red, *black;           /* [ncol] */
    /* stencil coefficients: row = 2c+alpha, col = 2b+beta */
    cx *G1, *G2, *G3;           /* G1 [ns][dof][dof]; G2, G3 [2][ns][dof][dof] */
    cx* w;                      /* interpolator columns [ntest][ns*dof] */
    fg_ws kc, crs;
    fg_ws* blk;                 /* one local-GMRES workspace per same-colour block */
    cx *sap_r, *sap_Dx, *sap_buf;
    cx *t_Dpsi, *t_r, *t_P, *t_eta1, *t_psi1;
} lvl_t;

struct or_ctx {
    or_params p;
    int L;
    cx* U;                      /* [ns0][2] */
    lvl_t lv[OR_MAXL];
    fg_ws out;
    cx *tmp0, *tmp1;
};

static void* xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

static void ws_init(fg_ws* W, int n, int m, int restarts, double tol) {
    W->n = n; W->m = m; W->restarts = restarts; W->tol = tol;
    W->V = xcalloc((size_t)(m + 1) * n, sizeof(cx));
    W->Z = xcalloc((size_t)m * n, sizeof(cx));
    W->H = xcalloc((size_t)(m + 1) * m, sizeof(cx));
    W->g = xcalloc(m + 1, sizeof(cx));
    W->sn = xcalloc(m, sizeof(cx));
    W->cn = xcalloc(m, sizeof(cx));
    W->eta = xcalloc(m, sizeof(cx));
    W->r = xcalloc(n, sizeof(cx));
    W->w = xcalloc(n, sizeof(cx));
    W->Dx = xcalloc(n, sizeof(cx));
}
static void ws_free(fg_ws* W) {
    free(W->V); free(W->Z); free(W->H); free(W->g); free(W->sn); free(W->cn);
    free(W->eta); free(W->r); free(W->w); free(W->Dx);
}

void or_default_params(or_params* p) {
    memset(p, 0, sizeof(*p));
    p->m0 = -0.18840579710144945;     /* src/main.cpp:64 */
    p->nu1 = 0; p->nu2 = 2;
    p->sap_m = 5; p->sap_restarts = 5; p->sap_gmres_tol = 1e-3; p->sap_tol = 1e-10;
    p->kc_m = 5; p->kc_restarts = 2; p->kc_tol = 0.1;
    p->crs_m = 20; p->crs_restarts = 20; p->crs_tol = 0.1;
    p->out_m = 20; p->out_restarts = 50; p->out_tol = 1e-10;
    p->tv_sap_iters = 4;
    p->seed = 50;
    p->cycle = 1;                      /* src/main.cpp:58 */
    p->nit = 0;                        /* src/main.cpp:129 */
}

/* Neighbour of site n on an nx x nt periodic lattice (include/boundary.h:36-40). */
static inline int nb_right(int n, int mu, int nx, int nt) {
    int x = n / nt, t = n % nt;
    return mu == 0 ? x * nt + (t + 1) % nt : ((x + 1) % nx) * nt + t;
}
static inline int nb_left(int n, int mu, int nx, int nt) {
    int x = n / nt, t = n % nt;
    return mu == 0 ? x * nt + (t + nt - 1) % nt : ((x + nx - 1) % nx) * nt + t;
}
/* Antiperiodic time boundary: -1 for the hop that wraps in t. */
static inline double sg_right(int n, int mu, int nt) { return (mu == 0 && n % nt == nt - 1) ? -1.0 : 1.0; }
static inline double sg_left(int n, int mu, int nt) { return (mu == 0 && n % nt == 0) ? -1.0 : 1.0; }

int or_nsites(const or_ctx* c, int l) { return c->lv[l].ns; }
int or_dof(const or_ctx* c, int l) { return c->lv[l].dof; }

/* Level-0 coefficients from U (src/level.cpp:87-125, PDF eq.40):
 * G2_mu = 1/2 (1 - sigma_mu) U_mu(n),  G3_mu = 1/2 (1 + sigma_mu) conj(U_mu(n - mu)),
 * sigma_0 = sigma_x (time), sigma_1 = sigma_y (space). */
static void make_fine_links(or_ctx* c) {
    lvl_t* L0 = &c->lv[0];
    const cx Mm[2][2][2] = {{{1, -1}, {-1, 1}}, {{1, I}, {-I, 1}}};
    const cx Pm[2][2][2] = {{{1, 1}, {1, 1}}, {{1, -I}, {I, 1}}};
    for (int n = 0; n < L0->ns; n++)
        for (int mu = 0; mu < 2; mu++) {
            cx u = c->U[2 * n + mu];
            cx ul = conj(c->U[2 * nb_left(n, mu, L0->nx, L0->nt) + mu]);
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) {
                    L0->G2[((size_t)mu * L0->ns + n) * 4 + a * 2 + b] = 0.5 * Mm[mu][a][b] * u;
                    L0->G3[((size_t)mu * L0->ns + n) * 4 + a * 2 + b] = 0.5 * Pm[mu][a][b] * ul;
                }
        }
}

or_ctx* or_create(const or_params* p, const double* U) {
    int L = p->levels;
    if (L < 2 || L > OR_MAXL || p->nx <= 0 || p->nt <= 0) return NULL;
    or_ctx* c = xcalloc(1, sizeof(or_ctx));
    c->p = *p;
    c->L = L;
    /* geometry per level (include/params.h:30-65) */
    for (int l = 0; l < L; l++) {
        lvl_t* v = &c->lv[l];
        v->nx = l == 0 ? p->nx : p->bx[l - 1];
        v->nt = l == 0 ? p->nt : p->bt[l - 1];
        v->ns = v->nx * v->nt;
        v->colors = l == 0 ? 1 : p->ntest[l - 1];
        v->dof = 2 * v->colors;
        if (l < L - 1) {
            v->bx = p->bx[l]; v->bt = p->bt[l]; v->ntest = p->ntest[l];
            v->sbx = p->sapx[l]; v->sbt = p->sapt[l];
        } else {
            v->bx = v->bt = 1; v->ntest = 1; v->sbx = v->sbt = 1;
        }
        if (v->bx <= 0 || v->bt <= 0 || v->sbx <= 0 || v->sbt <= 0 ||
            v->nx % v->bx || v->nt % v->bt || v->nx % v->sbx || v->nt % v->sbt) {
            or_destroy(c);
            return NULL;
        }
        v->nagg = 2 * v->bx * v->bt;
        v->xel = v->nx / v->bx; v->tel = v->nt / v->bt; v->spb = v->xel * v->tel;
        v->sxel = v->nx / v->sbx; v->stel = v->nt / v->sbt; v->sspb = v->sxel * v->stel;
        v->nsb = v->sbx * v->sbt;
        v->ncol = v->nsb != 1 ? v->nsb / 2 : 1;
    }
    size_t ns0 = c->lv[0].ns;
    c->U = xcalloc(ns0 * 2, sizeof(cx));
    memcpy(c->U, U, ns0 * 2 * sizeof(cx));
    for (int l = 0; l < L; l++) {
        lvl_t* v = &c->lv[l];
        size_t n = (size_t)v->ns * v->dof, dd = (size_t)v->dof * v->dof;
        v->G1 = xcalloc(v->ns * dd, sizeof(cx));
        v->G2 = xcalloc(2 * v->ns * dd, sizeof(cx));
        v->G3 = xcalloc(2 * v->ns * dd, sizeof(cx));
        if (l < L - 1) {
            /* aggregates a = (bx*BlocksT + bt)*2 + s, entries ordered (x, t, c) (src/level.cpp:3-34) */
            v->agg = xcalloc((size_t)v->nagg * v->colors * v->spb, sizeof(int));
            for (int X = 0; X < v->bx; X++)
                for (int T = 0; T < v->bt; T++)
                    for (int s = 0; s < 2; s++) {
                        int a = (X * v->bt + T) * 2 + s, cnt = 0;
                        for (int x = X * v->xel; x < (X + 1) * v->xel; x++)
                            for (int t = T * v->tel; t < (T + 1) * v->tel; t++)
                                for (int col = 0; col < v->colors; col++)
                                    v->agg[(size_t)a * v->colors * v->spb + cnt++] = (x * v->nt + t) * v->dof + 2 * col + s;
                    }
            v->w = xcalloc((size_t)v->ntest * n, sizeof(cx));
            v->t_Dpsi = xcalloc(n, sizeof(cx));
            v->t_r = xcalloc(n, sizeof(cx));
            v->t_P = xcalloc(n, sizeof(cx));
            size_t nc = (size_t)c->lv[l + 1].ns * c->lv[l + 1].dof;
            v->t_eta1 = xcalloc(nc, sizeof(cx));
            v->t_psi1 = xcalloc(nc, sizeof(cx));
            /* Schwarz blocks and red/black colouring (src/sap.cpp:3-39) */
            v->sap_sites = xcalloc((size_t)v->nsb * v->sspb, sizeof(int));
            v->red = xcalloc(v->ncol, sizeof(int));
            v->black = xcalloc(v->ncol, sizeof(int));
            for (int X = 0; X < v->sbx; X++)
                for (int T = 0; T < v->sbt; T++) {
                    int blk = X * v->sbt + T, cnt = 0;
                    for (int x = X * v->sxel; x < (X + 1) * v->sxel; x++)
                        for (int t = T * v->stel; t < (T + 1) * v->stel; t++)
                            v->sap_sites[(size_t)blk * v->sspb + cnt++] = x * v->nt + t;
                    int red;
                    if (v->sbt % 2 == 0) red = (X % 2 == 0) ? (blk % 2 == 0) : (blk % 2 != 0);
                    else red = blk % 2 == 0;
                    if (blk / 2 < v->ncol) {
                        if (red) v->red[blk / 2] = blk;
                        else v->black[blk / 2] = blk;
                    }
                }
            v->blk = xcalloc(v->ncol, sizeof(fg_ws));
            for (int b = 0; b < v->ncol; b++)
                ws_init(&v->blk[b], v->sspb * v->dof, p->sap_m, p->sap_restarts, p->sap_gmres_tol);
            v->sap_r = xcalloc(n, sizeof(cx));
            v->sap_Dx = xcalloc(n, sizeof(cx));
            v->sap_buf = xcalloc(n, sizeof(cx));
        }
        if (l >= 1) ws_init(&v->kc, (int)n, p->kc_m, p->kc_restarts, p->kc_tol);
        if (l == L - 1) ws_init(&v->crs, (int)n, p->crs_m, p->crs_restarts, p->crs_tol);
    }
    ws_init(&c->out, (int)(ns0 * 2), p->out_m, p->out_restarts, p->out_tol);
    c->tmp0 = xcalloc(ns0 * 2, sizeof(cx));
    c->tmp1 = xcalloc(ns0 * 2, sizeof(cx));
    make_fine_links(c);
#ifdef _OPENMP
    if (p->threads > 0) omp_set_num_threads(p->threads);
#endif
    return c;
}

void or_destroy(or_ctx* c) {
    if (!c) return;
    for (int l = 0; l < c->L; l++) {
        lvl_t* v = &c->lv[l];
        free(v->agg); free(v->sap_sites); free(v->red); free(v->black);
        free(v->G1); free(v->G2); free(v->G3); free(v->w);
        if (v->blk) { for (int b = 0; b < v->ncol; b++) ws_free(&v->blk[b]); free(v->blk); }
        free(v->sap_r); free(v->sap_Dx); free(v->sap_buf);
        free(v->t_Dpsi); free(v->t_r); free(v->t_P); free(v->t_eta1); free(v->t_psi1);
        if (v->kc.V) ws_free(&v->kc);
        if (v->crs.V) ws_free(&v->crs);
    }
    if (c->out.V) ws_free(&c->out);
    free(c->U); free(c->tmp0); free(c->tmp1);
    free(c);
}

/* ------------------------------------------------------------------ */
/* Operators                                                          */
/* ------------------------------------------------------------------ */

/* D_phi (src/dirac_operator.cpp:3-27). */
static void dphi(or_ctx* c, const cx* in, cx* out) {
    const lvl_t* v = &c->lv[0];
    const cx* U = c->U;
    const double mm = c->p.m0 + 2.0;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < v->ns; n++) {
        int r0 = nb_right(n, 0, v->nx, v->nt), r1 = nb_right(n, 1, v->nx, v->nt);
        int l0 = nb_left(n, 0, v->nx, v->nt), l1 = nb_left(n, 1, v->nx, v->nt);
        cx f0 = U[2 * n] * sg_right(n, 0, v->nt), f1 = U[2 * n + 1] * sg_right(n, 1, v->nt);
        cx b0 = conj(U[2 * l0]) * sg_left(n, 0, v->nt), b1 = conj(U[2 * l1 + 1]) * sg_left(n, 1, v->nt);
        const cx *pr0 = in + 2 * r0, *pr1 = in + 2 * r1, *pl0 = in + 2 * l0, *pl1 = in + 2 * l1;
        out[2 * n] = mm * in[2 * n] - 0.5 * (f0 * (pr0[0] - pr0[1]) + f1 * (pr1[0] + I * pr1[1]) +
                                             b0 * (pl0[0] + pl0[1]) + b1 * (pl1[0] - I * pl1[1]));
        out[2 * n + 1] = mm * in[2 * n + 1] - 0.5 * (f0 * (-pr0[0] + pr0[1]) + f1 * (-I * pr1[0] + pr1[1]) +
                                                     b0 * (pl0[0] + pl0[1]) + b1 * (I * pl1[0] + pl1[1]));
    }
}

/* D_dagger_phi (src/dirac_operator.cpp:30-54): projectors swapped. */
static void ddag(or_ctx* c, const cx* in, cx* out) {
    const lvl_t* v = &c->lv[0];
    const cx* U = c->U;
    const double mm = c->p.m0 + 2.0;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < v->ns; n++) {
        int r0 = nb_right(n, 0, v->nx, v->nt), r1 = nb_right(n, 1, v->nx, v->nt);
        int l0 = nb_left(n, 0, v->nx, v->nt), l1 = nb_left(n, 1, v->nx, v->nt);
        cx f0 = U[2 * n] * sg_right(n, 0, v->nt), f1 = U[2 * n + 1] * sg_right(n, 1, v->nt);
        cx b0 = conj(U[2 * l0]) * sg_left(n, 0, v->nt), b1 = conj(U[2 * l1 + 1]) * sg_left(n, 1, v->nt);
        const cx *pr0 = in + 2 * r0, *pr1 = in + 2 * r1, *pl0 = in + 2 * l0, *pl1 = in + 2 * l1;
        out[2 * n] = mm * in[2 * n] - 0.5 * (b0 * (pl0[0] - pl0[1]) + b1 * (pl1[0] + I * pl1[1]) +
                                             f0 * (pr0[0] + pr0[1]) + f1 * (pr1[0] - I * pr1[1]));
        out[2 * n + 1] = mm * in[2 * n + 1] - 0.5 * (b0 * (-pl0[0] + pl0[1]) + b1 * (-I * pl1[0] + pl1[1]) +
                                                     f0 * (pr0[0] + pr0[1]) + f1 * (I * pr1[0] + pr1[1]));
    }
}

/* Level stencil (src/level.cpp:161-180):
 * out = (m0+2) v - G1 v - sum_mu [G2_mu s+ v(n+mu) + G3_mu s- v(n-mu)],
 * columns visited in the reference order (beta outer, b inner). */
static void dop(or_ctx* c, int l, const cx* in, cx* out) {
    const lvl_t* v = &c->lv[l];
    const int D = v->dof, C = v->colors;
    const double mm = c->p.m0 + 2.0;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < v->ns; n++) {
        int R[2], Lf[2];
        double sR[2], sL[2];
        for (int mu = 0; mu < 2; mu++) {
            R[mu] = nb_right(n, mu, v->nx, v->nt); Lf[mu] = nb_left(n, mu, v->nx, v->nt);
            sR[mu] = sg_right(n, mu, v->nt); sL[mu] = sg_left(n, mu, v->nt);
        }
        for (int row = 0; row < D; row++) {
            cx acc = mm * in[(size_t)n * D + row];
            const cx* g1 = v->G1 + ((size_t)n * D + row) * D;
            for (int be = 0; be < 2; be++)
                for (int b = 0; b < C; b++) {
                    int col = 2 * b + be;
                    acc -= g1[col] * in[(size_t)n * D + col];
                    for (int mu = 0; mu < 2; mu++) {
                        const cx* g2 = v->G2 + (((size_t)mu * v->ns + n) * D + row) * D;
                        const cx* g3 = v->G3 + (((size_t)mu * v->ns + n) * D + row) * D;
                        acc -= g2[col] * sR[mu] * in[(size_t)R[mu] * D + col] +
                               g3[col] * sL[mu] * in[(size_t)Lf[mu] * D + col];
                    }
                }
            out[(size_t)n * D + row] = acc;
        }
    }
}

/* Block-local stencil with zero Dirichlet boundary (src/level.cpp:304-368):
 * neighbours outside Schwarz block `blk` read as zero; signs of the global
 * lattice are kept.  in/out are block-local [sspb][dof]. */
static void dlocal(or_ctx* c, int l, int blk, const cx* in, cx* out) {
    const lvl_t* v = &c->lv[l];
    const int D = v->dof, C = v->colors;
    const double mm = c->p.m0 + 2.0;
    const int* sites = v->sap_sites + (size_t)blk * v->sspb;
    for (int m = 0; m < v->sspb; m++) {
        int n = sites[m];
        int R[2], Lf[2];
        double sR[2], sL[2];
        for (int mu = 0; mu < 2; mu++) {
            int nn[2] = {nb_right(n, mu, v->nx, v->nt), nb_left(n, mu, v->nx, v->nt)};
            for (int k = 0; k < 2; k++) {
                /* getMandBlock (include/level.h:48-59) takes x = n / Nx, t = n % Nt;
                 * kept as is: identical to the true coordinates on square lattices */
                int x = nn[k] / v->nx, t = nn[k] % v->nt;
                int b2 = (x / v->sxel) * v->sbt + t / v->stel;
                int loc = (x % v->sxel) * v->stel + t % v->stel;
                if (k == 0) R[mu] = b2 == blk ? loc : -1;
                else Lf[mu] = b2 == blk ? loc : -1;
            }
            sR[mu] = sg_right(n, mu, v->nt); sL[mu] = sg_left(n, mu, v->nt);
        }
        for (int row = 0; row < D; row++) {
            cx acc = mm * in[(size_t)m * D + row];
            const cx* g1 = v->G1 + ((size_t)n * D + row) * D;
            for (int be = 0; be < 2; be++)
                for (int b = 0; b < C; b++) {
                    int col = 2 * b + be;
                    acc -= g1[col] * in[(size_t)m * D + col];
                    for (int mu = 0; mu < 2; mu++) {
                        const cx* g2 = v->G2 + (((size_t)mu * v->ns + n) * D + row) * D;
                        const cx* g3 = v->G3 + (((size_t)mu * v->ns + n) * D + row) * D;
                        cx pr = R[mu] >= 0 ? in[(size_t)R[mu] * D + col] : 0;
                        cx pl = Lf[mu] >= 0 ? in[(size_t)Lf[mu] * D + col] : 0;
                        acc -= g2[col] * sR[mu] * pr + g3[col] * sL[mu] * pl;
                    }
                }
            out[(size_t)m * D + row] = acc;
        }
    }
}

/* P v (src/level.cpp:182-209): out[n][2c+s] = sum_cc w_cc[n][2c+s] v[X(n)][2cc+s]. */
static void prolong(or_ctx* c, int l, const cx* vc, cx* out) {
    const lvl_t* v = &c->lv[l];
    const int Dc = c->lv[l + 1].dof, E = v->colors * v->spb;
    const size_t n = (size_t)v->ns * v->dof;
    memset(out, 0, n * sizeof(cx));
    for (int cc = 0; cc < v->ntest; cc++)
        for (int a = 0; a < v->nagg; a++) {
            int nc = a / 2, s = a % 2;
            const int* ent = v->agg + (size_t)a * E;
            for (int i = 0; i < E; i++) out[ent[i]] += v->w[cc * n + ent[i]] * vc[(size_t)nc * Dc + 2 * cc + s];
        }
}

/* P^dagger v (src/level.cpp:212-237). */
static void restrict_(or_ctx* c, int l, const cx* vf, cx* out) {
    const lvl_t* v = &c->lv[l];
    const int Dc = c->lv[l + 1].dof, E = v->colors * v->spb;
    const size_t n = (size_t)v->ns * v->dof;
    memset(out, 0, (size_t)c->lv[l + 1].ns * Dc * sizeof(cx));
    for (int cc = 0; cc < v->ntest; cc++)
        for (int a = 0; a < v->nagg; a++) {
            int nc = a / 2, s = a % 2;
            const int* ent = v->agg + (size_t)a * E;
            cx acc = out[(size_t)nc * Dc + 2 * cc + s];
            for (int i = 0; i < E; i++) acc += conj(v->w[cc * n + ent[i]]) * vf[ent[i]];
            out[(size_t)nc * Dc + 2 * cc + s] = acc;
        }
}

/* ------------------------------------------------------------------ */
/* FGMRES (src/fgmres.cpp:5-115, include/fgmres.h:93-120)              */
/* ------------------------------------------------------------------ */
typedef void (*op_fn)(or_ctx*, const void*, const cx*, cx*);

/* Givens rotation of column j (src/fgmres.cpp:90-104). */
static void givens(fg_ws* W, int j) {
    const int m = W->m;
    cx* H = W->H;   /* H[i*m + j] */
    for (int i = 0; i < j; i++) {
        cx tmp = conj(W->cn[i]) * H[i * m + j] + conj(W->sn[i]) * H[(i + 1) * m + j];
        H[(i + 1) * m + j] = -W->sn[i] * H[i * m + j] + W->cn[i] * H[(i + 1) * m + j];
        H[i * m + j] = tmp;
    }
    cx hjj = H[j * m + j], hj1 = H[(j + 1) * m + j];
    cx den = csqrt(conj(hjj) * hjj + conj(hj1) * hj1);
    W->sn[j] = hj1 / den;
    W->cn[j] = hjj / den;
    H[j * m + j] = conj(W->cn[j]) * hjj + conj(W->sn[j]) * hj1;
    H[(j + 1) * m + j] = 0;
}

/* Right-preconditioned restarted FGMRES with modified Gram-Schmidt.
 * pre == NULL is the identity.  pre receives Z_j with whatever it holds (the
 * K-cycle uses it as its initial guess, include/amg.h:37-39).  Stopping: the
 * Arnoldi estimate |g_{j+1}| < tol ||b|| ends a cycle early; the true residual
 * is checked after each cycle.  maxIt persists across cycles (src/fgmres.cpp:14). */
static int fgmres_body(or_ctx* c, fg_ws* W, op_fn op, const void* oarg, op_fn pre, const void* parg,
                       const cx* b, const cx* x0, cx* x, double* hist, int hist_cap, int* nhist);
static int fgmres(or_ctx* c, fg_ws* W, op_fn op, const void* oarg, op_fn pre, const void* parg,
                  const cx* b, const cx* x0, cx* x, double* hist, int hist_cap, int* nhist) {
    int it = fgmres_body(c, W, op, oarg, pre, parg, b, x0, x, hist, hist_cap, nhist);
    W->calls++;
    W->iters += it;
    return it;
}
static int fgmres_body(or_ctx* c, fg_ws* W, op_fn op, const void* oarg, op_fn pre, const void* parg,
                       const cx* b, const cx* x0, cx* x, double* hist, int hist_cap, int* nhist) {
    const int n = W->n, m = W->m;
    memset(W->V, 0, sizeof(cx) * (size_t)(m + 1) * n);
    memset(W->Z, 0, sizeof(cx) * (size_t)m * n);
    memset(W->H, 0, sizeof(cx) * (size_t)(m + 1) * m);
    memset(W->g, 0, sizeof(cx) * (m + 1));
    memset(W->sn, 0, sizeof(cx) * m); memset(W->cn, 0, sizeof(cx) * m); memset(W->eta, 0, sizeof(cx) * m);
    memset(W->r, 0, sizeof(cx) * n); memset(W->w, 0, sizeof(cx) * n); memset(W->Dx, 0, sizeof(cx) * n);
    if (x != x0) memcpy(x, x0, sizeof(cx) * n);
    op(c, oarg, x, W->Dx);
    for (int i = 0; i < n; i++) W->r[i] = b[i] - W->Dx[i];
    double nb = vnorm(b, n), err = vnorm(W->r, n);
    int maxIt = m;
    for (int k = 0; k < W->restarts; k++) {
        cx beta = err;
        cx ib = 1.0 / beta;
        for (int i = 0; i < n; i++) W->V[i] = ib * W->r[i];
        W->g[0] = beta;
        for (int j = 0; j < m; j++) {
            cx* Vj = W->V + (size_t)j * n;
            cx* Zj = W->Z + (size_t)j * n;
            if (pre) pre(c, parg, Vj, Zj);
            else memcpy(Zj, Vj, sizeof(cx) * n);
            op(c, oarg, Zj, W->w);
            for (int i = 0; i <= j; i++) {
                const cx* Vi = W->V + (size_t)i * n;
                cx h = vdot(W->w, Vi, n);
                W->H[i * m + j] = h;
                for (int q = 0; q < n; q++) W->w[q] -= h * Vi[q];
            }
            cx hn = vnorm(W->w, n);
            W->H[(j + 1) * m + j] = hn;
            if (creal(hn) > 0) {
                cx ih = 1.0 / hn;
                cx* Vn = W->V + (size_t)(j + 1) * n;
                for (int q = 0; q < n; q++) Vn[q] = ih * W->w[q];
            }
            givens(W, j);
            W->g[j + 1] = -W->sn[j] * W->g[j];
            W->g[j] = conj(W->cn[j]) * W->g[j];
            if (hist && nhist && *nhist < hist_cap) hist[(*nhist)++] = cabs(W->g[j + 1]);
            if (cabs(W->g[j + 1]) < W->tol * nb) { maxIt = j + 1; break; }
        }
        /* back substitution (src/fgmres.cpp:107-115) */
        for (int i = maxIt - 1; i >= 0; i--) {
            cx e = W->g[i];
            for (int q = i + 1; q < maxIt; q++) e -= W->H[i * m + q] * W->eta[q];
            W->eta[i] = e / W->H[i * m + i];
        }
        for (int q = 0; q < n; q++) {
            cx s = x[q];
            for (int j = 0; j < maxIt; j++) s = s + W->eta[j] * W->Z[(size_t)j * n + q];
            x[q] = s;
        }
        op(c, oarg, x, W->Dx);
        for (int i = 0; i < n; i++) W->r[i] = b[i] - W->Dx[i];
        err = vnorm(W->r, n);
        if (err < W->tol * nb) return k * m + maxIt;
    }
    return W->restarts * m;
}

/* operator adaptors */
typedef struct { int l, blk; } oparg_t;
static void op_dop(or_ctx* c, const void* a, const cx* in, cx* out) { dop(c, ((const oparg_t*)a)->l, in, out); }
static void op_dlocal(or_ctx* c, const void* a, const cx* in, cx* out) {
    const oparg_t* o = a;
    dlocal(c, o->l, o->blk, in, out);
}
static void op_dphi(or_ctx* c, const void* a, const cx* in, cx* out) { (void)a; dphi(c, in, out); }

/* ------------------------------------------------------------------ */
/* SAP (src/sap.cpp:44-174)                                           */
/* ------------------------------------------------------------------ */
static int sap(or_ctx* c, int l, const cx* vrhs, cx* x, int nu) {
    lvl_t* v = &c->lv[l];
    const int D = v->dof;
    const size_t n = (size_t)v->ns * D, nb = (size_t)v->sspb * D;
    double vn = vnorm(vrhs, (int)n), err = 0;
    dop(c, l, x, v->sap_Dx);
    for (size_t i = 0; i < n; i++) v->sap_r[i] = vrhs[i] - v->sap_Dx[i];
    for (int it = 0; it < nu; it++) {
        for (int colour = 0; colour < 2; colour++) {
            const int* list = colour == 0 ? v->red : v->black;
            memset(v->sap_buf, 0, n * sizeof(cx));
#pragma omp parallel for schedule(dynamic)
            for (int bi = 0; bi < v->ncol; bi++) {
                int blk = list[bi];
                fg_ws* W = &v->blk[bi];
                const int* sites = v->sap_sites + (size_t)blk * v->sspb;
                /* temp = I_B^T r ; x_B = GMRES(D_B, temp, x0 = temp) (src/sap.cpp:44-62) */
                cx* xb = malloc(nb * sizeof(cx));
                cx* bb = malloc(nb * sizeof(cx));
                for (int m = 0; m < v->sspb; m++)
                    for (int d = 0; d < D; d++) bb[(size_t)m * D + d] = v->sap_r[(size_t)sites[m] * D + d];
                oparg_t oa = {l, blk};
                fgmres(c, W, op_dlocal, &oa, NULL, NULL, bb, bb, xb, NULL, 0, NULL);
                for (int m = 0; m < v->sspb; m++)
                    for (int d = 0; d < D; d++) v->sap_buf[(size_t)sites[m] * D + d] = xb[(size_t)m * D + d];
                free(xb); free(bb);
            }
            for (size_t i = 0; i < n; i++) x[i] += v->sap_buf[i];
            dop(c, l, x, v->sap_Dx);
            for (size_t i = 0; i < n; i++) v->sap_r[i] = vrhs[i] - v->sap_Dx[i];
        }
        err = vnorm(v->sap_r, (int)n);
        if (err < c->p.sap_tol * vn) return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* K-cycle (src/amg.cpp:135-178, include/amg.h:19-40)                  */
/* ------------------------------------------------------------------ */
static void kcycle(or_ctx* c, int l, const cx* eta, cx* psi);
static void pre_kcycle(or_ctx* c, const void* a, const cx* in, cx* out) { kcycle(c, ((const oparg_t*)a)->l, in, out); }

static void kcycle(or_ctx* c, int l, const cx* eta, cx* psi) {
    lvl_t* v = &c->lv[l];
    oparg_t oa = {l, 0};
    if (l == c->L - 1) {
        /* coarsest: GMRES(crs_m, crs_restarts, crs_tol) from x0 = eta (src/amg.cpp:136-139) */
        fgmres(c, &v->crs, op_dop, &oa, NULL, NULL, eta, eta, psi, NULL, 0, NULL);
        return;
    }
    const size_t n = (size_t)v->ns * v->dof;
    if (c->p.nu1 > 0) sap(c, l, eta, psi, c->p.nu1);
    dop(c, l, psi, v->t_Dpsi);
    for (size_t i = 0; i < n; i++) v->t_r[i] = eta[i] - v->t_Dpsi[i];
    restrict_(c, l, v->t_r, v->t_eta1);
    oparg_t oc = {l + 1, 0};
    fgmres(c, &c->lv[l + 1].kc, op_dop, &oc, pre_kcycle, &oc, v->t_eta1, v->t_eta1, v->t_psi1, NULL, 0, NULL);
    prolong(c, l, v->t_psi1, v->t_P);
    for (size_t i = 0; i < n; i++) psi[i] += v->t_P[i];
    if (c->p.nu2 > 0) sap(c, l, eta, psi, c->p.nu2);
}

/* V-cycle (src/amg.cpp:90-132): smoothing, coarse correction by a recursive
 * V-cycle from a zero guess, smoothing; GMRES from x0 = eta at the coarsest. */
static void vcycle(or_ctx* c, int l, const cx* eta, cx* psi) {
    lvl_t* v = &c->lv[l];
    oparg_t oa = {l, 0};
    if (l == c->L - 1) {
        fgmres(c, &v->crs, op_dop, &oa, NULL, NULL, eta, eta, psi, NULL, 0, NULL);
        return;
    }
    const size_t n = (size_t)v->ns * v->dof;
    const size_t nc = (size_t)c->lv[l + 1].ns * c->lv[l + 1].dof;
    if (c->p.nu1 > 0) sap(c, l, eta, psi, c->p.nu1);
    dop(c, l, psi, v->t_Dpsi);
    for (size_t i = 0; i < n; i++) v->t_r[i] = eta[i] - v->t_Dpsi[i];
    restrict_(c, l, v->t_r, v->t_eta1);
    memset(v->t_psi1, 0, sizeof(cx) * nc);
    vcycle(c, l + 1, v->t_eta1, v->t_psi1);
    prolong(c, l, v->t_psi1, v->t_P);
    for (size_t i = 0; i < n; i++) psi[i] += v->t_P[i];
    if (c->p.nu2 > 0) sap(c, l, eta, psi, c->p.nu2);
}

static void cycle(or_ctx* c, int l, const cx* eta, cx* psi) {
    if (c->p.cycle == 0) vcycle(c, l, eta, psi);
    else kcycle(c, l, eta, psi);
}

/* Outer preconditioner: zero, then the cycle at level 0 (include/amg.h:134-144). */
static void pre_outer(or_ctx* c, const void* a, const cx* in, cx* out) {
    (void)a;
    memset(out, 0, sizeof(cx) * (size_t)c->lv[0].ns * 2);
    cycle(c, 0, in, out);
}

/* ------------------------------------------------------------------ */
/* Setup (src/amg.cpp:3-46, src/level.cpp:242-301, 373-417)           */
/* ------------------------------------------------------------------ */

/* Aggregate-wise modified Gram-Schmidt, then normalisation. */
static void orthonormalize(or_ctx* c, int l) {
    lvl_t* v = &c->lv[l];
    const size_t n = (size_t)v->ns * v->dof;
    const int E = v->colors * v->spb;
    for (int a = 0; a < v->nagg; a++) {
        const int* ent = v->agg + (size_t)a * E;
        for (int nt = 0; nt < v->ntest; nt++) {
            cx* wn = v->w + nt * n;
            for (int ntt = 0; ntt < nt; ntt++) {
                const cx* wm = v->w + ntt * n;
                cx proj = 0;
                for (int j = 0; j < E; j++) proj += wn[ent[j]] * conj(wm[ent[j]]);
                for (int j = 0; j < E; j++) wn[ent[j]] -= proj * wm[ent[j]];
            }
            cx nrm = 0;
            for (int j = 0; j < E; j++) nrm += wn[ent[j]] * conj(wn[ent[j]]);
            nrm = sqrt(creal(nrm));
            for (int j = 0; j < E; j++) wn[ent[j]] /= nrm;
        }
    }
}

static inline int agg_block(const lvl_t* v, int n) {
    int x = n / v->nt, t = n % v->nt;
    return (x / v->xel) * v->bt + t / v->tel;
}

/* Galerkin coarse links (src/level.cpp:242-301).  Hops that stay in block X
 * go to A, hops into X+mu to B_mu, into X-mu to C_mu; no boundary sign is
 * folded in (the coarse stencil applies its own). */
static void coarse_links(or_ctx* c, int l) {
    lvl_t* v = &c->lv[l];
    lvl_t* nx = &c->lv[l + 1];
    const int D = v->dof, C = v->colors, Nv = v->ntest, Dc = nx->dof;
    const size_t n = (size_t)v->ns * D;
    memset(nx->G1, 0, sizeof(cx) * (size_t)nx->ns * Dc * Dc);
    memset(nx->G2, 0, sizeof(cx) * 2 * (size_t)nx->ns * Dc * Dc);
    memset(nx->G3, 0, sizeof(cx) * 2 * (size_t)nx->ns * Dc * Dc);
#pragma omp parallel for schedule(dynamic)
    for (int X = 0; X < nx->ns; X++) {
        int bxi = X / v->bt, bti = X % v->bt;
        int XR[2], XL[2];
        for (int mu = 0; mu < 2; mu++) { XR[mu] = nb_right(X, mu, nx->nx, nx->nt); XL[mu] = nb_left(X, mu, nx->nx, nx->nt); }
        for (int al = 0; al < 2; al++)
            for (int be = 0; be < 2; be++)
                for (int p = 0; p < Nv; p++)
                    for (int s = 0; s < Nv; s++) {
                        size_t row = 2 * p + al, col = 2 * s + be;
                        cx A = 0, B[2] = {0, 0}, Cc[2] = {0, 0};
                        const cx *wp = v->w + p * n, *ws = v->w + s * n;
                        for (int x = bxi * v->xel; x < (bxi + 1) * v->xel; x++)
                            for (int t = bti * v->tel; t < (bti + 1) * v->tel; t++) {
                                int site = x * v->nt + t;
                                for (int cc = 0; cc < C; cc++)
                                    for (int b = 0; b < C; b++) {
                                        size_t fr = 2 * cc + al, fc = 2 * b + be;
                                        cx wpc = conj(wp[(size_t)site * D + fr]);
                                        A += wpc * v->G1[((size_t)site * D + fr) * D + fc] * ws[(size_t)site * D + fc];
                                        for (int mu = 0; mu < 2; mu++) {
                                            int r = nb_right(site, mu, v->nx, v->nt), lf = nb_left(site, mu, v->nx, v->nt);
                                            int br = agg_block(v, r), bl = agg_block(v, lf);
                                            cx wG2 = wpc * v->G2[(((size_t)mu * v->ns + site) * D + fr) * D + fc];
                                            cx wG3 = wpc * v->G3[(((size_t)mu * v->ns + site) * D + fr) * D + fc];
                                            if (br == X) A += wG2 * ws[(size_t)r * D + fc];
                                            else if (br == XR[mu]) B[mu] += wG2 * ws[(size_t)r * D + fc];
                                            if (bl == X) A += wG3 * ws[(size_t)lf * D + fc];
                                            else if (bl == XL[mu]) Cc[mu] += wG3 * ws[(size_t)lf * D + fc];
                                        }
                                    }
                            }
                        nx->G1[((size_t)X * Dc + row) * Dc + col] = A;
                        for (int mu = 0; mu < 2; mu++) {
                            nx->G2[(((size_t)mu * nx->ns + X) * Dc + row) * Dc + col] = B[mu];
                            nx->G3[(((size_t)mu * nx->ns + X) * Dc + row) * Dc + col] = Cc[mu];
                        }
                    }
    }
}

static void smooth_ortho_links(or_ctx* c, int smooth) {
    for (int l = 0; l < c->L - 1; l++) {
        lvl_t* v = &c->lv[l];
        size_t n = (size_t)v->ns * v->dof;
        if (smooth) {
            cx* zero = xcalloc(n, sizeof(cx));
            for (int i = 0; i < v->ntest; i++) sap(c, l, zero, v->w + i * n, c->p.tv_sap_iters);
            free(zero);
        }
        orthonormalize(c, l);
        coarse_links(c, l);
    }
}

/* Bootstrap (src/amg.cpp:52-84): per level, each test vector v becomes
 * v + cycle_l(v - D_l v) from a zero guess, using the current hierarchy; then
 * the level's interpolator and the next level's links are rebuilt. */
static void bootstrap(or_ctx* c) {
    for (int it = 0; it < c->p.nit; it++)
        for (int l = 0; l < c->L - 1; l++) {
            lvl_t* v = &c->lv[l];
            const size_t n = (size_t)v->ns * v->dof;
            cx* tvn = xcalloc((size_t)v->ntest * n, sizeof(cx));
            cx* rhs = xcalloc(n, sizeof(cx));
            for (int i = 0; i < v->ntest; i++) {
                const cx* w = v->w + i * n;
                cx* t = tvn + i * n;
                dop(c, l, w, rhs);
                for (size_t q = 0; q < n; q++) rhs[q] = w[q] - rhs[q];
                cycle(c, l, rhs, t);
                for (size_t q = 0; q < n; q++) t[q] += w[q];
            }
            memcpy(v->w, tvn, sizeof(cx) * (size_t)v->ntest * n);
            free(tvn);
            free(rhs);
            orthonormalize(c, l);
            coarse_links(c, l);
        }
}

void or_setup(or_ctx* c, int draw_order) {
    mt_t rng;
    mt_seed(&rng, c->p.seed);
    lvl_t* v0 = &c->lv[0];
    size_t n0 = (size_t)v0->ns * v0->dof;
    for (int i = 0; i < v0->ntest; i++)
        for (size_t q = 0; q < n0; q++) v0->w[i * n0 + q] = mt_complex(&rng, draw_order);
    /* levels 1..L-2: restriction of the raw level-(l-1) vectors, then fresh draws */
    for (int l = 1; l < c->L - 1; l++) {
        lvl_t* v = &c->lv[l];
        size_t n = (size_t)v->ns * v->dof, np = (size_t)c->lv[l - 1].ns * c->lv[l - 1].dof;
        for (int i = 0; i < v->ntest; i++) {
            if (i < c->lv[l - 1].ntest) restrict_(c, l - 1, c->lv[l - 1].w + i * np, v->w + i * n);
            else
                for (size_t q = 0; q < n; q++) v->w[i * n + q] = mt_complex(&rng, draw_order);
        }
    }
    smooth_ortho_links(c, 1);
    bootstrap(c);
}

void or_setup_from_tv(or_ctx* c, const double* const* tv, int smooth) {
    for (int l = 0; l < c->L - 1; l++) {
        lvl_t* v = &c->lv[l];
        memcpy(v->w, tv[l], sizeof(cx) * (size_t)v->ntest * v->ns * v->dof);
    }
    smooth_ortho_links(c, smooth);
    bootstrap(c);
}

void or_get_tv(or_ctx* c, int l, int i, double* out) {
    lvl_t* v = &c->lv[l];
    size_t n = (size_t)v->ns * v->dof;
    memcpy(out, v->w + i * n, n * sizeof(cx));
}

void or_get_links(or_ctx* c, int l, int which, int mu, double* out) {
    lvl_t* v = &c->lv[l];
    size_t sz = (size_t)v->ns * v->dof * v->dof;
    const cx* src = which == 0 ? v->G1 : (which == 1 ? v->G2 + mu * sz : v->G3 + mu * sz);
    memcpy(out, src, sz * sizeof(cx));
}

/* ------------------------------------------------------------------ */
/* Public wrappers                                                    */
/* ------------------------------------------------------------------ */
void or_dphi(or_ctx* c, const double* in, double* out) { dphi(c, (const cx*)in, (cx*)out); }
void or_ddag(or_ctx* c, const double* in, double* out) { ddag(c, (const cx*)in, (cx*)out); }
void or_dop(or_ctx* c, int l, const double* in, double* out) { dop(c, l, (const cx*)in, (cx*)out); }
void or_prolong(or_ctx* c, int l, const double* vc, double* out) { prolong(c, l, (const cx*)vc, (cx*)out); }
void or_restrict(or_ctx* c, int l, const double* vf, double* out) { restrict_(c, l, (const cx*)vf, (cx*)out); }
int or_sap(or_ctx* c, int l, const double* v, double* x, int nu) { return sap(c, l, (const cx*)v, (cx*)x, nu); }

void or_vcycle(or_ctx* c, int l, const double* eta, double* psi) { vcycle(c, l, (const cx*)eta, (cx*)psi); }

void or_kcycle0(or_ctx* c, const double* eta, double* psi) {
    memset(psi, 0, sizeof(cx) * (size_t)c->lv[0].ns * 2);
    kcycle(c, 0, (const cx*)eta, (cx*)psi);
}

/* AlgebraicMG::applyMultilevel (src/amg.cpp:181-231): x <- cycle(0, b, x),
 * residual with the level-0 stencil, stop when |r| < tol |b|.  Returns the
 * cycles used, or -it when it did not converge. */
int or_multigrid(or_ctx* c, const double* bd, double* xd, int it, double tol, double* relres) {
    const cx* b = (const cx*)bd;
    cx* x = (cx*)xd;
    const size_t n = (size_t)c->lv[0].ns * 2;
    cx* Dx = xcalloc(n, sizeof(cx));
    const double nb = vnorm(b, (int)n);
    double err = 0;
    int ret = -it;
    for (int i = 0; i < it; i++) {
        cycle(c, 0, b, x);
        dop(c, 0, x, Dx);
        for (size_t q = 0; q < n; q++) Dx[q] = b[q] - Dx[q];
        err = vnorm(Dx, (int)n);
        if (err < tol * nb) {
            ret = i + 1;
            break;
        }
    }
    if (relres) *relres = nb > 0 ? err / nb : 0.0;
    free(Dx);
    return ret;
}

int or_solve(or_ctx* c, const double* b, double* x, double* hist, int hist_cap, int* nhist,
             double* true_relres) {
    size_t n = (size_t)c->lv[0].ns * 2;
    memset(c->tmp0, 0, n * sizeof(cx));
    if (nhist) *nhist = 0;
    int it = fgmres(c, &c->out, op_dphi, NULL, pre_outer, NULL, (const cx*)b, c->tmp0, (cx*)x, hist, hist_cap, nhist);
    if (true_relres) {
        dphi(c, (const cx*)x, c->tmp1);
        for (size_t i = 0; i < n; i++) c->tmp1[i] = ((const cx*)b)[i] - c->tmp1[i];
        *true_relres = vnorm(c->tmp1, (int)n) / vnorm((const cx*)b, (int)n);
    }
    return it;
}

/* per level l: [kc calls, kc iters, coarsest calls, coarsest iters, SAP block calls, SAP block iters] */
void or_stats(or_ctx* c, long* out) {
    for (int l = 0; l < c->L; l++) {
        lvl_t* v = &c->lv[l];
        long bc = 0, bi = 0;
        for (int b = 0; v->blk && b < v->ncol; b++) { bc += v->blk[b].calls; bi += v->blk[b].iters; }
        out[6 * l + 0] = v->kc.calls; out[6 * l + 1] = v->kc.iters;
        out[6 * l + 2] = v->crs.calls; out[6 * l + 3] = v->crs.iters;
        out[6 * l + 4] = bc; out[6 * l + 5] = bi;
    }
}

int or_gmres_fine(or_ctx* c, const double* b, double* x, int m, int restarts, double tol) {
    size_t n = (size_t)c->lv[0].ns * 2;
    fg_ws W;
    ws_init(&W, (int)n, m, restarts, tol);
    memset(c->tmp0, 0, n * sizeof(cx));
    int it = fgmres(c, &W, op_dphi, NULL, NULL, NULL, (const cx*)b, c->tmp0, (cx*)x, NULL, 0, NULL);
    ws_free(&W);
    return it;
}

/* ------------------------------------------------------------------ */
/* Comparison solvers on the fine lattice                             */
/* ------------------------------------------------------------------ */

/* Bi-CGstab on D_phi from x0 = 0 (src/bi_cgstab.cpp:5-87).  hist gets |s| of
 * every iteration.  Returns the iterations (k+1 at convergence) or max_iter. */
int or_bicgstab(or_ctx* c, const double* bd, double* xd, int max_iter, double tol, double* hist, int hist_cap,
                int* nhist) {
    const int n = c->lv[0].ns * 2;
    const cx* b = (const cx*)bd;
    cx* x = (cx*)xd;
    cx *r = xcalloc(n, sizeof(cx)), *rt = xcalloc(n, sizeof(cx)), *d = xcalloc(n, sizeof(cx)),
       *s = xcalloc(n, sizeof(cx)), *t = xcalloc(n, sizeof(cx)), *Ad = xcalloc(n, sizeof(cx));
    memset(x, 0, sizeof(cx) * n);
    dphi(c, x, Ad);
    for (int i = 0; i < n; i++) r[i] = b[i] - Ad[i];
    memcpy(rt, r, sizeof(cx) * n);
    const double nb = vnorm(b, n);
    cx alpha = 0, omega = 0, rho = 0, rho_prev = 0, beta;
    int k = 0, ret = max_iter, nh = 0;
    while (k < max_iter) {
        rho = vdot(r, rt, n);
        if (k == 0) memcpy(d, r, sizeof(cx) * n);
        else {
            beta = alpha * rho / (omega * rho_prev);
            for (int i = 0; i < n; i++) d[i] = r[i] + beta * (d[i] - omega * Ad[i]);
        }
        dphi(c, d, Ad);
        alpha = rho / vdot(Ad, rt, n);
        for (int i = 0; i < n; i++) s[i] = r[i] - alpha * Ad[i];
        const double err = vnorm(s, n);
        if (hist && nh < hist_cap) hist[nh] = err;
        nh++;
        if (err < tol * nb) {
            for (int i = 0; i < n; i++) x[i] = x[i] + alpha * d[i];
            ret = k + 1;
            break;
        }
        dphi(c, s, t);
        omega = vdot(s, t, n) / vdot(t, t, n);
        for (int i = 0; i < n; i++) r[i] = s[i] - omega * t[i];
        for (int i = 0; i < n; i++) x[i] = x[i] + alpha * d[i] + omega * s[i];
        rho_prev = rho;
        k++;
    }
    if (nhist) *nhist = nh < hist_cap ? nh : hist_cap;
    free(r); free(rt); free(d); free(s); free(t); free(Ad);
    return ret;
}

/* Conjugate gradient for (D D^dagger) x = phi from x0 = phi
 * (src/conjugate_gradient.cpp:5-65).  Returns the number of D D^dagger
 * applications inside the loop (the reference prints one less) or -max_iter. */
int or_cg(or_ctx* c, const double* pd, double* xd, int max_iter, double tol) {
    const int n = c->lv[0].ns * 2;
    const cx* phi = (const cx*)pd;
    cx* x = (cx*)xd;
    cx *r = xcalloc(n, sizeof(cx)), *d = xcalloc(n, sizeof(cx)), *Ad = xcalloc(n, sizeof(cx)),
       *tmp = xcalloc(n, sizeof(cx));
    memcpy(x, phi, sizeof(cx) * n);
    ddag(c, x, tmp);
    dphi(c, tmp, Ad);
    for (int i = 0; i < n; i++) r[i] = phi[i] - Ad[i];
    memcpy(d, r, sizeof(cx) * n);
    cx rn = vdot(r, r, n);
    const double np = vnorm(phi, n);
    int k = 0, ret = -max_iter;
    while (k < max_iter) {
        ddag(c, d, tmp);
        dphi(c, tmp, Ad);
        const cx alpha = rn / vdot(d, Ad, n);
        for (int i = 0; i < n; i++) x[i] += alpha * d[i];
        for (int i = 0; i < n; i++) r[i] -= alpha * Ad[i];
        const double e2 = creal(vdot(r, r, n));
        if (sqrt(e2) < tol * np) {
            ret = k + 1;
            break;
        }
        const cx beta = e2 / rn;
        for (int i = 0; i < n; i++) {
            d[i] *= beta;
            d[i] += r[i];
        }
        rn = e2;
        k++;
    }
    free(r); free(d); free(Ad); free(tmp);
    return ret;
}

/* FGMRES with one SAP sweep as preconditioner, from x0 = 0
 * (include/sap.h:194-225): the preconditioner zeroes its output, then
 * SAP(in, out, 1) on the fine level. */
static void pre_sap(or_ctx* c, const void* a, const cx* in, cx* out) {
    (void)a;
    memset(out, 0, sizeof(cx) * (size_t)c->lv[0].ns * 2);
    sap(c, 0, in, out, 1);
}

int or_fgmres_sap(or_ctx* c, const double* b, double* x, double* hist, int hist_cap, int* nhist) {
    size_t n = (size_t)c->lv[0].ns * 2;
    fg_ws W;
    ws_init(&W, (int)n, c->p.out_m, c->p.out_restarts, c->p.out_tol);
    memset(c->tmp0, 0, n * sizeof(cx));
    int it = fgmres(c, &W, op_dphi, NULL, pre_sap, NULL, (const cx*)b, c->tmp0, (cx*)x, hist, hist_cap, nhist);
    ws_free(&W);
    return it;
}

/* Fermion force of S = phi^+ (D D^+)^-1 phi (FP64 CPU gradient): Y = D^-1 phi,
 * X = D^+^-1 Y = s3 D^-1 s3 Y by two multigrid solves, then
 * dS/dtheta_mu(n) = Re[i s U_mu(n) X(n)^+(1-sigma_mu)Y(n+mu) - i s U_mu(n)^* X(n+mu)^+(1+sigma_mu)Y(n)]
 * from dU/dtheta = iU on the hops of src/dirac_operator.cpp:3-27.  Returns S. */
double or_force(or_ctx* c, const double* phid, double* F) {
    const int ns = c->lv[0].ns, n = 2 * ns;
    const lvl_t* v = &c->lv[0];
    cx *Y = xcalloc(n, sizeof(cx)), *X = xcalloc(n, sizeof(cx)), *t = xcalloc(n, sizeof(cx));
    double rr;
    or_solve(c, phid, (double*)Y, NULL, 0, NULL, &rr);
    for (int i = 0; i < ns; i++) { t[2 * i] = Y[2 * i]; t[2 * i + 1] = -Y[2 * i + 1]; }
    or_solve(c, (const double*)t, (double*)X, NULL, 0, NULL, &rr);
    for (int i = 0; i < ns; i++) X[2 * i + 1] = -X[2 * i + 1];
    for (int s = 0; s < ns; s++)
        for (int mu = 0; mu < 2; mu++) {
            const int r = nb_right(s, mu, v->nx, v->nt);
            const double sg = sg_right(s, mu, v->nt);
            const cx u = c->U[2 * s + mu];
            cx t1, t2;
            if (mu == 0) {
                t1 = conj(X[2 * s] - X[2 * s + 1]) * (Y[2 * r] - Y[2 * r + 1]);
                t2 = conj(X[2 * r] + X[2 * r + 1]) * (Y[2 * s] + Y[2 * s + 1]);
            } else {
                t1 = conj(X[2 * s] + I * X[2 * s + 1]) * (Y[2 * r] + I * Y[2 * r + 1]);
                t2 = conj(X[2 * r] - I * X[2 * r + 1]) * (Y[2 * s] - I * Y[2 * s + 1]);
            }
            F[2 * s + mu] = creal(I * sg * (u * t1 - conj(u) * t2));
        }
    double S = creal(vdot(Y, Y, n));
    free(Y); free(X); free(t);
    return S;
}
