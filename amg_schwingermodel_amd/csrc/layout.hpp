// Launch-independent layout shared by the host orchestration and the kernels:
// workgroup sizes, compiled restart limits and the device-resident scalar
// state of the level-0 SAP block GMRES and of the outer FGMRES.
#pragma once
#include "common.hpp"

namespace amgs {

constexpr int FT = 256;       // threads per workgroup for fine-lattice kernels
constexpr int SAP_MAXM = 8;   // max restart length of the SAP block GMRES
constexpr int OUT_MAXM = 32;  // max restart length of the outer FGMRES
constexpr int OUT_SPT = 4;    // tiles of FT sites per workgroup in the outer Arnoldi step A

// Scalar state of every SAP item (one Schwarz block x one right-hand side),
// in device memory between the launches of one SAP sweep.
struct SapState {
    int *maxIt, *jbreak, *done;
    double *normb, *s;          // s: [item][M+1]
    double* sig0;               // FP16 storage: power-of-two scale of the stored Vt_0 [item]
    c64 *Hraw, *Hrot;           // [item][(M+1)*M]  (row i, column j at i*M+j)
    c64 *G;                     // [item][(M+1)*(M+1)]  G[i][k] = <V_i, V_k>, k < i
    c64 *gpre, *gpost, *cn, *sn;
    c64 *qa;                    // [item][nchunk][2M]   <w,Vt_i>, <Vt_j,Vt_k>
    double *qb;                 // [item][nchunk]       |w|^2
    double *qr;                 // [item][nchunk][2]    |r|^2, |b|^2
};

// Scalar state of the outer FGMRES, one slot per right-hand side.
struct OuterState {
    int *maxIt, *jbreak, *done, *active, *iters, *nhist;
    double *normb, *s, *hist;  // s [rhs][M+1]; hist [rhs][hcap]
    c64 *R, *G, *g, *cn, *sn, *eta;  // R [rhs][M][M], G [rhs][M+1][M+1]
    c64* qa;     // [rhs][ntile_a][2M+1]
    double* qb;  // [rhs][ntile]
    double* qr;  // [rhs][ntile][2]
    int M, ntile, hcap;
    int ntile_a;  // workgroups per rhs of the Arnoldi step A (OUT_SPT tiles of FT sites each)
    // several gauge configurations: rhs r's links at U + r * ucs, m0 + 2 = mmv[r]
    size_t ucs;
    const double* mmv;
};

}  // namespace amgs
