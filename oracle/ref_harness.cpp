// Driver for the reference DD-alpha-AMG program, used ONLY to produce golden
// vectors for the oracle and the HIP path (test infrastructure, never shipped,
// never run on a GPU host).  It replaces the reference's src/main.cpp:47-179:
// instead of reading HMC configurations from hard-coded external paths it reads
// a gauge field and vectors from raw little-endian complex128 files and dumps
// the reference's outputs as raw complex128 files.
//
// The reference is compiled unmodified from its own sources (see Makefile).
// Its compile-time lattice is the shipped include/config.h: NS=NT=256, LEVELS=4.
//
// Modes
//   setup <params.dat> <m0> <tv_sap_iters> <U.bin> <vecdir> <outdir>
//       - D_phi / D_dagger_phi on vecdir/psi.bin           (src/dirac_operator.cpp:3-54)
//       - AlgebraicMG setUpPhase(0)                        (src/amg.cpp:3-88)
//       - interpolator columns and coarse links per level  (src/level.cpp:242-301, 373-417)
//       - D_operator, P_v, Pt_v, SAP on vecdir/v<l>.bin    (src/level.cpp:161-237, src/sap.cpp:64-174)
//       - k_cycle(0) on vecdir/v0.bin                      (src/amg.cpp:135-178)
//   solve <params.dat> <m0> <tv_sap_iters> <U.bin> <b.bin> <outdir> [restarts]
//       - FGMRES_AMG(Ntot,2,20,restarts,1e-10).fgmres       (include/amg.h:99-145, src/fgmres.cpp:5-88)
//
// Solver overrides follow the reference main (src/main.cpp:61-64,129):
// K-cycle, Nit = 0.  SAP coverage is made complete on one rank by setting
// SAPV::sap_blocks_per_proc to a large value (src/sap.cpp:77-78).
#include "params.h"
#include "boundary.h"
#include "amg.h"
#include "bi_cgstab.h"
#include "conjugate_gradient.h"
#include <iostream>
#include <sstream>
#include "mpi.h"
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include <unistd.h>

static std::vector<c_double> read_bin(const std::string& path, size_t n) {
    std::vector<c_double> v(n);
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path.c_str()); exit(2); }
    size_t got = fread(v.data(), sizeof(c_double), n, f);
    fclose(f);
    if (got != n) { fprintf(stderr, "%s: expected %zu complex, got %zu\n", path.c_str(), n, got); exit(2); }
    return v;
}

static void write_bin(const std::string& path, const c_double* p, size_t n) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(2); }
    fwrite(p, sizeof(c_double), n, f);
    fclose(f);
}

static spinor to_spinor(const std::vector<c_double>& flat, int n1, int n2) {
    spinor s(n1, c_vector(n2, 0));
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) s[i][j] = flat[(size_t)i * n2 + j];
    return s;
}

static void write_spinor(const std::string& path, const spinor& s) {
    std::vector<c_double> flat;
    for (auto& row : s) flat.insert(flat.end(), row.begin(), row.end());
    write_bin(path, flat.data(), flat.size());
}

static void common_init(const char* params, double m0, int tv_iters) {
    readParameters(params);
    Coordinates();
    boundary();
    // cycle type and bootstrap iterations: K-cycle and Nit = 0 as main() runs
    // them (src/main.cpp:58,129) unless AMGS_REF_CYCLE / AMGS_REF_NIT say otherwise
    const char* cy = getenv("AMGS_REF_CYCLE");
    const char* nit = getenv("AMGS_REF_NIT");
    AMGV::cycle = cy ? atoi(cy) : 1;
    AMGV::Nit = nit ? atoi(nit) : 0;
    AMGV::SAP_test_vectors_iterations = tv_iters;
    mass::m0 = m0;
    // SAP blocks per rank: every block on each rank (the default here), or
    // AMGS_REF_BPP per rank as the reference's own runs distribute them
    // (1 block per rank and colour with mpirun -np 8, src/variables.cpp:51,
    // script.sh:38)
    const char* bpp = getenv("AMGS_REF_BPP");
    SAPV::sap_blocks_per_proc = bpp ? atoi(bpp) : 1 << 20;
}

static GaugeConf load_conf(const std::string& path) {
    GaugeConf g(LV::Nx, LV::Nt);
    auto flat = read_bin(path, (size_t)LV::Ntot * 2);
    g.setGconf(to_spinor(flat, LV::Ntot, 2));
    return g;
}

static int mode_setup(int argc, char** argv) {
    if (argc < 8) return 1;
    double m0 = atof(argv[3]);
    common_init(argv[2], m0, atoi(argv[4]));
    GaugeConf G = load_conf(argv[5]);
    std::string vd = argv[6], od = argv[7];
    const int L = AMGV::levels;

    {   // fine operator in the D_phi form used by the outer FGMRES
        spinor psi = to_spinor(read_bin(vd + "/psi.bin", (size_t)LV::Ntot * 2), LV::Ntot, 2);
        spinor out(LV::Ntot, c_vector(2, 0));
        D_phi(G.Conf, psi, out, m0);
        write_spinor(od + "/Dphi.bin", out);
        D_dagger_phi(G.Conf, psi, out, m0);
        write_spinor(od + "/Ddagphi.bin", out);
    }

    AlgebraicMG amg(G, m0, AMGV::nu1, AMGV::nu2);
    amg.setUpPhase(AMGV::Nit);

    for (int l = 0; l < L - 1; l++) {
        Level* lv = amg.levels[l];
        for (int i = 0; i < lv->Ntest; i++)
            write_spinor(od + "/tv_l" + std::to_string(l) + "_" + std::to_string(i) + ".bin",
                         lv->interpolator_columns[i]);
    }
    for (int l = 1; l < L; l++) {
        Level* lv = amg.levels[l];
        write_bin(od + "/G1_l" + std::to_string(l) + ".bin", lv->G1.data(), lv->G1.size());
        write_bin(od + "/G2_l" + std::to_string(l) + ".bin", lv->G2.data(), lv->G2.size());
        write_bin(od + "/G3_l" + std::to_string(l) + ".bin", lv->G3.data(), lv->G3.size());
    }
    for (int l = 0; l < L; l++) {
        int ns = LevelV::Nsites[l], dof = LevelV::DOF[l];
        spinor v = to_spinor(read_bin(vd + "/v" + std::to_string(l) + ".bin", (size_t)ns * dof), ns, dof);
        spinor out(ns, c_vector(dof, 0));
        amg.levels[l]->D_operator(v, out);
        write_spinor(od + "/Dop_l" + std::to_string(l) + ".bin", out);
        if (l < L - 1) {
            int nsc = LevelV::Nsites[l + 1], dofc = LevelV::DOF[l + 1];
            spinor vc = to_spinor(read_bin(vd + "/v" + std::to_string(l + 1) + ".bin", (size_t)nsc * dofc), nsc, dofc);
            spinor pout(ns, c_vector(dof, 0));
            amg.levels[l]->P_v(vc, pout);
            write_spinor(od + "/P_l" + std::to_string(l) + ".bin", pout);
            spinor rout(nsc, c_vector(dofc, 0));
            amg.levels[l]->Pt_v(v, rout);
            write_spinor(od + "/Pt_l" + std::to_string(l) + ".bin", rout);
            spinor x = to_spinor(read_bin(vd + "/x" + std::to_string(l) + ".bin", (size_t)ns * dof), ns, dof);
            amg.levels[l]->sap_l.SAP(v, x, 2, SAPV::sap_blocks_per_proc, false);
            write_spinor(od + "/SAP_l" + std::to_string(l) + ".bin", x);
        }
    }
    {
        spinor eta = to_spinor(read_bin(vd + "/v0.bin", (size_t)LV::Ntot * 2), LV::Ntot, 2);
        spinor psi(LV::Ntot, c_vector(2, 0));
        amg.k_cycle(0, eta, psi);
        write_spinor(od + "/kcycle.bin", psi);
        spinor psv(LV::Ntot, c_vector(2, 0));
        amg.v_cycle(0, eta, psv);
        write_spinor(od + "/vcycle.bin", psv);
    }
    return 0;
}

static int mode_solve(int argc, char** argv) {
    if (argc < 8) return 1;
    double m0 = atof(argv[3]);
    common_init(argv[2], m0, atoi(argv[4]));
    GaugeConf G = load_conf(argv[5]);
    spinor b = to_spinor(read_bin(argv[6], (size_t)LV::Ntot * 2), LV::Ntot, 2);
    std::string od = argv[7];
    int restarts = argc > 8 ? atoi(argv[8]) : FGMRESV::fgmres_restarts;
    if (chdir(od.c_str()) != 0) { fprintf(stderr, "chdir %s failed\n", od.c_str()); return 2; }
    spinor x0(LV::Ntot, c_vector(2, 0)), x(LV::Ntot, c_vector(2, 0));
    double t0 = MPI_Wtime();
    FGMRES_AMG f(LV::Ntot, 2, FGMRESV::fgmres_restart_length, restarts, FGMRESV::fgmres_tolerance, G, m0);
    double t1 = MPI_Wtime();
    int it = f.fgmres(b, x0, x, true, true);
    double t2 = MPI_Wtime();
    write_spinor("x.bin", x);
    FILE* fo = fopen("result.txt", "w");
    fprintf(fo, "iterations %d\nsetup_s %.6f\nsolve_s %.6f\n", it, t1 - t0, t2 - t1);
    fclose(fo);
    return 0;
}

// applyMultilevel(it, b, x, tol) from x0 (src/amg.cpp:181-231); writes x.bin
static int mode_multigrid(int argc, char** argv) {
    if (argc < 11) return 1;
    double m0 = atof(argv[3]);
    common_init(argv[2], m0, atoi(argv[4]));
    GaugeConf G = load_conf(argv[5]);
    spinor b = to_spinor(read_bin(argv[6], (size_t)LV::Ntot * 2), LV::Ntot, 2);
    spinor x = to_spinor(read_bin(argv[7], (size_t)LV::Ntot * 2), LV::Ntot, 2);
    std::string od = argv[8];
    const int it = atoi(argv[9]);
    const double tol = atof(argv[10]);
    AlgebraicMG amg(G, m0, AMGV::nu1, AMGV::nu2);
    amg.setUpPhase(AMGV::Nit);
    amg.applyMultilevel(it, b, x, tol, false);
    write_spinor(od + "/x.bin", x);
    return 0;
}

// comparison solvers on the fine lattice (src/bi_cgstab.cpp:5-87,
// src/conjugate_gradient.cpp:5-65, include/sap.h:194-225, include/fgmres.h:126-148):
// writes x.bin and result.txt (iterations), residual files where the
// reference writes them
static int mode_krylov(int argc, char** argv) {
    if (argc < 10) return 1;
    const std::string kind = argv[1];
    double m0 = atof(argv[3]);
    common_init(argv[2], m0, atoi(argv[4]));
    GaugeConf G = load_conf(argv[5]);
    spinor b = to_spinor(read_bin(argv[6], (size_t)LV::Ntot * 2), LV::Ntot, 2);
    std::string od = argv[7];
    const int p1 = atoi(argv[8]), p2 = atoi(argv[9]);
    if (chdir(od.c_str()) != 0) return 2;
    spinor x0(LV::Ntot, c_vector(2, 0)), x(LV::Ntot, c_vector(2, 0));
    long iters = -1;
    if (kind == "bicg") {
        x = bi_cgstab(&D_phi, LV::Ntot, 2, G.Conf, b, x0, m0, p1, 1e-10, false, true);
    } else if (kind == "cg") {
        CG::max_iter = p1;
        std::ostringstream cap;
        std::streambuf* old = std::cout.rdbuf(cap.rdbuf());
        const int ok = conjugate_gradient(G.Conf, b, x, m0);
        std::cout.rdbuf(old);
        std::istringstream in(cap.str());
        std::string w1, w2;
        in >> w1 >> w2 >> iters;  // "Converged in k iterations"
        if (!ok) iters = -1;
    } else if (kind == "fgmres_sap") {
        sap.set_params(G.Conf, m0);
        FGMRES_SAP f(LV::Ntot, 2, p1, p2, 1e-10, G.Conf, m0);
        iters = f.fgmres(b, x0, x, false, true);
    } else if (kind == "gmres") {
        FGMRES_fine_level f(LV::Ntot, 2, p1, p2, 1e-10, G.Conf, m0);
        iters = f.fgmres(b, x0, x, false, true);
    } else {
        return 1;
    }
    write_spinor("x.bin", x);
    FILE* fo = fopen("result.txt", "w");
    fprintf(fo, "iterations %ld\nsetup_s 0\nsolve_s 0\n", iters);
    fclose(fo);
    return 0;
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int rc = 1;
    if (argc > 1 && std::string(argv[1]) == "setup") rc = mode_setup(argc, argv);
    else if (argc > 1 && std::string(argv[1]) == "solve") rc = mode_solve(argc, argv);
    else if (argc > 1 && std::string(argv[1]) == "multigrid") rc = mode_multigrid(argc, argv);
    else if (argc > 1) rc = mode_krylov(argc, argv);
    if (rc == 1) fprintf(stderr, "usage: ref_harness setup|solve ...\n");
    MPI_Finalize();
    return rc;
}
