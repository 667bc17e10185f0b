// amgs_solve: command-line driver with the reference's file and console
// surface (src/main.cpp:51-131, src/tests.cpp:67-82, src/variables.cpp:170-221,
// src/fgmres.cpp:65-79), running the solve through libamgs on the GPU.
//
//   reads   parameters.dat (6 columns), a gauge field and a right-hand side in
//           the "x t mu re im" text format (or generates them from seeds);
//   prints  the parameter banner, the set-up time, "FGMRES converged in k
//           cycles Error e" / "With n iterations" and the job time;
//   writes  Residuals_FGMRES_<Nx>x<Nt>_Levels<L>_<K|V>cycle.res on convergence
//           and, with --x-out, the solution in the right-hand-side format.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/amgs.h"

namespace {

void usage() {
    std::fprintf(stderr,
                 "usage: amgs_solve --nx N --nt N [options]\n"
                 "  --params FILE      parameters.dat (default ../parameters.dat)\n"
                 "  --conf FILE        gauge field, x t mu re im per line\n"
                 "  --beta B --sweeps S --seed K   generate the field instead (Metropolis, cold start)\n"
                 "  --rhs FILE         right-hand side, x t mu re im per line\n"
                 "  --rhs-seed K       Gaussian right-hand side(s) instead (default 1)\n"
                 "  --nrhs R           solve R Gaussian right-hand sides in one batch\n"
                 "  --m0 M             bare mass (default -0.18840579710144945)\n"
                 "  --cycle K|V        multigrid cycle (default K)\n"
                 "  --nit N            bootstrap iterations of the set-up (default 0)\n"
                 "  --precision fp64|mixed|mixed16\n"
                 "  --restart M --restarts K --tol T   outer FGMRES (default 20, 50, 1e-10)\n"
                 "  --x-out FILE       write the solution of right-hand side 0\n"
                 "  --quiet            no parameter banner\n");
}

struct Opts {
    int nx = 0, nt = 0, nrhs = 1, nit = 0, restart = 20, restarts = 50, sweeps = 500, precision = AMGS_PREC_FP64;
    double m0 = -0.18840579710144945, beta = 2.0, tol = 1e-10;
    unsigned long long seed = 1, rhs_seed = 1;
    bool vcycle = false, quiet = false;
    std::string params = "../parameters.dat", conf, rhs, xout;
};

bool parse(int argc, char** argv, Opts& o) {
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto val = [&]() -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "amgs_solve: %s needs a value\n", a.c_str());
                return nullptr;
            }
            return argv[++i];
        };
        const char* v = nullptr;
        if (a == "--quiet") { o.quiet = true; continue; }
        if (a == "-h" || a == "--help") return false;
        if (!(v = val())) return false;
        if (a == "--nx") o.nx = std::atoi(v);
        else if (a == "--nt") o.nt = std::atoi(v);
        else if (a == "--params") o.params = v;
        else if (a == "--conf") o.conf = v;
        else if (a == "--beta") o.beta = std::atof(v);
        else if (a == "--sweeps") o.sweeps = std::atoi(v);
        else if (a == "--seed") o.seed = std::strtoull(v, nullptr, 10);
        else if (a == "--rhs") o.rhs = v;
        else if (a == "--rhs-seed") o.rhs_seed = std::strtoull(v, nullptr, 10);
        else if (a == "--nrhs") o.nrhs = std::atoi(v);
        else if (a == "--m0") o.m0 = std::atof(v);
        else if (a == "--cycle") {
            if (std::strcmp(v, "K") && std::strcmp(v, "V")) { std::fprintf(stderr, "amgs_solve: --cycle K or V\n"); return false; }
            o.vcycle = v[0] == 'V';
        } else if (a == "--nit") o.nit = std::atoi(v);
        else if (a == "--precision") {
            if (!std::strcmp(v, "fp64")) o.precision = AMGS_PREC_FP64;
            else if (!std::strcmp(v, "mixed")) o.precision = AMGS_PREC_MIXED;
            else if (!std::strcmp(v, "mixed16")) o.precision = AMGS_PREC_MIXED16;
            else { std::fprintf(stderr, "amgs_solve: --precision fp64, mixed or mixed16\n"); return false; }
        } else if (a == "--restart") o.restart = std::atoi(v);
        else if (a == "--restarts") o.restarts = std::atoi(v);
        else if (a == "--tol") o.tol = std::atof(v);
        else if (a == "--x-out") o.xout = v;
        else {
            std::fprintf(stderr, "amgs_solve: unknown option %s\n", a.c_str());
            return false;
        }
    }
    if (o.nx <= 0 || o.nt <= 0) {
        std::fprintf(stderr, "amgs_solve: --nx and --nt are required\n");
        return false;
    }
    if (o.nrhs < 1 || (!o.rhs.empty() && o.nrhs != 1)) {
        std::fprintf(stderr, "amgs_solve: --nrhs must be >= 1 (and 1 with --rhs)\n");
        return false;
    }
    return true;
}

// the reference's parameter banner (src/variables.cpp:170-221), from the
// numbers the library uses
void banner(const amgs_params& p) {
    const char* line = "---------------------------------------------------------------------------------------";
    std::printf("******************* AMG for the Dirac matrix in the Schwinger model *******************\n");
    std::printf("| Nx = %d Nt = %d\n", p.nx, p.nt);
    std::printf("| Lattice dimension = %d\n", p.nx * p.nt);
    std::printf("| Number of entries of the Dirac matrix = (%d)^2\n", 2 * p.nx * p.nt);
    std::printf("| Bare mass parameter m0 = %g\n%s\n", p.m0, line);
    std::printf("* Blocks, aggregates and test vectors at each level\n");
    for (int l = 0; l < p.levels - 1; l++)
        std::printf("| Level %d Block X %d Block T %d Ntest %d Nagg %d Number of lattice blocks %d Schwarz Block T %d "
                    "Schwarz Block X %d\n",
                    l, p.bx[l], p.bt[l], p.ntest[l], 2 * p.bx[l] * p.bt[l], p.bx[l] * p.bt[l], p.sapt[l], p.sapx[l]);
    std::printf("%s\n* Sites and degrees of freedom at each level\n", line);
    int nx = p.nx, nt = p.nt, dof = 2;
    for (int l = 0; l < p.levels; l++) {
        std::printf("| Level %d Nsites %d Nxsites %d NtSites %d DOF %d Colors %d\n", l, nx * nt, nx, nt, dof,
                    l == 0 ? 1 : dof / 2);
        if (l < p.levels - 1) {
            nx = p.bx[l];
            nt = p.bt[l];
            dof = 2 * p.ntest[l];
        }
    }
    std::printf("%s\n", line);
    std::printf("| GMRES restart length for SAP blocks = %d\n| GMRES restarts for SAP blocks = %d\n"
                "| GMRES tolerance for SAP blocks = %g\n%s\n",
                p.sap_m, p.sap_restarts, p.sap_gmres_tol, line);
    std::printf("* AMG parameters\n| Cycle = %s\n| Number of levels = %d\n", p.cycle == 0 ? "V-cycle" : "K-cycle",
                p.levels);
    std::printf("| nu1 (pre-smoothing) = %d nu2 (post-smoothing) = %d\n", p.nu1, p.nu2);
    std::printf("| Number of iterations for improving the interpolator = %d\n", p.nit);
    std::printf("| Restart length of GMRES at the coarse level = %d\n| Restarts of GMRES at the coarse level = %d\n"
                "| GMRES tolerance for the coarse level solution = %g\n",
                p.crs_m, p.crs_restarts, p.crs_tol);
    std::printf("* FGMRES with AMG preconditioning parameters\n| FGMRES restart length = %d\n| FGMRES restarts = %d\n"
                "| FGMRES tolerance = %g\n",
                p.out_m, p.out_restarts, p.out_tol);
    std::printf("*****************************************************************************************************\n");
}

int fail(amgs_ctx* c, const char* what, int rc) {
    std::fprintf(stderr, "amgs_solve: %s failed (%d): %s\n", what, rc, c ? amgs_last_error(c) : "");
    if (c) amgs_destroy(c);
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    Opts o;
    if (!parse(argc, argv, o)) {
        usage();
        return 1;
    }
    amgs_params p;
    amgs_default_params(&p);
    int rc = amgs_read_parameters(o.params.c_str(), &p);
    if (rc != AMGS_OK) {
        std::fprintf(stderr, "amgs_solve: cannot read %s (%d)\n", o.params.c_str(), rc);
        return 1;
    }
    p.nx = o.nx;
    p.nt = o.nt;
    p.m0 = o.m0;
    p.cycle = o.vcycle ? 0 : 1;
    p.nit = o.nit;
    p.precision = o.precision;
    p.max_rhs = o.nrhs;
    p.out_m = o.restart;
    p.out_restarts = o.restarts;
    p.out_tol = o.tol;

    const size_t ns = (size_t)o.nx * o.nt;
    std::vector<double> U(4 * ns), b(4 * ns * o.nrhs), x(4 * ns * o.nrhs);
    if (!o.conf.empty()) {
        if ((rc = amgs_read_conf(o.conf.c_str(), o.nx, o.nt, U.data())) != AMGS_OK) return fail(nullptr, "read conf", rc);
    } else {
        double plaq = 0;
        if ((rc = amgs_gauge_metropolis(o.nx, o.nt, o.beta, o.sweeps, o.seed, U.data(), &plaq)) != AMGS_OK)
            return fail(nullptr, "gauge generation", rc);
        if (!o.quiet) std::printf("| Generated gauge field: beta = %g, %d sweeps, plaquette = %.6f\n", o.beta, o.sweeps, plaq);
    }
    if (!o.rhs.empty()) {
        if ((rc = amgs_read_rhs(o.rhs.c_str(), o.nx, o.nt, b.data())) != AMGS_OK) return fail(nullptr, "read rhs", rc);
    } else {
        for (int r = 0; r < o.nrhs; r++)
            if ((rc = amgs_gaussian((int64_t)2 * ns, o.rhs_seed + r, b.data() + 4 * ns * r)) != AMGS_OK)
                return fail(nullptr, "rhs generation", rc);
    }
    if (!o.quiet) banner(p);

    std::printf("--------------FGMRES with AMG --------------\n");
    const auto t0 = std::chrono::steady_clock::now();
    amgs_ctx* c = nullptr;
    if ((rc = amgs_create(&p, &c)) != AMGS_OK) return fail(c, "amgs_create", rc);
    if ((rc = amgs_set_gauge(c, U.data())) != AMGS_OK) return fail(c, "amgs_set_gauge", rc);
    if ((rc = amgs_setup(c)) != AMGS_OK) return fail(c, "amgs_setup", rc);
    const auto t1 = std::chrono::steady_clock::now();
    std::printf("[MPI Process 0] Elapsed time for Set-up phase = %g seconds\n",
                std::chrono::duration<double>(t1 - t0).count());
    const int hcap = o.restart * o.restarts;
    std::vector<int> iters(o.nrhs), nhist(o.nrhs);
    std::vector<double> relres(o.nrhs), hist((size_t)o.nrhs * hcap);
    rc = amgs_solve(c, o.nrhs, b.data(), x.data(), iters.data(), relres.data(), hist.data(), hcap, nhist.data());
    if (rc < 0) return fail(c, "amgs_solve", rc);
    const auto t2 = std::chrono::steady_clock::now();
    for (int r = 0; r < o.nrhs; r++) {
        double nb = 0;
        for (size_t i = 0; i < 4 * ns; i++) nb += b[4 * ns * r + i] * b[4 * ns * r + i];
        const double err = relres[r] * std::sqrt(nb);
        if (o.nrhs > 1) std::printf("right-hand side %d: ", r);
        if (relres[r] < o.tol) {
            std::printf("FGMRES converged in %d cycles Error %g\n", (iters[r] - 1) / o.restart + 1, err);
            std::printf("With %d iterations\n", iters[r]);
        } else {
            std::printf("FGMRES did not converge in %d cycles Error %g\n", o.restarts, err);
        }
    }
    if (relres[0] < o.tol) {
        char name[256];
        std::snprintf(name, sizeof name, "Residuals_FGMRES_%dx%d_Levels%d_%s.res", o.nx, o.nt, p.levels,
                      o.vcycle ? "Vcycle" : "Kcycle");
        if ((rc = amgs_write_residuals(name, hist.data(), nhist[0])) != AMGS_OK) return fail(c, "residual file", rc);
    }
    if (!o.xout.empty() && (rc = amgs_write_rhs(o.xout.c_str(), o.nx, o.nt, x.data())) != AMGS_OK)
        return fail(c, "write x", rc);
    std::printf("[MPI process 0] time elapsed during the job: %.4fs.\n", std::chrono::duration<double>(t2 - t0).count());
    amgs_destroy(c);
    return 0;
}
