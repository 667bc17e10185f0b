// Host-only parts of libamgs: parameter file, workload generators, file
// formats, and the test-vector random stream.  No HIP calls here.
#include "host_util.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/amgs.h"

namespace amgs {

// ---------------------------------------------------------------------------
// test-vector stream (src/amg.cpp:8-35): std::mt19937(seed) (its output is
// fixed by the C++ standard) through generate_canonical<double, 53> as the
// standard specifies it for a 32-bit engine: two words, w0 + w1 * 2^32, over
// 2^64, kept below 1; then uniform_real_distribution(-1, 1): u * 2 - 1.
// Each complex entry consumes two numbers; g++ evaluates the imaginary
// operand of dist(g) + I*dist(g) first, so the first number is the imaginary
// part (pinned against the reference's own test vectors).
// ---------------------------------------------------------------------------
struct TvStream::Impl {
    std::mt19937 g;
    explicit Impl(uint32_t seed) : g(seed) {}
};
TvStream::TvStream(uint32_t seed) : p(new Impl(seed)) {}
TvStream::~TvStream() { delete p; }
double TvStream::uniform() {
    double sum = (double)p->g();
    sum += (double)p->g() * 4294967296.0;
    double u = sum / 18446744073709551616.0;
    if (u >= 1.0) u = std::nextafter(1.0, 0.0);
    return u * 2.0 + (-1.0);
}
void TvStream::fill(double* out, size_t ncomplex) {
    for (size_t i = 0; i < ncomplex; i++) {
        const double im = uniform();
        const double re = uniform();
        out[2 * i] = re;
        out[2 * i + 1] = im;
    }
}

// ---------------------------------------------------------------------------
// counter-based random numbers for the generators: splitmix64 finaliser of
// (seed, stream, counter); order-independent, so threads may run in any order
// ---------------------------------------------------------------------------
static inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t seed, uint64_t stream, uint64_t ctr) {
    const uint64_t h = mix64(seed ^ mix64(stream * 0x632be59bd9b4e019ull + ctr));
    return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0);  // (0, 1)
}

bool all_finite(const double* p, size_t n) {
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad) if (n > (1u << 16))
    for (long long i = 0; i < (long long)n; i++) bad |= !std::isfinite(p[i]);
    return bad == 0;
}

void par_copy(void* dst, const void* src, size_t bytes) {
    const size_t B = size_t(1) << 20;
    const long long nb = (long long)((bytes + B - 1) / B);
#pragma omp parallel for schedule(static) if (nb > 1)
    for (long long i = 0; i < nb; i++) {
        const size_t o = (size_t)i * B;
        std::memcpy(static_cast<char*>(dst) + o, static_cast<const char*>(src) + o, std::min(B, bytes - o));
    }
}

}  // namespace amgs

using namespace amgs;

extern "C" int amgs_gaussian(int64_t n, uint64_t seed, double* out) {
    if (n < 0 || (!out && n > 0)) return AMGS_ERR_ARG;
    for (int64_t i = 0; i < n; i++) {
        const double a = u01(seed, 1, 2 * (uint64_t)i), b = u01(seed, 1, 2 * (uint64_t)i + 1);
        const double r = std::sqrt(-2.0 * std::log(a));
        out[2 * i] = r * std::cos(2.0 * M_PI * b);
        out[2 * i + 1] = r * std::sin(2.0 * M_PI * b);
    }
    return AMGS_OK;
}

// Quenched U(1) Metropolis with the plaquette action S = beta sum_x (1 - cos P(x)),
// P(x,t) = th_t(x,t) + th_x(x,t+1) - th_t(x+1,t) - th_x(x,t), cold start.
// A sweep visits time links by x parity, then space links by t parity; links
// in one such class share no plaquette, so the order inside a class is free.
extern "C" int amgs_gauge_metropolis(int nx, int nt, double beta, int sweeps, uint64_t seed, double* U, double* plaq) {
    if (nx < 2 || nt < 2 || (nx & 1) || (nt & 1) || sweeps < 0 || !U) return AMGS_ERR_ARG;
    const int ns = nx * nt;
    std::vector<double> th((size_t)ns * 2, 0.0);  // th[n*2+mu], mu 0 = t, 1 = x
    auto I = [&](int x, int t) { return ((x + nx) % nx) * nt + (t + nt) % nt; };
    auto P = [&](int x, int t) {
        return th[2 * I(x, t)] + th[2 * I(x, t + 1) + 1] - th[2 * I(x + 1, t)] - th[2 * I(x, t) + 1];
    };
    const double step = 1.0;
    for (int sw = 0; sw < sweeps; sw++) {
        for (int cls = 0; cls < 4; cls++) {
            const int mu = cls >> 1, par = cls & 1;
#pragma omp parallel for schedule(static)
            for (int n = 0; n < ns; n++) {
                const int x = n / nt, t = n % nt;
                if (((mu == 0 ? x : t) & 1) != par) continue;
                const uint64_t ctr = ((uint64_t)sw * 2 + mu) * (uint64_t)ns + n;
                double& l = th[2 * n + mu];
                const double old = l;
                // the two plaquettes containing the link, before and after
                double s_old, s_new;
                if (mu == 0) {
                    s_old = std::cos(P(x, t)) + std::cos(P(x - 1, t));
                    l = old + step * (2.0 * u01(seed, 2, 2 * ctr) - 1.0);
                    s_new = std::cos(P(x, t)) + std::cos(P(x - 1, t));
                } else {
                    s_old = std::cos(P(x, t)) + std::cos(P(x, t - 1));
                    l = old + step * (2.0 * u01(seed, 2, 2 * ctr) - 1.0);
                    s_new = std::cos(P(x, t)) + std::cos(P(x, t - 1));
                }
                const double dS = -beta * (s_new - s_old);
                if (!(dS <= 0.0 || u01(seed, 2, 2 * ctr + 1) < std::exp(-dS))) l = old;
            }
        }
    }
    double acc = 0;
    for (int x = 0; x < nx; x++)
        for (int t = 0; t < nt; t++) acc += std::cos(P(x, t));
    if (plaq) *plaq = acc / ns;
    for (int n = 0; n < ns; n++)
        for (int mu = 0; mu < 2; mu++) {
            U[4 * n + 2 * mu] = std::cos(th[2 * n + mu]);
            U[4 * n + 2 * mu + 1] = std::sin(th[2 * n + mu]);
        }
    return AMGS_OK;
}

extern "C" int amgs_default_params(amgs_params* p) {
    if (!p) return AMGS_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    p->m0 = -0.18840579710144945;
    p->nu1 = 0;
    p->nu2 = 2;
    p->sap_m = 5;
    p->sap_restarts = 5;
    p->sap_gmres_tol = 1e-3;
    p->sap_tol = 1e-10;
    p->kc_m = 5;
    p->kc_restarts = 2;
    p->kc_tol = 0.1;
    p->crs_m = 20;
    p->crs_restarts = 20;
    p->crs_tol = 0.1;
    p->out_m = 20;
    p->out_restarts = 50;
    p->out_tol = 1e-10;
    p->tv_sap_iters = 4;
    p->tv_seed = 50;
    p->precision = AMGS_PREC_FP64;
    p->max_rhs = 1;
    p->cycle = 1;
    p->nit = 0;
    return AMGS_OK;
}

extern "C" int amgs_read_parameters(const char* path, amgs_params* p) {
    if (!path || !p) return AMGS_ERR_ARG;
    std::ifstream f(path);
    if (!f) return AMGS_ERR_IO;
    int level, bx, bt, nv, sx, st, rows = 0;
    while (f >> level >> bx >> bt >> nv >> sx >> st) {
        if (level < 0 || level >= AMGS_MAX_LEVELS - 1) return AMGS_ERR_GEOMETRY;
        p->bx[level] = bx;
        p->bt[level] = bt;
        p->ntest[level] = nv;
        p->sapx[level] = sx;
        p->sapt[level] = st;
        rows = level + 1 > rows ? level + 1 : rows;
    }
    if (rows == 0) return AMGS_ERR_IO;
    p->levels = rows + 1;
    return AMGS_OK;
}

// "x t mu re im" lines (src/gauge_conf.cpp:25-28, src/variables.cpp:125-127)
static int read_xtmu(const char* path, int nx, int nt, double* out) {
    std::ifstream f(path);
    if (!f) return AMGS_ERR_IO;
    int x, t, mu;
    double re, im;
    while (f >> x >> t >> mu >> re >> im) {
        if (x < 0 || x >= nx || t < 0 || t >= nt || mu < 0 || mu > 1) return AMGS_ERR_IO;
        const size_t i = ((size_t)x * nt + t) * 2 + mu;
        out[2 * i] = re;
        out[2 * i + 1] = im;
    }
    return AMGS_OK;
}
static int write_xtmu(const char* path, int nx, int nt, const double* v) {
    FILE* f = std::fopen(path, "w");
    if (!f) return AMGS_ERR_IO;
    for (int n = 0; n < nx * nt; n++)
        for (int mu = 0; mu < 2; mu++)
            std::fprintf(f, "%d%30d%30d%30.17e%30.17e\n", n / nt, n % nt, mu, v[4 * n + 2 * mu], v[4 * n + 2 * mu + 1]);
    std::fclose(f);
    return AMGS_OK;
}
extern "C" int amgs_read_conf(const char* path, int nx, int nt, double* U) { return read_xtmu(path, nx, nt, U); }
extern "C" int amgs_write_conf(const char* path, int nx, int nt, const double* U) { return write_xtmu(path, nx, nt, U); }
extern "C" int amgs_read_rhs(const char* path, int nx, int nt, double* b) { return read_xtmu(path, nx, nt, b); }
extern "C" int amgs_write_rhs(const char* path, int nx, int nt, const double* b) { return write_xtmu(path, nx, nt, b); }

// Residuals_FGMRES_*.res: "n" then the value in %30.17e (src/variables.cpp:101-115)
extern "C" int amgs_write_residuals(const char* path, const double* hist, int n) {
    FILE* f = std::fopen(path, "w");
    if (!f) return AMGS_ERR_IO;
    for (int i = 0; i < n; i++) std::fprintf(f, "%d%30.17e\n", i, hist[i]);
    std::fclose(f);
    return AMGS_OK;
}
