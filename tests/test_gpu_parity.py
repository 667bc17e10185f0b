"""GPU (libamgs, HIP on gfx950) against the CPU oracle on the same inputs.

Tolerances follow the reference's own parity bar: operator applies 1e-13
relative in FP64; full solves the same iteration count +-1 and solutions
within 1e-9 relative.  Set-up results (test vectors, coarse links) carry the
rounding of many reductions in a different order and of the SAP smoothing
(a Krylov method, whose rounding grows with each sweep), so they are compared
at 1e-9 relative; from the same raw test vectors the coarse operators, P and
P^dagger are compared at the operator tolerance 1e-13.
"""
import os
import subprocess

import numpy as np
import pytest

import amg_schwingermodel_amd as amgs
import oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    # name: (nx, nt, rows, m0, beta)
    "32_3lvl": (32, 32, [(0, 4, 4, 4, 4, 4), (1, 2, 2, 4, 2, 2)], -0.10, 2.0),
    "64_2lvl": (64, 64, [(0, 4, 4, 10, 4, 4)], -0.18, 2.0),
    "32_4lvl_odd": (32, 32, [(0, 4, 2, 3, 2, 1), (1, 2, 2, 2, 1, 2), (2, 1, 2, 2, 1, 1)], -0.05, 3.0),
    # non-square lattices: Schwarz-block membership by the reference's
    # getMandBlock rule (x = n / Nx), and block-flagged residuals by true
    # coordinates
    "32x16_2lvl": (32, 16, [(0, 8, 4, 4, 4, 2)], 0.10, 2.0),
    "16x32_3lvl": (16, 32, [(0, 4, 8, 4, 2, 4), (1, 2, 2, 2, 2, 2)], 0.0, 2.0),
    # one level-0 Schwarz block: it is both the red and the black block
    "32_1blk_3lvl": (32, 32, [(0, 4, 4, 4, 1, 1), (1, 2, 2, 4, 2, 2)], -0.10, 2.0),
    # Schwarz blocks of odd extent (15 x 15; no site pairs in the FP16 path)
    # and an odd 5 x 5 coarsest lattice
    "30_2lvl_odd_ext": (30, 30, [(0, 5, 5, 4, 2, 2)], -0.05, 2.0),
}


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make(case, setup=True, common=None, **kw):
    """GPU solver (parameters kw + common) and oracle (common) on one field."""
    nx, nt, rows, m0, beta = CASES[case]
    common = common or {}
    U, _ = amgs.gauge_metropolis(nx, nt, beta, 30, 1234)
    s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, **common, **kw), U)
    o = O.Oracle(O.default_params(nx, nt, rows, m0=m0, threads=8, **common), U)
    if setup:
        s.setup()
        o.setup(draw_order=1)
    return s, o, rows


def cvec(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def assert_history(hist, histo, b, tol=1e-10):
    """The |g_{j+1}| histories (src/fgmres.cpp:43-44) over their whole common
    prefix (the iteration counts may differ by one): 1e-6 relative, with an
    absolute floor of 1e-3 tol |b| for the entries at the tolerance."""
    k = min(len(hist), len(histo))
    assert k >= min(len(histo), 1) and abs(len(hist) - len(histo)) <= 1, (len(hist), len(histo))
    np.testing.assert_allclose(hist[:k], histo[:k], rtol=1e-6, atol=1e-3 * tol * np.linalg.norm(b))


@pytest.mark.parametrize("case", list(CASES))
def test_fine_operator(case):
    s, o, _ = make(case, setup=False)
    v = cvec(s.size(0), 1)
    assert rel(s.apply_D(0, v), o.dphi(v)) < 1e-13
    assert rel(s.apply_D(0, v, dagger=True), o.ddag(v)) < 1e-13
    # batched apply equals single applies
    V = np.stack([cvec(s.size(0), 2 + i) for i in range(3)])
    out = s.apply_D(0, V, nrhs=3).reshape(3, -1)
    for i in range(3):
        assert rel(out[i], o.dphi(V[i])) < 1e-13


@pytest.mark.parametrize("case", list(CASES))
def test_setup_hierarchy(case):
    s, o, rows = make(case)
    L = len(rows) + 1
    for l in range(L - 1):
        for i in range(rows[l][3]):
            assert rel(s.tv(l, i), o.tv(l, i)) < 1e-9, (l, i)
    for l in range(1, L):
        for which in range(3):
            for mu in range(2 if which else 1):
                assert rel(s.links(l, which, mu), o.links(l, which, mu)) < 1e-9, (l, which, mu)


@pytest.mark.parametrize("case", list(CASES))
def test_level_operators_from_same_tv(case):
    """Same raw test vectors on both sides, no smoothing: orthonormalisation,
    coarse links, P, P^dagger and the coarse stencils at operator tolerance."""
    s, o, rows = make(case, setup=False)
    L = len(rows) + 1
    tvs = [np.concatenate([cvec(o.size(l), 100 * l + i) for i in range(rows[l][3])]) for l in range(L - 1)]
    s.setup_from_tv(tvs, smooth=False)
    o.setup_from_tv(tvs, smooth=False)
    for l in range(L):
        v = cvec(s.size(l), 7 + l)
        assert rel(s.apply_D(l, v), o.dop(l, v)) < 1e-13, l
        if l < L - 1:
            vc = cvec(s.size(l + 1), 17 + l)
            assert rel(s.apply_P(l, vc), o.prolong(l, vc)) < 1e-13, l
            assert rel(s.apply_Pt(l, v), o.restrict(l, v)) < 1e-13, l


@pytest.mark.parametrize("case", list(CASES))
def test_sap(case):
    s, o, rows = make(case)
    L = len(rows) + 1
    for l in range(L - 1):
        v = cvec(s.size(l), 31 + l)
        x0 = cvec(s.size(l), 41 + l)
        xs = s.sap(l, v, x0, 2)
        xo, _ = o.sap(l, v, x0, 2)
        assert rel(xs, xo) < 1e-9, l


@pytest.mark.parametrize("case", list(CASES))
def test_kcycle(case):
    s, o, _ = make(case)
    eta = cvec(s.size(0), 51)
    assert rel(s.kcycle(eta), o.kcycle0(eta)) < 1e-9


@pytest.mark.parametrize("case", list(CASES))
def test_full_solve(case):
    s, o, _ = make(case)
    b = cvec(s.size(0), 61)
    x, it, rr, hist = s.solve(b)
    xo, ito, histo, rro = o.solve(b)
    assert abs(it - ito) <= 1, (it, ito)
    assert rr < 1e-10
    assert rel(x, xo) < 1e-9
    assert_history(hist, histo, b)


def test_batched_solve_matches_single():
    s, o, _ = make("32_3lvl", max_rhs=4)
    B = np.stack([cvec(s.size(0), 70 + i) for i in range(3)])
    X, its, rrs, _ = s.solve(B)
    for i in range(3):
        x1, it1, rr1, _ = s.solve(B[i])
        assert its[i] == it1
        assert rel(X[i], x1) < 1e-12
        assert rrs[i] < 1e-10


def test_mixed_precision_converges():
    s, o, _ = make("64_2lvl", precision=amgs.PREC_MIXED, max_rhs=2)
    b = cvec(s.size(0), 81)
    x, it, rr, _ = s.solve(b)
    xo, ito, _, _ = o.solve(b)
    assert rr < 1e-10
    assert rel(x, xo) < 1e-8
    assert it <= ito + 3


@pytest.mark.parametrize("case", list(CASES))
def test_mixed16_precision_converges(case):
    """FP16 Krylov basis in the level-0 SAP block GMRES (FP32 arithmetic,
    FP64 outer FGMRES): the batched solve reaches the FP64 tolerance, stays
    within 1e-8 of the oracle's solution and needs at most 3 more outer
    iterations; every right-hand side of the batch matches its single solve."""
    s, o, _ = make(case, precision=amgs.PREC_MIXED16, max_rhs=3)
    B = np.stack([cvec(s.size(0), 85 + i) for i in range(3)])
    X, its, rrs, _ = s.solve(B)
    for i in range(3):
        xo, ito, _, _ = o.solve(B[i])
        assert rrs[i] < 1e-10
        assert rel(X[i], xo) < 1e-8
        assert its[i] <= ito + 3, (its[i], ito)
    x1, it1, rr1, _ = s.solve(B[1])
    assert abs(it1 - its[1]) <= 1 and rr1 < 1e-10
    assert rel(x1, X[1]) < 1e-8


@pytest.mark.parametrize("case", list(CASES))
def test_vcycle(case):
    """V-cycle preconditioner (src/amg.cpp:90-132): one application at level 0
    and the full solve."""
    s, o, _ = make(case, common=dict(cycle=0))
    eta = cvec(s.size(0), 91)
    assert rel(s.kcycle(eta), o.vcycle(0, eta)) < 1e-9
    b = cvec(s.size(0), 92)
    x, it, rr, hist = s.solve(b)
    xo, ito, histo, _ = o.solve(b)
    assert abs(it - ito) <= 1, (it, ito)
    assert rr < 1e-10
    assert rel(x, xo) < 1e-9
    assert_history(hist, histo, b)


@pytest.mark.parametrize("cycle", [1, 0])
@pytest.mark.parametrize("case", list(CASES))
def test_bootstrap(case, cycle):
    """Bootstrap set-up with Nit = 2 (src/amg.cpp:52-84): test vectors, coarse
    links and the solve."""
    s, o, rows = make(case, common=dict(cycle=cycle, nit=2))
    L = len(rows) + 1
    for l in range(L - 1):
        for i in range(rows[l][3]):
            assert rel(s.tv(l, i), o.tv(l, i)) < 1e-9, (l, i)
    for l in range(1, L):
        assert rel(s.links(l, 0), o.links(l, 0)) < 1e-9, l
        assert rel(s.links(l, 1, 1), o.links(l, 1, 1)) < 1e-9, l
    b = cvec(s.size(0), 93)
    x, it, rr, hist = s.solve(b)
    xo, ito, histo, _ = o.solve(b)
    assert abs(it - ito) <= 1, (it, ito)
    assert rr < 1e-10
    assert rel(x, xo) < 1e-9
    assert_history(hist, histo, b)


def test_cli_matches_library(tmp_path):
    """amgs_solve reads parameters.dat, the conf and rhs files, prints the
    reference's convergence lines and writes the residual file and x."""
    nx, nt, rows, m0, beta = CASES["32_3lvl"]
    U, _ = amgs.gauge_metropolis(nx, nt, beta, 30, 1234)
    b = cvec(nx * nt * 2, 95)
    L = amgs.lib()
    assert L.amgs_write_conf(str(tmp_path / "U.ctxt").encode(), nx, nt, amgs._p(U)) == 0
    assert L.amgs_write_rhs(str(tmp_path / "b.rhs").encode(), nx, nt, amgs._p(b)) == 0
    (tmp_path / "parameters.dat").write_text("\n".join(" ".join(map(str, r)) for r in rows) + "\n")
    r = subprocess.run([amgs.cli_path(), "--nx", str(nx), "--nt", str(nt), "--params", "parameters.dat", "--conf",
                        "U.ctxt", "--rhs", "b.rhs", "--m0", repr(m0), "--x-out", "x.rhs"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    its = [int(l.split()[1]) for l in lines if l.startswith("With ") and l.endswith(" iterations")]
    s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0), U)
    s.setup()
    x, it, rr, hist = s.solve(b)
    assert its == [it]
    assert any(l.startswith("FGMRES converged in") for l in lines)
    xf = np.zeros_like(b)
    assert L.amgs_read_rhs(str(tmp_path / "x.rhs").encode(), nx, nt, amgs._p(xf)) == 0
    assert rel(xf, x) < 1e-12
    res = np.loadtxt(tmp_path / f"Residuals_FGMRES_{nx}x{nt}_Levels3_Kcycle.res", ndmin=2)
    np.testing.assert_allclose(res[:, 1], hist, rtol=1e-12)


@pytest.mark.parametrize("cycle", [1, 0])
def test_standalone_multigrid(cycle):
    """Stand-alone multigrid (src/amg.cpp:181-231) from a non-zero guess:
    fixed cycle count against the oracle, then convergence to 1e-10."""
    s, o, _ = make("32_3lvl", common=dict(cycle=cycle), max_rhs=2)
    b = cvec(s.size(0), 96)
    x0 = cvec(s.size(0), 97)
    x, cyc, rr = s.multigrid(b, x0, max_cycles=3, tol=1e-14)
    xo, cyco, rro = o.multigrid(b, x0, it=3, tol=1e-14)
    assert cyc == cyco == -3
    assert rel(x, xo) < 1e-9
    assert abs(rr - rro) <= 1e-6 * rro
    B = np.stack([b, cvec(s.size(0), 98)])
    X, cycs, rrs = s.multigrid(B, max_cycles=200, tol=1e-10)
    for i in range(2):
        xo, cyco, _ = o.multigrid(B[i], it=200, tol=1e-10)
        assert cyco > 0 and abs(int(cycs[i]) - cyco) <= 1, (cycs[i], cyco)
        assert rrs[i] < 1e-10
        assert rel(X[i], xo) < 1e-8


@pytest.mark.parametrize("case", ["32_3lvl", "64_2lvl"])
def test_comparison_solvers(case):
    """Bi-CGstab, CG on D D^dagger, FGMRES+SAP and plain GMRES on the fine
    lattice against the oracle (no multigrid set-up needed), batched."""
    nx, nt, rows, m0, beta = CASES[case]
    m0 = 0.1  # well-conditioned enough for the unpreconditioned methods
    U, _ = amgs.gauge_metropolis(nx, nt, beta, 30, 1234)
    s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, max_rhs=2), U)
    o = O.Oracle(O.default_params(nx, nt, rows, m0=m0, threads=8), U)
    B = np.stack([cvec(s.size(0), 110 + i) for i in range(2)])
    X, its, rrs, hists = s.bicgstab(B)
    for i in range(2):
        xo, ito, ho = o.bicgstab(B[i])
        assert abs(int(its[i]) - ito) <= max(1, ito // 50), (its[i], ito)
        assert rrs[i] < 1.5e-10 and rel(X[i], xo) < 1e-8
        k = min(len(hists[i]), len(ho), 10)
        np.testing.assert_allclose(hists[i][:k], ho[:k], rtol=1e-6)
    X, its, rrs = s.cg_normal(B)
    for i in range(2):
        xo, ito = o.cg(B[i])
        assert ito > 0 and abs(int(its[i]) - ito) <= max(1, ito // 50), (its[i], ito)
        assert rrs[i] < 1.5e-10 and rel(X[i], xo) < 1e-8
    for meth, ometh in ((s.fgmres_sap, o.fgmres_sap), (s.gmres, o.gmres_fine)):
        X, its, rrs, _ = meth(B)
        for i in range(2):
            res = ometh(B[i])
            xo, ito = res[0], res[1]
            assert abs(int(its[i]) - ito) <= 1, (meth.__name__, its[i], ito)
            assert rrs[i] < 1e-10 and rel(X[i], xo) < 1e-9


@pytest.mark.parametrize("form", ["DdagD", "DDdag"])
@pytest.mark.parametrize("prec", [amgs.PREC_FP64, amgs.PREC_MIXED])
def test_force(prec, form):
    """Adjoint fermion force on the GPU against the oracle's FP64 CPU gradient,
    and against central differences (h = 1e-5) of the GPU action on 16
    random links.  The oracle evaluates phi^+ (D D^+)^-1 phi; the D^+ D form
    equals it at sigma_3 phi."""
    nx, nt, rows, m0, beta = CASES["32_3lvl"]
    U, _ = amgs.gauge_metropolis(nx, nt, beta, 30, 1234)
    # tight solves so that the finite differences resolve the force
    s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, precision=prec, max_rhs=2, out_tol=1e-13), U)
    s.setup()
    o = O.Oracle(O.default_params(nx, nt, rows, m0=m0, threads=8), U)
    o.setup(draw_order=1)
    phi = np.stack([cvec(s.size(0), 120 + i) for i in range(2)])
    s3 = np.tile([1.0, -1.0], nx * nt)
    S, F = s.force(phi, form)
    for i in range(2):
        So, Fo = o.force(phi[i] if form == "DDdag" else s3 * phi[i])
        assert abs(S[i] - So) < 1e-9 * So
        assert np.abs(F[i] - Fo).max() < 1e-8 * np.abs(Fo).max()
    th = np.angle(U)
    rng = np.random.default_rng(121)
    h = 1e-5
    for _ in range(16):
        n, mu = int(rng.integers(nx * nt)), int(rng.integers(2))
        Sp = []
        for sg in (1, -1):
            t = th.copy()
            t[n, mu] += sg * h
            sp = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, precision=prec, out_tol=1e-13),
                             np.exp(1j * t))
            sp.setup()
            Sp.append(sp.force(phi[0], form)[0])
            sp.close()
        fd = (Sp[0] - Sp[1]) / (2 * h)
        assert abs(fd - F[0][n, mu]) < 1e-4 * np.abs(F[0]).max() + 1e-6 * abs(fd), (n, mu, fd, F[0][n, mu])


@pytest.mark.parametrize("prec", [amgs.PREC_FP64, amgs.PREC_MIXED16])
@pytest.mark.parametrize("case", ["32_3lvl", "32x16_2lvl"])
def test_restart_lengths_at_compiled_limits(case, prec):
    """SAP GMRES(8, 2), K-cycle FGMRES(8, 1), coarsest GMRES(24, 5) and outer
    FGMRES(32, 10): the largest restart lengths the kernels are compiled for
    (every SAP Arnoldi step J = 0..7, nine FP16 basis slots), against the
    oracle with the same parameters."""
    common = dict(sap_m=8, sap_restarts=2, kc_m=8, kc_restarts=1, crs_m=24, crs_restarts=5, out_m=32,
                  out_restarts=10)
    s, o, rows = make(case, common=common, precision=prec, max_rhs=2)
    for l in range(len(rows)):
        v = cvec(s.size(l), 35 + l)
        x0 = cvec(s.size(l), 45 + l)
        assert rel(s.sap(l, v, x0, 2), o.sap(l, v, x0, 2)[0]) < 1e-9, l
    B = np.stack([cvec(s.size(0), 95 + i) for i in range(2)])
    X, its, rrs, hists = s.solve(B)
    for i in range(2):
        xo, ito, histo, _ = o.solve(B[i])
        assert rrs[i] < 1e-10
        if prec == amgs.PREC_FP64:
            assert abs(its[i] - ito) <= 1, (its[i], ito)
            assert rel(X[i], xo) < 1e-9
            assert_history(hists[i], histo, B[i])
        else:
            assert its[i] <= ito + 3, (its[i], ito)
            assert rel(X[i], xo) < 1e-8


@pytest.mark.parametrize("smooth", [dict(nu1=1, nu2=3, sap_tol=0.3), dict(nu1=2, nu2=2, sap_tol=0.05)])
@pytest.mark.parametrize("case", ["32_3lvl", "32_4lvl_odd"])
def test_smoothing_parameters(case, smooth):
    """Pre-smoothing (nu1 > 0), three post-smoothing sweeps and a loose SAP
    tolerance, so that sweeps end early on the tolerance test after a black
    colour (src/sap.cpp:64-174): SAP of every level over three sweeps, one
    K-cycle and the full solve against the oracle; the FP16-basis solve
    converges to the FP64 tolerance."""
    s, o, rows = make(case, common=dict(smooth))
    for l in range(len(rows)):
        v = cvec(s.size(l), 98 + l)
        x0 = cvec(s.size(l), 108 + l)
        assert rel(s.sap(l, v, x0, 3), o.sap(l, v, x0, 3)[0]) < 1e-9, l
    eta = cvec(s.size(0), 118)
    assert rel(s.kcycle(eta), o.kcycle0(eta)) < 1e-9
    b = cvec(s.size(0), 97)
    x, it, rr, _ = s.solve(b)
    xo, ito, _, _ = o.solve(b)
    assert abs(it - ito) <= 1, (it, ito)
    assert rr < 1e-10
    assert rel(x, xo) < 1e-9
    s16, _, _ = make(case, setup=True, common=dict(smooth), precision=amgs.PREC_MIXED16)
    x16, it16, rr16, _ = s16.solve(b)
    assert rr16 < 1e-10 and it16 <= ito + 3, (it16, ito)
    assert rel(x16, xo) < 1e-8


def test_solve_multi_uneven_configurations():
    """Concurrent solves of independent configurations of different lattice
    sizes and shapes, blockings, masses and couplings (amgs_solve_multi): each equals its
    own solve on the same context, bit for bit, and converges."""
    import torch

    confs = [(32, 32, [(0, 4, 4, 4, 4, 4), (1, 2, 2, 4, 2, 2)], -0.10, 2.0, amgs.PREC_FP64, 1),
             (64, 64, [(0, 4, 4, 10, 4, 4)], -0.15, 3.0, amgs.PREC_MIXED, 2),
             (48, 48, [(0, 4, 4, 6, 4, 4), (1, 2, 2, 6, 2, 2)], -0.05, 1.0, amgs.PREC_MIXED16, 1),
             (16, 32, [(0, 4, 8, 4, 2, 4), (1, 2, 2, 2, 2, 2)], 0.0, 2.0, amgs.PREC_MIXED16, 2)]
    solvers, bs, xs, nr = [], [], [], []
    for i, (nx, nt, rows, m0, beta, prec, k) in enumerate(confs):
        U, _ = amgs.gauge_metropolis(nx, nt, beta, 30, 300 + i)
        s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, precision=prec, max_rhs=k), U)
        s.setup()
        solvers.append(s)
        b = np.stack([cvec(s.size(0), 310 + 10 * i + r) for r in range(k)])
        bs.append(torch.from_numpy(b.view(np.float64).copy()).to("cuda:0"))
        xs.append(torch.zeros_like(bs[-1]))
        nr.append(k)
    torch.cuda.synchronize()
    st, its, rrs = amgs.solve_multi(solvers, [b.data_ptr() for b in bs], [x.data_ptr() for x in xs], nr)
    assert list(st) == [amgs.AMGS_OK] * len(confs)
    for i, s in enumerate(solvers):
        nx, nt, rows, m0, beta, prec, k = confs[i]
        x_multi = xs[i].cpu().numpy().view(np.complex128).reshape(nr[i], -1)
        b_host = bs[i].cpu().numpy().view(np.complex128).reshape(nr[i], -1)
        x1, it1, rr1, _ = s.solve(b_host)
        assert np.array_equal(np.asarray(it1).reshape(-1), its[i])
        assert np.array_equal(x1.reshape(nr[i], -1), x_multi)
        assert np.all(rrs[i] < 1e-10)
        # every member against the oracle on its own configuration
        U, _ = amgs.gauge_metropolis(nx, nt, beta, 30, 300 + i)
        o = O.Oracle(O.default_params(nx, nt, rows, m0=m0, threads=8), U)
        o.setup(draw_order=1)
        for r in range(k):
            xo, ito, _, _ = o.solve(b_host[r])
            if prec == amgs.PREC_FP64:
                assert abs(int(its[i][r]) - ito) <= 1 and rel(x_multi[r], xo) < 1e-9, (i, r, its[i][r], ito)
            else:
                assert int(its[i][r]) <= ito + 3 and rel(x_multi[r], xo) < 1e-8, (i, r, its[i][r], ito)
        s.close()


def test_solve_multi_staggered_starts():
    """Sub-batches of one workload as concurrent contexts with staggered
    starts (amgs_solve_multi_staggered, graph replay, FP16 SAP basis): the
    same iterations and solutions, bit for bit, as simultaneous starts, for
    several offsets; offsets outside 0..1000 ms are rejected."""
    import torch

    nx = nt = 64
    rows = [(0, 4, 4, 10, 4, 4), (1, 2, 2, 10, 2, 2)]
    U, _ = amgs.gauge_metropolis(nx, nt, 3.0, 30, 77)
    sizes = [3, 2]
    solvers = []
    for k in sizes:
        s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=-0.10, precision=amgs.PREC_MIXED16, max_rhs=k), U)
        s.setup()
        s.set_graph(True)
        solvers.append(s)
    b = np.stack([cvec(nx * nt * 2, 900 + r) for r in range(sum(sizes))])
    bd = torch.from_numpy(b.view(np.float64).copy()).to("cuda:0")
    bp = [bd[0].data_ptr(), bd[sizes[0]].data_ptr()]

    def run(start_ms):
        xd = torch.zeros_like(bd)
        torch.cuda.synchronize()
        st, its, rrs = amgs.solve_multi(solvers, bp, [xd[0].data_ptr(), xd[sizes[0]].data_ptr()], sizes,
                                        start_ms=start_ms)
        torch.cuda.synchronize()
        assert list(st) == [amgs.AMGS_OK] * 2
        return np.concatenate(its), np.concatenate(rrs), xd.cpu().numpy()

    its0, rr0, x0 = run(None)
    assert np.all(rr0 < 1e-10)
    for start in ([0.0, 0.0], [0.0, 2.5], [4.0, 0.0]):
        its1, rr1, x1 = run(start)
        assert np.array_equal(its1, its0), start
        assert np.array_equal(rr1, rr0), start
        assert np.array_equal(x1, x0), start
    for bad in ([0.0, -1.0], [0.0, 2000.0], [float("nan"), 0.0]):
        with pytest.raises(amgs.AmgsError):
            run(bad)
    for s in solvers:
        s.close()


@pytest.mark.parametrize("prec", [amgs.PREC_FP64, amgs.PREC_MIXED, amgs.PREC_MIXED16])
@pytest.mark.parametrize("case", ["32_3lvl", "16x32_3lvl"])
def test_graph_replay_matches_eager(case, prec):
    """Whole outer restart cycles replayed from captured HIP graphs (the host
    waits once per cycle, the break test stays on the device) give the eager
    result bit for bit, over many replays, with eager solves in between and
    for two batch sizes."""
    s, o, _ = make(case, precision=prec, max_rhs=2)
    B = np.stack([cvec(s.size(0), 400 + i) for i in range(2)])
    xe1, ite1, rre1, _ = s.solve(B[0])
    xe2, ite2, _, _ = s.solve(B)
    s.set_graph(True)
    for rep in range(2):
        xg1, itg1, rrg1, _ = s.solve(B[0])
        assert itg1 == ite1 and np.array_equal(xg1, xe1), rep
        xg2, itg2, _, _ = s.solve(B)
        assert np.array_equal(itg2, ite2) and np.array_equal(xg2, xe2), rep
        s.set_graph(False)
        xe, ite, _, _ = s.solve(B[0])
        assert np.array_equal(xe, xe1)
        s.set_graph(True)
    assert rre1 < 1e-10


@pytest.mark.parametrize("prec", [amgs.PREC_FP64, amgs.PREC_MIXED, amgs.PREC_MIXED16])
def test_graph_replay_across_restarts(prec):
    """FGMRES(4): several outer restarts, so the captured cycle graph is
    replayed cycle after cycle, and a 2-rhs batch whose members converge at
    different steps (one finished, its kernels skipping it for the rest of a
    cycle and in later cycles, while the other goes on): bit for bit the eager
    solve (which leaves a cycle when both are done), against the oracle.  The
    comparison solvers' FGMRES with an SAP-sweep or no preconditioner, graphed
    in the same context, equal their eager runs too."""
    s, o, _ = make("32_3lvl", common=dict(out_m=4), precision=prec, max_rhs=2)
    # a prolongated coarse vector and its image under D: the oracle needs 11
    # and 9 outer iterations for them (cycles of 4: they stop at different steps)
    pc = s.apply_P(0, cvec(s.size(1), 5))
    B = np.stack([pc, s.apply_D(0, pc)])
    xe, ite, rre, he = s.solve(B)
    assert ite[0] != ite[1] and max(ite) > 8, ite  # restarts happened, members differ
    s.set_graph(True)
    for rep in range(2):
        xg, itg, rrg, hg = s.solve(B)
        assert np.array_equal(itg, ite) and np.array_equal(xg, xe), rep
        for i in range(2):
            assert np.array_equal(hg[i], he[i]), (rep, i)
    for i in range(2):
        xo, ito, _, _ = o.solve(B[i])
        assert rre[i] < 1e-10
        if prec == amgs.PREC_FP64:
            assert abs(int(ite[i]) - ito) <= 1 and rel(xe[i], xo) < 1e-9
        else:
            assert int(ite[i]) <= ito + 3 and rel(xe[i], xo) < 1e-8
    # other outer preconditioners get graphs of their own
    for solver in (s.fgmres_sap, s.gmres):
        s.set_graph(True)
        rg = solver(B[1])
        s.set_graph(False)
        re_ = solver(B[1])
        assert rg[1] == re_[1] and np.array_equal(rg[0], re_[0]), solver
    s.set_graph(True)
    xg, itg, _, _ = s.solve(B)
    assert np.array_equal(itg, ite) and np.array_equal(xg, xe)


@pytest.mark.parametrize("prec", [amgs.PREC_FP64, amgs.PREC_MIXED16])
def test_coarse_member_groups_bitwise(prec):
    """The coarse-level solve of a K-cycle with 1, 2, 4 or 8 workgroups per
    right-hand side (the Schwarz blocks of each level-1 colour split over the
    members, their results exchanged behind a member barrier): the solve is
    the same bit for bit, eager and from a captured graph, and matches the
    oracle.  64x64 with the 512^2 workload's coarse blocking (8x8 level-1
    sites in 16 Schwarz blocks, 8 per colour)."""
    nx = nt = 64
    rows = [(0, 8, 8, 6, 4, 4), (1, 4, 4, 6, 4, 4)]
    U, _ = amgs.gauge_metropolis(nx, nt, 3.0, 30, 1234)
    s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=-0.05, precision=prec, max_rhs=3), U)
    s.setup()
    B = np.stack([cvec(s.size(0), 600 + i) for i in range(3)])
    ref = None
    for g in (1, 2, 4, 8):
        s.set_tuning(3, coarse_grp=g)
        assert s.tuning(3)[1] == g
        for graph in (False, True):
            s.set_graph(graph)
            X, its, rrs, hists = s.solve(B)
            if ref is None:
                ref = (X, its, hists)
                continue
            assert np.array_equal(its, ref[1]) and np.array_equal(X, ref[0]), (g, graph)
            for i in range(3):
                assert np.array_equal(hists[i], ref[2][i]), (g, graph, i)
    o = O.Oracle(O.default_params(nx, nt, rows, m0=-0.05, threads=8), U)
    o.setup(draw_order=1)
    xo, ito, histo, _ = o.solve(B[0])
    if prec == amgs.PREC_FP64:
        assert abs(int(ref[1][0]) - ito) <= 1 and rel(ref[0][0], xo) < 1e-9
        assert_history(ref[2][0], histo, B[0])
    else:
        assert int(ref[1][0]) <= ito + 3 and rel(ref[0][0], xo) < 1e-8
    # the autotuner picks one of the candidates; the default fills the chip
    s.set_tuning(3, coarse_grp=0)
    assert s.tuning(3)[1] == 8
    s.autotune(3)
    assert s.tuning(3)[1] in (1, 2, 4, 8)
    s.close()


@pytest.mark.parametrize("case", ["32_3lvl", "32_4lvl_odd", "16x32_3lvl", "30_2lvl_odd_ext"])
def test_batched_coarse_operator_mfma(case):
    """The coarse operator on a block of right-hand sides as MFMA block-GEMMs:
    FP64 against the oracle's Level::D_operator at the operator tolerance for
    batches of 1, 5, 16 and 33 vectors (partial and several 16-rhs tiles,
    odd dof 4 and 6, non-square levels), independent of the autotuned launch
    shape bit for bit; the FP32 form on the FP32 hierarchy within FP32
    rounding."""
    s, o, rows = make(case, setup=False, precision=amgs.PREC_MIXED)
    L = len(rows) + 1
    tvs = [np.concatenate([cvec(o.size(l), 100 * l + i) for i in range(rows[l][3])]) for l in range(L - 1)]
    s.setup_from_tv(tvs, smooth=False)
    o.setup_from_tv(tvs, smooth=False)
    for l in range(1, L):
        for nrhs in (1, 5, 16, 33):
            X = np.stack([cvec(s.size(l), 700 + 10 * l + i) for i in range(nrhs)])
            Y = s.apply_D_batched(l, X)
            for i in range(nrhs):
                assert rel(Y[i], o.dop(l, X[i])) < 1e-13, (l, nrhs, i)
            w, us = s.autotune_batched_D(l, nrhs)
            assert w in (1, 2, 4, 8) and us > 0
            assert np.array_equal(s.apply_D_batched(l, X), Y), (l, nrhs)
            Y32 = s.apply_D_batched(l, X, precision=amgs.PREC_MIXED)
            for i in range(nrhs):
                assert rel(Y32[i], Y[i]) < 5e-6, (l, nrhs, i)
    s.close()


def test_error_paths():
    """Status codes in place of the reference's exit(1) / silent returns:
    NaN or Inf in the gauge field or the right-hand side -> AMGS_ERR_NAN;
    a solve that does not reach the tolerance -> AMGS_NOT_CONVERGED with the
    reference's iteration count restarts*m (src/fgmres.cpp:84-87) and x
    still written; a zero right-hand side -> x = 0 after 0 iterations, also
    inside a batch, without disturbing the other member."""

    L = amgs.lib()
    s, o, _ = make("32_3lvl", common=dict(out_m=4, out_restarts=2, out_tol=1e-14), max_rhs=2)
    n = s.size(0)
    bad = np.ones((n // 2, 2), np.complex128)
    bad[7, 1] = np.nan
    assert L.amgs_set_gauge(s.h, amgs._p(bad.reshape(-1))) == -6
    assert b"NaN" in L.amgs_last_error(s.h)
    b = cvec(n, 130)
    bnan = b.copy()
    bnan[5] = np.inf
    x = np.zeros(n, np.complex128)
    it = np.zeros(1, np.int32)
    rr = np.zeros(1)
    rc = L.amgs_solve(s.h, 1, amgs._p(bnan), amgs._p(x), it.ctypes.data_as(amgs._I), amgs._p(rr), None, 0, None)
    assert rc == -6
    # not converged: 2 restarts of FGMRES(4) cannot reach 1e-14
    x, it, rr, hist = s.solve(b)
    assert it == 2 * 4 and rr > 1e-14 and len(hist) == 8
    xo, ito, histo, _ = o.solve(b)
    assert ito == 8 and rel(x, xo) < 1e-9
    assert_history(hist, histo, b, tol=1e-14)
    its1 = np.zeros(1, np.int32)
    rr1 = np.zeros(1)
    rc = L.amgs_solve(s.h, 1, amgs._p(b), amgs._p(x), its1.ctypes.data_as(amgs._I), amgs._p(rr1), None, 0, None)
    assert rc == amgs.AMGS_NOT_CONVERGED and its1[0] == 8
    s.close()
    # zero right-hand side
    s, o, _ = make("32_3lvl", max_rhs=2)
    z = np.zeros(n, np.complex128)
    x, it, rr, hist = s.solve(z)
    assert it == 0 and rr == 0.0 and len(hist) == 0 and not np.any(x)
    B = np.stack([z, b])
    X, its, rrs, _ = s.solve(B)
    x1, it1, _, _ = s.solve(b)
    assert its[0] == 0 and not np.any(X[0])
    assert its[1] == it1 and np.array_equal(X[1], x1) and rrs[1] < 1e-10
    s.close()


def test_kernel_timing_records_sap_steps():
    """Per-launch timing of SAP steps A and B (amgs_kernel_timing, the event
    pair recorded by each launch's own dispatch): launches, device time and
    algorithmic bytes are recorded, the bytes per launch of step A lie within
    what its sites can move (FP16 basis: at most (16 + 8 J) bytes per block
    site and rhs for J <= m - 1, plus the colour's links), and the timed solve
    gives the untimed solve's result bit for bit."""
    s, o, _ = make("32_3lvl", precision=amgs.PREC_MIXED16, max_rhs=2)
    B = np.stack([cvec(s.size(0), 500 + i) for i in range(2)])
    x0, it0, _, _ = s.solve(B)
    s.kernel_timing(0, True)
    s.kernel_timing(1, True)
    x1, it1, rr1, _ = s.solve(B)
    na, msa, bya = s.kernel_stats(0)
    nb, msb, byb = s.kernel_stats(1)
    s.kernel_timing(0, False)
    s.kernel_timing(1, False)
    assert np.array_equal(it1, it0) and np.array_equal(x1, x0)
    assert na > 0 and nb > 0 and msa > 0 and msb > 0 and bya > 0 and byb > 0
    half = s.size(0) // 2 // 2          # sites of one colour (size(0) counts 2 spins per site)
    m = 5                               # SAP restart length of make()'s parameters
    bound = (16 + 8 * (m - 1)) * half * 2 + 16 * half
    assert bya / na <= bound, (bya / na, bound)
    assert rr1.max() < 1e-10


def test_full_size_256_three_levels():
    """The 256x256, 3-level workload (aggregation 8x8 then 4x4 blocks, 10 test
    vectors, 4x4 SAP blocks; beta = 2 Metropolis links from 500 sweeps,
    m0 = -0.18, one Gaussian right-hand side) at full size: the plain FP64 GPU
    solve against the oracle (iterations +-1, solution within 1e-9), and both
    mixed-precision forms to the FP64 tolerance within 1e-8 of it."""
    n = 256
    rows = [(0, 8, 8, 10, 4, 4), (1, 4, 4, 10, 4, 4)]
    U, _ = amgs.gauge_metropolis(n, n, 2.0, 500, 20260511)
    b = amgs.gaussian(n * n * 2, 2001)
    s = amgs.Solver(amgs.default_params(n, n, rows, m0=-0.18), U)
    s.setup()
    x, it, rr, _ = s.solve(b)
    o = O.Oracle(O.default_params(n, n, rows, m0=-0.18, threads=16), U)
    o.setup(draw_order=1)
    xo, ito, _, rro = o.solve(b)
    print(f"256^2 3-level: GPU FP64 {it} iterations relres {rr:.2e}; oracle {ito} iterations relres {rro:.2e}")
    assert abs(it - ito) <= 1, (it, ito)
    assert rr < 1e-10 and rel(x, xo) < 1e-9
    for prec in (amgs.PREC_MIXED, amgs.PREC_MIXED16):
        sm = amgs.Solver(amgs.default_params(n, n, rows, m0=-0.18, precision=prec), U)
        sm.setup()
        xm, itm, rrm, _ = sm.solve(b)
        print(f"  precision {prec}: {itm} iterations, relres {rrm:.2e}, error vs oracle {rel(xm, xo):.2e}")
        assert rrm < 1e-10 and rel(xm, xo) < 1e-8 and itm <= ito + 3
        sm.close()


BLOCK_CASES = {
    # 32x32-site Schwarz blocks (one pair of sites per thread), and 64x64 (two)
    "64_32blk": (64, 64, [(0, 4, 4, 6, 2, 2)], -0.10, 2.0),
    "128_64blk_3lvl": (128, 128, [(0, 8, 8, 6, 2, 2), (1, 2, 2, 6, 2, 2)], -0.10, 2.0),
}


@pytest.mark.parametrize("case", list(BLOCK_CASES))
@pytest.mark.parametrize("prec", [amgs.PREC_FP64, amgs.PREC_MIXED, amgs.PREC_MIXED16])
def test_sap_block_form_matches_chunked_form(case, prec):
    """The level-0 SAP with one workgroup per (Schwarz block, rhs) item -- a
    colour's whole block GMRES in one launch, the new Krylov vector in an LDS
    tile, W in registers -- against the chunked launches and the oracle: the
    FP64 smoother within 1e-12 of the chunked form and 1e-9 of the oracle;
    full solves in each precision with both forms against the oracle."""
    nx, nt, rows, m0, beta = BLOCK_CASES[case]
    U, _ = amgs.gauge_metropolis(nx, nt, beta, 30, 77)
    s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, precision=prec, max_rhs=2), U)
    s.setup()
    o = O.Oracle(O.default_params(nx, nt, rows, m0=m0, threads=8), U)
    o.setup(draw_order=1)
    B = np.stack([cvec(s.size(0), 700 + i) for i in range(2)])
    v, x0 = cvec(s.size(0), 710), cvec(s.size(0), 711)
    xo, _ = o.sap(0, v, x0, 2)
    res = {}
    for wgs in (1, 2048):  # 1: the block form
        for n in (1, 2):
            s.set_tuning(n, sap_wgs=wgs)
        s.kernel_timing(3, True)
        xs = s.sap(0, v, x0, 2)
        X, its, rrs, _ = s.solve(B)
        launches = s.kernel_stats(3)[0]
        s.kernel_timing(3, False)
        assert (launches > 0) == (wgs == 1), (wgs, launches)
        assert rel(xs, xo) < 1e-9, wgs
        res[wgs] = (xs, X, its)
        for i in range(2):
            xoi, ito, _, _ = o.solve(B[i])
            assert rrs[i] < 1e-10
            if prec == amgs.PREC_FP64:
                assert abs(int(its[i]) - ito) <= 1 and rel(X[i], xoi) < 1e-9, (wgs, i)
            else:
                assert int(its[i]) <= ito + 3 and rel(X[i], xoi) < 1e-8, (wgs, i)
    assert rel(res[1][0], res[2048][0]) < 1e-12
    if prec == amgs.PREC_FP64:
        assert np.array_equal(res[1][2], res[2048][2]) and rel(res[1][1], res[2048][1]) < 1e-11
    s.close()


STREAM_CASES = {
    # block shapes of the streaming level-0 SAP: 16x16 (one run of 2 rows per
    # 8 lanes), 32x32, 64x64 (4 rows per thread) and 128x128 (16 rows per
    # thread, the 512^2 benchmark's blocks)
    "64_16blk": (64, 64, [(0, 4, 4, 6, 4, 4)], -0.10, 2.0),
    "64_32blk": (64, 64, [(0, 4, 4, 6, 2, 2)], -0.10, 2.0),
    "128_64blk_3lvl": (128, 128, [(0, 8, 8, 6, 2, 2), (1, 2, 2, 6, 2, 2)], -0.10, 2.0),
    "256_128blk_3lvl": (256, 256, [(0, 8, 8, 6, 2, 2), (1, 2, 2, 6, 2, 2)], -0.10, 2.0),
}


@pytest.mark.parametrize("case", list(STREAM_CASES))
def test_sap_streaming_form(case):
    """The FP16-basis level-0 SAP streaming over block rows (tuning 2: one
    workgroup per (block, rhs) item, rows in registers, t-neighbours by lane
    shuffles, the last norm of a cycle from the Gram data) against the
    block-tile form (1), the chunked launches (2048) and the FP64 oracle.  The
    three FP16 forms differ only in rounding (order of the inner products, the
    last norm's evaluation), so their smoother outputs agree with each other
    and with the oracle to the FP16 level (1e-3; measured 1e-5 .. 2e-4, a
    wrong neighbour shows as 1e-2 or more).  Restart lengths 1, 2, 5 and 1, 2,
    5 cycles; then full mixed-precision solves with the streaming form against
    the oracle."""
    nx, nt, rows, m0, beta = STREAM_CASES[case]
    U, _ = amgs.gauge_metropolis(nx, nt, beta, 30, 77)
    o = O.Oracle(O.default_params(nx, nt, rows, m0=m0, threads=8), U)
    o.setup(draw_order=1)
    for m, restarts in ((1, 1), (2, 2), (5, 1), (5, 5)):
        s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, precision=amgs.PREC_MIXED16, max_rhs=2,
                                            sap_m=m, sap_restarts=restarts), U)
        s.setup()
        V = np.stack([cvec(s.size(0), 720 + i) for i in range(2)])
        X0 = np.stack([cvec(s.size(0), 730 + i) for i in range(2)])
        out = {}
        for wgs in (2, 1, 2048):
            s.set_tuning(2, sap_wgs=wgs)
            s.kernel_timing(3, True)
            out[wgs] = s.sap_inner(V, X0, 2, nrhs=2).reshape(2, -1)
            launches = s.kernel_stats(3)[0]
            s.kernel_timing(3, False)
            assert (launches > 0) == (wgs in (1, 2)), (wgs, launches)
        po = O.default_params(nx, nt, rows, m0=m0, threads=8, sap_m=m, sap_restarts=restarts)
        oo = O.Oracle(po, U)
        for i in range(2):
            xo, _ = oo.sap(0, V[i], X0[i], 2)
            for wgs in (2, 1, 2048):
                assert rel(out[wgs][i], xo) < 1e-3, (m, restarts, wgs, i)
            assert rel(out[2][i], out[1][i]) < 1e-3 and rel(out[2][i], out[2048][i]) < 1e-3, (m, restarts, i)
        s.close()
    s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, precision=amgs.PREC_MIXED16, max_rhs=2), U)
    s.setup()
    for n in (1, 2):
        s.set_tuning(n, sap_wgs=2)
    B = np.stack([cvec(s.size(0), 740 + i) for i in range(2)])
    s.kernel_timing(3, True)
    X, its, rrs, _ = s.solve(B)
    assert s.kernel_stats(3)[0] > 0
    s.kernel_timing(3, False)
    for i in range(2):
        xo, ito, _, _ = o.solve(B[i])
        assert rrs[i] < 1e-10
        assert int(its[i]) <= ito + 3 and rel(X[i], xo) < 1e-8, (i, its[i], ito)
    s.close()


def test_autotune_keeps_results_deterministic():
    """The SAP launch-shape autotuner picks one of its candidate workgroup
    targets; afterwards solves converge as before and repeat bit for bit."""
    s, o, _ = make("64_2lvl", precision=amgs.PREC_MIXED16, max_rhs=2)
    B = np.stack([cvec(s.size(0), 500 + i) for i in range(2)])
    x0, it0, rr0, _ = s.solve(B)
    best, ms = s.autotune(2)
    # 1: one workgroup per (block, rhs) item (its own kernel where the geometry allows)
    # 2: one workgroup per item streaming over block rows (k_sap_run)
    assert best in (1, 2, 512, 1024, 1536, 2048, 2560, 3072, 4096) and ms > 0
    x1, it1, rr1, _ = s.solve(B)
    x2, it2, rr2, _ = s.solve(B)
    assert np.array_equal(x1, x2) and np.array_equal(it1, it2)
    assert np.all(rr1 < 1e-10) and np.all(np.abs(np.asarray(it1) - np.asarray(it0)) <= 1)
    assert rel(x1, x0) < 1e-8
