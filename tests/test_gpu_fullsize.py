"""The full-size workloads on the GPU against the CPU oracle.

512x512, 3 levels (parameters.dat rows "0 8 8 10 4 4" / "1 4 4 10 4 4"),
beta = 3 Metropolis links (500 sweeps from a cold start, seeded), Gaussian
right-hand sides: the workload of the timed solve, with the gauge field and
right-hand sides that the benchmark generates.

* m0 = -0.12: the plain FP64 GPU solve against the oracle (iterations +-1,
  x within 1e-9, the whole |g| history), and the mixed-precision batch of 32
  right-hand sides (FP16 SAP basis, graph replay): every member's FP64
  relative residual recomputed on the CPU with the oracle's D, and two
  members within 1e-8 of the oracle's solutions.
* m0 = -0.20 (the named mass): below the critical mass of this ensemble.
  Ten FP64 outer iterations follow the oracle's residual history to within
  the oracle's own sensitivity to rounding (the iteration is chaotic there),
  and the solve stagnates (forty iterations stay far above 1e-10).

1024x1024, 4 levels ("0 8 8 12 4 4" / "1 4 4 12 4 4" / "2 2 2 12 2 2"),
beta = 4: the mixed-precision batch of 16 right-hand sides at M0_LARGE, the
most negative bare mass on a 0.01 grid at which the method converges on this
ensemble (the named m0 = -0.23 lies below the critical mass), with every
member's residual recomputed on the CPU, and the plain FP64 solve's iteration
count against the oracle's.
"""
import os

import numpy as np
import pytest

import amg_schwingermodel_amd as amgs
import oracle as O

pytestmark = pytest.mark.gpu

ROWS = [(0, 8, 8, 10, 4, 4), (1, 4, 4, 10, 4, 4)]
N, BETA, SWEEPS, SEED = 512, 3.0, 500, 20260512
M0, M0_NAMED = -0.12, -0.20
NRHS = 32
ROWS_L = [(0, 8, 8, 12, 4, 4), (1, 4, 4, 12, 4, 4), (2, 2, 2, 12, 2, 2)]
N_L, BETA_L, SEED_L, NRHS_L = 1024, 4.0, 20260513, 16
M0_LARGE = -0.09
THREADS = min(16, os.cpu_count() or 1)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def relres_cpu(o, x, b):
    """FP64 |b - D x| / |b| with the oracle's D (src/dirac_operator.cpp:3-27)."""
    return np.linalg.norm(b - o.dphi(x)) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def c3():
    U, plaq = amgs.gauge_metropolis(N, N, BETA, SWEEPS, SEED)
    B = np.stack([amgs.gaussian(N * N * 2, 1000 + i) for i in range(NRHS)])
    o = O.Oracle(O.default_params(N, N, ROWS, m0=M0, threads=THREADS), U)
    o.setup(draw_order=1)
    cache = {}

    def oracle_solve(i):
        if i not in cache:
            cache[i] = o.solve(B[i])
        return cache[i]

    return dict(U=U, B=B, o=o, solve=oracle_solve)


def test_metric_config_fp64_matches_oracle(c3):
    s = amgs.Solver(amgs.default_params(N, N, ROWS, m0=M0), c3["U"])
    s.setup()
    s.set_graph(True)
    b = c3["B"][0]
    x, it, rr, hist = s.solve(b)
    xo, ito, histo, rro = c3["solve"](0)
    print(f"512^2 m0={M0}: GPU FP64 {it} iterations relres {rr:.2e}; oracle {ito} relres {rro:.2e}; "
          f"x error {rel(x, xo):.2e}")
    assert abs(it - ito) <= 1, (it, ito)
    assert rr < 1e-10 and rel(x, xo) < 1e-9
    k = min(len(hist), len(histo))
    assert abs(len(hist) - len(histo)) <= 1
    np.testing.assert_allclose(hist[:k], histo[:k], rtol=1e-6, atol=1e-13 * np.linalg.norm(b))
    s.close()


def test_metric_config_mixed_batch(c3):
    import torch

    s = amgs.Solver(amgs.default_params(N, N, ROWS, m0=M0, max_rhs=NRHS, precision=amgs.PREC_MIXED16), c3["U"])
    s.setup()
    s.autotune(NRHS)
    s.set_graph(True)
    B = c3["B"]
    bd = torch.from_numpy(B.view(np.float64).copy()).to("cuda:0")
    xd = torch.zeros_like(bd)
    st, its, rrs = amgs.solve_multi([s], [bd.data_ptr()], [xd.data_ptr()], [NRHS])
    assert st[0] == amgs.AMGS_OK
    X = xd.cpu().numpy().view(np.complex128).reshape(NRHS, -1)
    o = c3["o"]
    rr_cpu = np.array([relres_cpu(o, X[i], B[i]) for i in range(NRHS)])
    print(f"512^2 mixed16 x {NRHS}: iterations {its[0].min()}..{its[0].max()}, relres (CPU) max {rr_cpu.max():.2e}")
    assert np.all(rr_cpu < 1e-10), rr_cpu
    np.testing.assert_allclose(rr_cpu, rrs[0], rtol=1e-3)
    for i in (0, 1):
        xo, ito, _, _ = c3["solve"](i)
        assert rel(X[i], xo) < 1e-8, (i, rel(X[i], xo))
        assert int(its[0][i]) <= ito + 3, (i, its[0][i], ito)
    s.close()


def test_named_mass_history_and_stagnation(c3):
    """m0 = -0.20: ten outer FGMRES iterations (one cycle of FGMRES(10)) on the
    GPU in FP64 against the oracle, and the mixed solve over forty
    iterations, which stays far above the tolerance (D has eigenvalues with
    negative real part on this ensemble; the spectrum itself is checked on
    smaller lattices of the same coupling in the CPU tests).

    Below the critical mass the reference's iteration is chaotic: the inner
    solves (the coarsest GMRES runs all its restarts without reaching its
    tolerance) amplify rounding, so that a relative change of 1e-15 in b
    moves the oracle's own |g| history by about 1e-2 (on 64^2 too, where the
    CPU tests check it).  The GPU history is held to that bound: within ten
    times the oracle's deviation from itself under such a perturbation."""
    U, b = c3["U"], c3["B"][0]
    kw = dict(m0=M0_NAMED, out_m=10, out_restarts=1)
    s = amgs.Solver(amgs.default_params(N, N, ROWS, **kw), U)
    s.setup()
    x, it, rr, hist = s.solve(b)
    o = O.Oracle(O.default_params(N, N, ROWS, threads=THREADS, **kw), U)
    o.setup(draw_order=1)
    xo, ito, histo, rro = o.solve(b)
    pert = np.random.default_rng(3).standard_normal(b.size) * (1e-15 * np.linalg.norm(b) / np.sqrt(b.size))
    _, _, histp, rrp = o.solve(b + pert)
    dev_gpu = np.max(np.abs(hist - histo) / histo)
    dev_self = np.max(np.abs(histp - histo) / histo)
    print(f"512^2 m0={M0_NAMED}, 10 iterations: relres GPU {rr:.4e}, oracle {rro:.4e}, oracle(b + 1e-15) "
          f"{rrp:.4e}; |g| history deviation GPU vs oracle {dev_gpu:.2e}, oracle vs itself {dev_self:.2e}")
    assert it == ito == 10 and len(hist) == len(histo) == 10
    assert dev_self > 1e-6  # chaotic: no implementation can match the history to 1e-6
    assert dev_gpu < 10 * dev_self
    assert abs(rr - rro) < 0.05 * rro and rr > 0.05
    s.close()
    sm = amgs.Solver(amgs.default_params(N, N, ROWS, m0=M0_NAMED, out_m=20, out_restarts=2,
                                         precision=amgs.PREC_MIXED16), U)
    sm.setup()
    sm.set_graph(True)
    _, itm, rrm, _ = sm.solve(b)
    print(f"512^2 m0={M0_NAMED}, {itm} mixed iterations: relres {rrm:.3e}")
    assert itm == 40 and rrm > 1e-4
    sm.close()


@pytest.mark.timeout(900)
def test_large_lattice_near_critical():
    U, plaq = amgs.gauge_metropolis(N_L, N_L, BETA_L, SWEEPS, SEED_L)
    n0 = N_L * N_L * 2
    B = np.stack([amgs.gaussian(n0, 5000 + i) for i in range(NRHS_L)])
    s = amgs.Solver(amgs.default_params(N_L, N_L, ROWS_L, m0=M0_LARGE, max_rhs=NRHS_L,
                                        precision=amgs.PREC_MIXED16), U)
    s.setup()
    s.set_graph(True)
    X, its, rrs, _ = s.solve(B)
    o = O.Oracle(O.default_params(N_L, N_L, ROWS_L, m0=M0_LARGE, threads=THREADS), U)
    rr_cpu = np.array([relres_cpu(o, X[i], B[i]) for i in range(NRHS_L)])
    print(f"1024^2 m0={M0_LARGE} mixed16 x {NRHS_L}: iterations {its.min()}..{its.max()}, "
          f"relres (CPU) max {rr_cpu.max():.2e}")
    assert np.all(rr_cpu < 1e-10), rr_cpu
    s.close()
    sp = amgs.Solver(amgs.default_params(N_L, N_L, ROWS_L, m0=M0_LARGE), U)
    sp.setup()
    sp.set_graph(True)
    x, it, rr, _ = sp.solve(B[0])
    o.setup(draw_order=1)
    xo, ito, _, rro = o.solve(B[0])
    print(f"1024^2 m0={M0_LARGE} FP64: GPU {it} iterations relres {rr:.2e}; oracle {ito} relres {rro:.2e}; "
          f"x error {rel(x, xo):.2e}; mixed members {its.min()}..{its.max()}")
    assert abs(it - ito) <= 1 and rr < 1e-10 and rel(x, xo) < 1e-9
    assert int(its[0]) <= ito + 3
    sp.close()
