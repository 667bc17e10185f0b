"""Several gauge configurations in one context (amgs_create_multi): every
kernel launch covers all configurations of one lattice shape, right-hand side
i solved with configuration i's links, mass and multigrid hierarchy.

* Each member of a multi-configuration solve equals the solve of a context
  of its own (FP64: bit for bit; mixed precision: the same tolerance), and
  the oracle's.
* The uneven batch: 48 seeded configurations (beta in [1, 4], m0 in
  [-0.22, -0.10], lattices drawn from 128^2, 256^2 and 512^2, 3 levels with
  the 8x8 / 4x4 aggregation), one right-hand side each, as three
  multi-configuration contexts solved in one amgs_solve_multi call.  Members
  above their critical mass converge to 1e-10 (recomputed on the CPU for a
  sample, solutions against the oracle); members below it return
  AMGS_NOT_CONVERGED after the full iteration budget, as the oracle does.
* The fermion force of the 256^2 members against the oracle's FP64 gradient,
  and against central differences (h = 1e-5) of the action on 64 random
  links, whose 128 perturbed gauge fields are solved as one
  multi-configuration batch.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import amg_schwingermodel_amd as amgs
import oracle as O

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def cvec(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


@pytest.mark.parametrize("prec", [amgs.PREC_FP64, amgs.PREC_MIXED16])
def test_members_equal_their_own_contexts(prec):
    nx = nt = 32
    rows = [(0, 4, 4, 4, 4, 4), (1, 2, 2, 4, 2, 2)]
    cfgs = [(2.0, -0.10, 11), (3.0, -0.05, 12), (1.0, 0.02, 13)]
    Us = [amgs.gauge_metropolis(nx, nt, beta, 30, seed)[0] for beta, _, seed in cfgs]
    m0s = [m0 for _, m0, _ in cfgs]
    B = np.stack([cvec(nx * nt * 2, 800 + i) for i in range(3)])
    sm = amgs.Solver(amgs.default_params(nx, nt, rows, precision=prec), np.stack(Us), m0s=m0s)
    assert amgs.lib().amgs_num_configs(sm.h) == 3
    sm.setup()
    X, its, rrs, hists = sm.solve(B)
    Dm = sm.apply_D(0, B, nrhs=3).reshape(3, -1)
    for i, (U, m0) in enumerate(zip(Us, m0s)):
        s1 = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, precision=prec), U)
        s1.setup()
        x1, it1, rr1, h1 = s1.solve(B[i])
        assert np.array_equal(Dm[i], s1.apply_D(0, B[i]))
        o = O.Oracle(O.default_params(nx, nt, rows, m0=m0, threads=8), U)
        o.setup(draw_order=1)
        xo, ito, _, _ = o.solve(B[i])
        assert rrs[i] < 1e-10
        if prec == amgs.PREC_FP64:
            assert its[i] == it1 and np.array_equal(X[i], x1), i
            assert np.array_equal(hists[i], h1), i
            assert abs(int(its[i]) - ito) <= 1 and rel(X[i], xo) < 1e-9, i
        else:
            assert abs(int(its[i]) - it1) <= 1 and rel(X[i], x1) < 1e-8, i
            assert int(its[i]) <= ito + 3 and rel(X[i], xo) < 1e-8, i
        s1.close()
    # the single-hierarchy component entry points refuse a multi-configuration context
    with pytest.raises(amgs.AmgsError):
        sm.apply_P(0, np.zeros(sm.size(1), np.complex128))
    sm.close()
