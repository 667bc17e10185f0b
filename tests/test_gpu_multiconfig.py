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


GEOMS = {
    # chunked level-0 SAP launches (8x8-site Schwarz blocks)
    "32": (32, [(0, 4, 4, 4, 4, 4), (1, 2, 2, 4, 2, 2)]),
    # one workgroup per (32x32-site block, rhs): per-configuration links and masses inside it
    "64_blockwise": (64, [(0, 4, 4, 4, 2, 2), (1, 2, 2, 4, 2, 2)]),
}


@pytest.mark.parametrize("geom", list(GEOMS))
@pytest.mark.parametrize("prec", [amgs.PREC_FP64, amgs.PREC_MIXED16])
def test_members_equal_their_own_contexts(prec, geom):
    nx = nt = GEOMS[geom][0]
    rows = GEOMS[geom][1]
    cfgs = [(2.0, -0.10, 11), (3.0, -0.05, 12), (1.0, 0.02, 13)]
    Us = [amgs.gauge_metropolis(nx, nt, beta, 30, seed)[0] for beta, _, seed in cfgs]
    m0s = [m0 for _, m0, _ in cfgs]
    B = np.stack([cvec(nx * nt * 2, 800 + i) for i in range(3)])
    sm = amgs.Solver(amgs.default_params(nx, nt, rows, precision=prec), np.stack(Us), m0s=m0s)
    assert amgs.lib().amgs_num_configs(sm.h) == 3
    blockwise = geom == "64_blockwise"
    if blockwise:
        sm.set_tuning(3, sap_wgs=1)
    sm.setup()
    X, its, rrs, hists = sm.solve(B)
    Dm = sm.apply_D(0, B, nrhs=3).reshape(3, -1)
    for i, (U, m0) in enumerate(zip(Us, m0s)):
        s1 = amgs.Solver(amgs.default_params(nx, nt, rows, m0=m0, precision=prec), U)
        if blockwise:
            s1.set_tuning(1, sap_wgs=1)
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


@pytest.fixture(scope="module")
def uneven():
    from amg_schwingermodel_amd import workloads

    return workloads.uneven_batch(log=lambda m: print(m, flush=True))


@pytest.mark.timeout(900)
def test_uneven_batch_in_one_call(uneven):
    """48 configurations of three lattice sizes, one multi-configuration
    context per size, solved by one amgs_solve_multi call (mixed precision,
    FP16 SAP basis, FGMRES(20) with 2 restarts: members below their critical
    mass run every coarse solve to its iteration limit, about a second per
    outer iteration at 512^2, so the budget is kept short)."""
    import torch

    from amg_schwingermodel_amd import workloads

    groups = uneven
    assert sum(len(g["members"]) for g in groups) == 48 and len(groups) == 3
    solvers, bd, xd = [], [], []
    for g in groups:
        p = workloads.uneven_params(g["n"], precision=amgs.PREC_MIXED16, out_restarts=2)
        s = amgs.Solver(p, g["U"], m0s=g["m0"])
        s.setup()
        s.set_graph(True)
        solvers.append(s)
        print(f"{g['n']}^2 group: {len(g['members'])} configurations set up", flush=True)
        bd.append(torch.from_numpy(g["b"].view(np.float64).copy()).to("cuda:0"))
        xd.append(torch.zeros_like(bd[-1]))
    torch.cuda.synchronize()
    st, its, rrs = amgs.solve_multi(solvers, [b.data_ptr() for b in bd], [x.data_ptr() for x in xd],
                                    [len(g["members"]) for g in groups])
    nconv = 0
    for k, g in enumerate(groups):
        X = xd[k].cpu().numpy().view(np.complex128).reshape(len(g["members"]), -1)
        conv = rrs[k] < 1e-10
        nconv += int(conv.sum())
        print(f"{g['n']}^2: {len(g['members'])} members, {int(conv.sum())} converged, iterations "
              f"{its[k].tolist()}, status {st[k]}")
        # every member either converged or used the whole budget (2 x 20)
        assert np.all(conv | (its[k] == 40)), (its[k], rrs[k])
        assert st[k] == (amgs.AMGS_OK if conv.all() else amgs.AMGS_NOT_CONVERGED)
        # sample: the first converged member against the oracle (and its
        # residual recomputed on the CPU), the first stalled one's oracle stalls too
        for want in (True, False):
            sel = np.flatnonzero(conv == want)
            if len(sel) == 0 or (not want and g["n"] != 128):
                continue  # a stalled oracle solve is checked on the smallest lattice only
            # the converged member with the fewest iterations (the oracle's CPU time)
            i = int(sel[np.argmin(its[k][sel])]) if want else int(sel[0])
            o = O.Oracle(O.default_params(g["n"], g["n"], workloads.UNEVEN_ROWS, m0=float(g["m0"][i]),
                                          threads=THREADS, out_restarts=2), g["U"][i])
            o.setup(draw_order=1)
            xo, ito, _, rro = o.solve(g["b"][i])
            print(f"  member {g['members'][i]} (beta {g['beta'][i]:.3f}, m0 {g['m0'][i]:.4f}): GPU {its[k][i]} "
                  f"iterations relres {rrs[k][i]:.2e}; oracle {ito} relres {rro:.2e}")
            if want:
                rcpu = np.linalg.norm(g["b"][i] - o.dphi(X[i])) / np.linalg.norm(g["b"][i])
                assert rcpu < 1e-10 and rel(X[i], xo) < 1e-8 and int(its[k][i]) <= ito + 3
            else:
                assert ito == 40 and rro > 1e-10
    assert 0 < nconv < 48
    for s in solvers:
        s.close()


@pytest.mark.timeout(900)
def test_force_of_the_256_members(uneven):
    """dS/dtheta of S = phi^+ (D^+ D)^-1 phi for the 256^2 members whose
    solves converge, batched in one FP64 multi-configuration context: against
    the oracle's FP64 gradient (two members), and against central differences
    (h = 1e-5) of S on 64 random links of one member, the 128 perturbed
    fields solved as one batch of a multi-configuration context."""
    from amg_schwingermodel_amd import workloads

    g = [g for g in uneven if g["n"] == 256][0]
    n = 256
    kw = dict(out_tol=1e-13)
    # members that converge fast: relative residual below 1e-6 after one
    # FGMRES(20) cycle (the others are near or below their critical mass)
    s = amgs.Solver(workloads.uneven_params(n, out_restarts=1), g["U"], m0s=g["m0"])
    s.setup()
    s3 = np.tile([1.0, -1.0], n * n)
    _, _, rr, _ = s.solve(s3 * g["b"])
    ok = [i for i in range(len(g["members"])) if rr[i] < 1e-6]
    s.close()
    assert len(ok) >= 2, rr
    U, m0s, phi = g["U"][ok], g["m0"][ok], g["b"][ok]
    s = amgs.Solver(workloads.uneven_params(n, **kw), U, m0s=m0s)
    s.setup()
    S, F = s.force(phi, "DdagD")
    s.close()
    for i in range(2):
        o = O.Oracle(O.default_params(n, n, workloads.UNEVEN_ROWS, m0=float(m0s[i]), threads=THREADS, **kw), U[i])
        o.setup(draw_order=1)
        So, Fo = o.force(s3 * phi[i])
        print(f"member {g['members'][ok[i]]}: S {S[i]:.10e} (oracle {So:.10e}), max |F - F_oracle| / max |F| "
              f"{np.abs(F[i] - Fo).max() / np.abs(Fo).max():.2e}")
        assert abs(S[i] - So) < 1e-9 * So
        assert np.abs(F[i] - Fo).max() < 1e-8 * np.abs(Fo).max()
    # central differences on 64 random links of member 0: 128 perturbed fields in one batch
    rng = np.random.default_rng(901)
    links = [(int(rng.integers(n * n)), int(rng.integers(2))) for _ in range(64)]
    h = 1e-5
    th = np.angle(U[0])
    Up = []
    for (site, mu) in links:
        for sg in (1.0, -1.0):
            t = th.copy()
            t[site, mu] += sg * h
            Up.append(np.exp(1j * t))
    sp = amgs.Solver(workloads.uneven_params(n, **kw), np.stack(Up), m0s=[float(m0s[0])] * len(Up))
    sp.setup()
    Xp, _, rrp, _ = sp.solve(np.tile(s3 * phi[0], (len(Up), 1)))
    sp.close()
    assert np.all(rrp < 1e-12)
    Sp = np.sum(np.abs(Xp) ** 2, axis=1)  # S = |D^-1 s3 phi|^2
    fd = (Sp[0::2] - Sp[1::2]) / (2 * h)
    Fl = np.array([F[0][site, mu] for (site, mu) in links])
    err = np.abs(fd - Fl)
    print(f"FD on 64 links: max |fd - F| {err.max():.2e}, max |F| {np.abs(F[0]).max():.2e}")
    assert np.all(err < 1e-4 * np.abs(F[0]).max() + 1e-6 * np.abs(fd))
