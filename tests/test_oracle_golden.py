"""The CPU oracle against golden vectors from the compiled reference.

Fixtures hold, per reference output, its SHA-256, norm and sampled entries
(tests/golden/make_golden.py).  The oracle follows the reference's operation
order, so the expected result is bit-for-bit equality; the sampled entries are
also checked to 1e-13 so a failure says how far off it is.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from tests.golden import inputs

GOLD = os.path.join(os.path.dirname(__file__), "golden")
N = 256


def load(name):
    z = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return z, meta


def build(meta, **kw):
    rows = [tuple(r) for r in meta["rows"]]
    U = inputs.gauge_field(N, N, *meta["field"])
    assert inputs.sha(U) == meta["U_sha"], "gauge-field generator drifted"
    vecs = inputs.case_vectors(N, N, rows, meta["vec"])
    for k, v in vecs.items():
        assert inputs.sha(v) == meta["vec_sha"][k], f"input {k} drifted"
    p = O.default_params(N, N, rows, m0=meta["m0"], tv_sap_iters=meta["tv"], threads=8, **kw)
    return O.Oracle(p, U), vecs, rows


def check(z, name, arr):
    arr = np.ascontiguousarray(arr, dtype=np.complex128)
    idx = z[name + "__idx"]
    ref = z[name + "__val"]
    scale = max(float(z[name + "__norm"]) / np.sqrt(arr.size), 1e-300)
    err = np.max(np.abs(arr[idx] - ref)) / scale
    assert err < 1e-13, f"{name}: sampled error {err:.3e}"
    assert abs(np.linalg.norm(arr) - float(z[name + "__norm"])) <= 1e-13 * float(z[name + "__norm"]) + 1e-300
    assert inputs.sha(arr) == str(z[name + "__sha"]), f"{name}: not bitwise equal (sampled error {err:.3e})"


def links_ref_layout(o, l):
    """Oracle links in the reference's index order: G1[x][a][b][c][bb], G2/G3[...][mu]."""
    d = o.dof(l)
    C = d // 2
    ns = o.nsites(l)
    G1 = o.links(l, 0).reshape(ns, C, 2, C, 2).transpose(0, 2, 4, 1, 3)
    G2 = np.stack([o.links(l, 1, mu) for mu in (0, 1)]).reshape(2, ns, C, 2, C, 2).transpose(1, 3, 5, 2, 4, 0)
    G3 = np.stack([o.links(l, 2, mu) for mu in (0, 1)]).reshape(2, ns, C, 2, C, 2).transpose(1, 3, 5, 2, 4, 0)
    return G1.reshape(-1), G2.reshape(-1), G3.reshape(-1)


@pytest.mark.parametrize("case", ["setup_a", "setup_b", "setup_boot", "setup_vboot"])
def test_setup_components(case):
    """setup_boot / setup_vboot: bootstrap set-up (Nit = 1, src/amg.cpp:52-84)
    with the K-cycle and with the V-cycle."""
    z, meta = load(case)
    o, v, rows = build(meta, cycle=meta.get("cycle", 1), nit=meta.get("nit", 0))
    L = len(rows) + 1
    check(z, "Dphi", o.dphi(v["psi"]))
    check(z, "Ddagphi", o.ddag(v["psi"]))
    o.setup(draw_order=1)   # g++ evaluates the imaginary draw of dist(g) + I*dist(g) first
    for l in range(L - 1):
        for i in range(rows[l][3]):
            check(z, f"tv_l{l}_{i}", o.tv(l, i))
    for l in range(1, L):
        G1, G2, G3 = links_ref_layout(o, l)
        check(z, f"G1_l{l}", G1)
        check(z, f"G2_l{l}", G2)
        check(z, f"G3_l{l}", G3)
    for l in range(L):
        check(z, f"Dop_l{l}", o.dop(l, v[f"v{l}"]))
        if l < L - 1:
            check(z, f"P_l{l}", o.prolong(l, v[f"v{l + 1}"]))
            check(z, f"Pt_l{l}", o.restrict(l, v[f"v{l}"]))
            x, _ = o.sap(l, v[f"v{l}"], v[f"x{l}"], 2)
            check(z, f"SAP_l{l}", x)
    check(z, "kcycle", o.kcycle0(v["v0"]))
    if "vcycle__sha" in z.files:
        check(z, "vcycle", o.vcycle(0, v["v0"]))


def test_draw_order_matters():
    """The other draw order must NOT reproduce the reference test vectors."""
    z, meta = load("setup_a")
    o, v, rows = build(meta)
    o.setup(draw_order=0)
    assert inputs.sha(o.tv(0, 0)) != str(z["tv_l0_0__sha"])


@pytest.mark.parametrize("case", ["solve_a", "solve_c", "solve_v", "solve_boot"])
def test_full_solve(case):
    """solve_v: V-cycle preconditioner; solve_boot: K-cycle after Nit = 2."""
    z, meta = load(case)
    o, v, rows = build(meta, out_restarts=meta["restarts"], cycle=meta.get("cycle", 1), nit=meta.get("nit", 0))
    o.setup(draw_order=1)
    x, it, hist, rr = o.solve(v["b"])
    assert it == int(z["iterations"])
    if "residuals" in z.files:  # the reference writes them only on convergence (src/fgmres.cpp:65-79)
        np.testing.assert_array_equal(hist, z["residuals"])
    check(z, "x", x)


@pytest.mark.parametrize("case", ["mg_k", "mg_v"])
def test_standalone_multigrid(case):
    """AlgebraicMG::applyMultilevel (src/amg.cpp:181-231): three K- or
    V-cycles from a non-zero initial guess."""
    z, meta = load(case)
    o, v, rows = build(meta, cycle=meta["cycle"])
    o.setup(draw_order=1)
    x, cyc, _ = o.multigrid(v["b"], v["x0"], it=meta["cycles"], tol=1e-14)
    assert cyc == -meta["cycles"]
    check(z, "x", x)


def test_bicgstab():
    """bi_cgstab (src/bi_cgstab.cpp:5-87): iterations, |s| history and x."""
    z, meta = load("kry_bicg")
    o, v, _ = build(meta)
    x, it, hist = o.bicgstab(v["b"], max_iter=meta["p"][0], tol=1e-10)
    assert it == len(z["residuals"])
    np.testing.assert_array_equal(hist, z["residuals"])
    check(z, "x", x)


def test_cg_normal_equations():
    """conjugate_gradient on D D^dagger (src/conjugate_gradient.cpp:5-65); the
    reference prints the loop counter, one less than the passes counted here."""
    z, meta = load("kry_cg")
    o, v, _ = build(meta)
    x, it = o.cg(v["b"], max_iter=meta["p"][0], tol=1e-10)
    assert it == int(z["iterations"]) + 1
    check(z, "x", x)


def test_gmres_fine():
    """FGMRES_fine_level (include/fgmres.h:126-148): GMRES(20) on D."""
    z, meta = load("kry_gmres")
    o, v, _ = build(meta)
    x, it = o.gmres_fine(v["b"], m=meta["p"][0], restarts=meta["p"][1])
    assert it == int(z["iterations"])
    check(z, "x", x)


def test_fgmres_sap():
    """FGMRES_SAP (include/sap.h:194-225).  The reference's fine-level SAP
    applies D in the D_phi form while the oracle's level-0 SAP uses the stencil
    form (the same operator, other rounding), so x agrees to 1e-10, not bitwise."""
    z, meta = load("kry_fsap")
    o, v, _ = build(meta, out_m=meta["p"][0], out_restarts=meta["p"][1])
    x, it, hist = o.fgmres_sap(v["b"])
    assert it == int(z["iterations"])
    np.testing.assert_allclose(hist, z["residuals"], rtol=1e-6)
    idx = z["x__idx"]
    assert np.max(np.abs(x[idx] - z["x__val"])) < 1e-10 * float(z["x__norm"]) / np.sqrt(x.size) * 100
    assert abs(np.linalg.norm(x) - float(z["x__norm"])) < 1e-10 * float(z["x__norm"])
