"""amgs_sap_dev_noalloc (the coarse-level SAP with its workspace in a static
device array, no allocation) against amgs_sap (the context's workspace) on
the same inputs, through the C ABI: two calls in a row, the second with a
smaller batch, and a batch beyond the static workspace rejected."""
import ctypes

import numpy as np
import pytest

import amg_schwingermodel_amd as amgs

pytestmark = pytest.mark.gpu


def cvec(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def noalloc_sap(s, level, V, X0, nu):
    import torch

    nrhs = V.shape[0]
    vd = torch.from_numpy(np.ascontiguousarray(V).view(np.float64).copy()).to("cuda:0")
    xd = torch.from_numpy(np.ascontiguousarray(X0).view(np.float64).copy()).to("cuda:0")
    torch.cuda.synchronize()
    rc = amgs.lib().amgs_sap_dev_noalloc(s.h, level, nrhs, ctypes.c_void_p(vd.data_ptr()),
                                         ctypes.c_void_p(xd.data_ptr()), nu)
    return rc, xd.cpu().numpy().view(np.complex128).reshape(nrhs, -1)


@pytest.mark.parametrize("prec", [amgs.PREC_FP64, amgs.PREC_MIXED16])
def test_sap_noalloc_matches_sap(prec):
    nx = nt = 32
    rows = [(0, 4, 4, 4, 4, 4), (1, 2, 2, 4, 2, 2)]
    U, _ = amgs.gauge_metropolis(nx, nt, 2.0, 30, 1234)
    s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=-0.10, precision=prec, max_rhs=3), U)
    s.setup()
    n1 = s.size(1)
    # first call: 3 right-hand sides, 2 sweeps
    V = np.stack([cvec(n1, 10 + i) for i in range(3)])
    X0 = np.stack([cvec(n1, 20 + i) for i in range(3)])
    ref = s.sap(1, V.reshape(-1), X0.reshape(-1), 2, nrhs=3).reshape(3, -1)
    rc, got = noalloc_sap(s, 1, V, X0, 2)
    assert rc == amgs.AMGS_OK
    assert np.array_equal(got, ref)
    # second call right after: a smaller batch (1 rhs, 1 sweep) on new inputs
    V1 = cvec(n1, 30)[None]
    X1 = np.zeros_like(V1)
    ref1 = s.sap(1, V1.reshape(-1), X1.reshape(-1), 1, nrhs=1).reshape(1, -1)
    rc, got1 = noalloc_sap(s, 1, V1, X1, 1)
    assert rc == amgs.AMGS_OK
    assert np.array_equal(got1, ref1)
    # the result is a real smoothing step: the residual falls
    r0 = np.linalg.norm(V1[0] - s.apply_D(1, X1[0]))
    r1 = np.linalg.norm(V1[0] - s.apply_D(1, got1[0]))
    assert r1 < 0.5 * r0
    s.close()


def test_sap_noalloc_rejects_oversized_batch_and_bad_level():
    nx = nt = 32
    rows = [(0, 4, 4, 4, 4, 4), (1, 2, 2, 4, 2, 2)]
    U, _ = amgs.gauge_metropolis(nx, nt, 2.0, 30, 1234)
    s = amgs.Solver(amgs.default_params(nx, nt, rows, m0=-0.10, max_rhs=1), U)
    s.setup()
    import torch

    L = amgs.lib()
    # valid device buffers large enough for every call below, so that a wrong
    # check could never address memory outside them
    nbig = 100000
    buf_v = torch.zeros(nbig * s.size(1) * 2, dtype=torch.float64, device="cuda:0")
    buf_x = torch.zeros_like(buf_v)
    v, x = ctypes.c_void_p(buf_v.data_ptr()), ctypes.c_void_p(buf_x.data_ptr())
    torch.cuda.synchronize()
    # level 0 and the coarsest level have no coarse SAP: rejected before any launch
    assert L.amgs_sap_dev_noalloc(s.h, 0, 1, v, x, 1) == -1
    assert L.amgs_sap_dev_noalloc(s.h, 2, 1, v, x, 1) == -1
    # a batch whose workspace exceeds the static array: rejected with a message
    assert L.amgs_sap_dev_noalloc(s.h, 1, nbig, v, x, 1) == -1
    assert b"static" in L.amgs_last_error(s.h)
    s.close()
