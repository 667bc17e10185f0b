"""Deterministic inputs for the golden fixtures.

Gauge fields and vectors are built only from numpy's PCG64 ``random()``
(uniform doubles, whose stream numpy keeps stable) through a Box-Muller
transform written here, so the same seeds give the same bytes anywhere.  Each
fixture records the SHA-256 of the inputs it was made from; tests check it
before comparing outputs.
"""
import hashlib

import numpy as np


def gaussian(rng, n):
    u1 = rng.random(n)
    u2 = rng.random(n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def cgauss(rng, n):
    return (gaussian(rng, n) + 1j * gaussian(rng, n)).astype(np.complex128)


def gauge_field(nx, nt, seed, width):
    """U(1) links exp(i theta), theta ~ N(0, width^2), shape [nx*nt, 2]."""
    rng = np.random.default_rng(seed)
    th = width * gaussian(rng, nx * nt * 2)
    return np.exp(1j * th).astype(np.complex128).reshape(nx * nt, 2)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def level_shapes(nx, nt, rows):
    """(nsites, dof) per level for parameters.dat rows (block counts)."""
    shapes = [(nx * nt, 2)]
    for (l, bx, bt, nv, sx, st) in rows:
        shapes.append((bx * bt, 2 * nv))
    return shapes


def case_vectors(nx, nt, rows, seed):
    """Random test vectors: psi (fine), v_l and x_l per level, b (fine)."""
    rng = np.random.default_rng(seed)
    sh = level_shapes(nx, nt, rows)
    out = {"psi": cgauss(rng, nx * nt * 2)}
    for l, (ns, d) in enumerate(sh):
        out[f"v{l}"] = cgauss(rng, ns * d)
        out[f"x{l}"] = cgauss(rng, ns * d)
    out["b"] = cgauss(rng, nx * nt * 2)
    return out
