"""Produce the golden fixtures from the compiled reference.

Run in the development container only (needs /root/reference, MPICH from
/opt/conda and ``make -C oracle ref``).  The reference's compile-time lattice is
its shipped include/config.h (256x256, 4 levels), so every case is 256^2 with
four levels; parameters.dat rows, m0 and solver limits vary per case.

For every reference output the fixture keeps its SHA-256, its 2-norm and 64
sampled entries (all entries when it has at most 512).  Inputs are rebuilt
from seeds by tests/golden/inputs.py; their SHA-256 is stored too.

    python tests/golden/make_golden.py [case ...]
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import inputs  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
N = 256

ROWS_A = [(0, 8, 8, 4, 4, 4), (1, 4, 4, 4, 4, 4), (2, 2, 2, 4, 2, 2)]
# odd geometry: SAP blocks spanning all of t at level 0, one aggregation block
# in x at level 2 (intra-block wrap), single Schwarz block at level 2
ROWS_B = [(0, 4, 2, 3, 2, 1), (1, 2, 2, 2, 1, 2), (2, 1, 2, 2, 1, 1)]

CASES = {
    "setup_a": dict(mode="setup", rows=ROWS_A, m0=-0.18, tv=4, field=(11, 0.5), vec=12),
    "setup_b": dict(mode="setup", rows=ROWS_B, m0=-0.10, tv=2, field=(21, 0.4), vec=22),
    "solve_a": dict(mode="solve", rows=ROWS_A, m0=0.05, tv=4, field=(11, 0.5), vec=12, restarts=50),
    "solve_c": dict(mode="solve", rows=ROWS_A, m0=-0.18, tv=4, field=(11, 0.5), vec=12, restarts=1),
    # bootstrap set-up (Nit > 0) with the K-cycle and with the V-cycle, and
    # V-cycle / bootstrapped K-cycle solves
    "setup_boot": dict(mode="setup", rows=ROWS_A, m0=-0.18, tv=4, field=(11, 0.5), vec=12, cycle=1, nit=1),
    "setup_vboot": dict(mode="setup", rows=ROWS_A, m0=-0.18, tv=4, field=(11, 0.5), vec=12, cycle=0, nit=1),
    "solve_v": dict(mode="solve", rows=ROWS_A, m0=0.05, tv=4, field=(11, 0.5), vec=12, restarts=50, cycle=0),
    "solve_boot": dict(mode="solve", rows=ROWS_A, m0=0.05, tv=4, field=(11, 0.5), vec=12, restarts=50, cycle=1,
                       nit=2),
    # stand-alone multigrid from a non-zero guess, a fixed number of cycles
    "mg_k": dict(mode="multigrid", rows=ROWS_A, m0=-0.05, tv=4, field=(11, 0.5), vec=12, cycles=3, cycle=1),
    "mg_v": dict(mode="multigrid", rows=ROWS_A, m0=-0.05, tv=4, field=(11, 0.5), vec=12, cycles=3, cycle=0),
    # comparison solvers on the fine lattice (p1, p2: max_iter or restart length, restarts)
    "kry_bicg": dict(mode="bicg", rows=ROWS_A, m0=0.05, tv=4, field=(11, 0.5), vec=12, p=(10000, 0)),
    "kry_cg": dict(mode="cg", rows=ROWS_A, m0=0.3, tv=4, field=(11, 0.5), vec=12, p=(100000, 0)),
    "kry_fsap": dict(mode="fgmres_sap", rows=ROWS_A, m0=0.05, tv=4, field=(11, 0.5), vec=12, p=(20, 50)),
    "kry_gmres": dict(mode="gmres", rows=ROWS_A, m0=0.3, tv=4, field=(11, 0.5), vec=12, p=(20, 100)),
}

NSAMP = 64


def pack(name, arr, out, rng):
    out[name + "__sha"] = np.array(inputs.sha(arr))
    out[name + "__norm"] = np.array(np.linalg.norm(arr))
    if arr.size <= 512:
        idx = np.arange(arr.size)
    else:
        idx = np.sort(rng.choice(arr.size, NSAMP, replace=False))
    out[name + "__idx"] = idx.astype(np.int64)
    out[name + "__val"] = arr[idx]


def run_case(name, cfg):
    rows = cfg["rows"]
    U = inputs.gauge_field(N, N, *cfg["field"])
    vecs = inputs.case_vectors(N, N, rows, cfg["vec"])
    out = {}
    meta = dict(cfg, rows=[list(r) for r in rows], U_sha=inputs.sha(U),
                vec_sha={k: inputs.sha(v) for k, v in vecs.items()})
    rng = np.random.default_rng(1000)
    with tempfile.TemporaryDirectory() as td:
        U.tofile(os.path.join(td, "U.bin"))
        vd = os.path.join(td, "vec")
        od = os.path.join(td, "out")
        os.makedirs(vd)
        os.makedirs(od)
        for k, v in vecs.items():
            v.tofile(os.path.join(vd, k + ".bin"))
        with open(os.path.join(td, "params.dat"), "w") as f:
            f.write("\n".join(" ".join(str(x) for x in r) for r in rows) + "\n")
        if cfg["mode"] == "setup":
            cmd = [HARNESS, "setup", "params.dat", repr(cfg["m0"]), str(cfg["tv"]), "U.bin", "vec", "out"]
        elif cfg["mode"] in ("bicg", "cg", "fgmres_sap", "gmres"):
            cmd = [HARNESS, cfg["mode"], "params.dat", repr(cfg["m0"]), str(cfg["tv"]), "U.bin", "vec/b.bin", "out",
                   str(cfg["p"][0]), str(cfg["p"][1])]
        elif cfg["mode"] == "multigrid":
            cmd = [HARNESS, "multigrid", "params.dat", repr(cfg["m0"]), str(cfg["tv"]), "U.bin", "vec/b.bin",
                   "vec/x0.bin", "out", str(cfg["cycles"]), "1e-14"]
        else:
            cmd = [HARNESS, "solve", "params.dat", repr(cfg["m0"]), str(cfg["tv"]), "U.bin",
                   "vec/b.bin", "out", str(cfg["restarts"])]
        env = dict(os.environ, AMGS_REF_CYCLE=str(cfg.get("cycle", 1)), AMGS_REF_NIT=str(cfg.get("nit", 0)))
        subprocess.run(cmd, cwd=td, check=True, stdout=subprocess.DEVNULL, env=env)
        for fn in sorted(os.listdir(od)):
            path = os.path.join(od, fn)
            if fn.endswith(".bin"):
                pack(fn[:-4], np.fromfile(path, np.complex128), out, rng)
            elif fn.endswith(".res"):
                out["residuals"] = np.loadtxt(path, ndmin=2)[:, 1]
            elif fn == "result.txt":
                kv = dict(line.split() for line in open(path))
                out["iterations"] = np.array(int(kv["iterations"]))
                meta["ref_setup_s"] = float(kv["setup_s"])
                meta["ref_solve_s"] = float(kv["solve_s"])
    out["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "done", {k: meta[k] for k in meta if k.startswith("ref_")})


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    for name in sys.argv[1:] or list(CASES):
        run_case(name, CASES[name])
