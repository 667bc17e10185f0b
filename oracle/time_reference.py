"""Times the reference itself, compiled here from its own sources, on the
benchmark workloads (CPU baseline; this container only, never on a GPU host).

For each workload the recipe in oracle/Makefile builds the reference for that
lattice size and level count (its own configure step: config.h from
config.h.in), the workload's gauge field and right-hand side are written as
raw complex128 files, and the reference's FGMRES_AMG solve runs under
`mpirun -np 8` with one Schwarz block per rank and colour as the reference's
own runs distribute them (script.sh:38, src/variables.cpp:51), and on one rank
with every block.  The reference's own timers give set-up (the FGMRES_AMG
constructor) and solve (fgmres) separately (src/tests.cpp:75-79).

    python oracle/time_reference.py [--out oracle/reference_timings.json] [--cases C1,C2,C3]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

CASES = {
    # name: (n, levels, parameters.dat rows, m0, beta, gauge seed, rhs seed)
    "C1": (64, 2, [(0, 4, 4, 10, 4, 4)], -0.18, 2.0, 20260510, 1000),
    "C2": (256, 3, [(0, 8, 8, 10, 4, 4), (1, 4, 4, 10, 4, 4)], -0.18, 2.0, 20260511, 2001),
    "C3": (512, 3, [(0, 8, 8, 10, 4, 4), (1, 4, 4, 10, 4, 4)], -0.12, 3.0, 20260512, 1000),
}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run_case(name, nproc_list=(8, 1)):
    import amg_schwingermodel_amd as amgs

    n, levels, rows, m0, beta, gseed, bseed = CASES[name]
    exe = os.path.join(HERE, "_ref", f"harness_{n}_{levels}")
    subprocess.run(["make", "-s", "-C", HERE, f"_ref/harness_{n}_{levels}"], check=True)
    U, plaq = amgs.gauge_metropolis(n, n, beta, 500, gseed)
    b = amgs.gaussian(n * n * 2, bseed)
    out = {"lattice": f"{n}x{n}", "levels": levels, "rows": [list(r) for r in rows], "m0": m0, "beta": beta,
           "plaquette": plaq}
    mpirun = os.path.join(os.environ.get("MPI_DIR", "/opt/conda"), "bin", "mpirun")
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "params.dat"), "w") as f:
            f.write("\n".join(" ".join(map(str, r)) for r in rows) + "\n")
        U.astype(np.complex128).tofile(os.path.join(td, "U.bin"))
        b.astype(np.complex128).tofile(os.path.join(td, "b.bin"))
        for npr in nproc_list:
            od = os.path.join(td, f"out{npr}")
            os.makedirs(od, exist_ok=True)
            # blocks per rank: one per rank and colour with 8 ranks, all on one rank
            env = dict(os.environ, AMGS_REF_BPP=str(1 if npr > 1 else 1 << 20))
            cmd = [exe, "solve", os.path.join(td, "params.dat"), repr(m0), "4", os.path.join(td, "U.bin"),
                   os.path.join(td, "b.bin"), od]
            if npr > 1:
                cmd = [mpirun, "-np", str(npr)] + cmd
            subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, env=env)
            kv = dict(line.split() for line in open(os.path.join(od, "result.txt")))
            out[f"np{npr}"] = {"iterations": int(kv["iterations"]), "setup_s": float(kv["setup_s"]),
                               "solve_s": float(kv["solve_s"])}
            print(f"{name} np={npr}: {out[f'np{npr}']}", flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "reference_timings.json"))
    ap.add_argument("--cases", default="C1,C2,C3")
    a = ap.parse_args()
    res = {"cpu": cpu_model(), "cores": os.cpu_count(), "cases": {}}
    for name in a.cases.split(","):
        res["cases"][name] = run_case(name)
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
