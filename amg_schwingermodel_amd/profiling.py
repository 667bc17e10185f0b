"""HBM traffic of the level-0 SAP kernels, from rocprofv3 hardware counters.

Two passes of the same short workload, one per counter (FETCH_SIZE, then
WRITE_SIZE, both in KiB per dispatch), run from the repository root:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- \\
        python3 -m amg_schwingermodel_amd.profiling sample [run|block|chunked]
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- \\
        python3 -m amg_schwingermodel_amd.profiling sample [run|block|chunked]
    python3 -m amg_schwingermodel_amd.profiling summarize OUT/fetch OUT/write OUT.json [kernel]

The workload first copies a 256 MiB device buffer three times (a plain copy
kernel of known size: 256 MiB read and 256 MiB written per launch), then runs
two outer iterations of the mixed-precision batched solve (512x512, 3 levels,
32 right-hand sides, m0 = -0.12, beta = 3; FP16 SAP Krylov basis, eager
launches), the configuration of the benchmark's roofline pass.  `run`
(default) runs the level-0 SAP as one workgroup per Schwarz block and rhs
streaming over block rows (k_sap_run, one launch per colour), `block` as the
block-tile kernel (k_sap_block), `chunked` as the five-kernel chunked form
(k_sap_init / arnoldi_a / arnoldi_b / update / resid).  The sample prints the
algorithmic bytes per launch that the library counts for the kernel.

`summarize` scales each counter by the ratio of the copy's known bytes to what
the counter reads for the copy, reports every SAP kernel it finds and puts the
named kernel's (default k_sap_run) bytes per launch at the top.
"""
import csv
import glob
import json
import os
import sys

import numpy as np

COPY_BYTES = 256 << 20
ROWS = [(0, 8, 8, 10, 4, 4), (1, 4, 4, 10, 4, 4)]
# kernel name prefixes reported by summarize, with the library's timing class
KERNELS = {"k_sap_run<": 3, "k_sap_block<float": 3, "k_sap_arnoldi_a<float": 0, "k_sap_arnoldi_b<float": 1}
# sample forms: the level-0 SAP tuning value that selects each
FORMS = {"run": 2, "block": 1, "chunked": 1024}


def sample(form="block", nx=512, nrhs=32, m0=-0.12):
    import torch

    import amg_schwingermodel_amd as amgs

    dev = torch.device("cuda:0")
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    U, _ = amgs.gauge_metropolis(nx, nx, 3.0, 500, 20260512)
    B = np.stack([amgs.gaussian(nx * nx * 2, 1000 + i) for i in range(nrhs)])
    s = amgs.Solver(amgs.default_params(nx, nx, ROWS, m0=m0, max_rhs=nrhs, precision=amgs.PREC_MIXED16,
                                        out_m=2, out_restarts=1), U)
    s.setup()
    s.set_tuning(nrhs, sap_wgs=FORMS[form])
    for kc in set(KERNELS.values()):
        s.kernel_timing(kc, True)
    s.solve(B)
    rec = {}
    for kc in sorted(set(KERNELS.values())):
        n, ms, nbytes = s.kernel_stats(kc)
        if n:
            rec[f"class {kc}"] = {"launches": n, "alg_bytes_per_launch": nbytes / n, "avg_us": ms * 1e3 / n}
    print(json.dumps({"form": form, "kernels": rec}), flush=True)
    s.close()


def _per_kernel(d, counter):
    """{kernel name: [values]} of one counter (KiB per dispatch) from a rocprofv3 csv directory."""
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {d}")
    out = {}
    for f in files:
        for r in csv.DictReader(open(f)):
            if r.get("Counter_Name") != counter:
                continue
            out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return out


def summarize(fetch_dir, write_dir, out_json, main="k_sap_run<"):
    res = {"copy": {}}
    counts = {}
    for key, d, counter in (("fetch", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        k = _per_kernel(d, counter)
        copy = [v for name, vals in k.items() if "copy" in name.lower() or "elementwise" in name.lower()
                for v in vals if v * 1024 > 0.25 * COPY_BYTES]
        if not copy:
            raise RuntimeError(f"no copy kernel of {COPY_BYTES} bytes found in {d}")
        scale = COPY_BYTES / (np.median(copy) * 1024)
        res["copy"][key] = {"kib_per_launch": float(np.median(copy)), "scale": float(scale)}
        for prefix in KERNELS:
            vals = [v for name, vs in k.items() if prefix in name for v in vs]
            if vals:
                c = counts.setdefault(prefix, {})
                c["launches"] = len(vals)
                c[f"{key}_bytes_per_launch"] = float(np.mean(vals) * 1024 * scale)
    for prefix, c in counts.items():
        if "fetch_bytes_per_launch" in c and "write_bytes_per_launch" in c:
            c["hbm_bytes_per_launch"] = c["fetch_bytes_per_launch"] + c["write_bytes_per_launch"]
    res["kernels"] = counts
    # the main kernel's bytes at the top (what the benchmark's roofline reads)
    if main in counts and "hbm_bytes_per_launch" in counts[main]:
        res["kernel"] = main
        res["hbm_bytes_per_launch"] = counts[main]["hbm_bytes_per_launch"]
    with open(out_json, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    if sys.argv[1:2] == ["sample"]:
        sample(*(sys.argv[2:3] or ["run"]))
    elif sys.argv[1:2] == ["summarize"] and len(sys.argv) in (5, 6):
        summarize(*sys.argv[2:6])
    else:
        print(__doc__)
        sys.exit(2)
