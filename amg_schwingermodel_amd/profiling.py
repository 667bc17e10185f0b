"""HBM traffic of the SAP Arnoldi steps A and B, from rocprofv3 hardware counters.

Two passes of the same short workload, one per counter (FETCH_SIZE, then
WRITE_SIZE, both in KiB per dispatch), run from the repository root:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- \\
        python3 -m amg_schwingermodel_amd.profiling sample
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- \\
        python3 -m amg_schwingermodel_amd.profiling sample
    python3 -m amg_schwingermodel_amd.profiling summarize OUT/fetch OUT/write \\
        profiles/sap_arnoldi_a_traffic.json

The workload first copies a 256 MiB device buffer three times (a plain copy
kernel of known size: 256 MiB read and 256 MiB written per launch), then runs
two outer iterations of the mixed-precision batched solve (512x512, 3 levels,
32 right-hand sides, m0 = -0.12, beta = 3; FP16 SAP Krylov basis, eager
launches), the configuration of the benchmark's roofline pass.  Each counter is scaled by the
ratio of the copy's known bytes to what the counter reads for the copy.
"""
import csv
import glob
import json
import os
import sys

import numpy as np

COPY_BYTES = 256 << 20
ROWS = [(0, 8, 8, 10, 4, 4), (1, 4, 4, 10, 4, 4)]


def sample(nx=512, nrhs=32, m0=-0.12):
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
    s.kernel_timing(0, True)
    s.solve(B)
    n, ms, nbytes = s.kernel_stats(0)
    print(json.dumps({"sap_a_launches": n, "sap_a_alg_bytes_per_launch": nbytes / max(n, 1)}), flush=True)
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


def summarize(fetch_dir, write_dir, out_json):
    res = {}
    for key, d, counter in (("fetch", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        k = _per_kernel(d, counter)
        copy = [v for name, vals in k.items() if "copy" in name.lower() or "elementwise" in name.lower()
                for v in vals if v * 1024 > 0.25 * COPY_BYTES]
        if not copy:
            raise RuntimeError(f"no copy kernel of {COPY_BYTES} bytes found in {d}")
        scale = COPY_BYTES / (np.median(copy) * 1024)
        sap = [v for name, vals in k.items() if "k_sap_arnoldi_a<float" in name for v in vals]
        res[key] = {"copy_kib_per_launch": float(np.median(copy)), "scale": float(scale),
                    "sap_a_launches": len(sap), "sap_a_bytes_per_launch": float(np.mean(sap) * 1024 * scale)}
    res["hbm_bytes_per_launch"] = res["fetch"]["sap_a_bytes_per_launch"] + res["write"]["sap_a_bytes_per_launch"]
    res["kernel"] = "k_sap_arnoldi_a<float, J>, all J"
    # step B of the same run, with the same copy scale
    b = {}
    for key, d, counter in (("fetch", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        vals = [v for name, vs in _per_kernel(d, counter).items() if "k_sap_arnoldi_b<float" in name for v in vs]
        b["launches"] = len(vals)
        b[f"{key}_bytes_per_launch"] = float(np.mean(vals) * 1024 * res[key]["scale"])
    b["hbm_bytes_per_launch"] = b["fetch_bytes_per_launch"] + b["write_bytes_per_launch"]
    b["kernel"] = "k_sap_arnoldi_b<float, true>"
    res["step_b"] = b
    with open(out_json, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    if sys.argv[1:2] == ["sample"]:
        sample()
    elif sys.argv[1:2] == ["summarize"] and len(sys.argv) == 5:
        summarize(*sys.argv[2:5])
    else:
        print(__doc__)
        sys.exit(2)
