"""Batched FP32 C2C FFT throughput on one MI355X.

Workload: forward complex<float> FFT, N=4096, 65536 contiguous transforms
(2 GiB in, 2 GiB out), re/im i.i.d. N(0,1) from seed 0, resident in HBM before
the timed region.  A step is one batched launch through the C ABI
(fyfft_execute).  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4096] [--batch 65536]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=16384,
                    help="transforms in the CPU baseline sample")
    return ap.parse_args()


def read_traffic(n, prec):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE), committed under profiles/; None if absent."""
    path = os.path.join(ROOT, "profiles", f"traffic_fft{n}_{prec}.json")
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(n, sample, prec):
    import numpy as np
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((sample, n)) + 1j * rng.standard_normal((sample, n))).astype(np.complex64)
    _, sec = oracle.fft32_batch_timed(x, False, threads)
    flops = 5.0 * n * math.log2(n) * sample
    return {"value": round(flops / sec / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
            "kind": "port",
            "sample": f"{sample} of the {n}-point transforms, FP32 scalar C port of the "
                      f"reference ladder (oracle/), {threads} threads, {sec:.2f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    import fyfft_amd as fy
    import oracle

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n, batch = a.n, a.batch
    prec = fy.F32 if a.precision == "f32" else fy.F64
    rdt = torch.float32 if prec == fy.F32 else torch.float64
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    x = torch.view_as_complex(torch.randn(batch, n, 2, dtype=rdt, device=dev, generator=g))
    y = torch.empty_like(x)
    plan = fy.Plan(n, batch, prec)
    stream = torch.cuda.current_stream(dev)
    h = stream.cuda_stream

    for _ in range(a.warmup):
        plan.execute_ptr(x.data_ptr(), y.data_ptr(), fy.FORWARD, h)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(a.steps):
        plan.execute_ptr(x.data_ptr(), y.data_ptr(), fy.FORWARD, h)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = start.elapsed_time(stop)
    ms = ms_total / a.steps

    # accuracy of the timed output on a sample of rows vs the FP64 oracle
    rows = [0, 1, batch // 2, batch - 1]
    xs = x[rows].cpu().numpy().astype(np.complex128)
    ys = y[rows].cpu().numpy()
    err = oracle.rel_l2(ys, oracle.fft_rows(xs))

    flops = 5.0 * n * math.log2(n) * batch
    elem = 8 if prec == fy.F32 else 16
    alg_bytes = 2.0 * elem * n * batch
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    traffic = read_traffic(n, a.precision)
    out = {
        "metric": "batched C2C GFLOP/s (5NlogN) + % HBM roofline vs plain GPU form, rel-L2<=1e-6",
        "value": round(flops / (ms * 1e-3) / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == fy.F32 else "f64",
        "data": "synthetic N(0,1) re/im, seed 0, resident in HBM",
        "config": {"workload": f"forward C2C FFT N={n} x {batch} contiguous transforms "
                               f"({'complex<float>' if prec == fy.F32 else 'complex<double>'})",
                   "n": n, "batch": batch, "plan": plan.describe()},
        "rel_l2_vs_fp64": err,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic},
        "wall_s": round(wall, 3),
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, min(a.cpu_sample, batch), a.precision)
    plan.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
