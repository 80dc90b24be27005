"""Batched FP32 C2C FFT throughput on one MI355X.

Default workload (the headline): forward complex<float> FFT, N=4096, 65536
contiguous transforms (2 GiB in, 2 GiB out), re/im i.i.d. N(0,1) from seed 0,
resident in HBM before the timed region.  A step is one batched launch
through the C ABI (fyfft_execute).  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4096] [--batch 65536]

The same line carries `configs`: every other configuration of the metric,
timed in the same run (20 timed steps after 5 warm-up steps each): the
N=1024 single-transform latency (host and device pointers), mixed N=6720
(FP32, FP64), the prime lengths (Rader 65537, Bluestein 10007 and 1000003),
the ragged grouped form against one launch per transform, the spectral
convolution forward+backward and the BF16/FP16 storage forms; each with ms,
roofline fraction and rel-L2 against the FP64 oracle.  --no-configs skips
them (profiling runs).  `cpu_baseline` is the oracle's C port timed on this
host; `cpu_reference_survey` holds the compiled reference's own CPU figures
from the survey container (the reference does not build here).

Other configurations (one JSON line each, same fields where they apply):
  --workload mixed     N=6720 complex<float> fwd+inv, batch 16384 (--precision f64: batch 4096)
  --workload prime     Rader N=65537 x 2048, Bluestein N=10007 x 1024 and N=1000003 x 16
  --workload ragged    20000 ragged transforms in one grouped launch vs one launch per transform
  --workload spectral  spectral convolution [256, 64, 4096], 512 modes, forward + backward
  --workload lowprec   BF16 and FP16 storage, N=4096 x 65536 (FP32-arithmetic and MFMA forms)
  --workload real      real-input R2C and real-output C2R, N=8192 x 65536 and N=2^20 x 256 (float)
  --workload nd        2-D 8192 x 8192 and 3-D 512^3 complex<float>, and a strided (column) batch
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
    ap.add_argument("--no-baseline", action="store_true",
                    help="ragged: skip the one-launch-per-transform baselines (kernel-trace runs)")
    ap.add_argument("--fp32-only", action="store_true",
                    help="ragged: only the FP32 grouped leg (no tuned and BF16 legs; profiling runs)")
    ap.add_argument("--no-configs", action="store_true",
                    help="headline only: skip the other configurations (profiling runs)")
    ap.add_argument("--cpu-sample", type=int, default=65536,
                    help="transforms in the CPU baseline sample (65536 x 4096 points: "
                         "about 20 s of CPU time over 16 threads)")
    ap.add_argument("--workload", choices=["c2c", "mixed", "prime", "ragged", "spectral", "lowprec", "real", "nd"],
                    default="c2c")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="3 FFT launches and 3 plain copies of the same size, for rocprofv3 --pmc")
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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, sample, prec):
    """The oracle's FP32 C port of the reference ladder, timed live on this
    host: all cores up to the 16 of one GPU's host share."""
    import numpy as np
    import oracle
    nproc = os.cpu_count() or 1
    threads = max(1, min(16, nproc))
    rng = np.random.default_rng(0)
    x = rng.standard_normal((sample, 2 * n), dtype=np.float32).view(np.complex64)
    _, sec = oracle.fft32_batch_timed(x, False, threads)
    flops = 5.0 * n * math.log2(n) * sample
    return {"value": round(flops / sec / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
            "kind": "port",
            "sample": f"{sample} of the {n}-point transforms, FP32 scalar C port of the "
                      f"reference ladder (oracle/), {threads} threads of nproc={nproc} "
                      f"({cpu_model()}), {sec:.2f} s wall"}


# The compiled reference (AVX2) measured in the survey container: it cannot
# be built on this host (its sources need Windows headers), so its own
# figures are reported as measured there, labelled, never re-synthesised.
CPU_REFERENCE_SURVEY = {
    "label": "survey container: 8 vCPU Intel Xeon @ 2.7 GHz, g++ 11.4 -O2 -mavx2 -mfma, "
             "one FFT1D_C2C_invoker per thread, out of place, best of 3 (not this host)",
    "unit": "GFLOP/s",
    "rows": [
        {"n": 1024, "strategy": "kernel_direct", "us_per_transform_1t": 3.40, "gflops_1t": 15.05, "gflops_8t": 81.6},
        {"n": 4096, "strategy": "kernel_direct", "us_per_transform_1t": 15.74, "gflops_1t": 15.61, "gflops_8t": 69.5},
        {"n": 16384, "strategy": "kernel_direct", "us_per_transform_1t": 77.1, "gflops_1t": 14.88, "gflops_8t": 70.5},
        {"n": 6720, "strategy": "mix_radix", "us_per_transform_1t": 126.5, "gflops_1t": 3.38, "gflops_8t": 14.2},
        {"n": 10007, "strategy": "rader", "us_per_transform_1t": 595, "gflops_1t": 1.12, "gflops_8t": 8.58},
        {"n": 10006, "strategy": "bluestein", "us_per_transform_1t": 566, "gflops_1t": 1.17, "gflops_8t": 7.99},
    ],
    "unsupported_by_reference": ["N=65537 (falls to an infeasible O(N^2) DFT)", "N=1000003 and ragged lengths "
                                 "> 80000 (beyond max_input_length)", "FP64, batching, BF16, autograd"],
}


def settle(fn, seconds=0.2):
    """Untimed calls of fn() for about `seconds` of wall time, so the GPU's
    clocks have left their idle state before a measurement starts (a cold
    start otherwise costs the first few milliseconds of a run up to 10 %)."""
    import torch
    if seconds <= 0:
        return
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    per = max(time.perf_counter() - t0, 1e-5)
    group = max(1, min(64, int(0.01 / per)))
    while time.perf_counter() - t0 < seconds:
        for _ in range(group):
            fn()
        torch.cuda.synchronize()


def time_steps(fn, steps, warmup, stream, settle_s=0.2):
    """ms per step of fn() on `stream` (HIP events on that stream), after a
    clock settle and `warmup` untimed steps."""
    import torch
    settle(fn, settle_s)
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def line(metric_extra, **kw):
    base = {"metric": "batched C2C GFLOP/s (5NlogN) + % HBM roofline vs plain GPU form, rel-L2<=1e-6",
            "unit": "GFLOP/s", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "data": "synthetic, seeded, resident in HBM"}
    base.update(metric_extra)
    base.update(kw)
    return base


def flops_of(n):
    return 5.0 * n * math.log2(n) if n > 1 else 0.0


def bench_transforms(a, cases, label):
    """cases: list of (n, batch, prec, directions).  One JSON line with a
    per-case table."""
    import numpy as np
    import torch
    import fyfft_amd as fy
    import oracle
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    rows, tot_flops, tot_ms, tot_bytes = [], 0.0, 0.0, 0.0
    for n, batch, prec, dirs in cases:
        rdt = torch.float32 if prec == fy.F32 else torch.float64
        g = torch.Generator(device=dev).manual_seed(0)
        if label == "prime":
            ph = torch.rand(batch, n, dtype=rdt, device=dev, generator=g) * (2 * math.pi)
            x = torch.polar(torch.ones_like(ph), ph)
        else:
            x = torch.view_as_complex(torch.randn(batch, n, 2, dtype=rdt, device=dev, generator=g))
        y = torch.empty_like(x)
        plan = fy.Plan(n, batch, prec)
        h = stream.cuda_stream
        for d in dirs:
            ms = time_steps(lambda: plan.execute_ptr(x.data_ptr(), y.data_ptr(), d, h), a.steps, a.warmup, stream)
            r = [0, batch - 1]
            xs = x[r].cpu().numpy().astype(np.complex128)
            err = oracle.rel_l2(y[r].cpu().numpy(), oracle.fft_rows(xs, inverse=d != fy.FORWARD))
            elem = 8 if prec == fy.F32 else 16
            byt = 2.0 * elem * n * batch
            rows.append({"n": n, "batch": batch, "dtype": "f32" if prec == fy.F32 else "f64",
                         "direction": ["forward", "inverse"][d], "plan": plan.describe(),
                         "ms": round(ms, 4), "GFLOP/s": round(flops_of(n) * batch / ms / 1e6, 1),
                         "hbm_frac": round(byt / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                         "rel_l2_vs_fp64": err})
            tot_flops += flops_of(n) * batch
            tot_ms += ms
            tot_bytes += byt
        plan.close()
        del x, y
        torch.cuda.empty_cache()
    return line({"config": {"workload": label, "cases": rows}}, value=round(tot_flops / tot_ms / 1e6, 2),
                steps=a.steps, warmup=a.warmup, ms_per_step=round(tot_ms, 4),
                dtype="/".join(sorted({r["dtype"] for r in rows})),
                roofline={"bound": "hbm", "achieved": round(tot_bytes / (tot_ms * 1e-3) / 1e9, 1),
                          "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": round(tot_bytes / (tot_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                          "traffic": None})


def bench_real(a):
    """R2C and C2R plans; algorithmic bytes per row: n reals + (n/2+1) complex."""
    import numpy as np
    import torch
    import fyfft_amd as fy
    import oracle
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    rows, tot_flops, tot_ms, tot_bytes = [], 0.0, 0.0, 0.0
    for n, batch in ((8192, 65536), (1 << 20, 256)):
        g = torch.Generator(device=dev).manual_seed(0)
        x = torch.randn(batch, n, device=dev, generator=g)
        X = torch.empty(batch, n // 2 + 1, dtype=torch.complex64, device=dev)
        y = torch.empty_like(x)
        plan = fy.RealPlan(n, batch, fy.F32)
        h = stream.cuda_stream
        byt = 4.0 * n * batch + 8.0 * (n // 2 + 1) * batch
        for d, src, dst in ((fy.FORWARD, x, X), (fy.INVERSE, X, y)):
            ms = time_steps(lambda: plan.execute_ptr(src.data_ptr(), dst.data_ptr(), d, h), a.steps, a.warmup,
                            stream)
            r = [0, batch - 1]
            xs = x[r].cpu().numpy().astype(np.complex128)
            ref = oracle.fft_rows(xs)[:, : n // 2 + 1]
            err = (oracle.rel_l2(X[r].cpu().numpy(), ref) if d == fy.FORWARD
                   else oracle.rel_l2(y[r].cpu().numpy(), x[r].cpu().numpy()))
            flops = 0.5 * flops_of(n) * batch  # half of the complex transform's nominal count
            rows.append({"n": n, "batch": batch, "dtype": "f32",
                         "direction": "R2C" if d == fy.FORWARD else "C2R (vs input)",
                         "plan": plan.describe(), "ms": round(ms, 4),
                         "GFLOP/s": round(flops / ms / 1e6, 1),
                         "hbm_frac": round(byt / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                         "rel_l2_vs_fp64": err})
            tot_flops += flops
            tot_ms += ms
            tot_bytes += byt
        plan.close()
        del x, X, y
        torch.cuda.empty_cache()
    return line({"config": {"workload": "real", "cases": rows}}, value=round(tot_flops / tot_ms / 1e6, 2),
                steps=a.steps, warmup=a.warmup, ms_per_step=round(tot_ms, 4), dtype="f32",
                roofline={"bound": "hbm", "achieved": round(tot_bytes / (tot_ms * 1e-3) / 1e9, 1),
                          "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": round(tot_bytes / (tot_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                          "traffic": None})


def bench_nd(a):
    """Multi-dimensional and strided plans; algorithmic bytes 16 per complex
    element (read once, written once), as for the 1-D plans."""
    import numpy as np
    import torch
    import fyfft_amd as fy
    import oracle
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    rows, tot_flops, tot_ms, tot_bytes = [], 0.0, 0.0, 0.0
    cases = [("2-D", (8192, 8192), 1), ("3-D", (512, 512, 512), 1), ("columns", (4096,), 65536)]
    for label, dims, batch in cases:
        g = torch.Generator(device=dev).manual_seed(0)
        total = int(np.prod(dims)) * batch
        x = torch.view_as_complex(torch.randn(total, 2, device=dev, generator=g))
        y = torch.empty_like(x)
        if label == "columns":
            # 65536 interleaved transforms of length 4096: element j of transform b at b + j*65536
            plan = fy.ManyPlan(dims[0], batch, batch, 1, batch, 1, fy.F32)
        else:
            plan = fy.NdPlan(dims, batch, fy.F32)
        h = stream.cuda_stream
        ms = time_steps(lambda: plan.execute_ptr(x.data_ptr(), y.data_ptr(), fy.FORWARD, h), a.steps, a.warmup,
                        stream)
        if label == "columns":
            cols = [0, batch - 1]
            xs = x.view(dims[0], batch)[:, cols].T.cpu().numpy().astype(np.complex128)
            err = oracle.rel_l2(y.view(dims[0], batch)[:, cols].T.cpu().numpy(), oracle.fft_rows(xs))
        else:
            # size-independent check: the inverse plan recovers the input
            z = torch.empty_like(x)
            plan.execute_ptr(y.data_ptr(), z.data_ptr(), fy.INVERSE, h)
            torch.cuda.synchronize()
            err = float(torch.linalg.vector_norm(z - x) / torch.linalg.vector_norm(x))
            del z
        flops = 5.0 * total * math.log2(int(np.prod(dims)))
        byt = 16.0 * total
        rows.append({"case": label, "dims": list(dims), "batch": batch, "plan": plan.describe()[:300],
                     "ms": round(ms, 4), "GFLOP/s": round(flops / ms / 1e6, 1),
                     "hbm_frac": round(byt / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                     "check": "rel-L2 vs FP64 oracle (2 transforms)" if label == "columns"
                              else "rel-L2 of inverse(forward(x)) vs x", "rel_l2": err})
        tot_flops += flops
        tot_ms += ms
        tot_bytes += byt
        plan.close()
        del x, y
        torch.cuda.empty_cache()
    return line({"config": {"workload": "nd", "cases": rows}}, value=round(tot_flops / tot_ms / 1e6, 2),
                steps=a.steps, warmup=a.warmup, ms_per_step=round(tot_ms, 4), dtype="f32",
                roofline={"bound": "hbm", "achieved": round(tot_bytes / (tot_ms * 1e-3) / 1e9, 1),
                          "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": round(tot_bytes / (tot_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                          "traffic": None})


def bench_ragged(a, compact_out=False):
    import numpy as np
    import torch
    import fyfft_amd as fy
    import oracle
    from fyfft_amd.workloads import ragged_lengths, l_max_for
    dev = torch.device("cuda:0")
    lens = ragged_lengths(count=20000, seed=0)
    tot = int(lens.sum())
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.view_as_complex(torch.randn(tot, 2, device=dev, generator=g))
    y = torch.empty_like(x)
    s = torch.cuda.Stream(dev)
    t0 = time.perf_counter()
    gp = fy.GroupedPlan(lens, fy.F32, fy.GROUPED_GRAPH)
    t_plan = time.perf_counter() - t0
    with torch.cuda.stream(s):
        ms = time_steps(lambda: gp.execute(x, y, fy.FORWARD, stream=s), a.steps, a.warmup, s)
    # accuracy on a sample of transforms of every class
    offs = gp.offsets
    pick = sorted(set([0, len(lens) - 1] + list(np.argsort(lens)[-3:]) + list(range(0, len(lens), 997))))
    errs = []
    yh = y.cpu().numpy()
    xh = x.cpu().numpy()
    for i in pick:
        n, o = int(lens[i]), int(offs[i])
        errs.append(oracle.rel_l2(yh[o:o + n], oracle.fft_ext(xh[o:o + n].astype(np.complex128))))
    ms_base = float("nan")
    if not getattr(a, "no_baseline", False):
        bp = fy.GroupedPlan(lens, fy.F32, fy.GROUPED_PER_TRANSFORM)
        ms_base = time_steps(lambda: bp.execute(x, y, fy.FORWARD), 2, 1, torch.cuda.current_stream(dev), 0)
        bp.close()
    flops = sum(flops_of(int(n)) for n in lens)
    byt = 16.0 * tot
    out = line({"config": {"workload": "ragged grouped C2C, 20000 complex<float> transforms, lengths "
                                       "log-uniform in [16, %d] snapped to 60%% smooth / 25%% Rader / 15%% "
                                       "Bluestein (seed 0)" % round(l_max_for(20000, 2.0 ** 28)),
                           "total_elements": tot, "distinct_lengths": len(set(lens.tolist())),
                           "plan": gp.describe(), "plan_seconds": round(t_plan, 2)}},
               value=round(flops / ms / 1e6, 2), steps=a.steps, warmup=a.warmup,
               ms_per_step=round(ms, 4), dtype="f32",
               rel_l2_vs_fp64_max=max(errs),
               roofline={"bound": "hbm", "achieved": round(byt / (ms * 1e-3) / 1e9, 1),
                         "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(byt / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4), "traffic": None},
               per_transform_baseline={"ms_per_step": round(ms_base, 3),
                                       "speedup_grouped": round(ms_base / ms, 2)})
    gp.close()
    if getattr(a, "fp32_only", False):
        return out
    # the same plan with every sub-plan autotuned (FYFFT_GROUPED_TUNED): the
    # lengths carrying the most elements first, within the tuning budget
    t0 = time.perf_counter()
    gt = fy.GroupedPlan(lens, fy.F32, fy.GROUPED_GRAPH | fy.GROUPED_TUNED)
    t_tune = time.perf_counter() - t0
    with torch.cuda.stream(s):
        ms_t = time_steps(lambda: gt.execute(x, y, fy.FORWARD, stream=s), a.steps, a.warmup, s)
    yh = y.cpu().numpy()
    errs_t = [oracle.rel_l2(yh[int(offs[i]):int(offs[i]) + int(lens[i])],
                            oracle.fft_ext(xh[int(offs[i]):int(offs[i]) + int(lens[i])].astype(np.complex128)))
              for i in pick]
    out["tuned"] = {"ms_per_step": round(ms_t, 4), "roofline_frac": round(byt / (ms_t * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                    "rel_l2_vs_fp64_max": max(errs_t), "plan_seconds": round(t_tune, 1),
                    "speedup_over_untuned": round(ms / ms_t, 3)}
    gt.close()
    del y
    # BF16 storage of the same workload: FP32 arithmetic, 8 bytes per element
    xb = torch.view_as_real(x).to(torch.bfloat16)
    yb = torch.empty_like(xb)
    gpb = fy.GroupedPlan(lens, fy.BF16, fy.GROUPED_GRAPH)
    with torch.cuda.stream(s):
        ms16 = time_steps(lambda: gpb.execute(xb, yb, fy.FORWARD, stream=s), a.steps, a.warmup, s)
    xq = xb.float().cpu().numpy().astype(np.float64)
    yq = yb.float().cpu().numpy().astype(np.float64)
    worst = 0.0
    for i in pick:
        n, o = int(lens[i]), int(offs[i])
        ex = oracle.fft_ext(xq[o:o + n, 0] + 1j * xq[o:o + n, 1])
        ext = np.stack([ex.real, ex.imag], -1)
        floor = np.linalg.norm(torch.from_numpy(ext).to(torch.bfloat16).double().numpy() - ext) / np.linalg.norm(ext)
        err = np.linalg.norm(yq[o:o + n] - ext) / np.linalg.norm(ext)
        worst = max(worst, err / floor)
    ms16_base = float("nan")
    if not getattr(a, "no_baseline", False):
        bpb = fy.GroupedPlan(lens, fy.BF16, fy.GROUPED_PER_TRANSFORM)
        ms16_base = time_steps(lambda: bpb.execute(xb, yb, fy.FORWARD), 2, 1, torch.cuda.current_stream(dev), 0)
        bpb.close()
    out["bf16"] = {"ms_per_step": round(ms16, 4), "value": round(flops / ms16 / 1e6, 2),
                   "roofline_frac": round(8.0 * tot / (ms16 * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                   "error_vs_bf16_rounding_floor_max": round(worst, 4),
                   "per_transform_baseline_ms": round(ms16_base, 3),
                   "speedup_grouped": round(ms16_base / ms16, 2), "plan": gpb.describe()}
    gpb.close()
    if compact_out:
        return {"ms_per_step": out["ms_per_step"], "value": out["value"], "unit": out["unit"],
                "roofline_frac": out["roofline"]["frac"], "rel_l2_vs_fp64_max": out["rel_l2_vs_fp64_max"],
                "per_transform_baseline_ms": out["per_transform_baseline"]["ms_per_step"],
                "speedup_grouped": out["per_transform_baseline"]["speedup_grouped"],
                "workload": out["config"]["workload"], "plan": out["config"]["plan"], "tuned": out["tuned"],
                "bf16": out["bf16"]}
    return out


def bench_spectral(a):
    import torch
    from fyfft_amd.autograd import SpectralConv1d
    dev = torch.device("cuda:0")
    B, C, L, K = 256, 64, 4096, 512
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.randn(B, C, L, device=dev, generator=g, requires_grad=True)
    layer = SpectralConv1d(C, C, K, device=dev, generator=g)
    gy = torch.randn(B, C, L, device=dev, generator=g)
    stream = torch.cuda.current_stream(dev)

    def step():
        x.grad = None
        layer.weight_kio.grad = None
        y = layer(x)
        y.backward(gy)
    ms = time_steps(step, a.steps, a.warmup, stream)
    fft_flops = 2 * 2 * B * C * flops_of(L)          # fwd: 2 transforms, bwd: 2 transforms
    mix_flops = 3 * 8.0 * B * C * C * K              # one forward and two backward GEMMs
    return line({"config": {"workload": f"spectral convolution forward+backward, x float32 "
                                        f"[{B}, {C}, {L}], W complex64 [{C}, {C}, {K}]"}},
                value=round((fft_flops + mix_flops) / ms / 1e6, 2), steps=a.steps,
                warmup=a.warmup, ms_per_step=round(ms, 4), dtype="f32",
                unit="GFLOP/s")


def bench_lowprec(a):
    import numpy as np
    import torch
    import fyfft_amd as fy
    import oracle
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    n, batch = a.n, a.batch
    rows, tot_ms, tot_bytes, tot_flops = [], 0.0, 0.0, 0.0
    for name, tdt, prec, form in (("bf16", torch.bfloat16, fy.BF16, fy.FORM_AUTO),
                                  ("fp16", torch.float16, fy.F16, fy.FORM_AUTO),
                                  ("bf16", torch.bfloat16, fy.BF16, fy.FORM_MFMA),
                                  ("fp16", torch.float16, fy.F16, fy.FORM_MFMA)):
        g = torch.Generator(device=dev).manual_seed(0)
        x = torch.randn(batch, n, 2, device=dev, generator=g).to(tdt)
        y = torch.empty_like(x)
        with fy.Plan(n, batch, prec, form) as p:
            h = stream.cuda_stream
            ms = time_steps(lambda: p.execute_ptr(x.data_ptr(), y.data_ptr(), fy.FORWARD, h),
                            a.steps, a.warmup, stream)
            desc = p.describe()
        r = [0, batch - 1]
        xq = x[r].float().cpu().numpy().astype(np.float64)
        ex = oracle.fft_rows(xq[..., 0] + 1j * xq[..., 1])
        ext = torch.from_numpy(np.stack([ex.real, ex.imag], -1))
        err = float((y[r].cpu().double() - ext).norm() / ext.norm())
        floor = float((ext.to(tdt).double() - ext).norm() / ext.norm())
        byt = 2.0 * 4 * n * batch
        rows.append({"dtype": name, "form": "mfma" if form == fy.FORM_MFMA else "auto", "plan": desc,
                     "ms": round(ms, 4),
                     "GFLOP/s": round(flops_of(n) * batch / ms / 1e6, 1),
                     "hbm_frac": round(byt / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                     "rel_l2_vs_fp64": err, "storage_rounding_floor": floor})
        tot_ms += ms
        tot_bytes += byt
        tot_flops += flops_of(n) * batch
        del x, y
    return line({"config": {"workload": f"low-precision storage, forward C2C N={n} x {batch}: FP32 "
                                        "arithmetic with 16-bit storage (auto form) and radix-16 stages "
                                        "on MFMA with split hi/lo fp16 operands (mfma form)",
                            "cases": rows}},
                value=round(tot_flops / tot_ms / 1e6, 2), steps=a.steps, warmup=a.warmup,
                ms_per_step=round(tot_ms, 4), dtype="bf16/fp16",
                roofline={"bound": "hbm", "achieved": round(tot_bytes / (tot_ms * 1e-3) / 1e9, 1),
                          "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": round(tot_bytes / (tot_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                          "traffic": None})


def bench_latency(a):
    """The reference's own configuration: one N = 1024 forward transform of
    complex<float> on HOST buffers (dispatch<false>, synchronous), as a
    drop-in caller runs it; plus the same transform on device pointers."""
    import numpy as np
    import torch
    import fyfft_amd as fy
    import oracle
    n, calls = 1024, 2000
    rng = np.random.default_rng(0)
    x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
    inv = fy.FFT1D_C2C_invoker(n)
    xf, out = x.view(np.float32), np.empty(2 * n, np.float32)
    for _ in range(100):
        inv.dispatch(xf, out, invert=False)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        for _ in range(calls // 5):
            inv.dispatch(xf, out, invert=False)
        ts.append((time.perf_counter() - t0) / (calls // 5))
    host_us = min(ts) * 1e6
    err = oracle.rel_l2(out.view(np.complex64), oracle.dft_naive(x.astype(np.complex128)))
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    with fy.Plan(n, 1, fy.F32) as p:
        h = torch.cuda.current_stream().cuda_stream
        for _ in range(100):
            p.execute_ptr(xd.data_ptr(), yd.data_ptr(), fy.FORWARD, h)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            p.execute_ptr(xd.data_ptr(), yd.data_ptr(), fy.FORWARD, h)
            torch.cuda.synchronize()
        dev_us = (time.perf_counter() - t0) / calls * 1e6
    return {"workload": "one forward C2C N=1024 complex<float>, batch 1 (latency-bound)",
            "us_per_call_host_pointers": round(host_us, 2),
            "us_per_call_device_pointers": round(dev_us, 2),
            "reference_cpu_us_1t_survey": 3.40, "rel_l2_vs_fp64": err,
            "note": "host path = copy in, launch, copy out, synchronise; the reference's 3.40 us is "
                    "one AVX2 core in the survey container"}


def bench_configs(a):
    """Every configuration of the metric, compact, for the default line."""
    import argparse as _ap
    import torch
    import fyfft_amd as fy
    sub = _ap.Namespace(**vars(a))
    sub.steps, sub.warmup = 20, 5
    keep = ("n", "batch", "dtype", "direction", "form", "plan", "ms", "hbm_frac", "rel_l2_vs_fp64",
            "storage_rounding_floor")

    def compact(line_):
        rows = line_["config"].get("cases", [])
        return {"ms_per_step": line_["ms_per_step"], "value": line_["value"], "unit": line_["unit"],
                "roofline_frac": line_.get("roofline", {}).get("frac"),
                "cases": [{k: r[k] for k in keep if k in r} for r in rows]}

    out = {}
    steps = [
        ("config1_latency", lambda: bench_latency(sub)),
        ("mixed", lambda: compact(bench_transforms(sub, [(6720, 16384, fy.F32, (fy.FORWARD, fy.INVERSE)),
                                                        (6720, 4096, fy.F64, (fy.FORWARD, fy.INVERSE))],
                                                   "mixed"))),
        ("prime", lambda: compact(bench_transforms(sub, [(65537, 2048, fy.F32, (fy.FORWARD,)),
                                                        (10007, 1024, fy.F32, (fy.FORWARD,)),
                                                        (1000003, 16, fy.F32, (fy.FORWARD,))], "prime"))),
        ("ragged", lambda: bench_ragged(sub, compact_out=True)),
        ("spectral", lambda: {k: v for k, v in bench_spectral(sub).items()
                              if k in ("value", "unit", "ms_per_step", "config")}),
        ("lowprec", lambda: compact(bench_lowprec(sub))),
    ]
    for name, fn in steps:
        t0 = time.perf_counter()
        try:
            out[name] = fn()
        except Exception as e:  # report, keep the other configurations
            out[name] = {"error": f"{type(e).__name__}: {e}"}
        out[name]["bench_seconds"] = round(time.perf_counter() - t0, 1)
        torch.cuda.empty_cache()
        print(f"# config {name} done in {out[name]['bench_seconds']} s", file=sys.stderr, flush=True)
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.workload != "c2c":
        import torch
        import fyfft_amd as fy
        torch.cuda.set_device(0)
        if a.workload == "mixed":
            cases = [(6720, 16384, fy.F32, (fy.FORWARD, fy.INVERSE)),
                     (6720, 4096, fy.F64, (fy.FORWARD, fy.INVERSE))]
            out = bench_transforms(a, cases, "mixed")
        elif a.workload == "prime":
            cases = [(65537, 2048, fy.F32, (fy.FORWARD,)), (10007, 1024, fy.F32, (fy.FORWARD,)),
                     (1000003, 16, fy.F32, (fy.FORWARD,))]
            out = bench_transforms(a, cases, "prime")
        elif a.workload == "ragged":
            out = bench_ragged(a)
        elif a.workload == "lowprec":
            out = bench_lowprec(a)
        elif a.workload == "real":
            out = bench_real(a)
        elif a.workload == "nd":
            out = bench_nd(a)
        else:
            out = bench_spectral(a)
        print(json.dumps(out), flush=True)
        return
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
    if a.pmc_probe:
        # counter calibration: the FFT kernel and an elementwise copy kernel
        # (read + write of the same 2*elem*n*batch bytes)
        xr, yr = torch.view_as_real(x), torch.view_as_real(y)
        for _ in range(3):
            plan.execute_ptr(x.data_ptr(), y.data_ptr(), fy.FORWARD, h)
            torch.mul(xr, 1.0, out=yr)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": plan.describe(), "bytes_each_way": x.numel() * x.element_size()}))
        return

    # half a second of untimed calls: on a freshly started GPU 0.2 s has
    # occasionally left the first timed steps ~10 % slow
    settle(lambda: plan.execute_ptr(x.data_ptr(), y.data_ptr(), fy.FORWARD, h), 0.5)
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
    # the same workload on the plain (generic) kernels
    with fy.Plan(n, batch, prec, fy.FORM_PLAIN) as pp:
        ms_plain = time_steps(lambda: pp.execute_ptr(x.data_ptr(), y.data_ptr(), fy.FORWARD, h),
                              max(10, a.steps // 4), max(3, a.warmup // 4), stream)
        out["plain_form"] = {"plan": pp.describe(), "ms_per_step": round(ms_plain, 5),
                             "hbm_frac": round(alg_bytes / (ms_plain * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                             "speedup_of_value": round(ms_plain / ms, 3)}
    plan.close()
    del x, y
    torch.cuda.empty_cache()
    if not a.no_configs:
        out["configs"] = bench_configs(a)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, min(a.cpu_sample, batch), a.precision)
        out["cpu_reference_survey"] = CPU_REFERENCE_SURVEY
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
