"""Benchmark: FFNO3D forward + input VJP throughput (Mvoxels/s) of the HIP engine on one MI355X.

Workload (one "step"): one forward and one input VJP (cotangent fixed) of FFNO3D at the production shape -- B=8
patches, D=128, H=W=64, width 64, vertical modes 32 (z_min_wavelength 1.28 at z_reference_length 40), 8 layers, rank 8,
hidden 128, species H + Ca II -- on a synthetic stratified atmosphere (seed 0) with seeded weights (seed 0). Inputs are
resident in HBM before timing.

Beside it, in the same run: the plain FP32 GPU path (stock torch ops, baseline/plain_fp32.py) on the same inputs
and weights (its speed and the HIP path's relative error in log10 n against it), and the CPU oracle in float32 on a
bounded sample (the CPU workload: B=1, D=64, 32 x 32, width 32, 4 layers) timed on this host's cores.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(B=8, D=128, H=64, W=64, width=64, n_layers=8, z_min_wavelength=1.28, hidden=128, rank=8)
VOLUME = dict(B=1, D=256, H=504, W=504)       # the whole-volume case: same network, one 504 x 504 x 256 atmosphere
CPU_WORKLOAD = dict(B=1, D=64, H=32, W=32, width=32, n_layers=4, z_min_wavelength=3.5, hidden=128, rank=8)
# the production configuration's width and vertical resolution (1001 vertical modes) on the same tile
PRODUCTION = dict(WORKLOAD, width=96, z_min_wavelength=0.04)


def model_config(w):
    from ffnoml_amd.atoms import species
    return dict(in_channels=6, out_channels=12, width=w["width"], n_layers=w["n_layers"], hidden_spectral=w["hidden"],
                hidden_vertical=w["hidden"], freq_mix_rank=w["rank"], z_reference_length=40.0,
                z_min_wavelength=w["z_min_wavelength"], z_feature_wavelength=0.3, species=species(["H", "CA"]))


def make(w, device):
    import numpy as np
    from ffnoml_amd import synthetic
    x, z, mx, sx = synthetic.atmosphere(w["B"], w["D"], w["H"], w["W"], seed=0)
    model = synthetic.random_model(0, **model_config(w))
    model.set_normalization(mx, sx, np.zeros(12), np.ones(12))
    ct = synthetic.cotangent((w["B"], 12, w["D"], w["H"], w["W"]), 0)
    return model.to(device), x.to(device), z.to(device), ct.to(device)


def launch_bytes(name, a):
    """Algorithmic HBM bytes of one launch: the float32 tensors its kernel must read and write once (activations
    [N, C], gates [N, 2C], inputs [N, 6], outputs [N, Cout], depth weights and heights [N]); weights, GroupNorm
    statistics and per-mode tables are not counted."""
    f = 4
    if name == "ffno_tw_gemm":
        g = a[0]._obj
        return f * g.batch * g.n * (g.k + g.m)
    if name == "ffno_mix":
        m = a[0]._obj
        return f * 2 * 2 * m.rows * m.modes * m.C
    if name in ("ffno_gn_finalize", "ffno_gn_grad_finalize"):
        return 0
    rows = a[0]
    vox = rows.B * rows.voxels
    unit = vox * rows.C * f

    def nz(p):
        return 1 if (p.value if hasattr(p, "value") else p) else 0

    def asm(x):
        return nz(x.base) + nz(x.t)

    n_out = 12
    table = {
        "ffno_lift": lambda: vox * f * 7 + unit,
        "ffno_block_in": lambda: unit * (asm(a[1]) + nz(a[2]) + 2) + vox * f * 2,
        "ffno_post": lambda: unit * 2 + vox * f,
        "ffno_fuse": lambda: unit * (4 + 2 * nz(a[15])) + vox * f,
        "ffno_resmlp": lambda: unit * (asm(a[1]) + nz(a[2]) + 1) + vox * f,
        "ffno_head": lambda: unit * (asm(a[1]) + nz(a[2])) + vox * f * (6 + n_out * nz(a[8])),
        "ffno_gn_grad_stats": lambda: unit * 2 + vox * f,
        "ffno_mlp_bwd": lambda: unit * (3 + asm(a[3])) + vox * f,
        "ffno_fuse_bwd": lambda: unit * 8 + vox * f,
        "ffno_fuse_mlp_bwd": lambda: unit * 8 + vox * f,
        "ffno_fuse_gates_bwd": lambda: unit * 11,
        "ffno_gates_bwd": lambda: unit * 5 + vox * f * 2,
        "ffno_head_bwd": lambda: unit * 2 + vox * f * (12 + n_out),
        "ffno_dot_reduce": lambda: unit * 2 + vox * f * 2,
        "ffno_lift_bwd": lambda: unit * 2 + vox * f * 13,
    }
    return table[name]()


def kernel_of(name, a):
    """The kernel (template instance, as rocprof names it, with * for the column-group count) an entry point call
    launches, for the entry points that launch one of several instances."""
    if name == "ffno_tw_gemm":
        g = a[0]._obj
        three = "true" if g.a_l3 else "false"
        if g.m_pad <= 128 and g.k_pad <= 128:
            return f"k_tw_small<{g.m_pad // 16}, {three}, *, {'true' if g.twiddle else 'false'}>"
        return f"k_tw_gemm<{three}>"
    if name == "ffno_mix":
        m = a[0]._obj
        return f"k_mix<{m.C // 16}, *, {'true' if m.parts == 3 else 'false'}, *>"
    return "k_" + name[5:]


def jvp_leg(model, x, z, dev, w, voxels, steps):
    """The tangent (JVP) at the benchmark shape: forward and directional derivative along a random input tangent, as
    the inversion's JVP requests run it (Engine.jvp, the hand-written tangent kernels)."""
    import torch
    from ffnoml_amd import synthetic
    eng = model.engine()
    tan = synthetic.perturbation(tuple(x.shape), seed=1).to(dev)
    with torch.no_grad():
        for _ in range(2):
            eng.jvp(x, z, 48e3, 48e3, tan)
        torch.cuda.synchronize()
        j0, j1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        nj = max(steps, 3)
        j0.record()
        for _ in range(nj):
            eng.jvp(x, z, 48e3, 48e3, tan)
        j1.record()
        torch.cuda.synchronize()
    jms = j0.elapsed_time(j1) / nj
    return {"workload": f"FFNO3D forward + JVP (tangent), B=8 D=128 H=W=64 width={w['width']} layers=8",
            "ms_per_jvp": round(jms, 3), "value": round(voxels / (jms * 1e-3) / 1e6, 3), "unit": "Mvoxels/s",
            "calls": nj}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-baselines", action="store_true", help="skip the plain FP32 GPU and CPU legs")
    ap.add_argument("--no-volume", action="store_true", help="skip the whole-volume 504 x 504 x 256 forward + VJP")
    ap.add_argument("--no-forward-only", action="store_true", help="skip the forward-only leg at the benchmark shape (and the tangent leg with it)")
    ap.add_argument("--no-production", action="store_true",
                    help="skip the production-width leg (width 96, 1001 vertical modes, on the benchmark tile)")
    ap.add_argument("--service-requests", type=int, default=400,
                    help="requests of the serving workload (Poisson arrivals at 40/s, seed 1; 0 skips it; the full "
                         "workload is 2000)")
    ap.add_argument("--service-slots", type=int, default=3, help="execution slots of the serving leg's batcher")
    ap.add_argument("--service-log", default=None,
                    help="write the serving leg's per-batch phase times, admissions and request times to this JSON file")
    ap.add_argument("--plain-service-requests", type=int, default=16,
                    help="the first requests of the same trace served one at a time by the plain FP32 GPU path")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device to device three times (the known-size reference for HBM counters)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 --pmc child runs (FETCH_SIZE, WRITE_SIZE, MFMA and VALU instructions)")
    ap.add_argument("--workload", choices=("benchmark", "production"), default="benchmark",
                    help="the timed step's configuration: the benchmark's (width 64, 32 vertical modes; the line's "
                         "metric) or, for profiling, the production width (96, 1001 vertical modes) on the same tile")
    ap.add_argument("--only-volume", action="store_true",
                    help="run the whole-volume leg alone (for profiling it) and print its record")
    ap.add_argument("--only-jvp", action="store_true",
                    help="run the forward + JVP leg alone at the benchmark shape (for profiling it) and print its record")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported (requests shard whole across GPUs "
              f"with no exchange; run one process per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.pmc_child:
        pmc_child()
        return
    # the counter runs go first, as child processes, before this process touches the GPU
    counters = None if args.no_traffic or args.only_volume or args.only_jvp else measure_counters()

    import torch
    from ffnoml_amd import _lib

    def log(msg):
        print(f"bench: {msg}", file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    if args.only_volume:
        model = make(WORKLOAD, dev)[0]
        print(json.dumps({"whole_volume": volume_leg(model, dev)}), flush=True)
        return
    copy_cal = copy_calibration(dev) if args.copy_calibration else None
    w = WORKLOAD if args.workload == "benchmark" else PRODUCTION
    voxels = w["B"] * w["D"] * w["H"] * w["W"]
    model, x, z, ct = make(w, dev)
    if args.only_jvp:
        print(json.dumps({"jvp": jvp_leg(model, x, z, dev, w, voxels, args.steps)}), flush=True)
        return

    def step():
        xg = x.detach().requires_grad_(True)
        y = model(xg, z, 48e3, 48e3)
        (xb,) = torch.autograd.grad(y, xg, grad_outputs=ct)
        return y, xb

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        y, xb = step()
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(end) / args.steps
    value = voxels / (ms * 1e-3) / 1e6
    log(f"fwd+VJP {ms:.1f} ms per step, {value:.2f} Mvox/s")

    # per-launch timing: HIP events around every launch of one more step, recorded on the stream the launches go
    # to (torch's current stream, which the engine passes to every entry point); grouped by kernel instance, the one
    # with the largest total time is the dominant kernel of the roofline
    events = []
    real_call = _lib.call

    def timed_call(n, *a):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        nbytes = launch_bytes(n, a)
        kern = kernel_of(n, a)
        e0.record()
        real_call(n, *a)
        e1.record()
        events.append((n, kern, nbytes, e0, e1))

    _lib.call = timed_call
    try:
        step()
        torch.cuda.synchronize()
    finally:
        _lib.call = real_call
    per, per_kernel = {}, {}
    for n, kern, nbytes, e0, e1 in events:
        dt = e0.elapsed_time(e1)
        t, b, c = per.get(n, (0.0, 0, 0))
        per[n] = (t + dt, b + nbytes, c + 1)
        t, b, c, _ = per_kernel.get(kern, (0.0, 0, 0, n))
        per_kernel[kern] = (t + dt, b + nbytes, c + 1, n)
    timed_total = sum(v[0] for v in per.values())
    kern = max(per_kernel, key=lambda k: per_kernel[k][0])
    ktot, kbytes, klaunches, entry = per_kernel[kern]
    kms = ktot / klaunches
    achieved = kbytes / (ktot * 1e-3) / 1e9
    ent_t, ent_b, ent_c = per[max(per, key=lambda k: per[k][0])]

    def kernel_row(k):
        t, b, c, n = per_kernel[k]
        row = {"entry_point": n, "launches": c, "ms_per_step": round(t, 2), "avg_us": round(1e3 * t / c, 1),
               "algorithmic_gb_per_launch": round(b / c / 1e9, 4), "hbm_frac": round(b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                                                                    4)}
        row.update(instructions_of(counters, k, c, t))
        return row

    out = {
        "metric": "FFNO3D fwd+VJP Mvoxels/s vs plain FP32 (rel err<=1e-3); service p50/p99 ms",
        "value": round(value, 3),
        "unit": "Mvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "wall_s_timed": round(wall, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 activations, split-bf16 MFMA (fp32 accumulate)",
        "data": "synthetic (stratified random-field atmosphere, seed 0; seeded random weights)",
        "config": {"workload": f"FFNO3D forward + input VJP, B=8 D=128 H=W=64 width={w['width']} "
                               f"modes(D)={int(40.0 / w['z_min_wavelength']) + 1} layers=8",
                   "global_batch": w["B"], "voxels_per_step": voxels, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": kern, "entry_point": entry, "launches": klaunches,
                     "avg_ms": round(kms, 4), "share_of_step": round(ktot / timed_total, 3),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": int(kbytes / klaunches),
                     "traffic": traffic_of(counters, kern),
                     **{k: v for k, v in instructions_of(counters, kern, klaunches, ktot).items()}},
        "entry_points_ms_per_step": {k: round(v[0], 2) for k, v in sorted(per.items(), key=lambda kv: -kv[1][0])},
        "kernels": {k: kernel_row(k) for k in sorted(per_kernel, key=lambda k: -per_kernel[k][0])[:12]},
    }
    if ent_t > ktot:
        out["roofline"]["largest_entry_point"] = {
            "name": max(per, key=lambda k: per[k][0]), "launches": ent_c, "ms_per_step": round(ent_t, 2),
            "achieved_gb_s": round(ent_b / (ent_t * 1e-3) / 1e9, 1),
            "note": "aggregates several kernel instances; the roofline above is the largest single kernel"}

    # the forward alone at the same shape
    fms = None
    if not args.no_forward_only:
        with torch.no_grad():
            for _ in range(2):
                yf = model(x, z, 48e3, 48e3)
            torch.cuda.synchronize()
            f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            nf = max(args.steps, 5)
            f0.record()
            for _ in range(nf):
                yf = model(x, z, 48e3, 48e3)
            f1.record()
            torch.cuda.synchronize()
        fms = f0.elapsed_time(f1) / nf
        out["forward_only"] = {"workload": f"FFNO3D forward, B=8 D=128 H=W=64 width={w['width']} "
                                           f"modes(D)={int(40.0 / w['z_min_wavelength']) + 1} layers=8",
                               "ms_per_forward": round(fms, 3), "value": round(voxels / (fms * 1e-3) / 1e6, 3),
                               "unit": "Mvoxels/s", "forwards": nf}
        del yf
        log(f"forward {fms:.1f} ms, {out['forward_only']['value']:.2f} Mvox/s")
        out["jvp"] = jvp_leg(model, x, z, dev, w, voxels, args.steps)
        log(f"jvp {out['jvp']['ms_per_jvp']:.1f} ms, {out['jvp']['value']:.2f} Mvox/s")

    if not args.no_baselines:
        from baseline.plain_fp32 import PlainFFNO3D
        plain = PlainFFNO3D(model).to(dev)

        def plain_step():
            xg = x.detach().requires_grad_(True)
            yp = plain(xg, z, 48e3, 48e3)
            (xbp,) = torch.autograd.grad(yp, xg, grad_outputs=ct)
            return yp, xbp

        def median_ms(fn, n=3):
            fn()
            torch.cuda.synchronize()
            ts = []
            for _ in range(n):
                s0 = time.perf_counter()
                r = fn()
                torch.cuda.synchronize()
                ts.append((time.perf_counter() - s0) * 1e3)
            return sorted(ts)[len(ts) // 2], r

        pms, (yp, xbp) = median_ms(plain_step)
        rel = float(((y.detach() - yp.detach()).abs() / yp.detach().abs().clamp_min(1.0)).max())
        grel = float((xb - xbp).norm() / xbp.norm())
        y0_plain = yp[:1].detach().double().cpu()
        del yp, xbp
        out["plain_fp32_gpu"] = {"value": round(voxels / (pms * 1e-3) / 1e6, 3), "unit": "Mvoxels/s",
                                 "ms_per_step": round(pms, 1), "speedup": round(pms / ms, 2),
                                 "rel_err_log10n_vs_hip": rel, "grad_rel_err_vs_hip": grel, "timing": "median of 3"}
        if fms is not None:
            with torch.no_grad():
                pfms, _ = median_ms(lambda: plain(x, z, 48e3, 48e3))
            out["forward_only"]["plain_fp32_ms"] = round(pfms, 1)
            out["forward_only"]["speedup_vs_plain_fp32"] = round(pfms / fms, 2)
        log(f"plain FP32: fwd+VJP {pms:.0f} ms")
        del plain
        out["cpu_baseline"], out["accuracy_vs_fp64"] = cpu_leg(model, x[:1].cpu(), z[:1].cpu(),
                                                               y[:1].detach().double().cpu(), y0_plain)
        if fms is not None:
            out["forward_only"]["rel_err_log10n_vs_fp64"] = out["accuracy_vs_fp64"]["hip_rel_err_log10n"]
    if copy_cal:
        out["copy_calibration"] = copy_cal
    del y, xb
    if not args.no_production:
        out["production_width"] = production_leg(dev, log)
    if args.service_requests > 0:
        out["service"] = service_leg(model, dev, args.service_requests,
                                     0 if args.no_baselines else args.plain_service_requests, args.service_slots,
                                     args.service_log)
    if not args.no_volume:
        del x, z, ct
        model.engine().clear_caches()
        torch.cuda.empty_cache()
        out["whole_volume"] = volume_leg(model, dev)
    print(json.dumps(out), flush=True)


def production_leg(dev, log):
    """The production configuration's width and vertical resolution (width 96, 1001 vertical modes at
    z_min_wavelength 0.04, more modes than the 128 levels) on the benchmark tile B=8, D=128, 64 x 64, 8 layers:
    forward and forward + input VJP, median of 3 after one warm-up, and the float64 oracle's error on item 0 is not
    recomputed here (tests/test_gpu_production.py checks this configuration's parity)."""
    import torch
    w = PRODUCTION
    model, x, z, ct = make(w, dev)
    voxels = w["B"] * w["D"] * w["H"] * w["W"]

    def fwd():
        with torch.no_grad():
            return model(x, z, 48e3, 48e3)

    def fwd_vjp():
        xg = x.detach().requires_grad_(True)
        return torch.autograd.grad(model(xg, z, 48e3, 48e3), xg, grad_outputs=ct)[0]

    res = {"workload": "FFNO3D B=8 D=128 H=W=64, width 96, 8 layers, 1001 vertical modes (z_min_wavelength 0.04)"}
    for name, fn in (("forward", fwd), ("fwd_vjp", fwd_vjp)):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        t = sorted(ts)[1]
        res[name + "_ms"] = round(t, 1)
        res[name + "_mvox_s"] = round(voxels / (t * 1e-3) / 1e6, 3)
        log(f"production width: {name} {t:.0f} ms")
    res["checkpoint_stride"] = model.engine().last_stride
    res["blocks_kept_whole"] = model.engine().last_keep
    del model, x, z, ct
    torch.cuda.empty_cache()
    return res


def pmc_child():
    """Run under rocprofv3 --pmc by measure_counters(): one forward + VJP step of the workload, then three 1 GiB
    device copies (the calibration kernel of known size)."""
    import torch
    dev = torch.device("cuda:0")
    model, x, z, ct = make(WORKLOAD, dev)
    xg = x.detach().requires_grad_(True)
    y = model(xg, z, 48e3, 48e3)
    torch.autograd.grad(y, xg, grad_outputs=ct)
    torch.cuda.synchronize()
    copy_calibration(dev)
    torch.cuda.synchronize()


MFMA_FLOP = 2 * 16 * 16 * 32          # one v_mfma_f32_16x16x32_bf16 (the only MFMA the library issues)
BF16_PEAK_TFLOPS = 2500.0               # MI355X dense bf16 MFMA
PMC_PASSES = (("FETCH_SIZE",), ("WRITE_SIZE",), ("SQ_INSTS_MFMA", "SQ_INSTS_VALU"))


def measure_counters(timeout=240):
    """Child runs of this script (--pmc-child) under rocprofv3, one pass per counter set: FETCH_SIZE, then WRITE_SIZE
    (KiB per dispatch), then the MFMA and VALU instruction counts. Returns {counter: {kernel: [values per
    dispatch]}} (kernel = its name with template arguments, as rocprof reports it) or None when rocprofv3 is missing
    or a pass fails."""
    import csv
    import re
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None
    out = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for counters in PMC_PASSES:
            d = os.path.join(tmp, counters[0])
            cmd = ["timeout", "-s", "KILL", str(timeout), exe, "--pmc", *counters, "--output-format", "csv", "-d", d,
                   "-o", "run", "--", sys.executable, "-u", os.path.abspath(__file__), "--pmc-child"]
            print(f"bench: rocprofv3 --pmc {' '.join(counters)}", file=sys.stderr, flush=True)
            r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), stdout=subprocess.DEVNULL,
                               stderr=subprocess.DEVNULL)
            path = os.path.join(d, "run_counter_collection.csv")
            if r.returncode != 0 or not os.path.exists(path):
                print(f"bench: counter pass {counters} failed ({r.returncode}); counters not measured", file=sys.stderr)
                return None
            for row in csv.DictReader(open(path)):
                if row["Counter_Name"] not in counters:
                    continue
                k = re.sub(r"\(.*", "", row["Kernel_Name"].replace("void ", "").replace(
                    "ffno::(anonymous namespace)::", ""))
                out.setdefault(row["Counter_Name"], {}).setdefault(k, []).append(float(row["Counter_Value"]))
    return out


def _matching(per, kern):
    """Values of the rocprof kernels a kernel_of() name stands for ('*' matches one template argument; a name without
    template arguments matches every instance)."""
    import fnmatch
    if "<" not in kern:
        return [v for k, vals in per.items() if k == kern or k.startswith(kern + "<") for v in vals]
    return [v for k, vals in per.items() if fnmatch.fnmatchcase(k, kern) for v in vals]


def traffic_of(counters, kern):
    """HBM bytes per launch of a kernel: the average FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by bytes
    moved / counter value of the 1 GiB device copy (__amd_rocclr_copyBuffer) in the same run."""
    if not counters or "FETCH_SIZE" not in counters:
        return None
    gib_kib = float(1 << 20)
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        per = counters[counter]
        cal = [v for v in per.get("__amd_rocclr_copyBuffer", []) if v > 0.25 * gib_kib]
        vals = _matching(per, kern)
        if not cal or not vals:
            return None
        scale = gib_kib / (sum(cal) / len(cal))
        res[counter] = (sum(vals) / len(vals), scale, len(vals))
    (f, fs, n), (w, ws, _) = res["FETCH_SIZE"], res["WRITE_SIZE"]
    return {"bytes_per_launch": int((f * fs + w * ws) * 1024), "fetch_kib": round(f, 1), "write_kib": round(w, 1),
            "fetch_scale": round(fs, 4), "write_scale": round(ws, 4), "dispatches": n,
            "method": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, one fwd+VJP step; scaled by the 1 GiB copy"}


def instructions_of(counters, kern, launches, total_ms):
    """Executed MFMA and VALU wave-instructions per launch of a kernel (rocprofv3 --pmc SQ_INSTS_MFMA SQ_INSTS_VALU,
    one fwd+VJP step) and the MFMA utilisation they give at the launch duration measured here: executed MFMA FLOP
    (2 x 16 x 16 x 32 per v_mfma_f32_16x16x32_bf16) / duration / the 2.5 PFLOP/s dense bf16 peak."""
    if not counters or "SQ_INSTS_MFMA" not in counters:
        return {}
    mf, va = _matching(counters["SQ_INSTS_MFMA"], kern), _matching(counters["SQ_INSTS_VALU"], kern)
    if not mf or not va:
        return {}
    m, v = sum(mf) / len(mf), sum(va) / len(va)
    sec = total_ms * 1e-3 / launches
    return {"mfma_insts_per_launch": int(m), "valu_insts_per_launch": int(v),
            "mfma_frac": round(m * MFMA_FLOP / sec / (BF16_PEAK_TFLOPS * 1e12), 4)}


SERVICE = dict(rate=40.0, seed=1, hw=(32, 64, 128, 256), depths=(96, 128, 192), vjp_share=0.3)


def service_trace(n, seed=SERVICE["seed"]):
    """The serving workload's requests: Poisson arrivals at 40/s; H = W drawn uniformly from 32, 64, 128, 256 and D
    from 96, 128, 192; 30 % forward + VJP, the rest forward only."""
    import numpy as np
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / SERVICE["rate"], n))
    hw = rng.choice(SERVICE["hw"], n)
    d = rng.choice(SERVICE["depths"], n)
    vjp = rng.random(n) < SERVICE["vjp_share"]
    return [dict(t=float(t[i]), D=int(d[i]), H=int(hw[i]), W=int(hw[i]), op="VJP" if vjp[i] else "PREDICT", seed=i)
            for i in range(n)]


def service_inputs(trace, variants=4):
    """SI arrays per request (features (6, nz, nx, ny), z (nz, nx, ny) in m, cotangent (nz, nx, ny, 12)): one
    stratified synthetic atmosphere per tile shape (the generator of the benchmark's inputs; the periodic random
    fields of a 256 x 256 tile are its 128 x 128 field repeated 2 x 2), in `variants` versions per shape, each rolled
    horizontally by its own random offset with its temperature scaled by 10^(0.01 N(0,1)); request i gets version
    i % variants. VJP cotangents are one random field per shape. Arrays are shared between requests (read only), so
    host memory stays bounded for the 2000-request workload."""
    import numpy as np
    from ffnoml_amd import synthetic
    base, pool, cots = {}, {}, {}
    out = []
    for r in trace:
        key = (r["D"], r["H"], r["W"])
        if key not in base:
            rep = max(1, r["H"] // 128)
            feats, zc = synthetic.atmosphere_features(1, r["D"], r["H"] // rep, r["W"] // rep, seed=sum(key))
            si = np.tile(feats[0], (1, 1, rep, rep))
            si[0] = 10.0 ** si[0]
            si[1:4] *= 1e3
            z = np.broadcast_to(zc[:, None, None] * 1e6, si.shape[1:]).astype(np.float32).copy()
            base[key] = (si.astype(np.float32), z)
        si, z = base[key]
        v = r["seed"] % variants
        if key + (v,) not in pool:
            rng = np.random.default_rng(1000 * sum(key) + v)
            f = np.roll(si, (int(rng.integers(r["H"])), int(rng.integers(r["W"]))), axis=(2, 3))
            f[0] *= np.float32(10.0 ** (0.01 * rng.standard_normal()))
            pool[key + (v,)] = f
        cot = None
        if r["op"] == "VJP":
            if key not in cots:
                cots[key] = np.random.default_rng(sum(key)).standard_normal(key + (12,), dtype=np.float32) * 1e-15
            cot = cots[key]
        out.append((pool[key + (v,)], z, cot))
    return out


def service_leg(model, dev, n, n_plain, slots=3, log_path=None):
    """The serving workload through ffnoml_amd.serving.ContinuousBatcher on the benchmark's model (width 64,
    8 layers): requests submitted at their arrival times from a client thread, batched per shape bucket, admitted by
    HBM footprint, run on concurrent worker streams; latency = completion - arrival. Beside it, the first n_plain
    requests of the same trace served one at a time, in arrival order, by the plain FP32 GPU path (stock torch ops,
    SI chain and autograd VJP as the reference runtime does them)."""
    import threading
    import numpy as np
    import torch
    from ffnoml_amd.backend import GPUBackend
    from ffnoml_amd.serving import ContinuousBatcher

    def log(msg):
        print(f"service: {msg}", file=sys.stderr, flush=True)

    trace = service_trace(n)
    t0 = time.perf_counter()
    inputs = service_inputs(trace)
    log(f"{n} requests generated in {time.perf_counter() - t0:.1f} s")
    backend = GPUBackend(None, model=model, level_names=[f"L{i}" for i in range(12)], device=str(dev))
    batcher = ContinuousBatcher({"m": backend}, slots=slots, max_batch=8)
    stats0 = torch.cuda.memory_stats(dev)
    futures = [None] * n
    try:
        def client():
            start = time.perf_counter()
            for i, (r, (f, z, cot)) in enumerate(zip(trace, inputs)):
                delay = start + r["t"] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                futures[i] = batcher.submit(r["op"], "m", f, z, 48e3, 48e3, cot)

        th = threading.Thread(target=client)
        w0 = time.perf_counter()
        th.start()
        th.join()
        for i, fu in enumerate(futures):
            fu.result()
            if (i + 1) % 20 == 0:
                log(f"{i + 1} of {n} answered, {time.perf_counter() - w0:.1f} s")
        wall = time.perf_counter() - w0
        done = sorted(batcher.completed, key=lambda q: q.t_submit)
        lat = np.array([(q.t_done - q.t_submit) * 1e3 for q in done])
        vox = sum(q.voxels for q in done)
        res = {"workload": f"{n} requests, Poisson 40/s (seed 1), H=W in 32..256, D in 96..192, 30% forward+VJP",
               "requests": n, "p50_ms": round(float(np.percentile(lat, 50)), 1),
               "p99_ms": round(float(np.percentile(lat, 99)), 1), "mean_ms": round(float(lat.mean()), 1),
               "throughput_mvox_s": round(vox / wall / 1e6, 3), "wall_s": round(wall, 2),
               "offered_mvox_s": round(vox / trace[-1]["t"] / 1e6, 3), "peak_batches_in_flight": batcher.peak_inflight,
               "graphs": len(batcher.graphs), "graph_captures": batcher.captures,
               "graph_replays": batcher.replays, "slots": batcher.slots, "max_batch": batcher.max_batch,
               "queue": queue_summary(batcher.queue_samples, w0), "phases": phase_summary(batcher.batch_log, wall)}
        # the caching allocator's events over the window: a failed hipMalloc frees the cache and retries, and a
        # flush synchronises every stream, so every worker waits for the longest batch in flight
        stats1 = torch.cuda.memory_stats(dev)
        res["allocator"] = {k: int(stats1.get(k, 0) - stats0.get(k, 0)) for k in
                            ("num_alloc_retries", "num_sync_all_streams", "num_device_alloc", "num_device_free")}
        if log_path:
            os.makedirs(os.path.dirname(os.path.abspath(log_path)), exist_ok=True)
            with open(log_path, "w") as fh:
                json.dump({"t0": w0, "batches": batcher.batch_log, "queue": batcher.queue_samples,
                           "requests": [(q.t_submit, q.t_start, q.t_done, q.op, q.voxels) for q in done]}, fh)
        log(f"p50 {res['p50_ms']} ms, p99 {res['p99_ms']} ms, {res['throughput_mvox_s']} Mvox/s")
    finally:
        batcher.close()
    if n_plain:
        # the engine's plans and per-mode tables of the trace's shapes go too: the plain path needs the memory
        import gc
        model.engine().clear_caches()
        gc.collect()
        torch.cuda.empty_cache()
        res["plain_fp32_one_at_a_time"] = plain_service(model, dev, trace[:n_plain], inputs[:n_plain], backend)
        log(f"plain: {res['plain_fp32_one_at_a_time']}")
        sub = [(q.t_done - q.t_submit) * 1e3 for q in done[:n_plain]]
        res["first_requests_p50_ms_in_the_full_run"] = round(float(np.percentile(sub, 50)), 1)
    return res


def phase_summary(log, wall):
    """Per operation: batches, mean batch size, Mvoxels, and the seconds the workers spent staging inputs (stack and
    copy in), waiting for the device work and copying out, summed over batches (the workers overlap, so the sums
    can exceed the serving window)."""
    out = {}
    for t, op, nb, vox, stage, compute, output in log:
        o = out.setdefault(op, dict(batches=0, requests=0, mvox=0.0, staging_s=0.0, compute_s=0.0, output_s=0.0))
        o["batches"] += 1
        o["requests"] += nb
        o["mvox"] += vox / 1e6
        o["staging_s"] += stage
        o["compute_s"] += compute
        o["output_s"] += output
    for o in out.values():
        o["mean_batch"] = round(o["requests"] / o["batches"], 2)
        for k in ("mvox", "staging_s", "compute_s", "output_s"):
            o[k] = round(o[k], 3)
    return out


def queue_summary(samples, t0, points=20):
    """Requests waiting at each admission: mean, max and a coarse series over the serving window."""
    import numpy as np
    if not samples:
        return None
    t = np.array([s[0] - t0 for s in samples])
    q = np.array([s[1] for s in samples])
    edges = np.linspace(0.0, max(float(t.max()), 1e-9), points + 1)
    series = [int(q[(t >= a) & (t <= b)].max()) if np.any((t >= a) & (t <= b)) else None
              for a, b in zip(edges[:-1], edges[1:])]
    return {"mean_waiting": round(float(q.mean()), 1), "max_waiting": int(q.max()),
            "max_waiting_per_interval": series, "interval_s": round(float(edges[1] - edges[0]), 2)}


def plain_service(model, dev, trace, inputs, backend, budget_s=90.0):
    """One request at a time, in arrival order, through the plain FP32 GPU path: predict = SI chain + forward under
    no_grad; VJP = forward + torch.autograd.grad of <cotangent, populations> (ffno_fsdp_runtime.py:294-348). Stops
    after budget_s seconds, or before a request that the rate of the same operation's requests so far says would end
    past it (a bounded sample: the requests answered by then)."""
    import numpy as np
    import torch
    from baseline.plain_fp32 import PlainFFNO3D
    plain = PlainFFNO3D(model).to(dev)
    lat, vox = [], 0
    rate = {}                                   # op -> (seconds, voxels) of its answered requests
    start = time.perf_counter()
    for r, (f, z, cot) in zip(trace, inputs):
        now = time.perf_counter() - start
        n_vox = r["D"] * r["H"] * r["W"]
        if r["op"] in rate:
            sec, done_vox = rate[r["op"]]
            if max(now, r["t"]) + sec / done_vox * n_vox > budget_s:
                print(f"service: plain sample stops before request {len(lat) + 1} ({r['op']} {r['D']}x{r['H']}x"
                      f"{r['W']}: about {sec / done_vox * n_vox:.0f} s at this path's rate)", file=sys.stderr,
                      flush=True)
                break
        if now < r["t"]:
            time.sleep(r["t"] - now)
        t_req = time.perf_counter()
        x_si = torch.from_numpy(f).to(dev)[None]
        z_si = torch.from_numpy(z).to(dev)[None]
        if r["op"] == "PREDICT":
            with torch.no_grad():
                xm, zm = backend.model_inputs(x_si, z_si)
                n_pop = backend.linear(backend.log_populations(plain(xm, zm, 48e3, 48e3)))
            n_pop.permute(0, 2, 3, 4, 1).cpu()
        else:
            x_si.requires_grad_(True)
            z_si.requires_grad_(True)
            xm, zm = backend.model_inputs(x_si, z_si)
            n_pop = backend.linear(backend.log_populations(plain(xm, zm, 48e3, 48e3)))
            c = torch.from_numpy(cot).to(dev).permute(3, 0, 1, 2)[None]
            gx, gz = torch.autograd.grad(n_pop, (x_si, z_si), grad_outputs=c)
            gx.cpu(), gz.cpu()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - start - r["t"]) * 1e3)
        vox += n_vox
        sec, done_vox = rate.get(r["op"], (0.0, 0))
        rate[r["op"]] = (sec + time.perf_counter() - t_req, done_vox + n_vox)
        print(f"service: plain request {len(lat)} ({r['op']} {r['D']}x{r['H']}x{r['W']}) answered, "
              f"{time.perf_counter() - start:.1f} s", file=sys.stderr, flush=True)
        if time.perf_counter() - start > budget_s:
            break
    wall = time.perf_counter() - start
    return {"requests": len(lat), "p50_ms": round(float(np.percentile(lat, 50)), 1),
            "p99_ms": round(float(np.percentile(lat, 99)), 1), "throughput_mvox_s": round(vox / wall / 1e6, 3),
            "wall_s": round(wall, 2)}


def copy_calibration(dev, nbytes=1 << 30):
    """Three device-to-device copies of nbytes (torch's copy kernel): the reference of known size that scales the
    HBM byte counters of a profile, and the copy bandwidth this GPU reaches."""
    import torch
    src = torch.ones(nbytes // 4, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 3
    del src, dst
    return {"bytes_per_copy": 2 * nbytes, "ms": round(ms, 4), "gb_s": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


def volume_leg(model, dev):
    """One 504 x 504 x 256 atmosphere (65 M voxels) through the same network on one GPU, no slabbing: forward alone,
    then forward + input VJP (the adjoint chooses its checkpoint stride from the free HBM), peak HBM of each.
    Checks at this size, where the float64 oracle cannot run: (1) the network is equivariant to circular horizontal
    shifts (1x1 layers, column-local vertical branch, FFT-periodic horizontal branch, shift-invariant GroupNorm
    sums), so f(roll(x)) must equal roll(f(x)) up to rounding -- a wrong twiddle or index of the 504-point transforms
    breaks it; (2) the adjoint against a central difference of the forward: the cotangent is all ones and the
    direction d shifts every voxel of channel c by (c + 1) / 6 (normalised units), so that both sides add up
    coherently over the 65 M voxels while the forward's rounding (about 1e-3 in log10 n, irregular in x) averages
    out: sum(f(x + e d) - f(x - e d)) / 2e against <x_bar, d>, e = 1e-2."""
    import torch
    from ffnoml_amd import synthetic

    def log(msg):
        print(f"volume: {msg}", file=sys.stderr, flush=True)

    v = VOLUME
    voxels = v["B"] * v["D"] * v["H"] * v["W"]
    t0 = time.perf_counter()
    x, z, _, _ = synthetic.atmosphere(v["B"], v["D"], v["H"], v["W"], seed=0)
    x, z = x.to(dev), z.to(dev)
    ct = torch.ones(v["B"], 12, v["D"], v["H"], v["W"], device=dev)
    log(f"inputs ready in {time.perf_counter() - t0:.1f} s")
    res = {"workload": "FFNO3D 504x504x256, B=1, width 64, 8 layers, vertical modes 32", "voxels": voxels}

    def timed(fn):
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        r = fn()
        e1.record()
        torch.cuda.synchronize()
        return r, e0.elapsed_time(e1), torch.cuda.max_memory_allocated(dev)

    log("forward")
    with torch.no_grad():
        y, ms, peak = timed(lambda: model(x, z, 48e3, 48e3))
    res["forward_ms"] = round(ms, 1)
    res["forward_mvox_s"] = round(voxels / (ms * 1e-3) / 1e6, 3)
    res["forward_peak_hbm_gb"] = round(peak / 1e9, 2)
    res["output_finite"] = bool(torch.isfinite(y).all())
    log(f"forward {ms:.0f} ms, peak {peak / 1e9:.1f} GB")
    sh, sw = 137, 311
    with torch.no_grad():
        yr = model(torch.roll(x, (sh, sw), (3, 4)), torch.roll(z, (sh, sw), (2, 3)), 48e3, 48e3)
        res["shift_equivariance_rel_err"] = float((yr - torch.roll(y, (sh, sw), (3, 4))).abs().max() /
                                                  y.abs().max().clamp_min(1.0))
    del yr
    log(f"shift equivariance {res['shift_equivariance_rel_err']:.2e}")

    def fwd_vjp():
        xg = x.detach().requires_grad_(True)
        yy = model(xg, z, 48e3, 48e3)
        (g,) = torch.autograd.grad(yy, xg, grad_outputs=ct)
        return g

    log("forward + vjp")
    xb, ms, peak = timed(fwd_vjp)
    res["fwd_vjp_ms"] = round(ms, 1)
    res["fwd_vjp_mvox_s"] = round(voxels / (ms * 1e-3) / 1e6, 3)
    res["fwd_vjp_peak_hbm_gb"] = round(peak / 1e9, 2)
    res["checkpoint_stride"] = model.engine().last_stride
    res["blocks_kept_whole"] = model.engine().last_keep
    log(f"fwd+vjp {ms:.0f} ms, peak {peak / 1e9:.1f} GB, stride {res['checkpoint_stride']}")
    # again, as an inversion loop calls it: the caching allocator now holds the adjoint's buffers (the first call
    # also pays for allocating them from nearly full HBM)
    _, ms2, _ = timed(fwd_vjp)
    res["fwd_vjp_repeat_ms"] = round(ms2, 1)
    res["fwd_vjp_repeat_mvox_s"] = round(voxels / (ms2 * 1e-3) / 1e6, 3)
    log(f"fwd+vjp again {ms2:.0f} ms")
    d = ((torch.arange(6, device=dev, dtype=torch.float32) + 1) / 6).view(1, 6, 1, 1, 1)
    eps = 1e-2
    with torch.no_grad():
        lp = model(x + eps * d, z, 48e3, 48e3).double().sum()
        lm = model(x - eps * d, z, 48e3, 48e3).double().sum()
        fd = float((lp - lm) / (2 * eps))
        ad = float((xb.double() * d.double()).sum())
    res["vjp_dot"], res["central_difference"] = ad, fd
    res["vjp_vs_central_difference_rel_err"] = abs(fd - ad) / max(abs(fd), 1e-30)
    log(f"vjp vs central difference: {ad:.6e} vs {fd:.6e}")
    return res


def cpu_leg(model, x0, z0, y0_hip, y0_plain):
    """The CPU leg: (1) the float32 CPU oracle (oracle/ffno_oracle.py) timed on the CPU workload, forward + input
    VJP, on all host cores; (2) the float64 oracle as the checker of the benchmark's first batch item (B=1, D=128,
    64 x 64, same weights): the relative error in log10 n, max |y - y64| / max(|y64|, 1), of the HIP path and of the
    plain FP32 GPU path."""
    import torch
    from ffnoml_amd import synthetic
    from oracle import ffno_oracle as O
    import numpy as np
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    w = CPU_WORKLOAD
    cfg = model_config(w)
    ocfg_of = lambda c: O.model_config(**{k: c[k] for k in (  # noqa: E731
        "width", "n_layers", "hidden_spectral", "hidden_vertical", "freq_mix_rank", "z_reference_length",
        "z_min_wavelength", "z_feature_wavelength", "species", "out_channels")})
    x, z, mx, sx = synthetic.atmosphere(w["B"], w["D"], w["H"], w["W"], seed=0)
    small = synthetic.random_model(0, **cfg)
    small.set_normalization(mx, sx, np.zeros(12), np.ones(12))
    sd = O.to_dtype({k: v.detach() for k, v in small.state_dict().items()}, torch.float32)
    ocfg = ocfg_of(cfg)
    ct = synthetic.cotangent((w["B"], 12, w["D"], w["H"], w["W"]), 0)
    O.vjp(sd, ocfg, x, z, 48e3, 48e3, ct)
    times = []
    budget = time.perf_counter() + 20.0
    while len(times) < 3 and (not times or time.perf_counter() < budget):
        t0 = time.perf_counter()
        O.vjp(sd, ocfg, x, z, 48e3, 48e3, ct)
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    voxels = w["B"] * w["D"] * w["H"] * w["W"]
    baseline = {"value": round(voxels / t / 1e6, 4), "unit": "Mvoxels/s", "cores": cores, "kind": "port",
                "sample": f"float32 oracle forward + input VJP, B=1 D=64 H=W=32 width 32, 4 layers, 12 vertical "
                          f"modes; median of {len(times)} after 1 warm-up, {t:.2f} s each"}
    t0 = time.perf_counter()
    sd64 = O.to_dtype({k: v.detach().cpu() for k, v in model.state_dict().items()}, torch.float64)
    with torch.no_grad():
        y64 = O.forward(sd64, ocfg_of(model_config(WORKLOAD)), x0.double(), z0.double(), 48e3, 48e3)
    scale = y64.abs().clamp_min(1.0)
    accuracy = {"checker": "float64 oracle, batch item 0 of the benchmark step (B=1 D=128 H=W=64 width 64, 8 layers)",
                "hip_rel_err_log10n": float(((y0_hip - y64).abs() / scale).max()),
                "plain_fp32_rel_err_log10n": float(((y0_plain - y64).abs() / scale).max()),
                "hip_max_abs_dex": float((y0_hip - y64).abs().max()),
                "plain_fp32_max_abs_dex": float((y0_plain - y64).abs().max()),
                "checker_s": round(time.perf_counter() - t0, 1)}
    return baseline, accuracy

if __name__ == "__main__":
    main()
