"""Benchmark: 512^2 projections/s and ray-samples/s on one MI355X, plain vs
surrogate.

Default workload (one step = one render of all views, inputs resident in
HBM): the tessellated gyroid, 6^3 cells of one period each, thickness 0.1,
rho 1, under z += 0.1 tanh(8x); 1024x1024 pixels, 256 views at azimuth
k*360/256, polar 90, simple integrator, ds 0.001.  The plain form (guarded
FP32 march, exact against the FP64 oracle) gives `value`; the surrogate form
(hash-grid MLP trained on 2^24 exact samples, 16x16 tiles checked against
the exact march and re-rendered above 1e-3) is timed beside it, and so are
the same views under the hierarchical integrator, the reference's default
(`hierarchical` object).  Other workloads: --workload
voxel|csg|lattice|gyroid|batch|surrogate.

Prints one JSON line with the throughput, the dominant kernel's roofline
position (HIP-event kernel time, algorithmic bytes or flops per sample)
and the FP64 CPU oracle timed on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import xray_projection_render_amd as X  # noqa: E402
from xray_projection_render_amd import workloads as W  # noqa: E402

METRIC = "512² projections/s & ray-samples/s; plain vs surrogate at ≤1e-3 error"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6

# Algorithmic cost per in-box sample (cost table of the benchmark definition:
# FMA = 2 flops, add/mul/div/sqrt = 1, compares free).
FLOPS_PER_SAMPLE = {"csg": 42.0, "voxel": 54.0, "lattice": 196.0, "gyroid": 35.0}
BYTES_PER_SAMPLE = {"voxel": 32.0}  # 8 float32 corner gathers


def make_workload(name: str) -> W.Workload:
    if name == "voxel":
        return W.voxel()
    if name == "csg":
        return W.csg()
    if name == "lattice":
        return W.lattice()
    if name == "gyroid":
        return W.gyroid_lattice(resolution=512, n_views=16)
    raise SystemExit(f"unknown workload {name}")


def inbox_samples(w: W.Workload, lo: np.ndarray, hi: np.ndarray) -> int:
    """Count of the reference's sample points s_k (FP64 accumulated) that
    fall inside the scene's clip box, over every ray of the workload."""
    smin, smax = w.R - 1.74, w.R + 1.74
    s = []
    x = smin
    while x < smax:
        s.append(x)
        x += w.ds
    s = np.asarray(s)
    res = w.resolution
    half = res / 2.0
    f = 1.0 / math.tan(w.fov / 2 * math.pi / 180)
    ii, jj = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    v = np.stack([ii / half - 1.0, jj / half - 1.0, np.full(ii.shape, -f),
                  np.ones(ii.shape)], -1).reshape(-1, 4)
    total = 0
    for az, po in w.angles:
        eye, cam = X.camera_from_angles(az, po, w.R)
        t = v @ cam.T
        p = t[:, :3] / t[:, 3:4]
        d = p - eye
        d = d / np.linalg.norm(d, axis=1, keepdims=True)
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = (lo - eye) / d
            t2 = (hi - eye) / d
        tn = np.nanmax(np.minimum(t1, t2), axis=1)
        tf = np.nanmin(np.maximum(t1, t2), axis=1)
        ok = tn <= tf
        a = np.searchsorted(s, tn[ok], side="left")
        b = np.searchsorted(s, tf[ok], side="right")
        total += int(np.sum(b - a))
    return total


def inbox_samples_device(w: W.Workload, lo: np.ndarray, hi: np.ndarray) -> int:
    """inbox_samples with the per-view arithmetic in FP64 on the device
    (torch), for workloads with many large views."""
    import torch
    dev = "cuda:0"
    smin, smax = w.R - 1.74, w.R + 1.74
    s, x = [], smin
    while x < smax:
        s.append(x)
        x += w.ds
    s = torch.tensor(s, dtype=torch.float64, device=dev)
    res = w.resolution
    half = res / 2.0
    f = 1.0 / math.tan(w.fov / 2 * math.pi / 180)
    ar = torch.arange(res, dtype=torch.float64, device=dev)
    ii, jj = torch.meshgrid(ar, ar, indexing="ij")
    v = torch.stack([ii / half - 1.0, jj / half - 1.0, torch.full_like(ii, -f),
                     torch.ones_like(ii)], -1).reshape(-1, 4)
    lo_t = torch.tensor(lo, dtype=torch.float64, device=dev)
    hi_t = torch.tensor(hi, dtype=torch.float64, device=dev)
    total = 0
    for az, po in w.angles:
        eye, cam = X.camera_from_angles(az, po, w.R)
        cam_t = torch.tensor(cam, dtype=torch.float64, device=dev)
        eye_t = torch.tensor(eye, dtype=torch.float64, device=dev)
        t = v @ cam_t.T
        p = t[:, :3] / t[:, 3:4]
        d = p - eye_t
        d = d / torch.linalg.norm(d, dim=1, keepdim=True)
        t1 = (lo_t - eye_t) / d
        t2 = (hi_t - eye_t) / d
        tn = torch.nan_to_num(torch.minimum(t1, t2), nan=-math.inf).amax(dim=1)
        tf = torch.nan_to_num(torch.maximum(t1, t2), nan=math.inf).amin(dim=1)
        ok = tn <= tf
        a = torch.searchsorted(s, tn[ok].contiguous(), right=False)
        b = torch.searchsorted(s, tf[ok].contiguous(), right=True)
        total += int((b - a).sum().item())
    return total


def cpu_baseline(w: W.Workload, budget_s: float = 12.0):
    """FP64 oracle (all host cores, dynamic row scheduling) on a bounded
    sample of the workload: whole views in order until ~budget_s of work
    (first a centre row band of view 0 to size the work).

    Returns the baseline record and the oracle images it rendered as a list
    of (view, row0, rows[view rows, res]) for the error check."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if w.voxels is not None:
        nz, nx, ny = w.voxels.shape
        node = O.voxel_node(w.voxels.astype(np.float64).reshape(-1), nx, ny, nz)
    else:
        node = O.build_object(w.obj)
    d = O.build_deformation(w.deformation) if w.deformation else None
    sc = O.Scene(node, d)
    cores = os.cpu_count() or 1
    res = w.resolution
    az, po = w.angles[0]
    band = max(1, res // 32)
    b0 = res // 2 - band // 2
    t = time.perf_counter()
    img = sc.render([az], [po], res, w.ds, R=w.R, fov=w.fov, integrator=w.integrator,
                    nthreads=cores, rows=(b0, b0 + band))
    probe = time.perf_counter() - t
    images = [(0, b0, img[0, b0:b0 + band])]
    if probe * res / band > budget_s:  # even one view is too much: centre rows
        rows = max(1, int(band * budget_s / max(probe, 1e-9)))
        r0 = max(0, res // 2 - rows // 2)
        rows = min(rows, res - r0)
        t = time.perf_counter()
        img = sc.render([az], [po], res, w.ds, R=w.R, fov=w.fov, integrator=w.integrator,
                        nthreads=cores, rows=(r0, r0 + rows))
        spent = time.perf_counter() - t
        images = [(0, r0, img[0, r0:r0 + rows])]
        pix, what = rows * res, f"view 0 rows {r0}..{r0 + rows - 1} of {res}"
    else:
        views, spent = 0, 0.0
        images = []
        while spent < budget_s and views < len(w.angles):
            a, p = w.angles[views]
            t = time.perf_counter()
            img = sc.render([a], [p], res, w.ds, R=w.R, fov=w.fov, integrator=w.integrator,
                            nthreads=cores)
            spent += time.perf_counter() - t
            images.append((views, 0, img[0]))
            views += 1
        pix, what = views * res * res, f"views 0..{views - 1} of {len(w.angles)}, full images"
    rec = {"value": pix / spent / (512 * 512), "unit": "512² projections/s",
           "cores": cores, "kind": "port",
           "sample": f"FP64 oracle (oracle/xrp_oracle.c, full [R-1.74, R+1.74] segment, "
                     f"no clipping), {w.name}: {what}, {spent:.1f} s"}
    return rec, images


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` as recorded by a
    rocprofv3 --pmc summary under profiles/ (a line
    `traffic_bytes_per_launch: <kernel> | <workload> | <bytes>`), or None."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for name in sorted(os.listdir(d)):
        with open(os.path.join(d, name), errors="replace") as fh:
            for line in fh:
                if not line.startswith("traffic_bytes_per_launch:"):
                    continue
                parts = [x.strip() for x in line.split(":", 1)[1].split("|")]
                if len(parts) == 3 and parts[0] == kernel and parts[1] == workload:
                    return float(parts[2])
    return None


PRIM_FLOPS = {"sphere": 8, "box": 3, "cube": 3, "parallelepiped": 18, "cylinder": 20,
              "gyroid": 11}
BATCH_MIN_DS = 0.004  # floor on the per-scene MinFeatureSize/5 step (K <= 870)


def batch_flops_per_sample(obj) -> float:
    """Benchmark cost table for a flat object_collection: base 9, the
    primitives, and n-1 adds for the collection."""
    kids = obj["objects"]
    return 9.0 + sum(PRIM_FLOPS[o["type"]] for o in kids) + (len(kids) - 1)


def bench_batch(args) -> int:
    """32 seeded scenes of 1-200 primitives, 128^2-512^2, 4-64 views each,
    in one graph-captured persistent launch (simple integrator)."""
    import torch  # noqa: F401  (initialises the device the same way as the main path)
    ws = W.batch()
    ctx = X.Context(0)
    scenes, items, flops, samples = [], [], 0.0, 0
    for w in ws:
        sc = X.Scene.from_dict(w.obj)
        # the reference's rule alone (main.go:350-353) with --batch-ds reference
        ds = sc.auto_ds() if args.batch_ds == "reference" else max(sc.auto_ds(), BATCH_MIN_DS)
        w.ds = ds
        scenes.append(sc)
        items.append((sc, w.angles, w.resolution, ds, {"integrator": "simple"}))
        lo, hi, empty = sc.clip_box()
        if empty:
            n = 0
        elif np.all(np.isfinite(lo)) and np.all(np.isfinite(hi)):
            n = inbox_samples(w, lo, hi)
        else:  # unbounded (gyroids): every sample of the segment
            n = w.pixels * X.sample_count(w.R - 1.74, w.R + 1.74, ds)
        samples += n
        flops += n * batch_flops_per_sample(w.obj)
    b = X.Batch(ctx, items)
    print(f"batch: {len(ws)} scenes planned", file=sys.stderr, flush=True)
    # The autotuner times 10 tile shapes on the whole batch (opt-in here);
    # without it, the fastest shape measured on MI355X for this batch: one
    # 8x8 block of rays per work item, 2 workgroups per CU (2.99 s per
    # launch; 2x2 blocks: 3.00 s, 4x1: 3.00 s, 8x2: 3.19 s).
    if args.tune:
        shape = b.tune(repeats=1)
    else:
        shape = (1, 1, 2)
        b.set_shape(*shape)
    for _ in range(max(0, args.warmup - 1)):
        b.run()
        print("batch: warm-up run done", file=sys.stderr, flush=True)
    kms = []
    t0 = time.perf_counter()
    for step in range(args.steps):
        kms.append(b.run()["kernel_ms"])
        print(f"batch: run {step + 1}/{args.steps} {kms[-1]:.0f} ms", file=sys.stderr,
              flush=True)
    elapsed = time.perf_counter() - t0
    st = b.run()
    print(f"batch: {args.steps} timed runs, {ms_per_step_of(elapsed, args.steps):.1f} ms each",
          file=sys.stderr, flush=True)
    lat = b.latencies()
    pixels = sum(w.pixels for w in ws)
    ms_per_step = 1000.0 * elapsed / args.steps
    value = pixels / (512.0 * 512.0) / (ms_per_step / 1000.0)
    k_ms = float(np.mean(kms))
    achieved = flops / (k_ms / 1000.0) / 1e12
    ds_rule = ("MinFeatureSize/5" if args.batch_ds == "reference"
               else f"max(MinFeatureSize/5, {BATCH_MIN_DS})")
    line = {
        "metric": METRIC, "value": value, "unit": "512² projections/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"batch: {len(ws)} scenes, {pixels / 512 ** 2:.1f} 512^2 "
                               f"projections, ds={ds_rule}",
                   "precision": "guarded", "integrator": "simple",
                   "tile_shape": list(shape)},
        "ray_samples_per_s": samples / (ms_per_step / 1000.0),
        "inbox_samples": samples,
        "reevaluated": st["reevaluated"],
        "latency_ms": {"p50": float(np.percentile(lat, 50)),
                       "p99": float(np.percentile(lat, 99)), "max": float(lat.max())},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": measured_traffic("batch_kernel<PREC_GUARDED,INT_SIMPLE>",
                                                 "batch"),
                     "kernel": "batch_kernel<PREC_GUARDED,INT_SIMPLE>", "kernel_ms": k_ms,
                     "per_sample": "cost table per scene", "samples_per_launch": samples},
    }
    if args.batch_ds == "floored":
        valu_issue(line["roofline"], "batch")
    if line["roofline"]["frac"] > 1.0:
        line["roofline"]["note"] = ("the cost-table rate counts the samples a safe radius "
                                    "settles without evaluating them, so it can pass the peak")
    if not args.no_cpu:
        line["cpu_baseline"] = batch_cpu_baseline(ws)
    print(json.dumps(line))
    b.close()
    for sc in scenes:
        sc.close()
    ctx.close()
    return 0


def ms_per_step_of(elapsed: float, steps: int) -> float:
    return 1000.0 * elapsed / max(1, steps)


def batch_cpu_baseline(ws, every: int = 4) -> dict:
    """FP64 oracle time for the whole batch, estimated on the host cores:
    one centre row per core of view 0 for every `every`-th scene, timed; the
    seconds per unit of work (rays x samples x cost-table flops, full
    segment as the oracle marches it) then scale to every scene."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    cores = os.cpu_count() or 1

    def work(w, rays):
        return rays * X.sample_count(w.R - 1.74, w.R + 1.74, w.ds) * batch_flops_per_sample(w.obj)

    spent, sampled = 0.0, 0.0
    for w in ws[::every]:
        sc = O.Scene(O.build_object(w.obj), None)
        rows = min(cores, w.resolution)
        r0 = w.resolution // 2 - rows // 2
        az, po = w.angles[0]
        t = time.perf_counter()
        sc.render([az], [po], w.resolution, w.ds, R=w.R, fov=w.fov, integrator="simple",
                  nthreads=cores, rows=(r0, r0 + rows))
        spent += time.perf_counter() - t
        print(f"cpu baseline: scene {ws.index(w)} {rows} rows of {w.resolution}^2 "
              f"{time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
        sampled += work(w, rows * w.resolution)
    est = spent / sampled * sum(work(w, w.pixels) for w in ws)
    pixels = sum(w.pixels for w in ws)
    return {"value": pixels / (512.0 * 512.0) / est, "unit": "512² projections/s",
            "cores": cores, "kind": "port",
            "sample": f"FP64 oracle, {cores} centre rows of view 0 of every {every}th scene "
                      f"({spent:.1f} s), scaled by rays x samples x cost-table flops to the "
                      f"{len(ws)} scenes"}


BF16_PEAK_TFLOPS = 2500.0
SURROGATE_FLOPS_PER_SAMPLE = 2.0 * (32 * 64 + 64 * 64 + 64)  # MLP multiply-adds
SURROGATE_GATHER_BYTES = 16 * 8 * 4  # hash-grid corners per sample x 2 BF16 features


def bench_surrogate(args) -> int:
    """Tessellated gyroid (6^3 cells, thickness 0.1) under z += 0.1 tanh(8x),
    1024^2, ds 0.001: the exact guarded march against the learned surrogate
    (2^24 exact training samples, step 2 ds, 16x16 tiles checked at 4 pixels
    each against the exact march and re-rendered above 1e-3)."""
    import torch
    n_views = args.views or 16
    w = W.gyroid_lattice(resolution=1024, n_views=n_views, ds=0.001, shear=True)
    ctx = X.Context(0)
    sc = X.Scene.from_dict(w.obj, w.deformation)
    res = w.resolution
    exact = torch.empty((n_views, res, res), dtype=torch.float32, device="cuda:0")
    sur_img = torch.empty_like(exact)
    kw = dict(integrator="simple", R=w.R, fov=w.fov)
    for _ in range(max(1, args.warmup)):
        X.render_device(ctx, sc, w.angles, res, w.ds, exact.data_ptr(), **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st_exact = X.render_device(ctx, sc, w.angles, res, w.ds, exact.data_ptr(), **kw)
    torch.cuda.synchronize()
    exact_ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    print(f"surrogate bench: exact march {exact_ms:.1f} ms per step", file=sys.stderr,
          flush=True)

    sur = X.Surrogate(ctx, sc, log2_table=19, n_min=16, n_max=512)
    tr = sur.train(samples=1 << 24, batch=65536, epochs=4)
    print(f"surrogate bench: trained, loss {tr['first_loss']:.3g} -> {tr['final_loss']:.3g}",
          file=sys.stderr, flush=True)
    ds_sur, tol = 2 * w.ds, 1e-3
    for _ in range(max(1, args.warmup)):
        sur.render_device(w.angles, res, w.ds, ds_sur, tol, sur_img.data_ptr(), **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = sur.render_device(w.angles, res, w.ds, ds_sur, tol, sur_img.data_ptr(), **kw)
    torch.cuda.synchronize()
    sur_ms = 1000.0 * (time.perf_counter() - t0) / args.steps
    diff = (sur_img - exact).abs()
    err_max, err_mean = float(diff.max()), float(diff.mean())
    frac_over = float((diff > tol).float().mean())
    pure = sur.render_device(w.angles, res, w.ds, ds_sur, 0.0, sur_img.data_ptr(), **kw)
    pure_err = float((sur_img - exact).abs().max())
    proj = w.pixels / (512.0 * 512.0)
    sur_samples = pure["network_samples"]  # evaluations issued (each wave's step range)
    achieved = SURROGATE_FLOPS_PER_SAMPLE * sur_samples / (pure["surrogate_ms"] / 1000.0) / 1e12
    line = {
        "metric": METRIC, "value": proj / (sur_ms / 1000.0), "unit": "512² projections/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": sur_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 mfma (surrogate) + f32/f64 (exact check)", "data": "synthetic",
        "config": {"workload": f"surrogate: gyroid 6^3 + tanh shear, {n_views} views x "
                               f"{res}^2, ds={w.ds:g}, surrogate step {ds_sur:g}, tile "
                               f"tolerance {tol:g}", "integrator": "simple"},
        "plain": {"value": proj / (exact_ms / 1000.0), "ms_per_step": exact_ms,
                  "kernel_ms": st_exact["kernel_ms"], "samples": st_exact["samples"],
                  "reevaluated": st_exact["reevaluated"]},
        "speedup_vs_plain": exact_ms / sur_ms,
        "surrogate": {"render": st, "pure_surrogate_ms": pure["surrogate_ms"],
                      "pure_surrogate_max_abs_error": pure_err, "train": tr},
        "error_vs_plain": {"max_abs": err_max, "mean_abs": err_mean,
                           "fraction_above_tolerance": frac_over, "tolerance": tol},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / BF16_PEAK_TFLOPS,
                     "traffic": measured_traffic("sur_render_kernel",
                                                 f"surrogate {n_views} views"),
                     "kernel": "sur_render_kernel", "kernel_ms": pure["surrogate_ms"],
                     "per_sample": f"{SURROGATE_FLOPS_PER_SAMPLE:.0f} flop (MLP)",
                     "samples_per_launch": sur_samples},
    }
    print(json.dumps(line))
    sur.close()
    sc.close()
    ctx.close()
    return 0


def valu_issue(roof: dict, workload: str) -> None:
    """Executed VALU issue of the roofline's kernel, from the VALU
    instructions per launch a rocprofv3 summary under profiles/ records
    (wave instructions x 4 cycles over the 1024 SIMDs' cycles at 2.4 GHz in
    the measured launch time), when one does."""
    insts = measured_counter("valu_insts_per_launch", roof["kernel"], workload)
    if insts and roof.get("kernel_ms"):
        roof["valu_issue_frac"] = insts * 4.0 / (roof["kernel_ms"] / 1000.0 * 1024 * 2.4e9)


def measured_counter(name: str, kernel: str, workload: str):
    """A per-launch figure recorded by a rocprofv3 summary under profiles/
    as a line `<name>: <kernel> | <workload> | <value>`, or None."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for fn in sorted(os.listdir(d)):
        with open(os.path.join(d, fn), errors="replace") as fh:
            for line in fh:
                if not line.startswith(name + ":"):
                    continue
                parts = [x.strip() for x in line.split(":", 1)[1].split("|")]
                if len(parts) == 3 and parts[0] == kernel and parts[1] == workload:
                    return float(parts[2])
    return None


C5_WORKLOAD = "gyroid6_tanh_256x1024"


def march_kernel_name(st: dict, integrator: str = "simple") -> str:
    """The guarded march's dominant kernel as the render ran it: samples per
    node visit, and the first pass of the two-pass form (whose other
    launches are timed with it in kernel_ms)."""
    if integrator != "simple":
        if st.get("two_pass"):
            return f"march_hier_try_kernel<{st['block']},2>"
        return "march_kernel<PREC_GUARDED,INT_HIERARCHICAL>"
    if st.get("two_pass"):
        return f"march_try_kernel<{st['block']}>"
    return f"march_kernel<PREC_GUARDED,INT_SIMPLE,{st['block']}>"


def bench_c5(args) -> int:
    """The metric's configuration: tessellated gyroid 6^3 + z += 0.1 tanh(8x),
    256 views x 1024^2, ds 0.001.  Plain form = the guarded FP32 march
    (`value`); surrogate form = hash-grid MLP + exact tile fallback at 1e-3."""
    import torch
    n_views = args.views or 256
    w = W.gyroid_lattice(resolution=1024, n_views=n_views, ds=0.001, shear=True)
    res = w.resolution
    ctx = X.Context(0)
    sc = X.Scene.from_dict(w.obj, w.deformation)
    lo, hi, empty = sc.clip_box()
    exact = torch.empty((n_views, res, res), dtype=torch.float32, device="cuda:0")
    integ = args.integrator or "simple"
    kw = dict(integrator=integ, precision=args.precision, R=w.R, fov=w.fov)
    for _ in range(args.warmup):
        st = X.render_device(ctx, sc, w.angles, res, w.ds, exact.data_ptr(), **kw)
    torch.cuda.synchronize()
    print("c5: plain warm-up done", file=sys.stderr, flush=True)
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = X.render_device(ctx, sc, w.angles, res, w.ds, exact.data_ptr(), **kw)
        kernel_ms.append(st["kernel_ms"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    print(f"c5: plain {args.steps} steps, {ms_per_step:.1f} ms each", file=sys.stderr,
          flush=True)
    proj = w.pixels / (512.0 * 512.0)
    value = proj / (ms_per_step / 1000.0)
    n_inbox = inbox_samples_device(w, lo, hi)
    k_ms = float(np.mean(kernel_ms))
    kernel = march_kernel_name(st, integ)
    achieved = FLOPS_PER_SAMPLE["gyroid"] * n_inbox / (k_ms / 1000.0) / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": measured_traffic(kernel, C5_WORKLOAD), "kernel": kernel,
            "kernel_ms": k_ms, "per_sample": f"{FLOPS_PER_SAMPLE['gyroid']:.0f} flop + 8 "
                                             "transcendental (cost table)",
            "samples_per_launch": n_inbox, "evaluated_samples": st["samples"] -
            st["skipped_samples"], "skipped_samples": st["skipped_samples"]}
    if integ != "simple":
        roof["fine_samples"] = st["fine_samples"]
    valu_issue(roof, C5_WORKLOAD)
    line = {
        "metric": METRIC, "value": value, "unit": "512² projections/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"gyroid 6^3 + z += 0.1 tanh(8x): {n_views} views x {res}^2, "
                               f"ds={w.ds:g}", "form": "plain (guarded FP32 march)",
                   "precision": {1: "guarded", 2: "fp64"}.get(st["precision_used"], "auto"),
                   "integrator": integ},
        "ray_samples_per_s": n_inbox / (ms_per_step / 1000.0),
        "inbox_samples": n_inbox,
        "reevaluated": st["reevaluated"],
        "roofline": roof,
    }
    if integ == "simple" and not args.no_hier:
        line["hierarchical"] = c5_hier_leg(args, ctx, sc, w, n_inbox, skip_cpu=args.no_cpu)
    if not args.no_surrogate:
        line["surrogate"] = c5_surrogate_leg(args, ctx, sc, w, exact, kw, ms_per_step)
    if not args.no_cpu:
        line["cpu_baseline"], images = cpu_baseline(w)
        err, n = 0.0, 0
        for v, r0, ref in images:
            got = exact[v, r0:r0 + ref.shape[0]].double().cpu().numpy()
            err = max(err, float(np.max(np.abs(got - ref))))
            n += ref.size
        line["error_vs_oracle"] = {"max_abs": err, "pixels": n, "tolerance": 1e-5}
    print(json.dumps(line))
    sc.close()
    ctx.close()
    return 0


def c5_hier_leg(args, ctx, sc, w, n_inbox: int, skip_cpu: bool) -> dict:
    """The same views under the hierarchical integrator, the reference's
    default (main.go:40, api.go:128-132): 1 warm-up and up to 3 timed
    renders, and (with the CPU leg) its error against the FP64 oracle on a
    band of view 0."""
    import torch
    res = w.resolution
    img = torch.empty((len(w.angles), res, res), dtype=torch.float32, device="cuda:0")
    kw = dict(integrator="hierarchical", precision=args.precision, R=w.R, fov=w.fov)
    st = X.render_device(ctx, sc, w.angles, res, w.ds, img.data_ptr(), **kw)
    torch.cuda.synchronize()
    steps = max(1, min(args.steps, 3))
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = X.render_device(ctx, sc, w.angles, res, w.ds, img.data_ptr(), **kw)
        kernel_ms.append(st["kernel_ms"])
    torch.cuda.synchronize()
    ms = 1000.0 * (time.perf_counter() - t0) / steps
    print(f"c5: hierarchical {steps} steps, {ms:.1f} ms each", file=sys.stderr, flush=True)
    out = {"integrator": "hierarchical", "steps": steps, "ms_per_step": ms,
           "value": w.pixels / (512.0 * 512.0) / (ms / 1000.0), "unit": "512² projections/s",
           "kernel": march_kernel_name(st, "hierarchical"),
           "kernel_ms": float(np.mean(kernel_ms)), "coarse_samples": n_inbox,
           "fine_samples": st["fine_samples"], "two_pass": st["two_pass"],
           "checked_rays": st["deferred_rays"], "redo_rays": st["redo_rays"],
           "traffic": measured_traffic(march_kernel_name(st, "hierarchical"), C5_WORKLOAD)}
    if not skip_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        osc = O.Scene(O.build_object(w.obj), O.build_deformation(w.deformation))
        r0, rows = res // 2 - 2, 4
        az, po = w.angles[0]
        ref = osc.render([az], [po], res, w.ds, R=w.R, fov=w.fov, integrator="hierarchical",
                         nthreads=os.cpu_count() or 1, rows=(r0, r0 + rows))[0, r0:r0 + rows]
        got = img[0, r0:r0 + rows].double().cpu().numpy()
        out["error_vs_oracle"] = {"max_abs": float(np.max(np.abs(got - ref))),
                                  "pixels": int(ref.size), "tolerance": 1e-5,
                                  "sample": f"view 0 rows {r0}..{r0 + rows - 1}"}
    del img
    return out


def c5_surrogate_leg(args, ctx, sc, w, exact, kw, plain_ms: float) -> dict:
    """Trains the surrogate (2^24 exact samples) and renders every view in
    both check modes: certified (the exact march judges every pixel; tiles
    beyond 1e-3 take the exact values: the form held to 1e-3) and sampled
    (16 pixels per tile; the escapes are measured).  Errors are taken
    against the plain render, itself checked against the oracle."""
    import torch
    res = w.resolution
    sur = X.Surrogate(ctx, sc, log2_table=19, n_min=16, n_max=512)
    tr = sur.train(samples=1 << 24, batch=65536, epochs=4)
    print(f"c5: surrogate trained, loss {tr['first_loss']:.3g} -> {tr['final_loss']:.3g}",
          file=sys.stderr, flush=True)
    img = torch.empty_like(exact)
    ds_sur, tol = 2 * w.ds, 1e-3
    out = {"config": f"hash grid 16 x 2^19 x 2, MLP 32-64-64-1 (BF16 MFMA), step {ds_sur:g} "
                     f"(midpoint), tolerance {tol:g}", "train": tr}
    l2 = X.diag_l2_bandwidth(ctx)
    for mode, t in (("certified", -tol), ("sampled", tol)):
        sur.render_device(w.angles, res, w.ds, ds_sur, t, img.data_ptr(), **kw)  # warm-up
        torch.cuda.synchronize()
        steps = max(1, args.surrogate_steps)
        t0 = time.perf_counter()
        for _ in range(steps):
            st = sur.render_device(w.angles, res, w.ds, ds_sur, t, img.data_ptr(), **kw)
        torch.cuda.synchronize()
        ms = 1000.0 * (time.perf_counter() - t0) / steps
        diff = (img - exact).abs()
        out[mode] = {
            "value": w.pixels / (512.0 * 512.0) / (ms / 1000.0), "unit": "512² projections/s",
            "ms_per_step": ms, "steps": steps, "speedup_vs_plain": plain_ms / ms,
            "error_vs_plain": {"max_abs": float(diff.max()), "mean_abs": float(diff.mean()),
                               "pixels_above_tolerance": int((diff > tol).sum()),
                               "tolerance": tol},
            "tiles": st["tiles"], "tiles_exact": st["tiles_failed"],
            "check_pixels": st["check_pixels"], "network_samples": st["network_samples"],
            "exact_samples": st["exact_samples"], "network_ms": st["surrogate_ms"],
            "check_ms": st["check_ms"], "fallback_ms": st["fallback_ms"]}
        # the network march against its two ceilings: BF16 MFMA for the MLP,
        # and the hashed gathers (16 levels x 8 corners x 4 B per network
        # sample) against the measured L2 read bandwidth
        net_s = st["surrogate_ms"] / 1000.0
        if net_s > 0:
            tf = st["network_samples"] * SURROGATE_FLOPS_PER_SAMPLE / net_s / 1e12
            gb = st["network_samples"] * SURROGATE_GATHER_BYTES / net_s / 1e9
            out[mode]["roofline"] = {
                "bound": "mfma", "achieved": tf, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tf / BF16_PEAK_TFLOPS, "kernel": "sur_render_kernel",
                "kernel_ms": st["surrogate_ms"],
                "gathers": {"achieved": gb, "unit": "GB/s", "l2_peak_measured": l2,
                            "frac": gb / l2 if l2 > 0 else None}}
        print(f"c5: surrogate {mode}: {ms:.0f} ms, max error {float(diff.max()):.2e}",
              file=sys.stderr, flush=True)
    sur.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c5")
    ap.add_argument("--no-hier", action="store_true",
                    help="c5: skip the hierarchical-integrator leg")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="c5: time the plain form only")
    ap.add_argument("--surrogate-steps", type=int, default=1,
                    help="c5: timed renders of the surrogate form")
    ap.add_argument("--precision", default="auto")
    ap.add_argument("--integrator", default="", help="override the workload's integrator")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--views", type=int, default=0, help="views of the surrogate workload")
    ap.add_argument("--tune", action="store_true", help="autotune the batch tile shape first")
    ap.add_argument("--batch-ds", default="floored", choices=["floored", "reference"],
                    help="batch step: the reference's MinFeatureSize/5 rule, floored at "
                         f"{BATCH_MIN_DS} (default), or the rule alone")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device 3 times: the reference "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(256 * 1024 * 1024 // 4, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": 3 * a.numel() * 4}))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this renderer runs on one GPU; views shard across "
                          "GPUs with no exchange, but only --gpus 1 is measured"}))
        return 2
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "surrogate":
        return bench_surrogate(args)
    if args.workload == "c5":
        return bench_c5(args)

    import torch
    w = make_workload(args.workload)
    if args.integrator:
        w.integrator = args.integrator
    ctx = X.Context(0)
    scene = (X.Scene.from_voxels(w.voxels, w.deformation) if w.voxels is not None
             else X.Scene.from_dict(w.obj, w.deformation))
    lo, hi, empty = scene.clip_box()
    out = torch.empty((len(w.angles), w.resolution, w.resolution), dtype=torch.float32,
                      device="cuda:0")
    kw = dict(integrator=w.integrator, precision=args.precision, R=w.R, fov=w.fov)

    for _ in range(args.warmup):
        st = X.render_device(ctx, scene, w.angles, w.resolution, w.ds, out.data_ptr(), **kw)
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = X.render_device(ctx, scene, w.angles, w.resolution, w.ds, out.data_ptr(), **kw)
        kernel_ms.append(st["kernel_ms"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    projections = w.pixels / (512.0 * 512.0)
    value = projections / (ms_per_step / 1000.0)

    n_inbox = inbox_samples(w, lo, hi) if not empty else 0
    k_ms = float(np.mean(kernel_ms))
    samples_per_s = n_inbox / (ms_per_step / 1000.0)
    prec = {1: "guarded", 2: "fp64"}.get(st["precision_used"], "auto")
    if args.workload in BYTES_PER_SAMPLE:
        achieved = BYTES_PER_SAMPLE[args.workload] * n_inbox / (k_ms / 1000.0) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    else:
        achieved = FLOPS_PER_SAMPLE[args.workload] * n_inbox / (k_ms / 1000.0) / 1e12
        peak = FP32_PEAK_TFLOPS
        roof = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None}
    if w.voxels is not None and prec != "fp64" and w.integrator == "simple":
        layout = os.environ.get("XRP_VOXEL_LAYOUT", "quad_slide")
        roof["kernel"] = {"linear": "voxel_kernel<VOX_LINEAR,CACHE_NONE>",
                          "brick4": "voxel_kernel<VOX_BRICK4,CACHE_NONE>",
                          "cached": "voxel_kernel<VOX_BRICK4,CACHE_CELL>",
                          "linear_cached": "voxel_kernel<VOX_LINEAR,CACHE_CELL>",
                          "pair": "voxel_kernel<VOX_BRICK4_PAIR,CACHE_CELL>",
                          "quad": "voxel_kernel<VOX_BRICK4_QUAD,CACHE_CELL>",
                          "quad_slide": "voxel_kernel<VOX_BRICK4_QUAD,CACHE_SLIDE>",
                          }.get(layout, "voxel_lds_kernel")
    else:
        roof["kernel"] = "march_kernel<%s,%s>" % (
            "PREC_FP64" if prec == "fp64" else "PREC_GUARDED",
            "INT_SIMPLE" if w.integrator == "simple" else "INT_HIERARCHICAL")
        if st.get("block"):
            roof["kernel"] = march_kernel_name(st, w.integrator)
    roof["skipped_samples"] = st.get("skipped_samples", 0)
    if roof["bound"] == "hbm":
        # the gathers mostly hit L2: also state them against the measured L2
        # read bandwidth (xrp_diag_l2_bandwidth)
        l2 = X.diag_l2_bandwidth(ctx)
        roof["l2_peak_measured"] = l2
        roof["l2_frac"] = achieved / l2 if l2 > 0 else None
        if l2 > 0 and achieved > HBM_PEAK_GBS:
            # more algorithmic bytes than HBM could deliver: the cache path
            # binds, so the fraction is taken against the L2 read ceiling
            roof["hbm_frac"] = roof["frac"]
            roof.update(bound="l2", peak=l2, frac=achieved / l2)
    roof["kernel_ms"] = k_ms
    valu_issue(roof, args.workload)
    roof["per_sample"] = (f"{BYTES_PER_SAMPLE[args.workload]:.0f} B" if args.workload
                          in BYTES_PER_SAMPLE else f"{FLOPS_PER_SAMPLE[args.workload]:.0f} flop")
    roof["samples_per_launch"] = n_inbox

    line = {
        "metric": METRIC, "value": value, "unit": "512² projections/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {len(w.angles)} views x {w.resolution}^2, "
                               f"ds={w.ds:g}" + (f", grid {w.voxels.shape[0]}^3 f32"
                                                 if w.voxels is not None else ""),
                   "precision": prec, "integrator": w.integrator},
        "ray_samples_per_s": samples_per_s,
        "inbox_samples": n_inbox,
        "reevaluated": st["reevaluated"],
        "kernel_stats": {k: st[k] for k in ("staged_chunks", "staged_voxels",
                                            "fallback_chunks")},
        "roofline": roof,
    }
    roof["traffic"] = measured_traffic(roof["kernel"], w.name)
    if not args.no_cpu:
        line["cpu_baseline"], images = cpu_baseline(w)
        # the same pixels from the last timed GPU render
        err, n = 0.0, 0
        for v, r0, ref in images:
            got = out[v, r0:r0 + ref.shape[0]].double().cpu().numpy()
            err = max(err, float(np.max(np.abs(got - ref))))
            n += ref.size
        line["error_vs_oracle"] = {"max_abs": err, "pixels": n, "tolerance": 1e-5}
    print(json.dumps(line))
    scene.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
