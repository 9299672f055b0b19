#!/usr/bin/env python3
"""Benchmark of the FMM evaluation: potential + gradient of 16M Plummer points.

Workload (default): N = 2^24 points of a Plummer sphere (scale radius 1,
truncated at r = 50), charges uniform in (0,1], expansion order p = 12, leaf
capacity 64, adaptive octree; output phi and grad phi in FP64.  One step is one
evaluation of all N points with charges and outputs resident in HBM.

Prints one JSON line: points/sec, the per-stage device times, the error against
an FP64 direct sum at 10k sampled targets, the roofline of the dominant kernel
(M2L) and the reference CPU FMM timed on this host's cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from simplefmm_amd import api, workloads as W  # noqa: E402

PEAK_FP64 = 78.6e12    # MI355X dense FP64 (vector and MFMA), FLOP/s
PEAK_FP32 = 157.3e12   # MI355X dense FP32 MFMA, FLOP/s
PEAK_F16 = 2.5e15      # MI355X dense FP16 MFMA, FLOP/s
PEAK_HBM = 8.0e12      # bytes/s

WORKLOADS = {
    # name: (generator, n, seed, p, cap); batch256: 256 independent cluster problems,
    # sizes log-uniform in [1e4, 1e6], p from {6, 8, 10, 12} (n, seed, p unused)
    "plummer16m_p12": ("plummer", 1 << 24, 3, 12, 64),
    "batch256": ("clusters", 0, 11, 0, 64),
    "uniform4m_p10": ("uniform_cube", 1 << 22, 1, 10, 64),
    "uniform100k_p8": ("uniform_cube", 100_000, 1, 8, 64),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def m2l_flops_complex(p: int) -> int:
    """Per (pair, component) flops of rotate-translate-rotate M2L on complex
    full-m expansions (8 n_rot + 5 n_coax + 12 (p+1)^2)."""
    n_rot = (p + 1) * (2 * p + 1) * (2 * p + 3) // 3
    n_coax = sum((p + 1 - abs(k)) ** 2 for k in range(-p, p + 1))
    return 8 * n_rot + 5 * n_coax + 12 * (p + 1) ** 2


def m2l_flops_real(p: int) -> int:
    """Per (pair, component) flops of the same M2L on the real form with
    conjugate symmetry (what the kernel executes): two polar rotations and the
    coaxial translation, each sum_n (n+1)^2 + n^2 multiply-adds, plus the two
    azimuthal 2x2 rotations of every (n, m>0)."""
    mac = sum((n + 1) ** 2 + n * n for n in range(p + 1))
    azi = 4 * p * (p + 1) // 2
    return 2 * (3 * mac + 2 * azi)


def m2l_flops_split(p: int) -> int:
    """Per (pair, component) flops of the mixed path's M2L as it runs on the f16
    matrix cores (p <= 12): the dense real-form class operator, a cos block of
    (p+1)(p+2)/2 and a sin block of p(p+1)/2 rows and columns, applied as three
    split products (hi x hi, lo x hi, hi x lo), 2 flops per multiply-add."""
    nc = (p + 1) * (p + 2) // 2
    ns = p * (p + 1) // 2
    return 3 * 2 * (nc * nc + ns * ns)


def cpu_baseline(gen: str, p: int, cap: int, sample_n: int, seed: int):
    """Reference CPU FMM (oracle/_ref, built from the reference sources) on a
    bounded sample of the workload, all host cores as MPI ranks."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness_v4")
    mpiexec = "/opt/conda/bin/mpiexec"
    if not (os.path.exists(exe) and os.path.exists(mpiexec)):
        return None, "reference CPU build not present"
    cores = len(os.sched_getaffinity(0)) or 1  # the CPUs this process may run on
    pos, q = getattr(W, gen)(sample_n, seed=seed)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        W.write_harness_input(fin, pos, q)
        try:
            r = subprocess.run([mpiexec, "-n", str(cores), exe, fin, "-", str(p), str(cap), "20", "time", "-", "1"],
                               capture_output=True, text=True, timeout=240)
        except subprocess.TimeoutExpired:
            return None, "reference CPU run timed out"
    if r.returncode != 0:
        return None, f"reference CPU run failed ({r.returncode}): {r.stderr.strip()[-200:]}"
    rec = json.loads([ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1])
    t_eval = rec["rinv_s"] + rec["rinv_curl_s"] + rec["rinv_curl2_s"]
    return {
        "value": sample_n / t_eval, "unit": "points/sec", "cores": cores, "kind": "reference",
        "sample": (f"{gen} N={sample_n} p={p} cap={cap}: reference FMM3D<double,3,{p},Vector,true>, "
                   f"phi by rinv + grad phi by two rinv<true> (curl identity), {cores} MPI ranks; "
                   f"eval {t_eval:.2f}s, plan build {rec['ctor_s']:.2f}s, tree {rec['tree_s']:.2f}s"),
    }, None


def traffic_from_profile(workload: str):
    """HBM bytes per M2L launch {"fp64": B, "mixed": B} from the committed counter
    summary of this workload (profiles/traffic_<workload>.md, line
    `m2l_traffic_bytes: fp64 <B> mixed <B>`), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("m2l_traffic_bytes:"):
                tok = line.split()
                return {tok[i]: float(tok[i + 1]) for i in range(1, len(tok) - 1, 2)}
    return None


def run_batch(args) -> int:
    """Batched workload: every problem of the batch evaluated by one batch launch
    per step (sfmm_batch_*: tuned number of concurrent streams), against the same
    evaluations issued one plan at a time on the plans' own streams."""
    import torch

    if not torch.cuda.is_available():
        log("[bench] no GPU visible")
        return 3
    specs = W.batch_specs(256)
    dev = torch.device("cuda:0")
    t0 = time.time()
    probs = []
    for n, p, seed in specs:
        pos, q = W.clusters(n, seed=seed)
        part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 64)
        probs.append(dict(n=n, p=p, pos=pos, q=q, part=part, dq=torch.from_numpy(q).to(dev),
                          du=torch.zeros((n, 4), dtype=torch.float64, device=dev)))
    total = sum(pb["n"] for pb in probs)
    log(f"[bench] batch: {len(probs)} problems, {total} points, trees {time.time() - t0:.1f}s")
    res = {}
    paths = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    for prec in paths:
        t_a = time.time()
        fmms = [api.FMM3D(pb["part"], pb["p"], ncharge=1, precision=api.MIXED if prec == "mixed" else api.FP64)
                for pb in probs]
        items = [(f, api.POTGRAD, 1, pb["dq"].data_ptr(), pb["du"].data_ptr()) for f, pb in zip(fmms, probs)]
        batch = api.Batch(items, streams=0)
        t_b = time.time()
        for _ in range(args.warmup):
            batch.launch()
        batch.synchronize()
        ta = time.perf_counter()
        for _ in range(args.steps):
            batch.launch()
        batch.synchronize()
        ms = (time.perf_counter() - ta) * 1e3 / args.steps
        # the same evaluations issued one plan at a time
        ta = time.perf_counter()
        for _ in range(args.steps):
            for f, pb in zip(fmms, probs):
                f.evaluate_device(api.POTGRAD, 1, pb["dq"].data_ptr(), pb["du"].data_ptr())
            for f in fmms:
                f.synchronize()
        ms_eager = (time.perf_counter() - ta) * 1e3 / args.steps
        # accuracy on the largest problem against the FP64 direct sum
        big = max(range(len(probs)), key=lambda k: probs[k]["n"])
        pb = probs[big]
        pb["du"].zero_()
        torch.cuda.synchronize()
        fmms[big].evaluate_device(api.POTGRAD, 1, pb["dq"].data_ptr(), pb["du"].data_ptr())
        fmms[big].synchronize()
        tg = np.sort(np.random.default_rng(7).choice(pb["n"], min(args.targets, pb["n"]), replace=False))
        ref = fmms[big].direct(api.POTGRAD, pb["q"], tg)
        U = pb["du"].cpu().numpy()

        def rel(a, b):
            return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))

        res[prec] = {"points_per_sec": total / (ms * 1e-3), "ms_per_step": ms, "streams": batch.streams,
                     "eager_ms_per_step": ms_eager, "plan_build_s": t_b - t_a,
                     "error_largest": {"n": pb["n"], "p": pb["p"], "targets": int(len(tg)),
                                       "rel_l2_phi": rel(U[tg, 0], ref[:, 0]),
                                       "rel_l2_grad": rel(U[tg, 1:], ref[:, 1:])}}
        log(f"[bench] batch {prec}: {ms:.2f} ms/step batched ({batch.streams} streams), "
            f"{ms_eager:.2f} ms eager; largest problem error {res[prec]['error_largest']}")
        del batch, fmms
    head = res["mixed"] if "mixed" in res else res["fp64"]
    print(json.dumps({
        "metric": "points/sec (pot+grad), batch of independent problems, 1 MI355X",
        "value": head["points_per_sec"], "unit": "points/sec", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if "mixed" not in res else "f64+f32", "data": "synthetic",
        "config": {"workload": "batch256", "problems": len(probs), "points": total, "generator": "clusters",
                   "orders": [6, 8, 10, 12], "leaf_capacity": 64, "output": "phi+grad",
                   "parallelism": "single GPU"},
        "paths": res,
        "speedup_mixed_vs_fp64": (res["mixed"]["points_per_sec"] / res["fp64"]["points_per_sec"]
                                  if len(res) == 2 else None),
    }), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plummer16m_p12", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="override the particle count")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--targets", type=int, default=10_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 18)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="first run a device-to-device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (multi-GPU sharding is not built)",
              file=sys.stderr)
        return 2

    if args.workload == "batch256":
        return run_batch(args)
    gen, n, seed, p, cap = WORKLOADS[args.workload]
    if args.n:
        n = args.n
    log(f"[bench] workload {args.workload}: {gen} n={n} p={p} cap={cap}")

    # CPU baseline first, before this process touches the GPU
    cpu, cpu_note = (None, "skipped (--no-cpu)") if args.no_cpu else cpu_baseline(gen, p, cap, args.cpu_sample, seed)
    log(f"[bench] cpu baseline: {cpu if cpu else cpu_note}")

    import torch  # device memory and streams only

    if not torch.cuda.is_available():
        log("[bench] no GPU visible")
        return 3
    if args.copy_bytes:
        # a plain copy of known size: the HBM counters of its kernel calibrate
        # FETCH_SIZE / WRITE_SIZE for the kernels below
        src = torch.ones(args.copy_bytes // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.copy_bytes} bytes")
        del src, dst
    pos, q = getattr(W, gen)(n, seed=seed)
    # the tree on the host (all CPU threads) and on the GPU: the same tree, array for array
    t0 = time.time()
    host_part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, cap)
    t_host = time.time() - t0
    t0 = time.time()
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, cap, on_device=True)
    t_dev = time.time() - t0
    same_tree = part.digest() == host_part.digest()
    del host_part
    info = part.info()
    log(f"[bench] tree: host {t_host:.2f}s, device {t_dev:.2f}s, identical {same_tree} (depth {info.depth}, "
        f"{info.n_node} nodes, {info.n_leaf} leaves, {info.m2l_pairs} M2L pairs, {info.near_pairs} near pairs)")
    if not same_tree:
        raise SystemExit("device tree differs from the host tree")

    dev = torch.device("cuda:0")
    dq = torch.from_numpy(q).to(dev)
    du = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    tg = np.sort(np.random.default_rng(12345).choice(n, min(args.targets, n), replace=False))
    ref = None

    def rel(a, b):
        return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))

    def run_path(prec: str) -> dict:
        nonlocal ref
        t_a = time.time()
        fmm = api.FMM3D(part, p, ncharge=1, precision=api.MIXED if prec == "mixed" else api.FP64)
        t_b = time.time()
        # the kernels' tile shapes, timed on this plan (outside the timed steps)
        fmm.tune(api.POTGRAD, 1, dq.data_ptr())
        tiles = fmm.tiles()
        log(f"[bench] {prec}: tiles {tiles} (tuned in {time.time() - t_b:.2f}s)")

        def step():
            fmm.evaluate_device(api.POTGRAD, 1, dq.data_ptr(), du.data_ptr())

        for _ in range(args.warmup):
            step()
        fmm.synchronize()
        torch.cuda.synchronize()
        fmm.record_stages(args.steps)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(args.steps):
            step()
        fmm.synchronize()
        torch.cuda.synchronize()
        tb = time.perf_counter()
        ms = (tb - ta) * 1e3 / args.steps
        stages = fmm.recorded_stage_ms(args.steps)
        fmm.record_stages(0)
        mean = {k: float(np.mean([s_[k] for s_ in stages])) for k in api.STAGES}
        log(f"[bench] {prec}: {ms:.2f} ms/step; stages " + ", ".join(f"{k} {v:.2f}" for k, v in mean.items()))
        # accuracy: one more evaluation into a zeroed buffer vs the FP64 direct sum
        du.zero_()
        torch.cuda.synchronize()
        step()
        fmm.synchronize()
        U = du.cpu().numpy()
        if ref is None:
            ref = fmm.direct(api.POTGRAD, q, tg)
        err_phi = rel(U[tg, 0], ref[:, 0])
        err_grad = rel(U[tg, 1:], ref[:, 1:])
        log(f"[bench] {prec}: rel L2 vs direct on {len(tg)} targets: phi {err_phi:.3e} grad {err_grad:.3e}")
        m2l_ms = mean["M2L"]
        # the M2L kernel's matrix-core precision: f16 (split products) for the mixed
        # path at p <= 12, FP32 above, FP64 on the plain path
        if prec == "mixed":
            split = p <= 12
            peak, kernel = (PEAK_F16, "m2l_dense_kernel (f16 split)") if split else (PEAK_FP32, "m2l_pipe_kernel (fp32)")
        else:
            peak, kernel = PEAK_FP64, "m2l_pipe_kernel (fp64)"
            split = False
        # algorithmic flops: the complex full-m count of rotate-translate-rotate per
        # (pair, component); executed: what the kernel's arithmetic needs without padding
        flops_complex = info.m2l_pairs * 4 * m2l_flops_complex(p)
        flops_exec = info.m2l_pairs * 4 * (m2l_flops_split(p) if split else m2l_flops_real(p))
        achieved = flops_complex / (m2l_ms * 1e-3)
        del fmm
        return {
            "points_per_sec": n / (ms * 1e-3), "ms_per_step": ms, "plan_build_s": t_b - t_a, "tiles": tiles,
            "error": {"targets": int(len(tg)), "rel_l2_phi": err_phi, "rel_l2_grad": err_grad,
                      "within_1e-6": bool(max(err_phi, err_grad) <= 1e-6)},
            "stages_ms": mean,
            "roofline": {"kernel": kernel, "bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                         "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                         "flops_per_launch": flops_complex, "avg_launch_ms": m2l_ms,
                         "executed_flops_per_launch": flops_exec,
                         "executed_tflops": flops_exec / (m2l_ms * 1e-3) / 1e12,
                         "executed_frac": flops_exec / (m2l_ms * 1e-3) / peak},
        }

    paths = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    res = {pr: run_path(pr) for pr in paths}
    traffic = traffic_from_profile(args.workload) if not args.n else None
    for pr in paths:
        res[pr]["roofline"]["traffic"] = traffic.get(pr) if traffic else None
    head = res["mixed"] if "mixed" in res else res["fp64"]
    line = {
        "metric": "points/sec (pot+grad) at rel err <=1e-6: mixed vs FP64 GPU, 1 MI355X",
        "value": head["points_per_sec"],
        "unit": "points/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if "mixed" not in res else "f64+f32+f16",
        "data": "synthetic",
        "precision_path": "mixed" if "mixed" in res else "fp64",
        "config": {"workload": args.workload, "generator": gen, "n": n, "p": p, "leaf_capacity": cap,
                   "max_level": 20, "output": "phi+grad", "parallelism": "single GPU"},
        "error": head["error"],
        "stages_ms": head["stages_ms"],
        "roofline": head["roofline"],
        "paths": res,
        "speedup_mixed_vs_fp64": (res["mixed"]["points_per_sec"] / res["fp64"]["points_per_sec"]
                                  if len(res) == 2 else None),
        "tree": {"depth": info.depth, "nodes": info.n_node, "leaves": info.n_leaf, "m2l_pairs": info.m2l_pairs,
                 "near_pairs": info.near_pairs, "tree_build_device_s": t_dev, "tree_build_host_s": t_host,
                 "identical": same_tree},
        "cpu_baseline": cpu if cpu else {"value": None, "note": cpu_note},
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
