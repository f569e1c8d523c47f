"""Benchmark: gain-matrix seconds of the symmetric-BEM EEG forward path on one
GPU, mixed precision vs plain FP64.

Workload (default): the 4-layer nested-sphere head with 10242 vertices per
surface (radii 0.80/0.84/0.90/1.0, conductivities 0.33/1.79/0.0042/0.33,
102408 unknowns), 5000 seeded dipoles in r <= 0.72 and 256 electrodes.
A step = the whole gain computation on the device: HeadMat, DipSourceMat
and Head2EEGMat assembly, the HeadMat factorisation (FP32 MFMA for the mixed
path), the adjoint solve with FP64 iterative refinement, and the gain GEMM.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run(args):
    import torch
    import openmeeg_amd as om
    from openmeeg_amd import spheres as S

    torch.cuda.set_device(0)
    meshes, cond, dip, el = S.make_config(args.config)
    geom = om.make_nested_geometry(meshes, cond)
    n = geom.size
    log("bench: config %s n=%d dipoles=%d electrodes=%d" % (args.config, n, len(dip), len(el)))

    def step(precision):
        G, info = om.gain_eeg_pipeline(geom, dip, el, precision=precision, domain_name="Brain")
        return G, info

    for w in range(args.warmup):
        t0 = time.time()
        G, info = step(args.precision)
        del G
        log("bench: warmup %d done in %.2f s" % (w, time.time() - t0))
    torch.cuda.synchronize()
    t0 = time.time()
    infos = []
    Gm = None
    for s in range(args.steps):
        Gm, info = step(args.precision)
        infos.append(info)
        log("bench: step %d stages %s" % (s, {k: round(v, 4) for k, v in info["times"].items()}))
    torch.cuda.synchronize()
    dt = (time.time() - t0) / max(1, args.steps)
    Gm_h = om.gain_to_numpy(Gm, len(el), len(dip))
    del Gm
    torch.cuda.empty_cache()

    # plain FP64 path on the same inputs (outside the timed region)
    err64 = None
    t64 = None
    if args.precision != "fp64":
        torch.cuda.synchronize()
        t1 = time.time()
        G64, info64 = step("fp64")
        torch.cuda.synchronize()
        t64 = time.time() - t1
        G64h = om.gain_to_numpy(G64, len(el), len(dip))
        del G64
        torch.cuda.empty_cache()
        err64 = float(np.linalg.norm(Gm_h - G64h) / np.linalg.norm(G64h))
        log("bench: fp64 path %.3f s, stages %s, mixed vs fp64 rel err %.3e"
            % (t64, {k: round(v, 4) for k, v in info64["times"].items()}, err64))

    stages = infos[-1]["times"]
    fact_s = stages["factor"]
    flops = n ** 3 / 3.0
    dtype_f = "f32" if args.precision == "mixed" else "f64"
    achieved = flops / fact_s / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_TFLOPS[dtype_f],
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_TFLOPS[dtype_f], 4), "traffic": None,
                "kernel": "signed-Cholesky factorisation (%s MFMA), n^3/3 flop" % dtype_f}
    out = {
        "metric": "gain-matrix seconds, mixed vs FP64 on 1 GPU at <=1e-6 rel err vs FP64",
        "value": round(dt, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mfma-factor+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "%s nested-sphere head, n=%d unknowns, %d dipoles, %d electrodes"
                   % (args.config, n, len(dip), len(el)), "precision": args.precision},
        "fp64_seconds": None if t64 is None else round(t64, 4),
        "rel_err_vs_fp64": err64,
        "stages_s": {k: round(v, 4) for k, v in stages.items()},
        "refine_iterations": infos[-1]["refine_iterations"],
        "roofline": roofline,
        "cpu_baseline": None,
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    out = run(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
