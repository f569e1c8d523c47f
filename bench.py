"""Benchmark: gain-matrix seconds of the symmetric-BEM EEG forward path on one
GPU, mixed precision vs plain FP64.

Workload (default): the 4-layer nested-sphere head with 10242 vertices per
surface (radii 0.80/0.84/0.90/1.0, conductivities 0.33/1.79/0.0042/0.33,
102408 unknowns), 5000 seeded dipoles in r <= 0.72 and 256 electrodes.
A step = the whole gain computation on the device: HeadMat, DipSourceMat
and Head2EEGMat assembly, the HeadMat factorisation (split-FP16 MFMA trailing
updates for the default mixed path, FP32 MFMA with --precision mixed, FP64
with --precision fp64), the adjoint solve with FP64 iterative refinement,
and the gain GEMM.

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
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "f16": 2500.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(config, dtype_f):
    """HBM bytes per launch of a trailing-update kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "factor_update_%s_%s_traffic.md" % (dtype_f, config))
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("traffic_bytes_per_launch:"):
            return float(ln.split(":")[1])
    return None


def roofline_of(precision, kern, config):
    """Roofline entry of a path's dominant kernel, the trailing update of the
    signed Cholesky, from the event timers: achieved = algorithmic flop of
    the update (2 x rank per reached lower-triangle output element, the
    FP32-equivalent flop for the split-FP16 form) / launch time.  The split
    form issues 3 FP16 MFMA products per algorithmic product; its issued
    rate is reported beside it (mfma_issued_*), not as `achieved`."""
    if precision == "mixed16":
        dom, peak, mult, tkey = "factor_update_f32s", PEAK_TFLOPS["f16"], 3.0, "f32s"
        kname = "k_syrk_split (split-FP16 trailing update of the signed Cholesky)"
    elif precision == "mixed":
        dom, peak, mult, tkey = "factor_update_f32_t128", PEAK_TFLOPS["f32"], 1.0, "f32"
        kname = "k_gemm<float,false,true,true,4,false> (interior tiles of the FP32 trailing update)"
    else:
        dom, peak, mult, tkey = "factor_update_f64_t128", PEAK_TFLOPS["f64"], 1.0, "f64"
        kname = "k_gemm_f64_dma<true, false> (interior 128 x 128 tiles of the FP64 trailing update)"
    calls, sec, work = kern[dom]
    achieved = work / sec / 1e12 if sec > 0 else 0.0
    out = {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
           "frac": round(achieved / peak, 4), "traffic": measured_traffic(config, tkey),
           "kernel": kname, "launches": calls, "avg_launch_ms": round(sec / max(calls, 1) * 1e3, 4),
           "algorithmic_flop_per_launch": work / max(calls, 1)}
    if mult != 1.0:
        out["mfma_issued_tflops"] = round(mult * achieved, 2)
        out["mfma_issued_frac"] = round(mult * achieved / peak, 4)
    return out


def batched_extra(om, S, torch, reps=3):
    """32 uneven perturbed heads (642..2562 vertices per surface, 200..2000
    dipoles, 64..256 electrodes): the single-head mixed16 pipeline head after
    head vs one batch (tuned, eager, then graph replays)."""
    hs = S.make_batched_heads(n_heads=32, seed=4)
    heads = [(om.make_nested_geometry(m, c), d, e) for m, c, d, e in hs]
    sizes = [g.size for g, _, _ in heads]
    ref = []
    for rep in range(2):   # the second round is timed (device tables built in the first)
        torch.cuda.synchronize()
        t0 = time.time()
        for g, d, e in heads:
            G, _ = om.gain_eeg_pipeline(g, d, e, precision="mixed16")
            if rep == 1:
                ref.append(om.gain_to_numpy(G, len(e), len(d)))
            del G
        torch.cuda.synchronize()
        t_seq = time.time() - t0
    log("bench: batched heads, sequential pipeline %.3f s" % t_seq)
    b = om.BatchedGain(heads, precision="mixed")
    tune = b.tune()
    torch.cuda.synchronize()
    t0 = time.time()
    b.run(graph=False)
    torch.cuda.synchronize()
    t_eager = time.time() - t0
    t0 = time.time()
    b.run(graph=True)      # captures, then replays
    torch.cuda.synchronize()
    t_capture = time.time() - t0
    t0 = time.time()
    for _ in range(reps):
        b.run(graph=True)
    torch.cuda.synchronize()
    t_graph = (time.time() - t0) / reps
    err = max(float(np.linalg.norm(b.gain(i) - ref[i]) / np.linalg.norm(ref[i])) for i in range(len(heads)))
    cfg = b.config()
    log("bench: batched eager %.3f s, first graph pass %.3f s, graph replay %.3f s, max rel diff vs pipeline %.2e, %s"
        % (t_eager, t_capture, t_graph, err, cfg))
    del b
    return {"heads": len(heads), "unknowns_min_max": [min(sizes), max(sizes)], "precision": "mixed",
            "sequential_pipeline_s": round(t_seq, 4), "batch_eager_s": round(t_eager, 4),
            "batch_graph_replay_s": round(t_graph, 4), "speedup_vs_sequential": round(t_seq / t_graph, 2),
            "max_rel_diff_vs_pipeline": err, "config": cfg,
            "tuning": [{"tile": t, "grid_cap": c, "split16": sp, "s": round(x, 6)} for t, c, sp, x in tune]}


def gradient_extra(om, S, torch):
    """Least-squares dipole fit on the 2562-vertex 3-layer head, 4096 dipoles,
    128 electrodes, SNR-10 targets: one forward and one backward pass."""
    meshes, cond, dip, el = S.make_config("c3")
    g = om.make_nested_geometry(meshes, cond)
    model = om.EEGForward(g, el, precision="mixed")
    rng = np.random.default_rng(3)
    with torch.no_grad():
        Gt = model.gain(torch.tensor(dip[:, :3]), torch.tensor(dip[:, 3:]))
    T = torch.tensor(S.noisy_targets(Gt.cpu().numpy(), rng, 10.0), device="cuda")
    out = {}
    for rep in range(2):
        p = torch.tensor(dip[:, :3] * 0.99, requires_grad=True)
        q = torch.tensor(dip[:, 3:], requires_grad=True)
        s = torch.tensor(model.conductivities(), requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.time()
        G = model.gain(p, q, s)
        L = 0.5 * ((G - T) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.time()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.time()
        out = {"forward_s": round(t1 - t0, 4), "backward_s": round(t2 - t1, 4), "n": g.size,
               "dipoles": len(dip), "electrodes": len(el), "precision": "mixed"}
    log("bench: c3 gradient %s" % out)
    return out


def meg_extra(om, S, torch, eeg_seconds=None):
    """EEG + MEG forward on the benchmark head (c2): one factor and one solve
    with the 256 electrode rows and 300 tilted magnetometer rows
    (GainEEGMEGadjoint), mixed16, against the same pipeline in FP64."""
    meshes, cond, dip, el = S.make_config("c2")
    g = om.make_nested_geometry(meshes, cond)
    p, o, w = S.sphere_squids(np.random.default_rng(5), 300)
    sq = om.Sensors(p, orientations=o, weights=w)
    ns, nd = 300, len(dip)
    Ge, Gm, info = om.gain_eeg_meg_pipeline(g, dip, el, sq, precision="mixed16", domain_name="Brain")   # warm
    del Ge, Gm
    torch.cuda.synchronize()
    t0 = time.time()
    Ge, Gm, info = om.gain_eeg_meg_pipeline(g, dip, el, sq, precision="mixed16", domain_name="Brain")
    torch.cuda.synchronize()
    t = time.time() - t0
    Gm = om.gain_to_numpy(Gm, ns, nd)
    del Ge
    torch.cuda.empty_cache()
    _, G64, _ = om.gain_eeg_meg_pipeline(g, dip, el, sq, precision="fp64", domain_name="Brain")
    G64 = om.gain_to_numpy(G64, ns, nd)
    err = float(np.linalg.norm(Gm - G64) / np.linalg.norm(G64))
    out = {"seconds": round(t, 4), "electrodes": len(el), "magnetometers": ns, "dipoles": nd, "n": g.size,
           "stages_s": {k: round(v, 4) for k, v in info["times"].items()}, "refine_iterations":
           info["refine_iterations"], "meg_rel_err_vs_fp64": err}
    if eeg_seconds:
        out["added_to_eeg_only_s"] = round(t - eeg_seconds, 4)
    log("bench: c2 EEG+MEG %s" % out)
    return out


def surf_extra(om, S, torch, eeg_seconds=None, cpu=True):
    """Surface sources on the benchmark head (c2): SurfSourceMat of a
    10242-vertex source sphere (radius 0.7, inside the brain) and the
    surface-source EEG gain pipeline (256 electrodes), mixed16 against the
    same pipeline in FP64.  CPU: the oracle's SurfSourceMat on a 642-vertex
    source sphere against the same head, scaled by the source triangle count
    (the S, N and D work is proportional to it)."""
    from openmeeg_amd import _lib
    from openmeeg_amd.api import _ssm_device
    L = _lib.lib()
    meshes, cond, dip, el = S.make_config("c2")
    g = om.make_nested_geometry(meshes, cond)
    P, T = S.geodesic_sphere(32)
    mesh = om.Mesh(0.7 * P, T)
    ne, nsrc = len(el), len(P)
    D = _ssm_device(g, mesh)                                        # warm
    del D
    torch.cuda.synchronize()
    t0 = time.time()
    D = _ssm_device(g, mesh)
    torch.cuda.synchronize()
    t_ssm = time.time() - t0
    del D
    # the same call again with the per-kernel event timers on
    _lib.check(L.omh_profile_enable(1))
    D = _ssm_device(g, mesh)
    torch.cuda.synchronize()
    del D
    kern = {}
    for k in ("s_block", "d_pairs"):
        import ctypes as C
        calls, sec, work = C.c_int64(), C.c_double(), C.c_double()
        _lib.check(L.omh_profile_read(k.encode(), C.byref(calls), C.byref(sec), C.byref(work)))
        if calls.value:
            kern[k] = {"launches": calls.value, "seconds": round(sec.value, 4),
                       "gauss_evaluations_per_s": work.value / sec.value}
    _lib.check(L.omh_profile_enable(0))
    Ge, _, info = om.gain_eeg_surf_pipeline(g, mesh, el, precision="mixed16")   # warm
    del Ge
    torch.cuda.synchronize()
    t0 = time.time()
    Ge, _, info = om.gain_eeg_surf_pipeline(g, mesh, el, precision="mixed16")
    torch.cuda.synchronize()
    t = time.time() - t0
    Ge = om.gain_to_numpy(Ge, ne, nsrc)
    torch.cuda.empty_cache()
    G64, _, info64 = om.gain_eeg_surf_pipeline(g, mesh, el, precision="fp64")
    G64 = om.gain_to_numpy(G64, ne, nsrc)
    err = float(np.linalg.norm(Ge - G64) / np.linalg.norm(G64))
    out = {"ssm_seconds": round(t_ssm, 4), "ssm_kernels": kern, "pipeline_seconds": round(t, 4),
           "source_vertices": nsrc, "source_triangles": len(T), "electrodes": ne, "n": g.size,
           "stages_s": {k: round(v, 4) for k, v in info["times"].items()},
           "fp64_ssm_seconds": round(info64["times"]["ssm"], 4),
           "refine_iterations": info["refine_iterations"], "rel_err_vs_fp64": err}
    if cpu:
        from oracle import oracle as O
        og = O.nested_geometry(meshes, cond)
        Ps, Ts = S.geodesic_sphere(8)
        st = {}
        O.surf_source_mat(og, 0.7 * Ps, Ts, stats=st)
        wall = st["pair_seconds"]
        out["cpu_ssm"] = {"value": round(wall * len(T) / len(Ts), 1), "unit": "s", "cores": 1, "kind": "port",
                          "sample": "oracle SurfSourceMat (its C pair assembly, S N D, without the Python table "
                                    "set-up) with a %d-vertex source sphere against the same head (%.2f s), scaled "
                                    "by the source triangle count %d/%d" % (len(Ps), wall, len(T), len(Ts))}
    log("bench: c2 surface sources %s" % out)
    return out


def internal_extra(om, S, torch, eeg_seconds=None, cpu=True):
    """Internal potentials on the benchmark head (c2): Surf2VolMat for 20000
    points inside the brain (radius < 0.75) and the internal-potential gain
    of the 5000 benchmark dipoles at 1000 of them through one factor and one
    solve (gain_internal_pot_pipeline), mixed16 against FP64."""
    from openmeeg_amd.api import _surf2vol_device
    meshes, cond, dip, el = S.make_config("c2")
    g = om.make_nested_geometry(meshes, cond)
    rng = np.random.default_rng(9)
    d = rng.standard_normal((20000, 3))
    pts = d * (0.75 * rng.random(20000) ** (1.0 / 3.0) / np.linalg.norm(d, axis=1))[:, None]
    M, kept = _surf2vol_device(g, pts)                                   # warm
    del M
    torch.cuda.synchronize()
    t0 = time.time()
    M, kept = _surf2vol_device(g, pts)
    torch.cuda.synchronize()
    t_h2ip = time.time() - t0
    del M
    sub = pts[:1000]
    G, kp, info = om.gain_internal_pot_pipeline(g, dip, sub, precision="mixed16", domain_name="Brain")   # warm
    del G
    torch.cuda.synchronize()
    t0 = time.time()
    G, kp, info = om.gain_internal_pot_pipeline(g, dip, sub, precision="mixed16", domain_name="Brain")
    torch.cuda.synchronize()
    t = time.time() - t0
    G = om.gain_to_numpy(G, len(kp), len(dip))
    torch.cuda.empty_cache()
    G64, _, _ = om.gain_internal_pot_pipeline(g, dip, sub, precision="fp64", domain_name="Brain")
    G64 = om.gain_to_numpy(G64, len(kp), len(dip))
    err = float(np.linalg.norm(G - G64) / np.linalg.norm(G64))
    nb = len(meshes[0][1])
    out = {"surf2vol_seconds": round(t_h2ip, 4), "points": len(kept), "n": g.size,
           "surf2vol_point_triangle_pairs_per_s": len(kept) * float(nb) / t_h2ip,
           "pipeline_seconds": round(t, 4), "pipeline_points": len(kp), "dipoles": len(dip),
           "refine_iterations": info["refine_iterations"], "rel_err_vs_fp64": err}
    if cpu:
        from oracle import oracle as O
        og = O.nested_geometry(meshes, cond)
        O.surf2vol(og, pts[:1])                         # tables and point location set up
        t0 = time.time()
        O.surf2vol(og, pts[:200])
        wall = time.time() - t0
        out["cpu_surf2vol"] = {"value": round(wall * len(kept) / 200.0, 1), "unit": "s", "cores": 1, "kind": "port",
                               "sample": "oracle Surf2VolMat for 200 of the points (%.2f s: point location, then the "
                                         "P1 double layer and OperatorS per triangle in C), scaled by the point "
                                         "count %d/200" % (wall, len(kept))}
    log("bench: c2 internal potentials %s" % out)
    return out


def c1_extra(om, S, torch, cpu=True):
    """The 2562-vertex three-sphere head (n = 17926, 2000 dipoles, 128
    electrodes) end to end: the device gain in FP64 and mixed16, the device
    HeadMat inverse (SymMatrix::inverse, om_minverser's computation), and
    the oracle's whole pipeline measured on this host's cores (HeadMat,
    DipSourceMat, Head2EEGMat, LAPACK dsytrf + dsytrs, gain)."""
    meshes, cond, dip, el = S.make_config("c1")
    g = om.make_nested_geometry(meshes, cond)
    ne, nd = len(el), len(dip)
    out = {"n": g.size, "dipoles": nd, "electrodes": ne}
    G64 = None
    for prec in ("fp64", "mixed16"):
        for rep in range(2):
            torch.cuda.synchronize()
            t0 = time.time()
            G, info = om.gain_eeg_pipeline(g, dip, el, precision=prec, domain_name="Brain")
            torch.cuda.synchronize()
            t = time.time() - t0
        out["gpu_%s_s" % prec] = round(t, 4)
        out["gpu_%s_stages_s" % prec] = {k: round(v, 4) for k, v in info["times"].items()}
        Gh = om.gain_to_numpy(G, ne, nd)
        del G
        if prec == "fp64":
            G64 = Gh
        else:
            out["gpu_mixed16_rel_err_vs_fp64"] = float(np.linalg.norm(Gh - G64) / np.linalg.norm(G64))
    from openmeeg_amd.api import _head_mat_device, SymMatrix
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.time()
        hm = SymMatrix._from_device(_head_mat_device(g), g.size, g.n_vertex_unknowns)
        torch.cuda.synchronize()
        t1 = time.time()
        inv = hm.inverse()
        torch.cuda.synchronize()
        t2 = time.time()
        del hm, inv
        torch.cuda.empty_cache()
    out["gpu_headmat_s"] = round(t1 - t0, 4)
    out["gpu_inverse_s"] = round(t2 - t1, 4)
    if cpu:
        from oracle import oracle as O
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t0 = time.time()
        Gc, st, og = O.time_pipeline(meshes, cond, dip, el, "Brain")
        wall = time.time() - t0
        out["cpu_pipeline"] = {"value": round(sum(st.values()), 2), "unit": "s", "cores": cores, "kind": "port",
                               "stages_s": {k: round(v, 3) for k, v in st.items()}, "wall_s": round(wall, 2),
                               "rel_err_vs_gpu_fp64": float(np.linalg.norm(Gc - G64) / np.linalg.norm(Gc))}
        out["_cpu_stages"] = (st, og, nd, ne)
    log("bench: c1 %s" % {k: v for k, v in out.items() if not k.startswith("_")})
    return out


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

    from openmeeg_amd import _lib
    L = _lib.lib()

    def prof(name):
        import ctypes as C
        calls, sec, work = C.c_int64(), C.c_double(), C.c_double()
        _lib.check(L.omh_profile_read(name.encode(), C.byref(calls), C.byref(sec), C.byref(work)))
        return calls.value, sec.value, work.value

    for w in range(args.warmup):
        t0 = time.time()
        G, info = step(args.precision)
        del G
        log("bench: warmup %d done in %.2f s" % (w, time.time() - t0))
    torch.cuda.synchronize()
    _lib.check(L.omh_profile_enable(1))
    t0 = time.time()
    infos = []
    Gm = None
    for s in range(args.steps):
        Gm, info = step(args.precision)
        infos.append(info)
        log("bench: step %d stages %s" % (s, {k: round(v, 4) for k, v in info["times"].items()}))
    torch.cuda.synchronize()
    dt = (time.time() - t0) / max(1, args.steps)
    kern = {k: prof(k) for k in ("factor_update_f32s", "factor_update_f32_t128", "factor_update_f32_t64",
                                  "factor_update_f64_t128",
                                  "diag_factor_f32", "diag_factor_f64", "s_block", "d_pairs", "pairs_far")}
    _lib.check(L.omh_profile_enable(0))
    Gm_h = om.gain_to_numpy(Gm, len(el), len(dip))
    del Gm
    torch.cuda.empty_cache()

    # plain FP64 path on the same inputs (outside the timed region)
    err64 = None
    t64 = None
    stages64 = None
    roofline64 = None
    if args.precision != "fp64" and not args.skip_fp64:
        torch.cuda.synchronize()
        _lib.check(L.omh_profile_enable(1))
        t1 = time.time()
        G64, info64 = step("fp64")
        torch.cuda.synchronize()
        t64 = time.time() - t1
        roofline64 = roofline_of("fp64", {"factor_update_f64_t128": prof("factor_update_f64_t128")}, args.config)
        _lib.check(L.omh_profile_enable(0))
        stages64 = info64["times"]
        G64h = om.gain_to_numpy(G64, len(el), len(dip))
        del G64
        torch.cuda.empty_cache()
        err64 = float(np.linalg.norm(Gm_h - G64h) / np.linalg.norm(G64h))
        log("bench: fp64 path %.3f s, stages %s, mixed vs fp64 rel err %.3e"
            % (t64, {k: round(v, 4) for k, v in stages64.items()}, err64))

    stages = infos[-1]["times"]
    roofline = roofline_of(args.precision, kern, args.config)
    others = {}
    for k, (c, sc, wk) in kern.items():
        if c:
            others[k] = {"launches": c, "seconds": round(sc / args.steps, 4), "work_per_s": wk / sc}
    extras = {}
    if not args.no_extras:
        _lib.check(L.omh_release_scratch())
        torch.cuda.empty_cache()
        extras["batched_32_heads"] = batched_extra(om, S, torch)
        torch.cuda.empty_cache()
        extras["gradient_c3"] = gradient_extra(om, S, torch)
        _lib.check(L.omh_release_scratch())
        torch.cuda.empty_cache()
        extras["eeg_meg_c2"] = meg_extra(om, S, torch, dt)
        _lib.check(L.omh_release_scratch())
        torch.cuda.empty_cache()
        extras["surface_sources_c2"] = surf_extra(om, S, torch, cpu=not args.no_cpu_baseline)
        _lib.check(L.omh_release_scratch())
        torch.cuda.empty_cache()
        extras["internal_potential_c2"] = internal_extra(om, S, torch, cpu=not args.no_cpu_baseline)
        _lib.check(L.omh_release_scratch())
        torch.cuda.empty_cache()
    cpu = None
    if not args.no_c1:
        _lib.check(L.omh_release_scratch())
        torch.cuda.empty_cache()
        c1 = c1_extra(om, S, torch, cpu=not args.no_cpu_baseline)
        cpu_c1 = c1.pop("_cpu_stages", None)
        extras["c1_end_to_end"] = c1
        if cpu_c1 is not None and args.config != "c1":
            from oracle import oracle as O
            st, og1, nd1, ne1 = cpu_c1
            og = O.nested_geometry(meshes, cond)
            est = O.extrapolate_stages(st, og1, nd1, ne1, og, len(dip), len(el))
            cpu = {"value": round(sum(est.values()), 1), "unit": "s", "cores": c1["cpu_pipeline"]["cores"],
                   "kind": "port",
                   "sample": "the oracle's whole pipeline measured on the n=%d three-sphere head (2000 dipoles, "
                             "128 electrodes: %.1f s, stages in extras.c1_end_to_end.cpu_pipeline) and scaled to "
                             "this workload stage by stage (HeadMat by S+D pair count, DipSourceMat by dipoles x "
                             "source-mesh triangles, factor by n^3, solve by n^2 x electrodes); an extrapolation. "
                             "The oracle factors with LAPACK's blocked dsytrf, the reference's om_minverser with "
                             "the packed, unblocked DSPTRF + DSPTRI, which is many times slower"
                             % (og1.size, c1["cpu_pipeline"]["value"]),
                   "stages_s": {k: round(v, 1) for k, v in est.items()}}
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
        "dtype": {"mixed16": "f16x3-split-mfma-factor+f64-refinement", "mixed": "f32-mfma-factor+f64-refinement",
                  "fp64": "f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": "%s nested-sphere head, n=%d unknowns, %d dipoles, %d electrodes"
                   % (args.config, n, len(dip), len(el)), "precision": args.precision},
        "fp64_seconds": None if t64 is None else round(t64, 4),
        "rel_err_vs_fp64": err64,
        "stages_s": {k: round(v, 4) for k, v in stages.items()},
        "fp64_stages_s": None if stages64 is None else {k: round(v, 4) for k, v in stages64.items()},
        "refine_iterations": infos[-1]["refine_iterations"],
        "roofline": roofline,
        "roofline_fp64": roofline64,
        "kernels": others,
        "cpu_baseline": cpu,
        "extras": extras or None,
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--precision", default="mixed16", choices=["mixed16", "mixed", "fp64"],
                    help="mixed16: split-FP16 MFMA factor + FP64 refinement (default); mixed: FP32 MFMA factor")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true", help="do not run the plain FP64 comparison")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device copy of this many bytes (counter calibration) and exit")
    ap.add_argument("--workload", default="gain", choices=["gain", "batched", "gradient", "meg", "surf", "internal"],
                    help="gain: the headline measurement (default); batched / gradient / meg / surf / internal: "
                         "only that extra")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-heads, gradient, EEG+MEG, surface-source and internal-potential "
                         "measurements")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the 2562-vertex head end to end (device gains and inverse, oracle pipeline on the "
                         "host), which also provides the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        import torch
        n = args.copy_probe // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8, "launches": 3}), flush=True)
        return
    if args.workload != "gain":
        # one of the extra workloads alone (profiling); not the headline line
        import torch
        import openmeeg_amd as om
        from openmeeg_amd import spheres as S
        torch.cuda.set_device(0)
        fn = {"batched": batched_extra, "gradient": gradient_extra, "meg": meg_extra,
              "surf": surf_extra, "internal": internal_extra}[args.workload]
        print(json.dumps({"workload": args.workload, "result": fn(om, S, torch)}), flush=True)
        return
    out = run(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
