"""Benchmark: simulation steps/s of the IPC Newton step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode VARIANT] [--all-variants]
                    [--pcg-tol 1e-6] [--pcg-tol-search 2e-2]

The value is the steps/s of --mode (default "chosen+inexact": the learned
Newton warm start with the FP32 Hessian and FP64 residual refinement, search
directions solved to --pcg-tol-search).  The plain FP64 method (x = x_n start,
FP64 Hessian, every direction solved to --pcg-tol: the reference's method with
PCG for the direct solve) is timed in the same run for the ratio, and so is
the plain method with the inexact search directions.  Every variant uses the
same Newton test, on a direction solved to --pcg-tol (the inexact variants
continue the solve of the direction that passes the test to --pcg-tol and test
again before they accept), and checks every step for intersections.

Workload: the compressed slab, a 50 x 80 x 50-vertex lattice (200,000
vertices, 948,395 tets, E = 1e6, nu = 0.45) between two Dirichlet plates that
close at 0.2 slab heights per second with a 30 degree sinusoidal twist,
dt = 0.005, dHat = 1e-3 bbox diagonal, Newton tolerance 1e-2 dt bbox.  W
warm-up steps (the plates reach the slab) run untimed, then K steps are timed
between two device synchronisations (every step also waits for its own
results on the host: the line search, the Newton test and the collision check).  The position state is resident in
HBM for the whole run.

Printed: one JSON line with steps/s, the roofline figures of the dominant
kernel (HIP-event durations of its launches over --roofline-steps further
steps of the same run, against its algorithmic bytes; the HBM traffic per
launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE, read from
profiles/k_spmv_traffic_compressed_slab.md), the kernel classes (SpMV,
assembly with the FP64 VALU fraction of its per-tet pass, CCD with a probe of
its narrow phase, broad phase), the whole 200-step trajectory of the value's
form and of plain FP64 (full_trajectory), and a CPU baseline: the C oracle
(OpenMP on every host core, a direct sparse solve per Newton iteration) on
two stacked cubes for 50 steps, beside the plain GPU path on the same scene.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TIMER_NAMES = ("spmv", "elastic_tet_pass", "ccd_query", "contact_query", "energy", "contact_assembly",
               "pcg_vector", "grid_build", "assembly_gather", "spmv_residual", "ccd_narrow")
FP64_VALU_PEAK_TFLOPS = 78.6
FP32_VALU_PEAK_TFLOPS = 157.3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(mode):
    """HBM bytes per SpMV launch from the committed rocprofv3 counter summary
    (a line 'traffic_bytes_per_launch: plain=N fp32=M'), or None."""
    import re
    path = os.path.join(ROOT, "profiles", "k_spmv_traffic_compressed_slab.md")
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*plain=(\d+)\s+fp32=(\d+)", txt)
    if not m:
        return None
    return int(m.group(1) if mode == "plain" else m.group(2))


def algorithmic_bytes(sim, precision):
    """Compulsory HBM bytes per launch of the kernel classes."""
    s = sim.setup
    nv, nt = sim.nv, len(s.tets)
    nu, nl = C.c_int(), C.c_int()
    from fastipc_amd import _lib
    _lib.call("fipc_bsr_storage", sim.ctx, C.byref(nu), C.byref(nl))
    nu, nl = nu.value, nl.value
    # FP32 Hessian: FP32 block values and FP32 iteration vectors p, q
    sa = 4 if precision == 1 else 8
    width = (2, 3, 4, 4, 4, 4, 4)
    contact = sum(n * ((3 * w) ** 2 * 8 + 4 * w) for n, w in zip(sim.counts, width))
    return {
        # stored (upper) block values + column index, (block, column) pair per
        # lower block, both row-pointer arrays, p read, q written, contact rows
        # (dense H_r + indices); the transposed re-read of a stored block is
        # not compulsory traffic
        "spmv": nu * (9 * sa + 4) + nl * 8 + 8 * (nv + 1) + 2 * 3 * nv * sa + contact,
        # per tet: 4 indices, D_m^-1, vol, lambda, mu, 10 contribution
        # positions; per vertex x; the stored block values written once (FP32
        # in FP32 mode) and the gradient
        "elastic_assembly": nt * (16 + 72 + 24 + 40) + nv * 24 + nu * 9 * sa + 3 * nv * 8,
    }


# the variants of the step: (FP32 Hessian, warm start: None | "learned" |
# "extrapolation", inexact search directions)
VARIANTS = {
    "plain": (False, None, False),        # the reference's method on the GPU: FP64, x = x_n start
    "fp32": (True, None, False),
    "fp32+extrapolation": (True, "extrapolation", False),
    "chosen": (True, "learned", False),   # learned warm start + FP32 Hessian
    # search directions solved to --pcg-tol-search, the direction that passes
    # the Newton test continued to --pcg-tol before the test is repeated
    "plain+inexact": (False, None, True),
    # the same form with the zero network's start (second-order
    # extrapolation of the correction): what the learned residual adds
    "fp32+extrapolation+inexact": (True, "extrapolation", True),
    "chosen+inexact": (True, "learned", True),
}
WEIGHTS = os.path.join(ROOT, "fastipc_amd", "weights", "ws_slab.npz")


def make_sim(args, variant, sc):
    """The solver of a variant on scene sc: (sim, warm-start hook, network,
    weights metadata)."""
    from fastipc_amd.newton import NewtonIPC
    from fastipc_amd import warmstart as W

    fp32, ws, inexact = VARIANTS[variant]
    sim = NewtonIPC(sc, precision=1 if fp32 else 0, pcg_tol=args.pcg_tol,
                    pcg_tol_search=args.pcg_tol_search if inexact else None,
                    certify=inexact and fp32 and not getattr(args, "no_certify", False))
    hook, net, warm_meta = None, None, None
    if ws == "learned":
        if not os.path.exists(WEIGHTS):
            raise FileNotFoundError(f"{WEIGHTS} is missing (python -m fastipc_amd.warmstart train ...)")
        net = W.WarmStartMLP.load(WEIGHTS)
        warm_meta = {k: net.meta.get(k) for k in ("family", "size", "seeds", "steps", "iters", "rows", "onpolicy")}
        warm_meta["rounds"] = len(net.meta.get("rounds", [])) or 1
        hook = W.LearnedWarmStart(net, W.scale_for(sc))
    elif ws == "extrapolation":
        # the zero network: the start repeats the previous step's correction
        net = W.WarmStartMLP(seed=0)
        net.set_params(np.zeros(net.nparams, np.float32))
        hook = W.LearnedWarmStart(net, W.scale_for(sc))
    return sim, hook, net, warm_meta


def run_gpu(args, variant):
    from fastipc_amd import _lib
    from fastipc_amd import scene as S

    sc = S.compressed_slab(50, 80, 50, steps=args.warmup + args.steps + args.roofline_steps)
    fp32, ws, inexact = VARIANTS[variant]
    precision = 1 if fp32 else 0
    t0 = time.perf_counter()
    sim, hook, net, warm_meta = make_sim(args, variant, sc)
    setup_s = time.perf_counter() - t0
    newton = []
    accepted = 0
    for i in range(args.warmup):
        t1 = time.perf_counter()
        st = sim.step(warm_start=hook)
        log(f"warmup step {i}: {time.perf_counter() - t1:.3f}s newton {st.newton_iters} pcg {st.pcg_iters} "
            f"contacts {st.counts[-1] if st.counts else None}")
    _lib.call("fipc_sync", sim.ctx)
    phases0 = dict(sim.timers.total)
    cert0 = sim.certified_solves
    t0 = time.perf_counter()
    pcg = []
    for i in range(args.steps):
        t1 = time.perf_counter()
        st = sim.step(warm_start=hook)
        log(f"[{variant}] timed step {i}: {time.perf_counter() - t1:.3f}s newton {st.newton_iters} "
            f"pcg {st.pcg_iters} warm start {'kept' if st.warm_accepted else '-'}")
        accepted += int(st.warm_accepted)
        newton.append(st.newton_iters)
        pcg.extend(st.pcg_iters)
        if sim.check_collision():
            raise RuntimeError("intersection after a step")
    _lib.call("fipc_sync", sim.ctx)
    wall = time.perf_counter() - t0
    cert1 = sim.certified_solves
    # host wall time by phase of the timed steps (NewtonIPC's phase timers,
    # two levels deep), ms per step
    phases = {k: round((v - phases0.get(k, 0.0)) / args.steps * 1e3, 3) for k, v in sim.timers.total.items()
              if k.count("/") <= 2 and v > phases0.get(k, 0.0)}
    if args.traffic_probe:
        # counter-calibration run: a plain device copy of known size
        cms = C.c_double()
        _lib.call("fipc_stream_copy", sim.ctx, 256 << 20, C.byref(cms))
        log(f"traffic probe: copied {256 << 20} bytes device to device in {cms.value:.3f} ms "
            f"({2 * (256 << 20) / cms.value / 1e6:.0f} GB/s)")
        sim.close()
        return None
    # kernel durations: HIP events recorded on the context stream around every
    # launch of the timed classes, over further steps of the same run (so that
    # the event records do not perturb the steps/s above)
    _lib.call("fipc_enable_timing", sim.ctx, 1)
    for i in range(args.roofline_steps):
        st = sim.step(warm_start=hook)
        log(f"event-timed step {i}: newton {st.newton_iters} pcg {st.pcg_iters}")
        if sim.check_collision():
            raise RuntimeError("intersection after a step")
    _lib.call("fipc_sync", sim.ctx)
    nt_ = len(TIMER_NAMES)
    ms = np.zeros(nt_)
    la = (C.c_long * nt_)()
    _lib.call("fipc_kernel_times", sim.ctx, nt_, _lib.dptr(ms), la)
    _lib.call("fipc_enable_timing", sim.ctx, 0)
    counts = st.counts[-1] if st.counts else [0] * 7
    kt = {TIMER_NAMES[i]: {"ms_total": float(ms[i]), "launches": int(la[i]),
                           "us_avg": float(ms[i] / la[i] * 1e3) if la[i] else 0.0} for i in range(nt_)}
    # the assembly is the per-tet pass plus the gathers (one launch each per Newton iteration)
    kt["elastic_assembly"] = {"ms_total": kt["elastic_tet_pass"]["ms_total"] + kt["assembly_gather"]["ms_total"],
                              "launches": kt["elastic_tet_pass"]["launches"],
                              "us_avg": kt["elastic_tet_pass"]["us_avg"] + kt["assembly_gather"]["us_avg"]}
    nbytes = algorithmic_bytes(sim, precision)
    # dominant kernel with a byte model: the single kernel with the most time
    # (the SpMV is one kernel; the assembly class is the tet pass plus two
    # gathers, so it competes through its largest kernel, the tet pass)
    single = {"spmv": kt["spmv"]["ms_total"],
              "elastic_assembly": max(kt["elastic_tet_pass"]["ms_total"], kt["assembly_gather"]["ms_total"])}
    dom = max(nbytes, key=lambda k: single[k])
    avg_s = kt[dom]["us_avg"] * 1e-6
    achieved = nbytes[dom] / avg_s / 1e9 if avg_s > 0 else 0.0
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic("fp32" if precision == 1 else "plain") if dom == "spmv" else None,
            "bytes_per_launch": nbytes[dom],
            "avg_launch_us": round(kt[dom]["us_avg"], 2), "launches": kt[dom]["launches"]}
    others = {k: {"achieved_GBs": round(nbytes[k] / (kt[k]["us_avg"] * 1e-6) / 1e9, 1) if kt[k]["us_avg"] else None,
                  "avg_launch_us": round(kt[k]["us_avg"], 2)} for k in nbytes}
    # FP64 VALU fraction of the per-tet pass (SVD, projected dP/dF, 12x12
    # blocks): its FP64 operations per launch from the committed rocprofv3
    # counters over its live event-timed duration
    flops = valu_flops()
    # the FP32 mode's tet pass is the tiled kernel, the FP64 mode's k_elastic
    tet = "k_elastic_tile" if precision == 1 else "k_elastic"
    if flops.get(tet) and kt["elastic_tet_pass"]["us_avg"]:
        # the pass runs FP64 (SVD, projections, gradient) and, in FP32 mode,
        # FP32 (the 12x12 contraction, the sums over a chunk's tets): its VALU
        # fraction is the time each share would take at its own peak, over
        # the measured duration
        dur = kt["elastic_tet_pass"]["us_avg"] * 1e-6
        f = flops[tet]
        f32 = flops.get(tet + "_fp32", 0)
        tf = f / dur / 1e12
        frac = (f / (FP64_VALU_PEAK_TFLOPS * 1e12) + f32 / (FP32_VALU_PEAK_TFLOPS * 1e12)) / dur
        others["elastic_assembly"].update({"tet_pass_fp64_flops_per_launch": f, "tet_pass_fp64_tflops": round(tf, 2),
                                           "tet_pass_fp32_flops_per_launch": f32,
                                           "tet_pass_valu_frac": round(frac, 4)})
    # CCD pass (point-triangle and edge-edge candidate queries on the swept
    # boxes, then the narrow phases: two timed query and two narrow-phase
    # launches per pass); its compulsory unique bytes are x and the direction
    # at the surface vertices, the primitive indices and 16 bytes per
    # candidate pair of the last pass
    pairs = np.zeros(2, np.int32)
    _lib.call("fipc_ccd_counts", sim.ctx, _lib.iptr(pairs))
    s_ = sim.setup
    nbp, nbe, nbt = len(s_.bp), len(s_.be), len(s_.bt)
    ccd_bytes = nbp * 48 + nbp * 4 + nbe * 8 + nbt * 12 + int(pairs.sum()) * 16
    passes = kt["ccd_query"]["launches"] // 2
    if passes:
        pass_us = (kt["ccd_query"]["ms_total"] + kt["ccd_narrow"]["ms_total"]) / passes * 1e3
        others["ccd"] = {"achieved_GBs": round(ccd_bytes / (pass_us * 1e-6) / 1e9, 1), "avg_pass_us": round(pass_us, 2),
                         "hbm_frac": round(ccd_bytes / (pass_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 5),
                         "unique_bytes_per_pass": ccd_bytes, "candidate_pairs_last_pass": [int(p) for p in pairs],
                         "bound": "latency of the grid queries (no pair reaches the narrow phase in these steps)"}
    others["ccd_narrow_phase"] = ccd_narrow_probe(sim, flops)
    others["broad_phase"] = broad_phase_probe(sim)
    res = {
        "variant": variant,
        "warm_start_kept": accepted,
        "certified_solves": cert1 - cert0,
        "warm_start_weights": warm_meta,
        "steps_per_s": args.steps / wall,
        "ms_per_step": wall / args.steps * 1e3,
        "newton_iters_per_step": float(np.mean(newton)),
        "pcg_iters_per_newton": float(np.mean(pcg)) if pcg else 0.0,
        "contacts_last_step": {n: int(c) for n, c in zip(("PP", "PE", "PT", "EE", "EEM", "PPM", "PEM"), counts)},
        "roofline": roof,
        "kernel_classes": others,
        "kernel_times_event_steps": kt,
        "host_phases_ms_per_step": phases,
        "setup_s": setup_s,
        "nv": sim.nv,
        "nt": len(sc.tets),
    }
    sim.close()
    if net is not None:
        net.close()
    return res


def valu_flops():
    """FP64 (and FP32, keyed <kernel>_fp32) operations per launch of the
    VALU-heavy kernels, from the committed rocprofv3 counter summary (lines
    'fp64_flops_per_launch: <kernel>=<N>' and 'fp32_flops_per_launch: ...'):
    64 lanes x (ADD + MUL + TRANS + 2 FMA) VALU instructions."""
    import re
    path = os.path.join(ROOT, "profiles", "valu_compressed_slab.md")
    try:
        txt = open(path).read()
    except OSError:
        return {}
    out = {k: int(v) for k, v in re.findall(r"fp64_flops_per_launch:\s*(\w+)=(\d+)", txt)}
    out.update({k + "_fp32": int(v) for k, v in re.findall(r"fp32_flops_per_launch:\s*(\w+)=(\d+)", txt)})
    return out


def ccd_narrow_probe(sim, flops):
    """The CCD narrow phase with work to do: the benchmark state with a
    seeded smooth direction of amplitude 2 dHat on the free vertices (the
    benchmark's own directions leave no pair below the lower bound), so
    that the pairs near the contact surfaces reach the root finding.  Pairs
    per second, the narrow phase's event-timed duration and its FP64 VALU
    fraction (operations per launch from the committed counters of this same
    probe)."""
    from fastipc_amd import _lib
    rng = np.random.default_rng(2024)
    x = sim.x
    free = ~sim.setup.dirichlet.astype(bool)
    k = rng.normal(size=(3, 3)) * 2.0
    ph = rng.uniform(0, 2 * np.pi, size=3)
    p = np.zeros_like(x)
    p[free] = 2.0 * sim.setup.dHat * np.sin(x[free] @ k + ph)
    sim.set_vector(1, p)
    _lib.call("fipc_enable_timing", sim.ctx, 1)
    reps = 5
    for _ in range(reps):
        alpha = sim.compute_intersection_free_step_size()
    nt_ = len(TIMER_NAMES)
    ms = np.zeros(nt_)
    la = (C.c_long * nt_)()
    _lib.call("fipc_kernel_times", sim.ctx, nt_, _lib.dptr(ms), la)
    _lib.call("fipc_enable_timing", sim.ctx, 0)
    pairs = np.zeros(2, np.int32)
    _lib.call("fipc_ccd_counts", sim.ctx, _lib.iptr(pairs))
    i = TIMER_NAMES.index("ccd_narrow")
    us = ms[i] / la[i] * 1e3 if la[i] else 0.0  # per launch (two launches per pass: PT, EE)
    out = {"pairs_per_pass": [int(v) for v in pairs], "alpha": alpha, "avg_launch_us": round(us, 2),
           "pairs_per_s": round(float(pairs.sum()) / (2 * us * 1e-6), 1) if us else None,
           "direction": "seeded smooth field, amplitude 2 dHat, on the benchmark state"}
    f = flops.get("k_ccd_pt", 0) + flops.get("k_ccd_ee", 0)
    if f and us:
        tf = f / (2 * us * 1e-6) / 1e12
        out.update({"fp64_flops_per_pass": f, "fp64_tflops": round(tf, 3), "valu_frac": round(tf / FP64_VALU_PEAK_TFLOPS, 5)})
    return out


def broad_phase_probe(sim, reps=10):
    """One constraint search (NEWTON_IPC.py:337-357) at the benchmark state
    with fresh grids (the state is rewritten, which invalidates them): host
    wall time per pass including its two synchronisations, against the
    compulsory bytes: the surface positions and primitive indices read, 8
    bytes per broad-phase candidate pair and the rows written."""
    x = sim.x
    ts = []
    for _ in range(reps):
        sim.set_state(x)
        t0 = time.perf_counter()
        counts = sim.find_constraints()
        ts.append(time.perf_counter() - t0)
    # one more pass, untimed, to count the candidates
    sim.search_stats(True)
    sim.set_state(x)
    sim.find_constraints()
    cand = sim.search_counts()
    sim.search_stats(False)
    s_ = sim.setup
    nbp, nbe, nbt = len(s_.bp), len(s_.be), len(s_.bt)
    width = (2, 3, 4, 4, 4, 4, 4)
    nbytes = nbp * (24 + 4) + nbe * 8 + nbt * 12 + 8 * sum(cand) + sum(4 * w * n for w, n in zip(width, counts))
    t = float(np.median(ts))
    gbs = nbytes / t / 1e9
    return {"pass_us": round(t * 1e6, 1), "candidates": cand, "rows": counts, "unique_bytes_per_pass": nbytes,
            "achieved_GBs": round(gbs, 2), "hbm_frac": round(gbs / HBM_PEAK_GBS, 6),
            "bound": "launch and synchronisation latency (about ten launches, two host waits)"}


def run_full_trajectory(args, variant, nsteps=200):
    """The benchmark scene's whole 200-step trajectory (both plates in
    contact from about step 80) with a fresh solver: steps 2..200 timed (the
    first is the warm-up, as the reference's Timer_Print excludes it), every
    step checked for intersections."""
    from fastipc_amd import _lib
    from fastipc_amd import scene as S

    sc = S.compressed_slab(50, 80, 50, steps=nsteps)
    sim, hook, net, _ = make_sim(args, variant, sc)
    sim.step(warm_start=hook)
    if sim.check_collision():
        raise RuntimeError("intersection after a step")
    newton, pcg = [], []
    _lib.call("fipc_sync", sim.ctx)
    t0 = time.perf_counter()
    for _ in range(1, nsteps):
        st = sim.step(warm_start=hook)
        newton.append(st.newton_iters)
        pcg.extend(st.pcg_iters)
        if sim.check_collision():
            raise RuntimeError("intersection after a step")
    _lib.call("fipc_sync", sim.ctx)
    wall = time.perf_counter() - t0
    cert1 = sim.certified_solves
    counts = sim.find_constraints()
    rows = sim.constraints(sort=False)
    nslab = sim.nv - 16
    self_rows = int(sum(np.all(r < nslab, axis=1).sum() for r in rows.values() if len(r)))
    newton = np.asarray(newton)
    out = {"steps_per_s": round((nsteps - 1) / wall, 4), "timed_steps": nsteps - 1,
           "newton_iters_per_step": round(float(newton.mean()), 3),
           "newton_iters_per_step_by_quarter": [round(float(q.mean()), 2) for q in np.array_split(newton, 4)],
           "pcg_iters_per_newton": round(float(np.mean(pcg)), 1),
           "contacts_last_step": dict(zip(("PP", "PE", "PT", "EE", "EEM", "PPM", "PEM"), counts)),
           "slab_slab_rows_last_step": self_rows, "intersections": 0}
    log(f"[{variant}] full trajectory: {json.dumps(out)}")
    sim.close()
    if net is not None:
        net.close()
    return out


def run_cpu_baseline(args):
    """The oracle (strict-IEEE C, OpenMP on every host core, CSR + a direct
    sparse solve per Newton iteration: the reference's structure) on the
    two-stacked-cubes workload: two 10^3-vertex cubes dropped on a ground
    box, 50 steps; steps/s over steps 2..50 (the first
    is excluded, as the reference's Timer_Print does)."""
    from fastipc_amd import scene as S
    from oracle import lib as OL
    from oracle.newton import OracleIPC, SOLVER

    sc = S.config(0)
    o = OracleIPC(sc)
    o.step()
    t0 = time.perf_counter()
    newton = 0
    for _ in range(sc.steps - 1):
        newton += o.step()["newton_iters"]
    wall = time.perf_counter() - t0
    return {"value": round((sc.steps - 1) / wall, 4), "unit": "steps/s", "cores": OL.lib().orc_num_threads(),
            "kind": "port",
            "sample": f"two stacked cubes ({sc.n_verts} verts, {sc.n_tets} tets), {sc.steps} steps, "
                      f"steps 2-{sc.steps} timed ({wall:.1f} s, {newton / (sc.steps - 1):.2f} Newton iterations "
                      f"per step), direct solver {SOLVER}; a different, far smaller scene than the GPU value's"}


def run_gpu_config0(args):
    """The plain FP64 GPU path on the CPU baseline's scene (two stacked cubes),
    counted the same way, for a same-scene comparison."""
    from fastipc_amd import _lib
    from fastipc_amd import scene as S
    from fastipc_amd.newton import NewtonIPC

    sc = S.config(0)
    sim = NewtonIPC(sc, pcg_tol=args.pcg_tol)
    sim.step()
    _lib.call("fipc_sync", sim.ctx)
    t0 = time.perf_counter()
    newton = 0
    for _ in range(sc.steps - 1):
        newton += sim.step().newton_iters
        if sim.check_collision():
            raise RuntimeError("intersection after a step")
    _lib.call("fipc_sync", sim.ctx)
    wall = time.perf_counter() - t0
    cert1 = sim.certified_solves
    sim.close()
    return {"steps_per_s": round((sc.steps - 1) / wall, 3), "newton_iters_per_step": round(newton / (sc.steps - 1), 2),
            "variant": "plain"}


def run_cpu_slab_extrapolation(args):
    """The oracle (strict-IEEE C, OpenMP, CSR + direct solve) on a bounded
    sample: the same slab scene at 10 x 16 x 10 vertices, steps after the
    plates touch, scaled to the full slab linearly in the tet count
    (optimistic for the CPU, whose direct solve grows faster than linearly:
    at 20 x 32 x 20 one SuperLU solve already takes 10-30 s here)."""
    from fastipc_amd import scene as S
    from oracle import lib as OL
    from oracle.newton import OracleIPC, SOLVER

    nx, ny, nz = 10, 16, 10
    sc = S.compressed_slab(nx, ny, nz)
    o = OracleIPC(sc)
    for _ in range(args.cpu_warmup):
        o.step()
    t0 = time.perf_counter()
    k = 0
    while k < args.cpu_steps:
        o.step()
        k += 1
        if time.perf_counter() - t0 > args.cpu_budget_s:
            break
    wall = time.perf_counter() - t0
    sps = k / wall
    full_tets = 49 * 79 * 49 * 5
    return {"value": sps * len(sc.tets) / full_tets, "unit": "steps/s", "cores": OL.lib().orc_num_threads(),
            "kind": "port",
            "sample": f"compressed_slab {nx}x{ny}x{nz} ({sc.n_verts} verts, {len(sc.tets)} tets), {k} steps after "
                      f"{args.cpu_warmup} warm-up steps at {sps:.3f} steps/s, scaled by tets to 948,395; "
                      f"direct solver {SOLVER}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=14)
    ap.add_argument("--mode", choices=tuple(VARIANTS), default="chosen+inexact",
                    help="the variant whose steps/s is the value (default: learned warm start + FP32 Hessian "
                         "+ inexact search directions)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 runs (exact and inexact directions) timed for the ratios")
    ap.add_argument("--all-variants", action="store_true", help="also time every other variant")
    ap.add_argument("--pcg-tol", type=float, default=1e-6)
    ap.add_argument("--pcg-tol-search", type=float, default=2e-2,
                    help="PCG tolerance of the search directions of the +inexact variants")
    ap.add_argument("--no-certify", action="store_true",
                    help="solve the deciding direction of the FP32 +inexact variants to --pcg-tol even when "
                         "its Newton test is certified earlier")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-full-trajectory", action="store_true",
                    help="skip the 200-step runs of the value's form and of plain FP64")
    ap.add_argument("--no-cpu-slab", action="store_true",
                    help="skip the extrapolated CPU rate of the slab (10x16x10 sample scaled by tets)")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-warmup", type=int, default=6)
    ap.add_argument("--cpu-budget-s", type=float, default=15.0)
    ap.add_argument("--roofline-steps", type=int, default=5,
                    help="steps after the timed ones whose kernels are timed with HIP events")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="counter-calibration run for rocprofv3 --pmc: the timed steps, then a 256 MiB "
                         "device copy instead of the timing step; prints no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not an implementation)", file=sys.stderr)
        sys.exit(2)
    g = run_gpu(args, args.mode)
    if g is None:
        return
    others = {}
    extra = [] if args.no_plain else [v for v in ("plain", "plain+inexact", "fp32+extrapolation+inexact")
                                      if v != args.mode]
    if args.all_variants:
        extra += [v for v in VARIANTS if v not in (args.mode, "plain")]
    for v in extra:
        r = run_gpu(args, v)
        others[v] = {"steps_per_s": round(r["steps_per_s"], 4), "ms_per_step": round(r["ms_per_step"], 3),
                     "newton_iters_per_step": round(r["newton_iters_per_step"], 2),
                     "pcg_iters_per_newton": round(r["pcg_iters_per_newton"], 1),
                     "warm_start_kept": r["warm_start_kept"]}
    full = {}
    if not args.no_full_trajectory:
        for v in dict.fromkeys((args.mode, "plain") + (() if args.no_plain else ("fp32+extrapolation+inexact",))):
            full[v] = run_full_trajectory(args, v)
        if "plain" in full and args.mode != "plain":
            full["speedup_vs_plain_fp64"] = round(full[args.mode]["steps_per_s"] / full["plain"]["steps_per_s"], 4)
    cpu = None if args.no_cpu_baseline else run_cpu_baseline(args)
    if cpu is not None:
        cpu["gpu_plain_fp64_same_scene"] = run_gpu_config0(args)
        if not args.no_cpu_slab:
            cpu["slab_extrapolation"] = run_cpu_slab_extrapolation(args)
    fp32, ws, inexact = VARIANTS[args.mode]
    line = {
        "metric": "sim steps/s at 1M tets, warm-start+FP32 vs plain FP64 GPU at equal residual",
        "value": round(g["steps_per_s"], 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(g["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-hessian+f64" if fp32 else "f64",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "compressed_slab 50x80x50 (200,000 verts, 948,395 tets), plates closing 0.2 H/s, "
                               "twist 30 deg, dt 0.005", "variant": args.mode, "warm_start": ws,
                   "warm_start_weights": g["warm_start_weights"], "nv": g["nv"], "nt": g["nt"],
                   "newton_iters_per_step": round(g["newton_iters_per_step"], 2),
                   "pcg_iters_per_newton": round(g["pcg_iters_per_newton"], 1),
                   "warm_start_kept": g["warm_start_kept"],
                   "newton_tol": "1e-2 dt bbox on max|p| (every variant)", "intersections": 0,
                   "pcg_rel_tol": args.pcg_tol,
                   "pcg_rel_tol_search_directions": args.pcg_tol_search if inexact else args.pcg_tol,
                   # the deciding direction stops before pcg_rel_tol once the
                   # exact direction provably passes the Newton test
                   "newton_test_certificate": bool(inexact and fp32 and not args.no_certify),
                   "certified_tests": g.get("certified_solves"),
                   "contacts_last_step": g["contacts_last_step"]},
        "same_run_variants": others,
        "speedup_vs_plain_fp64": round(g["steps_per_s"] / others["plain"]["steps_per_s"], 4) if "plain" in others else None,
        # the same ratio against the plain method given the inexact search
        # directions too: what the warm start and the FP32 Hessian add
        "speedup_vs_plain_fp64_inexact": (round(g["steps_per_s"] / others["plain+inexact"]["steps_per_s"], 4)
                                          if "plain+inexact" in others else None),
        # the whole 200-step trajectory of the benchmark scene (steps 2-200),
        # value's form and plain FP64, fresh solvers in the same run
        "full_trajectory": full or None,
        "roofline": g["roofline"],
        "cpu_baseline": cpu,
        "kernel_classes": g["kernel_classes"],
        "kernel_times": {k: {"launches": v["launches"], "us_avg": round(v["us_avg"], 2)}
                         for k, v in g["kernel_times_event_steps"].items() if v["launches"]},
        "host_phases_ms_per_step": g["host_phases_ms_per_step"],
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
