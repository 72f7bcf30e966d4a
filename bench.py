"""Benchmark: AMR Cahn-Hilliard time steps per second on one MI355X, mixed
FP32/FP64 with per-level batched launches against plain FP64.

Workloads (--workload):
  ch3d_amr_64_l3 (default): 3D, 64^3 periodic root with three refined levels
      (finest effective 512^3, 1883 patches of 16^3..64^3, 15.8 M cells),
      eps = 0.008, dt = 1e-5, a sphere of radius 0.3 perturbed by real
      spherical harmonics l <= 6 (seed 1004), RMS residual below 1e-8.
  ch3d_uniform_256: 256^3 uniform periodic root on [0,1]^3 with multigrid
      levels down to 2^3, eps = 0.015, dt = 5e-5, phi = 0.1 + 0.05*U(-1,1)
      (seed 1003), RMS residual below 1e-8.
  ch2d_amr_128_l2: 2D, 128^2 root with two refined levels, eps = 0.01,
      dt = 1e-5, a circle of radius 0.25 plus 0.01 U(-1,1) (seed 1002),
      residual below 1e-9.
  tm3d_amr_64_l3: the 3D tumour-growth system (phi, mu, nutrient n,
      pressure p; proliferation 1.0, apoptosis 0.1, consumption 10), 64^3
      root with three refined levels (435 patches, 4.4 M cells), eps = 0.01,
      dt = 1e-3, a tumour of radius 0.15 perturbed by real spherical
      harmonics l <= 6 (seed 1005), n = 1 and p = 0 on every face, residual
      below 1e-8.
The AMR meshes start as the reference's own regrid result for these
initial data (bsam_amd/workloads/*.json); every patch is seeded with the
same formula at its own level.

A "step" is one time step of the reference's time loop (bsamroutines.f90:
672-758): AMR(rootlevel) -- the regrid, on the device and host
(bsam_time_step_amr: gradient tagging, Berger-Rigoutsos clustering, new
levels, InitFields, TransferValues) with the grid parameters the meshes
were made with --, CopyQToQold on every level, MultigridIterations: FAS
V(2,2) cycles until the RMS residual (ErrorAFAS) is below the tolerance.
`value` times those steps.  The same paths on the workload's mesh held
fixed (no regrid) are reported in `frozen_mesh`.  Inputs are resident in
HBM before timing starts.

--precision both (default) times the mixed path (`value`) and the plain
FP64 path on the same steps, reports their ratio, and measures both
against a FP64 run converged to 1e-12 (max-norm over every level's
interior).  The JSON line also carries the roofline of the dominant kernel
(the finest-level colour sweep of the headline path) and a CPU baseline
(the C oracle of oracle/, one thread, on a bounded sample of the same
workload).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bsam1.2_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


# The finest-level colour sweep is the dominant kernel (8 launches per
# V-cycle); its name in kernel traces per workload kind and precision.
DOMINANT = {("uniform", "fp64"): "k_relax3d_march<double, false, true, 1, 0>",
            ("uniform", "mixed"): "k_relax3d_march<float, true, true, 1, 0>",
            ("amr", "fp64"): "k_relax_ch<3, false, false>",
            ("amr", "mixed"): "k_relax_lin<3, false, 0, false>",
            ("tumour", "fp64"): "k_relax_tm<3, false>",
            ("tumour", "mixed"): "k_relax_lin_tm<3, 0, false>"}


def measured_traffic(workload, precision, kernel):
    """HBM bytes per launch of `kernel` at its largest grid (the finest
    level) from the committed counter digest (rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs, scaled by the copy kernel), or None when no digest for
    this workload and precision is present."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{precision}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        dig = json.load(fh)
    hits = [v for v in dig["kernels"].values() if v["kernel"] == kernel]
    if not hits:
        return None
    return max(hits, key=lambda v: v["grid"])["bytes"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ch3d_amr_64_l3",
                    choices=["ch3d_amr_64_l3", "ch3d_uniform_256", "ch2d_amr_128_l2", "tm3d_amr_64_l3"])
    ap.add_argument("--n", type=int, default=256, help="root cells per side of the uniform workload")
    ap.add_argument("--batching", type=int, default=1,
                    help="1: one launch per level (default); 0: one launch per patch")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--autotune", type=int, default=1,
                    help="time the candidate launch shapes once per hierarchy before the warm-up (bsam_autotune)")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--tagging", type=int, default=1,
                    help="time the regrid tagging entry point on the AMR levels after the timed steps")
    ap.add_argument("--check", type=int, default=1,
                    help="measure the paths' max-norm error against FP64 converged to 1e-12")
    ap.add_argument("--kstat-steps", type=int, default=1,
                    help="extra instrumented steps for the kernel roofline (0: skip)")
    ap.add_argument("--regrid-steps", type=int, default=-1,
                    help="timed time steps with the reference's regrid (AMR every step) per path "
                         "(-1: as many as --steps; 0: frozen mesh only)")
    ap.add_argument("--sweep-images", type=int, default=1,
                    help="1: levels the workload installs get sweep-written ghost images (default); 0: none, "
                         "as the levels a regrid rebuilds (for counter runs of the headline's sweep)")
    ap.add_argument("--regrid-tables", type=int, default=1,
                    help="levels rebuilt by the regrid get flattened ghost tables (1) or keep the staged fill (0)")
    return ap.parse_args()


from bsam_amd.workload import Workload  # noqa: E402


def make_state(wl, precision, use_graph, batching, tol=None):
    from bsam_amd import gpu
    p = wl.params(precision, use_graph, batching)
    if tol is not None:
        p.qerrortol = tol
    S = gpu.Solver(p, wl.hier)
    wl.load(S)
    return S


def timed_path(wl, a, precision):
    """Warm-up, then exactly a.steps timed time steps of one path, then the
    instrumented steps.  Returns (solver, elapsed s, V-cycles per timed
    step, roofline or None)."""
    from bsam_amd import abi
    t_set = time.perf_counter()
    S = make_state(wl, precision, a.graph, a.batching)
    if not a.sweep_images:
        S.set_knob(abi.KNOB_SWEEP_IMAGES, 0, 0)
    t_tune = time.perf_counter()
    tune = S.autotune() if a.autotune else None
    t_end = time.perf_counter()
    if tune is not None:
        # host set-up of the hierarchy (bsam_set_hierarchy: work lists,
        # sibling discovery, flattened ghost tables) with the upload of the
        # initial fields, and the autotuning, in seconds
        tune["setup_s"] = round(t_tune - t_set, 3)
        tune["autotune_s"] = round(t_end - t_tune, 3)
    print(f"[bench] {wl.name} {precision} graph={a.graph} batching={a.batching}: warmup {a.warmup}",
          file=sys.stderr, flush=True)
    first = True
    for _ in range(a.warmup):
        S.time_step(first)
        first = False
    S.synchronize()
    cycles = []
    bad = 0  # timed steps that ended at maxvcycles above the tolerance
    t0 = time.perf_counter()
    for _ in range(a.steps):
        rc, ncyc, r = S.time_step(first)
        first = False
        cycles.append(ncyc)
        if rc != abi.BSAM_OK:
            bad += 1
            print(f"[bench] step did not converge (rc={rc}, residual {r})", file=sys.stderr)
    S.synchronize()
    elapsed = time.perf_counter() - t0
    print(f"[bench] {precision}: {1e3 * elapsed / a.steps:.2f} ms/step, V-cycles {cycles}", file=sys.stderr,
          flush=True)
    roof = None
    if a.kstat_steps > 0:
        # the sweep as the headline's regrid step runs it: the levels a
        # regrid rebuilds have no sweep-written ghost images (same bits)
        S.set_knob(abi.KNOB_SWEEP_IMAGES, 0, 0)
        S.kstat_enable(True)
        for _ in range(a.kstat_steps):
            S.time_step(False)
        nl, ms, b = S.kstat_read(abi.KSTAT_RELAX_FINEST)
        S.kstat_enable(False)
        S.set_knob(abi.KNOB_SWEEP_IMAGES, 0, a.sweep_images)
        avg_ms = ms / max(nl, 1)
        achieved = b / (avg_ms * 1e-3) / 1e9 if nl else None
        kind = "uniform" if wl.name.startswith("ch3d_uniform") else ("tumour" if wl.tumour else "amr")
        kname = DOMINANT.get((kind, precision), "k_relax_ch / k_relax_lin")
        roof = dict(bound="hbm", kernel=f"{kname} (finest-level colour sweep, {precision})", achieved=achieved,
                    peak=HBM_PEAK_GBS, unit="GB/s", frac=(achieved / HBM_PEAK_GBS) if achieved else None,
                    traffic=measured_traffic(wl.name, precision, kname),
                    avg_us=avg_ms * 1e3, launches=nl, bytes_per_launch=b, sweep_images=False)
    return S, elapsed, sum(cycles) / a.steps, roof, tune, bad


# The regrid parameters the AMR workloads' meshes were made with
# (tests/golden/make_amr_workloads.py): the probe module's gradient tagging
# at 0.05, one buffer cell, 8 minimum grid points, fill ratio 0.8.
REGRID = dict(errflagopt=11, qtolerance=0.05, ibuffer=1, minimumgridpoints=8, desiredfillratio=0.8)


def regrid_path(wl, a, precision):
    """The reference's time loop with its regrid (bsamroutines.f90:672-758):
    every step is AMR(rootlevel) on the device (tagging, clustering, new
    levels installed, InitFields, TransferValues), CopyQToQold,
    MultigridIterations (bsam_time_step_amr).  One untimed first step
    (with its FillDown), then a.regrid_steps timed steps; then a few AMR
    passes alone.  Returns a dict for the bench line."""
    from bsam_amd import abi
    try:
        grid = abi.grid_params(wl.finest, **REGRID)
        S = make_state(wl, precision, a.graph, a.batching)
        if a.autotune:
            S.autotune()
        S.set_knob(abi.KNOB_REGRID_TABLES, 0, a.regrid_tables)
        steps = a.steps if a.regrid_steps < 0 else a.regrid_steps
        print(f"[bench] {precision} with regrid: warmup {max(1, a.warmup)}", file=sys.stderr, flush=True)
        for w in range(max(1, a.warmup)):
            S.time_step_amr(grid, w == 0)
        S.synchronize()
        cycles, bad = [], 0
        t0 = time.perf_counter()
        for _ in range(steps):
            rc, n, r = S.time_step_amr(grid, False)
            cycles.append(n)
            bad += rc != abi.BSAM_OK
        S.synchronize()
        el = time.perf_counter() - t0
        t1 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            S.amr(grid)
        S.synchronize()
        amr_s = (time.perf_counter() - t1) / reps
        per = {}
        for d in S.hier:
            if d["level"] >= 0:
                per[str(d["level"])] = per.get(str(d["level"]), 0) + 1
        out = dict(value=steps / el, unit="time-steps/s", steps=steps, elapsed=el,
                   ms_per_step=1e3 * el / steps, amr_ms=1e3 * amr_s,
                   vcycles_per_step=sum(cycles) / len(cycles), nonconverged_steps=bad,
                   patches_per_level=per, grid=dict(REGRID, maxlevel=wl.finest),
                   ghost_tables="flattened" if a.regrid_tables else "staged")
        print(f"[bench] {precision} with regrid: {out['ms_per_step']:.1f} ms/step, AMR pass {out['amr_ms']:.1f} ms, "
              f"mesh {per}", file=sys.stderr, flush=True)
        S.close()
        return out
    except Exception as e:  # a failed measurement must not cost the bench line
        return dict(error=str(e))


def tune_choices(tune):
    """The chosen launch shapes per level of an autotune report."""
    if not tune:
        return None
    return {str(L["level"]): {k[:-7]: v for k, v in L.items() if k.endswith("_choice")} for L in tune["levels"]}


def fields(wl, S):
    return {k: S.get(k, "q") for k, d in enumerate(wl.hier) if d["level"] >= 0}


def max_err(wl, qa, qb):
    import numpy as np
    inner = (slice(1, -1),) * wl.nd
    e = [0.0] * wl.nv
    for k in qa:
        for v in range(wl.nv):
            e[v] = max(e[v], float(np.max(np.abs(qa[k][(v,) + inner] - qb[k][(v,) + inner]))))
    return dict(zip(["phi", "mu", "n", "p"], e))


QTOL_TAG = 1e-2  # qtolerance(0:5) of the reference's CH and BioGrowth griddata.dat


def tagging(wl, X, reps):
    """Regrid tagging (bsam_estimate_level_errors, EstimateLevelErrors) on
    the levels below the finest, as the reference's AMR(level) runs it for
    level < maxlevel after SetGhost(level, 0), from the initial state, for
    errflagopt 10 (the CH configuration's) and 1.  Returns per option the
    mean wall time of the calls over all levels (ms, flags copied to the
    host included) and the cells flagged."""
    try:
        return _tagging(wl, X, reps)
    except Exception as e:  # a failed measurement must not cost the bench line
        return dict(error=str(e))


def _tagging(wl, X, reps):
    out = {}
    levels = list(range(0, wl.finest))
    # errflaguser (10) is the CH/BioGrowth test; the tumour model has only 1
    from bsam_amd import abi
    opts = (1,) if X.params.model == abi.MODEL_TUMOUR else (10, 1)
    for opt in opts:
        wl.load(X)
        total, tagged = 0.0, 0
        for lv in levels:
            n = sum(d["mx"][0] * d["mx"][1] * d["mx"][2] for d in wl.hier if d["level"] == lv)
            X.set_ghost(lv, 0)
            if hasattr(X, "synchronize"):
                X.synchronize()
            X.estimate_level_errors(lv, opt, QTOL_TAG, n)  # untimed first call
            t0 = time.perf_counter()
            for _ in range(reps):
                _, k = X.estimate_level_errors(lv, opt, QTOL_TAG, n)
            total += (time.perf_counter() - t0) / reps
            tagged += k
        out[str(opt)] = dict(ms=round(1e3 * total, 3), tagged=tagged)
    return dict(levels=levels, qtolerance=QTOL_TAG, errflagopt=out)


def frame_write(wl, X):
    """One output frame (WriteQ) of the initial state, written to the null
    device so that the edit and the copies are timed and not the disk: wall
    time (s) and the frame's bytes (cell records of levels >= 0; headers
    are a few hundred bytes per patch)."""
    try:
        wl.load(X)
        if hasattr(X, "synchronize"):
            X.synchronize()
        nb = sum((27 * wl.nv - 1) * (d["mx"][0] + 2) * (d["mx"][1] + 2) * ((d["mx"][2] + 2) if wl.nd == 3 else 1)
                 for d in wl.hier if d["level"] >= 0)
        t0 = time.perf_counter()
        X.write_frame(os.devnull, 0.0)
        dt = time.perf_counter() - t0
        return dict(s=round(dt, 4), bytes=nb, gb_per_s=round(nb / dt / 1e9, 3), target=os.devnull)
    except Exception as e:  # a failed frame must not cost the bench line
        return dict(error=str(e))


# The reference (amdflang -O3, one thread) over the C oracle on the 128^3
# uniform workload in this build's container: 376 ms against 203 ms per
# V-cycle of the second time step.
REF_OVER_ORACLE = 1.85


def cpu_baseline(wl):
    """The C oracle built at -O3 (oracle/liboracle_o3.so, one thread) on the
    same workload: the first whole time step (FillDown, CopyQToQold,
    MultigridIterations to the tolerance) on the frozen mesh, timed.  The
    reference's regrid, which the GPU line includes, is not part of it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_o3.so")
    O = oracle.Oracle(wl.params("fp64", 0, 1), wl.hier)
    wl.load(O)
    t0 = time.perf_counter()
    rc, nc, r = O.time_step(True)
    dt = time.perf_counter() - t0
    tag = frame = None
    if wl.finest > 0:
        O2 = oracle.Oracle(wl.params("fp64", 0, 1), wl.hier)
        tag = tagging(wl, O2, 1)
        frame = frame_write(wl, O2)
    return dict(value=1.0 / dt, unit="time-steps/s", cores=1, kind="port",
                calibration=dict(reference_over_oracle=REF_OVER_ORACLE,
                                 reference_estimate=1.0 / (dt * REF_OVER_ORACLE),
                                 basis="amdflang -O3 reference vs this oracle per V-cycle, 128^3 uniform CH, "
                                       "one thread, in the build container"),
                tagging=tag, frame=frame,
                sample=f"time step 1 of the {wl.name} workload, whole ({nc} V-cycles to the tolerance, "
                       f"{dt:.2f} s, frozen mesh, no regrid), on the C oracle (oracle/bsam_oracle.c, "
                       f"gcc -O3, 1 thread, {_cpu_model()})")


def main():
    import faulthandler
    faulthandler.enable()  # a crash prints the Python stack that led to it
    # and the library's native stack before it (read when the library loads)
    os.environ.setdefault("BSAM_BACKTRACE", "1")
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    wl = Workload(a.workload, a.n)
    paths = ["mixed", "fp64"] if a.precision == "both" else [a.precision]
    res = {}
    for pr in paths:
        S, elapsed, vc, roof, tune, bad = timed_path(wl, a, pr)
        res[pr] = dict(elapsed=elapsed, vc=vc, roof=roof, q=fields(wl, S) if a.check else None, tune=tune,
                       bad=bad)
        S.close()
    head = paths[0]
    regrid = None
    if a.regrid_steps != 0 and wl.finest > 0:
        regrid = {pr: regrid_path(wl, a, pr) for pr in paths}
    check = None
    if a.check:
        # FP64 converged to 1e-12 on the same number of steps: the error
        # scale of stopping at the tolerance, for both paths
        nst = a.warmup + a.steps + a.kstat_steps
        print(f"[bench] FP64 to 1e-12 over {nst} steps for the error check", file=sys.stderr, flush=True)
        R = make_state(wl, "fp64", a.graph, a.batching, tol=1e-12)
        for s in range(nst):
            R.time_step(s == 0)
        qr = fields(wl, R)
        R.close()
        check = {"steps": nst, "reference": "FP64 path converged to RMS residual 1e-12"}
        for pr in paths:
            check[f"{pr}_vs_converged"] = max_err(wl, res[pr]["q"], qr)
        if len(paths) == 2:
            check["mixed_vs_fp64"] = max_err(wl, res["mixed"]["q"], res["fp64"]["q"])
    tag = None
    if a.tagging and wl.finest > 0:
        print("[bench] regrid tagging", file=sys.stderr, flush=True)
        T = make_state(wl, "fp64", 0, a.batching)
        tag = tagging(wl, T, 5)
        tag["frame"] = frame_write(wl, T)
        T.close()
    cpu = None
    if a.cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(wl)
    el = res[head]["elapsed"]
    frozen = {
        "path": {pr: {"value": a.steps / res[pr]["elapsed"], "ms_per_step": 1000.0 * res[pr]["elapsed"] / a.steps,
                      "vcycles_per_step": res[pr]["vc"], "vcycles_per_s": res[pr]["vc"] * a.steps / res[pr]["elapsed"],
                      "nonconverged_steps": res[pr]["bad"], "autotuned": tune_choices(res[pr]["tune"])}
                 for pr in paths},
        "note": "the mesh of the workload held fixed: SetGhost(root..finest), FillDown on the first step, "
                "CopyQToQold, MultigridIterations per step",
    }
    if len(paths) == 2:
        frozen["speedup_vs_fp64"] = res["fp64"]["elapsed"] / el
    # the headline: the reference's time step with its regrid (AMR every
    # step), as the metric counts it; the frozen mesh where no regrid runs
    rg = regrid if regrid and all("value" in regrid[pr] for pr in paths) else None
    if rg:
        value, ms, vc = rg[head]["value"], rg[head]["ms_per_step"], rg[head]["vcycles_per_step"]
        nsteps = rg[head]["steps"]
    else:
        value, ms, vc, nsteps = a.steps / el, 1000.0 * el / a.steps, res[head]["vc"], a.steps
    out = {
        "metric": ("AMR tumour-growth time-steps/s at ||r||<=1e-8: mixed+batched vs plain FP64" if wl.tumour else
                   "AMR CH time-steps/s at ||r||<=1e-8: mixed+batched vs plain FP64; HBM GB/s"
                   if wl.tol == 1e-8 else f"AMR CH time-steps/s at ||r||<={wl.tol:g}: mixed+batched vs plain FP64"),
        "value": value,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": wl.name, "mesh": wl.label,
                   "eps": wl.eps, "dt": wl.dt, "tol": wl.tol, "cycle": "FAS V(2,2)",
                   "path": ("mixed FP32/FP64 defect correction" if head == "mixed" else "plain FP64"),
                   "regrid": ("AMR every step on the device: gradient tagging 0.05, ibuffer 1, minimumgridpoints 8, "
                              "fill ratio 0.8, maxlevel %d; rebuilt levels keep the %s ghost fill"
                              % (wl.finest, "flattened" if a.regrid_tables else "staged")) if rg else
                             "none: the workload's mesh held fixed",
                   # the library runs the loop body eagerly on meshes AMR built
                   "hipgraph": ("frozen mesh only (the regrid step runs eagerly)" if rg and a.graph
                                else bool(a.graph)),
                   "launches": "one per level" if a.batching else "one per patch",
                   "autotuned": tune_choices(res[head]["tune"])},
        "vcycles_per_step": vc,
        "vcycles_per_s": vc * value,
        "roofline": res[head]["roof"],
        "cpu_baseline": cpu,
        "regrid_tagging": tag,
        "with_regrid": regrid,
        "frozen_mesh": frozen,
    }
    if len(paths) == 2:
        f = res["fp64"]
        out["plain_fp64"] = {"value": rg["fp64"]["value"] if rg else a.steps / f["elapsed"],
                             "ms_per_step": rg["fp64"]["ms_per_step"] if rg else 1000.0 * f["elapsed"] / a.steps,
                             "vcycles_per_step": rg["fp64"]["vcycles_per_step"] if rg else f["vc"],
                             "roofline": f["roof"], "autotuned": tune_choices(f["tune"])}
        out["speedup_vs_fp64"] = out["value"] / out["plain_fp64"]["value"]
    # a timed step that stopped at maxvcycles did not reach the metric's
    # tolerance: the line says so and is marked invalid
    out["nonconverged_steps"] = {pr: res[pr]["bad"] + (rg[pr]["nonconverged_steps"] if rg else 0) for pr in paths}
    if any(out["nonconverged_steps"].values()):
        out["valid"] = False
    if check is not None:
        out["max_norm_error"] = check
    if res[head]["tune"]:
        out["autotune_report"] = {pr: res[pr]["tune"] for pr in paths}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
