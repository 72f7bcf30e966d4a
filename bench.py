"""Benchmark: Cahn-Hilliard time steps per second on one MI355X.

Workloads (--workload):
  ch3d_uniform_256 (default): 256^3 uniform periodic root on [0,1]^3 with
      multigrid levels down to 2^3, eps = 0.015, dt = 5e-5,
      phi = 0.1 + 0.05*U(-1,1) (seed 1003), RMS residual below 1e-8.
  ch2d_amr_128_l2: 2D, 128^2 root with two refined levels, eps = 0.01,
      dt = 1e-5, a circle of radius 0.25 plus 0.01 U(-1,1) (seed 1002),
      residual below 1e-9.
  ch3d_amr_64_l3: 3D, 64^3 root with three refined levels (finest
      effective 512^3), eps = 0.008, dt = 1e-5, a sphere of radius 0.3
      perturbed by real spherical harmonics l <= 6 (seed 1004), residual
      below 1e-8.
The AMR meshes are the reference's own regrid result for these initial
data (bsam_amd/workloads/*.json), frozen: regrid is not part of the path.
Every patch is seeded with the same formula at its own level.

Each time step is solved by FAS V(2,2) cycles until the RMS residual is
below the tolerance.  A "step" is one time step of the reference's time loop
(bsamroutines.f90:673-758) after its regrid: SetGhost(root), on the first
step FillDown, CopyQToQold on every level, MultigridIterations.  Inputs are
resident in HBM before timing starts.

Prints one JSON line with the throughput, the roofline of the dominant
kernel (the finest-level colour sweep), and a CPU baseline (the C oracle of
oracle/, one thread, on a bounded sample of the same workload).
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
# V-cycle); its name in kernel traces per precision.
DOMINANT = {"fp64": "k_relax3d_march<double, false, true, 1>",
            "mixed": "k_relax3d_march<float, true, true, 1>"}


def measured_traffic(workload, precision, kernel):
    """HBM bytes per launch of `kernel` from the committed counter digest
    (rocprofv3 FETCH_SIZE / WRITE_SIZE runs, scaled by the copy kernel), or
    None when no digest for this workload and precision is present."""
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
    ap.add_argument("--workload", default="ch3d_uniform_256",
                    choices=["ch3d_uniform_256", "ch2d_amr_128_l2", "ch3d_amr_64_l3"])
    ap.add_argument("--n", type=int, default=256, help="root cells per side of the uniform workload")
    ap.add_argument("--batching", type=int, default=1,
                    help="1: one launch per level (default); 0: one launch per patch")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--check-fp64", type=int, default=1,
                    help="mixed mode: also run the FP64 path and report the max-norm difference")
    ap.add_argument("--kstat-steps", type=int, default=1,
                    help="extra instrumented steps for the kernel roofline (0: skip)")
    return ap.parse_args()


class Workload:
    """Hierarchy, run parameters and initial q of every patch."""

    def __init__(self, name, n):
        import numpy as np
        from bsam_amd import configs, hierarchy
        self.name = name
        if name == "ch3d_uniform_256":
            w = configs.WORKLOADS["ch3d_uniform_256"]
            self.nd, self.n = 3, n
            self.minlevel = -(int(round(np.log2(n))) - 1)
            self.hier = hierarchy.uniform(3, (n,) * 3, self.minlevel, [2] * 6)
            self.eps, self.dt, self.tol, self.steps = w["eps"], w["dt"], w["tol"], w["steps"]
            self.init = dict(mode=1, seed=w["init"]["seed"], base=w["init"]["base"], amp=w["init"]["amp"])
            self.label = f"{n}^3 periodic root, levels 0..{self.minlevel}"
            self.name = f"ch3d_uniform_{n}"
        else:
            w = json.load(open(os.path.join(ROOT, "bsam1.2_amd", "bsam_amd", "workloads", name + ".json")))
            self.nd, self.n, self.minlevel = w["ndims"], w["n"], w["minlevel"]
            base = hierarchy.uniform(self.nd, (self.n,) * self.nd, self.minlevel, w["mthbc"])
            self.hier = hierarchy.from_patch_boxes(base, w["boxes"], self.nd)
            self.eps, self.dt, self.tol, self.steps = w["eps"], w["dt"], w["tol"], w["steps"]
            i = w["init"]
            self.init = dict(mode=i["initmode"], seed=i["seed"], amp=i["amp"], radius=i["radius"],
                             pamp=i.get("pamp", 0.02), lmax=i.get("lmax", 6))
            self.label = (f"{self.n}^{self.nd} periodic root, levels {self.minlevel}..{w['maxlevel']}, "
                          f"patches per level {w['patches_per_level']}, cells per level {w['cells_per_level']}")
        self.root = [k for k, d in enumerate(self.hier) if d["level"] == 0][0]
        self.q0 = {}
        for k, d in enumerate(self.hier):
            if d["level"] < 0:
                continue
            q = np.zeros(hierarchy.field_shape(d, self.nd, 2, "q"))
            q[(0,) + (slice(1, -1),) * self.nd] = configs.init_phi(d, self.nd, eps=self.eps, **self.init)
            self.q0[k] = q
        self.finest = max(d["level"] for d in self.hier)

    def params(self, precision, use_graph, batching):
        from bsam_amd import abi, configs
        prec = abi.PREC_MIXED if precision == "mixed" else abi.PREC_FP64
        return configs.params(self.nd, self.dt, self.eps, self.tol, precision=prec, use_graph=use_graph,
                              batching=batching)

    def load(self, S):
        """Initial fields into solver S (GPU or oracle), ghosts set level by
        level, qold = q."""
        for k, q in self.q0.items():
            S.put(k, "q", q)
        for lv in range(0, self.finest + 1):
            S.set_ghost(lv, 0)
        for lv in range(0, self.finest + 1):
            S.copy_q_to_qold(lv)


def make_state(wl, precision, use_graph, batching):
    from bsam_amd import gpu
    S = gpu.Solver(wl.params(precision, use_graph, batching), wl.hier)
    wl.load(S)
    return S


def cpu_baseline(wl):
    """The second time step of the same workload on the C oracle (one
    thread); the first step, untimed, brings it to the state the timed GPU
    steps start from."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    O = oracle.Oracle(wl.params("fp64", 0, 1), wl.hier)
    wl.load(O)
    O.time_step(True)
    # time steps 2, 3, ... until about 10 s of CPU work or the workload's
    # step count
    t0 = time.perf_counter()
    nsteps = ncyc = 0
    while True:
        rc, nc, r = O.time_step(False)
        nsteps += 1
        ncyc += nc
        dt = time.perf_counter() - t0
        if dt >= 10.0 or nsteps + 1 >= wl.steps:
            break
    return dict(value=nsteps / dt, unit="time-steps/s", cores=1, kind="port",
                sample=f"time steps 2..{nsteps + 1} of the {wl.name} workload ({ncyc} V-cycles, {dt:.2f} s, "
                       f"{1e3 * dt / max(ncyc, 1):.1f} ms per V-cycle) on the C oracle "
                       f"(oracle/bsam_oracle.c, gcc -O2, 1 thread, {_cpu_model()})")


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    from bsam_amd import abi
    wl = Workload(a.workload, a.n)
    S = make_state(wl, a.precision, a.graph, a.batching)
    print(f"[bench] {wl.name} {a.precision} graph={a.graph} batching={a.batching}: warmup {a.warmup}",
          file=sys.stderr, flush=True)
    first = True
    cycles = []
    for _ in range(a.warmup):
        rc, ncyc, r = S.time_step(first)
        first = False
    S.synchronize()
    print(f"[bench] timing {a.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        rc, ncyc, r = S.time_step(first)
        first = False
        cycles.append(ncyc)
        if rc != abi.BSAM_OK:
            print(f"[bench] step did not converge (rc={rc}, residual {r})", file=sys.stderr)
    S.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_step = 1000.0 * elapsed / a.steps
    print(f"[bench] {ms_step:.2f} ms/step, V-cycles {cycles}", file=sys.stderr, flush=True)

    roof = None
    if a.kstat_steps > 0:
        S.kstat_enable(True)
        for _ in range(a.kstat_steps):
            S.time_step(False)
        nl, ms, b = S.kstat_read(abi.KSTAT_RELAX_FINEST)
        S.kstat_enable(False)
        avg_ms = ms / max(nl, 1)
        achieved = b / (avg_ms * 1e-3) / 1e9 if nl else None
        kname = DOMINANT[a.precision] if wl.name.startswith("ch3d_uniform") else \
            ("k_relax_ch<%d, false> / k_relax_lin (finest level, all patches)" % wl.nd)
        roof = dict(bound="hbm", kernel=f"{kname} (finest-level colour sweep)", achieved=achieved,
                    peak=HBM_PEAK_GBS, unit="GB/s", frac=(achieved / HBM_PEAK_GBS) if achieved else None,
                    traffic=measured_traffic(wl.name, a.precision, kname),
                    avg_us=avg_ms * 1e3, launches=nl, bytes_per_launch=b)
    err = None
    if a.precision == "mixed" and a.check_fp64:
        # the FP64 path on the same steps (bit-identical to the oracle in the
        # parity tests) as the reference for the mixed path's error
        import numpy as np
        print("[bench] fp64 reference run for the mixed-precision error", file=sys.stderr, flush=True)
        keep = [k for k, d in enumerate(wl.hier) if d["level"] >= 0]
        qm = {k: S.get(k, "q") for k in keep}
        S.close()
        R = make_state(wl, "fp64", a.graph, a.batching)
        first_r = True
        for _ in range(a.warmup + a.steps + a.kstat_steps):
            R.time_step(first_r)
            first_r = False
        inner = (slice(1, -1),) * wl.nd
        ephi = emu = 0.0
        for k in keep:
            qr = R.get(k, "q")
            ephi = max(ephi, float(np.max(np.abs(qm[k][(0,) + inner] - qr[(0,) + inner]))))
            emu = max(emu, float(np.max(np.abs(qm[k][(1,) + inner] - qr[(1,) + inner]))))
        err = {"phi": ephi, "mu": emu, "steps": a.warmup + a.steps + a.kstat_steps}
        R.close()
    cpu = None
    if a.cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(wl)
    vc = sum(cycles)
    out = {
        "metric": f"AMR CH time-steps/s at ||r||<={wl.tol:g} ({wl.nd}D CH, "
                  + ("plain FP64 path)" if a.precision == "fp64" else "mixed FP32/FP64 path)"),
        "value": a.steps / elapsed,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": wl.name, "mesh": wl.label,
                   "eps": wl.eps, "dt": wl.dt, "tol": wl.tol, "cycle": "FAS V(2,2)",
                   "precision": a.precision, "hipgraph": bool(a.graph),
                   "launches": "one per level" if a.batching else "one per patch"},
        "vcycles_per_step": vc / a.steps,
        "vcycles_per_s": vc / elapsed,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if err is not None:
        out["max_norm_error_vs_fp64"] = err
    print(json.dumps(out))


if __name__ == "__main__":
    main()
