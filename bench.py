"""Benchmark: 3D Cahn-Hilliard time steps per second on one MI355X.

Workload (default): 256^3 uniform periodic root on [0,1]^3 with multigrid
levels down to 2^3, eps = 0.015, dt = 5e-5, phi = 0.1 + 0.05*U(-1,1)
(seed 1003), each time step solved by FAS V(2,2) cycles until the RMS
residual is below 1e-8.  A "step" is one time step of the reference's time
loop without regrid (bsamroutines.f90:673-758): SetGhost(root), on the first
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
    ap.add_argument("--n", type=int, default=256, help="root cells per side (default 256)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--check-fp64", type=int, default=1,
                    help="mixed mode: also run the FP64 path and report the max-norm difference")
    ap.add_argument("--kstat-steps", type=int, default=1,
                    help="extra instrumented steps for the kernel roofline (0: skip)")
    return ap.parse_args()


def make_state(n, precision, use_graph):
    import numpy as np
    from bsam_amd import abi, configs, gpu, hierarchy
    w = configs.WORKLOADS["ch3d_uniform_256"]
    nd = 3
    minlevel = -(int(round(np.log2(n))) - 1)
    hier = hierarchy.uniform(nd, (n,) * nd, minlevel, [2] * 6)
    prec = abi.PREC_MIXED if precision == "mixed" else abi.PREC_FP64
    p = configs.params(nd, w["dt"], w["eps"], w["tol"], precision=prec, use_graph=use_graph)
    S = gpu.Solver(p, hier)
    root = len(hier) - 1
    phi = configs.init_phi(hier[root], nd, 1, w["init"]["seed"], w["init"]["base"], w["init"]["amp"])
    q0 = np.zeros((2,) + (n + 2,) * nd)
    q0[0, 1:-1, 1:-1, 1:-1] = phi
    S.put(root, "q", q0)
    S.set_ghost(0, 0)
    S.copy_q_to_qold(0)
    return S, hier, p, q0, minlevel


def cpu_baseline(n, hier, q0, minlevel):
    """The second time step of the same workload on the C oracle (one
    thread); the first step, untimed, brings it to the state the timed GPU
    steps start from."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from bsam_amd import configs
    w = configs.WORKLOADS["ch3d_uniform_256"]
    O = oracle.Oracle(configs.params(3, w["dt"], w["eps"], w["tol"]), hier)
    O.put(len(hier) - 1, "q", q0)
    O.set_ghost(0, 0)
    O.copy_q_to_qold(0)
    O.time_step(True)
    t0 = time.perf_counter()
    rc, ncyc, r = O.time_step(False)
    dt = time.perf_counter() - t0
    return dict(value=1.0 / dt, unit="time-steps/s", cores=1, kind="port",
                sample=f"second time step of the {n}^3 workload ({ncyc} V-cycles, {dt:.2f} s, "
                       f"{1e3 * dt / max(ncyc, 1):.0f} ms per V-cycle) on the C oracle "
                       f"(oracle/bsam_oracle.c, gcc -O2, 1 thread, {_cpu_model()})")


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    from bsam_amd import abi
    n = a.n
    S, hier, p, q0, minlevel = make_state(n, a.precision, a.graph)
    print(f"[bench] {n}^3 {a.precision} graph={a.graph}: warmup {a.warmup}", file=sys.stderr, flush=True)
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
        kname = DOMINANT[a.precision]
        roof = dict(bound="hbm", kernel=f"{kname} (finest-level colour sweep)", achieved=achieved,
                    peak=HBM_PEAK_GBS, unit="GB/s", frac=(achieved / HBM_PEAK_GBS) if achieved else None,
                    traffic=measured_traffic(f"ch3d_uniform_{n}", a.precision, kname),
                    avg_us=avg_ms * 1e3, launches=nl, bytes_per_launch=b)
    err = None
    if a.precision == "mixed" and a.check_fp64:
        # the FP64 path on the same steps (bit-identical to the oracle in the
        # parity tests) as the reference for the mixed path's error
        import numpy as np
        print("[bench] fp64 reference run for the mixed-precision error", file=sys.stderr, flush=True)
        root = len(hier) - 1
        qm = S.get(root, "q")
        S.close()
        R, _, _, _, _ = make_state(n, "fp64", a.graph)
        first_r = True
        for _ in range(a.warmup + a.steps + a.kstat_steps):
            R.time_step(first_r)
            first_r = False
        qr = R.get(root, "q")
        inner = (slice(1, -1),) * 3
        err = {"phi": float(np.max(np.abs(qm[(0,) + inner] - qr[(0,) + inner]))),
               "mu": float(np.max(np.abs(qm[(1,) + inner] - qr[(1,) + inner]))),
               "steps": a.warmup + a.steps + a.kstat_steps}
        R.close()
    cpu = None
    if a.cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(n, hier, q0, minlevel)
    vc = sum(cycles)
    out = {
        "metric": "AMR CH time-steps/s at ||r||<=1e-8 (3D CH, plain FP64 path)" if a.precision == "fp64"
                  else "AMR CH time-steps/s at ||r||<=1e-8 (3D CH, mixed FP32/FP64 path)",
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
        "config": {"workload": f"ch3d_uniform_{n}", "mesh": f"{n}^3 periodic root, levels 0..{minlevel}",
                   "eps": 0.015, "dt": 5e-5, "tol": 1e-8, "cycle": "FAS V(2,2)",
                   "precision": a.precision, "hipgraph": bool(a.graph)},
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
