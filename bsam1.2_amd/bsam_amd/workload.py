"""The bench workloads: a mesh hierarchy (uniform, or exported from the
reference's own regrid, `workloads/*.json`), run parameters and the seeded
initial q of every patch, loadable into the GPU solver or the oracle."""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


class Workload:
    """Hierarchy, run parameters and initial q of every patch."""

    def __init__(self, name, n):
        import numpy as np
        from bsam_amd import configs, hierarchy
        self.name = name
        self.tumour = None
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
            w = json.load(open(os.path.join(HERE, "workloads", name + ".json")))
            self.nd, self.n, self.minlevel = w["ndims"], w["n"], w["minlevel"]
            base = hierarchy.uniform(self.nd, (self.n,) * self.nd, self.minlevel, w["mthbc"])
            self.hier = hierarchy.from_patch_boxes(base, w["boxes"], self.nd)
            self.eps, self.dt, self.tol, self.steps = w["eps"], w["dt"], w["tol"], w["steps"]
            self.tumour = w.get("tumour")
            i = w["init"]
            self.init = dict(mode=i["initmode"], seed=i["seed"], amp=i["amp"], radius=i["radius"],
                             pamp=i.get("pamp", 0.02), lmax=i.get("lmax", 6))
            bc = "periodic" if w["mthbc"][0] == 2 else "physical-boundary"
            self.label = (f"{self.n}^{self.nd} {bc} root, levels {self.minlevel}..{w['maxlevel']}, "
                          f"patches per level {w['patches_per_level']}, cells per level {w['cells_per_level']}")
        self.nv = 4 if self.tumour else 2
        self.root = [k for k, d in enumerate(self.hier) if d["level"] == 0][0]
        self.q0 = {}
        for k, d in enumerate(self.hier):
            if d["level"] < 0:
                continue
            q = np.zeros(hierarchy.field_shape(d, self.nd, self.nv, "q"))
            q[(0,) + (slice(1, -1),) * self.nd] = configs.init_phi(d, self.nd, eps=self.eps, **self.init)
            if self.tumour:
                q[(2,) + (slice(1, -1),) * self.nd] = 1.0  # nutrient at its far-field value, pressure 0
            self.q0[k] = q
        self.finest = max(d["level"] for d in self.hier)

    def params(self, precision, use_graph, batching):
        from bsam_amd import abi, configs
        prec = abi.PREC_MIXED if precision == "mixed" else abi.PREC_FP64
        kw = {}
        if self.tumour:
            t = self.tumour
            kw = dict(model=abi.MODEL_TUMOUR, extra=(t["lamp"], t["lama"], t["kappa"]))
        return configs.params(self.nd, self.dt, self.eps, self.tol, precision=prec, use_graph=use_graph,
                              batching=batching, **kw)

    def load(self, S):
        """Initial fields into solver S (GPU or oracle), ghosts set level by
        level, qold = q."""
        for k, q in self.q0.items():
            S.put(k, "q", q)
        for lv in range(0, self.finest + 1):
            S.set_ghost(lv, 0)
        for lv in range(0, self.finest + 1):
            S.copy_q_to_qold(lv)
