"""Workload definitions: seeded initial data and run parameters.

The reference ships no seeded inputs (its QInit hooks are fixed shapes,
PROBLEM/CH/problem.f90:45-146), so the workloads define their own: a
counter-based generator keyed by (seed, level, global cell index) gives
U(-1,1) per cell, identical in this module, in the C oracle's callers and in
the probe PROBLEM module compiled against the reference
(oracle/ref_problem/problem.f90, SeededU).
"""
import math

import numpy as np

from . import abi, hierarchy

M32 = np.uint64(0xFFFFFFFF)
M16 = np.uint64(0xFFFF)


def _mulmod32(x, m):
    m = np.uint64(m)
    xl = x & M16
    xh = (x >> np.uint64(16)) & M16
    return (xl * m + (((xh * m) & M16) << np.uint64(16))) & M32


def _mix32(x):
    x = x & M32
    x = x ^ (x >> np.uint64(16))
    x = _mulmod32(x, 0x7FEB352D)
    x = x ^ (x >> np.uint64(15))
    x = _mulmod32(x, 0x846CA68B)
    x = x ^ (x >> np.uint64(16))
    return x


def seeded_u(seed, level, i, j, k):
    """U(-1,1) for global cell (i,j,k) (1-based) on `level`; arrays broadcast."""
    i = np.asarray(i, dtype=np.int64).astype(np.uint64)
    j = np.asarray(j, dtype=np.int64).astype(np.uint64)
    k = np.asarray(k, dtype=np.int64).astype(np.uint64)
    x = _mix32(np.uint64(seed))
    x = _mix32(x ^ np.uint64(level + 64))
    x = _mix32(x ^ i)
    x = _mix32(x ^ j)
    x = _mix32(x ^ k)
    return 2.0 * (x.astype(np.float64) / 4294967296.0) - 1.0


def sphere_radius(seed, dx, dy, dz, radius, pamp=0.02, lmax=6):
    """Radius of the perturbed sphere in direction (dx, dy, dz) about its
    centre: radius + pamp * sum_{l=1..lmax, |m|<=l} c_lm Y_lm(theta, psi),
    orthonormal real spherical harmonics (polar axis z), c_lm = U(-1,1) of
    the seeded generator keyed by level -32, i = l, j = m+8, k = 1."""
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    with np.errstate(invalid="ignore", divide="ignore"):
        ct = np.where(r > 0.0, dz / np.where(r > 0.0, r, 1.0), 1.0)
    psi = np.where(r > 0.0, np.arctan2(dy, dx), 0.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    s = np.zeros_like(r)
    for mm in range(0, lmax + 1):
        pmm = np.ones_like(r)
        fact = 1.0
        for _ in range(mm):
            pmm = -pmm * fact * st
            fact += 2.0
        pa, pb = np.zeros_like(r), pmm
        for l in range(mm, lmax + 1):
            if l == mm:
                pl = pmm
            else:
                pl = (ct * float(2 * l - 1) * pb - float(l + mm - 1) * pa) / float(l - mm)
                pa, pb = pb, pl
            if l < 1:
                continue
            nrm = 1.0
            for k in range(l - mm + 1, l + mm + 1):
                nrm = nrm / float(k)
            nrm = math.sqrt(float(2 * l + 1) / (4.0 * math.pi) * nrm)
            if mm == 0:
                s = s + float(seeded_u(seed, -32, l, 8, 1)) * nrm * pl
            else:
                ylm = math.sqrt(2.0) * nrm * pl
                s = s + float(seeded_u(seed, -32, l, mm + 8, 1)) * ylm * np.cos(float(mm) * psi)
                s = s + float(seeded_u(seed, -32, l, -mm + 8, 1)) * ylm * np.sin(float(mm) * psi)
    return radius + pamp * s


def init_phi(desc, ndims, mode, seed, base=0.0, amp=0.05, radius=0.25, eps=0.03,
             xc=(0.5, 0.5, 0.5), pamp=0.02, lmax=6):
    """Interior phi of one patch, shape ([mz,] my, mx), following the probe
    module's InitPhi (mode 1: base + amp U; mode 2: tanh circle/sphere +
    amp U; mode 3: tanh perturbed sphere + amp U)."""
    h = desc["dx"]
    lev = desc["level"]
    mx = desc["mx"]
    g0 = [desc["mglobal"][a][0] for a in range(3)]
    ax = [np.arange(1, mx[a] + 1) for a in range(ndims)]
    if ndims == 2:
        J, I = np.meshgrid(ax[1], ax[0], indexing="ij")
        K = np.ones_like(I)
        gi, gj, gk = I + g0[0] - 1, J + g0[1] - 1, K
    else:
        K, J, I = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        gi, gj, gk = I + g0[0] - 1, J + g0[1] - 1, K + g0[2] - 1
    u = seeded_u(seed, lev, gi, gj, gk)
    if mode == 1:
        return base + amp * u
    x = (I - 0.5) * h + desc["xlower"][0]
    y = (J - 0.5) * h + desc["xlower"][1]
    z = (K - 0.5) * h + desc["xlower"][2] if ndims == 3 else xc[2]
    r = np.sqrt((x - xc[0]) ** 2 + (y - xc[1]) ** 2 + (z - xc[2]) ** 2)
    if mode == 3:
        rs = sphere_radius(seed, x - xc[0], y - xc[1], z - xc[2], radius, pamp, lmax)
        return np.tanh((rs - r) / (math.sqrt(2.0) * eps)) + amp * u
    return np.tanh((radius - r) / (math.sqrt(2.0) * eps)) + amp * u


def params(ndims, dt, eps, qerrortol, maxvcycles=50, nsmoothingpasses=2, precision=abi.PREC_FP64,
           model=abi.MODEL_CH, batching=1, use_graph=0, errortype=1, extra=()):
    """Run parameters.  `extra` are the model constants after eps^2: BG
    (gamma, lambda, czero); tumour (proliferation, apoptosis, nutrient
    consumption)."""
    p = abi.Params()
    p.ndims = ndims
    p.nrvars = 4 if model == abi.MODEL_TUMOUR else 2
    p.mbc = 1
    p.maux = 0
    p.model = model
    p.nsmoothingpasses = nsmoothingpasses
    p.maxvcycles = maxvcycles
    p.errortype = errortype
    p.updateauxfreq = 1
    p.precision = precision
    p.batching = batching
    p.use_graph = use_graph
    p.dt = dt
    p.qerrortol = qerrortol
    p.model_params[0] = eps * eps
    for n, v in enumerate(extra):
        p.model_params[1 + n] = v
    return p


# The five workloads (seeds 1001..1005).  Sizes follow the workload text;
# `minlevel` coarsens the root down to 2 cells per side.
WORKLOADS = {
    "ch2d_uniform_128": dict(ndims=2, mx=(128, 128), minlevel=-6, eps=0.03, dt=1e-4, steps=10,
                             tol=1e-9, init=dict(mode=1, seed=1001, base=0.0, amp=0.05)),
    "ch3d_uniform_256": dict(ndims=3, mx=(256, 256, 256), minlevel=-7, eps=0.015, dt=5e-5, steps=10,
                             tol=1e-8, init=dict(mode=1, seed=1003, base=0.1, amp=0.05)),
}


def build_uniform(name, mx=None, minlevel=None):
    w = dict(WORKLOADS[name])
    if mx is not None:
        w["mx"] = mx
        w["minlevel"] = -int(round(math.log2(mx[0]))) + 1
    if minlevel is not None:
        w["minlevel"] = minlevel
    hier = hierarchy.uniform(w["ndims"], w["mx"], w["minlevel"], [2] * (2 * w["ndims"]))
    return w, hier
