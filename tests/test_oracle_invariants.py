"""The CPU oracle against what the mathematics fixes, independently of the
reference's golden vectors.

  * Stationarity: with f := L(q) for arbitrary q, a relaxation sweep leaves
    q unchanged (the pointwise Newton / Gauss-Seidel update is exact at the
    solution of its own equations) -- CH (CH/problem.f90:318-369, 554-614
    against Operator :393-422, 638-669), BioGrowth (BioGrowth/problem.f90:
    255-353) and the tumour model.  A relaxation and an operator that
    disagree in a sign, a coefficient or a neighbour fail it.
  * Closed forms: on a periodic mesh the undivided Laplacian of a Fourier
    mode is (sum_d 2 cos(k_d h) - 2) times the mode, so Operator(q) of Fourier
    modes is known exactly (ULap2D/3D, gridutilities.f90:50-72, 372-399).
  * Linear fields: the coarse-fine ghost interpolation (quadratic along a
    line, linear or bilinear across it: boundary.f90:54-63, 143-333,
    1048-1550), the prolongation (BiLin/TriLinProlongationP1, gridutilities.
    f90:156-195, 510-566) and the restriction (Restriction2D/3D :103-126,
    437-466) reproduce a globally linear field exactly.
  * Conservation: a converged Cahn-Hilliard step on a periodic mesh keeps
    the mean of phi (phi - dt Lap(mu) = phi_old sums to zero change).
"""
import numpy as np
import pytest

import oracle
from bsam_amd import abi, configs, hierarchy


def _root(hier):
    return [k for k, d in enumerate(hier) if d["level"] == 0][0]


def _inner(nd):
    return (slice(1, -1),) * nd


def _model_params(model, nd, dt=1e-3, eps=0.1, tol=1e-9):
    if model == "ch":
        return configs.params(nd, dt, eps, tol)
    if model == "bg":
        return configs.params(nd, dt, eps, tol, model=abi.MODEL_BG, extra=(1.5, 2.0, 0.5))
    return configs.params(nd, dt, eps, tol, model=abi.MODEL_TUMOUR, extra=(1.0, 0.1, 10.0))


@pytest.mark.parametrize("model,nd,mthbc", [("ch", 2, 2), ("ch", 3, 2), ("ch", 3, 10), ("bg", 2, 10),
                                            ("tumour", 2, 10), ("tumour", 3, 10), ("tumour", 3, 2)])
def test_relax_is_stationary_at_the_solution(model, nd, mthbc):
    n = 8
    hier = hierarchy.uniform(nd, (n,) * nd, -2, [mthbc] * (2 * nd))
    root = _root(hier)
    p = _model_params(model, nd)
    O = oracle.Oracle(p, hier)
    rng = np.random.default_rng(5)
    q = O.field(root, "q")
    q[...] = 0.0
    q[(slice(None),) + _inner(nd)] = 0.3 * rng.standard_normal(q[(slice(None),) + _inner(nd)].shape)
    if model == "tumour":
        q[(2,) + _inner(nd)] += 1.0
    O.put(root, "qold", q + 0.1 * rng.standard_normal(q.shape))
    O.set_ghost(0, 0)
    O.put(root, "f", np.zeros_like(O.get(root, "f")))
    O.residual(0)                       # rf = f - L(q) = -L(q)
    O.put(root, "f", -O.get(root, "rf"))
    before = O.get(root, "q")
    for ipass in (1, 2):
        O.relax_level(0, ipass)
        O.set_ghost(0, ipass)
    after = O.get(root, "q")
    scale = np.max(np.abs(before))
    assert np.max(np.abs(after - before)) <= 1e-12 * scale
    # and a plain perturbation of f is not stationary (the check has teeth)
    O.put(root, "f", O.get(root, "f") + 1e-3)
    O.relax_level(0, 1)
    assert np.max(np.abs(O.get(root, "q") - after)) > 1e-8


@pytest.mark.parametrize("model,nd", [("ch", 2), ("ch", 3), ("tumour", 2), ("tumour", 3)])
def test_operator_of_fourier_modes(model, nd):
    n = 16
    hier = hierarchy.uniform(nd, (n,) * nd, -3, [2] * (2 * nd))
    root = _root(hier)
    dt, eps = 2e-4, 0.05
    p = _model_params(model, nd, dt=dt, eps=eps)
    O = oracle.Oracle(p, hier)
    h = 1.0 / n
    xc = (np.arange(n) + 0.5) * h
    grids = np.meshgrid(*([xc] * nd), indexing="ij")  # (x, y[, z])
    ks = [(1, 2, 1), (3, 1, 2), (2, 2, 1), (1, 3, 3)]  # wave numbers of the four fields
    modes, lams = [], []
    for kv in ks:
        m = np.ones_like(grids[0])
        lam = 0.0
        for a in range(nd):
            m = m * np.cos(2 * np.pi * kv[a] * grids[a] + 0.3 * (a + 1))
            lam += 2.0 * np.cos(2 * np.pi * kv[a] * h) - 2.0
        modes.append(m)
        lams.append(lam)
    nv = p.nrvars
    q = np.zeros(hierarchy.field_shape(hier[root], nd, nv, "q"))
    amp = [0.4, 0.7, 0.2, 0.5]
    for v in range(nv):
        q[(v,) + _inner(nd)] = amp[v] * modes[v].T      # stored (var, [k,] j, i)
    qold_val = -0.3
    O.put(root, "q", q)
    O.put(root, "qold", np.full_like(q, qold_val))
    O.set_ghost(0, 0)
    O.put(root, "f", np.zeros_like(O.get(root, "f")))
    O.residual(0)
    got = -O.get(root, "rf")                            # L(q)
    phi, mu = amp[0] * modes[0], amp[1] * modes[1]
    L = [phi - dt / h**2 * lams[1] * mu, mu - phi**3 + eps**2 / h**2 * lams[0] * phi]
    if model == "tumour":
        nut, pr = amp[2] * modes[2], amp[3] * modes[3]
        t0 = 0.5 * (qold_val + 1.0)
        lamp, kappa = 1.0, 10.0
        L[0] = L[0] - dt * lamp * t0 * nut
        L += [-lams[2] / h**2 * nut + kappa * t0 * nut, -lams[3] / h**2 * pr - lamp * t0 * nut]
    for v in range(nv):
        ref = L[v].T
        assert np.max(np.abs(got[v] - ref)) <= 1e-11 * (1.0 + np.max(np.abs(ref))), (model, v)


def _amr_linear(nd, model):
    """16^d root with physical faces and two level-1 patches away from the
    root's boundary, sharing a face; every level >= 0 holds the same
    globally linear field in every variable."""
    base = hierarchy.uniform(nd, (16,) * nd, -3, [10] * (2 * nd))
    if nd == 2:
        boxes = [[1, [[9, 16], [9, 20]]], [1, [[17, 24], [9, 20]]]]
    else:
        boxes = [[1, [[9, 16], [9, 16], [9, 20]]], [1, [[17, 24], [9, 16], [9, 20]]]]
    hier = hierarchy.from_patch_boxes(base, boxes, nd)
    O = oracle.Oracle(_model_params(model, nd), hier)
    coef = np.array([0.7, -1.3, 0.4, 2.1][:nd])

    def lin(d, shape_field):
        h = d["dx"]
        ext = [d["mx"][a] + 2 for a in range(nd)]
        axes = [d["xlower"][a] + (np.arange(ext[a]) - 0.5) * h for a in range(nd)]
        g = np.meshgrid(*axes, indexing="ij")
        val = 0.25 + sum(coef[a] * g[a] for a in range(nd))
        return val.T                                       # ([k,] j, i)
    return O, hier, lin


@pytest.mark.parametrize("model,nd", [("ch", 2), ("ch", 3), ("tumour", 3)])
def test_coarse_fine_ghosts_reproduce_linear_fields(model, nd):
    O, hier, lin = _amr_linear(nd, model)
    nv = O.nv
    fine = [k for k, d in enumerate(hier) if d["level"] == 1]
    for k, d in enumerate(hier):
        if d["level"] < 0:
            continue
        q = np.empty(hierarchy.field_shape(d, nd, nv, "q"))
        for v in range(nv):
            q[v] = (v + 1) * lin(d, "q")
        if d["level"] == 1:
            q[(slice(None),)] = np.where(np.pad(np.ones(q.shape[1:], bool)[_inner(nd)], 1), q, np.nan)
        O.put(k, "q", q)
    O.set_ghost(1, 0)
    nface = 0
    for k in fine:
        d = hier[k]
        got = O.get(k, "q")
        for v in range(nv):
            ref = (v + 1) * lin(d, "q")
            ok = ~np.isnan(got[v])
            assert np.max(np.abs(got[v][ok] - ref[ok])) <= 1e-13, (k, v)
            # every face ghost is written
            idx = np.indices(got[v].shape)
            out = sum(((idx[a] == 0) | (idx[a] == got[v].shape[a] - 1)).astype(int) for a in range(nd))
            assert not np.isnan(got[v][out == 1]).any()
            nface += int((out == 1).sum())
    assert nface > 0


@pytest.mark.parametrize("nd", [2, 3])
def test_prolongation_and_restriction_of_linear_fields(nd):
    O, hier, lin = _amr_linear(nd, "ch")
    fine = [k for k, d in enumerate(hier) if d["level"] == 1]
    root = _root(hier)
    # prolongation: q = 0, qc = 0, parent = linear  ->  q = linear (ghosts too)
    O.put(root, "q", np.stack([lin(hier[root], "q"), -2.0 * lin(hier[root], "q")]))
    for k in fine:
        O.put(k, "q", np.zeros(hierarchy.field_shape(hier[k], nd, 2, "q")))
        O.put(k, "qc", np.zeros(hierarchy.field_shape(hier[k], nd, 2, "qc")))
    O.correct_fine(1)
    for k in fine:
        got = O.get(k, "q")
        ref = lin(hier[k], "q")
        assert np.max(np.abs(got[0] - ref)) <= 1e-13
        assert np.max(np.abs(got[1] + 2.0 * ref)) <= 1e-13
    # restriction: fine = linear  ->  parent cells under the patch = linear
    for k in fine:
        O.put(k, "q", np.stack([lin(hier[k], "q")] * 2))
    O.put(root, "q", np.zeros(hierarchy.field_shape(hier[root], nd, 2, "q")))
    O.restrict_solution(1)
    pr = O.get(root, "q")
    ref = lin(hier[root], "q")
    covered = 0
    for k in fine:
        mb = hier[k]["mbounds"]
        sl = tuple(slice(mb[a][0], mb[a][1] + 1) for a in range(nd))[::-1]
        assert np.max(np.abs(pr[0][sl] - ref[sl])) <= 1e-13
        covered += pr[0][sl].size
    assert covered > 0


@pytest.mark.parametrize("nd", [2, 3])
def test_ch_step_conserves_mass(nd):
    n = 16
    hier = hierarchy.uniform(nd, (n,) * nd, -3, [2] * (2 * nd))
    root = _root(hier)
    O = oracle.Oracle(configs.params(nd, 1e-4, 0.05, 1e-11), hier)
    phi = configs.init_phi(hier[root], nd, 1, 77, 0.1, 0.3)
    q = np.zeros(hierarchy.field_shape(hier[root], nd, 2, "q"))
    q[(0,) + _inner(nd)] = phi
    O.put(root, "q", q)
    O.set_ghost(0, 0)
    O.copy_q_to_qold(0)
    m0 = O.get(root, "q")[(0,) + _inner(nd)].mean()
    for s in range(2):
        rc, ncyc, r = O.time_step(s == 0)
        assert rc == abi.BSAM_OK
    m1 = O.get(root, "q")[(0,) + _inner(nd)].mean()
    assert abs(m1 - m0) <= 1e-10
    assert abs(m0 - 0.1) < 0.05   # the field is not trivially zero-mean
