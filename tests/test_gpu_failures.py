"""Failure handling of the HIP path, forced from the host so that every
kernel still runs to its normal end:

* maxvcycles exhausted: BSAM_NOT_CONVERGED, with the count and the last
  residual, as the reference's loop ends silently at maxvcycles
  (LIBRARY/afasroutines.f90:79-108), in both precisions;
* a non-finite value in the solution (a NaN uploaded into q): the residual
  norm is NaN after the first cycle and the call returns BSAM_EDIVERGED
  instead of iterating on; the context stays usable once a finite state is
  uploaded again;
* the mixed path's hand-over to FP64 V-cycles when the outer iteration
  stalls (forced with a stall ratio of 0): the step still converges, to the
  FP64 path's solution.
"""
import numpy as np
import pytest

import fixture_util as fu
from bsam_amd import abi
from bsam_amd.gpu import BsamError

pytestmark = pytest.mark.gpu
NAME = "ch3d_amr2"


def _solver(cls, precision, **kw):
    meta, z = fu.load(NAME)
    st = meta["steps"][0]
    S = cls(fu.step_params(meta, precision=precision, **kw), st["hierarchy"])
    fu.load_state(S, meta, z, 0)
    return S, meta, z, st


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_MIXED])
def test_maxvcycles_exhausted(precision, gpu_solver_cls):
    print(f"[gpu] not converged, precision {precision}", flush=True)
    meta, _ = fu.load(NAME)
    S, meta, z, st = _solver(gpu_solver_cls, precision, maxvcycles=2, use_graph=1)
    rc, n, r, hist = S.multigrid_iterations()
    assert rc == abi.BSAM_NOT_CONVERGED and n == 2 and len(hist) == 2, (rc, n, hist)
    assert r == hist[-1] > meta["tol"] and np.isfinite(r)
    S.close()


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_MIXED])
def test_nan_in_solution_diverges(precision, gpu_solver_cls):
    print(f"[gpu] NaN in q, precision {precision}", flush=True)
    S, meta, z, st = _solver(gpu_solver_cls, precision, use_graph=1)
    fine = [i for i in st["end"] if st["hierarchy"][i]["level"] == st["finestlevel"]][0]
    q = S.get(fine, "q")
    q[(0,) + (3,) * (q.ndim - 1)] = np.nan
    S.put(fine, "q", q)
    with pytest.raises(BsamError) as ei:
        S.multigrid_iterations()
    assert ei.value.code == abi.BSAM_EDIVERGED, ei.value
    assert "not finite" in str(ei.value)
    # a finite state again: the same context converges as before
    fu.load_state(S, meta, z, 0)
    rc, n, r, hist = S.multigrid_iterations()
    assert rc == abi.BSAM_OK and r < meta["tol"], (rc, n, r)
    S.close()


def test_mixed_hands_over_to_fp64(gpu_solver_cls):
    print("[gpu] mixed -> FP64 hand-over", flush=True)
    out = {}
    for mode in ("fp64", "mixed", "forced"):
        prec = abi.PREC_FP64 if mode == "fp64" else abi.PREC_MIXED
        S, meta, z, st = _solver(gpu_solver_cls, prec, use_graph=1)
        if mode == "forced":
            S.set_mixed_fallback(0.0)
        rc, n, r, hist = S.multigrid_iterations()
        assert rc == abi.BSAM_OK and r < meta["tol"], (mode, rc, n, r)
        if prec == abi.PREC_MIXED:
            assert S.mixed_fallbacks() == (1 if mode == "forced" else 0), mode
        out[mode] = {i: S.get(i, "q") for i in st["end"]}
        S.close()
    for mode in ("mixed", "forced"):
        err = max(float(np.max(np.abs(out[mode][i] - out["fp64"][i]))) for i in out["fp64"])
        print(f"[gpu] {mode} vs fp64 max-norm {err:.3e}", flush=True)
        assert err < 1e-7, (mode, err)
