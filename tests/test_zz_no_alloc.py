"""The allocation-free tagging call, bsam_estimate_level_errors_noalloc,
against bsam_estimate_level_errors on the same states: flags, counts and
(errflagopt 1) the fields the call changes, bit for bit, on two calls in a
row on one context, the second on a smaller level; a level beyond the fixed
scratch array is refused with nothing run."""
import numpy as np
import pytest

import fixture_util as fu
from bsam_amd import abi, configs, hierarchy
from bsam_amd.gpu import BsamError

pytestmark = pytest.mark.gpu


def _pair(gpu_cls, name):
    meta, z = fu.load(name)
    st = meta["steps"][0]
    out = []
    for _ in range(2):
        S = gpu_cls(fu.step_params(meta), st["hierarchy"])
        fu.load_state(S, meta, z, 0)
        out.append(S)
    return st["hierarchy"], out


@pytest.mark.parametrize("opt,qtol", [(10, 1e-2), (1, 1e-7)])
def test_noalloc_matches_tagging(opt, qtol, gpu_solver_cls):
    hier, (A, B) = _pair(gpu_solver_cls, "ch3d_amr2")
    lo, hi = hierarchy.levels(hier)
    sizes = {lv: sum(int(np.prod(d["mx"])) for d in hier if d["level"] == lv) for lv in range(1, hi + 1)}
    # the larger level first, then the smaller one
    order = sorted(sizes, key=lambda lv: -sizes[lv])
    assert sizes[order[0]] > sizes[order[1]]
    for lv in order:
        print(f"[gpu] noalloc tagging opt {opt} level {lv} ({sizes[lv]} cells)", flush=True)
        for X in (A, B):
            X.set_ghost(lv, 0)
        fa, na = A.estimate_level_errors(lv, opt, qtol)
        fb, nb = B.estimate_level_errors_noalloc(lv, opt, qtol)
        assert np.array_equal(fa, fb) and na == nb, (lv, opt, int(np.sum(fa != fb)))
        assert 0 < na < fa.size, (lv, na)
        if opt == 1:
            for k, d in enumerate(hier):
                if d["level"] in (lv, lv - 1):
                    for f in ("q", "qold"):
                        assert np.array_equal(A.get(k, f), B.get(k, f)), (lv, k, f)
    A.close()
    B.close()


def test_noalloc_refuses_a_level_beyond_the_scratch(gpu_solver_cls):
    """A 256^3 root (16.8 M cells) needs 64 MiB of flags: refused, and the
    growing entry point still tags it."""
    n = 256
    hier = hierarchy.uniform(3, (n,) * 3, -1, [2] * 6)
    S = gpu_solver_cls(configs.params(3, 5e-5, 0.015, 1e-8), hier)
    print("[gpu] noalloc refusal on 256^3", flush=True)
    with pytest.raises(BsamError) as ei:
        S.estimate_level_errors_noalloc(0, 10, 1e-2)
    assert ei.value.code == abi.BSAM_EINVAL and "at most 4194304" in str(ei.value)
    flags, nt = S.estimate_level_errors(0, 10, 1e-2)
    assert flags.size == n ** 3 and nt == 0
    S.close()
