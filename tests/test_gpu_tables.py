"""The flattened ghost tables planned on the device (plan_maps_device: the
symbolic run of SetGhost(level, ipass) and the selection of the seven
tables, cell by cell) against the host planner (plan_ghost_maps), which the
host self-test pins against the staged fill: for every refined level of
the reference's AMR fixtures and of the bench meshes, the two planners must
give the same tables rule for rule (counts, first-launch counts and a hash
of the bytes), or both refuse the level.  The bitwise GPU parity tests run
with the device planner (the default), so they check its tables' values as
well."""
import pytest

import fixture_util as fu
from bsam_amd import abi
from bsam_amd.gpu import BsamError

pytestmark = pytest.mark.gpu

FIXTURES = ["ch2d_amr", "ch2d_amr_phys", "ch3d_amr", "ch3d_amr_phys", "ch3d_amr2", "ch3d_amr2_phys",
            "tm2d_amr_phys", "tm3d_amr_phys"]
WORKLOADS = ["ch3d_amr_64_l3", "tm3d_amr_64_l3", "ch2d_amr_128_l2"]


def _cases():
    out = []
    for name in FIXTURES:
        meta, _ = fu.load(name)
        for s in range(len(meta["steps"])):
            out.append((name, s))
    return out + [(w, -1) for w in WORKLOADS]


def _mesh(name, step):
    if step < 0:
        from bsam_amd.workload import Workload
        wl = Workload(name, 256)
        return wl.params("fp64", 0, 1), wl.hier
    meta, _ = fu.load(name)
    return fu.step_params(meta), meta["steps"][step]["hierarchy"]


def _digests(cls, params, hier, planner):
    nroot = sum(1 for d in hier if d["level"] <= 0)
    assert all(d["level"] <= 0 for d in hier[:nroot])
    # install the levels at and below the root first, so that every refined
    # level is built by the install that follows, with the chosen planner
    S = cls(params, hier[:nroot])
    S.set_knob(abi.KNOB_TABLE_PLANNER, 0, planner)
    assert S.get_knob(abi.KNOB_TABLE_PLANNER, 0) == planner
    S.set_hierarchy(hier)
    out = {}
    for lv in range(1, max(d["level"] for d in hier) + 1):
        try:
            out[lv] = S.ghost_table_digest(lv)
        except BsamError as e:
            assert e.code == abi.BSAM_ESTATE
            print(f"  planner {planner} level {lv}: {e}", flush=True)
            out[lv] = "staged fill"
    S.close()
    return out


@pytest.mark.parametrize("name,step", _cases())
def test_device_planner_matches_host_planner(name, step, gpu_solver_cls):
    print(f"[gpu] ghost tables {name} step {step}", flush=True)
    params, hier = _mesh(name, step)
    host = _digests(gpu_solver_cls, params, hier, 0)
    dev = _digests(gpu_solver_cls, params, hier, 1)
    assert any(v != "staged fill" for v in host.values())
    for lv in host:
        assert dev[lv] == host[lv], (lv, dev[lv], host[lv])
