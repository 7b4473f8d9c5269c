"""GPU parity of the 2D fiber-frame transient (sgpu_frame2d_*) against the
oracle's restatement of Strut's Newmark / Newton driver on the same seeded
frames and ground motions.

Bar: identical Newton iteration counts at every step, and nodal
displacement histories within 1e-8 relative max error (max over the
history of |u_gpu - u_oracle| / max |u_oracle|, per DOF with a floor of
1e-12 m).
"""

import numpy as np
import pytest

from oracle import oracle
from strut_amd import _native, cases
from strut_amd.frame2d import load_frame2d_case
from strut_amd.transient import FrameBatch

pytestmark = pytest.mark.gpu

REL = 1e-8


def _rel_err(u_gpu, u_ref):
    scale = np.maximum(np.max(np.abs(u_ref), axis=0), 1e-12)
    return float(np.max(np.abs(u_gpu - u_ref) / scale))


SCHEDULES = [_native.SCHEDULE_ROUNDS, _native.SCHEDULE_PERSISTENT, _native.SCHEDULE_QUEUE, _native.SCHEDULE_AUTO]


@pytest.mark.parametrize("schedule", SCHEDULES, ids=["rounds", "persistent", "queue", "auto"])
def test_gpu_config0_plain_matches_oracle(schedule):
    fc = load_frame2d_case(cases.config0_case(steps=1000, seed=0))
    ref = oracle.run_transient2d(fc)
    assert ref["status"] == 0
    fb = FrameBatch(fc.materials)
    f = fb.add_frame(fc, record_all=True)
    i = fb.add_case_motion(f, fc)
    fb.finalize()
    fb.set_schedule(schedule)
    fb.run(rounds=64)
    got = fb.result(i)
    sec_evals = int(fb.section_evals()[i])
    fb.close()
    print(f"config0 1000 steps: status {got['status']}, iterations {got['iters'].sum()} vs {ref['iters'].sum()}, "
          f"section evaluations {sec_evals} vs {ref['section_evals']}")
    assert got["status"] == 0 and got["steps_done"] == 1000
    np.testing.assert_array_equal(got["iters"], ref["iters"])
    assert sec_evals == ref["section_evals"]  # the same local iterations in every element
    err = _rel_err(got["u"], ref["u"])
    print(f"config0 max relative displacement error {err:.3e}")
    assert err <= REL


@pytest.mark.parametrize("schedule", SCHEDULES, ids=["rounds", "persistent", "queue", "auto"])
def test_gpu_ensemble_of_uneven_instances_matches_oracle(schedule):
    """Two frames x several motions of different lengths in one launch."""
    specs = [(3, 1, 0, 0.4, 260), (3, 1, 1, 0.8, 180), (4, 3, 2, 0.6, 140), (4, 3, 3, 0.3, 220)]
    fcs = [load_frame2d_case(cases.rc_frame2d_case(s, b, steps=n, dt=0.01, seed=seed, pga_g=pga, strong=3.0))
           for s, b, seed, pga, n in specs]
    fb = FrameBatch(fcs[0].materials)
    frame_of = {}
    ids = []
    for fc, (s, b, *_rest) in zip(fcs, specs):
        if (s, b) not in frame_of:
            frame_of[(s, b)] = fb.add_frame(fc, record_all=True)
        ids.append(fb.add_case_motion(frame_of[(s, b)], fc))
    fb.finalize()
    fb.set_schedule(schedule)
    fb.run(rounds=16)
    for fc, iid in zip(fcs, ids):
        got = fb.result(iid)
        ref = oracle.run_transient2d(fc)
        assert ref["status"] == 0 and got["status"] == 0
        np.testing.assert_array_equal(got["iters"], ref["iters"])
        assert _rel_err(got["u"], ref["u"]) <= REL
    fb.close()


@pytest.mark.parametrize("column_type,beam_type", [("forceBeamColumn2d", "dispBeamColumn2d"),
                                                   ("dispBeamColumn2d", "dispBeamColumn2d")],
                         ids=["fbc-columns-dbc-beams", "dbc-all"])
def test_gpu_disp_beam_column_frames_match_oracle(column_type, beam_type):
    """dispBeamColumn2d elements (disp_beam_column2d.mojo:114-254) in the
    3-story frame, alone or with forceBeamColumn2d columns: identical Newton
    counts and displacements within 1e-8 of the oracle."""
    case = cases.rc_frame2d_case(3, 1, steps=300, dt=0.01, seed=0, pga_g=0.4,
                                 column_type=column_type, beam_type=beam_type)
    fc = load_frame2d_case(case)
    assert set(fc.elem_type.tolist()) == ({0, 1} if column_type != beam_type else {1})
    ref = oracle.run_transient2d(fc)
    fb = FrameBatch(fc.materials)
    f = fb.add_frame(fc, record_all=True)
    i = fb.add_case_motion(f, fc)
    fb.finalize()
    fb.run(rounds=32)
    got = fb.result(i)
    fb.close()
    n = ref["steps_done"]
    print(f"{column_type}/{beam_type}: status {got['status']} vs {ref['status']}, "
          f"iterations {got['iters'][:n].sum()} vs {ref['iters'][:n].sum()}")
    assert got["status"] == ref["status"] and got["steps_done"] == n
    np.testing.assert_array_equal(got["iters"][:n], ref["iters"][:n])
    err = _rel_err(got["u"][:n], ref["u"][:n])
    print(f"max relative displacement error {err:.3e}")
    assert err <= REL


@pytest.mark.parametrize("variant", ["all-elastic", "elastic-beams"])
def test_gpu_all_elastic_sections_take_the_exact_state(variant):
    """forceBeamColumn2d elements whose sections have only elastic fibers take
    the reference's exact state (force_beam_column2d.mojo:608-732, tried
    before the local iteration at :1577-1620): on the whole frame (a linear
    problem: two assemblies per step) or on its beams only, next to RC
    columns.  Identical Newton and section-evaluation counts, displacements
    within 1e-8 of the oracle, whose exact state is pinned against an
    independent linear Newmark solution."""
    import copy

    case = cases.config0_case(steps=300, seed=3)
    if variant == "all-elastic":
        case["materials"] = [{"id": m["id"], "type": "Elastic", "params": {"E": 3.0e7}} for m in case["materials"]]
    else:
        case["materials"].append({"id": 99, "type": "Elastic", "params": {"E": 3.0e7}})
        beam = copy.deepcopy(next(s for s in case["sections"] if s["id"] == 2))
        beam["id"] = 3
        for part in beam["params"].get("patches", []) + beam["params"].get("layers", []):
            part["material"] = 99
        case["sections"].append(beam)
        for e in case["elements"]:
            if e["section"] == 2:
                e["section"] = 3
    fc = load_frame2d_case(case)
    ref = oracle.run_transient2d(fc)
    assert ref["status"] == 0
    if variant == "all-elastic":
        assert np.all(ref["iters"] == 2)
    fb = FrameBatch(fc.materials)
    f = fb.add_frame(fc, record_all=True)
    i = fb.add_case_motion(f, fc)
    fb.finalize()
    fb.run(rounds=32)
    got = fb.result(i)
    sec_evals = int(fb.section_evals()[i])
    fb.close()
    print(f"{variant}: iterations {got['iters'].sum()} vs {ref['iters'].sum()}, section evaluations {sec_evals} vs "
          f"{ref['section_evals']}, max relative error {_rel_err(got['u'], ref['u']):.2e}", flush=True)
    assert got["status"] == 0 and got["steps_done"] == 300
    np.testing.assert_array_equal(got["iters"], ref["iters"])
    assert sec_evals == ref["section_evals"]
    assert _rel_err(got["u"], ref["u"]) <= REL
