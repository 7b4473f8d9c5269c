"""GPU parity of the 2D fiber beam-column elements outside any frame
(sgpu_elem2d_*: forceBeamColumn2d, dispBeamColumn2d) against the oracle's
element routines (oracle/fbc2d.c, oracle/dbc2d.c), through the C ABI.

Elements of both formulations and several geometries, RC sections
(Concrete02 core and cover, Steel02 bars), one launch per call; a history
of committed steps deep into the nonlinear range with three trial states
per step (as Newton iterations make them).  Bar: tangents, forces and the
force-based basic state within 1e-10 relative (Steel02's pow may differ
from the host's in the last ulp); elastic sections give the oracle's bits.
"""

import numpy as np
import pytest

from oracle import oracle
from strut_amd import cases
from strut_amd._native import StrutGpuError
from strut_amd.elements2d import Elements2d
from strut_amd.model import MaterialTable, beam_integration_rule, discretize_section
from strut_amd.sections import SectionRuntime

pytestmark = pytest.mark.gpu

XI, W = (np.array(a) for a in beam_integration_rule("Lobatto", 5))
GEOMS = [((0.0, 0.0), (0.0, 3.0)), ((0.0, 3.0), (6.0, 3.0)), ((1.0, 0.5), (4.0, 4.5)), ((6.0, 3.0), (0.0, 3.0))]
REL = 1e-10


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _setup(mats, cells_by_kind, specs):
    t, p = mats.arrays()
    rt = SectionRuntime(mats, device=0)
    defs = {k: rt.add_definition(2, c) for k, c in cells_by_kind.items()}
    dev = Elements2d(rt)
    refs = []
    for kind, sec_kind, (x1, x2) in specs:
        eid = dev.add(kind, defs[sec_kind], XI, W, x1, x2)
        sec = oracle.Section2d(cells_by_kind[sec_kind], t, p)
        sec.alloc(len(XI))
        refs.append({"kind": kind, "sec": sec, "x1": x1, "x2": x2, "q": np.zeros(3 + 2 * len(XI) + 3)})
        assert eid == len(refs) - 1
    rt.finalize()
    return rt, dev, refs


def _oracle_state(r, u):
    (x1, y1), (x2, y2) = r["x1"], r["x2"]
    if r["kind"] == "fbc":
        return oracle.fbc2d_eval(r["sec"], 0, XI, W, x1, y1, x2, y2, r["q"], u)
    return oracle.dbc2d_eval(r["sec"], 0, XI, W, x1, y1, x2, y2, u)


def _history(n_el, steps, seed):
    rng = np.random.default_rng(seed)
    shape = rng.normal(size=(n_el, 6)) * np.array([0.004, 0.004, 0.002, 0.004, 0.004, 0.002])
    shape[:, [0, 3]] *= 0.25  # mostly bending
    for s in range(steps):
        amp = (1.0 + 4.0 * s / steps) * np.sin(0.7 * (s + 1))
        yield [shape * amp * f for f in (0.6, 0.9, 1.0)]


@pytest.mark.parametrize("sections", ["rc", "elastic"])
def test_gpu_elem2d_matches_oracle_over_a_committed_history(sections):
    if sections == "rc":
        mats = MaterialTable(cases.rc_materials())
        cells = {"col": discretize_section(cases.rc_section(1, dim=2, **cases.FRAME2D_COLUMN), mats),
                 "beam": discretize_section(cases.rc_section(2, dim=2, **cases.FRAME2D_BEAM), mats)}
    else:
        mats = MaterialTable([{"id": 9, "type": "Elastic", "params": {"E": 30.0e6}}])
        sec = {"id": 1, "type": "FiberSection2d", "params": {"patches": [
            {"type": "rect", "material": 9, "num_subdiv_y": 12, "num_subdiv_z": 1,
             "y_i": -0.25, "z_i": -0.15, "y_j": 0.25, "z_j": 0.15}]}}
        cells = {"col": discretize_section(sec, mats), "beam": discretize_section(sec, mats)}
    specs = [(kind, "col" if g == 0 else "beam", GEOMS[g]) for g in range(len(GEOMS)) for kind in ("fbc", "dbc")]
    rt, dev, refs = _setup(mats, cells, specs)
    ids = np.arange(len(specs), dtype=np.int32)
    worst = {"k": 0.0, "f": 0.0, "q": 0.0}
    n_calls = 0
    for trials in _history(len(specs), steps=12, seed=3):
        for u in trials:
            k, f = dev.state(ids, u)
            for e, r in enumerate(refs):
                kr, fr = _oracle_state(r, u[e])
                worst["k"] = max(worst["k"], _rel(k[e], kr))
                worst["f"] = max(worst["f"], _rel(f[e], fr))
                if sections == "elastic":
                    np.testing.assert_array_equal(k[e], kr)
                    np.testing.assert_array_equal(f[e], fr)
            n_calls += 1
        dev.commit(ids)
        for r in refs:
            for ip in range(len(XI)):
                r["sec"].commit(ip)
    for e, r in enumerate(refs):
        if r["kind"] == "fbc":
            worst["q"] = max(worst["q"], _rel(dev.basic_state(e), r["q"]))
    print(f"{sections}: {n_calls} calls of {len(specs)} elements, worst relative k {worst['k']:.2e} "
          f"f {worst['f']:.2e} basic state {worst['q']:.2e}", flush=True)
    assert max(worst.values()) <= REL
    rt.close()


def test_gpu_elem2d_commit_syncs_the_basic_deformation():
    """With the converged end displacements, the commit sets a force-based
    element's basic deformation to the chord values of those displacements
    (helpers.mojo:1832-1883): axial elongation and the two end rotations
    relative to the chord."""
    mats = MaterialTable(cases.rc_materials())
    cells = {"col": discretize_section(cases.rc_section(1, dim=2, **cases.FRAME2D_COLUMN), mats)}
    rt, dev, refs = _setup(mats, cells, [("fbc", "col", ((0.0, 0.0), (3.0, 4.0)))])
    u = np.array([[0.001, -0.002, 0.003, 0.004, 0.001, -0.002]])
    dev.state([0], u)
    dev.commit([0], u)
    q = dev.basic_state(0)
    c, s, L = 0.6, 0.8, 5.0
    ul = [c * u[0, 0] + s * u[0, 1], -s * u[0, 0] + c * u[0, 1], u[0, 2],
          c * u[0, 3] + s * u[0, 4], -s * u[0, 3] + c * u[0, 4], u[0, 5]]
    chord = (ul[4] - ul[1]) / L
    np.testing.assert_allclose(q[-3:], [ul[3] - ul[0], ul[2] - chord, ul[5] - chord], rtol=1e-14, atol=1e-18)
    rt.close()


def test_gpu_elem2d_refuses_bad_calls():
    mats = MaterialTable(cases.rc_materials())
    cells2 = discretize_section(cases.rc_section(1, dim=2, **cases.FRAME2D_COLUMN), mats)
    cells3 = discretize_section(cases.rc_section(1, dim=3, b=0.4, h=0.4, n_core=4, n_bars=4), mats)
    rt = SectionRuntime(mats, device=0)
    d2 = rt.add_definition(2, cells2)
    d3 = rt.add_definition(3, cells3)
    dev = Elements2d(rt)
    with pytest.raises(StrutGpuError, match="2D fiber section"):
        dev.add("fbc", d3, XI, W, (0, 0), (0, 3))
    with pytest.raises(StrutGpuError, match="zero length"):
        dev.add("fbc", d2, XI, W, (1, 1), (1, 1))
    a = dev.add("fbc", d2, XI, W, (0, 0), (0, 3))
    with pytest.raises(StrutGpuError, match="finalize"):
        dev.state([a], np.zeros((1, 6)))
    rt.finalize()
    with pytest.raises(StrutGpuError, match="before sgpu_sections_finalize"):
        dev.add("dbc", d2, XI, W, (0, 0), (3, 0))
    with pytest.raises(StrutGpuError, match="out of range"):
        dev.state([a + 1], np.zeros((1, 6)))
    with pytest.raises(StrutGpuError, match="listed twice"):
        dev.state([a, a], np.zeros((2, 6)))
    k, f = dev.state([a], np.zeros((1, 6)))
    assert np.all(np.isfinite(k)) and np.abs(f).max() <= 1e-9
    rt.close()
