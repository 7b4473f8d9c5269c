"""Host-side checks of the 3D case loader and the Transformation tables
(strut_amd/frame3d.py, strut_amd/constraints.py): the cases the reference's
loader refuses are refused here too (loader.mojo:2230-2372: a retained node
that is itself constrained, a DOF constrained twice; input_types.mojo: the
drift recorder's denominator), and cases outside the 3D fiber path raise
instead of running differently."""

import copy

import numpy as np
import pytest

from strut_amd import cases
from strut_amd.constraints import mpc_tables, rigid_diaphragm
from strut_amd.frame3d import load_frame3d_case


def _case():
    return cases.rc_frame3d_case(1, 1, 1, steps=10, n_core=4, n_bars=4)


def test_tables_shape_and_rows():
    fc = load_frame3d_case(_case())
    n = fc.total_dofs
    assert len(fc.row_offsets) == n + 1 and fc.row_offsets[0] == 0
    rows = np.diff(fc.row_offsets)
    assert np.all(rows[fc.sp.astype(bool) & ~fc.slave] == 0)   # fixed DOFs expand to nothing
    assert np.all(rows[fc.slave] == 3)                          # diaphragm rows: ux, uy, rz of the master
    free = ~fc.constrained
    assert np.all(rows[free] == 1)
    assert np.array_equal(fc.dof_pool[fc.row_offsets[:-1][free]], np.flatnonzero(free))


def test_retained_node_that_is_constrained_is_refused():
    coords = {1: (0, 0, 0), 2: (1, 0, 0), 3: (2, 0, 0)}
    mps = rigid_diaphragm(coords, 1, [2], 3) + rigid_diaphragm(coords, 2, [3], 3)
    with pytest.raises(ValueError, match="retained_node cannot also be constrained"):
        mpc_tables(3, 6, np.zeros(18, dtype=bool), mps, {1: 0, 2: 1, 3: 2})


def test_dof_constrained_twice_is_refused():
    coords = {1: (0, 0, 0), 2: (1, 0, 0), 3: (2, 0, 0)}
    mps = rigid_diaphragm(coords, 1, [3], 3) + rigid_diaphragm(coords, 2, [3], 3)
    with pytest.raises(ValueError, match="already controlled"):
        mpc_tables(3, 6, np.zeros(18, dtype=bool), mps, {1: 0, 2: 1, 3: 2})


def test_unsupported_inputs_raise():
    c = _case()
    bad = copy.deepcopy(c)
    bad["elements"][0]["type"] = "elasticBeamColumn3d"
    with pytest.raises(ValueError, match="outside the 3D fiber-frame path"):
        load_frame3d_case(bad)
    bad = copy.deepcopy(c)
    bad["elements"][0]["geomTransf"] = "Corotational"
    with pytest.raises(ValueError, match="geomTransf"):
        load_frame3d_case(bad)
    bad = copy.deepcopy(c)
    bad["analysis"]["constraints"] = "Plain"
    with pytest.raises(ValueError, match="Transformation"):
        load_frame3d_case(bad)
    bad = copy.deepcopy(c)
    bad["patterns"] = bad["patterns"] * 2
    with pytest.raises(ValueError, match="at most three"):
        load_frame3d_case(bad)
    bad = copy.deepcopy(c)
    bad["recorders"].append({"type": "drift", "i_node": 1, "j_node": 2, "dof": 1, "perp_dirn": 3})  # same z
    with pytest.raises(ValueError, match="denominator"):
        load_frame3d_case(bad)
    bad = copy.deepcopy(c)
    bad["analysis"]["solver_chain"] = [{"algorithm": "Newton"}, {"algorithm": "ModifiedNewton"}]
    with pytest.raises(ValueError, match="solver_chain"):
        load_frame3d_case(bad)


def test_2d_loader_refuses_a_solver_chain():
    from strut_amd import cases
    from strut_amd.frame2d import load_frame2d_case

    c = cases.config0_case(steps=100)
    c["analysis"]["solver_chain"] = [{"algorithm": "Newton"}]
    with pytest.raises(ValueError, match="solver_chain"):
        load_frame2d_case(c)


def test_one_pattern_is_the_reference_schema():
    c = _case()
    one = copy.deepcopy(c)
    one["pattern"] = one.pop("patterns")[0]
    fc = load_frame3d_case(one)
    assert len(fc.patterns) == 1 and fc.patterns[0].direction == 1
