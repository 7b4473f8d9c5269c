"""Case-schema host logic: material normalisation (loader.mojo:915-1096)
and fiber discretisation (fiber2d.mojo:3504-3878) against closed forms."""

import math

import numpy as np
import pytest

from strut_amd import cases
from strut_amd.model import (CONCRETE02, STEEL02, MaterialTable, discretize_section, make_uniaxial_def)


def test_concrete02_defaults_and_sign_normalisation():
    d = make_uniaxial_def({"type": "Concrete02", "params": {"fpc": 30.0, "epsc0": 0.002, "fpcu": 5.0,
                                                             "epscu": 0.01}})
    assert d.mat_type == CONCRETE02
    fpc, epsc0, fpcu, epscu, rat, ft, ets = d.p[:7]
    assert (fpc, epsc0, fpcu, epscu) == (-30.0, -0.002, -5.0, -0.01)
    assert rat == 0.1 and ft == 3.0 and ets == pytest.approx(0.1 * -30.0 / -0.002)


def test_steel02_defaults_and_group_rules():
    d = make_uniaxial_def({"type": "Steel02", "params": {"Fy": 400.0, "E0": 2.0e5, "b": 0.02}})
    assert d.mat_type == STEEL02
    assert d.p[3:11] == (15.0, 0.925, 0.15, 0.0, 1.0, 0.0, 1.0, 0.0)
    with pytest.raises(ValueError):
        make_uniaxial_def({"type": "Steel02", "params": {"Fy": 400.0, "E0": 2.0e5, "b": 0.02, "R0": 18.0}})
    with pytest.raises(ValueError):
        make_uniaxial_def({"type": "Steel02", "params": {"Fy": 400.0, "E0": 2.0e5, "b": 1.0}})
    with pytest.raises(ValueError):
        make_uniaxial_def({"type": "Concrete02", "params": {"fpc": -30.0, "epsc0": -0.002, "fpcu": -40.0,
                                                           "epscu": -0.01}})


def test_rect_and_quadr_patches_integrate_exactly():
    mats = MaterialTable([{"id": 1, "type": "Elastic", "params": {"E": 1.0}}])
    rect = {"type": "rect", "material": 1, "num_subdiv_y": 6, "num_subdiv_z": 4, "y_i": 0.3, "z_i": 0.2,
            "y_j": -0.3, "z_j": -0.2}
    cells = discretize_section({"params": {"patches": [rect]}}, mats)
    assert len(cells) == 24
    assert cells["area"].sum() == pytest.approx(0.6 * 0.4, rel=1e-14)
    assert np.sum(cells["area"] * cells["y"] ** 2) == pytest.approx(0.4 * 0.6 ** 3 / 12 * (1 - 1 / 36), rel=1e-12)
    quad = {"type": "quadr", "material": 1, "num_subdiv_y": 4, "num_subdiv_z": 3, "y_i": -0.4, "z_i": 0.2,
            "y_j": -0.2, "z_j": -0.2, "y_k": 0.4, "z_k": -0.1, "y_l": 0.2, "z_l": 0.3}
    qc = discretize_section({"params": {"patches": [quad]}}, mats)
    poly = np.array([[-0.4, 0.2], [-0.2, -0.2], [0.4, -0.1], [0.2, 0.3]])
    shoelace = 0.5 * abs(np.dot(poly[:, 0], np.roll(poly[:, 1], -1)) - np.dot(poly[:, 1], np.roll(poly[:, 0], -1)))
    assert qc["area"].sum() == pytest.approx(shoelace, rel=1e-12)


def test_straight_layer_positions():
    mats = MaterialTable([{"id": 1, "type": "Elastic", "params": {"E": 1.0}}])
    lay = {"type": "straight", "material": 1, "num_bars": 4, "bar_area": 1e-4, "y_start": -0.3, "z_start": 0.0,
           "y_end": 0.3, "z_end": 0.6}
    c = discretize_section({"params": {"layers": [lay]}}, mats)
    np.testing.assert_allclose(c["y"], [-0.3, -0.1, 0.1, 0.3], atol=1e-15)
    one = dict(lay, num_bars=1)
    c1 = discretize_section({"params": {"layers": [one]}}, mats)
    assert c1["y"][0] == 0.0 and c1["z"][0] == 0.3


def test_rc_section_counts():
    mats = MaterialTable(cases.rc_materials())
    col = discretize_section(cases.rc_section(1, dim=2, **cases.FRAME2D_COLUMN), mats)
    assert len(col) == 100 + 2 * 2 * 12 + 2 * 12 * 2 + 8
    col3 = discretize_section(cases.rc_section(1, dim=3, **cases.FRAME3D_COLUMN), mats)
    assert len(col3) == 144 + 4 * 2 * 14 + 12
    gross = col["area"][col["mat"] != 2].sum()
    assert gross == pytest.approx(0.4 * 0.4, rel=1e-12)
    pop = cases.frame3d_section_population()
    assert pop["columns"] == 720 and pop["beams"] == 1200
