"""GPU parity of the Transformation-handler collapses (strut_amd/csrc/mpc.hip)
against the oracle (oracle/mpc.c, helpers.mojo:3288-3397) through the C ABI.
The device gathers each output entry from the terms that reach it in the
reference's addition order, so every result is the oracle's bits: a 3D frame
with rigid floor diaphragms (6 DOFs per node, fixed base, retained nodes'
uz/rx/ry fixed), an equalDOF model, and a larger multi-story frame."""

import numpy as np
import pytest

from oracle import oracle
from strut_amd import cases
from strut_amd.constraints import TransformationHandler, mpc_tables
from strut_amd.model import MaterialTable
from strut_amd.sections import SectionRuntime

pytestmark = pytest.mark.gpu


def _runtime():
    return SectionRuntime(MaterialTable([{"id": 1, "type": "Elastic", "params": {"E": 1.0}}]), device=0)


def _check(rt, tables, slave, constrained, K, F, u):
    ro, dp, cp = tables
    h = TransformationHandler(rt)
    h.set(ro, dp, cp, constrained, slave)
    np.testing.assert_array_equal(h.collapse_matrix(K), oracle.mpc_collapse_matrix(K, tables))
    np.testing.assert_array_equal(h.collapse_vector(F), oracle.mpc_collapse_vector(F, tables))
    np.testing.assert_array_equal(h.enforce(u), oracle.mpc_enforce(u, tables, constrained, slave))


@pytest.mark.parametrize("stories,bays_x,bays_y", [(2, 2, 1), (6, 3, 2)])
def test_gpu_diaphragm_frame_collapse_is_oracle_bits(stories, bays_x, bays_y):
    m = cases.frame3d_diaphragm_model(stories=stories, bays_x=bays_x, bays_y=bays_y)
    ro, dp, cp, slave, constrained = m["tables"]
    K = cases.synthetic_frame_stiffness(m, seed=stories)
    rng = np.random.default_rng(stories + 1)
    F = rng.normal(size=K.shape[0])
    u = rng.normal(size=K.shape[0])
    rt = _runtime()
    print(f"diaphragm frame {stories}x{bays_x}x{bays_y}: n={K.shape[0]}", flush=True)
    _check(rt, (ro, dp, cp), slave, constrained, K, F, u)
    rt.close()


def test_gpu_equaldof_collapse_is_oracle_bits():
    # a 2D frame (3 DOFs per node): two floors whose nodes share ux by
    # equalDOF with the left node, the base fixed
    coords = {i + 1: (float(i % 3), float(i // 3), 0.0) for i in range(9)}
    node_index = {n: n - 1 for n in coords}
    sp = np.zeros(27, dtype=bool)
    sp[:9] = True
    mps = [{"type": "equalDOF", "retained_node": r, "constrained_node": c, "dofs": [1]}
           for r, cs in ((4, (5, 6)), (7, (8, 9))) for c in cs]
    ro, dp, cp, slave, constrained = mpc_tables(9, 3, sp, mps, node_index)
    rng = np.random.default_rng(9)
    a = rng.normal(size=(27, 27))
    K = a @ a.T
    K[np.abs(K) < 0.5] = 0.0  # zero entries the collapse must skip
    F, u = rng.normal(size=27), rng.normal(size=27)
    rt = _runtime()
    _check(rt, (ro, dp, cp), slave, constrained, K, F, u)
    rt.close()


def test_gpu_mpc_rejects_a_slave_retaining_a_slave():
    from strut_amd._native import StrutGpuError

    ro = np.array([0, 1, 2, 3], dtype=np.int32)
    dp = np.array([0, 2, 1], dtype=np.int32)
    cp = np.ones(3)
    slave = np.array([0, 1, 1], dtype=bool)
    rt = _runtime()
    h = TransformationHandler(rt)
    with pytest.raises(StrutGpuError, match="retains slave"):
        h.set(ro, dp, cp, slave, slave)
    rt.close()
