"""sgpu_section_eval_device_noalloc (fixed device scratch, no allocation)
against sgpu_section_eval_device (library-allocated scratch), both through
the C ABI on device-resident points and results: two calls in a row, the
second with fewer points, must give the same bits (the same kernels run on
the same committed states), and a call beyond the fixed scratch is refused
before any launch."""

import ctypes

import numpy as np
import pytest

from strut_amd import _native, cases
from strut_amd._native import POINT_DTYPE, RESULT_DTYPE
from strut_amd.model import MaterialTable, discretize_section
from strut_amd.sections import SectionRuntime

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 1 << 20


def _runtime(n_inst):
    mats = MaterialTable(cases.rc_materials())
    rt = SectionRuntime(mats, device=0)
    defs = []
    for dim, shapes in ((2, (cases.FRAME2D_COLUMN, cases.FRAME2D_BEAM)),
                        (3, (cases.FRAME3D_COLUMN, cases.FRAME3D_BEAM))):
        for i, s in enumerate(shapes):
            cells = discretize_section(cases.rc_section(i + 1, dim=dim, **s), mats)
            d = rt.add_definition(dim, cells)
            rt.alloc_instances(d, n_inst)
            defs.append(d)
    rt.finalize()
    return rt, defs


def _points(rt, defs, n_inst, step, count, seed):
    did = np.repeat(defs, n_inst)
    inst = np.tile(np.arange(n_inst), len(defs))
    e, kz, ky = cases.synthetic_section_deformation(len(did), step, seed=seed, dim=3)
    pts = rt.points(did, inst, e, kz, ky)
    return np.ascontiguousarray(pts[:count])


def _eval(rt, fn, d_pts, d_res, n):
    import torch

    rc = fn(rt.ctx, n, ctypes.c_void_p(d_pts.data_ptr()), ctypes.c_void_p(d_res.data_ptr()))
    assert rc == 0, rt.lib.sgpu_last_error(rt.ctx)
    assert rt.lib.sgpu_synchronize(rt.ctx) == 0
    torch.cuda.synchronize()
    return d_res.cpu().numpy().view(RESULT_DTYPE).copy()


def test_noalloc_section_eval_matches_allocating_entry_point(gpu_available):
    import torch

    n_inst = 24
    rt, defs = _runtime(n_inst)
    lib = rt.lib
    dev = torch.device("cuda", 0)
    # a committed history first, so the laws run off their initial branches
    for step in range(4):
        ok, _ = rt.eval(_points(rt, defs, n_inst, step, len(defs) * n_inst, seed=11))
        assert ok
        rt.commit_all()
    for call, count in enumerate((len(defs) * n_inst, 37)):
        pts = _points(rt, defs, n_inst, 10 + call, count, seed=11)
        d_pts = torch.from_numpy(pts.view(np.uint8).copy()).to(dev)
        d_res_a = torch.zeros(count * RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        d_res_b = torch.full((count * RESULT_DTYPE.itemsize,), 255, dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        want = _eval(rt, lib.sgpu_section_eval_device, d_pts, d_res_a, count)
        trial_a = [rt.fiber_state(d, 0) for d in defs]
        got = _eval(rt, lib.sgpu_section_eval_device_noalloc, d_pts, d_res_b, count)
        trial_b = [rt.fiber_state(d, 0) for d in defs]
        assert np.all(want["valid"] == 1)
        for field in ("s", "k", "f", "valid"):
            assert np.array_equal(got[field], want[field]), (call, field)
        # and the trial fiber states it leaves are the same
        for a, b in zip(trial_a, trial_b):
            assert np.array_equal(a, b)
    rt.close()


def test_noalloc_section_eval_refuses_beyond_scratch(gpu_available):
    import torch

    n_inst = 4
    rt, defs = _runtime(n_inst)
    too_many = SCRATCH_DOUBLES // 2 + 1  # beyond the scratch for any definition with an aux slot
    dev = torch.device("cuda", 0)
    d_pts = torch.zeros(too_many * POINT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    d_res = torch.zeros(too_many * RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    rc = rt.lib.sgpu_section_eval_device_noalloc(rt.ctx, too_many, ctypes.c_void_p(d_pts.data_ptr()),
                                                 ctypes.c_void_p(d_res.data_ptr()))
    assert rc == -1
    assert b"fixed scratch" in rt.lib.sgpu_last_error(rt.ctx)
    rt.close()
