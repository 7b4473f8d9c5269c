"""The learned Newton predictor in the device transient (a reduced form of
the learned-predictor configuration): the 4-story x 3-bay frame, a network
trained on the device from trajectories the solver records on one seed
range, then held-out records of another seed range run with the plain
Newmark predictor and with the learned one.

Checked: the learned runs stop where the plain runs stop, use fewer Newton
iterations in total, and on records without a stalled Newton step their peak
inter-story drifts stay within 1e-6 relative of the plain FP64 solution;
proposals replaced by the Newmark predictor are counted.

A stalled step (more than 10 Newton iterations) converges to a point that
depends on where Newton starts: on held-out seed 9007 step 374 takes 19
iterations in the oracle, and the oracle itself, started 10% closer to the
converged displacement, moves that step's solution by 1.7e-7 relative and
the record's peak drift by 1.8e-6.  Such records are reported, not held to
the 1e-6 bar."""

import numpy as np
import pytest

from strut_amd import _native, cases
from strut_amd.frame2d import load_frame2d_case
from strut_amd.learned import train_predictor
from strut_amd.transient import FrameBatch

pytestmark = pytest.mark.gpu

STEPS = 400


def _case(frame: dict, seed: int) -> object:
    m = cases.config2_motion(seed, max_steps=STEPS)
    c = dict(frame)
    c["time_series"] = [{"type": "Path", "tag": 1, "dt": m["dt"], "factor": 1.0,
                         "values": [float(v) for v in m["values"]]}]
    c["analysis"] = dict(frame["analysis"], steps=m["steps"])
    return load_frame2d_case(c)


def _batch(fcs, record=False, net=None):
    fb = FrameBatch(fcs[0].materials)
    fid = fb.add_frame(fcs[0], record_all=True)
    ids = [fb.add_case_motion(fid, fc) for fc in fcs]
    if record:
        fb.record_samples(True)
    fb.finalize()
    if net is not None:
        fb.set_predictor(fid, net)
    fb.set_schedule(_native.SCHEDULE_ROUNDS)
    fb.run(rounds=16)
    return fb, fid, ids


def test_gpu_learned_predictor_gain_and_drift():
    frame = cases.config2_frame_case(4, 3, steps=100)
    pga = lambda s: cases.config2_motion_params(s)["pga_g"]
    train = [_case(frame, s) for s in range(7000, 7400) if pga(s) <= 0.6][:64]
    held = [_case(frame, s) for s in range(9000, 9100) if pga(s) <= 0.6][:6]
    fb, fid, _ = _batch(train, record=True)
    x, y = fb.samples(fid)
    print("training pairs", x.shape[0], flush=True)
    assert x.shape[1] == 5 * fb.info(fid)["free"] + 2
    net, losses = train_predictor(x, y, seed=1, epochs=40, batch=256, lr=1e-3)
    print("loss", losses[0], "->", losses[-1], flush=True)
    assert losses[-1] < 0.5 * losses[0]
    fb.close()
    plain, _, pid = _batch(held)
    learned, _, lid = _batch(held, net=net)
    it_p = it_l = 0
    held_to_bar = 0
    for k, (a, b) in enumerate(zip(pid, lid)):
        rp, rl = plain.result(a), learned.result(b)
        sl = learned.inst_stats(b)
        assert rp["status"] == rl["status"] and rp["steps_done"] == rl["steps_done"]
        n = rp["steps_done"]
        it_p += int(rp["iters"][:n].sum())
        it_l += int(rl["iters"][:n].sum())
        dp = cases.peak_story_drift(rp["u"][:n], 4, 3)
        dl = cases.peak_story_drift(rl["u"][:n], 4, 3)
        err = float(np.max(np.abs(dl - dp) / dp))
        stalled = int(rp["iters"][:n].max()) > 10
        print(f"record {k}: iterations plain {rp['iters'][:n].sum()} learned {rl['iters'][:n].sum()} "
              f"proposals {sl['proposals']} fallbacks {sl['fallbacks']} drift rel err {err:.2e}"
              f"{' (stalled step: reported only)' if stalled else ''}")
        assert sl["proposals"] >= n and sl["fallbacks"] < sl["proposals"]
        if not stalled:
            assert err <= 1e-6
            held_to_bar += 1
    print(f"total iterations plain {it_p} learned {it_l}: {100.0 * (1 - it_l / it_p):.1f}% fewer")
    assert it_l < it_p
    assert held_to_bar >= 4
    plain.close()
    learned.close()
    net.close()
