"""Decision tape of the FP64 differentiable form (pic_tape): a run that
replays the integer decisions of a recorded run stays on that run's smooth
piece of the step, whose derivative is what pic_adjoint computes.

* Replaying a run's own tape from the same state reproduces it bit for bit
  (the decisions it reads are the ones it would have made).
* Config 3 (64x16x16 cells, 32 ppc, 200 steps, checkpoint every 20, loss =
  field energy over steps 150..200, theta = beam drift and a_1..a_8 = 1e-3):
  every parameter's adjoint gradient within 1e-4 of the central difference
  at step 1e-5 of the forwards that replay the base run's tape -- the stated
  bar, against the map the adjoint differentiates -- both with the
  reference's periodic ghosts (ghost_mode 0, whose y-ghost layers of Ex and
  Ez are never updated: discontinuous where a particle crosses a y face,
  smooth on every piece the tape freezes) and with ghost_mode 1.
* The refusals: FP32 form, a context that has stepped, replay without a tape.
"""
import importlib

import numpy as np
import pytest

from helpers import load_workload

pytestmark = pytest.mark.gpu


def _ctx(picmod, parts, prec=0):
    prm = picmod.make_params((32, 8, 8), (3.2, 0.8, 0.8), 0.05, precision=prec,
                             differentiable=1, ghost_mode=1)
    g = picmod.Pic(prm)
    load_workload(g, parts)
    return g


def test_replay_of_own_tape_is_bitwise(picmod):
    _, parts = picmod.generate_workload(0)
    a = _ctx(picmod, parts)
    a.tape(a.TAPE_RECORD, 40)
    Wa = a.run(1, 40)
    words, cells = a.tape_get(40)
    assert words.shape == (40, a.id_count()) and (cells >= 0).all()
    assert ((words >> np.uint64(41)) & np.uint64(1)).all()           # every word recorded
    assert not ((words >> np.uint64(42)) & np.uint64(1)).any()       # and replayable
    b = _ctx(picmod, parts)
    b.tape_set(words, cells)
    b.tape(b.TAPE_REPLAY, 40)
    Wb = b.run(1, 40)
    assert np.array_equal(Wa, Wb)
    for i in range(6):
        assert np.array_equal(a.get_field(i), b.get_field(i))
    assert np.array_equal(a.cell_counts(), b.cell_counts())
    # after the tape's steps the context runs free again
    assert np.array_equal(a.run(41, 5), b.run(41, 5))
    a.close()
    b.close()


def test_tape_refusals(picmod):
    _, parts = picmod.generate_workload(0, scale_n=(8, 4, 4))
    g32 = _ctx(picmod, parts, prec=1)
    with pytest.raises(picmod.PicError) as e:
        g32.tape(g32.TAPE_RECORD, 5)
    assert e.value.code == 1
    g32.close()
    g = _ctx(picmod, parts)
    with pytest.raises(picmod.PicError) as e:
        g.tape(g.TAPE_REPLAY, 5)                      # no tape set
    assert e.value.code == 7
    g.run(1, 1)
    with pytest.raises(picmod.PicError) as e:
        g.tape(g.TAPE_RECORD, 5)                      # already stepped
    assert e.value.code == 7
    g.close()


@pytest.mark.parametrize("ghost_mode", [0, 1], ids=["reference-ghosts", "ghost-mode-1"])
def test_config3_gradient_against_frozen_finite_differences(picmod, ghost_mode):
    pytest.importorskip("torch")
    ag = importlib.import_module("0buf_amd.autograd")
    r = ag.fd_check_frozen(3, (1e-5,), ghost_mode=ghost_mode)
    assert r["loss_replay"] == r["loss"]
    g, fd = r["grad"], r["fd"][1e-5]
    rel = np.abs(g - fd) / np.abs(fd)
    print("frozen decisions, step 1e-5: rel " + " ".join(f"{v:.1e}" for v in rel) +
          f"  normwise {np.linalg.norm(g - fd) / np.linalg.norm(fd):.2e}", flush=True)
    assert rel.max() <= 1e-4, rel
