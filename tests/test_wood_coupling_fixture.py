"""The wood-coupling fixture on the CPU: its default scene is the one of
js/main.js:132-155 (16 pieces: seven all-char coals on the floor, two 5 cm
logs, seven 2.5 cm sticks), and the FP32 oracle replaying the recorded pieces
reproduces the recorded readbacks exactly (the gas side of the fixture is
deterministic, so a GPU replay compares against the oracle's frames)."""
import numpy as np

from oracle.pyoracle import Oracle
from wood_fixture import default_scene, load, replay


def test_default_scene_is_the_reference_recipe():
    p, n = default_scene()
    assert n == 16
    P = p.reshape(64, 20)[:n]
    assert np.all(P[:, 7] == 1)                         # alive
    coals, logs, sticks = P[:7], P[7:9], P[9:]
    assert np.allclose(coals[:, 8:12], 1050)            # preheated char coals
    assert np.all(coals[:, 16:20] == 1)                 # full char cover
    assert np.allclose(coals[:, 1], coals[:, 3], rtol=0, atol=2e-3)   # resting on the floor
    assert np.allclose(logs[:, 8:12], 880) and np.allclose(sticks[:, 8:12], 880)
    # sides after the char layer (0.3 x density, 25 % char yield) shrank them
    assert np.all((2 * logs[:, 3] > 0.047) & (2 * logs[:, 3] < 0.05))
    assert np.all((2 * sticks[:, 3] > 0.023) & (2 * sticks[:, 3] < 0.025))
    assert abs(float(logs[0, 0] + logs[1, 0]) / 2 - 0.35) < 1e-6       # symmetric about the centre
    assert np.all(P[:, 12:16] == 0)                     # no pyrolysis gas before the first wood step


def test_oracle_replay_reproduces_recorded_frames():
    d = load()
    o = Oracle(int(d["nx"]), int(d["ny"]), float(d["dx"]), float(d["dt"]))
    for k, v in d["params"].items():
        o.set_param(k, v)
    o.reset()
    frames = 6
    got = replay(o, d, frames)
    o.close()
    for f, (acc, rad, fw) in enumerate(got):
        assert np.array_equal(acc, d["acc"][f]), f
        assert np.array_equal(rad, d["rad"][f]), f
        assert np.array_equal(fw, d["faceW"][f]), f
    # the fire is burning and the wood sees it
    assert d["acc"][-1][512] > 0 and int(d["acc"][-1][514]) > 1500
    assert np.all(d["n"] == 16)
