"""Host side of the tile-skip network (proton_solver/tileskip.py), no GPU.

The weights file round-trips through numpy without pickling, malformed
weights are refused, the initial parameters have the layout the C ABI
declares, and the committed weights (when present) are finite float32 of
that size.
"""

import numpy as np
import pytest

from proton_solver import _native
from proton_solver import tileskip as TS


def test_parameter_layout_matches_the_library():
    p = TS.init_params(3)
    assert p.dtype == np.float32
    assert p.shape == (_native.load_library().m1_mlp_param_count(),)
    # biases start at zero, weights do not
    for a, b in [(704, 768), (4864, 4928), (9024, 9088), (9152, 9153)]:
        assert not p[a:b].any()
    assert np.std(p[:704]) > 0.1
    assert len(TS.FEATURES) == 11


def test_weights_round_trip(tmp_path):
    p = TS.init_params(5)
    f = tmp_path / "w.npy"
    TS.save_params(p, f)
    q = TS.load_params(f)
    assert q.dtype == np.float32 and np.array_equal(p, q)


def test_malformed_weights_are_refused(tmp_path):
    f = tmp_path / "bad.npy"
    np.save(f, np.zeros(10, dtype=np.float32))
    with pytest.raises(ValueError):
        TS.load_params(f)
    np.save(f, np.zeros(9153, dtype=np.float64))
    with pytest.raises(ValueError):
        TS.load_params(f)
    np.save(f, np.array([{"a": 1}], dtype=object), allow_pickle=True)
    with pytest.raises(ValueError):        # object arrays need unpickling: refused
        TS.load_params(f)


def test_missing_weights_name_the_training_command(tmp_path):
    with pytest.raises(FileNotFoundError, match="proton_solver.tileskip"):
        TS.load_model(0, tmp_path / "none.npy")


@pytest.mark.skipif(not TS.WEIGHTS_FILE.exists(), reason="no trained weights committed")
def test_committed_weights_are_well_formed():
    p = TS.load_params()
    assert np.isfinite(p).all()
    assert np.abs(p).max() < 100.0


def test_skip_options_defaults():
    o = TS.SkipOptions()
    assert 0.0 < o.decision < 1.0 and 1 <= o.horizon <= 255
    assert 1 <= o.infer_every <= o.horizon
    assert o.judge_every >= 1 and 0.0 < o.judge_tol <= 1e-3
