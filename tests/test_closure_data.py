"""The committed pair-approach dataset (lbcuda_amd/data/closure_pairs.npz,
made on an MI355X by `python -m lbcuda_amd.closure generate`): its format
and the physics it must show.  No GPU."""
import numpy as np

from lbcuda_amd import closure as C


def test_format():
    z = np.load(C.DATA)
    d = z["data"]
    assert list(z["cols"]) == C.COLS and d.shape[1] == len(C.COLS)
    assert np.all(np.isfinite(d)) and len(d) > 100
    assert np.abs(d[:, 4:10]).max() <= C.OUTLIER
    for run in np.unique(d[:, 10]):
        b = d[d[:, 10] == run, 11]
        assert np.all(np.diff(b) > 0)           # one row per gap bin, ascending
    # gaps (in coarse units) fall inside their bins
    g = d[:, 0] * d[:, 3]
    lo = C.GAP_TRAIN_MIN + C.GAP_BIN * d[:, 11]
    assert np.all((g >= lo - 1e-9) & (g <= lo + C.GAP_BIN + 1e-9))


def test_resolved_head_on_force_grows_as_gap_closes():
    """Lubrication: the 4x-resolved normal resistance of an approaching
    pair rises monotonically as the gap closes, and is repulsive."""
    d = C.load()
    for run in np.unique(d[:, 10]):
        r = d[d[:, 10] == run]
        if r[0, 2] > 1e-9:          # oblique runs: v_t > 0
            continue
        cn = r[np.argsort(r[:, 0]), 7]
        assert np.all(cn > 0)
        assert np.all(np.diff(cn) < 2e-2)       # non-increasing with gap, up to bin noise
        assert cn[0] > 1.2 * cn[-1]


def test_split_holds_out_every_third_bin():
    d = C.load()
    tr, ho = C.split(d)
    assert len(tr) + len(ho) == len(d)
    assert abs(len(ho) - len(d) / 3) < 7
