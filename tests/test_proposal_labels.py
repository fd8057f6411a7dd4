"""Label files of the proposal pipeline: rollouts split over several runs are
reassembled in set order, with the features rebuilt from the set indices; gaps and
misnamed files are refused.  CPU only."""
import os

import numpy as np
import pytest

from lenia_amd import proposal as P


def _write(d, lo, hi):
    n = hi - lo
    P.save_labels(os.path.join(d, f"labels_{lo}_{hi}.npz"),
                  dict(survival=np.arange(lo, hi, dtype=np.float32), concentration=np.full(n, 0.5, np.float32),
                       drift=np.zeros(n, np.float32)))


def test_labels_reassembled_in_set_order(tmp_path):
    _write(tmp_path, 6, 16)
    _write(tmp_path, 0, 6)
    files = sorted(str(p) for p in tmp_path.iterdir())
    r = P.load_labels(files, 100, 12)
    assert np.array_equal(r["survival"], np.arange(12, dtype=np.float32))
    assert np.array_equal(r["features"], P.features(P.sample(100, np.arange(12))))


def test_labels_gaps_and_names_refused(tmp_path):
    _write(tmp_path, 0, 4)
    _write(tmp_path, 5, 9)
    files = sorted(str(p) for p in tmp_path.iterdir())
    with pytest.raises(ValueError, match="uncovered"):
        P.load_labels(files, 100, 8)
    with pytest.raises(ValueError, match="cover 4 of 8"):
        P.load_labels(files[:1], 100, 8)
    bad = tmp_path / "other.npz"
    bad.write_bytes(b"")
    with pytest.raises(ValueError, match="labels_<lo>_<hi>"):
        P.load_labels([str(bad)], 100, 4)
