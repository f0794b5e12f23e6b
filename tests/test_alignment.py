"""Per-dataset shifts and the alignment-file override of an [[input]]
dataset (include/config.h:284-343), in the Python drivers and the C++ API.
Host logic only: runs without a GPU."""
import math

import numpy as np
import pytest

from laminocam_amd.drivers import Dataset, apply_alignment, broadcast_shifts, load_alignment


def _ds(n=4, shifts=None, gamma=-math.pi / 6, beta=0.0):
    return Dataset(projs=np.zeros((n, 2, 3), np.float32),
                   angles=np.arange(n) * 2 * math.pi / n, gamma=gamma, beta=beta, shifts=shifts)


def test_broadcast_shifts():
    assert np.array_equal(broadcast_shifts(None, 3), np.zeros((3, 2)))
    assert np.array_equal(broadcast_shifts([[1.5, -2.0]], 3), [[1.5, -2.0]] * 3)   # cor-offset
    s = np.arange(8.0).reshape(4, 2)
    assert np.array_equal(broadcast_shifts(s, 4), s)
    with pytest.raises(ValueError, match="3 entries but 4 projections"):
        broadcast_shifts(np.zeros((3, 2)), 4)


def test_alignment_overrides_angles_tilts_and_shifts():
    t = {"gamma_deg": 29.5, "beta_deg": -0.25, "corrected_angles_deg": [0.0, 91.0, 179.5],
         "shifts_px": [[0.5, 1.0], [-1.0, 0.0], [2.0, -2.5]]}
    out = apply_alignment(_ds(4), t)
    # the parser's sign flip on gamma (include/config.h:206, :311)
    assert out.gamma == -29.5 * math.pi / 180 and out.beta == -0.25 * math.pi / 180
    assert np.array_equal(out.angles, np.array([0.0, 91.0, 179.5]) * math.pi / 180)
    assert np.array_equal(out.shifts, [[0.5, 1.0], [-1.0, 0.0], [2.0, -2.5]])
    # no tilts in the table: the dataset's are kept
    out = apply_alignment(_ds(2, gamma=0.3, beta=0.1),
                          {"corrected_angles_deg": [1.0, 2.0], "shifts_px": [[0, 0], [1, 1]]})
    assert (out.gamma, out.beta) == (0.3, 0.1)


def test_alignment_keeps_broadcast_shifts_when_none_given():
    out = apply_alignment(_ds(3, shifts=[[0.25, -0.5]]), {"corrected_angles_deg": [0, 120, 240]})
    assert np.array_equal(out.shifts, [[0.25, -0.5]] * 3)
    # corrected angles of another count need their own shifts
    with pytest.raises(ValueError, match="shifts_px has 3 entries but corrected_angles_deg has 2"):
        apply_alignment(_ds(3), {"corrected_angles_deg": [0, 120]})


def test_alignment_errors():
    with pytest.raises(ValueError, match="missing 'corrected_angles_deg'"):
        apply_alignment(_ds(2), {"gamma_deg": 30.0})
    with pytest.raises(ValueError, match="pairs"):
        apply_alignment(_ds(2), {"corrected_angles_deg": [0, 1], "shifts_px": [[0.0], [1.0, 2.0]]})


def test_load_alignment(tmp_path):
    p = tmp_path / "align.toml"
    p.write_text('gamma_deg = 30.0\ncorrected_angles_deg = [0.0, 90.0]\n'
                 'shifts_px = [[1.0, 2.0], [3.0, 4.0]]\n')
    out = apply_alignment(_ds(2), load_alignment(str(p)))
    assert out.gamma == -30.0 * math.pi / 180 and np.array_equal(out.shifts, [[1, 2], [3, 4]])
