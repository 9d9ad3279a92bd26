"""Crop and display orientation: the oracle (oracle/orient_oracle.py) pinned
by hand-written cases of the eight EXIF symmetries (mov.rs:121-146) and by
their group laws; the library's host functions (prr_oriented_size,
prr_oriented_bayer) against it. No GPU needed."""
import itertools

import numpy as np
import pytest

from oracle import orient_oracle as O
from proresraw_rs_amd import orient as R

A = np.array([[1, 2, 3],
              [4, 5, 6]])

# What each orientation shows for the stored 2x3 picture A, written by hand
# from the names: mirrored left to right, turned 180 degrees, mirrored top to
# bottom, mirrored about the main diagonal, turned 90 degrees clockwise (the
# stored left column becomes the top row, read bottom to top), mirrored
# about the anti-diagonal, turned 90 degrees anticlockwise.
SHOWN = {
    1: [[1, 2, 3], [4, 5, 6]],
    2: [[3, 2, 1], [6, 5, 4]],
    3: [[6, 5, 4], [3, 2, 1]],
    4: [[4, 5, 6], [1, 2, 3]],
    5: [[1, 4], [2, 5], [3, 6]],
    6: [[4, 1], [5, 2], [6, 3]],
    7: [[6, 3], [5, 2], [4, 1]],
    8: [[3, 6], [2, 5], [1, 4]],
}


def test_each_orientation_by_hand():
    for o, want in SHOWN.items():
        assert O.orient(A, o).tolist() == want, o


def test_group_laws():
    rng = np.random.default_rng(1)
    a = rng.integers(0, 100, (5, 7))
    rot90, rot180, rot270 = 6, 3, 8
    twice = lambda o: O.orient(O.orient(a, o), o)  # noqa: E731
    for o in (1, 2, 3, 4, 5, 7):
        assert np.array_equal(twice(o), a), o   # involutions
    assert np.array_equal(twice(rot90), O.orient(a, rot180))
    assert np.array_equal(O.orient(O.orient(a, rot90), rot270), a)
    assert np.array_equal(O.orient(O.orient(a, 2), 4), O.orient(a, rot180))
    assert np.array_equal(O.orient(O.orient(a, 5), rot180), O.orient(a, 7))
    for o in range(1, 9):
        assert O.orient(a, o).shape == ((7, 5) if o >= 5 else (5, 7))


def test_crop_margins():
    a = np.arange(10 * 12).reshape(10, 12)
    assert np.array_equal(O.crop(a, (1, 2, 3, 4)), a[3:6, 1:10])
    # Margins that leave nothing are not applied.
    assert np.array_equal(O.crop(a, (6, 6, 0, 0)), a)
    assert np.array_equal(O.crop(a, (0, 0, 5, 5)), a)


def test_sizes_and_patterns_match_the_oracle():
    for (w, h), margins, o in itertools.product(
            [(12, 10), (7, 5), (4096, 2160), (2, 2)], [(0, 0, 0, 0), (1, 2, 3, 0), (8, 8, 0, 0)],
            range(1, 9)):
        a = np.zeros((h, w))
        assert R.oriented_size(w, h, margins, o) == O.crop_orient(a, margins, o).shape[::-1]
        if min(w, h) <= 64:
            for bayer in range(4):
                want = O.bayer_after(bayer, margins, w, h, o)
                assert R.oriented_bayer(bayer, margins, w, h, o) == want


def test_bayer_of_a_turned_rggb():
    # A quarter turn clockwise shows the stored (0, h-1) at the top left (a
    # green for even h) and the stored (0, h-2), a red, right of it: GRBG.
    assert O.bayer_after(0, (0, 0, 0, 0), 8, 6, 6) == 1
    # A half turn shows the stored bottom-right blue first: BGGR.
    assert O.bayer_after(0, (0, 0, 0, 0), 8, 6, 3) == 3
    # One column cropped off the left: GRBG.
    assert O.bayer_after(0, (1, 0, 0, 0), 8, 6, 1) == 1


def test_bad_arguments():
    with pytest.raises(ValueError):
        R.oriented_size(10, 10, None, 9)
    with pytest.raises(ValueError):
        R.oriented_bayer(4, None, 10, 10, 1)
