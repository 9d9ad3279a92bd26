"""Oracle (test infrastructure only): crop and display orientation of a
picture in NumPy, for the device kernel of csrc/prr_orient.hip.

The eight orientations are the EXIF symmetries named by the crate's
mov::Orientation (src/mov.rs:121-146: "what to do to the stored frame to
show it upright"; Rotate90 is "turned 90 degrees clockwise", Transpose
"mirrored about the main diagonal", Transverse "about the anti-diagonal"),
written with NumPy's own flips, transposes and quarter turns. The crop
margins are header::Crop's (src/header.rs:98-109), applied only when they
leave a picture (examples/prr2dng/dng.rs:190-196). Pinned in
tests/test_orient.py by hand-written small cases and the group laws of
the eight symmetries.
"""
import numpy as np


def crop(a, margins):
    """margins = (left, right, top, bottom)."""
    left, right, top, bottom = (int(m) for m in margins)
    h, w = a.shape[:2]
    if w > left + right and h > top + bottom:
        return a[top:h - bottom, left:w - right]
    return a


def orient(a, orientation):
    """The picture turned as EXIF orientation 1..8 says to show it upright."""
    o = int(orientation)
    if o == 1:
        return a
    if o == 2:
        return a[:, ::-1]
    if o == 3:
        return a[::-1, ::-1]
    if o == 4:
        return a[::-1, :]
    if o == 5:
        return np.swapaxes(a, 0, 1)
    if o == 6:
        return np.rot90(a, -1)       # clockwise
    if o == 7:
        return np.swapaxes(a[::-1, ::-1], 0, 1)
    if o == 8:
        return np.rot90(a, 1)        # anticlockwise
    raise ValueError(o)


def crop_orient(a, margins, orientation):
    return np.ascontiguousarray(orient(crop(a, margins), orientation))


def bayer_after(bayer, margins, width, height, orientation):
    """The pattern code (2 * red row + red column of the 2x2 cell, as
    BayerPattern::red_position, header.rs:66-73) of the cropped, turned CFA,
    by turning a picture of colour labels."""
    red_row, red_col = bayer >> 1, bayer & 1
    yy, xx = np.mgrid[0:height, 0:width]
    red = ((yy % 2) == red_row) & ((xx % 2) == red_col)
    r = crop_orient(red, margins, orientation)
    (y, x), = np.argwhere(r[:2, :2])
    return int(2 * y + x)
