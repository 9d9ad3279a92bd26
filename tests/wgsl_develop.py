"""A second, independent restatement of the crate's develop chain, written
from its WGSL twin (src/gpu/develop.wgsl:95-303) in NumPy float32 over
whole pictures at once, to pin the oracle's develop (oracle/prr_oracle.c,
written from src/demosaic.rs:413-720) bit for bit.

TEST INFRASTRUCTURE ONLY (a checker of the checker): imported by tests/.

The WGSL's `quot` and `root` (develop.wgsl:72-85) are its way of getting
correctly rounded division and square root on a GPU; here they are NumPy's
float32 `/` and `sqrt`, which are IEEE correctly rounded. `rounded(x)`
(develop.wgsl:66-68) is `x`. `fma` is computed exactly (one rounding) by
`fma32`, so every fused multiply-add the WGSL writes is one here. Each
function also counts the branches it takes (`Coverage`), so a test can
show that a comparison exercised every one.
"""
from dataclasses import dataclass, field

import numpy as np

F = np.float32


def fma32(a, b, c):
    """a * b + c for float32 operands with a single rounding to float32.

    The product of two float32 values is exact in float64 (48 significant
    bits); the float64 sum s = p + c is rounded once, and TwoSum gives its
    exact error e (s + e == p + c). Rounding s to float32 gives the right
    answer unless s lies exactly halfway between two float32 values (float64
    holds those midpoints exactly) while e != 0: then the exact sum lies on
    e's side of the midpoint and rounds to that neighbour."""
    a64 = np.asarray(a, np.float32).astype(np.float64)
    b64 = np.asarray(b, np.float32).astype(np.float64)
    c64 = np.asarray(c, np.float32).astype(np.float64)
    p = a64 * b64
    s = p + c64
    bp = s - p
    e = (p - (s - bp)) + (c64 - bp)
    r = s.astype(np.float32)
    d = s - r.astype(np.float64)
    other = np.nextafter(r, np.where(d > 0, F(np.inf), F(-np.inf)).astype(np.float32))
    half = (other.astype(np.float64) - r.astype(np.float64)) * 0.5
    midpoint = (d != 0) & (d == half)
    beyond = midpoint & (e != 0) & (np.sign(e) == np.sign(d))
    return np.where(beyond, other, r).astype(np.float32)


def linearize_lut(curve):
    """Curve::lut (curve.rs:14-33) for all 65,536 code values."""
    code = np.arange(65536, dtype=np.int64)
    pts = np.array(list(curve) + [65536], dtype=np.int64)
    seg, frac = code >> 13, code & 0x1FFF
    c0, c1 = pts[seg], pts[seg + 1]
    delta = (c1 - c0) & 0xFFFF
    return np.minimum((c0 * 8192 + delta * frac + 4096) >> 13, 65535)


@dataclass
class Params:
    """The develop shader's uniforms (develop.wgsl:12-40) with the stage
    applied: DemosaicParams fields plus the output matrix."""
    cfa: int
    black: int
    one_over_range: float
    noise: tuple
    clip: tuple
    black_over_range: float
    wb: tuple
    matrix: tuple = None

    @classmethod
    def of(cls, prm, stage, matrix=None):
        """From an oracle OrcParams and a stage (0 Demosaiced, 1
        WhiteBalanced, 2 Rec2020): Demosaiced develops without white
        balance (wb = clip = 1)."""
        wb = (1.0, 1.0) if stage == 0 else (prm.wb[0], prm.wb[1])
        clip = (1.0, 1.0) if stage == 0 else (prm.clip[0], prm.clip[1])
        return cls(prm.cfa_phase, prm.black, prm.one_over_range, tuple(prm.noise), clip,
                   prm.black_over_range, wb, tuple(matrix) if stage == 2 else None)


@dataclass
class Coverage:
    counts: dict = field(default_factory=dict)

    def add(self, name, mask):
        self.counts[name] = self.counts.get(name, 0) + int(np.count_nonzero(mask))


def variance(v, off, noise):
    """develop.wgsl:89-92: c + max(b, a * (off + v)^2)."""
    t = F(off) + v
    return F(noise[2]) + np.maximum(F(noise[1]), F(noise[0]) * (t * t))


def dot4(a, b):
    """develop.wgsl:100-102: fma(a.w, b.w, fma(a.z, b.z, fma(a.x, b.x, a.y * b.y)))."""
    return fma32(a[3], b[3], fma32(a[2], b[2], fma32(a[0], b[0], a[1] * b[1])))


def sum4(v):
    """develop.wgsl:107-109: w + (z + (x + y))."""
    return v[3] + (v[2] + (v[0] + v[1]))


def mirror(d, pos, length, c):
    """develop.wgsl:131-143, in its order of tests."""
    d0, d1, d2, d3 = d
    at0, at1 = pos == 0, pos == 1
    atl = (pos == length - 1) & ~at0 & ~at1
    atl2 = (pos == length - 2) & ~at0 & ~at1 & ~atl
    n0 = np.where(at0, d1, np.where(atl, d0, d0))
    n1 = np.where(at0, d1, np.where(atl, d0, d1))
    n2 = np.where(at0, d3, np.where(at1, c, np.where(atl, d2, d2)))
    n3 = np.where(at0, d3, np.where(atl, d2, np.where(atl2, c, d3)))
    return [n0, n1, n2, n3]


def green(norm, p, cov, plant=None, ratios=None):
    """Pass 1 (develop.wgsl:145-200) over a height x width picture of
    normalized values: (red or blue, green) per sensel. plant=
    "ratio_denominator" drops the factor 2 of the ratio's denominator (a
    mistake for tests to catch). A dict passed as `ratios` receives the
    horizontal and vertical curvature ratios ("h", "v"; meaningful at the
    red and blue sensels)."""
    h, w = norm.shape
    ys, xs = np.mgrid[0:h, 0:w]
    red_row, red_col = p.cfa >> 1, p.cfa & 1
    phase = (ys & 1) == red_row
    first = np.where(phase, red_col, 1 - red_col)
    site = (xs & 1) == first

    def at(dx, dy):
        return norm[np.clip(ys + dy, 0, h - 1), np.clip(xs + dx, 0, w - 1)]

    clip = np.where(phase, F(p.clip[0]), F(p.clip[1])).astype(np.float32)
    c = np.minimum(norm, clip)
    one = F(1.0)
    hh = [np.minimum(at(-1, 0), one), np.minimum(at(1, 0), one),
          np.minimum(at(-2, 0), clip), np.minimum(at(2, 0), clip)]
    vv = [np.minimum(at(0, -1), one), np.minimum(at(0, 1), one),
          np.minimum(at(0, -2), clip), np.minimum(at(0, 2), clip)]
    hh = mirror(hh, xs, w, c)
    vv = mirror(vv, ys, h, c)
    off = F(p.black) * F(p.one_over_range)
    var_h = [variance(v, off, p.noise) for v in hh]
    var_v = [variance(v, off, p.noise) for v in vv]
    sums = [hh[0] + hh[1], vv[0] + vv[1], hh[2] + hh[3], vv[2] + vv[3]]
    gate = c > np.sqrt(variance(c, off, p.noise))
    two = F(2.0)
    t = [c * two + sums[2], c * two + sums[3], c * -two + sums[2], c * -two + sums[3]]
    ratio = []
    for k, (d, vd) in enumerate(((hh, var_h), (vv, var_v))):
        sd = [np.sqrt(v) for v in vd]
        a = d[0] - d[1]
        b = d[2] - d[3]
        any_ = ((a * b > 0) | ((d[2] - c) * (d[3] - c) > 0) | (sd[0] + sd[1] > np.abs(a)) |
                (sd[2] + sd[3] > np.abs(b)))
        above = (d[0] > sd[0]) & (d[1] > sd[1]) & (d[2] > sd[2]) & (d[3] > sd[3])
        take = gate & any_ & above
        with np.errstate(all="ignore"):
            den = t[k] if plant == "ratio_denominator" else t[k] * two
            r = np.where(take, t[2 + k] / den, F(0.0)).astype(np.float32)
        name = "hv"[k]
        cov.add(f"ratio_{name}_pos", site & take & (r > 0))
        cov.add(f"ratio_{name}_neg", site & take & (r < 0))
        cov.add(f"ratio_{name}_zero_by_noise", site & gate & any_ & ~above)
        ratio.append(r)
        if ratios is not None:
            ratios[name] = r
    cov.add("green_gate_closed", site & ~gate)
    half, one = F(0.5), F(1.0)
    gh = (sums[0] * half) * (one - ratio[0])
    gv = (sums[1] * half) * (one - ratio[1])
    dh = [hh[0] - gh, hh[1] - gh, hh[2] - c, hh[3] - c]
    dv = [vv[0] - gv, vv[1] - gv, vv[2] - c, vv[3] - c]
    eh = dot4(dh, dh) + (var_h[0] + var_h[1])
    ev = dot4(dv, dv) + (var_v[0] + var_v[1])
    with np.errstate(all="ignore"):
        g = gv + ((gh - gv) * ev) / (eh + ev)
    # Mirror cases counted at red/blue sites only (where the green pass uses them).
    for tag, mask in (("mirror_h_", xs), ("mirror_v_", ys)):
        length = w if tag == "mirror_h_" else h
        cov.add(tag + "0", site & (mask == 0))
        cov.add(tag + "1", site & (mask == 1))
        cov.add(tag + "last", site & (mask == length - 1) & (mask > 1))
        cov.add(tag + "last2", site & (mask == length - 2) & (mask > 1))
    rb = np.where(site, c, F(0.0)).astype(np.float32)
    gr = np.where(site, g, norm).astype(np.float32)
    return rb, gr


def estimate(xx, gg, cg, cs, off, noise, cov, mask, plant=None):
    """develop.wgsl:210-234 on arrays of four neighbours. plant=
    "weight_swap" exchanges a and b in the weight (a mistake for tests to
    catch)."""
    quarter = [F(0.25)] * 4
    mean_x = dot4(xx, quarter)
    centre_ok = cs < cg
    var_x = [variance(v, off, noise) for v in xx]
    sx = [np.sqrt(v) for v in var_x]
    sg = [np.sqrt(variance(v, off, noise)) for v in gg]
    above = np.ones_like(centre_ok)
    for i in range(4):
        above &= (sx[i] < xx[i]) & (sg[i] < gg[i])
    with np.errstate(all="ignore"):
        lo = [((cg - cs) * (xx[i] - sx[i])) / (sg[i] + gg[i]) for i in range(4)]
        mid = [(cg * xx[i]) / gg[i] for i in range(4)]
        hi = [((cs + cg) * (sx[i] + xx[i])) / (gg[i] - sg[i]) for i in range(4)]
        spread = dot4([mid[i] - lo[i] for i in range(4)], [hi[i] - mid[i] for i in range(4)])
        mean_mid = dot4(mid, quarter)
        dm = [m - mean_mid for m in mid]
        b = fma32(spread, F(0.75), dot4(dm, dm))
        dx = [v - mean_x for v in xx]
        a = fma32(sum4(var_x), F(0.75), dot4(dx, dx))
        weight = (b if plant == "weight_swap" else a) / (b + a)
        res = fma32(mean_mid - mean_x, weight, mean_x)
    ok = centre_ok & above
    cov.add("estimate_fallback_centre", mask & ~centre_ok)
    cov.add("estimate_fallback_neighbour", mask & centre_ok & ~above)
    cov.add("estimate_weight_interior",
            mask & ok & (weight > 0) & (weight < 1) & (mean_mid != mean_x))
    return np.where(ok, res, mean_x).astype(np.float32)


def chroma(rb, gr, p, cov, plant=None):
    """Pass 2 (develop.wgsl:236-303): white-balanced camera RGB clipped at
    1.0, through the output matrix when there is one."""
    h, w = rb.shape
    ys, xs = np.mgrid[0:h, 0:w]
    xm = np.where(xs > 0, xs - 1, xs + 1)
    xp = np.where(xs < w - 1, xs + 1, xs - 1)
    ym = np.where(ys > 0, ys - 1, ys + 1)
    yp = np.where(ys < h - 1, ys + 1, ys - 1)
    off = F(p.black_over_range)
    c, g = rb, gr
    cs = np.sqrt(variance(g, off, p.noise))
    on_red_row = (ys & 1) == (p.cfa >> 1)
    red_col = p.cfa & 1
    rb_col = np.where(on_red_row, red_col, 1 - red_col)
    site = (xs & 1) == rb_col

    def at(x, y):
        return rb[y, x], gr[y, x]

    n = [at(xm, ym), at(xm, yp), at(xp, ym), at(xp, yp)]
    e = estimate([v[0] for v in n], [v[1] for v in n], g, cs, off, p.noise, cov, site, plant)

    def pair(u, v, mask):
        return estimate([u[0], v[0], u[0], v[0]], [u[1], v[1], u[1], v[1]], g, cs, off,
                        p.noise, cov, mask, plant)

    eh = pair(at(xm, ys), at(xp, ys), ~site)
    ev = pair(at(xs, ym), at(xs, yp), ~site)
    r = np.where(site, np.where(on_red_row, c, e), np.where(on_red_row, eh, ev))
    b = np.where(site, np.where(on_red_row, e, c), np.where(on_red_row, ev, eh))
    one = F(1.0)
    rgb = [np.minimum(r * F(p.wb[0]), one), np.minimum(g, one), np.minimum(b * F(p.wb[1]), one)]
    if p.matrix is not None:
        m = [F(v) for v in p.matrix]
        rgb = [fma32(rgb[2], m[3 * i + 2], fma32(rgb[0], m[3 * i], rgb[1] * m[3 * i + 1]))
               for i in range(3)]
    return [np.asarray(v, np.float32) for v in rgb]


def develop(codes, curve, p, cov=None, plant=None):
    """The whole chain on a height x width picture of code values: the
    folded linearize-and-normalize table (Conversion::new,
    demosaic.rs:293-298), the green pass and the chroma pass. Returns the
    three planes and the branch coverage."""
    cov = cov or Coverage()
    lin = linearize_lut(curve)
    lut = (F(p.one_over_range) * (lin - p.black).astype(np.float32)).astype(np.float32)
    norm = lut[np.asarray(codes, dtype=np.int64)]
    rb, gr = green(norm, p, cov, plant)
    return chroma(rb, gr, p, cov, plant), cov


def green_from_linear(linear, p, cov=None):
    """green_pass (demosaic.rs:453-464) on linear values: normalize, then
    pass 1."""
    cov = cov or Coverage()
    norm = (F(p.one_over_range) * (np.asarray(linear, np.int64) - p.black).astype(np.float32))
    return green(norm.astype(np.float32), p, cov), cov
