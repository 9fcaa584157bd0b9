"""The contact state's spatial form skips a tile when the boxes of its two
blocks of circles do not meet.  Each block's box is the extent of its
circles' centres widened by |r| and then by 1e-5 (|x| + |y| + |r|)max
(contact_sparse.h: scs_block_box), and a pair is a contact when the prox's
fp32 test fl(fma(dx, dx, fl(dy dy))) < fl(fl(r1 + r2)^2) passes
(prox_device.h: pair_hits).  These CPU tests emulate both in fp32 and check,
on pairs placed at the collision boundary, that every pair the test accepts
lies in boxes that meet -- the claim that makes skipping a tile exact --
and that the check catches boxes built with a plausible mistake."""
import numpy as np

F = np.float32


def pair_hits(c1, c2, r1, r2):
    """pair_hits in fp32: dx, dy, s rounded; fma(dx, dx, dy*dy) with one rounding."""
    dx = (c1[..., 0] - c2[..., 0]).astype(F)
    dy = (c1[..., 1] - c2[..., 1]).astype(F)
    dy2 = (dy * dy).astype(F)
    d2 = (dx.astype(np.float64) * dx.astype(np.float64) + dy2.astype(np.float64)).astype(F)
    s = (r1 + r2).astype(F)
    return d2 < (s * s).astype(F)


def block_box(c, r, margin=True):
    """scs_block_box for one circle (a block of one): the widened box in fp32."""
    ar = np.abs(r).astype(F)
    lx, hx = (c[..., 0] - ar).astype(F), (c[..., 0] + ar).astype(F)
    ly, hy = (c[..., 1] - ar).astype(F), (c[..., 1] + ar).astype(F)
    if margin:
        d = (F(1e-5) * (np.abs(c[..., 0]) + np.abs(c[..., 1]) + ar).astype(F)).astype(F)
        lx, hx, ly, hy = (lx - d).astype(F), (hx + d).astype(F), (ly - d).astype(F), (hy + d).astype(F)
    return lx, hx, ly, hy


def meet(b1, b2):
    return (b1[0] <= b2[1]) & (b2[0] <= b1[1]) & (b1[2] <= b2[3]) & (b2[2] <= b1[3])


def boundary_pairs(n, seed, scale=1.0, offset=0.0):
    """Pairs whose centre distance is within a few ulps of r1 + r2, along
    directions near the axes (where the box test is tightest)."""
    rng = np.random.default_rng(seed)
    c1 = (rng.uniform(0.0, 1.0, (n, 2)) * scale + offset).astype(F)
    r1 = rng.uniform(0.0, 0.05, n).astype(F) * F(scale)
    r2 = rng.uniform(0.0, 0.05, n).astype(F) * F(scale)
    ang = rng.choice([0.0, np.pi / 2, np.pi, 1.5 * np.pi], n) + rng.normal(0.0, 1e-4, n)
    dist = (r1.astype(np.float64) + r2) * (1.0 + rng.integers(-8, 9, n) * 2.0 ** -24)
    c2 = (c1 + np.stack([np.cos(ang), np.sin(ang)], 1) * dist[:, None]).astype(F)
    # some negative radii (the radius variable is not clamped)
    flip = rng.random(n) < 0.1
    r1 = np.where(flip, -r1, r1).astype(F)
    return c1, c2, r1, r2


def test_accepted_pairs_lie_in_meeting_boxes():
    for seed, scale, offset in ((0, 1.0, 0.0), (1, 1.0, 0.0), (2, 4.0, -2.0), (3, 1e-3, 0.5)):
        c1, c2, r1, r2 = boundary_pairs(200_000, seed, scale, offset)
        hit = pair_hits(c1, c2, r1, r2)
        ok = meet(block_box(c1, r1), block_box(c2, r2))
        assert hit.sum() > 1000
        assert not np.any(hit & ~ok), f"seed {seed}: {np.sum(hit & ~ok)} accepted pairs in boxes that do not meet"


def test_the_check_catches_boxes_without_the_radii():
    """A plausible mistake, boxes of the centres alone (the |r| widening
    dropped), leaves accepted pairs outside the boxes: the check above
    would catch it."""
    c1, c2, r1, r2 = boundary_pairs(200_000, 7)
    hit = pair_hits(c1, c2, r1, r2)
    zero = np.zeros_like(r1)
    assert np.any(hit & ~meet(block_box(c1, zero), block_box(c2, zero)))
    assert not np.any(hit & ~meet(block_box(c1, r1), block_box(c2, r2)))
