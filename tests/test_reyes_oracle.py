"""Pins the Reyes oracle (oracle/reyes_oracle.c) with cases whose answers the
mathematics fixes: the reference holds no Reyes fixtures or images.

Split: a linear patch splits into linear sub-patches whose control points are
known exactly (dyadic), in a known order with known dice grids.  Dice: a
Bezier patch reproduces its corner control points, and a patch in the z = 0
plane has the normal (0, 0, +-1) exactly, so its Phong colour has a closed
form.  Shade: axis-aligned rectangles cover exactly the samples strictly
inside them; the three-corner binning of ShadeStage::assignBin drops a known
set of pixels; ties go to the later micropolygon.
"""
import numpy as np
import pytest

IDENTITY = np.eye(4, dtype=np.float32).reshape(16)


def linear_patch(x0, y0, hx, hy, z=0.5):
    """Control points of a linear patch: CP[4r+c] = (x0 + c*hx, y0 + r*hy, z, 1)."""
    cp = np.zeros((1, 64), np.float32)
    for r in range(4):
        for c in range(4):
            cp[0, 4 * (4 * r + c):4 * (4 * r + c) + 4] = (x0 + c * hx, y0 + r * hy, z, 1.0)
    return cp


def next_pow2(x):
    return 1 << (x - 1).bit_length()


def test_split_linear_square_tree(oracle_mod):
    # 48x48 px square on a 256x256 screen (identity camera: px = (x+1)*128),
    # threshold 16: U, U (48 -> 24 -> 12) then V, V: 16 leaves at level 4,
    # emitted as u-interval * 4 + v-interval, each 12x12 px (nu = nv = 2).
    cp = linear_patch(-0.5, -0.5, 0.125, 0.125)
    leaves, nu, nv = oracle_mod.reyes_split(cp, 256, 256, IDENTITY, 16.0)
    assert leaves.shape[0] == 16
    assert (nu == 2).all() and (nv == 2).all()
    for k in range(16):
        iu, iv = k // 4, k % 4
        for r in range(4):
            for c in range(4):
                want = (-0.5 + 0.09375 * iu + 0.03125 * c, -0.5 + 0.09375 * iv + 0.03125 * r, 0.5, 1.0)
                got = leaves[k, 4 * (4 * r + c):4 * (4 * r + c) + 4]
                assert tuple(got) == pytest.approx(want, abs=0), (k, r, c)


def test_split_direction_prefers_u(oracle_mod):
    # 96 px wide, 24 px tall: U splits 96 -> 48 -> 24 -> 12, then V 24 -> 12
    cp = linear_patch(-0.5, -0.5, 0.25, 0.0625)
    leaves, nu, nv = oracle_mod.reyes_split(cp, 256, 256, IDENTITY, 16.0)
    assert leaves.shape[0] == 16
    xs = leaves[:, 0]  # corner CP0 x of each leaf, emission order
    # level 4 order: u-interval major, v-interval minor
    assert np.array_equal(xs, np.repeat(-0.5 + (0.75 / 8) * np.arange(8), 2).astype(np.float32))
    assert np.array_equal(leaves[:, 1], np.tile([-0.5, -0.5 + 0.09375], 8).astype(np.float32))


@pytest.mark.parametrize("hu,hv", [(1, 2), (4, 5), (8, 8), (12, 3), (20, 33), (40, 0.5)])
@pytest.mark.parametrize("thr", [4.0, 16.0])
def test_leaf_grid_table(oracle_mod, hu, hv, thr):
    # max_splits = 0: the patch is a leaf whatever its size; extents are
    # 3*hu and 3*hv pixels exactly (axis-aligned polylines)
    cp = linear_patch(-0.5, -0.5, hu / 128.0, hv / 128.0)
    leaves, nu, nv = oracle_mod.reyes_split(cp, 256, 256, IDENTITY, thr, max_splits=0)
    assert leaves.shape[0] == 1
    for h, n in ((hu, nu[0]), (hv, nv[0])):
        want = min(max(next_pow2(int(np.ceil(3 * h / thr)) + 1), 2), 32)
        assert n == want, (h, thr, n, want)


def test_split_depth_cap(oracle_mod):
    cp = linear_patch(-0.9, -0.9, 0.6, 0.6)  # ~460 px square at threshold 4
    for cap in (0, 1, 3, 6):
        leaves, _, _ = oracle_mod.reyes_split(cp, 256, 256, IDENTITY, 4.0, max_splits=cap)
        assert leaves.shape[0] == 2 ** cap


def dyadic_patch(seed):
    rng = np.random.default_rng(seed)
    cp = np.zeros((1, 64), np.float32)
    q = rng.integers(-32, 33, size=(16, 3)) / 64.0
    for i in range(16):
        cp[0, 4 * i:4 * i + 4] = (q[i, 0], q[i, 1], q[i, 2] * 0.25, 1.0)
    return cp


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_dice_corner_reproduction(oracle_mod, seed):
    # identity camera on a 2x2 screen: sx = x + 1, sy = y + 1 for w = 1; the
    # patch corners are the corner control points (B_i(0), B_i(1) are 0/1)
    cp = dyadic_patch(seed)
    up = oracle_mod.reyes_dice(cp, np.array([2], np.int32), np.array([2], np.int32), 8, 2, 2, IDENTITY, 0.0)
    assert up.shape[0] == 64
    P = cp[0].reshape(16, 4)
    q00, q70, q77, q07 = up[0], up[7 * 8 + 0], up[7 * 8 + 7], up[0 * 8 + 7]
    assert (q00["x"][0], q00["y"][0]) == (P[0, 0] + 1, P[0, 1] + 1)    # (u,v) = (0,0): CP 0
    assert (q70["x"][1], q70["y"][1]) == (P[3, 0] + 1, P[3, 1] + 1)    # (1,0): CP 3
    assert (q77["x"][2], q77["y"][2]) == (P[15, 0] + 1, P[15, 1] + 1)  # (1,1): CP 15
    assert (q07["x"][3], q07["y"][3]) == (P[12, 0] + 1, P[12, 1] + 1)  # (0,1): CP 12


def test_dice_linear_patch_interior(oracle_mod):
    # a linear patch diced 8x8: every corner lies on the regular grid (the
    # Bernstein basis sums to one and reproduces linear functions; float
    # rounding only)
    cp = linear_patch(-0.5, -0.25, 0.125, 0.0625, z=0.0)
    up = oracle_mod.reyes_dice(cp, np.array([2], np.int32), np.array([2], np.int32), 8, 2, 2, IDENTITY, 0.0)
    for k in range(8):
        for j in range(8):
            q = up[k * 8 + j]
            for cidx, (du, dv) in enumerate(((0, 0), (1, 0), (1, 1), (0, 1))):
                ex = -0.5 + 0.375 * (k + du) / 8 + 1
                ey = -0.25 + 0.1875 * (j + dv) / 8 + 1
                assert abs(q["x"][cidx] - ex) < 1e-6 and abs(q["y"][cidx] - ey) < 1e-6


def f32(x):
    return np.float32(x)


def quad_colour(c):
    s = f32(f32(f32(c + c) + c) + c)
    return f32(min(s, f32(4.0)) * f32(0.25))


@pytest.mark.parametrize("flip", [False, True])
def test_planar_patch_normal_and_colour(oracle_mod, flip):
    # patch in the z = 0 plane: dP/du and dP/dv have zero z, so the normal is
    # exactly (0, 0, +-1).  u along +x, v along +y gives +z; the shade stage
    # inverts it and dots with matcol (0.10, 0.35, 0.85) -> -0.85 -> 0, so every
    # vertex colour is 0.2; with v along -y the dot is +0.85 and the colour is
    # 0.2 + 0.85 * lightvec, saturated (shade.pikostage:64-85, 22-30).
    cp = linear_patch(-0.5, 0.5 if flip else -0.5, 0.125, -0.125 if flip else 0.125, z=0.0)
    up = oracle_mod.reyes_dice(cp, np.array([2], np.int32), np.array([2], np.int32), 4, 64, 64, IDENTITY, 0.0)
    if not flip:
        want = [quad_colour(f32(0.2))] * 3
    else:
        d = f32(0.85)
        want = [quad_colour(min(max(f32(f32(0.2) + f32(d * f32(l))), f32(0.0)), f32(1.0)))
                for l in (0.1773, 0.5773, 0.973)]
    for q in up:
        assert tuple(q["rgb"]) == tuple(want)
        assert q["z"] == 0.0


def upolys(quads):
    """quads: list of ((x0,y0),(x1,y1),(x2,y2),(x3,y3), z, (r,g,b))."""
    out = np.zeros(len(quads), oracle_mod_upoly())
    for k, (p0, p1, p2, p3, z, rgb) in enumerate(quads):
        out[k]["x"] = (p0[0], p1[0], p2[0], p3[0])
        out[k]["y"] = (p0[1], p1[1], p2[1], p3[1])
        out[k]["z"] = z
        out[k]["rgb"] = rgb
    return out


def oracle_mod_upoly():
    from oracle import oracle
    return oracle.UPOLY


def covered(z):
    return (z != 0x3F800000)


def offsets(spp):
    return [(0.5, 0.5)] if spp == 1 else [(0.25, 0.25), (0.75, 0.25), (0.25, 0.75), (0.75, 0.75)]


@pytest.mark.parametrize("spp", [1, 4])
@pytest.mark.parametrize("cw", [False, True])
@pytest.mark.parametrize("box", [(10, 3, 14, 9), (10.25, 3.5, 14.75, 9.25), (0.0, 0.0, 1.0, 1.0), (5.5, 5.5, 6.5, 6.5)])
def test_shade_rectangle_strict_interior(oracle_mod, spp, cw, box):
    a, b, c, d = box
    pts = [(a, b), (c, b), (c, d), (a, d)]
    if cw:
        pts = pts[::-1]
    up = upolys([(*pts, 0.5, (1.0, 0.0, 0.0))])
    _, z = oracle_mod.reyes_shade(up, 32, 16, spp)
    want = np.zeros((16, 32, spp), bool)
    for y in range(16):
        for x in range(32):
            for s, (ox, oy) in enumerate(offsets(spp)):
                want[y, x, s] = a < x + ox < c and b < y + oy < d
    assert np.array_equal(covered(z), want)


def test_shade_three_corner_binning(oracle_mod):
    # trapezoid (70,10) (90,10) (90,20) (60,20): the first three corners lie in
    # 64-px bin column 1 only, so pixels left of x = 64 are never tested
    # (computeBoundingBox uses screenPos[0..2], dice.pikostage:18-22).  Row
    # y = 10 + r covers x >= max(70 - r, 64) up to 89: 239 pixels (245 if
    # the fourth corner were binned too).
    up = upolys([((70, 10), (90, 10), (90, 20), (60, 20), 0.5, (1.0, 1.0, 1.0))])
    _, z = oracle_mod.reyes_shade(up, 128, 32, 1)
    cov = covered(z[..., 0])
    assert cov.sum() == 239
    for r in range(10):
        row = np.nonzero(cov[10 + r])[0]
        assert row[0] == max(70 - r, 64) and row[-1] == 89


def test_shade_depth_order_and_ties(oracle_mod):
    sq = ((2, 2), (6, 2), (6, 6), (2, 6))
    cases = [
        ([(0.5, (1, 0, 0)), (0.5, (0, 1, 0))], (0, 1, 0)),    # tie: the later one
        ([(0.3, (1, 0, 0)), (0.5, (0, 1, 0))], (1, 0, 0)),    # nearer first stays
        ([(0.5, (1, 0, 0)), (0.3, (0, 1, 0))], (0, 1, 0)),    # nearer later wins
        ([(1.0, (1, 0, 0))], (1, 0, 0)),                       # z = 1 ties the clear: drawn
        ([(1.5, (1, 0, 0))], None),                            # beyond the clear: not drawn
    ]
    bg = np.array(oracle_mod.from_abgr(0xff663313), np.float32)
    for quads, want in cases:
        up = upolys([(*sq, z, rgb) for z, rgb in quads])
        rgba, _ = oracle_mod.reyes_shade(up, 8, 8, 1)
        got = rgba[4, 4, :3]
        exp = bg if want is None else np.array(want, np.float32)
        assert np.array_equal(got, exp), (quads, got)
        assert rgba[4, 4, 3] == 1.0


def test_resolve_box_filter(oracle_mod):
    # a rectangle covering the left column of strata of pixel (3, 3) only:
    # samples 0 and 2 take the colour, 1 and 3 keep the clear colour
    up = upolys([((3.0, 3.0), (3.5, 3.0), (3.5, 4.0), (3.0, 4.0), 0.5, (0.5, 0.25, 1.0))])
    rgba, z = oracle_mod.reyes_shade(up, 8, 8, 4)
    assert list(covered(z[3, 3])) == [True, False, True, False]
    bg = np.array(oracle_mod.from_abgr(0xff663313), np.float32)
    col = np.array((0.5, 0.25, 1.0), np.float32)
    want = (((col + bg) + col) + bg) * np.float32(0.25)
    assert np.array_equal(rgba[3, 3, :3], want)
    assert np.array_equal(rgba[0, 0, :3], (((bg + bg) + bg) + bg) * np.float32(0.25))


def test_sphere_frame_is_finite_and_closed(oracle_mod):
    from piko_amd import scenes
    sc = scenes.sphere_patches(nlat=16, nlon=16, W=320, H=180)
    rgba, z, nup = oracle_mod.reyes_frame(sc)
    assert np.isfinite(rgba).all()
    leaves, _, _ = oracle_mod.reyes_split(sc.cp, sc.W, sc.H, sc.view_proj, sc.threshold)
    assert nup == 64 * leaves.shape[0]
    # the silhouette is a disc: the centre pixel and its neighbourhood are covered
    assert covered(z[80:100, 150:170]).all()


@pytest.mark.parametrize("disp", [0.5, -0.25])
def test_displacement_closed_form_on_planar_patch(oracle_mod, disp):
    # a linear patch in the z = 0 plane (u along +x, v along +y: unit normal
    # exactly (0, 0, 1)); displacement d = disp * (x^2 - y^2) along the normal
    # (dice.pikostage:206-229 with the displacement shader) moves each corner
    # to z = d, and x, y stay.  The identity camera gives screen z = z.  All
    # coordinates are dyadic with few bits, so every product and sum is exact
    # and the micropolygon's flat depth (((z0+z1)+z2)+z3)/4 has a closed form.
    from fractions import Fraction as F
    x0, y0, hx, hy = -0.5, -0.25, 0.125, 0.0625
    cp = linear_patch(x0, y0, hx, hy, z=0.0)
    up = oracle_mod.reyes_dice(cp, np.array([2], np.int32), np.array([2], np.int32), 8, 2, 2, IDENTITY, disp)
    assert up.shape[0] == 64
    for k in range(8):
        for j in range(8):
            q = up[k * 8 + j]
            zs = []
            for cidx, (du, dv) in enumerate(((0, 0), (1, 0), (1, 1), (0, 1))):
                x = F(x0) + F(3 * hx).limit_denominator() * F(k + du, 8)
                y = F(y0) + F(3 * hy).limit_denominator() * F(j + dv, 8)
                zs.append(F(disp) * (x * x - y * y))
                # x and y are unmoved: the normal has no x or y part
                assert q["x"][cidx] == np.float32(float(x + 1)) and q["y"][cidx] == np.float32(float(y + 1))
            assert q["z"] == np.float32(float(sum(zs) / 4)), (k, j)
