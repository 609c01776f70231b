"""Pins the oracle (CPU restatement of the reference CPU backend) against facts
that do not come from the oracle's own formulas: textbook point-in-triangle
tests, a max-over-corners reading of the trivial-reject test, rectangle pixel
counts, closed-form barycentric colours, tie order, the toABGR/fromABGR
round trip and bin-size invariance.  The reference holds no golden vectors
or tests of its own for this path (its tree has no tests directory)."""
from fractions import Fraction

import numpy as np
import pytest

from piko_amd import scenes


def orient2d(ax, ay, bx, by, px, py):
    """Textbook orientation determinant (b-a) x (p-a), exact in Python ints."""
    return (bx - ax) * (py - ay) - (by - ay) * (px - ax)


def make_scene(tris, W, H):
    """tris: list of ((x,y,z)*3, (c0,c1,c2)) in screen space."""
    pos = np.array([[c for v in t[0] for c in v] for t in tris], np.float32).reshape(-1, 9)
    col = np.array([t[1] for t in tris], np.uint32).reshape(-1, 3)
    return scenes.TriScene(pos, col, W, H)


def covered_pixels_textbook(xy, S, W, H):
    """Pixel centres inside the closed (counter-clockwise) triangle."""
    x0, y0, x1, y1, x2, y2 = (int(v) for v in xy)
    half = 1 << (S - 1)
    out = set()
    for py in range(H):
        for px in range(W):
            cx, cy = (px << S) + half, (py << S) + half
            if (orient2d(x0, y0, x1, y1, cx, cy) >= 0 and orient2d(x1, y1, x2, y2, cx, cy) >= 0
                    and orient2d(x2, y2, x0, y0, cx, cy) >= 0):
                out.add((px, py))
    return out


CLEAR = 0xff663313


def test_abgr_round_trip_all_bytes(oracle_mod):
    # fromABGR (math.h:323-330) then CPU toABGR (math.h:286-289) is the identity
    for b in range(256):
        c = (0xFF << 24) | (b << 16) | ((255 - b) << 8) | (b ^ 0x5A)
        assert oracle_mod.to_abgr(*oracle_mod.from_abgr(c)) == c


def test_to_abgr_truncates(oracle_mod):
    assert oracle_mod.to_abgr(1.0, 0.0, 0.0) == 0xFF0000FF
    assert oracle_mod.to_abgr(0.999, 0.5, 0.25) == (0xFF << 24) | (63 << 16) | (127 << 8) | 254


def test_fixpt_truncates_toward_zero(oracle_mod):
    assert oracle_mod.fixpt(1.99, 4) == 31
    assert oracle_mod.fixpt(-1.99, 4) == -31
    assert oracle_mod.fixpt(3.5, 8) == 896
    assert oracle_mod.fixpt(1e30, 8) == 1 << 29


@pytest.mark.parametrize("S", [4, 8])
def test_vertex_stage_culls(oracle_mod, S):
    ccw = ((2.0, 2.0, 0.5), (12.0, 3.0, 0.5), (5.0, 11.0, 0.5))
    cw = (ccw[0], ccw[2], ccw[1])
    flat = ((2.0, 2.0, 0.5), (6.0, 2.0, 0.5), (10.0, 2.0, 0.5))  # zero area
    # y-extent strictly between pixel-centre rows 2.5 and 3.5, wide in x: culled
    thin_y = ((1.0, 2.6, 0.5), (30.0, 2.7, 0.5), (15.0, 3.4, 0.5))
    # x-extent between pixel-centre columns but tall in y: kept (the reference
    # tests the Y axis twice, rasterMacros.h:74)
    thin_x = ((2.6, 1.0, 0.5), (3.4, 2.0, 0.5), (2.7, 30.0, 0.5))
    # vertex beyond the 225-pixel guard band: culled
    far = ((2.0, 2.0, 0.5), (64.0 + 226.0, 3.0, 0.5), (5.0, 11.0, 0.5))
    tris = [(t, (0xFFFFFFFF,) * 3) for t in (ccw, cw, flat, thin_y, thin_x, far)]
    sc = make_scene(tris, 64, 64)
    st = oracle_mod.vertex_stage(sc, S)
    assert list(st.keep) == [1, 0, 0, 0, 1, 0]


def random_tris(rng, n, W, H, S, big=False):
    tris = []
    for _ in range(n):
        while True:
            s = (W if big else 9)
            cx, cy = rng.uniform(-4, W + 4), rng.uniform(-4, H + 4)
            v = [(cx + rng.uniform(-s, s), cy + rng.uniform(-s, s)) for _ in range(3)]
            a = (v[1][0] - v[0][0]) * (v[2][1] - v[0][1]) - (v[2][0] - v[0][0]) * (v[1][1] - v[0][1])
            if abs(a) > 1e-3:
                break
        if a < 0:
            v = [v[0], v[2], v[1]]
        z = rng.uniform(0.1, 1.0, 3)
        cols = [int(c) for c in rng.integers(0, 2**32, 3, dtype=np.uint64)]
        tris.append((tuple((v[k][0], v[k][1], float(z[k])) for k in range(3)), tuple(cols)))
    return tris


@pytest.mark.parametrize("S,T", [(4, 8), (8, 8), (8, 16), (8, 32), (8, 64)])
def test_bin_walk_is_max_corner_test(oracle_mod, S, T):
    """RasterStage::assignBin (raster.pikostage:40-117) assigns bin b iff, for
    each edge, the maximum of the edge function over b's four corners is >= 0
    (within the rounded-bbox bin range); and it never misses a bin holding a
    covered pixel."""
    rng = np.random.default_rng(7 + S + T)
    W, H = 70, 45
    sc = make_scene(random_tris(rng, 60, W, H, S), W, H)
    st = oracle_mod.vertex_stage(sc, S)
    start, tris = oracle_mod.bins(st, W, H, S, T)
    nBx, nBy = -(-W // T), -(-H // T)
    got = {}
    for b in range(nBx * nBy):
        for t in tris[start[b]:start[b + 1]]:
            got.setdefault(int(t), []).append(b)
    bsb = S + (T.bit_length() - 1)
    B = 1 << bsb
    for i in range(st.n):
        if not st.keep[i]:
            assert int(i) not in got
            continue
        x0, y0, x1, y1, x2, y2 = (int(v) for v in st.xy[i])
        lo_x = min(x0, x1, x2) & ~0xFF
        lo_y = min(y0, y1, y2) & ~0xFF
        hi_x = (max(x0, x1, x2) + 0xFF) & ~0xFF
        hi_y = (max(y0, y1, y2) + 0xFF) & ~0xFF
        want = []
        for by in range(max(0, lo_y >> bsb), min(hi_y >> bsb, nBy - 1) + 1):
            for bx in range(max(0, lo_x >> bsb), min(hi_x >> bsb, nBx - 1) + 1):
                corners = [(bx * B + dx, by * B + dy) for dx in (0, B) for dy in (0, B)]
                ok = all(max(orient2d(ax, ay, bx_, by_, cx, cy) for cx, cy in corners) >= 0
                         for ax, ay, bx_, by_ in ((x0, y0, x1, y1), (x1, y1, x2, y2), (x2, y2, x0, y0)))
                if ok:
                    want.append(by * nBx + bx)
        assert got.get(i, []) == want, i
        for (px, py) in covered_pixels_textbook(st.xy[i], S, W, H):
            assert (py // T) * nBx + (px // T) in want
    # queues hold triangles in submission order
    for b in range(nBx * nBy):
        q = tris[start[b]:start[b + 1]]
        assert np.all(np.diff(q) > 0)


@pytest.mark.parametrize("S", [4, 8])
def test_coverage_matches_textbook(oracle_mod, S):
    rng = np.random.default_rng(11 + S)
    W, H = 40, 33
    for t in random_tris(rng, 25, W, H, S):
        sc = make_scene([t], W, H)
        st = oracle_mod.vertex_stage(sc, S)
        color, depth = oracle_mod.raster_bruteforce(st, W, H, S)
        written = {(x, y) for y in range(H) for x in range(W) if color[y, x] != CLEAR}
        want = covered_pixels_textbook(st.xy[0], S, W, H) if st.keep[0] else set()
        assert written == want


def test_rectangle_split_double_counts_diagonal(oracle_mod):
    """No fill rule (raster.pikostage:104,228): two triangles sharing a
    diagonal through pixel centres both cover those centres."""
    W = H = 16
    a = ((0.0, 0.0, 0.5), (8.0, 0.0, 0.5), (8.0, 8.0, 0.5))
    b = ((0.0, 0.0, 0.5), (8.0, 8.0, 0.5), (0.0, 8.0, 0.5))
    sc = make_scene([(a, (0xFF0000FF,) * 3), (b, (0xFF00FF00,) * 3)], W, H)
    st = oracle_mod.vertex_stage(sc, 8)
    f, cov = oracle_mod.fragment_stats(st, W, H, 8)
    assert cov == 64 and f == 64 + 8
    color, _ = oracle_mod.raster_bruteforce(st, W, H, 8)
    # equal z: the later triangle (b) wins the shared diagonal
    for k in range(8):
        assert color[k, k] == 0xFF00FF00


def test_depth_order_and_ties(oracle_mod):
    W = H = 16
    big = ((0.0, 0.0), (16.0, 0.0), (0.0, 16.0))

    def tri(z, c):
        return (tuple((x, y, z) for x, y in big), (c,) * 3)

    near, far = tri(0.25, 0xFF0000FF), tri(0.75, 0xFF00FF00)
    for order in ([near, far], [far, near]):
        color, depth = oracle_mod.raster_bruteforce(oracle_mod.vertex_stage(make_scene(order, W, H), 8), W, H, 8)
        assert color[0, 0] == 0xFF0000FF and depth[0, 0] == np.float32(0.25)
    # ties: depth test is <= (raster.pikostage:280-285), so the later wins
    t1, t2 = tri(0.5, 0xFF111111), tri(0.5, 0xFF222222)
    color, _ = oracle_mod.raster_bruteforce(oracle_mod.vertex_stage(make_scene([t1, t2], W, H), 8), W, H, 8)
    assert color[0, 0] == 0xFF222222
    # z beyond the cleared depth 1.0 never passes; z == 1.0 does
    c_far, c_one = tri(1.5, 0xFF333333), tri(1.0, 0xFF444444)
    color, depth = oracle_mod.raster_bruteforce(oracle_mod.vertex_stage(make_scene([c_far], W, H), 8), W, H, 8)
    assert color[0, 0] == CLEAR and depth[0, 0] == 1.0
    color, _ = oracle_mod.raster_bruteforce(oracle_mod.vertex_stage(make_scene([c_one], W, H), 8), W, H, 8)
    assert color[0, 0] == 0xFF444444


def test_gouraud_matches_exact_barycentrics(oracle_mod):
    """Colour = c2 + alpha (c0-c2) + beta (c1-c2) with alpha = e1/denom,
    beta = e2/denom (raster.pikostage:173-174, 261-293); compared with the
    exact rational value, truncated, within one step."""
    S, W, H = 8, 48, 40
    v = ((3.3, 2.1, 0.2), (44.7, 6.9, 0.9), (12.2, 37.5, 0.5))
    cols = (0xFF0000FF, 0xFF00FF00, 0xFFFF0000)  # red, green, blue vertices
    st = oracle_mod.vertex_stage(make_scene([(v, cols)], W, H), S)
    color, depth = oracle_mod.raster_bruteforce(st, W, H, S)
    x0, y0, x1, y1, x2, y2 = (int(t) for t in st.xy[0])
    denom = orient2d(x0, y0, x1, y1, x2, y2)
    half = 1 << (S - 1)
    n = 0
    for py in range(H):
        for px in range(W):
            if color[py, px] == CLEAR:
                continue
            cx, cy = (px << S) + half, (py << S) + half
            alpha = Fraction(orient2d(x1, y1, x2, y2, cx, cy), denom)   # weight of v0
            beta = Fraction(orient2d(x2, y2, x0, y0, cx, cy), denom)    # weight of v1
            gamma = 1 - alpha - beta
            want = [alpha * 255, beta * 255, gamma * 255]  # r from v0, g from v1, b from v2
            got = [(color[py, px] >> (8 * k)) & 255 for k in range(3)]
            for g, w in zip(got, want):
                assert abs(g - int(w)) <= 1
            zw = float(gamma * Fraction(0.5) + alpha * Fraction(float(np.float32(0.2)))
                       + beta * Fraction(float(np.float32(0.9))))
            assert abs(float(depth[py, px]) - zw) < 1e-5
            n += 1
    assert n > 300


@pytest.mark.parametrize("S", [4, 8])
def test_bin_size_invariance(oracle_mod, S):
    """Binned raster over any bin size equals the unbinned brute force."""
    sc = scenes.tris_uniform(5, 400, 96, 72)
    st = oracle_mod.vertex_stage(sc, S)
    ref_c, ref_d = oracle_mod.raster_bruteforce(st, 96, 72, S)
    for T in (8, 16, 32, 64):
        start, tris = oracle_mod.bins(st, 96, 72, S, T)
        c, d = oracle_mod.raster(st, 96, 72, S, T, start, tris)
        assert np.array_equal(c, ref_c) and np.array_equal(d.view(np.int32), ref_d.view(np.int32))


def test_order_invariance_without_ties(oracle_mod):
    """Permuting triangles with distinct z changes nothing (depth = min z)."""
    rng = np.random.default_rng(3)
    sc = scenes.tris_uniform(9, 300, 64, 64)
    perm = rng.permutation(sc.n)
    sc2 = scenes.TriScene(sc.pos[perm], sc.col[perm], 64, 64)
    a = oracle_mod.frame(sc, 8, 16)
    b = oracle_mod.frame(sc2, 8, 16)
    # ties are possible only on exactly equal z bits; uniform float z makes them
    # vanishingly rare at this size, so the frames agree
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_threaded_schedule_equals_single_thread(oracle_mod):
    sc = scenes.tris_uniform(1, 5000, 256, 256)
    a = oracle_mod.frame(sc, 8, 16, threads=1)
    b = oracle_mod.frame(sc, 8, 16, threads=8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_world_vertex_stage_projects_and_lights(oracle_mod):
    """vtxTransform (rasterMacros.h:14-21) with an identity matrix maps NDC
    (x,y) to ((x+1)W/2, (y+1)H/2); lighting is max(n.l,0)*mat + ambient,
    saturated and truncated (main.cpp:317-334)."""
    W, H = 64, 32
    pos = np.array([[-0.5, -0.5, 0.25, 0.5, -0.5, 0.25, 0.0, 0.5, 0.25]], np.float32)
    nrm = np.array([[0.0, 0.0, 1.0] * 3], np.float32)
    sc = scenes.WorldScene(pos, nrm, W, H, np.eye(4, dtype=np.float32).reshape(16),
                           np.array([0, 0, 1], np.float32), np.array([0.5, 0.25, 1.0], np.float32),
                           np.array([0.1, 0.1, 0.1], np.float32))
    st = oracle_mod.vertex_stage(sc, 8)
    assert list(st.xy[0]) == [16 * 256, 8 * 256, 48 * 256, 8 * 256, 32 * 256, 24 * 256]
    assert st.z[0, 0] == np.float32(0.25)
    r, g, b = (int(np.float32(0.5) * 255), int(np.float32(np.float32(0.25) + np.float32(0.1)) * 255), 255)
    want = (0xFF << 24) | (b << 16) | (g << 8) | int(np.float32(np.float32(0.5) + np.float32(0.1)) * 255)
    assert st.col[0, 0] == want
    assert r >= 0


def test_config1_scene_statistics(oracle_mod):
    sc = scenes.tris_uniform(1, 50000, 512, 512)
    p = sc.pos.reshape(-1, 3, 3).astype(np.float64)
    e1, e2 = p[:, 1, :2] - p[:, 0, :2], p[:, 2, :2] - p[:, 0, :2]
    area = 0.5 * (e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    assert np.all(area > 0)                         # emitted counter-clockwise
    assert 18.0 < np.median(area) < 22.0            # log-normal, median 20 px^2
    assert 0.9 < np.std(np.log(area)) < 1.1         # sigma = 1
    cen = p[:, :, :2].mean(axis=1)
    assert cen.min() >= -1 and cen.max() <= 513
    z = p[:, :, 2]
    assert z.min() >= 0.1 and z.max() < 1.0
    st = oracle_mod.vertex_stage(sc, 8)
    assert st.keep.sum() > 49000


def test_world_vertex_stage_perspective_divide(oracle_mod):
    """vtxTransform with w != 1 (rasterMacros.h:14-21): a perspective matrix
    with clip w = -z_eye and every quantity dyadic, so 1/w, x/w and the
    viewport maps are exact and the fixed-point coordinates have a closed
    form: x_fx = ((x / w + 1) * W / 2) * 2^S."""
    from fractions import Fraction as F
    W, H, S = 64, 32, 8
    # column-major: x' = x, y' = y, z' = z / 2 + 1, w' = -z
    m = np.zeros(16, np.float32)
    m[0], m[5], m[10], m[14], m[11] = 1.0, 1.0, 0.5, 1.0, -1.0
    verts = [(-0.5, -0.5, -2.0), (1.0, -1.0, -4.0), (0.25, 0.75, -1.0)]  # w = 2, 4, 1
    pos = np.array([sum(verts, ())], np.float32)
    nrm = np.array([[0.0, 0.0, 1.0] * 3], np.float32)
    sc = scenes.WorldScene(pos, nrm, W, H, m, np.array([0, 0, 1], np.float32), np.array([0.5, 0.5, 0.5], np.float32),
                           np.array([0.1, 0.1, 0.1], np.float32))
    st = oracle_mod.vertex_stage(sc, S)
    want_xy, want_z = [], []
    for x, y, z in verts:
        w = F(-z)
        want_xy += [int(((F(x) / w + 1) * F(W, 2)) * 2 ** S), int(((F(y) / w + 1) * F(H, 2)) * 2 ** S)]
        want_z.append(float((F(z) / 2 + 1) / w))
    assert list(st.xy[0]) == want_xy
    assert list(st.z[0]) == [np.float32(v) for v in want_z]
    # the triangle keeps its winding through the divide: counter-clockwise, kept
    assert st.keep[0] == 1


def test_threaded_schedule_under_thread_sanitizer(oracle_mod):
    """The threaded schedule (bins over 8 host threads, CPUBackend.cpp:522-571)
    built with ThreadSanitizer: no race is reported and the frame equals the
    single-threaded one at every bin size (oracle/tsan_main.c)."""
    import subprocess
    r = subprocess.run(["make", "-s", "-C", oracle_mod.HERE, "tsan"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stdout + r.stderr
    assert r.stdout.count("equal the single-threaded frame") == 4
