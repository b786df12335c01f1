"""Pins the CPU oracle (oracle/tomoj_oracle.c) before it is trusted as the checker.

TomoJ ships no tests, fixtures or golden vectors for this path, and its Java
code cannot run here (no JDK, dependencies not vendored).  Each oracle function
is therefore checked against what the mathematics fixes: closed forms,
invariants, special cases that reduce to an independent library (numpy's FFT),
and analytic line integrals.  Each check is chosen so that a plausible slip in
the restatement (a swapped index, a wrong sign of the tilt, a dropped edge
rule, a missing exit step) fails it.
"""
import math

import numpy as np
import pytest


def geom(O, W, H, D, **kw):
    return O.make_geom(W, H, D, nthreads=4, **kw)


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("theta", [-60.0, -17.5, 0.0, 33.0, 60.0])
def test_euler_single_axis_closed_form(oracle, theta):
    """E(0,theta,0) = [[c,0,-s],[0,1,0],[s,0,c]] (MatrixUtils.java:238-246)."""
    E = oracle.tilt_eulers([theta])[0]
    t = theta / 180.0 * math.pi
    ref = np.array([[math.cos(t), 0, -math.sin(t)], [0, 1, 0], [math.sin(t), 0, math.cos(t)]])
    assert np.allclose(E, ref, atol=1e-15)


def test_euler_general_is_rotation(oracle):
    E = oracle.euler_matrix(23.0, 41.0, -77.0)
    assert np.allclose(E @ E.T, np.eye(3), atol=1e-14)
    assert abs(np.linalg.det(E) - 1) < 1e-14


def test_python_euler_matches_oracle(oracle):
    """The package's Euler matrices (used to drive the GPU) equal the oracle's bit for bit."""
    from tomoj_amd.reconstruction import tilt_eulers
    tilts = np.arange(-60, 61, 1.0)
    assert np.array_equal(tilt_eulers(tilts), oracle.tilt_eulers(tilts))
    assert np.array_equal(tilt_eulers(tilts, 12.5, False), oracle.tilt_eulers(tilts, 12.5, False))


def test_fp_factors_single_axis(oracle):
    """dirX = tan(theta), dirY = 0, normIncrement = 1/|cos(theta)| (VoxelProjector3D.java:692-694)."""
    for th in (-60.0, -5.0, 25.0, 60.0):
        f = oracle.fp_factors(oracle.tilt_eulers([th])[0])
        t = th / 180 * math.pi
        assert f[6] == pytest.approx(math.tan(t), rel=1e-15)
        assert f[7] == 0.0
        assert f[8] == pytest.approx(1 / abs(math.cos(t)), rel=1e-15)


# ---------------------------------------------------------------------------
# samplers
# ---------------------------------------------------------------------------

def test_getpixel_linear_exact_and_edges(oracle):
    """Bilinear of a linear function is exact inside; the last column, last row and
    last slice return the nearest sample (TomoReconstruction2.java:163-164)."""
    W, H, D = 8, 6, 4
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    vol = (0.5 * x + 0.25 * y + 2.0 * z).astype(np.float32)
    assert oracle.getpixel(vol, 2.5, 3.25, 1) == pytest.approx(0.5 * 2.5 + 0.25 * 3.25 + 2, abs=1e-6)
    # x in [W-1, W): nearest = value at xbase
    assert oracle.getpixel(vol, 7.6, 2.0, 1) == pytest.approx(0.5 * 7 + 0.25 * 2 + 2)
    # last slice: no interpolation in x or y
    assert oracle.getpixel(vol, 2.5, 3.25, D - 1) == pytest.approx(0.5 * 2 + 0.25 * 3 + 2 * (D - 1))
    # slightly negative x: xbase 0, fraction 0
    assert oracle.getpixel(vol, -1e-12, 1.0, 0) == pytest.approx(0.25)


def test_ij_interpolation_closed_form(oracle):
    """ImageJ getInterpolatedValue: bilinear inside, clamped in the last-pixel band."""
    img = np.arange(20, dtype=np.float32).reshape(4, 5)  # value = 5*y + x
    assert oracle.ij_interpolated_value(img, 1.25, 2.5) == pytest.approx(5 * 2.5 + 1.25)
    assert oracle.ij_interpolated_value(img, 4.5, 1.0) == pytest.approx(5 + 4)       # x clamped
    assert oracle.ij_interpolated_value(img, 2.0, 3.75) == pytest.approx(15 + 2)     # y clamped
    assert oracle.ij_interpolated_value(img, 5.0, 1.0) == 0.0                         # outside


# ---------------------------------------------------------------------------
# forward projection
# ---------------------------------------------------------------------------

def test_fp_constant_volume_zero_tilt(oracle):
    """theta = 0, constant c: every ray stays inside, proj = c*D, norm = D, no exit step."""
    W, H, D = 12, 5, 7
    g = geom(oracle, W, H, D)
    vol = np.full((D, H, W), 0.75, np.float32)
    p, n = oracle.project(g, vol, oracle.tilt_eulers([0.0]))
    assert np.all(p == np.float32(0.75 * D)) and np.all(n == D)


@pytest.mark.parametrize("theta", [-50.0, 20.0, 37.0])
def test_fp_constant_volume_tilted_counts(oracle, theta):
    """Constant c = 0.5 (sums exact in float): proj = c * normInc * n_in and
    norm = normInc * (n_in + exited), n_in and 'exited' counted from the closed-form
    ray x(z) = x0 + (z - vcz) tan(theta) (pixels whose samples come within 1e-6
    of a border are skipped)."""
    W, H, D = 40, 3, 24
    g = geom(oracle, W, H, D)
    vol = np.full((D, H, W), 0.5, np.float32)
    E = oracle.tilt_eulers([theta])
    p, n = oracle.project(g, vol, E)
    t = theta / 180 * math.pi
    ni = 1 / abs(math.cos(t))
    pc0, vcx, vcz = (W - 1) / 2, (W - 1) / 2, (D - 1) / 2
    zs = np.arange(D)
    checked = 0
    for i in range(W):
        x = (i - pc0) / math.cos(t) + (zs - vcz) * math.tan(t) + vcx
        if np.any(np.abs(x) < 1e-6) or np.any(np.abs(x - W) < 1e-6):
            continue
        inside = (x >= 0) & (x < W)
        n_in = int(inside.sum())
        if n_in == 0:
            assert math.isnan(p[0, 1, i]) and n[0, 1, i] == 0
            continue
        last = np.nonzero(inside)[0].max()
        exited = 1 if last < D - 1 else 0
        assert p[0, 1, i] == pytest.approx(0.5 * ni * n_in, rel=1e-6)
        assert n[0, 1, i] == pytest.approx(ni * (n_in + exited), rel=1e-6)
        checked += 1
    assert checked > W // 2


def test_fp_linear_ramp(oracle):
    """f = a*x + b is interpolated exactly, except the nearest rule on the last slice
    and at x >= W-1: the projection equals normInc * sum_z f(x_z) with those rules."""
    W, H, D, th = 30, 4, 10, 25.0
    g = geom(oracle, W, H, D)
    x = np.arange(W, dtype=np.float64)
    vol = np.broadcast_to((0.125 * x + 1.0).astype(np.float32), (D, H, W)).copy()
    p, _ = oracle.project(g, vol, oracle.tilt_eulers([th]))
    t = th / 180 * math.pi
    ni = 1 / math.cos(t)
    for i in range(W):
        xs = (i - (W - 1) / 2) / math.cos(t) + (np.arange(D) - (D - 1) / 2) * math.tan(t) + (W - 1) / 2
        inside = (xs >= 0) & (xs < W)
        if not inside.any():
            continue
        vals = []
        for z, xv in zip(range(D), xs):
            if not inside[z]:
                continue
            xb = int(xv)
            if xb == W - 1 or z == D - 1:
                vals.append(0.125 * xb + 1.0)
            else:
                vals.append(0.125 * xv + 1.0)
        assert p[0, 2, i] == pytest.approx(ni * sum(vals), rel=2e-5), i


def test_fp64_matches_analytic_line_integrals(oracle):
    """Projection of a voxelised phantom matches the analytic line integrals in the
    same geometry (orientation, tilt sign and centres); a flipped tilt sign is far off."""
    W, H, D = 64, 8, 48
    ell = np.zeros((1, 16))
    ell[0, 0] = 1.0
    ell[0, 1:4] = [9.0, 0.0, -6.0]       # off-centre in x and z: sensitive to the tilt sign
    ell[0, 4:7] = [10.0, 30.0, 7.0]
    ell[0, 7:16] = np.eye(3).ravel()
    vol = oracle.phantom_volume(ell, W, H, D).astype(np.float64)
    g = geom(oracle, W, H, D)
    E = oracle.tilt_eulers([35.0])
    p, _ = oracle.project(g, vol, E, f64=True)
    ref = oracle.phantom_project(ell, E, W, H)
    wrong = oracle.phantom_project(ell, oracle.tilt_eulers([-35.0]), W, H)
    row = slice(2, 6)
    good = np.linalg.norm(p[0, row] - ref[0, row]) / np.linalg.norm(ref[0, row])
    bad = np.linalg.norm(p[0, row] - wrong[0, row]) / np.linalg.norm(ref[0, row])
    assert good < 0.08 and bad > 0.5


def test_analytic_sphere_closed_form(oracle):
    """Chord of a sphere of radius r at distance d from its centre: 2*sqrt(r^2-d^2)."""
    ell = np.zeros((1, 16))
    ell[0, 0] = 0.5
    ell[0, 4:7] = 5.0
    ell[0, 7:16] = np.eye(3).ravel()
    W = H = 15
    p = oracle.phantom_project(ell, oracle.tilt_eulers([0.0, 40.0]), W, H)
    for k in range(2):
        for (i, j) in [(7, 7), (9, 7), (7, 3), (10, 9), (12, 10)]:
            d2 = (i - 7) ** 2 + (j - 7) ** 2
            ref = 0.5 * 2 * math.sqrt(25 - d2) if d2 < 25 else 0.0
            assert p[k, j, i] == pytest.approx(ref, abs=1e-12)


# ---------------------------------------------------------------------------
# residual
# ---------------------------------------------------------------------------

def test_difference_known_answers(oracle):
    """diff = ((exp - th) / norm) * factor; NaN in exp or th leaves 0 and no error term."""
    exp = np.array([3.0, 1.0, np.nan, 2.0], np.float32)
    th = np.array([1.0, 1.5, 1.0, np.nan], np.float32)
    nm = np.array([2.0, 0.5, 1.0, 1.0], np.float32)
    d, e = oracle.difference(exp, th, nm, 0.5)
    assert list(d) == [0.5, -0.5, 0.0, 0.0]
    assert e == pytest.approx(0.5)


# ---------------------------------------------------------------------------
# backprojection
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("theta", [0.0, 30.0, -45.0])
def test_bp_constant_image(oracle, theta):
    """A constant image v adds exactly v to every voxel whose projection lands in
    [0,W) x [0,H) and nothing elsewhere."""
    W, H, D = 20, 6, 9
    g = geom(oracle, W, H, D)
    vol = np.zeros((D, H, W), np.float32)
    oracle.backproject(g, vol, np.full((1, H, W), 0.25, np.float32), oracle.tilt_eulers([theta]))
    t = theta / 180 * math.pi
    z, y, x = np.meshgrid(np.arange(D) - (D - 1) / 2, np.arange(H), np.arange(W) - (W - 1) / 2, indexing="ij")
    xx = x * math.cos(t) - z * math.sin(t) + (W - 1) / 2
    inside = (xx >= 0) & (xx < W)
    far = (np.abs(xx) > 1e-9) & (np.abs(xx - W) > 1e-9)
    assert np.all(vol[inside & far] == np.float32(0.25))
    assert np.all(vol[~inside & far] == 0)


def test_bp_delta_image_tent(oracle):
    """A single-pixel image spreads as the bilinear tent max(0, 1 - |xx - i0|) on row j0."""
    W, H, D, th = 24, 5, 11, 28.0
    g = geom(oracle, W, H, D)
    img = np.zeros((1, H, W), np.float32)
    img[0, 2, 9] = 1.0
    vol = np.zeros((D, H, W), np.float32)
    oracle.backproject(g, vol, img, oracle.tilt_eulers([th]))
    t = th / 180 * math.pi
    z, x = np.meshgrid(np.arange(D) - (D - 1) / 2, np.arange(W) - (W - 1) / 2, indexing="ij")
    xx = x * math.cos(t) - z * math.sin(t) + (W - 1) / 2
    tent = np.maximum(0, 1 - np.abs(xx - 9)).astype(np.float32)
    assert np.allclose(vol[:, 2, :], tent, atol=1e-6)
    assert np.all(vol[:, [0, 1, 3, 4], :] == 0)


def test_bp_scrubs_nan_and_inf(oracle):
    """NaN samples are skipped; a sum that becomes Inf is reset to 0 (VoxelProjector3D.java:291-300).
    ImageJ's bilinear formula multiplies the next row by its (zero) weight, so a NaN
    in the row below also makes the sample NaN: row 0 skips image 0 too."""
    W, H, D = 6, 3, 2
    g = geom(oracle, W, H, D)
    vol = np.zeros((D, H, W), np.float32)
    img = np.full((2, H, W), 1.0, np.float32)
    img[0, 1, :] = np.nan
    img[1, 2, :] = np.float32(3e38)
    vol[:, 2, :] = np.float32(3e38)
    oracle.backproject(g, vol, img, oracle.tilt_eulers([0.0, 0.0]))
    assert np.all(vol[:, 0, :] == 1.0)
    assert np.all(vol[:, 1, :] == 1.0)
    assert np.all(vol[:, 2, :] == 0.0)


# ---------------------------------------------------------------------------
# loops
# ---------------------------------------------------------------------------

def test_order_shuffle_perpendicular_chain(oracle):
    """Each next projection is the remaining one closest to perpendicular to the
    previous: from -60 the chain is -60, 30, -58, 32, ... (orderShuffleART)."""
    tilts = np.arange(-60, 61, 2.0)
    order = oracle.order_shuffle(oracle.tilt_eulers(tilts), 0)
    assert sorted(order) == list(range(61))
    assert list(order[:6]) == [0, 45, 1, 46, 2, 47]
    rem = set(range(61))
    prev = order[0]
    rem.discard(prev)
    for o in order[1:]:
        best = min(abs(math.cos(math.radians(tilts[prev] - tilts[r]))) for r in rem)
        assert abs(math.cos(math.radians(tilts[prev] - tilts[o]))) <= best + 1e-12
        rem.discard(o)
        prev = o


def test_python_order_shuffle_matches_oracle(oracle):
    """The library's orderShuffleART (tj_order_shuffle, host code) against the
    oracle: uniform tilts, and seeded uneven tilt sets with a repeated angle
    (equal |cos| distances, where the first minimum in index order wins)."""
    from tomoj_amd.reconstruction import order_shuffle
    E = oracle.tilt_eulers(np.arange(-60, 61, 1.0))
    for first in (0, 17, 60, 120):
        assert list(order_shuffle(E, first)) == list(oracle.order_shuffle(E, first))
    rng = np.random.default_rng(41)
    for n in (2, 7, 41, 80):
        tilts = np.sort(rng.uniform(-70.0, 70.0, n))
        if n > 2:
            tilts[n // 2] = tilts[n // 2 - 1]
        E = oracle.tilt_eulers(tilts)
        for first in (0, n // 3, n - 1):
            assert list(order_shuffle(E, first)) == list(oracle.order_shuffle(E, first))


def test_sirt_fixed_point(oracle):
    """If the measured projections are the projections of x*, SIRT leaves x* unchanged
    (every difference is exactly 0)."""
    W, H, D = 16, 4, 8
    g = geom(oracle, W, H, D)
    rng = np.random.default_rng(1)
    x = rng.random((D, H, W), dtype=np.float32)
    E = oracle.tilt_eulers(np.arange(-40, 41, 20.0))
    exps, _ = oracle.project(g, x, E)
    vol = x.copy()
    err = oracle.ossart(g, vol, exps, E, 1, 1.0, 0)
    assert np.array_equal(vol, x)
    assert err[0, 0] == 0.0


def test_convert_diff_to_error_known_answers(oracle):
    """convertDiffToErrorImage (VoxelProjector3D.java:315-324): float squares,
    NaN kept, overflow to +inf as Java's float multiply gives."""
    d = np.array([1.5, -2.0, np.nan, -0.0, 3e19, -1e-30], np.float32)
    got = oracle.convert_diff_to_error(d)
    assert got[0] == 2.25 and got[1] == 4.0 and np.isnan(got[2])
    assert got[3] == 0.0 and not np.signbit(got[3])
    assert got[4] == np.inf and got[5] == 0.0


@pytest.mark.parametrize("n", [1, 4])
def test_error_volume_constant_difference(oracle, n):
    """The error volume of one OS-SART iteration in closed form.  A zero
    volume, n copies of the theta = 0 projection, measured images all equal to
    c: the forward projection is 0 with ray norm D, every difference is
    (c / D) * (1/n) (the factor lambda / update is inside it), so the volume
    becomes n * (c / D) / n = c / D and the error volume, the backprojection of
    the squared differences image by image, n * (c / (D n))^2 = c^2 / (D^2 n)
    at every voxel."""
    W, H, D = 12, 3, 8
    c = 0.75
    g = geom(oracle, W, H, D)
    E = oracle.tilt_eulers([0.0] * n)
    exps = np.full((n, H, W), c, np.float32)
    vol = np.zeros((D, H, W), np.float32)
    ev = np.zeros((D, H, W), np.float32)
    oracle.ossart(g, vol, exps, E, 1, 1.0, n, errvol=ev)
    d = np.float32(np.float32(c / D) * np.float32(1.0 / n))
    assert np.allclose(vol, c / D, rtol=1e-6)
    assert np.allclose(ev, np.float32(d * d) * n, rtol=1e-6)
    # without an error volume the reconstruction is the same
    vol2 = np.zeros((D, H, W), np.float32)
    oracle.ossart(g, vol2, exps, E, 1, 1.0, n)
    assert np.array_equal(vol, vol2)


def test_error_volume_is_backprojection_of_squares(oracle):
    """One subset: the error volume equals the backprojection of the subset's
    squared differences, and the reconstruction does not see the squares."""
    W, H, D = 20, 4, 10
    g = geom(oracle, W, H, D)
    tilts = [-30.0, 10.0, 45.0]
    E = oracle.tilt_eulers(tilts)
    ell = oracle.ellipsoids(5, 3, W, H, D)
    exps = oracle.phantom_project(ell, E, W, H).astype(np.float32)
    exps[1, 2, 3] = np.nan
    vol = np.zeros((D, H, W), np.float32)
    ev = np.zeros((D, H, W), np.float32)
    oracle.ossart(g, vol, exps, E, 1, 1.0, 3, errvol=ev)
    order = oracle.order_shuffle(E, 0)
    factor = float(np.float32(1.0 / 3))
    diffs, _, _, _ = oracle.project_and_difference(g, np.zeros((D, H, W), np.float32), exps[order], E[order],
                                                   factor)
    ref_vol = oracle.backproject(g, np.zeros((D, H, W), np.float32), diffs, E[order])
    ref_ev = oracle.backproject(g, np.zeros((D, H, W), np.float32),
                                np.stack([oracle.convert_diff_to_error(d) for d in diffs]), E[order])
    assert np.array_equal(vol, ref_vol)
    assert np.array_equal(ev, ref_ev)
    assert (ev >= 0).all()


def test_sirt_residual_decreases(oracle):
    W, H, D = 32, 4, 16
    g = geom(oracle, W, H, D)
    tilts = np.arange(-60, 61, 5.0)
    E = oracle.tilt_eulers(tilts)
    ell = oracle.ellipsoids(7, 4, W, H, D)
    exps = oracle.phantom_project(ell, E, W, H).astype(np.float32)
    vol = np.zeros((D, H, W), np.float32)
    err = oracle.ossart(g, vol, exps, E, 6, 1.0, 0)
    assert np.all(np.diff(err[:, 0]) < 0)


def test_ossart_f32_close_to_f64(oracle):
    """The FP32-faithful loop and the FP64 twin agree to FP32 rounding level."""
    W, H, D = 24, 4, 12
    g = geom(oracle, W, H, D)
    E = oracle.tilt_eulers(np.arange(-60, 61, 10.0))
    ell = oracle.ellipsoids(3, 3, W, H, D)
    exps = oracle.phantom_project(ell, E, W, H)
    v32 = np.zeros((D, H, W), np.float32)
    v64 = np.zeros((D, H, W), np.float64)
    oracle.ossart(g, v32, exps.astype(np.float32), E, 3, 1.0, 4)
    oracle.ossart(g, v64, exps, E, 3, 1.0, 4, f64=True)
    assert np.linalg.norm(v32 - v64) / np.linalg.norm(v64) < 1e-5


# ---------------------------------------------------------------------------
# weighting
# ---------------------------------------------------------------------------

def test_fft2_matches_numpy(oracle):
    rng = np.random.default_rng(3)
    for (sy, sx) in [(8, 16), (6, 10), (32, 32)]:
        re = rng.standard_normal((sy, sx))
        im = rng.standard_normal((sy, sx))
        r, i = oracle.fft2(re, im, -1)
        ref = np.fft.fft2(re + 1j * im)
        assert np.allclose(r + 1j * i, ref, atol=1e-10)


def test_weighting_matches_numpy(oracle):
    rng = np.random.default_rng(4)
    img = rng.standard_normal((16, 32)).astype(np.float32)
    mask = rng.random((16, 32)).astype(np.float32)
    mask = 0.5 * (mask + np.roll(mask[::-1, ::-1], (1, 1), (0, 1)))  # Hermitian-symmetric
    out = oracle.weighting(img, mask)
    ref = np.real(np.fft.ifft2(np.fft.fft2(img.astype(np.float64)) * mask))
    assert np.allclose(out, ref, atol=1e-5)


def test_weighting_mask_special_cases(oracle):
    """Column ci = 0 sees sinc(0) = 1 for every tilt: weight = N * diameter.  For one
    tilt theta: mask = 1/(diameter * max(0.005, sinc(pi*diameter*ci*sin(theta)/sx)))."""
    sx, sy, dia = 32, 8, 16.0
    tilts = np.arange(-60, 61, 20.0)
    m = oracle.weighting_mask(sx, sy, tilts, dia)
    assert np.allclose(m[:, 0], np.float32(1 / (len(tilts) * dia)))
    th = 30.0
    m1 = oracle.weighting_mask(sx, sy, [th], dia)
    for i in [1, 3, 7, 16, 20]:
        ci = i if i <= sx / 2 else i - sx
        v = dia * math.pi * ci * math.sin(math.radians(th)) / sx
        w = max(math.sin(v) / v, 0.005)
        assert m1[2, i] == pytest.approx(1 / (w * dia), rel=1e-5)


def test_noise_statistics(oracle):
    z = oracle.gaussian_noise(12345, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_fista_step_closed_form(oracle):
    """FistaProjector3D.endOfIteration: t runs 0 -> 1 -> golden ratio -> ...,
    so the first step restores the start-of-iteration volume (factor -1), the
    second leaves the volume alone (factor 0), the third extrapolates by
    (phi - 1) / t3; NaN differences are skipped."""
    rng = np.random.default_rng(4)
    prev = rng.standard_normal(64).astype(np.float32)
    vol = rng.standard_normal(64).astype(np.float32)
    v = vol.copy()
    t = oracle.fista_step(v, prev, 0.0)
    assert t == 1.0 and np.allclose(v, prev, rtol=1e-6, atol=1e-6)
    for k in range(64):      # x + (float)(-1 * (float)(x - p))
        assert v[k] == np.float32(vol[k] + np.float32(-float(np.float32(vol[k] - prev[k]))))
    v = vol.copy()
    t2 = oracle.fista_step(v, prev, 1.0)
    assert t2 == (1 + 5 ** 0.5) / 2 and np.array_equal(v, vol)
    v = vol.copy()
    t3 = oracle.fista_step(v, prev, t2)
    assert t3 == (1 + (1 + 4 * t2 * t2) ** 0.5) / 2
    f = (t2 - 1) / t3
    for k in range(64):      # float difference, double scale, float add
        d = np.float32(vol[k] - prev[k])
        assert v[k] == np.float32(vol[k] + np.float32(f * float(d)))
    v = vol.copy()
    p2 = prev.copy()
    p2[5] = np.nan
    oracle.fista_step(v, p2, t2)
    assert v[5] == vol[5]


def test_normalisation_roi_and_statistics(oracle):
    """Projection supply with the identity alignment: the ROI is round(W cos t)
    columns centred by integer division (at least min_size), the statistics
    are the sample mean and standard deviation (n - 1) of the ROI."""
    assert oracle.roi_rect(512, 0.0) == (0, 512)
    assert oracle.roi_rect(512, 60.0) == (128, 384)              # round(256) = 256 columns
    assert oracle.roi_rect(100, 45.0) == (14, 85)                # round(70.71) = 71, (100-71)/2 = 14
    assert oracle.roi_rect(100, 89.0, min_size=20) == (40, 60)
    rng = np.random.default_rng(9)
    imgs = rng.standard_normal((3, 12, 40)).astype(np.float32) * 3 + 5
    tilts = [-50.0, 0.0, 30.0]
    out, means, sds = oracle.normalized_projections(imgs, tilts)
    for k, t in enumerate(tilts):
        x0, x1 = oracle.roi_rect(40, t)
        roi = imgs[k][:, x0:x1].astype(np.float64)
        assert np.isclose(means[k], roi.mean(), rtol=1e-13)
        assert np.isclose(sds[k], roi.std(ddof=1), rtol=1e-12)
        assert np.array_equal(out[k], ((imgs[k].astype(np.float64) - means[k]) / sds[k]).astype(np.float32))


def test_affine_inverse_states(oracle):
    """createInverse per transform state: translation and scale invert entry by
    entry (exact), the general state through the determinant (T Tinv = I)."""
    assert np.array_equal(oracle.affine_inverse([1, 0, 0, 1, 3.5, -2]), [1, 0, 0, 1, -3.5, 2])
    assert np.array_equal(oracle.affine_inverse([2, 0, 0, 4, 1, 2]), [0.5, 0, 0, 0.25, -0.5, -0.5])
    assert np.array_equal(oracle.affine_inverse([0, 1, -1, 0, 0, 0]), [0, -1, 1, 0, 0, 0])   # 90 degrees
    assert oracle.affine_inverse([0, 0, 0, 0, 1, 1]) is None
    rng = np.random.default_rng(4)
    for _ in range(20):
        m = rng.standard_normal(6)
        r = oracle.affine_inverse(m)
        A = np.array([[m[0], m[2], m[4]], [m[1], m[3], m[5]], [0, 0, 1]])
        B = np.array([[r[0], r[2], r[4]], [r[1], r[3], r[5]], [0, 0, 1]])
        assert np.allclose(A @ B, np.eye(3), atol=1e-12)


def test_affine_resample_closed_forms(oracle):
    """applyTransformOnImage: an integer translation shifts the image exactly,
    a 90-degree rotation of a square image permutes it, a 2x scale samples
    midpoints (exact bilinear of a linear image), and pixels that map outside
    take the fill (mean, 0 when normalised, NaN, or 0)."""
    rng = np.random.default_rng(5)
    h, w = 10, 14
    img = (rng.standard_normal((h, w)) * 3 + 8).astype(np.float32)
    out, _ = oracle.aligned_projections(img[None], [[1, 0, 0, 1, 3, -2]], normalize=False, fill=0)
    want = np.zeros_like(img)
    want[0:h - 2, 3:] = img[2:, :w - 3]
    assert np.array_equal(out[0], want)
    # rotation by 90 degrees: dst[j, i] = src[n - i, j] (i >= 1), NaN-filled column 0
    n = 12
    sq = (rng.standard_normal((n, n)) + 1).astype(np.float32)
    out, _ = oracle.aligned_projections(sq[None], [[0, 1, -1, 0, 0, 0]], normalize=False, fill=2)
    assert np.isnan(out[0][:, 0]).all()
    for i in range(1, n):
        assert np.array_equal(out[0][:, i], sq[n - i, :])
    # 2x scale of a linear image: bilinear samples are exact up to rounding
    lin = (0.25 * np.arange(w)[None, :] + 0.5 * np.arange(h)[:, None] + 1).astype(np.float32)
    out, _ = oracle.aligned_projections(lin[None], [[2, 0, 0, 2, 0, 0]], normalize=False, fill=0)
    jj, ii = np.mgrid[0:h, 0:w]
    xx, yy = 0.5 * (ii - w / 2) + w / 2, 0.5 * (jj - h / 2) + h / 2
    inner = (xx < w - 1) & (yy < h - 1)
    assert np.allclose(out[0][inner], (0.25 * xx + 0.5 * yy + 1)[inner], rtol=1e-6)
    # fills and normalisation with given statistics
    st = np.array([[2.0, 4.0]])
    for fill, normalize, val in [(1, False, 2.0), (1, True, 0.0), (0, False, 0.0)]:
        out, _ = oracle.aligned_projections(img[None], [[1, 0, 0, 1, 5, 0]], normalize=normalize, fill=fill,
                                            stats=st)
        assert (out[0][:, :5] == np.float32(val)).all()
    out, _ = oracle.aligned_projections(img[None], [[1, 0, 0, 1, 5, 0]], normalize=True, fill=1, stats=st)
    assert np.array_equal(out[0][:, 5:], ((img[:, :w - 5].astype(np.float64) - 2.0) / 4.0).astype(np.float32))


def test_affine_roi_rectangle(oracle):
    """getImageStatistics' polygon under a rotation-free transform: the identity
    gives the normalisation ROI over every row; a translation moves it."""
    for w, t in [(512, 0.0), (100, 45.0), (37, -61.0)]:
        x0, x1 = oracle.roi_rect(w, t)
        assert oracle.roi_rect_affine(w, 20, t, [1, 0, 0, 1, 0, 0]) == [x0, x1, 0, 20]
        assert oracle.roi_rect_affine(w, 20, t, [1, 0, 0, 1, 4, -3]) == [x0 + 4, x1 + 4, -3, 17]


# --------------------------------------------------------------------------
# total-variation denoising (TotalVariationDenoising_3D.denoise3D)
# --------------------------------------------------------------------------

def _tv_line_f64(vals, theta, g, dt, iterations):
    """One voxel-thick line worked out in double: the gradient has one
    non-zero component, so dd1 = 1 + dt/theta/g |du|; the last voxel keeps
    its value because TomoJ's divergence there is minus the voxel's own p,
    which stays 0 (its forward difference is 0)."""
    f = np.array(vals, np.float64)
    u = f.copy()
    p = np.zeros_like(f)
    for _ in range(iterations):
        du = np.append(np.diff(u), 0.0)
        p = (p - dt / theta * du) / (1 + dt / theta / g * np.abs(du))
        div = np.empty_like(p)
        div[0] = p[0]
        div[1:-1] = p[1:-1] - p[:-2]
        div[-1] = -p[-1]
        u = f - theta * div
    return u


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_tv_line_closed_form(oracle, axis):
    """A 3-voxel line along z, y or x: after one step the first voxel moves by
    dt (b-a)/dd1 towards its neighbour and the last one does not move; after
    three steps the line agrees with the double-precision working of the same
    recurrence.  Pins the axis each difference runs along, the sign of the
    p update, g's place in dd1, f (not u) as the base of the u update, and
    the boundary rule of TomoJ's divergence."""
    vals = [1.0, 3.0, 2.5]
    shape = [1, 1, 1]
    shape[axis] = 3
    vol = np.array(vals, np.float32).reshape(shape)
    theta, g, dt = 25.0, 2.0, 0.1
    one = oracle.tv_denoise3d(vol, theta, g, dt, 1).ravel()
    dd1 = 1 + dt / theta / g * 2.0
    assert one[0] == pytest.approx(1.0 + dt * 2.0 / dd1, rel=1e-6)
    assert one[2] == np.float32(2.5)
    three = oracle.tv_denoise3d(vol, theta, g, dt, 3).ravel()
    assert np.allclose(three, _tv_line_f64(vals, theta, g, dt, 3), rtol=1e-6)
    assert three[2] == np.float32(2.5)


def test_tv_constant_volume_is_fixed(oracle):
    vol = np.full((5, 4, 6), 0.75, np.float32)
    assert np.array_equal(oracle.tv_denoise3d(vol, 25, 1, 0.25, 7), vol)


def test_tv_slice_constant_volume_reduces_to_2d(oracle):
    """Equal slices: duz = 0 and p_z = 0 everywhere, so denoise3D must equal the
    separately structured 2-D routine (denoise2D) on each slice, bit for bit."""
    rng = np.random.default_rng(5)
    img = rng.standard_normal((11, 13)).astype(np.float32)
    vol = np.repeat(img[None], 4, axis=0)
    for theta, g, dt, n in [(25.0, 1.0, 0.25, 5), (12.0, 3.0, 0.1, 2)]:
        out = oracle.tv_denoise3d(vol, theta, g, dt, n)
        ref = oracle.tv_denoise2d(img, theta, g, dt, n)
        for z in range(4):
            assert np.array_equal(out[z], ref)


def test_tv_random_volume_double_precision(oracle):
    """A random volume against the same recurrence written with whole-array
    differences in double precision."""
    rng = np.random.default_rng(8)
    f = rng.standard_normal((6, 7, 9)).astype(np.float32)
    theta, g, dt, n = 20.0, 1.5, 0.2, 3
    out = oracle.tv_denoise3d(f, theta, g, dt, n)
    F = f.astype(np.float64)
    u = F.copy()
    p = np.zeros((3,) + F.shape)
    for _ in range(n):
        d = [np.zeros_like(u) for _ in range(3)]
        d[0][:, :, :-1] = np.diff(u, axis=2)   # x
        d[1][:, :-1, :] = np.diff(u, axis=1)   # y
        d[2][:-1, :, :] = np.diff(u, axis=0)   # z
        dd1 = 1 + dt / theta / g * np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
        p = (p - dt / theta * np.stack(d)) / dd1
        div = np.zeros_like(u)
        for c, ax in ((0, 2), (1, 1), (2, 0)):
            q = np.moveaxis(p[c], ax, 0)
            dv = np.empty_like(q)
            dv[0] = q[0]
            dv[1:-1] = q[1:-1] - q[:-2]
            dv[-1] = -q[-1]
            div += np.moveaxis(dv, 0, ax)
        u = F - theta * div
    assert np.abs(out - u).max() <= 1e-5 * np.abs(u).max()


def test_tv_lowers_total_variation(oracle):
    """Denoising a noisy smooth volume with TVMProjector3D's effective step
    (dt = regularizationLambda = 0.1) lowers its total variation and brings it
    closer to the clean volume (a sign error in the p update raises the
    variation)."""
    rng = np.random.default_rng(2)
    z, y, x = np.mgrid[0:12, 0:14, 0:16]
    clean = np.sin(x / 4.0) + np.cos(y / 5.0) + 0.1 * z
    noisy = (clean + 0.3 * rng.standard_normal(clean.shape)).astype(np.float32)

    def tv(v):
        v = v.astype(np.float64)
        return sum(np.abs(np.diff(v, axis=a)).sum() for a in range(3))

    out = oracle.tv_denoise3d(noisy, 25.0, 1.0, 0.1, 5)
    assert tv(out) < 0.5 * tv(noisy)
    assert np.abs(out - clean).mean() < np.abs(noisy - clean).mean()


# --------------------------------------------------------------------------
# Bayesian projector (BayesianVoxelProjector3D)
# --------------------------------------------------------------------------

def test_bayes_convolution_closed_forms(oracle):
    """The [1,-2,1]^3 stencil with zero outside the volume: a constant volume
    is separable, so the result is c times the product of the per-axis sums
    (0 inside, -1 on a face, -2 across a 1-voxel axis) and -lambda times that
    survives only at corners; a delta reproduces the stencil around it; a
    volume that does not vary along one axis gives 0 away from the borders."""
    lam = 0.25
    c = np.float32(3.0)
    vol = np.full((5, 4, 6), c, np.float32)
    d = oracle.bayes_convolve(vol, -lam)
    corner = np.zeros_like(d, bool)
    corner[np.ix_([0, 4], [0, 3], [0, 5])] = True
    assert np.all(d[corner] == np.float32(lam * c))
    assert np.all(d[~corner] == 0)
    flat = np.full((1, 3, 3), c, np.float32)   # D = 1: the z sum is -2
    assert oracle.bayes_convolve(flat, -lam)[0, 0, 0] == np.float32(-lam * c * -2 * -1 * -1)
    delta = np.zeros((5, 5, 5), np.float32)
    delta[2, 2, 2] = 1.0
    a = np.array([1.0, -2.0, 1.0])
    stencil = np.einsum("i,j,k->ijk", a, a, a)
    assert np.array_equal(oracle.bayes_convolve(delta, -lam)[1:4, 1:4, 1:4], (-lam * stencil).astype(np.float32))
    rng = np.random.default_rng(4)
    xz = rng.standard_normal((6, 1, 7)).astype(np.float32)
    assert np.all(oracle.bayes_convolve(np.repeat(xz, 5, axis=1), -lam)[:, 1:-1, :] == 0)


def test_bayes_convolution_random_volume(oracle):
    """A random volume against a zero-padded separable convolution in double."""
    rng = np.random.default_rng(6)
    vol = rng.standard_normal((7, 8, 9)).astype(np.float32)
    got = oracle.bayes_convolve(vol, -0.3)
    pad = np.pad(vol.astype(np.float64), 1)
    ref = np.zeros(vol.shape)
    a = (1.0, -2.0, 1.0)
    for i in range(3):
        for j in range(3):
            for k in range(3):
                ref += a[i] * a[j] * a[k] * pad[k:k + 7, j:j + 8, i:i + 9]
    assert np.allclose(got, (-0.3 * ref).astype(np.float32), rtol=1e-6, atol=1e-6)


def test_bayes_iteration_without_prior_is_sirt(oracle):
    """With lambda = 0 and alpha = 1 the distance volume holds only the
    backprojected differences: from a zero volume the first iteration equals
    a SIRT iteration bit for bit (the same sums in the same order), the next
    ones to float rounding."""
    W, H, D = 24, 5, 10
    E = oracle.tilt_eulers(np.arange(-60, 61, 15.0))
    exps = oracle.phantom_project(oracle.ellipsoids(5, 4, W, H, D), E, W, H).astype(np.float32)
    g = oracle.make_geom(W, H, D, nthreads=4)
    a = np.zeros((D, H, W), np.float32)
    b = np.zeros((D, H, W), np.float32)
    ea = oracle.bayes_iteration(g, a, exps, E, 1.0, 0.0)
    eb = oracle.ossart(g, b, exps, E, 1, 1.0, 0)[0]
    assert np.array_equal(a, b)
    assert np.allclose(ea, eb, rtol=1e-12)
    for _ in range(2):
        oracle.bayes_iteration(g, a, exps, E, 1.0, 0.0)
        oracle.ossart(g, b, exps, E, 1, 1.0, 0)
    assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max()


def test_bayes_end_step(oracle):
    """endOfIteration: rec += (float)alpha * dist, then MAX against 0 with
    positivity (NaN stays NaN)."""
    a = np.array([1.0, -3.0, np.nan], np.float32)
    d = np.array([2.0, 1.0, 1.0], np.float32)
    oracle._L().or_bayes_end(oracle._p(a), oracle._p(d), 3, 0.5, 1)
    assert a[0] == 2.0 and a[1] == 0.0 and np.isnan(a[2])
    assert np.array_equal(d, [1.0, 0.5, 0.5])
    b = np.array([-3.0], np.float32)
    oracle._L().or_bayes_end(oracle._p(b), oracle._p(np.array([1.0], np.float32)), 1, 0.5, 0)
    assert b[0] == -2.5


# --------------------------------------------------------------------------
# projection subsets (ReconstructionParameters.getAvailableIndexes)
# --------------------------------------------------------------------------

def test_available_indexes_known_lists(oracle):
    """EVEN / ODD keep that parity, at most N/2 images (so the last even image
    of an odd N is dropped); ALL keeps every image.  The library's host
    routine and the parameters object give the same lists."""
    import ctypes as C
    import tomoj_amd
    from tomoj_amd import parameters as RP
    cases = {
        (121, 11): list(range(0, 119, 2)), (121, 12): list(range(1, 120, 2)),
        (10, 11): [0, 2, 4, 6, 8], (10, 12): [1, 3, 5, 7, 9],
        (7, 11): [0, 2, 4], (7, 12): [1, 3, 5], (1, 11): [], (1, 12): [],
        (5, 10): [0, 1, 2, 3, 4], (5, 0): [0, 1, 2, 3, 4],
    }
    lib = tomoj_amd.load_library()
    for (n, t), want in cases.items():
        assert list(oracle.available_indexes(n, t)) == want
        out = (C.c_int * max(n, 1))()
        k = lib.tj_available_indexes(n, t, out)
        assert [out[i] for i in range(k)] == want
        p = RP.ReconstructionParameters.createOSSARTParameters(8, 8, 8, 1, 1.0, 0)
        p.setFscType(t)
        assert p.getAvailableIndexes(RP.TiltSeriesData(np.zeros((n, 2, 2), np.float32), [0.0] * n)) == want


def test_order_shuffle_available_uses_leading_matrices(oracle):
    """With a list of available images TomoJ chains positions 0..len-1 with
    the Euler matrices of IMAGES 0..len-1 (getEulerMatrix(im)) and emits the
    listed image at each position; the Python loop does the same."""
    from tomoj_amd.reconstruction import order_shuffle
    E = oracle.tilt_eulers(np.arange(-60, 61, 5.0))
    odd = oracle.available_indexes(len(E), 12)
    for first in (0, 3, len(odd) - 1):
        chain = oracle.order_shuffle(E[:len(odd)], first)
        want = [int(odd[k]) for k in chain]
        assert list(oracle.order_shuffle_available(E, odd, first)) == want
        assert order_shuffle(E, first, odd) == want
        # not the chain over the listed images' own matrices
        assert want != [int(odd[k]) for k in oracle.order_shuffle(E[odd], first)]


def test_osart_available_subsets_follow_full_count(oracle):
    """update 0 means N: with half the images listed no subset ever fills and
    the volume is untouched (err[0] = 0); with update 4 over 10 listed images
    the two full subsets run and the last two images are never projected."""
    W, H, D = 20, 4, 8
    E = oracle.tilt_eulers(np.arange(-60, 61, 6.0))          # 21 images
    exps = oracle.phantom_project(oracle.ellipsoids(8, 3, W, H, D), E, W, H).astype(np.float32)
    g = oracle.make_geom(W, H, D, nthreads=4)
    even = oracle.available_indexes(len(E), 11)              # 10 images
    vol = np.zeros((D, H, W), np.float32)
    err = oracle.ossart(g, vol, exps, E, 2, 1.0, 0, available=even)
    assert not vol.any() and (err[:, 0] == 0).all()
    got = np.zeros((D, H, W), np.float32)
    err = oracle.ossart(g, got, exps, E, 1, 1.0, 4, first_index=2, available=even)
    order = oracle.order_shuffle_available(E, even, 2)
    ref = np.zeros((D, H, W), np.float32)
    factor = float(np.float32(1.0 / 4))
    som = 0.0
    for s in (order[0:4], order[4:8]):
        diffs, errs, _, _ = oracle.project_and_difference(g, ref, exps[s], E[s], factor)
        oracle.backproject(g, ref, diffs, E[s])
        som += errs.sum()
    assert np.array_equal(got, ref)
    assert err[0, 0] == pytest.approx(np.sqrt(som / (len(E) * W * H)), rel=1e-12)


# --------------------------------------------------------------------------
# Fourier shell correlation (ResolutionEstimation.fsc)
# --------------------------------------------------------------------------

def test_fsc_mask_closed_form(oracle):
    """1 inside radius D, (1 + cos((d - D) pi / 20)) / 2 over the next 20
    voxels (0.5 at d = D + 10), 0 beyond."""
    D, H, W = 4, 3, 64
    m = oracle.fsc_mask(D, H, W)
    zc, yc, xc = 2, 1, 32                      # (W/2, H/2, D/2) rounded to voxels
    assert m[zc, yc, xc + 3] == 1.0             # dist sqrt(9 + 0.25) < 4
    assert m[zc, 1, xc + 14] == np.float32((1 + np.cos((np.sqrt(14 ** 2 + 0.25) - 4) * np.pi / 20)) / 2)
    assert m[zc, 1, xc + 30] == 0.0
    mm = oracle.fsc_mask(4, 4, 64)              # H even: dy = 0 on row 2
    assert mm[2, 2, 32 + 14] == np.float32(0.5)


def test_fsc_known_answers(oracle):
    """Identical halves correlate to 1 with no error, opposite ones to -1,
    a scale changes nothing; the first column is i/W, the third 2/sqrt(n) with
    n = 1 for the DC shell; the shell counts cover every frequency below R."""
    rng = np.random.default_rng(3)
    v = rng.standard_normal((8, 12, 16)).astype(np.float32)
    f = oracle.fsc(v, v)
    R = 16 // 2 + 1
    assert f.shape == (4, R)
    assert np.allclose(f[1], 1.0, atol=1e-12) and np.all(f[3] == 0)
    assert np.allclose(oracle.fsc(v, -v)[1], -1.0, atol=1e-12)
    assert np.allclose(oracle.fsc(v, 3 * v)[1], 1.0, atol=1e-12)
    assert np.array_equal(f[0], np.arange(R) / 16)
    assert f[2][0] == 2.0
    cnt = (2 / f[2]) ** 2
    assert cnt.sum() == pytest.approx(np.sum(np.round(np.sqrt(
        np.add.outer(np.add.outer(np.fft.fftfreq(8) ** 2, np.fft.fftfreq(12) ** 2), np.fft.fftfreq(16) ** 2))
        * 16 + 1e-9) < R))


def test_fsc_transform_convention(oracle):
    """The shell sums use the forward, unscaled transform over all three axes:
    for a tiny volume the DC shell's |E|^2 is the squared masked sum, and the
    whole FSC matches a direct DFT evaluation."""
    rng = np.random.default_rng(9)
    D, H, W = 2, 3, 4
    a = rng.standard_normal((D, H, W)).astype(np.float32)
    b = rng.standard_normal((D, H, W)).astype(np.float32)
    m = oracle.fsc_mask(D, H, W)
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")

    def dft(v):
        out = np.zeros((D, H, W), complex)
        for kz in range(D):
            for ky in range(H):
                for kx in range(W):
                    ph = np.exp(-2j * np.pi * (kz * z / D + ky * y / H + kx * x / W))
                    out[kz, ky, kx] = np.sum((v * m).astype(np.float64) * ph)
        return out

    E, O = dft(a), dft(b)
    f = oracle.fsc(a, b)
    # shell of every (kz, ky, kx), then the FSC of each shell
    fr = lambda n: np.where(np.arange(n) <= n // 2, np.arange(n), np.arange(n) - n) / n
    wi = np.floor(np.sqrt(np.add.outer(np.add.outer(fr(D) ** 2, fr(H) ** 2), fr(W) ** 2)) * W + 0.5)
    for s in range(W // 2 + 1):
        sel = wi == s
        if not sel.any():
            continue
        want = (E[sel] * np.conj(O[sel])).real.sum() / np.sqrt((abs(E[sel]) ** 2).sum() * (abs(O[sel]) ** 2).sum())
        assert f[1][s] == pytest.approx(want, rel=1e-9)


# --------------------------------------------------------------------------
# hysteresis threshold of the projection supply (TiltSeries.applyThreshold)
# --------------------------------------------------------------------------

def test_threshold_ramp_levels(oracle):
    """A 0..99 ramp: 256 bins over [0, 99]; with the 25% fraction the last bin
    whose running count stays below 25 is bin 61, level 61/256*99 = 23.6, so
    exactly the values 0..23 take the fill value."""
    img = np.arange(100, dtype=np.float32).reshape(10, 10)
    for fill, val in ((2, None), (0, 0.0), (1, np.float32(49.5))):
        out = oracle.threshold(img, 0.25, fill=fill)
        low = out.ravel()[:24]
        assert (np.isnan(low).all() if val is None else (low == val).all())
        assert np.array_equal(out.ravel()[24:], img.ravel()[24:])


def test_threshold_hysteresis_connectivity(oracle):
    """Low (0), middle (50) and high (100) values: the middle ones take the
    fill value only when connected to a filled pixel through index -+1 (which
    wraps from the end of one row to the start of the next) or -+width; a NaN
    already in the image seeds the NaN fill."""
    W, H = 8, 6
    img = np.full((H, W), 100.0, np.float32)
    img[0, 7] = 0.0                               # low
    for y, x in ((1, 0), (1, 1), (2, 1), (4, 5)):
        img[y, x] = 50.0                          # middle
    n_low, n_mid, n = 1, 4, W * H
    t, h = (n_low + 1.5) / n, (n_low + n_mid + 1.5) / n
    out = oracle.threshold(img, t, h, fill=0)
    want = img.copy()
    for y, x in ((0, 7), (1, 0), (1, 1), (2, 1)):
        want[y, x] = 0.0
    assert np.array_equal(out, want)              # (4, 5) is not connected
    nan_img = img.copy()
    nan_img[5, 5] = np.nan                        # index 45 = (4, 5) + width
    out = oracle.threshold(nan_img, t, h, fill=2)
    assert np.isnan(out[4, 5]) and np.isnan(out[1, 0]) and out[3, 3] == 100.0


def test_threshold_constant_image_unchanged(oracle):
    img = np.full((5, 7), 3.0, np.float32)
    assert np.array_equal(oracle.threshold(img, 0.5, 0.9, fill=2), img)


def test_polygon_roi_statistics(oracle):
    """The polygon ROI of a rotating alignment: an axis-aligned polygon gives
    the rectangle's statistics exactly; a rotated quadrilateral takes the
    pixels whose centres pass a convex-polygon sign test (no centre lies on an
    edge here), checked against numpy's mean and sample deviation; a polygon
    larger than the image takes every pixel."""
    rng = np.random.default_rng(12)
    h, w = 30, 40
    img = rng.standard_normal((h, w)).astype(np.float32)
    rect = oracle.poly_stats(img, [3, 20, 20, 3, 4, 4, 17, 17])
    assert rect == oracle.rect_stats(img, [3, 20, 4, 17])
    xs, ys = np.array([2, 33, 27, -4]), np.array([1, 9, 36, 24])   # odd dx - dy: no centre on an edge line
    yy, xx = np.mgrid[0:h, 0:w] + 0.5
    inside = np.ones((h, w), bool)
    for i in range(4):
        j = (i + 1) % 4
        cross = (xs[j] - xs[i]) * (yy - ys[i]) - (ys[j] - ys[i]) * (xx - xs[i])
        assert not (cross == 0).any()
        inside &= cross > 0
    m, sd = oracle.poly_stats(img, list(xs) + list(ys))
    sel = img[inside].astype(np.float64)
    assert m == pytest.approx(sel.mean(), rel=1e-12)
    assert sd == pytest.approx(sel.std(ddof=1), rel=1e-9)
    m, sd = oracle.poly_stats(img, [-5, 50, 50, -5, -5, -5, 40, 40])
    assert m == pytest.approx(img.astype(np.float64).mean(), rel=1e-12)


# --------------------------------------------------------------------------
# spectral SNR (ResolutionEstimation.SSNR)
# --------------------------------------------------------------------------

def test_ssnr_ring_sums_brute_force(oracle):
    """The ring sums against a direct evaluation: explicit DFT sums, each
    frequency's radius from convertToFrequency, and its rings walked one by one
    with ceil/floor (NaN in an input counts as 0)."""
    rng = np.random.default_rng(14)
    n, h, w = 2, 5, 6
    a = [rng.standard_normal((n, h, w)).astype(np.float32) for _ in range(4)]
    a[0][1, 2, 3] = np.nan
    sums = oracle.ssnr_sums(*a, ring_width=2)
    want = np.zeros((5, w))
    y, x = np.mgrid[0:h, 0:w]
    for t in range(n):
        se, st, ne, nt = (np.nan_to_num(b[t]) for b in a)
        imgs = (st, (se - st).astype(np.float32), nt, (ne - nt).astype(np.float32))
        for ky in range(h):
            for kx in range(w):
                fxx = (kx if kx <= w // 2 else kx - w) / w
                fyy = (ky if ky <= h // 2 else ky - h) / h
                wi = np.sqrt(fxx * fxx + fyy * fyy) * w
                ph = np.exp(-2j * np.pi * (ky * y / h + kx * x / w))
                pw = [abs(np.sum(im.astype(np.float64) * ph)) ** 2 for im in imgs]
                for l in range(max(0, int(np.ceil(wi - 2))), int(np.floor(wi))):
                    for r in range(4):
                        want[r][l] += pw[r]
                    want[4][l] += 1
    assert np.allclose(sums, want, rtol=1e-10, atol=1e-9)


def test_ssnr_rows_known_answers(oracle):
    """Equal signal and noise sums give SSNR 1 (corrected -1000) and equal
    signal/noise rows; a signal ten times the noise power ratio gives
    10 log10(10 - 1) dB; the frequency row is i/W over pixel size."""
    w = 8
    K = np.arange(1, w + 1, dtype=np.float64)
    base = np.array([np.full(w, 4.0), np.full(w, 2.0), np.full(w, 4.0), np.full(w, 2.0), K])
    r = oracle.ssnr_rows(base, pixel_size=2.0)
    assert r.shape == (9, w // 2 + 1)
    assert np.all(r[2] == -1000) and np.allclose(r[3], r[6])
    assert np.allclose(r[1], np.arange(w // 2 + 1) / w / 2.0)
    assert np.allclose(r[4], 10 * np.log10(4.0 / K[:5] / (w + 1)))
    strong = base.copy()
    strong[0] *= 10
    assert np.allclose(oracle.ssnr_rows(strong)[2], 10 * np.log10(9.0))


def test_vssnr_map_closed_form(oracle):
    """The signal's projections twice the noise's, the differences measured -
    projected equal (small integers, so every float sum is exact): ISSNR/alpha
    = 4 at every frequency, each map value 10 log10(4); with an odd width
    swapQuadrant2D leaves the last column at 0."""
    rng = np.random.default_rng(19)
    n, h, w = 2, 6, 7
    th_n = rng.integers(-8, 9, (n, h, w)).astype(np.float32)
    d = rng.integers(-8, 9, (n, h, w)).astype(np.float32)
    noi_e = (th_n + d).astype(np.float32)
    sig_th = 2 * th_n
    sig_e = (sig_th + d).astype(np.float32)
    maps = oracle.vssnr_maps(sig_e, sig_th, noi_e, th_n)
    assert np.all(maps[:, :, w - 1] == 0)
    assert np.allclose(maps[:, :, :w - 1], 10 * np.log10(4.0), atol=1e-4)


# --------------------------------------------------------------------------
# BgART: the 3-D cross median and the collapse filter
# --------------------------------------------------------------------------

def test_median_cross_boundary_zeros(oracle):
    """A constant volume: sorted[3] of the 7 cross values stays c while at most
    3 of them are TomoJ's zeros (beyond the stack, or a failed index test
    idx+1 < W*H, idx-1 > 0, idx-W > 0, idx+W < W*H) and is 0 from 4 on."""
    W, H, c = 5, 4, 2.5
    one = oracle.median_cross(np.full((1, H, W), c, np.float32))[0]
    # one slice: both z values are 0.  A second in-plane failure makes 4 zeros:
    # idx 0 and 1 fail idx-1 > 0 and idx-W > 0; idx W*H-1 fails idx+1 < W*H and
    # idx+W < W*H; every other voxel fails at most one test
    exp = np.full((H, W), c, np.float32)
    exp[0, 0] = exp[0, 1] = exp[H - 1, W - 1] = 0.0
    assert np.array_equal(one, exp)
    three = oracle.median_cross(np.full((3, H, W), c, np.float32))
    # the middle slice has both z neighbours, so never more than 2 zeros
    assert np.array_equal(three[1], np.full((H, W), c, np.float32))
    # the first slice has one z zero: idx 0, 1 and W*H-1 reach 3 zeros and keep c
    assert np.array_equal(three[0], np.full((H, W), c, np.float32))


def test_median_cross_linear_ramp_and_spike(oracle):
    """Away from the index tests the cross is symmetric about a linear
    function, so an integer ramp is a fixed point; a lone spike goes."""
    D, H, W = 5, 6, 7
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    ramp = (3 * x + 5 * y + 11 * z + 100).astype(np.float32)
    out = oracle.median_cross(ramp)
    inner = (slice(1, D - 1), slice(1, H - 1), slice(1, W - 1))
    assert np.array_equal(out[inner], ramp[inner])
    vol = np.full((D, H, W), 1.0, np.float32)
    vol[2, 3, 3] = 50.0
    assert np.array_equal(oracle.median_cross(vol)[1:-1, 1:-1, 1:-1], vol[1:-1, 1:-1, 1:-1] * 0 + 1.0)


def test_median_cross_signed_zero_order(oracle):
    """Arrays.sort(float[]) puts -0.0 before 0.0: on one slice of -0.0 the first
    row's voxels see four +0.0 and three -0.0 (sorted[3] = +0.0); inner voxels
    see two +0.0 (sorted[3] = -0.0)."""
    W, H = 4, 4
    out = oracle.median_cross(np.full((1, H, W), -0.0, np.float32))[0]
    assert not np.signbit(out[0, 1]) and np.signbit(out[2, 2])


def test_collapse_closed_form(oracle):
    """40 voxels at 0.5, 200 at 1.0, 16 at 5.0: min 0.5, max 5, bin 4.5/256;
    1.0 falls in bin (int)(256/4.5*0.5) = 28, so mv = 0.5 + 28*4.5/256 =
    0.9921875 and sigma (bins below 28: only bin 0) = 0.4921875."""
    vals = np.array([0.5] * 40 + [1.0] * 200 + [5.0] * 16, np.float32)
    vol = vals.reshape(4, 8, 8)
    mv, sig = 0.5 + 28 * 4.5 / 256, 28 * 4.5 / 256
    dark, got = oracle.collapse(vol, 1.0, True)
    assert got == mv
    exp = np.where(vol < mv + sig, np.float32(mv), vol)
    assert np.array_equal(dark, exp)
    light, _ = oracle.collapse(vol, 1.0, False)
    assert np.array_equal(light, np.where(vol > mv - sig, np.float32(mv), vol))   # 0.5 == th1 stays
    assert np.array_equal(light.ravel()[:40], vals[:40])
    # a given background replaces the histogram's; sigma follows it
    fixed, got = oracle.collapse(vol, 2.0, True, bg_value=1.0)
    assert got == 1.0
    assert np.array_equal(fixed, np.where(vol < 1.0 + 2.0 * 0.5, np.float32(1.0), vol))


def test_collapse_skips_non_finite(oracle):
    """NaN and +-inf are outside the statistics; NaN and +inf survive the dark
    collapse, -inf is below the threshold and collapses."""
    vals = np.array([0.5] * 40 + [1.0] * 200 + [5.0] * 13 + [np.nan, np.inf, -np.inf], np.float32)
    vol = vals.reshape(4, 8, 8)
    out, mv = oracle.collapse(vol, 1.0, True)
    assert mv == 0.5 + 28 * 4.5 / 256
    flat = out.ravel()
    assert np.isnan(flat[-3]) and flat[-2] == np.inf and flat[-1] == np.float32(mv)
