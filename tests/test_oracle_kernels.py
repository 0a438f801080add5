"""Cone kernels and cone parametrisation of the oracle against closed forms."""
import numpy as np
import pytest

from oracle.oracle import cone_kernel, cone_separated, disc_kernel, ray_kernel


def test_ray_kernel_integrates_to_one_and_endpoints():
    for Z in (0.01, 1.0, 7.5):
        z = np.linspace(0, Z, 200001)
        v = np.array([ray_kernel(t, Z) for t in z[:-1]] + [0.0])
        assert abs(np.trapezoid(v, z) - 1.0) < 1e-6
        assert ray_kernel(0.0, Z) == pytest.approx(1.5 / Z, rel=1e-15)
        assert ray_kernel(Z, Z) == 0.0 and ray_kernel(-1e-9, Z) == 0.0


def test_disc_kernel_integrates_to_one():
    for R in (0.003, 0.5, 2.0):
        r = np.linspace(0, R, 100001)
        v = np.array([disc_kernel(t * t, R) for t in r])
        assert abs(np.trapezoid(v * 2 * np.pi * r, r) - 1.0) < 1e-6
        assert disc_kernel(0.0, R) == pytest.approx(2 / (np.pi * R * R), rel=1e-15)
        assert disc_kernel(R * R, R) == 0.0


def _link(cam, p, n, scale, bw=2.0, rel=1.0, sampleness=True):
    cam, p, n = map(np.asarray, (cam, p, n))
    R = bw * scale
    if sampleness:
        d = (p - cam) / np.linalg.norm(p - cam)
        L = R * rel
        end = p + d * L * 0.5
        return cam, -n, end, R, L
    return cam, -n, p, R, 3.4028234663852886e38


def test_cone_parametrisation_invariants():
    cam, p, n = [0, 0, 3.0], [0.1, -0.05, 0.2], [0.0, 0.0, 1.0]
    a, en, e, R, L = _link(cam, p, n, 0.01, sampleness=False)
    k, tc, tcut, length, deg = cone_kernel(a, en, e, R, L, np.asarray(cam), False)
    assert not deg and tc == 0.0 and tcut == 0.0
    k, tc, tcut, length, deg = cone_kernel(a, en, e, R, L, e, False)
    assert tc == pytest.approx(1.0, abs=1e-15)
    assert length == pytest.approx(np.linalg.norm(np.subtract(p, cam)), rel=1e-14)   # complete cone: |cam - sample|
    # point on the end disc at radius < R is inside; beyond R is rejected
    u = np.cross(en, [1, 0, 0]); u /= np.linalg.norm(u)
    mid = np.asarray(cam) + 0.5 * (np.asarray(e) - np.asarray(cam))
    inside = mid + u * 0.49 * R
    outside = mid + u * 0.51 * R
    assert cone_kernel(a, en, e, R, L, inside, False)[0] > 0
    assert cone_kernel(a, en, e, R, L, outside, False)[0] == 0.0


def test_sampleness_kernel_peaks_at_sample_and_is_symmetric():
    cam, p, n = np.array([0.3, 0.2, 3.0]), np.array([0.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0])
    a, en, e, R, L = _link(cam, p, n, 0.01)
    k0, tc, tcut, length, deg = cone_kernel(a, en, e, R, L, p, True)
    assert tcut == pytest.approx(0.5, abs=1e-9) and length == pytest.approx(R)
    d = (p - cam) / np.linalg.norm(p - cam)
    # on the axis the disc kernel is 2/(pi (R tc)^2): the ray part k * tc^2 is symmetric in tCut about 0.5
    r = [cone_kernel(a, en, e, R, L, p + d * s * L * 0.2, True) for s in (-2, -1, 1, 2)]
    ray = [k * tc * tc for k, tc, _, _, _ in r]
    assert ray[0] == pytest.approx(ray[3], rel=1e-9) and ray[1] == pytest.approx(ray[2], rel=1e-9)
    assert [x[2] for x in r] == pytest.approx([0.1, 0.3, 0.7, 0.9], abs=1e-9)
    assert k0 * tc * tc > ray[1] > ray[0] > 0


def test_backfacing_link_is_degenerate():
    cam, p, n = [0, 0, 3.0], [0, 0, 0], [0, 0, -1.0]   # normal points away from the camera
    a, en, e, R, L = _link(cam, p, n, 0.01, sampleness=False)
    assert cone_kernel(a, en, e, R, L, [0, 0, 1.0], False)[4]


def test_culling_bound_is_conservative():
    rng = np.random.default_rng(0)
    for trial in range(300):
        cam = rng.normal(size=3) * 3
        p = rng.normal(size=3) * 0.3
        n = (cam - p) + rng.normal(size=3) * 0.5
        n /= np.linalg.norm(n)
        samp = bool(trial % 2)
        a, en, e, R, L = _link(cam, p, n, rng.uniform(0.005, 0.05), sampleness=samp)
        size = rng.uniform(0.005, 0.3)
        lo = p + rng.normal(size=3) * 0.1 - size / 2
        if not cone_separated(a, en, e, R, L, lo, size):
            continue
        # a culled box must contain no accepted point
        pts = lo + rng.uniform(0, size, size=(400, 3))
        for x in pts:
            assert cone_kernel(a, en, e, R, L, x, samp)[0] == 0.0
