"""CPU pins of the render oracle (oracle/render_oracle.hpp): the exp it uses
against math.exp, and whole images against closed forms of the ray march
(sums computed here in float64 from the geometry, not by re-running the
march)."""
import math

import numpy as np
import pytest

O = pytest.importorskip("oracle")


def ulps(a: float, b: float) -> int:
    ia = np.array([a], np.float32).view(np.int32)[0]
    ib = np.array([b], np.float32).view(np.int32)[0]
    return abs(int(ia) - int(ib))


def test_exp_within_two_ulps():
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.linspace(-87.3, 88.7, 40001), rng.uniform(-20, 20, 20000),
                         rng.uniform(-1e-3, 1e-3, 2000)]).astype(np.float32)
    worst = 0
    for x in xs:
        got = O.bexp(float(x))
        want = np.float32(math.exp(float(x)))
        worst = max(worst, ulps(got, float(want)))
    assert worst <= 2, worst
    assert O.bexp(0.0) == 1.0
    assert math.isnan(O.bexp(float("nan")))
    assert O.bexp(89.0) == float("inf") and O.bexp(-88.0) == 0.0


def base_settings(w=16, h=16):
    r = O._RenderSettings()
    r.width, r.height = w, h
    r.cam_focal, r.cam_aperture = 50.0, 20.0
    r.cam_u[:] = [1, 0, 0]
    r.cam_v[:] = [0, 1, 0]
    r.cam_w[:] = [0, 0, 1]
    r.cam_trans[:] = [0.5, 0.5, 3.0]
    r.p_step, r.s_step, r.cutoff = 0.01, 0.02, 1e-4
    r.light_count = 0
    r.density_scale, r.shadow_density_scale = 1.0, 1.0
    r.volume_color[:] = [0.8, 0.9, 1.0]
    r.scattering[:] = [0.5, 1.0, 1.5]
    r.absorption, r.gain, r.emission_scale = 0.25, 0.0, 1.0
    r.ramp_size = 1
    r.ramp_colors[0][:] = [1.0, 0.5, 0.25]
    r.temp_range[:] = [0.0, 1.0]
    r.ms_scale, r.ms_blur_iter = 0.0, 0
    r.ms_mask[:] = [0.0, 1.0]
    return r


def camera_ray(r, i, j):
    """Pinhole ray of pixel (i, j) in float64 (camera model of
    src/Camera.h:42-58)."""
    o = np.array(r.cam_trans[:], float)
    u, v, w = (np.array(x[:], float) for x in (r.cam_u, r.cam_v, r.cam_w))
    hw = 0.5 * r.cam_aperture
    hh = hw / (r.width / r.height)
    tl = o - r.cam_focal * w - hw * u + hh * v
    d = tl + (i + 0.5) / r.width * 2 * hw * u + (j + 0.5) / r.height * (-2 * hh) * v - o
    return o, d / np.linalg.norm(d)


def box_span(o, d):
    t0, t1 = -np.inf, np.inf
    for a in range(3):
        ta, tb = (0.0 - o[a]) / d[a], (1.0 - o[a]) / d[a]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    return max(t0, 0.0), t1


DX = 1.0 / 16.0


def render(T, r):
    return O.render(T, DX, (0, 0, 0), (1, 1, 1), r)


def test_empty_volume_is_black():
    img = render(np.zeros((16, 16, 16), np.float32), base_settings())
    assert not img.any()


@pytest.mark.parametrize("c", [0.3, 0.8])
def test_constant_emitting_medium_telescopes(c):
    """No lights: L = sum_n T_n E (1 - dT) with T_{n+1} = T_n dT, which is
    E (1 - dT^N) for N samples along the chord through the box."""
    r = base_settings()
    T = np.full((16, 16, 16), c, np.float32)
    img = render(T, r)
    sig = np.array(r.scattering[:]) + r.absorption
    dT = np.exp(-sig * c * r.p_step)
    E = np.array(r.ramp_colors[0][:]) * r.emission_scale
    checked = 0
    for j in range(r.height):
        for i in range(r.width):
            o, d = camera_ray(r, i, j)
            t0, t1 = box_span(o, d)
            if not t1 > t0:
                assert not img[j, i].any()
                continue
            start = t0 + DX * 0.5
            n = math.floor((t1 - start) / r.p_step) + 1
            frac = (t1 - start) / r.p_step % 1.0
            if min(frac, 1 - frac) < 1e-3:
                continue   # chord length within rounding of a step boundary
            want = E * (1 - dT ** n)
            np.testing.assert_allclose(img[j, i], want, rtol=2e-4, atol=1e-6)
            checked += 1
    assert checked > 100


def test_multiscatter_of_a_constant_field():
    """A constant field above the ramp's low end downsamples and blurs to
    itself (block mean, weights summing to 1), so the emission gains
    ms_color * c * ms_scale * smoothstep(mask, c) at every sample."""
    r = base_settings()
    r.ms_scale, r.ms_blur_iter = 0.5, 3
    r.ms_color[:] = [1.0, 0.2, 0.0]
    r.ms_mask[:] = [0.0, 2.0]
    c = 0.6
    img = render(np.full((16, 16, 16), c, np.float32), r)
    y = min(max((c - 0.0) / 2.0, 0.0), 1.0)
    sm = y * y * (3 - 2 * y)
    E = np.array(r.ramp_colors[0][:]) * r.emission_scale + np.array(r.ms_color[:]) * c * r.ms_scale * sm
    sig = np.array(r.scattering[:]) + r.absorption
    dT = np.exp(-sig * c * r.p_step)
    i, j = r.width // 2, r.height // 2
    o, d = camera_ray(r, i, j)
    t0, t1 = box_span(o, d)
    n = math.floor((t1 - (t0 + DX * 0.5)) / r.p_step) + 1
    np.testing.assert_allclose(img[j, i], E * (1 - dT ** n), rtol=2e-4)


def test_single_light_along_the_view_axis():
    """One light shining along -z (shadow rays toward +z, back to the
    camera's face of the box), no emission, constant medium: each primary
    sample at depth z sees exp(-sigma c s_step M(z)) with M(z) shadow samples
    on its chord to z = 1, so L = sum_n albedo illum T_n (1 - dT) S_n."""
    r = base_settings(8, 8)
    r.emission_scale = 0.0
    r.light_count = 1
    r.light_dir[0][:] = [0.0, 0.0, -1.0]
    r.light_color[0][:] = [1.0, 1.0, 1.0]
    r.light_intensity[0], r.light_exposure[0] = 2.0, 1.0
    r.cam_aperture = 0.5
    c = 0.5
    img = render(np.full((16, 16, 16), c, np.float32), r)
    sig = np.array(r.scattering[:]) + r.absorption
    albedo = np.array(r.volume_color[:]) * np.array(r.scattering[:]) / sig
    illum = np.array(r.light_color[0][:]) * 2.0 * 2.0
    dT = np.exp(-sig * c * r.p_step)
    i, j = 4, 4
    o, d = camera_ray(r, i, j)
    t0, t1 = box_span(o, d)
    L, Tn, t = np.zeros(3), np.ones(3), t0 + DX * 0.5
    while t <= t1:
        z = o[2] + t * d[2]
        m = math.floor(((1.0 - z) - DX * 0.5) / r.s_step) + 1 if (1.0 - z) >= DX * 0.5 else 0
        S = np.exp(-sig * c * r.s_step * m)
        L += albedo * S * Tn * illum * (1 - dT)
        Tn = Tn * dT
        t += r.p_step
    np.testing.assert_allclose(img[j, i], L, rtol=5e-3)
