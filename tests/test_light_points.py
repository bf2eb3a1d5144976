"""Sample points of area lights in the oracle (Light::get_point,
src/light.cc:106-109; SphLight::get_point :272-275, the position plus
libvmath's sphrand(radius), a uniform point on the sphere; BoxLight::get_point
:346-354, a uniform point of the box transformed by the light's matrix).

Pins for the distribution (the device draws the same points; the GPU render
tests compare its images with the oracle's):
- sphere light: every point is at the radius from the position; by
  Archimedes' theorem each coordinate of a uniform point on the sphere is
  uniform on [-r, r], and its azimuth is uniform: Kolmogorov-Smirnov tests;
- box light: mapped back through the inverse light matrix every point lies
  in the box, each local coordinate is uniform on [-dim/2, dim/2] (KS) and
  the coordinates are uncorrelated.
A dropped factor 2 in the Marsaglia map leaves the sphere; z = 1 - s instead
of 1 - 2s, or one draw reused for two coordinates, fails the KS or
correlation tests.
"""
import numpy as np
import pytest
from scipy import stats

import oracle
import sray_amd as S

XML = """<scene>
  <light type="sphere" color="1 1 1"><xform pos="1 2 3"/><sphere rad="0.5"/></light>
  <light type="box" color="1 1 1"><xform pos="-2 4 1" axisrot="30 0 1 0"/>
    <aabox min="-1 -0.25 -0.5" max="1 0.25 0.5"/></light>
  <light type="point" color="1 1 1"><xform pos="0 9 0"/></light>
</scene>"""

N = 20000


@pytest.fixture(scope="module")
def scene():
    return S.HostScene.load_xml(text=XML)


@pytest.mark.parametrize("prec", [32, 64])
def test_sphere_light_uniform_on_surface(scene, prec):
    o = oracle.OracleScene(scene.desc_ptr(), prec)
    p = o.light_points(0, N, seed=7, pixel=3) - np.array([1.0, 2.0, 3.0])
    r = np.linalg.norm(p, axis=1)
    assert np.allclose(r, 0.5, rtol=1e-5, atol=0)
    u = p / 0.5
    for k in range(3):
        assert stats.kstest(u[:, k], stats.uniform(loc=-1, scale=2).cdf).pvalue > 1e-3, k
    az = np.arctan2(u[:, 1], u[:, 0])
    assert stats.kstest(az, stats.uniform(loc=-np.pi, scale=2 * np.pi).cdf).pvalue > 1e-3
    assert abs(np.corrcoef(u[:, 0], u[:, 2])[0, 1]) < 0.03


@pytest.mark.parametrize("prec", [32, 64])
def test_box_light_uniform_in_box(scene, prec):
    o = oracle.OracleScene(scene.desc_ptr(), prec)
    p = o.light_points(1, N, seed=9, pixel=11)
    lt = scene.desc.lights[1]
    M = np.array(list(lt.xform), np.float64).reshape(3, 4)
    loc = np.linalg.solve(M[:, :3], (p - M[:, 3]).T).T
    half = np.array([1.0, 0.25, 0.5])
    assert np.all(np.abs(loc) <= half * (1 + 1e-5) + 1e-6)
    for k in range(3):
        assert stats.kstest(loc[:, k], stats.uniform(loc=-half[k], scale=2 * half[k]).cdf).pvalue > 1e-3, k
    c = np.corrcoef(loc.T)
    assert np.all(np.abs(c[np.triu_indices(3, 1)]) < 0.03)
    # the matrix is the light's placement: its centre maps to the position
    assert np.allclose(M[:, 3], [-2.0, 4.0, 1.0])


def test_point_light_and_sample_independence(scene):
    o = oracle.OracleScene(scene.desc_ptr(), 32)
    p = o.light_points(2, 10)
    assert np.array_equal(p, np.tile([0.0, 9.0, 0.0], (10, 1)))
    # different pixels and seeds draw different points, the same ones repeat
    a = o.light_points(0, 64, seed=1, pixel=0)
    assert np.array_equal(a, o.light_points(0, 64, seed=1, pixel=0))
    assert not np.any(np.all(a == o.light_points(0, 64, seed=1, pixel=1), axis=1))
    assert not np.any(np.all(a == o.light_points(0, 64, seed=2, pixel=0), axis=1))
    with pytest.raises(RuntimeError):
        o.light_points(3, 1)
