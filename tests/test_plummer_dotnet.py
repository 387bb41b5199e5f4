"""PlummerSphere.Generate reproduced body for body (cm_generate_plummer_dotnet)
with .NET's System.Random(seed) stream (ScenarioGenerators/PlummerSphere.cs:
41-116).  .NET is absent here: the stream is pinned by the known first draws
of System.Random(0), by properties of its seeding, and the generator by the
oracle's restatement and by Plummer statistics.  Host code only."""
import numpy as np
import pytest

import cmhip
import oracle as O

MBIG = 2147483647


def test_dotnet_random_known_answers():
    """new Random(0).Next() gives 1559595546, 1755192844, 1649316166, and
    new Random(0).NextDouble() 0.7262432699679598 (draw / (2^31 - 1))."""
    d = O.dotnet_random(0, 3)
    assert list(d) == [1559595546, 1755192844, 1649316166]
    assert float(d[0]) * (1.0 / MBIG) == 0.7262432699679598


def test_dotnet_random_seeding_properties():
    """The seed enters as |seed| (and int.MinValue as int.MaxValue); draws lie
    in [0, 2^31 - 2]; neighbouring seeds give unrelated streams."""
    assert (O.dotnet_random(-12345, 200) == O.dotnet_random(12345, 200)).all()
    assert (O.dotnet_random(-2 ** 31, 50) == O.dotnet_random(2 ** 31 - 1, 50)).all()
    d = O.dotnet_random(7, 100000)
    assert d.min() >= 0 and d.max() <= MBIG - 1
    assert abs(d.mean() / MBIG - 0.5) < 0.005
    a, b = O.dotnet_random(42, 1000), O.dotnet_random(43, 1000)
    assert (a != b).mean() > 0.99


@pytest.mark.parametrize("n,mass,scale,seed", [(1, 100.0, 2.0, 42), (7, 100.0, 2.0, 0),
                                               (5000, 100.0, 2.0, 42), (3000, 1.0, 1.0, 2024)])
def test_generator_matches_restatement_bit_for_bit(n, mass, scale, seed):
    soa = cmhip.generate_plummer_dotnet(n, mass, scale, seed)
    ref = O.plummer_dotnet(n, mass, scale, seed)
    for k, f in (("px", "pos_x"), ("py", "pos_y"), ("pz", "pos_z"), ("vx", "vel_x"),
                 ("vy", "vel_y"), ("vz", "vel_z"), ("mass", "mass")):
        np.testing.assert_array_equal(getattr(soa, f)[:n], ref[k])
    assert (soa.radius[:n] == 0.02).all() and (soa.is_active[:n] == 1).all()


def test_generator_defaults_follow_the_plummer_profile():
    """Reference defaults (M = 100, a = 2): half-mass radius
    a / sqrt(2^(2/3) - 1) = 1.305 a; speeds below the local escape speed; equal
    masses M / n; no centre-of-mass shift (PlummerSphere.Generate applies
    none), so the COM is only statistically near 0."""
    n = 20000
    soa = cmhip.generate_plummer_dotnet(n)
    x, y, z = soa.pos_x[:n], soa.pos_y[:n], soa.pos_z[:n]
    r = np.sqrt(x * x + y * y + z * z)
    assert abs(np.median(r) / (1.3048 * 2.0) - 1.0) < 0.03
    v = np.sqrt(soa.vel_x[:n] ** 2 + soa.vel_y[:n] ** 2 + soa.vel_z[:n] ** 2)
    assert (v <= np.sqrt(2.0 * 100.0 / np.sqrt(r * r + 4.0)) * (1 + 1e-12)).all()
    assert (soa.mass[:n] == 100.0 / n).all()
    assert abs(np.mean(x)) > 0.0


def test_generator_arguments():
    with pytest.raises(cmhip.CmError):
        cmhip.generate_plummer_dotnet(10, seed=-1)
    with pytest.raises(cmhip.CmError):
        cmhip.generate_plummer_dotnet(10, total_mass=0.0)
    assert cmhip.generate_plummer_dotnet(0).count == 0
