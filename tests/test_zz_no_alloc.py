"""cm_energy_noalloc, the entry point that allocates nothing, against
cm_energy through the C ABI: bit-identical reports for every potential mode on
two calls in a row, the second with a smaller body count, and the size
limit of its static scratch rejected with CM_ERR_INVALID_ARGUMENT."""
import numpy as np
import pytest

import cmhip

pytestmark = pytest.mark.gpu

EPS = 1e-4


def _same(a, b):
    for k in a:
        assert a[k] == b[k], (k, a[k], b[k])


def test_noalloc_energy_equals_energy_on_two_calls():
    be = cmhip.HipPhysicsBackend(60000, precision=cmhip.PREC_FP64, theta=0.5)
    for n, seed in ((60000, 3), (17000, 4)):
        soa = cmhip.generate_plummer(n, 1.0, 1.0, seed=seed, mass_fraction_cut=0.999)
        rng = np.random.default_rng(seed)
        soa.is_active[:n][rng.random(n) < 0.05] = 0
        be.upload(soa)
        print(f"noalloc energy: n={n}", flush=True)
        for mode in (0, 1, 2):
            ref = be.energy(EPS, mode)
            got = be.energy(EPS, mode, noalloc=True)
            _same(got, ref)
            assert got["active_count"] == int(soa.is_active[:n].sum())
            if mode:
                assert got["potential"] < 0.0


def test_noalloc_energy_rejects_more_bodies_than_its_scratch():
    n = (1 << 22) + 1
    print(f"noalloc limit: n={n}", flush=True)
    be = cmhip.HipPhysicsBackend(n, kind=cmhip.BACKEND_DIRECT)
    z = np.zeros(n)
    soa = cmhip.BodySoA.from_arrays(pos_x=np.arange(n, dtype=np.float64), pos_y=z, pos_z=z,
                                    mass=np.ones(n), is_active=np.ones(n, np.uint8))
    be.upload(soa)
    with pytest.raises(cmhip.CmError) as err:
        be.energy(EPS, 1, noalloc=True)
    assert err.value.code == cmhip.CM_ERR_INVALID_ARGUMENT
    # without a potential no rows are needed: any size is served
    e = be.energy(EPS, 0, noalloc=True)
    assert e["active_count"] == n and e["total_mass"] == float(n)
