"""CPU checks of the C ABI library: it loads and exports what include/cmhip.h declares."""
import os
import re
import subprocess

import cmhip

from conftest import PKG, ROOT


def declared_functions():
    hdr = open(os.path.join(ROOT, "include", "cmhip.h")).read()
    return sorted(set(re.findall(r"^(?:int32_t|const char\*)\s+(cm_\w+)\s*\(", hdr, re.M)))


def test_header_lists_every_export():
    assert sorted(cmhip.EXPORTS) == declared_functions()


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(PKG, "libcmhip.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(cm_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_library_loads_and_binds():
    L = cmhip.load_library()
    for name in cmhip.EXPORTS:
        assert getattr(L, name) is not None


def test_plummer_generator_statistics():
    """PlummerSphere semantics (PlummerSphere.cs:41-116): half-mass radius of a
    Plummer sphere is a / sqrt(2^(2/3) - 1) = 1.305 a; virial 2K/|W| ~ 1."""
    import numpy as np
    import oracle as O
    from conftest import soa_to_state
    soa = cmhip.generate_plummer(20000, 1.0, 1.0, seed=7)
    r = np.sqrt(soa.pos_x[:20000] ** 2 + soa.pos_y[:20000] ** 2 + soa.pos_z[:20000] ** 2)
    assert abs(np.median(r) / 1.3048 - 1.0) < 0.03
    s = soa_to_state(soa)
    assert abs(np.sum(s["mass"] * s["px"])) < 1e-12 and abs(np.sum(s["mass"] * s["vx"])) < 1e-12
    e = O.energy_exact(s, 1e-4, True, nthreads=8)
    assert abs(2 * e["ke"] / abs(e["pe"]) - 1.0) < 0.05


def test_plummer_generator_is_deterministic_and_truncates():
    a = cmhip.generate_plummer(5000, seed=3, mass_fraction_cut=0.999)
    b = cmhip.generate_plummer(5000, seed=3, mass_fraction_cut=0.999)
    assert (a.pos_x == b.pos_x).all() and (a.vel_z == b.vel_z).all()
    import numpy as np
    r = np.sqrt(a.pos_x ** 2 + a.pos_y ** 2 + a.pos_z ** 2)
    # 0.999 mass fraction: r <= a / sqrt(0.999^(-2/3) - 1) = 38.7 a (before COM shift)
    assert r.max() < 39.0
