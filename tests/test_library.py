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


def _handle_functions():
    """Entry points whose first parameter is an engine or ensemble handle."""
    hdr = open(os.path.join(ROOT, "include", "cmhip.h")).read()
    decl = re.findall(r"^(int32_t|const char\*)\s+(cm_\w+)\s*\(\s*void\*\s*(\w+)", hdr, re.M)
    return [(ret, name) for ret, name, _ in decl]


def test_null_handles_are_rejected_before_any_device_call():
    """Every call on a NULL engine or ensemble handle returns
    CM_ERR_NOT_INITIALIZED (the reference's ErrorCode::NotInitialized,
    engine/include/celestial/core/error.hpp:9-18) and touches no device: the
    check comes first, so this runs without a GPU."""
    import ctypes as C
    L = C.CDLL(os.path.join(PKG, "libcmhip.so"))
    funcs = _handle_functions()
    assert len(funcs) > 30
    for ret, name in funcs:
        fn = getattr(L, name)
        if ret == "const char*":
            fn.restype = C.c_char_p
            fn.argtypes = [C.c_void_p]
            assert fn(None) is not None, name      # a message, never a NULL pointer
            continue
        fn.restype = C.c_int32
        # the handle check precedes every other argument's use; the rest are
        # zero / NULL (pointer-sized integer registers)
        assert fn(None, *([None] * 8)) == cmhip.CM_ERR_NOT_INITIALIZED, name


def test_create_rejects_bad_arguments_without_a_device():
    import ctypes as C
    L = C.CDLL(os.path.join(PKG, "libcmhip.so"))
    L.cm_create.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]
    L.cm_create.restype = C.c_int32
    h = C.c_void_p()
    assert L.cm_create(0, 0, C.byref(h)) == cmhip.CM_ERR_INVALID_ARGUMENT
    assert L.cm_create(-5, 0, C.byref(h)) == cmhip.CM_ERR_INVALID_ARGUMENT
    assert L.cm_create(16, 0, None) == cmhip.CM_ERR_INVALID_ARGUMENT
    assert h.value is None
