"""CPU checks of the C ABI: the library loads, exports exactly what
include/gsap.h declares, and host-side validation rejects bad arguments
before any device work."""
import ctypes
import os
import re

import gsap_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "gsap.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gsap_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert declared() == sorted(gsap_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = gsap_amd.lib()
    missing = [s for s in declared() if not hasattr(L, s)]
    assert not missing, missing


def test_version_and_device_count_callable():
    L = gsap_amd.lib()
    assert b"gfx950" in L.gsap_version()
    assert L.gsap_device_count() >= 0


def test_create_validates_before_device():
    L = gsap_amd.lib()
    h = ctypes.c_void_p()
    d = gsap_amd.SapDesc()
    d.input_kind = 7
    assert L.gsap_sap_create(ctypes.byref(d), ctypes.byref(h)) == gsap_amd.ERR_INVALID_ARG
    assert b"input_kind" in L.gsap_last_error()
    d.input_kind = gsap_amd.INPUT_AABB
    d.axis_mode = 9
    assert L.gsap_sap_create(ctypes.byref(d), ctypes.byref(h)) == gsap_amd.ERR_INVALID_ARG
    d.axis_mode = gsap_amd.AXIS_PCA
    d.digit_bits = 5
    assert L.gsap_sap_create(ctypes.byref(d), ctypes.byref(h)) == gsap_amd.ERR_UNSUPPORTED
    d.digit_bits = 0
    d.sort_items = 3
    assert L.gsap_sap_create(ctypes.byref(d), ctypes.byref(h)) == gsap_amd.ERR_INVALID_ARG
    assert L.gsap_sap_create(None, ctypes.byref(h)) == gsap_amd.ERR_INVALID_ARG
    assert L.gsap_prb_create(0, 10, 10, ctypes.byref(h)) == gsap_amd.ERR_INVALID_ARG
    # null handles are rejected, not dereferenced
    assert L.gsap_sap_collide(None, None, None, 0, None, None, 0, None, None, None,
                              None) == gsap_amd.ERR_INVALID_ARG
    assert L.gsap_prb_step(None, None, None, ctypes.c_float(0.1)) == gsap_amd.ERR_INVALID_ARG
    assert L.gsap_radix_sort_u32(None, None, None, None, 0, 32) == gsap_amd.ERR_INVALID_ARG


def test_no_cpu_fallback_in_product():
    """The product package never imports the oracle."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "gsap_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                for bad in ("import oracle", "from oracle", "liboracle", "gsap_oracle", "or_sap("):
                    assert bad not in txt, (f, bad)


def test_dropin_math_headers_host_checks():
    """The Quaternion / Matrix3x3 drop-in types (include/gsap/quaternion.h,
    matrix3x3.h) checked on the host: quarter turns, R v = q v q^-1, R R^T = I,
    M M^-1 = I, composition order (tests/shim/header_math.cpp)."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "shim", "_build", "header_math")
    subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "shim"), "_build/header_math"],
                   check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
