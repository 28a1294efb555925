"""CPU-side checks of the C-ABI library: it loads, exports every entry point
include/sem_abi.h declares, and its host-side L0 routines (GLL, derivMat,
interpMat, semmesh) agree with the oracle. No GPU compute is called."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import sem_oracle as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "sem_abi.h")


def _lib():
    import semhip
    return semhip.load_library()


def test_library_exports_every_declared_symbol():
    lib = _lib()
    names = re.findall(r"^(?:int|const char\*)\s+(sem_\w+)\(", open(HDR).read(), re.M)
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert lib.sem_abi_version() == 1


def test_signature_table_covers_header():
    from semhip import _lib as L
    names = set(re.findall(r"^int\s+(sem_\w+)\(", open(HDR).read(), re.M))
    assert names == set(L.SIGNATURES)


@pytest.mark.parametrize("n", [2, 3, 5, 8, 9, 10, 16, 17])
def test_host_gll_and_derivmat_match_oracle(n):
    import semhip
    z, w = semhip.gausslobatto(n)
    zo, wo = o.gausslobatto(n)
    np.testing.assert_allclose(z, zo, rtol=0, atol=2e-16 * n)
    np.testing.assert_allclose(w, wo, rtol=4e-15 * n, atol=0)
    D = semhip.derivMat(z)
    np.testing.assert_allclose(D, o.derivMat(z), rtol=0, atol=1e-15 * n * n)
    # identical inputs and loop order -> identical bits
    np.testing.assert_array_equal(D, o.derivMat(z))


def test_host_interpmat_and_semmesh_match_oracle():
    import semhip
    zi, _ = o.gausslobatto(10)
    zo_, _ = o.gausslobatto(15)
    np.testing.assert_array_equal(semhip.interpMat(zo_, zi), o.interpMat(zo_, zi))
    np.testing.assert_array_equal(semhip.interpMat(2.0, [1.0, 0.0]), o.interpMat(2.0, [1.0, 0.0]))
    for E, n in [(4, 9), (5, 8), (256, 8)]:
        z, w = semhip.semmesh(E, n)
        zo2, wo2 = o.semmesh(E, n)
        np.testing.assert_allclose(z, zo2, rtol=0, atol=4e-16)
        np.testing.assert_allclose(w, wo2, rtol=1e-14)


def test_errors_are_reported_not_crashing():
    from semhip import _lib as L
    lib = _lib()
    st = lib.sem_gausslobatto(1, None, None)
    assert st == -1
    assert b"n >= 2" in lib.sem_last_error()
    with pytest.raises(L.SemError):
        L.call("sem_derivmat", 0, None, None)
