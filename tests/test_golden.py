"""The committed golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py from the oracle) against the oracle today
(regression pin) and, for the GLL tables, against the library's host
routines bit for bit."""
import os

import numpy as np
import pytest

from oracle import sem_oracle as o

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def fresh():
    from golden.make_golden import golden
    return golden()


def test_fixture_keys_complete(gold, fresh):
    assert set(gold) == set(fresh)


def test_integer_and_structural_fixtures_bit_exact(gold, fresh):
    for k in ("c1_l2g", "c1_mult", "c1_mask", "c1_pcg_iters"):
        np.testing.assert_array_equal(gold[k], fresh[k], err_msg=k)


def test_float_fixtures_reproduced(gold, fresh):
    for k, v in gold.items():
        if v.dtype.kind == "f":
            scale = max(np.abs(v).max(), 1e-300)
            assert np.abs(fresh[k] - v).max() <= 1e-13 * scale, k


def test_gll_tables_match_library_host_routines(gold):
    try:
        import semhip
    except OSError as e:
        pytest.skip(f"libsemhip.so not loadable: {e}")
    # nodes and weights: two independent Newton restatements (the library's in
    # long double) agree to 1 ulp (bit-level GLL parity is unpinned); derivMat
    # of the same nodes is bit-identical
    for p in (8, 9, 10, 16):
        z, w = semhip.gausslobatto(p)
        np.testing.assert_allclose(z, gold[f"gll{p}_z"], rtol=0, atol=2.3e-16)
        np.testing.assert_allclose(w, gold[f"gll{p}_w"], rtol=0, atol=2.3e-16)
        np.testing.assert_array_equal(semhip.derivMat(gold[f"gll{p}_z"]), gold[f"gll{p}_D"])


def test_bdf_fixtures_match_library_host_routines(gold):
    try:
        from semhip import cd
    except OSError as e:
        pytest.skip(f"libsemhip.so not loadable: {e}")
    for i in range(3):
        t = gold[f"bdf_t{i}"]
        a, b = cd.bdfExtK(list(t), k=len(t) - 1)
        np.testing.assert_array_equal(a, gold[f"bdf_a{i}"])
        np.testing.assert_array_equal(b, gold[f"bdf_b{i}"])
