"""klft_plaquette_noalloc (the plaquette entry point that allocates no device
memory) against klft_plaquette, through the C ABI.

Two calls in a row on one GPU, the second on a smaller lattice, so the static
scratch keeps the first call's contents when the second starts: the sum and
the per-site values must equal klft_plaquette's (same kernel, same launch
shape, so the same reduction order: bit-identical), and the small case also
matches the FP64 oracle to 1e-12.  Per-site output past the 2^20-site limit
is rejected with KLFT_E_ARG before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

import klft_amd as klft  # noqa: E402


def _noalloc(g, normalize, per_site):
    out = C.c_double()
    ps = np.full(g.volume, np.nan) if per_site else None
    rc = klft.lib.klft_plaquette_noalloc(g.h, int(normalize), C.byref(out),
                                         klft._dptr(ps) if per_site else None)
    return rc, out.value, ps


def _with_alloc(g, normalize, per_site):
    out = C.c_double()
    ps = np.full(g.volume, np.nan) if per_site else None
    rc = klft.lib.klft_plaquette(g.h, int(normalize), C.byref(out), klft._dptr(ps) if per_site else None)
    assert rc == 0, klft.lib.klft_last_error()
    return out.value, ps


@pytest.mark.parametrize("storage", ["f64", "f32", "f32_rows"])
def test_noalloc_matches_plaquette_two_calls(storage):
    big = klft.GaugeField(4, 3, (32, 32, 32, 32), storage).set_random_algebra(0.7, 91)
    dims_s = (4, 6, 4, 8)
    U = orc.random_algebra(4, 3, dims_s, 0.6, 92)
    small = klft.GaugeField(4, 3, dims_s, storage).upload(U)
    for g, label in [(big, "32^4 (at the per-site limit)"), (small, "4x6x4x8")]:
        for normalize in (1, 0):
            for per_site in (True, False):
                rc, v, ps = _noalloc(g, normalize, per_site)
                assert rc == 0, klft.lib.klft_last_error()
                v_ref, ps_ref = _with_alloc(g, normalize, per_site)
                print(f"{storage} {label} normalize={normalize} per_site={per_site}: "
                      f"{v!r} vs {v_ref!r}")
                assert v == v_ref
                if per_site:
                    np.testing.assert_array_equal(ps, ps_ref)
    bound = 1e-12 if storage == "f64" else 1e-6
    rc, v, _ = _noalloc(small, 1, False)
    assert rc == 0
    assert abs(v - orc.plaquette(U, 4, 3, dims_s)) < bound
    big.close()
    small.close()


def test_noalloc_rejects_per_site_past_limit():
    g = klft.GaugeField(4, 1, (34, 34, 34, 34), "f64").set_random_algebra(0.4, 5)
    rc, _, _ = _noalloc(g, 1, True)
    assert rc == -1  # KLFT_E_ARG
    assert "no-allocation limit" in klft.lib.klft_last_error().decode()
    rc, v, _ = _noalloc(g, 1, False)  # the sum alone has no site limit
    assert rc == 0, klft.lib.klft_last_error()
    assert v == _with_alloc(g, 1, False)[0]
    g.close()
