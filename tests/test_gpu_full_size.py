"""GPU parity at the benchmark's own size: SU(3) on 32^4 (and 8^4).

The oracle's FP64 kernels run these sizes in seconds on the host cores, so
the full-size checks compare with it directly: the seeded start field, the
per-site plaquette, one complete Metropolis sweep (nHits = 10, beta = 6.0,
delta = 0.1, the headline workload), the HMC force, and FP32 / FP32 two-row
plaquettes within 1e-6 of FP64.  Size-independent properties ride along:
links stay unitary to rounding, and the cold field's plaquette is exactly 1.
"""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

import klft_amd as klft  # noqa: E402


@pytest.fixture(scope="module", params=[8, 32])
def field(request):
    L = request.param
    dims = (L, L, L, L)
    U = orc.random_algebra(4, 3, dims, 0.5, 2024)
    return L, dims, U


def test_seeded_start_and_plaquette(field):
    L, dims, U = field
    g = klft.GaugeField(4, 3, dims, "f64").set_random_algebra(0.5, 2024)
    np.testing.assert_allclose(g.download(), U, rtol=0, atol=1e-14)
    got, ps = klft.GaugePlaquette(g, per_site=True)
    want, ps_o = orc.plaquette(U, 4, 3, dims, per_site=True)
    np.testing.assert_allclose(ps, ps_o, rtol=0, atol=1e-12)
    assert abs(got - want) <= 1e-12 * np.abs(ps_o).sum() / (np.prod(dims) * 6 * 3) + 1e-15


def test_cold_field_is_exactly_one(field):
    L, dims, _ = field
    g = klft.GaugeField(4, 3, dims, "f64").set_identity()
    got, ps = klft.GaugePlaquette(g, per_site=True)
    _, ps_o = orc.plaquette(orc.identity(4, 3, dims), 4, 3, dims, per_site=True)
    assert got == 1.0 and np.array_equal(ps, ps_o)


def test_full_sweep_matches_oracle(field):
    L, dims, U = field
    g = klft.GaugeField(4, 3, dims, "f64").upload(U)
    p = klft.MetropolisParams(nHits=10, beta=6.0, delta=0.1, seed=4242)
    acc, n = klft.sweep_Metropolis(g, p, 0, with_count=True)
    Uo = U.copy()
    acc_o, n_o = orc.metropolis_sweep(Uo, 4, 3, dims, 10, 6.0, 0.1, 4242, 0)
    assert n == n_o and abs(acc - acc_o) < 1e-15
    Ug = g.download()
    np.testing.assert_allclose(Ug, Uo, rtol=0, atol=1e-11)
    # unitarity to rounding after the sweep's restoreSUN
    V = Ug.reshape(-1, 3, 3)
    dev = np.abs(V @ np.conj(np.transpose(V, (0, 2, 1))) - np.eye(3)).max()
    assert dev < 1e-13


def test_force_matches_oracle(field):
    L, dims, U = field
    g = klft.GaugeField(4, 3, dims, "f64").upload(U)
    h = klft.HMC(g, beta=6.0, tau=1.0, nsteps=20, seed=5)
    np.testing.assert_allclose(h.force(), orc.force(U, 4, 3, dims, 6.0), rtol=0, atol=1e-12)


@pytest.mark.parametrize("storage", ["f32", "f32_rows"])
def test_fp32_plaquette_within_1e6(field, storage):
    L, dims, U = field
    g = klft.GaugeField(4, 3, dims, storage).set_random_algebra(0.5, 2024).reunitarize()
    assert abs(klft.GaugePlaquette(g) - orc.plaquette(U, 4, 3, dims)) < 1e-6


@pytest.mark.parametrize("storage", ["f64", "f32_rows"])
def test_new_fields_keep_their_start(storage):
    """Fields created back to back (each over memory the previous one freed)
    and started at once keep every link of their start: the creation's clear
    of the links is ordered before the initialiser on the library stream.
    Each link of an exp(i eps H) start is unitary, so no row may be zero."""
    dims = (24, 24, 24, 24)
    for rep in range(3):
        g = klft.GaugeField(4, 3, dims, storage).set_random_algebra(0.5, 2024 + rep)
        U = g.download()
        g.close()
        rows = np.sqrt((np.abs(U) ** 2).sum(axis=-1))
        assert np.isfinite(U).all()
        assert np.abs(rows - 1.0).max() < 1e-6, (rep, float(rows.min()))


def test_lattice_beyond_65536_reduction_blocks():
    """48^3 x 40 (4.4 M sites, 69,120 link blocks): every reduction grid is
    larger than 2^16 blocks.  The field is an 8^4 oracle field repeated
    periodically, so the plaquette, per-site plaquette, force and kinetic
    energy are the 8^4 oracle values repeated (sums scaled by the 1080 copies).
    """
    small = (8, 8, 8, 8)
    reps = (6, 6, 6, 5)
    dims = tuple(a * b for a, b in zip(small, reps))
    U8 = orc.random_algebra(4, 3, small, 0.5, 77)
    want_plaq, ps8 = orc.plaquette(U8, 4, 3, small, per_site=True)
    F8 = orc.force(U8, 4, 3, small, 6.0)
    g = klft.GaugeField(4, 3, dims, "f64").upload(np.tile(U8, reps + (1, 1, 1)))
    got, ps = klft.GaugePlaquette(g, per_site=True)
    assert abs(got - want_plaq) < 1e-12
    np.testing.assert_allclose(ps.reshape(dims), np.tile(ps8.reshape(small), reps), rtol=0, atol=1e-12)
    del ps
    h = klft.HMC(g, beta=6.0, tau=1.0, nsteps=20, seed=5)
    np.testing.assert_allclose(h.force(), np.tile(F8, reps + (1, 1)), rtol=0, atol=1e-12)
    P8 = np.random.default_rng(3).standard_normal(small + (4, 8))
    h.set_momenta(np.tile(P8, reps + (1, 1)))
    h8 = klft.HMC(klft.GaugeField(4, 3, small, "f64").upload(U8), beta=6.0, tau=1.0, nsteps=20, seed=5)
    h8.set_momenta(P8)
    assert abs(h.kinetic() - 1080 * h8.kinetic()) <= 1e-12 * abs(h.kinetic())


@pytest.mark.parametrize("variant", [4, 5])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_walking_plaquette_with_a_partial_segment(monkeypatch, storage, variant):
    """The dimension-0 walking plaquette (variants 4, 5) on 26 x 24^3: 13,824
    walkers per segment; FP64 links take 6 sites per walker in 5 segments,
    FP32 links 3 in 9, the last segment holding 2 in both.  Per-site values
    against the oracle (FP64 1e-12, FP32 links 5e-5 per site) and the sum."""
    dims = (26, 24, 24, 24)
    U = orc.random_algebra(4, 3, dims, 0.5, 808)
    monkeypatch.setenv("KLFT_VARIANT", f"plaquette={variant}")
    g = klft.GaugeField(4, 3, dims, storage).upload(U)
    got, ps = klft.GaugePlaquette(g, per_site=True)
    want, ps_o = orc.plaquette(U, 4, 3, dims, per_site=True)
    tol = 1e-12 if storage == "f64" else 5e-5
    np.testing.assert_allclose(ps, ps_o, rtol=0, atol=tol)
    assert abs(got - want) < (1e-12 if storage == "f64" else 1e-6)
