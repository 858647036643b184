"""GPU parity: the HIP library (through the C ABI) against the FP64 oracle.

Tolerances: FP64 per-site values 1e-12 (entries are O(1)); global sums are
compared relative to the sum of absolute per-site terms (the reduction order
differs between the GPU tree and the oracle's pairwise sum).  Index and
schedule data must be bit-exact.
"""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

import klft_amd as klft  # noqa: E402

SHAPES = [(2, (6, 4)), (3, (4, 6, 4)), (4, (4, 4, 6, 4))]


def seeded(nd, nc, dims, eps, seed):
    return orc.random_algebra(nd, nc, dims, eps, seed)


@pytest.mark.parametrize("nc", [1, 2, 3])
@pytest.mark.parametrize("nd,dims", SHAPES)
def test_upload_download_roundtrip(nd, nc, dims):
    U = seeded(nd, nc, dims, 0.7, 11)
    g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
    assert np.array_equal(g.download(), U)


@pytest.mark.parametrize("nc", [1, 2, 3])
@pytest.mark.parametrize("nd,dims", SHAPES)
def test_random_algebra_init_matches_oracle(nd, nc, dims):
    U = seeded(nd, nc, dims, 0.5, 7)
    g = klft.GaugeField(nd, nc, dims, "f64").set_random_algebra(0.5, 7)
    np.testing.assert_allclose(g.download(), U, rtol=0, atol=1e-14)


@pytest.mark.parametrize("storage", ["f64", "f32", "f32_rows"])
@pytest.mark.parametrize("nc", [1, 2, 3])
@pytest.mark.parametrize("nd,dims", SHAPES)
def test_random_sun_init_matches_oracle(nd, nc, dims, storage):
    """deviceGaugeField(L.., rng, delta): every link randSUN(delta)
    (GaugeField.hpp:96-112); FP64 within 1e-14, FP32 storage within the
    rounding of the stored precision."""
    U = orc.random_sun_field(nd, nc, dims, 0.5, 2468)
    g = klft.GaugeField(nd, nc, dims, storage).set_random_sun(0.5, 2468)
    tol = 1e-14 if storage == "f64" else 1e-6
    np.testing.assert_allclose(g.download(), U, rtol=0, atol=tol)


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("nc", [1, 2, 3])
@pytest.mark.parametrize("nd,dims", SHAPES)
def test_random_complex_and_constant_init(nd, nc, dims, storage):
    """deviceGaugeField(L.., rng) (GaugeField.hpp:118-141): every entry
    complex(drand(-1,1), drand(-1,1)), bit-exact in FP64; and
    deviceGaugeField(L.., complex_t) (:65-84): every entry the constant."""
    U = orc.random_complex_field(nd, nc, dims, 97)
    g = klft.GaugeField(nd, nc, dims, storage).set_random_complex(97)
    if storage == "f64":
        assert np.array_equal(g.download(), U)
    else:
        np.testing.assert_allclose(g.download(), U, rtol=0, atol=1e-7)
    g.set_constant(0.25 - 0.5j)
    assert np.array_equal(g.download(), np.full(U.shape, 0.25 - 0.5j))


@pytest.mark.parametrize("nc", [2, 3])
def test_general_matrix_inits_reject_two_row_storage(nc):
    g = klft.GaugeField(4, nc, (4, 4, 4, 4), "f32_rows")
    with pytest.raises(klft.KlftError):
        g.set_random_complex(1)
    with pytest.raises(klft.KlftError):
        g.set_constant(1.0)


@pytest.mark.parametrize("nd,dims", SHAPES + [(4, (8, 6, 4, 10))])
def test_schedule_bit_exact(nd, dims):
    U = orc.identity(nd, 1, dims)
    _, _, visit = orc.metropolis_sweep(U, nd, 1, dims, 0, 1.0, 0.1, 1, 0, visit=True)
    g = klft.GaugeField(nd, 1, dims, "f64").set_identity()
    v = klft.metropolis_schedule(g)
    assert np.array_equal(v, visit)
    assert set(np.unique(v)) == set(range(2 ** nd))


@pytest.mark.parametrize("nc", [1, 2, 3])
@pytest.mark.parametrize("nd,dims", SHAPES)
def test_plaquette_cold_is_one(nd, nc, dims):
    g = klft.GaugeField(nd, nc, dims, "f64").set_identity()
    assert klft.GaugePlaquette(g) == 1.0


@pytest.mark.parametrize("eps", [0.1, 0.5, 1.0])
@pytest.mark.parametrize("nc", [1, 2, 3])
@pytest.mark.parametrize("nd,dims", SHAPES)
def test_plaquette_parity(nd, nc, dims, eps):
    U = seeded(nd, nc, dims, eps, 100)
    ref, ref_ps = orc.plaquette(U, nd, nc, dims, per_site=True)
    g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
    got, ps = klft.GaugePlaquette(g, per_site=True)
    np.testing.assert_allclose(ps, ref_ps, rtol=0, atol=1e-12)
    norm = np.prod(dims) * (nd * (nd - 1) // 2) * nc
    assert abs(got - ref) <= 1e-12 * np.abs(ref_ps).sum() / norm + 1e-15
    raw = klft.GaugePlaquette(g, normalize=False)
    assert abs(raw - orc.plaquette(U, nd, nc, dims, normalize=False)) <= 1e-12 * np.abs(ref_ps).sum()


@pytest.mark.parametrize("nc", [1, 2, 3])
@pytest.mark.parametrize("nd,dims", SHAPES)
def test_staple_field_parity(nd, nc, dims):
    U = seeded(nd, nc, dims, 0.8, 5)
    S = orc.staple_field(U, nd, nc, dims)
    g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
    out = klft.stapleField(g).download()
    np.testing.assert_allclose(out, S, rtol=0, atol=1e-12)


@pytest.mark.parametrize("nc", [1, 2, 3])
@pytest.mark.parametrize("nd,dims", [(3, (4, 6, 8)), (4, (4, 4, 6, 8))])
def test_wilson_mu_nu_parity(nd, nc, dims):
    U = seeded(nd, nc, dims, 0.6, 21)
    g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
    pairs = [(1, 1), (2, 2), (3, 4), (4, 3), (2, 5), (0, 2)]
    for mu, nu in [(0, 1), (1, 2), (nd - 1, 0)]:
        ref, ref_ps = orc.wilson_mu_nu(U, nd, nc, dims, mu, nu, pairs, per_site=True)
        got, ps = klft.WilsonLoop_mu_nu(g, mu, nu, pairs, per_site=True)
        np.testing.assert_allclose(ps, ref_ps, rtol=0, atol=1e-12)
        scale = np.abs(ref_ps).sum() / np.prod(dims) / nc
        for (m, n, Lm, Ln, w), r in zip(got, ref):
            assert (m, n) == (mu, nu)
            assert abs(w - r) <= 1e-12 * max(scale, 1.0)


@pytest.mark.parametrize("nc", [1, 2, 3])
def test_wilson_temporal_parity(nc):
    """Every L, T in 1..8 on 8 x 6 x 10 x 12 (loops up to and beyond the
    extents: lines wrap periodically)."""
    nd, dims = 4, (8, 6, 10, 12)
    U = seeded(nd, nc, dims, 0.4, 3)
    pairs = [(L, T) for L in range(1, 9) for T in range(1, 9)]
    ref = orc.wilson_temporal(U, nd, nc, dims, pairs)
    g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
    got = klft.WilsonLoop_temporal(g, pairs)
    for (L, T, w), r, p in zip(got, ref, pairs):
        assert (L, T) == p
        assert abs(w - r) <= 1e-12


@pytest.mark.parametrize("storage,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_wilson_temporal_long_time_extent(storage, tol):
    """A time extent of 256: one FP64 column's mu-lines and nu-links do not
    both fit the kernel's LDS, so the nu-links are loaded at every step (FP32
    columns still stage them); both against the oracle."""
    nd, nc, dims = 4, 3, (4, 2, 2, 256)
    U = seeded(nd, nc, dims, 0.4, 8)
    pairs = [(1, 1), (2, 3), (3, 8), (4, 5)]
    ref = orc.wilson_temporal(U, nd, nc, dims, pairs)
    g = klft.GaugeField(nd, nc, dims, storage).upload(U)
    for (L, T, w), r in zip(klft.WilsonLoop_temporal(g, pairs), ref):
        assert abs(w - r) <= tol, (L, T, w, r)


def test_wilson_cold_is_one():
    g = klft.GaugeField(4, 3, (4, 4, 4, 8), "f64").set_identity()
    for L, T, w in klft.WilsonLoop_temporal(g, [(1, 1), (2, 3), (4, 8)]):
        assert abs(w - 1.0) < 1e-14


@pytest.mark.parametrize("nc,beta", [(1, 2.0), (2, 2.3), (3, 5.5)])
@pytest.mark.parametrize("nd,dims", SHAPES)
def test_metropolis_fp64_tracks_oracle(nd, nc, dims, beta):
    """Same keyed stream: the GPU sweep follows the oracle's decisions."""
    U = seeded(nd, nc, dims, 0.3, 9)
    g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
    p = klft.MetropolisParams(nHits=4, beta=beta, delta=0.3, seed=4242)
    Uo = U.copy()
    for sweep in range(3):
        acc_o, n_o = orc.metropolis_sweep(Uo, nd, nc, dims, 4, beta, 0.3, 4242, sweep)
        acc_g, n_g = klft.sweep_Metropolis(g, p, sweep, with_count=True)
        assert n_g == n_o, (sweep, n_g, n_o)
        assert acc_g == acc_o
    np.testing.assert_allclose(g.download(), Uo, rtol=0, atol=1e-11)


@pytest.mark.parametrize("nc", [1, 2, 3])
def test_hmc_momenta_and_force_parity(nc):
    nd, dims = 4, (4, 4, 4, 6)
    U = seeded(nd, nc, dims, 0.7, 31)
    g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
    h = klft.HMC(g, beta=2.0, tau=1.0, nsteps=5, seed=77)
    h.refresh_momenta(3)
    np.testing.assert_allclose(h.momenta(), orc.refresh_momenta(nd, nc, dims, 77, 3), rtol=0, atol=1e-14)
    np.testing.assert_allclose(h.force(), orc.force(U, nd, nc, dims, 2.0), rtol=0, atol=1e-12)
    P = orc.refresh_momenta(nd, nc, dims, 77, 3)
    assert abs(h.kinetic() - orc.kinetic(P, nd, nc, dims)) < 1e-10
    assert abs(klft.gauge_action(g, 2.0) - orc.gauge_action(U, nd, nc, dims, 2.0)) < 1e-9


@pytest.mark.parametrize("nc", [1, 2, 3])
def test_hmc_update_gauge_parity(nc):
    nd, dims = 4, (4, 4, 4, 4)
    U = seeded(nd, nc, dims, 0.7, 32)
    g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
    h = klft.HMC(g, beta=2.0, tau=1.0, nsteps=5, seed=78)
    P = orc.refresh_momenta(nd, nc, dims, 78, 0)
    h.set_momenta(P)
    h.update_gauge(0.1)
    Uo = U.copy()
    orc.update_gauge(Uo, P, nd, nc, dims, 0.1)
    np.testing.assert_allclose(g.download(), Uo, rtol=0, atol=1e-13)


@pytest.mark.parametrize("nc,beta", [(1, 1.0), (2, 2.3), (3, 5.0)])
def test_hmc_trajectory_parity(nc, beta):
    nd, dims = 4, (4, 4, 4, 4)
    U = seeded(nd, nc, dims, 0.5, 40)
    g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
    h = klft.HMC(g, beta=beta, tau=1.0, nsteps=10, seed=99)
    Uo = U.copy()
    for traj in range(2):
        acc_o, dH_o, _ = orc.hmc_trajectory(Uo, nd, nc, dims, beta, 1.0, 10, 99, traj)
        dH, acc = h.trajectory(traj)
        assert abs(dH - dH_o) < 1e-8 * max(1.0, abs(dH_o))
        assert acc == acc_o
    np.testing.assert_allclose(g.download(), Uo, rtol=0, atol=1e-10)


@pytest.mark.parametrize("storage", ["f32", "f32_rows"])
@pytest.mark.parametrize("nc", [1, 2, 3])
def test_fp32_plaquette_within_1e6(nc, storage):
    nd, dims = 4, (4, 4, 6, 8)
    U = seeded(nd, nc, dims, 0.5, 13)
    ref = orc.plaquette(U, nd, nc, dims)
    g = klft.GaugeField(nd, nc, dims, storage).set_random_algebra(0.5, 13).reunitarize()
    got = klft.GaugePlaquette(g)
    assert abs(got - ref) < 1e-6


@pytest.mark.parametrize("variant", [0, 4, 7, 11, 12, 13])
@pytest.mark.parametrize("nc,beta,storage", [(1, 2.0, "f64"), (2, 2.3, "f64"), (3, 5.5, "f64"),
                                             (3, 5.5, "f32"), (2, 2.3, "f32_rows")])
def test_metropolis_variants_agree(monkeypatch, variant, nc, beta, storage):
    """Every sweep variant (one lane per site; one lane with the low-register
    staple; one lane building the next staple during the hits; one lane with
    the site's own links in registers; site waves fed by staple waves, in two
    staple term orders) makes the same decisions: FP64 against the oracle,
    FP32 against variant 0."""
    nd, dims = 4, (4, 6, 4, 4)
    U = orc.random_algebra(nd, nc, dims, 0.3, 19)
    p = klft.MetropolisParams(nHits=7, beta=beta, delta=0.3, seed=99)
    monkeypatch.setenv("KLFT_VARIANT", f"metropolis={variant}")
    g = klft.GaugeField(nd, nc, dims, storage).upload(U)
    n_g = [klft.sweep_Metropolis(g, p, s, with_count=True)[1] for s in range(2)]
    if storage == "f64":
        Uo = U.copy()
        n_o = [orc.metropolis_sweep(Uo, nd, nc, dims, 7, beta, 0.3, 99, s)[1] for s in range(2)]
        assert n_g == n_o
        np.testing.assert_allclose(g.download(), Uo, rtol=0, atol=1e-11)
    else:
        monkeypatch.setenv("KLFT_VARIANT", "metropolis=0")
        g0 = klft.GaugeField(nd, nc, dims, storage).upload(U)
        n_0 = [klft.sweep_Metropolis(g0, p, s, with_count=True)[1] for s in range(2)]
        assert abs(n_g[0] - n_0[0]) <= 2 and abs(n_g[1] - n_0[1]) <= 4


@pytest.mark.parametrize("nd,nc,dims", [(4, 3, (12, 12, 8, 10)), (4, 2, (8, 8, 8, 8)), (3, 3, (10, 8, 12)),
                                        (2, 1, (16, 12)), (4, 1, (6, 6, 6, 6))])
@pytest.mark.parametrize("variant", [12, 13])
def test_staple_wave_sweep_matches_oracle(monkeypatch, nd, nc, dims, variant):
    """The site-wave / staple-wave kernel on lattices with several blocks and
    a partly filled last block (every wave of a block meets every barrier),
    over three sweeps: FP64 decisions identical to the oracle, fields within
    1e-11."""
    U = orc.random_algebra(nd, nc, dims, 0.4, 71)
    beta = {1: 2.0, 2: 2.3, 3: 5.5}[nc]
    p = klft.MetropolisParams(nHits=6, beta=beta, delta=0.3, seed=123)
    monkeypatch.setenv("KLFT_VARIANT", f"metropolis={variant}")
    g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
    n_g = [klft.sweep_Metropolis(g, p, s, with_count=True)[1] for s in range(3)]
    Uo = U.copy()
    n_o = [orc.metropolis_sweep(Uo, nd, nc, dims, 6, beta, 0.3, 123, s)[1] for s in range(3)]
    assert n_g == n_o
    np.testing.assert_allclose(g.download(), Uo, rtol=0, atol=1e-11)


@pytest.mark.parametrize("variants", [(0, 0, 0), (3, 3, 4), (4, 0, 4), (5, 3, 0)])
@pytest.mark.parametrize("nd,nc,dims", [(4, 3, (4, 4, 6, 4)), (3, 2, (4, 6, 4)), (4, 1, (6, 4, 4, 4))])
def test_measurement_and_force_variants_match_oracle(monkeypatch, variants, nd, nc, dims):
    """Every plaquette variant (plane by plane, all loads in flight at once,
    threads walking dimension 0 with all loads in flight or plane by plane),
    staple-field variant (one term at a time, all loads in flight) and force
    variant (one term at a time, the full staple) against the oracle."""
    U = seeded(nd, nc, dims, 0.6, 41)
    vp, vs, vf = variants
    monkeypatch.setenv("KLFT_VARIANT", f"plaquette={vp},staple_field={vs},force={vf}")
    g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
    got, ps = klft.GaugePlaquette(g, per_site=True)
    want, ps_o = orc.plaquette(U, nd, nc, dims, per_site=True)
    np.testing.assert_allclose(ps, ps_o, rtol=0, atol=1e-12)
    assert abs(got - want) < 1e-12
    np.testing.assert_allclose(klft.stapleField(g).download(), orc.staple_field(U, nd, nc, dims), rtol=0, atol=1e-12)
    h = klft.HMC(g, beta=2.5, tau=1.0, nsteps=4, seed=9)
    np.testing.assert_allclose(h.force(), orc.force(U, nd, nc, dims, 2.5), rtol=0, atol=1e-12)
    # the momentum update itself (P -= dtau F) through the same variant
    h.refresh_momenta(1)
    P0 = h.momenta()
    h.update_momentum(0.125)
    P1 = P0.copy()
    orc.update_momentum(P1, orc.force(U, nd, nc, dims, 2.5), nd, nc, dims, 0.125)
    np.testing.assert_allclose(h.momenta(), P1, rtol=0, atol=1e-12)


def test_removed_variant_is_an_error(monkeypatch):
    """A variant number that is no longer built (four lanes per site, 1) fails
    with the library's message instead of running something else."""
    monkeypatch.setenv("KLFT_VARIANT", "metropolis=1")
    g = klft.GaugeField(4, 2, (4, 4, 4, 4), "f64").set_identity()
    with pytest.raises(klft.KlftError, match="not built"):
        klft.sweep_Metropolis(g, klft.MetropolisParams(nHits=2, beta=2.0, delta=0.2, seed=1), 0)


@pytest.mark.parametrize("storage,dims,nc,variant", [
    ("f64", (12, 12, 8, 10), 3, 13), ("f32", (16, 16, 16, 16), 3, 13), ("f32_rows", (16, 16, 16, 16), 3, 12),
    ("f32", (32, 32, 32, 32), 3, 12), ("f32_rows", (32, 32, 32, 32), 3, 12), ("f64", (10, 8, 12), 3, 12),
    ("f64", (8, 8, 8, 8), 2, 13), ("f32", (6, 6, 6, 6), 1, 12)])
def test_paired_passes_equal_sequential_sweep(monkeypatch, storage, dims, nc, variant):
    """Passes three or more coordinate parities apart read none of each
    other's links, so a sweep that runs such passes two to a launch (4D:
    {3,4} {5,8} {6,9} {7,10} {11,12}; 3D: {3,4}) leaves the links and the
    accept counts bit for bit as the pass-by-pass sweep does.  32^4 FP32 pairs
    two 512-thread blocks per CU; the smaller lattices one."""
    nd = len(dims)
    beta = {1: 2.0, 2: 2.3, 3: 6.0}[nc]
    p = klft.MetropolisParams(nHits=10, beta=beta, delta=0.1, seed=4242)
    out = []
    for pairs in ("0", "1"):
        monkeypatch.setenv("KLFT_VARIANT", f"metropolis={variant},metropolis_pairs={pairs}")
        g = klft.GaugeField(nd, nc, dims, storage).set_random_algebra(0.5, 2024)
        n = [klft.sweep_Metropolis(g, p, s, with_count=True)[1] for s in range(3)]
        out.append((n, g.download()))
        g.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("storage,dims,nc,variant", [
    ("f64", (12, 12, 8, 10), 3, 13), ("f32", (16, 16, 16, 16), 3, 13), ("f32_rows", (16, 16, 16, 16), 3, 12),
    ("f32", (32, 32, 32, 32), 3, 12), ("f64", (10, 8, 12), 3, 12), ("f64", (8, 8, 8, 8), 2, 13),
    ("f32", (6, 6, 6, 6), 1, 12), ("f64", (8, 8, 8, 8), 3, 15)])
def test_run_pairing_across_sweeps_equals_sequential_sweeps(monkeypatch, storage, dims, nc, variant):
    """metropolis_run over several sweeps also runs independent passes of
    consecutive sweeps in one launch (4D: 13|0', 14|1', 15|2'); the links and
    every sweep's acceptance rate equal those of sweeps run one by one with
    no pairing, bit for bit (variant 15: the symmetric SU(3) proposal)."""
    nd = len(dims)
    beta = {1: 2.0, 2: 2.3, 3: 6.0}[nc]
    p = klft.MetropolisParams(nHits=10, beta=beta, delta=0.1, seed=4242)
    out = []
    for pairs in ("0", "1"):
        monkeypatch.setenv("KLFT_VARIANT", f"metropolis={variant},metropolis_pairs={pairs}")
        g = klft.GaugeField(nd, nc, dims, storage).set_random_algebra(0.5, 2024)
        if variant == 15:
            g.set_proposal("symmetric")
        if pairs == "0":
            rates = [klft.sweep_Metropolis(g, p, s) for s in range(3, 11)]
        else:
            # runs of 1, 2 and 5 sweeps: no window, one, several
            rates = [r for s0, n in ((3, 1), (4, 2), (6, 5)) for r in klft.metropolis_run(g, p, s0, n)]
        out.append((rates, g.download()))
        g.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("storage,nc", [("f32", 2), ("f64", 3)])
def test_long_run_equals_sequential_sweeps(monkeypatch, storage, nc):
    """A run of 300 sweeps (head, 299 windows of the one-sweep period, tail)
    leaves the links and all 300 acceptance rates bit for bit as 300 single
    sweeps with unpaired passes do."""
    dims = (4, 4, 4, 4)
    p = klft.MetropolisParams(nHits=4, beta={2: 2.3, 3: 5.5}[nc], delta=0.2, seed=77)
    out = []
    for pairs in ("0", "1"):
        monkeypatch.setenv("KLFT_VARIANT", f"metropolis_pairs={pairs}")
        g = klft.GaugeField(4, nc, dims, storage).set_random_algebra(0.5, 9)
        if pairs == "0":
            rates = [klft.sweep_Metropolis(g, p, s) for s in range(300)]
        else:
            rates = list(klft.metropolis_run(g, p, 0, 300))
        out.append((rates, g.download()))
        g.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("storage", ["f64", "f32", "f32_rows"])
def test_new_field_links_start_zeroed(storage):
    """A field's links read back as zeros before any initialiser runs."""
    g = klft.GaugeField(4, 3, (4, 4, 4, 6), storage)
    U = g.download()
    g.close()
    assert U.shape == (4, 4, 4, 6, 4, 3, 3)
    assert not np.any(U)
