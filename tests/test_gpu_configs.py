"""The benchmark configurations at their configured shapes, against the oracle.

* SU(2) HMC, 4D 16^4, beta = 2.3, tau = 1, 20 leapfrog steps, hot start
  randSUN(0.5) keyed by 2 * seed (lib/HMC_SU2.cpp:41-42): FP64 links follow
  the oracle trajectory by trajectory; FP32 links with FP64 dH / action sums
  stay within a stated dH bound with the same accept decisions.
* SU(3) 32^4 measurement kernels on fixed seeded fields exp(i eps H): the
  staple field per link and the temporal Wilson loops for every L, T in
  1..8 (64 pairs), FP64 within 1e-12 and FP32 / FP32 two-row within 1e-6.
* 64 independent chains, extents drawn (seeded) from {8, 10, 12, 16, 20} per
  dimension, U(1) and SU(2) alternating, beta = 1.5 + k / 63, nHits = 10,
  delta = 0.1: FP64 chains match the oracle chain by chain (decisions and
  links); FP32 chains agree with the FP64 ones in their mean plaquette.
"""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

import klft_amd as klft  # noqa: E402

PAIRS = [(L, T) for L in range(1, 9) for T in range(1, 9)]


# ---------------------------------------------------------------------------
# SU(2) 16^4 HMC
# ---------------------------------------------------------------------------
C2_DIMS, C2_BETA, C2_SEED = (16, 16, 16, 16), 2.3, 1234


@pytest.fixture(scope="module")
def c2_oracle():
    U = orc.random_sun_field(4, 2, C2_DIMS, 0.5, 2 * C2_SEED)
    start = U.copy()
    out = []
    for traj in range(3):
        acc, dH, plaq = orc.hmc_trajectory(U, 4, 2, C2_DIMS, C2_BETA, 1.0, 20, C2_SEED, traj)
        out.append((acc, dH, plaq, U.copy()))
    return start, out


def test_c2_hot_start_matches_oracle(c2_oracle):
    start, _ = c2_oracle
    g = klft.GaugeField(4, 2, C2_DIMS, "f64").set_random_sun(0.5, 2 * C2_SEED)
    np.testing.assert_allclose(g.download(), start, rtol=0, atol=1e-14)


def test_c2_fp64_trajectories_match_oracle(c2_oracle):
    start, ref = c2_oracle
    g = klft.GaugeField(4, 2, C2_DIMS, "f64").upload(start)
    h = klft.HMC(g, beta=C2_BETA, tau=1.0, nsteps=20, seed=C2_SEED)
    for traj, (acc_o, dH_o, plaq_o, U_o) in enumerate(ref):
        dH, acc = h.trajectory(traj)
        print(f"C2 f64 traj {traj}: dH {dH:.10g} (oracle {dH_o:.10g}) accept {acc}")
        assert acc == acc_o
        assert abs(dH - dH_o) < 1e-8 * max(1.0, abs(dH_o))
        assert abs(klft.GaugePlaquette(g) - plaq_o) < 1e-12
        np.testing.assert_allclose(g.download(), U_o, rtol=0, atol=1e-10)


@pytest.mark.parametrize("storage", ["f32", "f32_rows"])
def test_c2_fp32_trajectories_track_oracle(c2_oracle, storage):
    """FP32 links (+ FP64 dH and action sums): |dH32 - dH64| < 1e-4 (1 +
    |dH64|) per trajectory and the same accept decisions while the FP64
    decision is not within 1 % of its threshold; plaquette within 1e-5."""
    start, ref = c2_oracle
    g = klft.GaugeField(4, 2, C2_DIMS, storage).upload(start)
    h = klft.HMC(g, beta=C2_BETA, tau=1.0, nsteps=20, seed=C2_SEED)
    for traj, (acc_o, dH_o, plaq_o, _) in enumerate(ref):
        dH, acc = h.trajectory(traj)
        rel = abs(dH - dH_o) / (1.0 + abs(dH_o))
        print(f"C2 {storage} traj {traj}: dH {dH:.8g} (oracle {dH_o:.8g}), rel err {rel:.2e}, accept {acc}")
        assert rel < 1e-4
        if acc != acc_o:
            u = orc.stream_doubles(C2_SEED, 7, 0, traj, 0, 0, 0, 1)[0]
            assert abs(u - np.exp(-dH_o)) <= 0.01 * np.exp(-dH_o)
            return
        assert abs(klft.GaugePlaquette(g) - plaq_o) < 1e-5


# ---------------------------------------------------------------------------
# SU(3) 32^4 measurement kernels
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module", params=[0.1, 1.0])
def c3(request):
    eps = request.param
    dims = (32, 32, 32, 32)
    U = orc.random_algebra(4, 3, dims, eps, 3000 + int(10 * eps))
    return eps, dims, U, orc.wilson_temporal(U, 4, 3, dims, PAIRS)


@pytest.mark.parametrize("storage", ["f64", "f32", "f32_rows"])
def test_c3_wilson_temporal_all_pairs_32(c3, storage):
    eps, dims, U, ref = c3
    g = klft.GaugeField(4, 3, dims, storage).upload(U)
    got = np.array([w for _, _, w in klft.WilsonLoop_temporal(g, PAIRS)])
    err = float(np.abs(got - ref).max())
    bound = 1e-12 if storage == "f64" else 1e-6
    print(f"C3 W_temp 32^4 eps={eps} {storage}: max err {err:.2e} (bound {bound:.0e})")
    assert err < bound


def test_c3_staple_field_32(c3):
    """stapleField at 32^4 per link: FP64 within 1e-12, FP32 within 1e-5."""
    eps, dims, U, _ = c3
    S = orc.staple_field(U, 4, 3, dims)
    for storage, bound in [("f64", 1e-12), ("f32", 1e-5), ("f32_rows", 1e-5)]:
        g = klft.GaugeField(4, 3, dims, storage).upload(U)
        s = klft.stapleField(g)
        err = float(np.abs(s.download() - S).max())
        print(f"C3 staple field 32^4 eps={eps} {storage}: max err {err:.2e} (bound {bound:.0e})")
        assert err < bound
        s.close()
        g.close()


# ---------------------------------------------------------------------------
# 64-chain batch
# ---------------------------------------------------------------------------
def c4_chains():
    rng = np.random.default_rng(64)
    ext = [8, 10, 12, 16, 20]
    return [(1 if k % 2 == 0 else 2, tuple(int(x) for x in rng.choice(ext, 4)), 1.5 + k / 63.0)
            for k in range(64)]


def test_c4_fp64_batch_matches_oracle_chain_by_chain():
    chains = c4_chains()
    b = klft.Batch(chains, "f64")
    refs = []
    for i, (nc, dims, _) in enumerate(chains):
        U = orc.random_algebra(4, nc, dims, 0.3, 700 + i)
        b.upload(i, U)
        refs.append(U)
    acc = b.sweep(10, 0.1, 2025, 0)
    for i, (nc, dims, beta) in enumerate(chains):
        a, _ = orc.metropolis_sweep_defect(refs[i], np.ones((*dims, 4)), 4, nc, dims, 10, beta, 0.1, 2025, 0, i)
        assert acc[i] == a, (i, acc[i], a)
        np.testing.assert_allclose(b.download(i), refs[i], rtol=0, atol=1e-11)


def test_c4_fp32_batch_matches_oracle_chain_by_chain():
    """FP32 links: each of the 64 chains takes the decisions of the oracle run
    with the FP32 draws (24-bit uniforms) and the chain's stream index, in
    FP64 arithmetic; accept counts within 0.5 % (at least 2) per chain, links
    that agree to 1e-4 at least 97 % of each chain's links and within 5e-6 on
    those (a decision that FP32 rounding moves across its threshold changes
    that link and its neighbours' later hits)."""
    chains = c4_chains()
    b = klft.Batch(chains, "f32")
    refs = []
    for i, (nc, dims, _) in enumerate(chains):
        U = orc.random_algebra(4, nc, dims, 0.3, 900 + i)
        b.upload(i, U)
        refs.append(U)
    acc = b.sweep(10, 0.1, 2026, 0)
    worst_frac, worst_err, ndiff = 1.0, 0.0, 0
    for i, (nc, dims, beta) in enumerate(chains):
        a, n_o = orc.metropolis_sweep_narrow(refs[i], 4, nc, dims, 10, beta, 0.1, 2026, 0, chain=i)
        n_g = int(round(acc[i] * np.prod(dims) * 4 * 10))
        ndiff += int(n_g != n_o)
        assert abs(n_g - n_o) <= max(2, 0.005 * n_o), (i, n_g, n_o)
        d = np.abs(b.download(i) - refs[i]).reshape(-1, nc * nc).max(axis=1)
        agree = d < 1e-4
        worst_frac = min(worst_frac, float(agree.mean()))
        worst_err = max(worst_err, float(d[agree].max()))
    print(f"C4 FP32 batch vs oracle (FP32 draws): chains with a differing count {ndiff}/64, "
          f"links agreeing >= {worst_frac:.4%} per chain, max |dU| on them {worst_err:.2e}")
    assert worst_frac >= 0.97 and worst_err < 5e-6


def test_c4_fp32_batch_tracks_fp64():
    """FP32 links + FP64 plaquette sums against FP64 links, same chains and
    keyed streams, from the cold start: after 100 graph-replayed sweeps, 30
    plaquettes every 5 sweeps per chain; the per-chain means agree within
    5 combined standard errors (6 bins of 5) plus 1e-4."""
    chains = c4_chains()
    b32, b64 = klft.Batch(chains, "f32"), klft.Batch(chains, "f64")
    b32.set_identity()
    b64.set_identity()
    b32.run(10, 0.1, 77, 0, 100)
    b64.run(10, 0.1, 77, 0, 100)
    p32, p64 = [], []
    for s in range(30):
        b32.run(10, 0.1, 77, 100 + 5 * s, 5)
        b64.run(10, 0.1, 77, 100 + 5 * s, 5)
        p32.append(b32.plaquette())
        p64.append(b64.plaquette())
    p32, p64 = np.array(p32), np.array(p64)

    def binned(p):
        bins = p.reshape(6, 5, -1).mean(axis=1)
        return bins.mean(axis=0), bins.std(axis=0, ddof=1) / np.sqrt(6)
    m32, e32 = binned(p32)
    m64, e64 = binned(p64)
    z = np.abs(m32 - m64) / (np.hypot(e32, e64) + 1e-12)
    print(f"C4 FP32 vs FP64 per-chain plaquette: max |z| {z.max():.2f}, max |diff| {np.abs(m32 - m64).max():.2e}")
    assert np.all(np.abs(m32 - m64) < 5 * np.hypot(e32, e64) + 1e-4)
