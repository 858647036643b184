"""GPU parity of the batched-chain and PTBC replica engine against the oracle.

Chains are keyed by their index in the Philox counter, so chain i of a batch
reproduces the oracle sweep with chain = i (orc.metropolis_sweep_defect with
all coupling factors 1).  FP64 storage makes the comparison exact up to
rounding; FP32 batches are checked statistically.
"""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu
import klft_amd as klft  # noqa: E402

CHAINS = [(1, (4, 6, 4, 4), 1.5), (2, (6, 4, 4, 8), 2.1), (3, (4, 4, 6, 4), 5.2), (2, (4, 4, 4, 4), 2.5),
          (1, (8, 4, 4, 6), 2.0)]


def start(nc, dims, i):
    return orc.random_algebra(4, nc, dims, 0.4, 100 + i)


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_batch_fp64_matches_oracle_per_chain(monkeypatch, variant):
    """The pass kernels (0: one lane per site, 1: site waves fed by staple
    waves, 2: the same with each direction's upper and lower staple terms back
    to back) against the oracle chain by chain."""
    monkeypatch.setenv("KLFT_VARIANT", f"batch={variant}")
    b = klft.Batch(CHAINS, "f64")
    for i, (nc, dims, _) in enumerate(CHAINS):
        b.upload(i, start(nc, dims, i))
    refs = [start(nc, dims, i) for i, (nc, dims, _) in enumerate(CHAINS)]
    for sweep in range(2):
        acc = b.sweep(6, 0.3, 777, sweep)
        for i, (nc, dims, beta) in enumerate(CHAINS):
            ones = np.ones((*dims, 4))
            a, n = orc.metropolis_sweep_defect(refs[i], ones, 4, nc, dims, 6, beta, 0.3, 777, sweep, i)
            assert acc[i] == a, (sweep, i, acc[i], a)
    for i, (nc, dims, _) in enumerate(CHAINS):
        np.testing.assert_allclose(b.download(i), refs[i], rtol=0, atol=1e-11)
    np.testing.assert_allclose(b.plaquette(), [orc.plaquette(refs[i], 4, nc, dims)
                                               for i, (nc, dims, _) in enumerate(CHAINS)], atol=1e-13)


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_batch_graph_replay_equals_direct_sweeps(monkeypatch, storage, variant):
    monkeypatch.setenv("KLFT_VARIANT", f"batch={variant}")
    b1 = klft.Batch(CHAINS, storage)
    b2 = klft.Batch(CHAINS, storage)
    for i, (nc, dims, _) in enumerate(CHAINS):
        U = start(nc, dims, i)
        b1.upload(i, U)
        b2.upload(i, U)
    for s in range(5):
        b1.sweep(4, 0.2, 31, 10 + s)
    b2.run(4, 0.2, 31, 10, 5)
    for i in range(len(CHAINS)):
        assert np.array_equal(b1.download(i), b2.download(i))


def test_batch_fp32_plaquette_tracks_fp64():
    """FP32 links with FP64 plaquette sums: the same chains stay within
    statistical agreement of the FP64 ones after thermalisation."""
    chains = [(1, (8, 8, 8, 8), 1.0 + 0.1 * k) for k in range(4)] + [(2, (8, 8, 8, 8), 2.0 + 0.1 * k) for k in range(4)]
    b32, b64 = klft.Batch(chains, "f32"), klft.Batch(chains, "f64")
    b32.set_identity()
    b64.set_identity()
    b32.run(10, 0.5, 5, 0, 150)
    b64.run(10, 0.5, 5, 0, 150)
    p32, p64 = [], []
    for s in range(40):
        b32.run(10, 0.5, 5, 150 + 5 * s, 5)
        b64.run(10, 0.5, 5, 150 + 5 * s, 5)
        p32.append(b32.plaquette())
        p64.append(b64.plaquette())
    p32, p64 = np.array(p32), np.array(p64)
    err = np.sqrt(p32.var(axis=0) / 8 + p64.var(axis=0) / 8) + 1e-4
    assert np.all(np.abs(p32.mean(axis=0) - p64.mean(axis=0)) < 5 * err)


@pytest.mark.parametrize("variant,dims", [(0, (4, 4, 4, 4)), (1, (4, 4, 4, 4)), (1, (12, 8, 8, 8)), (2, (12, 8, 8, 8))])
def test_ptbc_fp64_step_matches_oracle(monkeypatch, variant, dims):
    """Replica sweeps and swaps against the oracle, for both pass kernels; the
    12 x 8^3 replicas take two blocks per pass, the second partly filled."""
    monkeypatch.setenv("KLFT_VARIANT", f"batch={variant}")
    nrep, nc, beta = 4, 3, 5.5
    L, pos = 2, [1, 0, 2]
    c = np.linspace(1.0, 0.0, nrep)
    seed = 4321
    p = klft.PTBC(nrep, nc, dims, beta, L, pos, c, "f64")
    refs = [orc.random_algebra(4, nc, dims, 0.3, 50 + r) for r in range(nrep)]
    for r in range(nrep):
        p.upload(r, refs[r])
    conf_of = list(range(nrep))   # conf_of[k]: configuration holding coupling k
    for step in range(3):
        swaps = p.step(5, 0.2, seed, step)
        cval = {conf_of[k]: c[k] for k in range(nrep)}
        for r in range(nrep):
            C = orc.defect_field(4, dims, L, pos, cval[r])
            orc.metropolis_sweep_defect(refs[r], C, 4, nc, dims, 5, beta, 0.2, seed, step, r)
        for r in range(nrep):
            np.testing.assert_allclose(p.download(r), refs[r], rtol=0, atol=1e-11)

        def S(U, ck):
            return -(beta / nc) * orc.plaquette_defect(U, orc.defect_field(4, dims, L, pos, ck), 4, nc, dims,
                                                       normalize=False)
        want = []
        for k in range(nrep - 1):
            a, b_ = conf_of[k], conf_of[k + 1]
            dS = S(refs[a], c[k + 1]) + S(refs[b_], c[k]) - S(refs[a], c[k]) - S(refs[b_], c[k + 1])
            u = orc.stream_doubles(seed, 8, k, step, 0, 0, 0, 1)[0]
            acc = dS <= 0 or u < np.exp(-dS)
            want.append(int(acc))
            if acc:
                conf_of[k], conf_of[k + 1] = b_, a
        assert list(swaps) == want
        assert list(p.labels()) == conf_of
    for r in range(nrep):
        for k in range(nrep):
            want = -(beta / nc) * orc.plaquette_defect(refs[r], orc.defect_field(4, dims, L, pos, c[k]), 4, nc,
                                                       dims, normalize=False)
            assert abs(p.action(r, k) - want) < 1e-9 * abs(want)


def test_ptbc_fp32_step_matches_oracle_with_fp32_draws():
    """C5's setting at 8^4 (SU(3), 2^3 defect, c_r linear from 1 to 0, nHits
    = 8, delta = 0.2) in FP32: each replica's sweep takes the decisions of the
    oracle run with the FP32 draws and the replica's stream index (links that
    agree to 1e-4: at least 97 %, within 5e-6 on those), and the FP64 swap
    pass on the FP32 fields takes the oracle's swap decisions (Metropolis.hpp
    :72-113 with the defect scaling of PTBC)."""
    nrep, nc, beta, dims = 4, 3, 6.0, (8, 8, 8, 8)
    L, pos = 2, [0, 0, 0]
    c = np.linspace(1.0, 0.0, nrep)
    seed = 5150
    p = klft.PTBC(nrep, nc, dims, beta, L, pos, c, "f32")
    refs = [orc.random_algebra(4, nc, dims, 0.3, 80 + r) for r in range(nrep)]
    for r in range(nrep):
        p.upload(r, refs[r])
    conf_of = list(range(nrep))
    worst_frac, worst_err = 1.0, 0.0
    for step in range(3):
        swaps = p.step(8, 0.2, seed, step)
        cval = {conf_of[k]: c[k] for k in range(nrep)}
        for r in range(nrep):
            C = orc.defect_field(4, dims, L, pos, cval[r])
            orc.metropolis_sweep_defect(refs[r], C, 4, nc, dims, 8, beta, 0.2, seed, step, r, narrow=True)
            d = np.abs(p.download(r) - refs[r]).reshape(-1, nc * nc).max(axis=1)
            agree = d < 1e-4
            worst_frac = min(worst_frac, float(agree.mean()))
            worst_err = max(worst_err, float(d[agree].max()))

        def S(U, ck):
            return -(beta / nc) * orc.plaquette_defect(U, orc.defect_field(4, dims, L, pos, ck), 4, nc, dims,
                                                       normalize=False)
        want = []
        for k in range(nrep - 1):
            a, b_ = conf_of[k], conf_of[k + 1]
            dS = S(refs[a], c[k + 1]) + S(refs[b_], c[k]) - S(refs[a], c[k]) - S(refs[b_], c[k + 1])
            u = orc.stream_doubles(seed, 8, k, step, 0, 0, 0, 1)[0]
            acc = dS <= 0 or u < np.exp(-dS)
            want.append(int(acc))
            if acc:
                conf_of[k], conf_of[k + 1] = b_, a
        assert list(swaps) == want, (step, list(swaps), want)
        assert list(p.labels()) == conf_of
    print(f"PTBC FP32 vs oracle (FP32 draws): links agreeing >= {worst_frac:.4%}, max |dU| on them "
          f"{worst_err:.2e}; swap decisions identical")
    assert worst_frac >= 0.97 and worst_err < 5e-6


def test_ptbc_c5_shape_runs_in_fp32():
    """The 16-replica SU(3) 16^4 set: one graph-captured step (sweep of every
    replica, FP64 swap pass) completes and keeps the labels a permutation."""
    nrep = 16
    c = np.linspace(1.0, 0.0, nrep)
    p = klft.PTBC(nrep, 3, (16, 16, 16, 16), 6.0, 2, [0, 0, 0], c, "f32")
    p.set_identity()
    for step in range(3):
        p.step(8, 0.2, 9, step)
    assert sorted(p.labels()) == list(range(nrep))
    pl = p.plaquette()
    assert np.all((pl > 0.3) & (pl <= 1.0))


@pytest.mark.parametrize("nrep,nc,dims,L,pos", [(70, 2, (4, 4, 4, 4), 2, [1, 3, 0]),
                                               (5, 3, (4, 6, 4, 4), 4, [0, 2, 1])])
def test_ptbc_many_replicas_and_full_length_defect(nrep, nc, dims, L, pos):
    """More than 64 replicas (the defect sums' grid covers every replica), and
    a defect as long as a spatial extent (L = L1 = L3 = 4: no plaquette is
    counted twice): swap decisions, labels and every replica's action under
    every tested coupling against the oracle's defect-scaled action."""
    beta, seed = 2.3 if nc == 2 else 5.5, 97
    c = np.linspace(1.0, 0.0, nrep)
    p = klft.PTBC(nrep, nc, dims, beta, L, pos, c, "f64")
    refs = [orc.random_algebra(4, nc, dims, 0.3, 900 + r) for r in range(nrep)]
    for r in range(nrep):
        p.upload(r, refs[r])

    def S(U, ck):
        return -(beta / nc) * orc.plaquette_defect(U, orc.defect_field(4, dims, L, pos, ck), 4, nc, dims,
                                                   normalize=False)
    for r in [0, nrep // 2, nrep - 1]:
        for k in [0, nrep - 1]:
            want = S(refs[r], c[k])
            assert abs(p.action(r, k) - want) < 1e-9 * abs(want), (r, k)
    conf_of = list(range(nrep))
    swaps = p.step(3, 0.2, seed, 0)
    cval = {conf_of[k]: c[k] for k in range(nrep)}
    for r in range(nrep):
        orc.metropolis_sweep_defect(refs[r], orc.defect_field(4, dims, L, pos, cval[r]), 4, nc, dims, 3, beta, 0.2,
                                    seed, 0, r)
    want = []
    for k in range(nrep - 1):
        a, b_ = conf_of[k], conf_of[k + 1]
        dS = S(refs[a], c[k + 1]) + S(refs[b_], c[k]) - S(refs[a], c[k]) - S(refs[b_], c[k + 1])
        u = orc.stream_doubles(seed, 8, k, 0, 0, 0, 0, 1)[0]
        acc = dS <= 0 or u < np.exp(-dS)
        want.append(int(acc))
        if acc:
            conf_of[k], conf_of[k + 1] = b_, a
    assert list(swaps) == want
    assert list(p.labels()) == conf_of


@pytest.mark.parametrize("storage", ["f32", "f32_rows"])
def test_batch_run_pairs_across_sweeps_at_bench_shape(storage):
    """16 SU(3) chains of 16^4 (the batched bench leg; passes pair at two
    blocks per CU): a graph-replayed run of 4 sweeps, whose passes of
    consecutive sweeps share launches (13|0', 14|1', 15|2'), leaves every
    chain's links bit for bit as 4 single sweeps do."""
    chains = [(3, (16, 16, 16, 16), 5.5 + 0.05 * i) for i in range(16)]
    out = []
    for mode in ("sweeps", "run"):
        b = klft.Batch(chains, storage)
        b.set_identity()
        if mode == "sweeps":
            for s in range(4):
                b.sweep(10, 0.1, 4242, 7 + s)
        else:
            b.run(10, 0.1, 4242, 7, 4)
        out.append(([b.download(i) for i in (0, 5, 15)], b.plaquette()))
        b.close()
    for u0, u1 in zip(out[0][0], out[1][0]):
        assert np.array_equal(u0, u1)
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][1].max() < 1.0


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("storage", ["f64", "f32", "f32_rows"])
def test_batch_paired_passes_equal_sequential(monkeypatch, storage, variant):
    """Independent passes two to a launch (the default while both passes'
    blocks fit on the GPU) leave every chain's links and accept counts bit for
    bit as the pass-by-pass sweep does; the same for a PTBC step with its
    swap decisions."""
    chains = [(3, (8, 8, 8, 8), 6.0), (2, (8, 6, 4, 10), 2.3), (1, (6, 6, 6, 6), 2.0), (3, (16, 8, 8, 8), 5.5)]
    if storage == "f32_rows":
        chains = [c for c in chains if c[0] > 1]
    out = []
    for pairs in ("0", "1"):
        monkeypatch.setenv("KLFT_VARIANT", f"batch={variant},batch_pairs={pairs}")
        b = klft.Batch(chains, storage)
        for i, (nc, dims, _) in enumerate(chains):
            b.upload(i, start(nc, dims, i))
        acc = [list(b.sweep(10, 0.2, 55, s)) for s in range(2)]
        b.run(10, 0.2, 55, 2, 2)  # graph-captured sweeps
        out.append((acc, [b.download(i) for i in range(len(chains))]))
        b.close()
    assert out[0][0] == out[1][0]
    for u0, u1 in zip(out[0][1], out[1][1]):
        assert np.array_equal(u0, u1)
    res = []
    for pairs in ("0", "1"):
        monkeypatch.setenv("KLFT_VARIANT", f"batch={variant},batch_pairs={pairs}")
        p = klft.PTBC(6, 3, (8, 8, 8, 8), 5.5, 2, [0, 0, 0], np.linspace(1.0, 0.0, 6),
                      "f32" if storage == "f32_rows" else storage)
        for r in range(6):
            p.upload(r, start(3, (8, 8, 8, 8), 40 + r))
        sw = [list(p.step(8, 0.2, 9, s)) for s in range(3)]
        res.append((sw, list(p.labels()), [p.download(i) for i in range(6)]))
        p.close()
    assert res[0][0] == res[1][0] and res[0][1] == res[1][1]
    for u0, u1 in zip(res[0][2], res[1][2]):
        assert np.array_equal(u0, u1)
