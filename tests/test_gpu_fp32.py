"""FP32 and FP32 two-row storage against the FP64 oracle.

The links are held and multiplied in FP32 (two-row storage keeps the first
Nc-1 rows and rebuilds the last as conj(r0 x r1)); FP64 is kept for every
global sum.  Each check compares the FP32 result with the oracle's FP64
result on the same FP64 input field, so the difference is the precision
error of the FP32 path.  The bounds are stated per test (max |difference|
over all entries); every test prints the measured error.

Reference semantics: GaugeMonomial.hpp:27-37 (force), HamiltonianField.hpp:
32-38 (exp update), Gauge_Util.hpp:27-97 (staple field), WilsonLoop.hpp:
170-304 (Wilson loops), HMC.hpp:48-71 (trajectory), GaugePlaquette.hpp:
89-135 (plaquette).
"""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

import klft_amd as klft  # noqa: E402

STORAGES = ["f32", "f32_rows"]
EPS = [0.1, 0.5, 1.0]
PAIRS = [(L, T) for L in range(1, 9) for T in range(1, 9)]


def report(name, err, bound):
    print(f"{name}: max |FP32 - FP64 oracle| = {err:.3e} (bound {bound:.0e})")
    assert err < bound, (name, err, bound)


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("eps", EPS)
@pytest.mark.parametrize("nc", [1, 2, 3])
def test_fp32_force(nc, eps, storage):
    """F_a = (beta/Nc) Im Tr(T_a U Sigma): bound 2e-5 * beta/Nc (entries of
    the staple are up to 2(nd-1) = 6 in magnitude)."""
    if nc == 1 and storage == "f32_rows":
        pytest.skip("U(1) has no two-row form")
    nd, dims, beta = 4, (6, 4, 4, 6), 6.0
    U = orc.random_algebra(nd, nc, dims, eps, 300 + nc)
    g = klft.GaugeField(nd, nc, dims, storage).upload(U)
    h = klft.HMC(g, beta=beta, tau=1.0, nsteps=4, seed=3)
    err = float(np.abs(h.force() - orc.force(U, nd, nc, dims, beta)).max())
    report(f"force nc={nc} eps={eps} {storage}", err, 2e-5 * beta / nc)


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("nc", [1, 2, 3])
def test_fp32_exp_update(nc, storage):
    """U <- exp(i dtau P) U with dtau = 0.1, P ~ N(0,1) per generator; FP32
    links are re-unitarised after the update (Gram-Schmidt, FP64 norms).
    Bound 2e-6 per entry."""
    if nc == 1 and storage == "f32_rows":
        pytest.skip("U(1) has no two-row form")
    nd, dims = 4, (4, 4, 6, 4)
    U = orc.random_algebra(nd, nc, dims, 0.7, 410 + nc)
    P = orc.refresh_momenta(nd, nc, dims, 78, 0)
    g = klft.GaugeField(nd, nc, dims, storage).upload(U)
    h = klft.HMC(g, beta=2.0, tau=1.0, nsteps=5, seed=78)
    h.set_momenta(P)
    h.update_gauge(0.1)
    Uo = U.copy()
    orc.update_gauge(Uo, P, nd, nc, dims, 0.1)
    report(f"exp update nc={nc} {storage}", float(np.abs(g.download() - Uo).max()), 2e-6)


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("eps", EPS)
def test_fp32_staple_field(eps, storage):
    """stapleField of an SU(3) field: bound 1e-5 per entry (entries up to 6)."""
    nd, nc, dims = 4, 3, (6, 6, 4, 4)
    U = orc.random_algebra(nd, nc, dims, eps, 520)
    g = klft.GaugeField(nd, nc, dims, storage).upload(U)
    err = float(np.abs(klft.stapleField(g).download() - orc.staple_field(U, nd, nc, dims)).max())
    report(f"staple field eps={eps} {storage}", err, 1e-5)


@pytest.fixture(scope="module", params=EPS)
def field16(request):
    eps = request.param
    dims = (16, 16, 16, 16)
    U = orc.random_algebra(4, 3, dims, eps, 3000 + int(10 * eps))
    return eps, dims, U, orc.wilson_temporal(U, 4, 3, dims, PAIRS)


def test_fp64_wilson_temporal_all_pairs_16(field16):
    """Temporal Wilson loops for every L, T in 1..8 (64 pairs) at 16^4, FP64
    storage: within 1e-12 of the oracle (loops up to half the extent)."""
    eps, dims, U, ref = field16
    g = klft.GaugeField(4, 3, dims, "f64").upload(U)
    got = np.array([w for _, _, w in klft.WilsonLoop_temporal(g, PAIRS)])
    err = float(np.abs(got - ref).max())
    print(f"W_temp 16^4 eps={eps} f64: max err {err:.2e}")
    assert err < 1e-12


@pytest.mark.parametrize("storage", STORAGES)
def test_fp32_wilson_temporal_all_pairs_16(field16, storage):
    """The same 64 pairs from FP32 / FP32 two-row links, FP64 sums: bound
    1e-6 per pair (the plaquette bound of the FP32 path)."""
    eps, dims, U, ref = field16
    g = klft.GaugeField(4, 3, dims, storage).upload(U)
    got = np.array([w for _, _, w in klft.WilsonLoop_temporal(g, PAIRS)])
    report(f"W_temp 16^4 eps={eps} {storage}", float(np.abs(got - ref).max()), 1e-6)


@pytest.mark.parametrize("storage", STORAGES)
def test_fp32_wilson_mu_nu_and_plaquette_16(field16, storage):
    """Planar loops in plane (0, 1) for L, T in 1..8 and the plaquette,
    FP32 links with FP64 sums: bound 1e-6."""
    eps, dims, U, _ = field16
    g = klft.GaugeField(4, 3, dims, storage).upload(U)
    got = np.array([r[4] for r in klft.WilsonLoop_mu_nu(g, 0, 1, PAIRS)])
    ref = orc.wilson_mu_nu(U, 4, 3, dims, 0, 1, PAIRS)
    report(f"W_mu_nu(0,1) 16^4 eps={eps} {storage}", float(np.abs(got - ref).max()), 1e-6)
    report(f"plaquette 16^4 eps={eps} {storage}", abs(klft.GaugePlaquette(g) - orc.plaquette(U, 4, 3, dims)), 1e-6)


def _decisions_agree(acc32, acc64, dH64, seed, traj, rel):
    """FP32 and FP64 accept decisions agree unless the FP64 decision is
    within `rel` of its threshold (u against exp(-dH))."""
    if acc32 == acc64:
        return True
    u = orc.stream_doubles(seed, 7, 0, traj, 0, 0, 0, 1)[0]
    return abs(u - np.exp(-dH64)) <= rel * np.exp(-dH64)


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("nc,beta", [(1, 1.0), (2, 2.3), (3, 5.0)])
def test_fp32_hmc_trajectory(nc, beta, storage):
    """Two leapfrog trajectories (tau = 1, 10 steps) with FP32 links and FP64
    dH / action sums, against the FP64 oracle: |dH32 - dH64| < 2e-3 (1 +
    |dH64|), the accept decisions identical unless the oracle's is within 1 %
    of its threshold, and the final links within 1e-4."""
    if nc == 1 and storage == "f32_rows":
        pytest.skip("U(1) has no two-row form")
    nd, dims, seed = 4, (4, 4, 4, 4), 99
    U = orc.random_algebra(nd, nc, dims, 0.5, 40)
    g = klft.GaugeField(nd, nc, dims, storage).upload(U)
    h = klft.HMC(g, beta=beta, tau=1.0, nsteps=10, seed=seed)
    Uo = U.copy()
    worst = 0.0
    for traj in range(2):
        acc_o, dH_o, _ = orc.hmc_trajectory(Uo, nd, nc, dims, beta, 1.0, 10, seed, traj)
        dH, acc = h.trajectory(traj)
        worst = max(worst, abs(dH - dH_o) / (1.0 + abs(dH_o)))
        assert _decisions_agree(acc, acc_o, dH_o, seed, traj, 0.01), (traj, dH, dH_o)
        if acc != acc_o:
            return
    report(f"HMC dH nc={nc} {storage} (relative to 1 + |dH|)", worst, 2e-3)
    report(f"HMC links after 2 trajectories nc={nc} {storage}", float(np.abs(g.download() - Uo).max()), 1e-4)


# FP32 Metropolis decision by decision.  The FP32 kernels draw 24-bit
# uniforms, one stream word each; orc.metropolis_sweep_narrow runs the same
# draws through the reference's algorithm in FP64 (Metropolis.hpp:72-113,
# SUN.hpp:163-251), so the two chains take the same accept decisions except
# where FP32 rounding of dS moves a decision across its threshold, which
# changes that link and, through its staples, its neighbours' later hits.
# Bounds: the accept counts of each sweep within 0.5 % (at least 2), links
# that agree to 1e-4 at least 97 % of all links, and on those the largest
# difference below 5e-6 (FP32 rounding and re-unitarisation; measured: identical
# counts, every link within 7e-7).
MET_CASES = [(4, 3, (6, 6, 6, 6), 5.5, 0.2, False), (4, 3, (6, 6, 6, 6), 5.5, 0.2, True),
             (4, 2, (8, 8, 8, 8), 2.3, 0.3, False), (4, 1, (8, 8, 8, 8), 2.0, 0.5, False),
             (3, 3, (8, 6, 8), 5.0, 0.2, False)]


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("nd,nc,dims,beta,delta,sym", MET_CASES)
def test_fp32_metropolis_tracks_oracle_decisions(storage, nd, nc, dims, beta, delta, sym):
    U = orc.random_algebra(nd, nc, dims, 0.5, 314)
    g = klft.GaugeField(nd, nc, dims, storage).upload(U)
    if sym:
        g.set_proposal("symmetric")
    p = klft.MetropolisParams(nHits=8, beta=beta, delta=delta, seed=2718)
    Uo = U.copy()
    counts = []
    for s in range(2):
        n_g = klft.sweep_Metropolis(g, p, s, with_count=True)[1]
        n_o = orc.metropolis_sweep_narrow(Uo, nd, nc, dims, 8, beta, delta, 2718, s, sym=sym)[1]
        counts.append((n_g, n_o))
    d = np.abs(g.download() - Uo).reshape(-1, nc * nc).max(axis=1)
    agree = d < 1e-4
    frac = float(agree.mean())
    err = float(d[agree].max())
    print(f"FP32 Metropolis {storage} nd={nd} nc={nc} sym={sym}: counts (GPU, oracle) {counts}, "
          f"links agreeing {frac:.4%}, max |dU| on them {err:.2e}")
    for n_g, n_o in counts:
        assert abs(n_g - n_o) <= max(2, 0.005 * n_o), counts
    assert frac >= 0.97, frac
    assert err < 5e-6, err
