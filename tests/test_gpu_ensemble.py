"""Ensemble-level checks on the GPU.

* HMC reversibility (Integrator.hpp:26-53): integrate, flip the momenta,
  integrate again -> the links and momenta return.
* <exp(-dH)> = 1 (HMC.hpp:48-71, Monomial.hpp:44-46) on SU(2) 16^4 at
  beta = 2.3, FP64 and FP32 links.
* SU(2) HMC and SU(2) Metropolis (whose randSUN<2> proposal is symmetric:
  r and r^dagger = (alpha, -n) are equally likely) sample the same
  exp(-S): equal mean plaquettes at beta = 2.3 within errors.
* FP32 / FP32 two-row SU(3) Metropolis ensembles agree with FP64.

Three independent SU(3) Metropolis chains (different seeds) on 12^4 at
beta = 3.0, nHits = 10, from Haar-random links: after 300 sweeps the
plaquette is recorded every 10 sweeps for 1600 sweeps.  (At beta = 6 the
reference's delta-independent SU(3) proposal is accepted about 1 % of the
time and a chain is still far from equilibrium after hundreds of sweeps.)
The FP32 averages must agree with the FP64 one within three combined
standard errors (the records are binned in 16 blocks of 10 against
autocorrelation).  The chains are keyed, so the outcome is reproducible run
to run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import klft_amd as klft  # noqa: E402

DIMS = (12, 12, 12, 12)
THERM, EVERY, NMEAS = 300, 10, 160


def chain(storage, seed):
    g = klft.GaugeField(4, 3, DIMS, storage).set_haar(seed)
    p = klft.MetropolisParams(nHits=10, beta=3.0, delta=0.1, seed=seed)
    klft.metropolis_run(g, p, 0, THERM, rates=False)
    vals = []
    for k in range(NMEAS):
        klft.metropolis_run(g, p, THERM + EVERY * k, EVERY, rates=False)
        vals.append(klft.GaugePlaquette(g))
    return np.array(vals)


def mean_err(v, nbin=16):
    b = v.reshape(nbin, -1).mean(axis=1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nbin)


def test_su3_fp32_ensembles_match_fp64():
    m64, e64 = mean_err(chain("f64", 11))
    for storage, seed in [("f32", 12), ("f32_rows", 13)]:
        m, e = mean_err(chain(storage, seed))
        print(f"<P> f64 {m64:.6f}({e64 * 1e6:.0f}e-6)  {storage} {m:.6f}({e * 1e6:.0f}e-6)  "
              f"diff {(m - m64) / np.hypot(e, e64):+.2f} sigma")
        assert abs(m - m64) < 3 * np.hypot(e, e64)


# ---------------------------------------------------------------------------
# HMC reversibility
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("storage,bound", [("f64", 1e-12), ("f32", 2e-5), ("f32_rows", 2e-5)])
@pytest.mark.parametrize("nc,beta,dims", [(2, 2.3, (16, 16, 16, 16)), (3, 5.5, (8, 8, 8, 8))])
def test_hmc_reversibility(nc, beta, dims, storage, bound):
    """klft_hmc_integrate (tau = 1, 20 steps), P -> -P, integrate again: the
    links return to the start within `bound` per entry and the momenta to
    -P0 within 100 * bound (the final restoreGauge of each integration and
    rounding are the only breaks of exact reversibility)."""
    import oracle as orc
    U0 = orc.random_algebra(4, nc, dims, 0.5, 515)
    g = klft.GaugeField(4, nc, dims, storage).upload(U0)
    if storage != "f64":
        U0 = g.download()  # the FP32 start as stored
    h = klft.HMC(g, beta=beta, tau=1.0, nsteps=20, seed=8)
    h.refresh_momenta(0)
    P0 = h.momenta()
    h.integrate()
    h.set_momenta(-h.momenta())
    h.integrate()
    du = float(np.abs(g.download() - U0).max())
    dp = float(np.abs(h.momenta() + P0).max())
    print(f"reversibility nc={nc} {storage}: max |dU| {du:.2e}, max |dP| {dp:.2e}")
    assert du < bound and dp < 100 * bound


# ---------------------------------------------------------------------------
# <exp(-dH)> = 1 and HMC vs Metropolis, SU(2) 16^4 at beta = 2.3
# ---------------------------------------------------------------------------
SU2_DIMS, SU2_BETA = (16, 16, 16, 16), 2.3
# With the reference's generators sigma_a (Tr T_a T_b = 2 delta_ab) a
# leapfrog step moves twice as far as with sigma_a / 2: at the configured 20
# steps |dH| is about 10 and <exp(-dH)> has no usable error bar, so these
# ensembles take 80 steps (acceptance about 90 %).
NSTEPS, THERM_T, NTRAJ = 80, 100, 300


def hmc_chain(storage, seed):
    g = klft.GaugeField(4, 2, SU2_DIMS, storage).set_random_sun(0.5, 2 * seed)
    h = klft.HMC(g, beta=SU2_BETA, tau=1.0, nsteps=NSTEPS, seed=seed)
    ex, plaq, acc = [], [], []
    for i in range(THERM_T + NTRAJ):
        dH, ok = h.trajectory(i)
        if i >= THERM_T:
            ex.append(np.exp(-dH))
            plaq.append(klft.GaugePlaquette(g))
            acc.append(ok)
    return np.array(ex), np.array(plaq), float(np.mean(acc))


@pytest.fixture(scope="module")
def su2_hmc():
    return {st: hmc_chain(st, 4321 + k) for k, st in enumerate(["f64", "f32"])}


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_exp_minus_dH_is_one(su2_hmc, storage):
    ex, _, acc = su2_hmc[storage]
    m, e = mean_err(ex, nbin=10)
    print(f"SU(2) 16^4 {storage}: <exp(-dH)> = {m:.4f} +- {e:.4f}, acceptance {acc:.2f}")
    assert abs(m - 1.0) < 4 * e


def test_su2_hmc_and_metropolis_plaquettes_agree(su2_hmc):
    """Metropolis (nHits = 10, delta = 0.3, cold start, 300 thermalisation
    sweeps, 200 measurements every 5 sweeps) against the FP64 HMC chain:
    means within 4 combined standard errors (10 bins each)."""
    _, plaq_h, _ = su2_hmc["f64"]
    g = klft.GaugeField(4, 2, SU2_DIMS, "f64").set_identity()
    p = klft.MetropolisParams(nHits=10, beta=SU2_BETA, delta=0.3, seed=2718)
    klft.metropolis_run(g, p, 0, 300, rates=False)
    vals = []
    for k in range(200):
        klft.metropolis_run(g, p, 300 + 5 * k, 5, rates=False)
        vals.append(klft.GaugePlaquette(g))
    mh, eh = mean_err(plaq_h, nbin=10)
    mm, em = mean_err(np.array(vals), nbin=10)
    z = (mh - mm) / np.hypot(eh, em)
    print(f"SU(2) 16^4 beta 2.3: <P> HMC {mh:.5f}({eh * 1e5:.0f}e-5)  Metropolis {mm:.5f}({em * 1e5:.0f}e-5)  "
          f"diff {z:+.2f} sigma")
    assert abs(z) < 4


def test_su2_fp32_hmc_plaquette_matches_fp64(su2_hmc):
    _, p64, _ = su2_hmc["f64"]
    _, p32, _ = su2_hmc["f32"]
    m64, e64 = mean_err(p64, nbin=10)
    m32, e32 = mean_err(p32, nbin=10)
    z = (m32 - m64) / np.hypot(e32, e64)
    print(f"SU(2) 16^4 HMC <P>: f64 {m64:.5f}({e64 * 1e5:.0f}e-5)  f32 {m32:.5f}({e32 * 1e5:.0f}e-5)  {z:+.2f} sigma")
    assert abs(z) < 4
