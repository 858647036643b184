"""The opt-in symmetric SU(3) Metropolis proposal (klft_field_set_proposal)
against the oracle and against exactly known ensembles.

* FP64 sweeps with the symmetric law equal the oracle's decision by decision
  (2D, 3D, 4D; the reference-law default is untouched by the setting on
  other fields).
* 2D SU(3): the plaquette of the chain matches the closed form
  sum_k det[I_{k+i-j}(beta/3)] (FP64 and FP32 links), which the reference's
  asymmetric randSUN<3> misses (SUN.hpp:192-251).
* 4D SU(3) 8^4: symmetric Metropolis and HMC (HMC.hpp:48-71) sample the same
  Wilson action, so their mean plaquettes agree within statistical errors.
"""
import numpy as np
import pytest

import oracle as orc
from test_oracle import binned_mean, sun_plaquette_2d

pytestmark = pytest.mark.gpu

import klft_amd as klft  # noqa: E402


@pytest.mark.parametrize("nd,dims", [(2, (8, 6)), (3, (4, 6, 4)), (4, (4, 4, 6, 4))])
def test_symmetric_proposal_matches_oracle(nd, dims):
    U = orc.random_algebra(nd, 3, dims, 0.4, 808)
    g = klft.GaugeField(nd, 3, dims, "f64").upload(U).set_proposal("symmetric")
    p = klft.MetropolisParams(nHits=6, beta=5.5, delta=0.35, seed=19)
    Uo = U.copy()
    for sweep in range(3):
        _, n = klft.sweep_Metropolis(g, p, sweep, with_count=True)
        _, n_o = orc.metropolis_sweep_sym(Uo, nd, 3, dims, 6, 5.5, 0.35, 19, sweep)
        assert n == n_o, (sweep, n, n_o)
        assert 0 < n < 6 * nd * int(np.prod(dims))
    np.testing.assert_allclose(g.download(), Uo, rtol=0, atol=1e-11)
    # the default law on another field still follows the reference's randSUN
    h = klft.GaugeField(nd, 3, dims, "f64").upload(U)
    Ur = U.copy()
    _, n = klft.sweep_Metropolis(h, p, 0, with_count=True)
    _, n_r = orc.metropolis_sweep(Ur, nd, 3, dims, 6, 5.5, 0.35, 19, 0)
    assert n == n_r
    np.testing.assert_allclose(h.download(), Ur, rtol=0, atol=1e-11)


def _chain_plaquettes(g, p, therm, nmeas, every):
    klft.metropolis_run(g, p, 0, therm, rates=False)
    out = []
    for i in range(nmeas):
        klft.metropolis_run(g, p, therm + i * every, every, rates=False)
        out.append(klft.GaugePlaquette(g))
    return np.array(out)


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_symmetric_2d_su3_closed_form(storage):
    beta = 4.0
    g = klft.GaugeField(2, 3, (32, 32), storage).set_identity().set_proposal("symmetric")
    p = klft.MetropolisParams(nHits=10, beta=beta, delta=0.6, seed=2718)
    P = _chain_plaquettes(g, p, 200, 400, 2)
    m, e = binned_mean(P, 20)
    exact = sun_plaquette_2d(3, beta)
    print(f"2D SU(3) beta={beta} {storage} symmetric: <P> = {m:.5f} +- {e:.5f}, closed form {exact:.5f}, "
          f"{(m - exact) / e:+.2f} sigma")
    assert abs(m - exact) < 4 * e + 1e-3


def test_symmetric_metropolis_equals_hmc_su3_4d():
    dims, beta = (8, 8, 8, 8), 5.0
    gm = klft.GaugeField(4, 3, dims, "f64").set_identity().set_proposal("symmetric")
    p = klft.MetropolisParams(nHits=10, beta=beta, delta=0.3, seed=99)
    Pm = _chain_plaquettes(gm, p, 300, 400, 3)
    # HMC from a cold start rejects everything (the leapfrog error is large
    # far from equilibrium), so its chain starts from a field thermalised by
    # an independent symmetric Metropolis chain
    gh = klft.GaugeField(4, 3, dims, "f64").set_identity().set_proposal("symmetric")
    klft.metropolis_run(gh, klft.MetropolisParams(nHits=10, beta=beta, delta=0.3, seed=7), 0, 300, rates=False)
    h = klft.HMC(gh, beta=beta, tau=1.0, nsteps=40, seed=5)
    Ph, acc = [], []
    for i in range(450):
        _, ok = h.trajectory(i)
        acc.append(ok)
        if i >= 50:
            Ph.append(klft.GaugePlaquette(gh))
    mm, em = binned_mean(Pm, 20)
    mh, eh = binned_mean(np.array(Ph), 20)
    z = (mm - mh) / np.hypot(em, eh)
    print(f"4D SU(3) 8^4 beta={beta}: symmetric Metropolis <P> = {mm:.5f} +- {em:.5f}, "
          f"HMC <P> = {mh:.5f} +- {eh:.5f} (acceptance {np.mean(acc[50:]):.2f}), {z:+.2f} sigma")
    assert np.mean(acc[50:]) > 0.3
    assert abs(mm - mh) < 4 * np.hypot(em, eh) + 1e-3
