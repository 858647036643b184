"""Ensemble-level agreement of FP32 and FP32 two-row links with FP64.

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

klft = pytest.importorskip("klft_amd")

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
