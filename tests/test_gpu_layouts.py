"""GPU parity under the optional tiled site order (KLFT_TILES=1).

The site order inside a sublattice is fixed when the library first builds a
lattice geometry, so each case runs in a child process with the variable set.
The child uploads an oracle field, runs the plaquette, three Wilson loops,
one Metropolis sweep (every sweep variant) and one HMC force, and writes the
results; the parent
compares them with the oracle at the FP64 bars of test_gpu_parity.py.
"""
import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import klft_amd as klft
nd, nc, dims, variant, path = int(sys.argv[2]), int(sys.argv[3]), json.loads(sys.argv[4]), sys.argv[5], sys.argv[6]
import os
os.environ["KLFT_VARIANT"] = "metropolis=" + variant
U = np.load(path + ".U.npy")
g = klft.GaugeField(nd, nc, dims, "f64").upload(U)
plaq = klft.GaugePlaquette(g)
wl = [w for (_, _, _, _, w) in klft.WilsonLoop_mu_nu(g, 0, 1, [(1, 1), (2, 3), (3, 5)])]
p = klft.MetropolisParams(nHits=5, beta=5.5 if nc == 3 else 2.2, delta=0.3, seed=31)
acc, n = klft.sweep_Metropolis(g, p, 0, with_count=True)
h = klft.HMC(g, beta=5.5, tau=1.0, nsteps=4, seed=5)
F = h.force()
np.save(path + ".Uout.npy", g.download())
np.save(path + ".F.npy", F)
json.dump({"plaq": plaq, "n": n, "wl": wl}, open(path + ".json", "w"))
"""


@pytest.mark.parametrize("variant", ["0", "7", "11", "12", "13"])
@pytest.mark.parametrize("nd,nc,dims", [(4, 3, (4, 8, 4, 8)), (3, 2, (4, 6, 8)), (4, 1, (8, 4, 4, 4))])
def test_tiled_site_order_matches_oracle(tmp_path, nd, nc, dims, variant):
    U = orc.random_algebra(nd, nc, dims, 0.4, 23)
    base = str(tmp_path / "case")
    np.save(base + ".U.npy", U)
    env = dict(os.environ, KLFT_TILES="1")
    r = subprocess.run([sys.executable, "-c", CHILD, str(ROOT), str(nd), str(nc), json.dumps(list(dims)), variant,
                        base], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = json.load(open(base + ".json"))
    assert abs(got["plaq"] - orc.plaquette(U, nd, nc, dims)) < 1e-12
    ref = orc.wilson_mu_nu(U, nd, nc, dims, 0, 1, [(1, 1), (2, 3), (3, 5)])
    for w, r in zip(got["wl"], ref):
        assert abs(w - r) < 1e-12
    Uo = U.copy()
    beta = 5.5 if nc == 3 else 2.2
    _, n_o = orc.metropolis_sweep(Uo, nd, nc, dims, 5, beta, 0.3, 31, 0)
    assert got["n"] == n_o
    Ug = np.load(base + ".Uout.npy")
    np.testing.assert_allclose(Ug, Uo, rtol=0, atol=1e-11)
    # the force of the field the GPU holds, so the sweep's rounding is not counted twice
    F = orc.force(np.ascontiguousarray(Ug), nd, nc, dims, 5.5)
    np.testing.assert_allclose(np.load(base + ".F.npy"), F, rtol=0, atol=1e-12)
