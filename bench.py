"""Benchmark of the klft gauge-update hot path on one MI355X.

Workload (the headline metric's configuration): SU(3), 4D lattice 32^4,
checkerboarded Metropolis with nHits = 10, beta = 6.0, delta = 0.1, from a
seeded exp(i 0.5 H) start.  A "step" is one full sweep (16 sublattice passes,
4 * 32^4 link updates).  `value` is the FP64 (plain method) link-updates/s;
the FP32 and FP32 two-row variants, HMC trajectories/s (tau = 1, 20 leapfrog
steps) and measurement-kernel times are reported beside it.

roofline: the dominant kernel is the Metropolis pass.  Algorithmic bytes per
pass = 36 * (V / 16) * B_U (32 of the 64 (parity class, direction) link groups
read, 4 read and written; B_U = 144 / 72 / 48 bytes for FP64 / FP32 / FP32
two-row SU(3) links), divided by the pass's average duration from HIP events
on the library stream.  traffic: HBM bytes per pass from the committed
rocprofv3 summary (profiles/), when present.

cpu_baseline: the FP64 CPU oracle (the build's own restatement of the
reference's OpenMP path, not the reference) runs one sweep of the same
lattice with all host cores allotted (at most 16).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0
METRIC = "SU(3) link-updates/s + HMC traj/s at 32^4; % HBM roofline; FP32 vs FP64"
BYTES_PER_LINK = {"f64": 144, "f32": 72, "f32_rows": 48}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--nhits", type=int, default=10)
    ap.add_argument("--beta", type=float, default=6.0)
    ap.add_argument("--delta", type=float, default=0.1)
    ap.add_argument("--hmc-traj", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="Metropolis lines only")
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy the FP64 link field 3x first (known-size kernel that scales PMC byte counters)")
    ap.add_argument("--only", default=None, help="comma list of precisions to run (f64,f32,f32_rows)")
    return ap.parse_args()


def sync_time(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def metropolis_leg(klft, st, a, dims):
    V = dims[0] * dims[1] * dims[2] * dims[3]
    g = klft.GaugeField(4, 3, dims, st).set_random_algebra(0.5, 2024)
    p = klft.MetropolisParams(nHits=a.nhits, beta=a.beta, delta=a.delta, seed=4242)
    if a.warmup:
        klft.metropolis_run(g, p, 0, a.warmup, rates=False)
    klft.profile(True)
    t = sync_time(lambda: klft.metropolis_run(g, p, a.warmup, a.steps, rates=False))
    ms, n = klft.profile_read("metropolis_pass")
    klft.profile(False)
    acc = klft.sweep_Metropolis(g, p, a.warmup + a.steps)
    plaq = klft.GaugePlaquette(g)
    lu = 4 * V * a.steps / t
    pass_us = 1e3 * ms / max(n, 1)
    alg = 36 * (V // 16) * BYTES_PER_LINK[st]
    return g, {"link_updates_per_s": lu, "ms_per_sweep": 1e3 * t / a.steps, "pass_us": pass_us,
               "pass_launches": n, "alg_bytes_per_pass": alg,
               "achieved_GBs": alg / (pass_us * 1e-6) / 1e9, "acceptance": acc, "plaquette": plaq}


def hmc_leg(klft, st, a, dims):
    g = klft.GaugeField(4, 3, dims, st).set_random_algebra(0.5, 77)
    h = klft.HMC(g, beta=a.beta, tau=1.0, nsteps=20, seed=99)
    h.trajectory(0)
    klft.profile(True)
    res = []
    t = sync_time(lambda: res.extend(h.trajectory(i + 1) for i in range(a.hmc_traj)))
    fms, fn = klft.profile_read("force")
    ums, un = klft.profile_read("update_gauge")
    klft.profile(False)
    return {"traj_per_s": a.hmc_traj / t, "ms_per_traj": 1e3 * t / a.hmc_traj,
            "dH": [r[0] for r in res], "accepted": [r[1] for r in res],
            "force_us": 1e3 * fms / max(fn, 1), "update_gauge_us": 1e3 * ums / max(un, 1)}


def measure_leg(klft, g, st, dims):
    V = dims[0] * dims[1] * dims[2] * dims[3]
    klft.GaugePlaquette(g)
    klft.profile(True)
    for _ in range(5):
        klft.GaugePlaquette(g)
    pms, pn = klft.profile_read("plaquette")
    pairs = [(L, T) for L in range(1, 9) for T in range(1, 9)]
    klft.WilsonLoop_temporal(g, pairs)
    wms, wn = klft.profile_read("wilson")
    klft.profile(False)
    b = BYTES_PER_LINK[st]
    return {"plaquette_us": 1e3 * pms / max(pn, 1),
            "plaquette_alg_GBs": 4 * V * b / (1e-3 * pms / max(pn, 1)) / 1e9,
            "wilson_temporal_64pairs_ms": wms / max(wn, 1)}


def cpu_leg(dims, a):
    """One FP64 sweep of the same lattice on the host cores with the oracle."""
    import numpy as np  # noqa: F401
    sys.path.insert(0, str(ROOT / "tests"))
    cores = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    cores = int(os.environ["OMP_NUM_THREADS"])
    import oracle as orc
    U = orc.random_algebra(4, 3, dims, 0.5, 2024)
    t = sync_time(lambda: orc.metropolis_sweep(U, 4, 3, dims, a.nhits, a.beta, a.delta, 4242, 0))
    V = dims[0] * dims[1] * dims[2] * dims[3]
    return {"value": 4 * V / t, "unit": "link-updates/s", "cores": cores, "kind": "port",
            "sample": f"one FP64 SU(3) Metropolis sweep of {dims[0]}^4 (nHits={a.nhits}), "
                      f"{t:.1f} s, OpenMP oracle"}


def traffic_from_profiles():
    f = ROOT / "profiles" / "metropolis_pass_su3_32_traffic.json"
    if f.exists():
        try:
            return json.loads(f.read_text())
        except Exception:
            return None
    return None


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (every box has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    import klft_amd as klft

    dims = (a.L, a.L, a.L, a.L)
    V = a.L ** 4
    legs, extra = {}, {}
    g64 = None
    if a.copy_probe:
        src = klft.GaugeField(4, 3, dims, "f64").set_identity()
        dst = klft.GaugeField(4, 3, dims, "f64")
        for _ in range(3):
            dst.copy_from(src)
        print(f"[bench] copy probe: 3 x {4 * V * 144} bytes read + written", file=sys.stderr, flush=True)
        src.close()
        dst.close()
    precisions = a.only.split(",") if a.only else ["f64", "f32", "f32_rows"]
    for st in precisions:
        g, legs[st] = metropolis_leg(klft, st, a, dims)
        print(f"[bench] metropolis {st}: {legs[st]['link_updates_per_s']:.3e} lu/s, "
              f"pass {legs[st]['pass_us']:.1f} us", file=sys.stderr, flush=True)
        if not a.no_extra:
            legs[st].update(measure_leg(klft, g, st, dims))
        if st == "f64":
            g64 = g
        else:
            g.close()
    if not a.no_extra:
        for st in precisions:
            extra[st] = hmc_leg(klft, st, a, dims)
            print(f"[bench] hmc {st}: {extra[st]['traj_per_s']:.2f} traj/s", file=sys.stderr, flush=True)
    del g64

    main_leg = legs["f64"] if "f64" in legs else legs[precisions[0]]
    traffic = None
    tj = traffic_from_profiles()
    if tj and "f64" in tj:
        traffic = tj["f64"].get("bytes_per_pass")
    roofline = {"bound": "hbm", "kernel": "metropolis_pass (FP64 SU(3))",
                "achieved": round(main_leg["achieved_GBs"], 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(main_leg["achieved_GBs"] / PEAK_HBM_GBS, 4), "traffic": traffic}
    cpu = None
    if not a.no_cpu:
        cpu = cpu_leg(dims, a)
        print(f"[bench] cpu oracle: {cpu['value']:.3e} lu/s on {cpu['cores']} cores",
              file=sys.stderr, flush=True)
    out = {
        "metric": METRIC, "value": main_leg["link_updates_per_s"], "unit": "link-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": main_leg["ms_per_sweep"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded exp(i 0.5 H) start)",
        "config": {"workload": f"SU(3) Metropolis sweep, 4D {a.L}^4, nHits={a.nhits}, beta={a.beta}, "
                               f"delta={a.delta}", "lattice": list(dims), "group": "SU(3)",
                   "link_updates_per_step": 4 * V, "precision": "FP64 plain method"},
        "roofline": roofline, "cpu_baseline": cpu,
        "variants": {st: {k: (round(v, 6) if isinstance(v, float) else v) for k, v in d.items()}
                     for st, d in legs.items()},
        "hmc": {st: {k: (round(v, 6) if isinstance(v, float) else v) for k, v in d.items()}
                for st, d in extra.items()},
        "fp32_over_fp64": {st: legs[st]["link_updates_per_s"] / main_leg["link_updates_per_s"]
                           for st in legs},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
