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
on the library stream.  traffic: HBM bytes per pass and the VALU busy share
from rocprofv3 runs of this bench command on this build (profiles/); bound
is the roof the pass is closer to on those measurements.

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
    ap.add_argument("--hmc-traj", type=int, default=10)
    ap.add_argument("--hmc-warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="Metropolis lines only")
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy the FP64 link field 3x first (known-size kernel that scales PMC byte counters)")
    ap.add_argument("--only", default=None, help="comma list of precisions to run (f64,f32,f32_rows)")
    ap.add_argument("--config", default=None,
                    help="per-configuration reports instead of the headline line: c1,...,c5 or all")
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
    # the timed sweeps run as a caller runs them, without per-launch events
    t = sync_time(lambda: klft.metropolis_run(g, p, a.warmup, a.steps, rates=False))
    # pass durations: HIP events around every pass launch, in sweeps of their own
    nprof = max(2, min(a.steps, 5))
    klft.profile(True)
    klft.metropolis_run(g, p, a.warmup + a.steps, nprof, rates=False)
    ms, n = klft.profile_read("metropolis_pass")
    # launches that run two independent passes at once (when both fit)
    ms2, n2 = klft.profile_read("metropolis_pass_pair")
    klft.profile(False)
    acc = klft.sweep_Metropolis(g, p, a.warmup + a.steps + nprof)
    plaq = klft.GaugePlaquette(g)
    lu = 4 * V * a.steps / t
    pass_us = 1e3 * ms / max(n, 1)
    sweep_kernel_us = 1e3 * (ms + ms2) / nprof
    alg = 36 * (V // 16) * BYTES_PER_LINK[st]
    out = {"link_updates_per_s": lu, "ms_per_sweep": 1e3 * t / a.steps, "pass_us": pass_us,
           "pass_launches": n, "alg_bytes_per_pass": alg,
           "achieved_GBs": 16 * alg / (sweep_kernel_us * 1e-6) / 1e9, "acceptance": acc, "plaquette": plaq}
    if n2:
        out.update({"pair_us": 1e3 * ms2 / n2, "pair_launches": n2, "sweep_kernel_us": sweep_kernel_us})
    return g, out


def hmc_leg(klft, st, a, dims):
    """HMC trajectories (tau = 1, 20 leapfrog steps = 21 fused MD steps) at
    beta.  Before them, a.hmc_warmup untimed trajectories with 80 leapfrog
    steps (a step a quarter as long, so most are accepted) carry the field
    from the exp(i 0.5 H) start towards equilibrium; at 20 steps on 32^4 the
    energy error of a trajectory is O(100) even there, so the timed
    trajectories' acceptance is reported but is not what they measure."""
    import numpy as np
    g = klft.GaugeField(4, 3, dims, st).set_random_algebra(0.5, 77)
    hw = klft.HMC(g, beta=a.beta, tau=1.0, nsteps=80, seed=98)
    wres = [hw.trajectory(i) for i in range(a.hmc_warmup)]
    hw.close()
    plaq_warm = klft.GaugePlaquette(g)
    h = klft.HMC(g, beta=a.beta, tau=1.0, nsteps=20, seed=99)
    res = []
    t = sync_time(lambda: res.extend(h.trajectory(a.hmc_warmup + i) for i in range(a.hmc_traj)))
    # kernel durations from two further trajectories with HIP events on
    klft.profile(True)
    for i in range(2):
        h.trajectory(a.hmc_warmup + a.hmc_traj + i)
    mms, mn = klft.profile_read("md_step")
    fms, fn = klft.profile_read("force")
    klft.profile(False)
    dH = np.array([r[0] for r in res])
    out = {"traj_per_s": a.hmc_traj / t, "ms_per_traj": 1e3 * t / a.hmc_traj, "trajectories": a.hmc_traj,
           "warmup_trajectories": a.hmc_warmup, "warmup_nsteps": 80,
           "warmup_acceptance": float(np.mean([r[1] for r in wres])) if wres else None,
           "plaquette_after_warmup": plaq_warm, "acceptance": float(np.mean([r[1] for r in res])),
           "mean_dH": float(dH.mean()), "mean_abs_dH": float(np.abs(dH).mean()),
           "plaquette": klft.GaugePlaquette(g),
           "md_step_us": 1e3 * mms / max(mn, 1), "final_half_step_force_us": 1e3 * fms / max(fn, 1)}
    h.close()
    g.close()
    return out


def batched_leg(klft, a):
    """The batched engine at the headline's volume: 16 independent SU(3) 16^4
    chains (1,048,576 sites, as 32^4), and at four times it (64 chains), with
    FP32 links + FP64 sums (and FP64 links for comparison), one launch
    per sublattice pass, or per pair of independent passes, over all
    chains, sweeps replayed from one captured
    HIP graph; and the PTBC replica set (16 x SU(3) 16^4, 2^3 defect at
    x0 = 0, c_r = 1 - r/15, nHits = 8, delta = 0.2) stepped with its swap
    pass.  Algorithmic bytes per sweep = 36 * V_total * B_U (16 passes of
    36 (V/16) B_U)."""
    import numpy as np
    out = {}
    # 16 chains: the headline's volume; 64 chains: four times it, two blocks
    # of a pass per CU, the engine with the GPU full
    for nch in (16, 64):
        chains = [(3, (16, 16, 16, 16), a.beta)] * nch
        vtot = nch * 16 ** 4
        for st in ["f32", "f32_rows", "f64"]:
            b = klft.Batch(chains, st)
            b.set_identity()
            b.run(a.nhits, a.delta, 4242, 0, 3)
            nsw = 20 if nch == 16 else 8
            t = sync_time(lambda: b.run(a.nhits, a.delta, 4242, 3, nsw))
            alg = 36 * vtot * BYTES_PER_LINK[st]
            out[f"chains{nch}_su3_16^4_{st}"] = {
                "sweeps_per_s": nsw / t, "link_updates_per_s": 4 * vtot * nsw / t, "ms_per_sweep": 1e3 * t / nsw,
                "alg_bytes_per_sweep": alg, "achieved_GBs": alg * nsw / t / 1e9,
                "frac_hbm": alg * nsw / t / 1e9 / PEAK_HBM_GBS}
            b.close()
            print(f"[bench] batched {nch} x 16^4 {st}: {4 * vtot * nsw / t:.3e} lu/s", file=sys.stderr, flush=True)
    n = 16
    vtot = n * 16 ** 4
    p = klft.PTBC(n, 3, (16, 16, 16, 16), 6.0, 2, [0, 0, 0], 1.0 - np.arange(n) / 15.0, "f32")
    p.set_identity()
    for s in range(5):
        p.step(8, 0.2, 11, s)
    nst = 30
    sw = np.zeros(n - 1)

    def run():
        for s in range(nst):
            sw[:] += p.step(8, 0.2, 11, 5 + s)
    t = sync_time(run)
    alg = 36 * vtot * BYTES_PER_LINK["f32"]
    out["ptbc16_su3_16^4_f32"] = {
        "steps_per_s": nst / t, "replica_link_updates_per_s": 4 * vtot * nst / t, "ms_per_step": 1e3 * t / nst,
        "alg_bytes_per_step": alg, "achieved_GBs": alg * nst / t / 1e9, "frac_hbm": alg * nst / t / 1e9 / PEAK_HBM_GBS,
        "swap_acceptance_min_max": [float(sw.min() / nst), float(sw.max() / nst)]}
    p.close()
    return out


def measure_leg(klft, g, st, dims):
    V = dims[0] * dims[1] * dims[2] * dims[3]
    pairs = [(L, T) for L in range(1, 9) for T in range(1, 9)]
    # one untimed call of each first (the first launch of a kernel also
    # loads its code object)
    klft.GaugePlaquette(g)
    klft.WilsonLoop_temporal(g, pairs)
    klft.profile(True)
    for _ in range(5):
        klft.GaugePlaquette(g)
    pms, pn = klft.profile_read("plaquette")
    for _ in range(3):
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


def pass_profile():
    """Counters of the Metropolis pass kernel measured with rocprofv3 on this
    build (profiles/metropolis_pass_su3_32.json, made from the same bench
    command): HBM bytes per pass from FETCH_SIZE / WRITE_SIZE scaled by the
    copy-kernel factor, VALUBusy, and the kernel-trace average duration."""
    f = ROOT / "profiles" / "metropolis_pass_su3_32.json"
    if f.exists():
        try:
            return json.loads(f.read_text())
        except Exception:
            return None
    return None


# ---------------------------------------------------------------------------
# per-configuration reports (--config c1..c5): one JSON line each
# ---------------------------------------------------------------------------
C1_YAML = """MetropolisParams:
  Ndims: 4
  Nd: 4
  Nc: 1
  L0: 8
  L1: 8
  L2: 8
  L3: 8
  nHits: 10
  nSweep: {nsweep}
  seed: 32091
  beta: 2.0
  delta: 0.1
GaugeObservableParams:
  measurement_interval: 10
  measure_plaquette: true
  measure_wilson_loop_temporal: true
  measure_wilson_loop_mu_nu: true
  W_temp_L_T_pairs: [[2, 2], [3, 4], [4, 3], [4, 4]]
  W_mu_nu_pairs: [[0, 1], [1, 2], [3, 2]]
  W_Lmu_Lnu_pairs: [[2, 2], [3, 3], [4, 3]]
  plaquette_filename: "plaquette.out"
  W_temp_filename: "W_temp.out"
  W_mu_nu_filename: "W_mu_nu.out"
  write_to_file: true
"""


def config_c1(klft, a):
    """U(1) 8^4, the README input: the metropolis driver end to end (1000
    sweeps, 99 measurements, three files), timed, beside the oracle running the
    same keyed chain on the host cores."""
    import subprocess
    import tempfile
    nsweep = 1000
    with tempfile.TemporaryDirectory() as d:
        Path(d, "input.yaml").write_text(C1_YAML.format(nsweep=nsweep))
        t = sync_time(lambda: subprocess.run([str(ROOT / "klft_amd" / "metropolis"), "-f", "input.yaml"], cwd=d,
                                             check=True, capture_output=True))
        rows = Path(d, "plaquette.out").read_text().splitlines()[1:]
    plaq = [float(r.split(", ")[1]) for r in rows]
    out = {"config": "C1 U(1) 8^4 Metropolis, README input, 1000 sweeps", "gpu_driver_s": t,
           "sweeps_per_s": nsweep / t, "link_updates_per_s": 4 * 8 ** 4 * nsweep / t,
           "measurements": len(plaq), "plaquette_last": plaq[-1], "plaquette_mean_2nd_half": float(sum(plaq[49:]) / len(plaq[49:]))}
    if not a.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle as orc
        dims = (8, 8, 8, 8)
        U = orc.identity(4, 1, dims)
        cpu_plaq = []

        def run():
            for s in range(nsweep):
                orc.metropolis_sweep(U, 4, 1, dims, 10, 2.0, 0.1, 32091, s)
                if s % 10 == 0 and s:
                    cpu_plaq.append(orc.plaquette(U, 4, 1, dims))
                    orc.wilson_temporal(U, 4, 1, dims, [(2, 2), (3, 4), (4, 3), (4, 4)])
                    for mu, nu in [(0, 1), (1, 2), (3, 2)]:
                        orc.wilson_mu_nu(U, 4, 1, dims, mu, nu, [(2, 2), (3, 3), (4, 3)])
        tc = sync_time(run)
        out["cpu_baseline"] = {"value": 4 * 8 ** 4 * nsweep / tc, "unit": "link-updates/s",
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
                               "sample": f"the same 1000-sweep keyed chain with measurements, {tc:.1f} s"}
        out["max_abs_plaquette_diff_vs_cpu"] = max(abs(x - y) for x, y in zip(plaq, cpu_plaq))
    return out


def _hmc_chain(klft, st, nsteps, ntraj, therm):
    """One SU(2) 16^4 HMC chain (beta 2.3, tau 1, Haar hot start): rates,
    acceptance, binned plaquette and <exp(-dH)> after `therm` trajectories."""
    import numpy as np
    g = klft.GaugeField(4, 2, (16, 16, 16, 16), st).set_random_sun(0.5, 2 * 1234)
    h = klft.HMC(g, beta=2.3, tau=1.0, nsteps=nsteps, seed=1234)
    dH, acc, plaq = [], [], []
    t0 = time.perf_counter()
    for i in range(ntraj):
        d, ok = h.trajectory(i)
        dH.append(d)
        acc.append(ok)
        plaq.append(klft.GaugePlaquette(g))
    t = time.perf_counter() - t0
    e = np.exp(-np.array(dH[therm:]))
    p = np.array(plaq[therm:])
    nb = 10
    eb = np.array([b.mean() for b in np.array_split(e, nb)])
    pb = np.array([b.mean() for b in np.array_split(p, nb)])
    h.close()
    g.close()
    return {"traj_per_s": ntraj / t, "acceptance": float(np.mean(acc[therm:])),
            "plaquette": float(p.mean()), "plaquette_err": float(pb.std(ddof=1) / np.sqrt(nb)),
            "exp_minus_dH": float(e.mean()), "exp_minus_dH_err": float(eb.std(ddof=1) / np.sqrt(nb)),
            "mean_dH": float(np.mean(dH[therm:])), "mean_abs_dH": float(np.abs(dH[therm:]).mean())}


def config_c2(klft, a):
    """SU(2) 16^4 HMC, beta 2.3, Haar hot start, tau 1: the configured 20
    steps (21 gauge updates, 200 trajectories) with FP64 and with FP32 links
    (+ FP64 dH and action sums), and 80 steps (400 trajectories), where the
    step is small enough for <exp(-dH)> to be measured with a useful error.
    The generators are the reference's sigma_a (Tr T_a T_b = 2 delta_ab), so
    a step moves twice as far as with sigma_a / 2."""
    import numpy as np
    res = {"config": "C2 SU(2) 16^4 HMC beta=2.3 tau=1, n=20 (200 traj) and n=80 (400 traj)"}
    for st in ["f64", "f32"]:
        res[st] = _hmc_chain(klft, st, 20, 200, 50)
    d = res["f32"]["plaquette"] - res["f64"]["plaquette"]
    res["plaquette_fp32_minus_fp64_sigma"] = d / np.hypot(res["f32"]["plaquette_err"], res["f64"]["plaquette_err"])
    for st in ["f64", "f32"]:
        res[st + "_n80"] = _hmc_chain(klft, st, 80, 400, 100)
    return res


def config_c3(klft, a):
    """SU(3) 32^4 fixed fields exp(i eps H), eps in {0.1, 0.5, 1.0}: plaquette,
    staple field, HMC force and temporal Wilson loops (L, T = 1..8); FP32 and
    FP32 two-row (re-unitarised) against FP64, with kernel times and GB/s."""
    import numpy as np
    dims = (32, 32, 32, 32)
    V = 32 ** 4
    pairs = [(L, T) for L in range(1, 9) for T in range(1, 9)]
    res = {"config": "C3 SU(3) 32^4 measurement + force kernels, eps in {0.1,0.5,1.0}"}
    for eps in [0.1, 0.5, 1.0]:
        row = {}
        ref = {}
        for st in ["f64", "f32", "f32_rows"]:
            g = klft.GaugeField(4, 3, dims, st).set_random_algebra(eps, 3000 + int(10 * eps))
            if st != "f64":
                g.reunitarize()
            # untimed first calls (a kernel's first launch loads its code object)
            klft.GaugePlaquette(g)
            W = np.array([w for _, _, w in klft.WilsonLoop_temporal(g, pairs)])
            klft.profile(True)
            P = [klft.GaugePlaquette(g) for _ in range(5)][-1]
            pms, pn = klft.profile_read("plaquette")
            klft.WilsonLoop_temporal(g, pairs)
            wms, _ = klft.profile_read("wilson")
            s = klft.stapleField(g)
            for _ in range(3):
                klft.lib.klft_staple_field(g.h, s.h)
            sms, sn = klft.profile_read("staple_field")
            h = klft.HMC(g, beta=6.0, tau=1.0, nsteps=20, seed=1)
            F = h.force()
            for _ in range(3):
                h.update_momentum(0.0)
            fms, fn = klft.profile_read("force")
            klft.profile(False)
            Sd = s.download()
            b = BYTES_PER_LINK[st]
            bs = 2 * 9 * (8 if st == "f64" else 4)
            entry = {"plaquette": P, "plaquette_us": 1e3 * pms / pn, "plaquette_GBs": 4 * V * b / (1e-3 * pms / pn) / 1e9,
                     "staple_us": 1e3 * sms / sn, "staple_GBs": 4 * V * (b + bs) / (1e-3 * sms / sn) / 1e9,
                     "force_us": 1e3 * fms / fn,
                     "force_GBs": 4 * V * (b + 2 * 8 * (8 if st == 'f64' else 4)) / (1e-3 * fms / fn) / 1e9,
                     "wilson_64pairs_ms": wms}
            if st == "f64":
                ref = {"P": P, "W": W, "S": Sd, "F": F}
            else:
                entry["plaquette_err_vs_fp64"] = abs(P - ref["P"])
                entry["wilson_max_abs_err_vs_fp64"] = float(np.abs(W - ref["W"]).max())
                entry["staple_max_abs_err_vs_fp64"] = float(np.abs(Sd - ref["S"]).max())
                entry["force_max_abs_err_vs_fp64"] = float(np.abs(F - ref["F"]).max())
            row[st] = entry
            h.close()
            s.close()
            g.close()
        res[f"eps={eps}"] = row
    return res


def config_c4(klft, a):
    """64 independent 4D chains, extents from {8,10,12,16,20} per dimension
    (seeded), U(1)/SU(2) alternating, beta 1.5..2.5, nHits=10, delta=0.1,
    FP32 links + FP64 plaquette sums: one graph-captured batched sweep vs 64
    per-chain sweeps."""
    import numpy as np
    rng = np.random.default_rng(64)
    ext = [8, 10, 12, 16, 20]
    chains = [(1 if k % 2 == 0 else 2, tuple(int(x) for x in rng.choice(ext, 4)), 1.5 + k / 63.0)
              for k in range(64)]
    vol = sum(int(np.prod(c[1])) for c in chains)
    b = klft.Batch(chains, "f32")
    b.set_identity()
    b.run(10, 0.1, 7, 0, 3)
    nsw = 20
    t = sync_time(lambda: b.run(10, 0.1, 7, 3, nsw))
    plaq = b.plaquette()
    fields = [klft.GaugeField(4, nc, dims, "f32").set_identity() for nc, dims, _ in chains]
    params = [klft.MetropolisParams(nHits=10, beta=beta, delta=0.1, seed=7) for _, _, beta in chains]
    for f, p in zip(fields, params):
        klft.metropolis_run(f, p, 0, 1, rates=False)

    def per_chain():
        for f, p in zip(fields, params):
            klft.metropolis_run(f, p, 1, nsw, rates=False)
    t1 = sync_time(per_chain)
    return {"config": "C4 64 chains, extents {8..20}^4 mixed, U(1)/SU(2), beta 1.5..2.5, FP32",
            "total_sites": vol, "batched_sweeps_per_s": nsw / t,
            "batched_link_updates_per_s": 4 * vol * nsw / t, "per_chain_sweeps_per_s": nsw / t1,
            "per_chain_link_updates_per_s": 4 * vol * nsw / t1, "speedup_batched": t1 / t,
            "plaquette_min": float(plaq.min()), "plaquette_max": float(plaq.max())}


def config_c5(klft, a):
    """16 SU(3) replicas of 16^4 with a 2x2x2 defect at x0 = 0, c_r = 1 - r/15,
    nHits=8, delta=0.2, FP32 sweeps, FP64 swap dS; one graph-captured step per
    replica sweep + swap pass."""
    import numpy as np
    n = 16
    c = 1.0 - np.arange(n) / 15.0
    p = klft.PTBC(n, 3, (16, 16, 16, 16), 6.0, 2, [0, 0, 0], c, "f32")
    p.set_identity()
    for s in range(5):
        p.step(8, 0.2, 11, s)
    nst = 40
    sw = np.zeros(n - 1)

    def run():
        for s in range(nst):
            sw[:] += p.step(8, 0.2, 11, 5 + s)
    t = sync_time(run)
    return {"config": "C5 PTBC 16 x SU(3) 16^4, defect 2^3 at x0=0, c_r = 1 - r/15, FP32",
            "steps_per_s": nst / t, "replica_link_updates_per_s": n * 4 * 16 ** 4 * nst / t,
            "swap_acceptance_per_pair": [float(x) for x in sw / nst],
            "plaquette_per_replica": [float(x) for x in p.plaquette()], "labels": [int(x) for x in p.labels()]}


def run_configs(klft, a):
    names = ["c1", "c2", "c3", "c4", "c5"] if a.config == "all" else a.config.split(",")
    fns = {"c1": config_c1, "c2": config_c2, "c3": config_c3, "c4": config_c4, "c5": config_c5}
    for nm in names:
        print(f"[bench] config {nm}", file=sys.stderr, flush=True)
        r = fns[nm](klft, a)
        print(json.dumps(r, default=float), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (every box has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    import klft_amd as klft

    if a.config:
        run_configs(klft, a)
        return

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
    batched = {}
    if not a.no_extra:
        for st in precisions:
            extra[st] = hmc_leg(klft, st, a, dims)
            print(f"[bench] hmc {st}: {extra[st]['traj_per_s']:.2f} traj/s", file=sys.stderr, flush=True)
        batched = batched_leg(klft, a)
    del g64

    main_st = "f64" if "f64" in legs else precisions[0]
    main_leg = legs[main_st]
    prof = (pass_profile() or {}).get(main_st, {})
    traffic = prof.get("bytes_per_pass")
    roofline = {"bound": "hbm", "kernel": f"metropolis_pass ({main_st} SU(3))",
                "achieved": round(main_leg["achieved_GBs"], 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(main_leg["achieved_GBs"] / PEAK_HBM_GBS, 4), "traffic": traffic}
    if traffic:
        # the roof that binds: HBM on the bytes the pass actually moves, or
        # VALU issue (rocprofv3 VALUBusy of the same kernel)
        tf = traffic / (main_leg["pass_us"] * 1e-6) / 1e12 / (PEAK_HBM_GBS / 1e3)
        roofline["traffic_GBs"] = round(traffic / (main_leg["pass_us"] * 1e-6) / 1e9, 1)
        roofline["traffic_frac"] = round(tf, 4)
        # the two floors under one pass: its measured bytes at 8 TB/s, and
        # the time its VALU is busy (VALUBusy x duration, with the VALU
        # instructions per wave behind it); the larger one is the bound
        roofline["hbm_traffic_floor_us"] = round(traffic / (PEAK_HBM_GBS * 1e9) * 1e6, 1)
        vb = prof.get("valu_busy_pct")
        if vb is not None:
            roofline["valu_busy_frac"] = round(vb / 100.0, 4)
            roofline["valu_active_us"] = round(vb / 100.0 * main_leg["pass_us"], 1)
            if prof.get("valu_insts_per_wave"):
                roofline["valu_insts_per_wave"] = prof["valu_insts_per_wave"]
            roofline["bound"] = "hbm" if tf >= vb / 100.0 else "valu"
        roofline["profile"] = "profiles/metropolis_pass_su3_32.json"
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
        "batched": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in batched.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
