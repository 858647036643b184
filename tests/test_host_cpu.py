"""Host-side checks that need no GPU: the C ABI library loads and exports every
declared symbol, the library's Philox header reproduces the known answers
and the oracle's streams, the drop-in YAML parser reads inputs as yaml-cpp
would (checked against PyYAML plus yaml-cpp's fallback rules), and the
metropolis driver's argument / input error paths (binaries/metropolis.cpp,
lib/Metropolis.cpp:61-68)."""
import ctypes
import json
import os
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
import yaml

import oracle as orc

ROOT = Path(__file__).resolve().parents[1]
HELP = Path(__file__).resolve().parent / "helpers"
BIN = Path(__file__).resolve().parent / "golden" / "_bin"
LIB = ROOT / "klft_amd" / "libklft_hip.so"


def declared_functions():
    src = (ROOT / "include" / "klft_hip.h").read_text()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(klft_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    assert LIB.exists(), "build the library first (python -m klft_amd.build)"
    lib = ctypes.CDLL(str(LIB))
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_marks_launches_with_the_reference_labels():
    """Every launch sits in a roctx range named with the reference's kernel
    label (Metropolis.hpp:161, GaugePlaquette.hpp:117, Gauge_Util.hpp:45,
    WilsonLoop.hpp:226, HamiltonianField.hpp:25-34): the library imports the
    roctx range calls and carries the label strings."""
    assert LIB.exists(), "build the library first (python -m klft_amd.build)"
    syms = subprocess.run(["nm", "-D", str(LIB)], capture_output=True, text=True, check=True).stdout
    assert re.search(r"U roctxRangePushA", syms) and re.search(r"U roctxRangePop", syms)
    blob = LIB.read_bytes()
    for label in [b"sweep_Metropolis_GaugeField_sublat_%d", b"GaugePlaquette_GaugeField", b"stapleField_GaugeField",
                  b"WilsonLoop_GaugeField_WLoop_munu", b"derivative+update_momentum+update_gauge"]:
        assert label in blob, label


def test_no_null_stream_memory_operations():
    """The library's kernels run on one non-blocking stream, which is not
    ordered after work on the null stream: a hipMemset or hipMemcpy there can
    still be running when a later kernel starts (a field's clear of its links
    once landed after the initialiser had written them, leaving zero links
    that the sweep turned into NaN).  Every memset and copy in the library is
    the Async form on the library stream."""
    bad = []
    for p in sorted((ROOT / "klft_amd" / "csrc").rglob("*")):
        if p.suffix not in (".hip", ".hpp", ".cpp", ".h"):
            continue
        src = re.sub(r"//[^\n]*", "", p.read_text())
        for m in re.finditer(r"\bhip(Memset|Memcpy)(D32|D16|D8|2D|3D|Peer|ToSymbol|FromSymbol)?\s*\(", src):
            bad.append(f"{p.relative_to(ROOT)}:{src.count(chr(10), 0, m.start()) + 1}: {m.group(0)}")
    assert not bad, bad


def test_python_binding_covers_the_header():
    import klft_amd
    klft_amd.load()
    assert set(declared_functions()) <= set(klft_amd.EXPORTED)


@pytest.fixture(scope="session")
def helpers():
    BIN.mkdir(parents=True, exist_ok=True)
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{ROOT / 'klft_amd' / 'csrc'}", "-o", str(BIN / "philox_dump"),
                    str(HELP / "philox_dump.cpp")], check=True)
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{ROOT / 'include'}", "-o", str(BIN / "yaml_dump"),
                    str(HELP / "yaml_dump.cpp")], check=True)
    return BIN


@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1])])
def test_library_philox_known_answers(helpers, ctr, key, want):
    out = subprocess.run([str(helpers / "philox_dump")] + [f"{v:x}" for v in ctr + key] + ["0"],
                         capture_output=True, text=True, check=True).stdout.split()
    assert [int(w, 16) for w in out] == want


def test_library_streams_equal_oracle_streams(helpers):
    for args in [(32091, 1, 7, 3, 9, 0, 2), (2 ** 40 + 5, 2, 1234, 1, 0, 0, 3), (1, 3, 0, 0, 0, 63, 1)]:
        out = subprocess.run([str(helpers / "philox_dump")] + [str(a) for a in args] + ["9"],
                             capture_output=True, text=True, check=True).stdout.split()
        np.testing.assert_array_equal(np.array(out, dtype=float), orc.stream_doubles(*args, 9))


# ---------------------------------------------------------------------------
# YAML: expected values from PyYAML with yaml-cpp's as<T>(fallback) rules
# (InputParser.hpp:33-125)
# ---------------------------------------------------------------------------
MP_KEYS = [("Ndims", int, 4), ("L0", int, 32), ("L1", int, 32), ("L2", int, 32), ("L3", int, 32),
           ("nHits", int, 10), ("nSweep", int, 1000), ("seed", int, 1234), ("Nd", "uint", 4), ("Nc", "uint", 2),
           ("beta", float, 1.0), ("delta", float, 0.1)]


def expected(text):
    doc = yaml.safe_load(text)
    if not isinstance(doc, dict) or "MetropolisParams" not in doc or "GaugeObservableParams" not in doc:
        return {"ok": False}
    mp = doc["MetropolisParams"] or {}
    gp = doc["GaugeObservableParams"] or {}

    def get(d, k, typ, fb):
        v = d.get(k) if isinstance(d, dict) else None
        if typ is int:
            return v if isinstance(v, int) and not isinstance(v, bool) else fb
        if typ == "uint":  # size_t keys: yaml-cpp rejects a sign
            return v if isinstance(v, int) and not isinstance(v, bool) and v >= 0 else fb
        if typ is float:
            return float(v) if isinstance(v, (int, float)) and not isinstance(v, bool) else fb
        if typ is bool:
            return v if isinstance(v, bool) else fb
        return v if isinstance(v, str) else fb

    e = {"ok": True}
    for k, t, fb in MP_KEYS:
        e[k] = get(mp, k, t, fb)
    e["L"] = [e.pop("L0"), e.pop("L1"), e.pop("L2"), e.pop("L3")]
    e["measurement_interval"] = get(gp, "measurement_interval", "uint", 0)
    for k in ["measure_plaquette", "measure_wilson_loop_temporal", "measure_wilson_loop_mu_nu", "write_to_file"]:
        e[k] = get(gp, k, bool, False)
    for k in ["W_temp_L_T_pairs", "W_mu_nu_pairs", "W_Lmu_Lnu_pairs"]:
        e[k] = [list(p) for p in gp.get(k) or []]
    for k in ["plaquette_filename", "W_temp_filename", "W_mu_nu_filename"]:
        e[k] = get(gp, k, str, "")
    return e


README_INPUT = """# input.yaml
MetropolisParams:    # parameters for the Metropolis algorithm
  Ndims: 4    # number of dimensions [2,3,4]
  Nd: 4       # number of link dimensions (this must be strictly same as Ndims)
  Nc: 1       # number of colors (defines SU(Nc)) [1,2,3]
  L0: 8       # lattice extent in 0 direction
  L1: 8       # lattice extent in 1 direction
  L2: 8       # lattice extent in 2 direction
  L3: 8       # lattice extent in 3 direction
  nHits: 10       # number of hits per sweep
  nSweep: 1000      # number of sweeps
  seed: 32091     # random seed
  beta: 2.0       # inverse coupling constant
  delta: 0.1      # step size for the Metropolis algorithm

GaugeObservableParams:
  measurement_interval: 10               # interval for measurements
  measure_plaquette: true                # measure the plaquette
  measure_wilson_loop_temporal: true    # measure the temporal Wilson loop
  measure_wilson_loop_mu_nu: true       # measure the spatial Wilson loop
  W_temp_L_T_pairs:      # pairs of (L, T) values for the temporal Wilson loop
    - [2, 2]
    - [3, 4]             # keep a non-decreasing order (as much as possible)
    - [4, 3]             # for maximum efficiency
    - [4, 4]
  W_mu_nu_pairs:      # pairs of (mu, nu) values for the planar Wilson loop
    - [0, 1]
    - [1, 2]
    - [3, 2]
  W_Lmu_Lnu_pairs:      # pairs of (Lmu, Lnu) values for the lengths of the
    - [2, 2]            # planar Wilson loop in the mu and nu directions
    - [3, 3]            # again, keep a non-decreasing order (as much as possible)
    - [4, 3]
  plaquette_filename: "plaquette.out"  # filename to output the plaquette
  W_temp_filename: "W_temp.out"        # filename to output the temporal Wilson loop
  W_mu_nu_filename: "W_mu_nu.out"      # filename to output the planar Wilson loop
  write_to_file: true                  # write the measurements to file
"""

YAML_CASES = {
    "readme": README_INPUT,
    "defaults": "MetropolisParams:\n  Nc: 3\nGaugeObservableParams:\n  measure_plaquette: yes\n",
    "flow_pairs": ("MetropolisParams: {Ndims: 2, L0: 16, L1: 16, beta: 2.3}\n"
                   "GaugeObservableParams:\n  W_temp_L_T_pairs: [[1, 1], [2, 3]]\n  write_to_file: false\n"
                   "  plaquette_filename: 'p.txt'\n"),
    "seq_same_indent": ("MetropolisParams:\n  Nc: 2\nGaugeObservableParams:\n  W_mu_nu_pairs:\n  - [0, 1]\n"
                        "  - [2, 3]\n  measurement_interval: 5\n"),
    "bad_values_fall_back": ("MetropolisParams:\n  L0: eight\n  beta: high\n  nHits: 3.5\n"
                             "GaugeObservableParams:\n  measure_plaquette: maybe\n"),
    "missing_observables": "MetropolisParams:\n  Nc: 1\n",
    "missing_metropolis": "GaugeObservableParams:\n  measure_plaquette: true\n",
    "empty_sections": "MetropolisParams:\nGaugeObservableParams:\n",
    "negative_interval": "MetropolisParams:\n  seed: -5\nGaugeObservableParams:\n  measurement_interval: -3\n",
}


@pytest.mark.parametrize("name", sorted(YAML_CASES))
def test_yaml_parser_matches_pyyaml(helpers, tmp_path, name):
    text = YAML_CASES[name]
    f = tmp_path / "in.yaml"
    f.write_text(text)
    out = subprocess.run([str(helpers / "yaml_dump"), str(f)], capture_output=True, text=True).stdout
    got = json.loads(out.strip().split("\n")[-1])
    want = expected(text)
    assert got == want, (got, want)


def test_yaml_malformed_pair_is_an_error(helpers, tmp_path):
    f = tmp_path / "bad.yaml"
    f.write_text("MetropolisParams:\n  Nc: 1\nGaugeObservableParams:\n  W_temp_L_T_pairs:\n    - [2, x]\n")
    out = subprocess.run([str(helpers / "yaml_dump"), str(f)], capture_output=True, text=True).stdout
    assert "Error parsing input file" in out and json.loads(out.strip().split("\n")[-1]) == {"ok": False}


# ---------------------------------------------------------------------------
# metropolis driver error paths (no GPU work happens before parsing)
# ---------------------------------------------------------------------------
METRO = ROOT / "klft_amd" / "metropolis"


@pytest.mark.skipif(not METRO.exists(), reason="drivers not built")
def test_metropolis_cli_help_and_errors(tmp_path):
    r = subprocess.run([str(METRO), "-h"], capture_output=True, text=True)
    assert r.returncode == 0 and "--filename" in r.stdout and "Metropolis for SU(N) gauge fields" in r.stdout
    r = subprocess.run([str(METRO), "-f", str(tmp_path / "nope.yaml")], capture_output=True, text=True)
    assert r.returncode == 255 and "Error parsing input file" in r.stdout
    f = tmp_path / "half.yaml"
    f.write_text("MetropolisParams:\n  Nc: 1\n")
    r = subprocess.run([str(METRO), "-f", str(f)], capture_output=True, text=True)
    assert r.returncode == 255 and "GaugeObservableParams not found" in r.stdout
    r = subprocess.run([str(METRO), "--bogus"], capture_output=True, text=True)
    assert r.returncode == 255
