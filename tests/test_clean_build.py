"""A clean checkout builds.

The files a fresh clone holds (tracked and not-ignored files: no built
library, no object files, no oracle build) are copied to a scratch directory;
__graft_entry__.build() compiles the oracle and every HIP source for gfx950
there, and the package then opens the fresh library and binds every symbol
declared in include/klft_hip.h.  No GPU is needed: nothing is launched.
"""
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _checkout_files():
    if shutil.which("git") is None or not (ROOT / ".git").exists():
        pytest.skip("not a git checkout")
    out = subprocess.run(["git", "ls-files", "--cached", "--others", "--exclude-standard", "-z"], cwd=ROOT,
                         capture_output=True, check=True).stdout
    return [f for f in out.decode().split("\0") if f and (ROOT / f).is_file()]


def test_clean_checkout_builds_and_loads(tmp_path):
    files = _checkout_files()
    assert "klft_amd/build.py" in files and "__graft_entry__.py" in files
    assert not any(f.endswith(".so") for f in files)
    for f in files:
        dst = tmp_path / f
        dst.parent.mkdir(parents=True, exist_ok=True)
        shutil.copy2(ROOT / f, dst)
    r = subprocess.run([sys.executable, "-c", "import __graft_entry__ as e; e.build()"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert (tmp_path / "klft_amd" / "libklft_hip.so").exists()
    check = ("import sys, re; sys.path.insert(0, '.');"
             "import klft_amd; klft_amd.load();"
             "src = re.sub(r'/\\*.*?\\*/', '', open('include/klft_hip.h').read(), flags=re.S);"
             "names = set(re.findall(r'\\b(klft_\\w+)\\s*\\(', src));"
             "missing = names - set(klft_amd.EXPORTED);"
             "assert not missing, missing;"
             "print(klft_amd.lib.klft_version().decode())")
    r = subprocess.run([sys.executable, "-c", check], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "gfx950" in r.stdout
