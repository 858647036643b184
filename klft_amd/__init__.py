"""klft_amd: MI355X-native klft gauge-update and measurement path.

Python view of the C ABI in include/klft_hip.h (libklft_hip.so, built in-tree
by ``python -m klft_amd.build``).  Names mirror the reference's C++ API
(deviceGaugeField, MetropolisParams, sweep_Metropolis, GaugePlaquette,
WilsonLoop_mu_nu, WilsonLoop_temporal, HMC ...), with the same argument
meaning and error behaviour: a failing call raises KlftError carrying the
library's message.  There is no CPU fallback: the library is loaded on first
use (so ``python -m klft_amd.build`` runs on a clean checkout), and every call
raises ImportError if it has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field as dc_field
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libklft_hip.so"

STORAGE_F64 = 0
STORAGE_F32 = 1
STORAGE_F32_ROWS = 2
STORAGES = {"f64": STORAGE_F64, "f32": STORAGE_F32, "f32_rows": STORAGE_F32_ROWS}
PROPOSALS = {"reference": 0, "symmetric": 1}


class KlftError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"klft error {code}: {msg}")
        self.code = code


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m klft_amd.build`")
    return C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)


class _Library:
    """libklft_hip.so, opened and given its ctypes signatures on first use."""

    def __init__(self):
        self._dll = None
        self.exported: list[str] = []

    def load(self) -> C.CDLL:
        if self._dll is None:
            dll = _load()
            names = []
            for name, (res, args) in _SIGS.items():
                fn = getattr(dll, name, None)
                if fn is None:
                    continue
                fn.restype = res
                fn.argtypes = args
                names.append(name)
            self.exported[:] = names
            self._dll = dll
        return self._dll

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        return getattr(self.load(), name)


lib = _Library()

_p = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class MetropolisParamsC(C.Structure):
    _fields_ = [("nhits", C.c_int), ("beta", C.c_double), ("delta", C.c_double), ("seed", C.c_uint64)]


class HmcParamsC(C.Structure):
    _fields_ = [("beta", C.c_double), ("tau", C.c_double), ("nsteps", C.c_int), ("seed", C.c_uint64)]


class ChainDescC(C.Structure):
    _fields_ = [("nc", C.c_int), ("dims", C.c_int * 4), ("beta", C.c_double)]


_SIGS = {
    "klft_init": (C.c_int, [C.c_int]),
    "klft_last_error": (C.c_char_p, []),
    "klft_version": (C.c_char_p, []),
    "klft_field_create": (C.c_int, [C.c_int, C.c_int, _ip, C.c_int, C.POINTER(_p)]),
    "klft_field_destroy": (C.c_int, [_p]),
    "klft_field_info": (C.c_int, [_p, _ip, _ip, _ip, _ip]),
    "klft_field_set_identity": (C.c_int, [_p]),
    "klft_field_set_random_sun": (C.c_int, [_p, C.c_double, C.c_uint64]),
    "klft_field_set_random_algebra": (C.c_int, [_p, C.c_double, C.c_uint64]),
    "klft_field_set_haar": (C.c_int, [_p, C.c_uint64]),
    "klft_field_set_random_complex": (C.c_int, [_p, C.c_uint64]),
    "klft_field_set_constant": (C.c_int, [_p, C.c_double, C.c_double]),
    "klft_field_upload": (C.c_int, [_p, _dp]),
    "klft_field_download": (C.c_int, [_p, _dp]),
    "klft_field_copy": (C.c_int, [_p, _p]),
    "klft_field_reunitarize": (C.c_int, [_p]),
    "klft_plaquette": (C.c_int, [_p, C.c_int, _dp, _dp]),
    "klft_plaquette_noalloc": (C.c_int, [_p, C.c_int, _dp, _dp]),
    "klft_staple_field": (C.c_int, [_p, _p]),
    "klft_wilson_loop_mu_nu": (C.c_int, [_p, C.c_int, C.c_int, _ip, C.c_int, C.c_int, _dp, _dp]),
    "klft_wilson_loop_temporal": (C.c_int, [_p, _ip, C.c_int, C.c_int, _dp]),
    "klft_metropolis_sweep": (C.c_int, [_p, C.POINTER(MetropolisParamsC), C.c_uint64, _dp, _i64p]),
    "klft_metropolis_run": (C.c_int, [_p, C.POINTER(MetropolisParamsC), C.c_uint64, C.c_int, _dp]),
    "klft_metropolis_schedule": (C.c_int, [_p, _i32p]),
    "klft_sweep_launches": (C.c_int, [C.c_int, _i32p, C.POINTER(C.c_int)]),
    "klft_run_launches": (C.c_int, [C.c_int, C.c_int, _i32p, C.c_int, C.POINTER(C.c_int)]),
    "klft_field_set_proposal": (C.c_int, [_p, C.c_int]),
    "klft_hmc_create": (C.c_int, [_p, C.POINTER(HmcParamsC), C.POINTER(_p)]),
    "klft_hmc_destroy": (C.c_int, [_p]),
    "klft_hmc_refresh_momenta": (C.c_int, [_p, C.c_uint64]),
    "klft_hmc_momenta_upload": (C.c_int, [_p, _dp]),
    "klft_hmc_momenta_download": (C.c_int, [_p, _dp]),
    "klft_hmc_force": (C.c_int, [_p, _dp]),
    "klft_hmc_update_momentum": (C.c_int, [_p, C.c_double]),
    "klft_hmc_update_gauge": (C.c_int, [_p, C.c_double]),
    "klft_hmc_kinetic": (C.c_int, [_p, _dp]),
    "klft_gauge_action": (C.c_int, [_p, C.c_double, _dp]),
    "klft_hmc_integrate": (C.c_int, [_p]),
    "klft_hmc_trajectory": (C.c_int, [_p, C.c_uint64, _dp, _ip]),
    "klft_batch_create": (C.c_int, [C.c_int, C.POINTER(ChainDescC), C.c_int, C.POINTER(_p)]),
    "klft_batch_destroy": (C.c_int, [_p]),
    "klft_batch_set_identity": (C.c_int, [_p]),
    "klft_batch_sweep": (C.c_int, [_p, C.c_int, C.c_double, C.c_uint64, C.c_uint64, _dp]),
    "klft_batch_run": (C.c_int, [_p, C.c_int, C.c_double, C.c_uint64, C.c_uint64, C.c_int]),
    "klft_batch_plaquette": (C.c_int, [_p, _dp]),
    "klft_batch_download": (C.c_int, [_p, C.c_int, _dp]),
    "klft_batch_upload": (C.c_int, [_p, C.c_int, _dp]),
    "klft_ptbc_create": (C.c_int, [C.c_int, C.c_int, _ip, C.c_int, C.c_double, C.c_int, _ip, _dp,
                                   C.POINTER(_p)]),
    "klft_ptbc_step": (C.c_int, [_p, C.c_int, C.c_double, C.c_uint64, C.c_uint64, _ip]),
    "klft_ptbc_labels": (C.c_int, [_p, _ip]),
    "klft_ptbc_action": (C.c_int, [_p, C.c_int, C.c_int, _dp]),
    "klft_metropolis_main": (C.c_int, [C.c_char_p]),
    "klft_tuning": (C.c_int, [C.c_int]),
    "klft_tune_cache_read": (C.c_int, [C.c_char_p]),
    "klft_tune_cache_write": (C.c_int, [C.c_char_p]),
    "klft_tune_lookup": (C.c_int, [C.c_char_p, _ip]),
    "klft_profile": (C.c_int, [C.c_int]),
    "klft_profile_read": (C.c_int, [C.c_char_p, _dp, _i64p]),
}

# names of _SIGS the loaded library exports (filled by the first load)
EXPORTED = lib.exported


def load() -> C.CDLL:
    """Open libklft_hip.so now (ImportError if it has not been built)."""
    return lib.load()


def _check(rc: int) -> None:
    if rc != 0:
        raise KlftError(rc, lib.klft_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _iarr(vals) -> "C.Array":
    vals = list(vals)
    return (C.c_int * max(1, len(vals)))(*vals)


def init(device: int = 0) -> None:
    _check(lib.klft_init(device))


def ngen(nc: int) -> int:
    return {1: 1, 2: 3, 3: 8}[nc]


class GaugeField:
    """deviceGaugeField<Nd,Nc> (GaugeField.hpp:30-239) on one MI355X.

    Host images use the reference's LayoutRight AoS order: complex array of
    shape (L0, ..., L_{nd-1}, nd, nc, nc).
    """

    def __init__(self, nd: int, nc: int, dims, storage="f64"):
        if isinstance(storage, str):
            storage = STORAGES[storage]
        dims = list(dims)[:nd]
        self.nd, self.nc, self.dims, self.storage = nd, nc, tuple(dims), storage
        h = _p()
        _check(lib.klft_field_create(nd, nc, _iarr(dims), storage, C.byref(h)))
        self.h = h

    @property
    def volume(self) -> int:
        return int(np.prod(self.dims))

    def shape(self):
        return (*self.dims, self.nd, self.nc, self.nc)

    def close(self):
        if getattr(self, "h", None):
            lib.klft_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_identity(self):
        _check(lib.klft_field_set_identity(self.h))
        return self

    def set_random_sun(self, delta: float, seed: int):
        _check(lib.klft_field_set_random_sun(self.h, delta, seed))
        return self

    def set_random_algebra(self, eps: float, seed: int):
        _check(lib.klft_field_set_random_algebra(self.h, eps, seed))
        return self

    def set_haar(self, seed: int):
        _check(lib.klft_field_set_haar(self.h, seed))
        return self

    def set_random_complex(self, seed: int):
        """deviceGaugeField(L.., rng) (GaugeField.hpp:118-141)."""
        _check(lib.klft_field_set_random_complex(self.h, seed))
        return self

    def set_constant(self, value: complex):
        """deviceGaugeField(L.., complex_t init) (GaugeField.hpp:65-84)."""
        value = complex(value)
        _check(lib.klft_field_set_constant(self.h, value.real, value.imag))
        return self

    def upload(self, U: np.ndarray):
        a = np.ascontiguousarray(U, dtype=np.complex128).reshape(-1)
        if a.size != self.volume * self.nd * self.nc * self.nc:
            raise ValueError("host image has the wrong size")
        _check(lib.klft_field_upload(self.h, a.ctypes.data_as(_dp)))
        return self

    def download(self) -> np.ndarray:
        a = np.empty(self.shape(), dtype=np.complex128)
        _check(lib.klft_field_download(self.h, a.ctypes.data_as(_dp)))
        return a

    def copy_from(self, other: "GaugeField"):
        _check(lib.klft_field_copy(self.h, other.h))
        return self

    def reunitarize(self):
        _check(lib.klft_field_reunitarize(self.h))
        return self

    def set_proposal(self, law: str):
        """SU(3) Metropolis proposal: "reference" (randSUN<3>, SUN.hpp:192-251)
        or "symmetric" (exp(i delta h.lambda), h_a ~ U(-1, 1); no counterpart
        in the reference)."""
        _check(lib.klft_field_set_proposal(self.h, PROPOSALS[law]))
        return self


def GaugePlaquette(g: GaugeField, normalize: bool = True, per_site: bool = False):
    """GaugePlaquette<rank,Nc> (GaugePlaquette.hpp:89-135)."""
    out = C.c_double()
    ps = np.empty(g.volume) if per_site else None
    _check(lib.klft_plaquette(g.h, int(normalize), C.byref(out), _dptr(ps) if per_site else None))
    return (out.value, ps) if per_site else out.value


def stapleField(g: GaugeField) -> GaugeField:
    """stapleField (Gauge_Util.hpp:27-97)."""
    st = STORAGE_F64 if g.storage == STORAGE_F64 else STORAGE_F32
    out = GaugeField(g.nd, g.nc, g.dims, st)
    _check(lib.klft_staple_field(g.h, out.h))
    return out


def WilsonLoop_mu_nu(g: GaugeField, mu: int, nu: int, pairs, normalize: bool = True,
                     per_site: bool = False):
    """WilsonLoop_mu_nu (WilsonLoop.hpp:170-245); returns [(mu, nu, Lmu, Lnu, W)]."""
    pairs = [tuple(p) for p in pairs]
    flat = _iarr([v for p in pairs for v in p])
    out = np.zeros(max(1, len(pairs)))
    ps = np.empty(g.volume) if per_site else None
    _check(lib.klft_wilson_loop_mu_nu(g.h, mu, nu, flat, len(pairs), int(normalize), _dptr(out),
                                      _dptr(ps) if per_site else None))
    res = [(float(mu), float(nu), float(p[0]), float(p[1]), float(out[i])) for i, p in enumerate(pairs)]
    return (res, ps) if per_site else res


def WilsonLoop_temporal(g: GaugeField, pairs, normalize: bool = True):
    """WilsonLoop_temporal (WilsonLoop.hpp:250-304); returns [(L, T, W)]."""
    pairs = [tuple(p) for p in pairs]
    flat = _iarr([v for p in pairs for v in p])
    out = np.zeros(max(1, len(pairs)))
    _check(lib.klft_wilson_loop_temporal(g.h, flat, len(pairs), int(normalize), _dptr(out)))
    return [(float(p[0]), float(p[1]), float(out[i])) for i, p in enumerate(pairs)]


@dataclass
class MetropolisParams:
    """MetropolisParams (Metropolis_Params.hpp:28-99), struct defaults."""
    Ndims: int = 4
    L0: int = 4
    L1: int = 4
    L2: int = 4
    L3: int = 4
    nHits: int = 10
    nSweep: int = 1000
    seed: int = 1234
    Nd: int = 4
    Nc: int = 2
    beta: float = 1.0
    delta: float = 0.1

    def c(self) -> MetropolisParamsC:
        return MetropolisParamsC(self.nHits, self.beta, self.delta, self.seed)


def sweep_Metropolis(g: GaugeField, params: MetropolisParams, sweep: int = 0,
                     with_count: bool = False):
    """sweep_Metropolis<rank,Nc> (Metropolis.hpp:118-178): returns the acceptance rate."""
    acc = C.c_double()
    n = C.c_int64()
    p = params.c()
    _check(lib.klft_metropolis_sweep(g.h, C.byref(p), sweep, C.byref(acc), C.byref(n)))
    return (acc.value, n.value) if with_count else acc.value


def metropolis_run(g: GaugeField, params: MetropolisParams, sweep0: int, n: int, rates: bool = True):
    out = np.zeros(max(1, n)) if rates else None
    p = params.c()
    _check(lib.klft_metropolis_run(g.h, C.byref(p), sweep0, n, _dptr(out) if rates else None))
    return out[:n] if rates else None


def metropolis_schedule(g: GaugeField) -> np.ndarray:
    v = np.empty(g.volume, dtype=np.int32)
    _check(lib.klft_metropolis_schedule(g.h, v.ctypes.data_as(_i32p)))
    return v


def sweep_launches(nd: int):
    """The launches of a sweep whose independent passes run two to a launch:
    a list of pass tuples, in launch order."""
    g = np.empty(2 << nd, dtype=np.int32)
    n = C.c_int()
    _check(lib.klft_sweep_launches(nd, g.ctypes.data_as(_i32p), C.byref(n)))
    return [tuple(int(p) for p in g[2 * i:2 * i + 2] if p >= 0) for i in range(n.value)]


def run_launches(nd: int, nsweeps: int):
    """The launches of metropolis_run over nsweeps sweeps: a list of
    ((pass, sweep), ...) tuples in launch order, sweeps relative to the
    run's first (independent passes of consecutive sweeps share launches)."""
    cap = (1 << nd) * max(nsweeps, 1)
    g = np.empty(4 * cap, dtype=np.int32)
    n = C.c_int()
    _check(lib.klft_run_launches(nd, nsweeps, g.ctypes.data_as(_i32p), cap, C.byref(n)))
    return [tuple((int(g[4 * i + 2 * k]), int(g[4 * i + 2 * k + 1])) for k in range(2) if g[4 * i + 2 * k] >= 0)
            for i in range(n.value)]


class HMC:
    """HMC::hmc_step driver (HMC.hpp:48-71) over a GaugeField."""

    def __init__(self, g: GaugeField, beta: float, tau: float, nsteps: int, seed: int):
        self.g = g
        self.params = HmcParamsC(beta, tau, nsteps, seed)
        h = _p()
        _check(lib.klft_hmc_create(g.h, C.byref(self.params), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.klft_hmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _mshape(self):
        return (*self.g.dims, self.g.nd, ngen(self.g.nc))

    def refresh_momenta(self, traj: int):
        _check(lib.klft_hmc_refresh_momenta(self.h, traj))

    def momenta(self) -> np.ndarray:
        a = np.empty(self._mshape())
        _check(lib.klft_hmc_momenta_download(self.h, _dptr(a)))
        return a

    def set_momenta(self, P: np.ndarray):
        a = np.ascontiguousarray(P, dtype=np.float64)
        _check(lib.klft_hmc_momenta_upload(self.h, _dptr(a)))

    def force(self) -> np.ndarray:
        a = np.empty(self._mshape())
        _check(lib.klft_hmc_force(self.h, _dptr(a)))
        return a

    def update_momentum(self, dtau: float):
        _check(lib.klft_hmc_update_momentum(self.h, dtau))

    def update_gauge(self, dtau: float):
        _check(lib.klft_hmc_update_gauge(self.h, dtau))

    def kinetic(self) -> float:
        o = C.c_double()
        _check(lib.klft_hmc_kinetic(self.h, C.byref(o)))
        return o.value

    def integrate(self):
        _check(lib.klft_hmc_integrate(self.h))

    def trajectory(self, traj: int):
        dH = C.c_double()
        acc = C.c_int()
        _check(lib.klft_hmc_trajectory(self.h, traj, C.byref(dH), C.byref(acc)))
        return dH.value, bool(acc.value)


def gauge_action(g: GaugeField, beta: float) -> float:
    o = C.c_double()
    _check(lib.klft_gauge_action(g.h, beta, C.byref(o)))
    return o.value


def profile(on: bool = True) -> None:
    """Enable kernel timing (HIP events on the library stream)."""
    _check(lib.klft_profile(int(on)))


def profile_read(name: str):
    """(total_ms, launches) recorded for kernel `name` since the last read."""
    ms = C.c_double()
    n = C.c_int64()
    _check(lib.klft_profile_read(name.encode(), C.byref(ms), C.byref(n)))
    return ms.value, n.value


def tuning(on: bool = True) -> None:
    _check(lib.klft_tuning(int(on)))


def tune_cache_read(path: str) -> None:
    _check(lib.klft_tune_cache_read(str(path).encode()))


def tune_cache_write(path: str) -> None:
    _check(lib.klft_tune_cache_write(str(path).encode()))


class Batch:
    """Many independent 4D chains updated by one launch per sublattice pass."""

    def __init__(self, chains, storage="f32"):
        """chains: list of (nc, (L0, L1, L2, L3), beta)."""
        if isinstance(storage, str):
            storage = STORAGES[storage]
        arr = (ChainDescC * len(chains))()
        for i, (nc, dims, beta) in enumerate(chains):
            arr[i].nc = nc
            for d in range(4):
                arr[i].dims[d] = dims[d]
            arr[i].beta = beta
        self.chains = [(nc, tuple(dims), beta) for nc, dims, beta in chains]
        h = _p()
        _check(lib.klft_batch_create(len(chains), arr, storage, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.klft_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_identity(self):
        _check(lib.klft_batch_set_identity(self.h))

    def sweep(self, nhits, delta, seed, sweep):
        acc = np.zeros(len(self.chains))
        _check(lib.klft_batch_sweep(self.h, nhits, delta, seed, sweep, _dptr(acc)))
        return acc

    def run(self, nhits, delta, seed, sweep0, n):
        _check(lib.klft_batch_run(self.h, nhits, delta, seed, sweep0, n))

    def plaquette(self):
        out = np.zeros(len(self.chains))
        _check(lib.klft_batch_plaquette(self.h, _dptr(out)))
        return out

    def download(self, i):
        nc, dims, _ = self.chains[i]
        a = np.empty((*dims, 4, nc, nc), dtype=np.complex128)
        _check(lib.klft_batch_download(self.h, i, a.ctypes.data_as(_dp)))
        return a

    def upload(self, i, U):
        a = np.ascontiguousarray(U, dtype=np.complex128)
        _check(lib.klft_batch_upload(self.h, i, a.ctypes.data_as(_dp)))


class PTBC(Batch):
    """Replica set with a defect whose boundary links carry coupling c[r]."""

    def __init__(self, nreplicas, nc, dims, beta, defect_length, defect_pos, c, storage="f32"):
        if isinstance(storage, str):
            storage = STORAGES[storage]
        self.chains = [(nc, tuple(dims), beta)] * nreplicas
        cv = np.ascontiguousarray(c, dtype=np.float64)
        h = _p()
        _check(lib.klft_ptbc_create(nreplicas, nc, _iarr(dims), storage, beta, defect_length, _iarr(defect_pos),
                                    _dptr(cv), C.byref(h)))
        self.h = h
        self.c = cv

    def step(self, nhits, delta, seed, step):
        sw = np.zeros(max(1, len(self.chains) - 1), dtype=np.int32)
        _check(lib.klft_ptbc_step(self.h, nhits, delta, seed, step, sw.ctypes.data_as(_ip)))
        return sw[:len(self.chains) - 1]

    def labels(self):
        out = np.zeros(len(self.chains), dtype=np.int32)
        _check(lib.klft_ptbc_labels(self.h, out.ctypes.data_as(_ip)))
        return out

    def action(self, replica, coupling):
        o = C.c_double()
        _check(lib.klft_ptbc_action(self.h, replica, coupling, C.byref(o)))
        return o.value
