"""ctypes view of the FP64 CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP library against it.  Arrays
use the reference's LayoutRight AoS image, complex (L0, .., L_{nd-1}, nd, nc, nc).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
SO = ROOT / "oracle" / "liboracle.so"


def _load():
    if not SO.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    return C.CDLL(str(SO))


L = _load()
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_u32 = C.c_uint32
_u64 = C.c_uint64

L.orc_philox4x32_10.argtypes = [C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]
L.orc_stream_doubles.argtypes = [_u64, _u32, _u32, _u32, _u32, _u32, _u32, C.c_int, _dp]
L.orc_matmul.argtypes = [C.c_int, _dp, _dp, _dp]
L.orc_dagger.argtypes = [C.c_int, _dp, _dp]
L.orc_restore_sun.argtypes = [C.c_int, _dp]
L.orc_rand_sun.argtypes = [C.c_int, C.c_double, _u64, _u32, _u32, _u32, _u32, _dp]
L.orc_rand_sun.restype = C.c_int
L.orc_exp_iq.argtypes = [C.c_int, _dp, _dp]
L.orc_field_identity.argtypes = [C.c_int, C.c_int, _ip, _dp]
L.orc_field_random_algebra.argtypes = [C.c_int, C.c_int, _ip, C.c_double, _u64, _dp]
L.orc_field_random_sun.argtypes = [C.c_int, C.c_int, _ip, C.c_double, _u64, _dp]
L.orc_field_random_complex.argtypes = [C.c_int, C.c_int, _ip, _u64, _dp]
L.orc_staple.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_long, C.c_int, _dp]
L.orc_staple_field.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp]
L.orc_plaquette.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int, _dp]
L.orc_plaquette.restype = C.c_double
L.orc_wilson_loop_mu_nu.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int, C.c_int, _ip, C.c_int,
                                    C.c_int, _dp, _dp]
L.orc_wilson_loop_temporal.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_int, C.c_int, _dp]
L.orc_metropolis_sweep.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int, C.c_double, C.c_double,
                                   _u64, _u32, _lp, _ip]
L.orc_metropolis_sweep.restype = C.c_double
L.orc_ngen.argtypes = [C.c_int]
L.orc_ngen.restype = C.c_int
L.orc_refresh_momenta.argtypes = [C.c_int, C.c_int, _ip, _u64, _u32, _dp]
L.orc_force.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_double, _dp]
L.orc_update_momentum.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, C.c_double]
L.orc_update_gauge.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, C.c_double]
L.orc_restore_field.argtypes = [C.c_int, C.c_int, _ip, _dp]
L.orc_kinetic.argtypes = [C.c_int, C.c_int, _ip, _dp]
L.orc_kinetic.restype = C.c_double
L.orc_gauge_action.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_double]
L.orc_gauge_action.restype = C.c_double
L.orc_leapfrog.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, C.c_double, C.c_double, C.c_int]
L.orc_hmc_trajectory.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_double, C.c_double, C.c_int, _u64,
                                 _u32, _dp, _dp]
L.orc_hmc_trajectory.restype = C.c_int
L.orc_defect_field.argtypes = [C.c_int, _ip, C.c_int, _ip, C.c_double, _dp]
L.orc_plaquette_defect.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, C.c_int]
L.orc_plaquette_defect.restype = C.c_double
L.orc_metropolis_sweep_defect.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, C.c_int, C.c_double,
                                          C.c_double, _u64, _u32, _u32, _lp]
L.orc_metropolis_sweep_defect.restype = C.c_double


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(v):
    v = list(v)
    return (C.c_int * max(1, len(v)))(*v)


def philox(ctr, key):
    c = (_u32 * 4)(*ctr)
    k = (_u32 * 2)(*key)
    o = (_u32 * 4)()
    L.orc_philox4x32_10(c, k, o)
    return list(o)


def stream_doubles(seed, tag, c0, c1, c2, chain, mu, n):
    out = np.empty(n)
    L.orc_stream_doubles(seed, tag, c0, c1, c2, chain, mu, n, _d(out))
    return out


def field_shape(nd, nc, dims):
    return (*dims[:nd], nd, nc, nc)


def identity(nd, nc, dims):
    U = np.empty(field_shape(nd, nc, dims), dtype=np.complex128)
    L.orc_field_identity(nd, nc, _i(dims), _d(U))
    return U


def random_algebra(nd, nc, dims, eps, seed):
    U = np.empty(field_shape(nd, nc, dims), dtype=np.complex128)
    L.orc_field_random_algebra(nd, nc, _i(dims), eps, seed, _d(U))
    return U


def random_sun_field(nd, nc, dims, delta, seed):
    U = np.empty(field_shape(nd, nc, dims), dtype=np.complex128)
    L.orc_field_random_sun(nd, nc, _i(dims), delta, seed, _d(U))
    return U


def random_complex_field(nd, nc, dims, seed):
    U = np.empty(field_shape(nd, nc, dims), dtype=np.complex128)
    L.orc_field_random_complex(nd, nc, _i(dims), seed, _d(U))
    return U


def plaquette(U, nd, nc, dims, normalize=True, per_site=False):
    U = np.ascontiguousarray(U)
    ps = np.empty(int(np.prod(dims[:nd]))) if per_site else None
    v = L.orc_plaquette(nd, nc, _i(dims), _d(U), int(normalize), _d(ps) if per_site else None)
    return (v, ps) if per_site else v


def staple(U, nd, nc, dims, site, mu):
    S = np.empty((nc, nc), dtype=np.complex128)
    L.orc_staple(nd, nc, _i(dims), _d(np.ascontiguousarray(U)), site, mu, _d(S))
    return S


def staple_field(U, nd, nc, dims):
    S = np.empty(field_shape(nd, nc, dims), dtype=np.complex128)
    L.orc_staple_field(nd, nc, _i(dims), _d(np.ascontiguousarray(U)), _d(S))
    return S


def wilson_mu_nu(U, nd, nc, dims, mu, nu, pairs, normalize=True, per_site=False):
    out = np.zeros(max(1, len(pairs)))
    ps = np.empty(int(np.prod(dims[:nd]))) if per_site else None
    L.orc_wilson_loop_mu_nu(nd, nc, _i(dims), _d(np.ascontiguousarray(U)), mu, nu,
                            _i([v for p in pairs for v in p]), len(pairs), int(normalize), _d(out),
                            _d(ps) if per_site else None)
    return (out[:len(pairs)], ps) if per_site else out[:len(pairs)]


def wilson_temporal(U, nd, nc, dims, pairs, normalize=True):
    out = np.zeros(max(1, len(pairs)))
    L.orc_wilson_loop_temporal(nd, nc, _i(dims), _d(np.ascontiguousarray(U)),
                               _i([v for p in pairs for v in p]), len(pairs), int(normalize), _d(out))
    return out[:len(pairs)]


L.orc_metropolis_sweep_sym.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int, C.c_double, C.c_double,
                                       C.c_uint64, C.c_uint32, C.POINTER(C.c_long)]
L.orc_metropolis_sweep_sym.restype = C.c_double
L.orc_rand_sym3.argtypes = [C.c_double, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _dp]


def metropolis_sweep_sym(U, nd, nc, dims, nhits, beta, delta, seed, sweep):
    """Metropolis sweep with the opt-in symmetric SU(3) proposal, in place on U."""
    assert U.flags["C_CONTIGUOUS"] and U.dtype == np.complex128
    n = C.c_long()
    acc = L.orc_metropolis_sweep_sym(nd, nc, _i(dims), _d(U), nhits, beta, delta, seed, sweep, C.byref(n))
    return acc, n.value


L.orc_metropolis_sweep_narrow.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_int, C.c_double, C.c_double,
                                          C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_long)]
L.orc_metropolis_sweep_narrow.restype = C.c_double


def metropolis_sweep_narrow(U, nd, nc, dims, nhits, beta, delta, seed, sweep, sym=False, chain=0):
    """Metropolis sweep with the FP32 links' draws (24-bit uniforms, one stream
    word each) in FP64 arithmetic, in place on U; returns (acc, naccepted).
    chain: the stream's chain index (a batched chain's position)."""
    assert U.flags["C_CONTIGUOUS"] and U.dtype == np.complex128
    n = C.c_long()
    acc = L.orc_metropolis_sweep_narrow(nd, nc, _i(dims), _d(U), nhits, beta, delta, seed, sweep, chain,
                                        int(sym), C.byref(n))
    return acc, n.value


def rand_sym3(delta, seed, sweep, site, hit, mu):
    out = np.empty((3, 3), dtype=np.complex128)
    L.orc_rand_sym3(delta, seed, sweep, site, hit, mu, _d(out))
    return out


def metropolis_sweep(U, nd, nc, dims, nhits, beta, delta, seed, sweep, visit=False):
    """In place on U (must be C-contiguous complex128). Returns (acc, naccepted[, visit])."""
    assert U.flags["C_CONTIGUOUS"] and U.dtype == np.complex128
    n = C.c_long()
    vis = np.full(int(np.prod(dims[:nd])), -1, dtype=np.int32) if visit else None
    acc = L.orc_metropolis_sweep(nd, nc, _i(dims), _d(U), nhits, beta, delta, seed, sweep,
                                 C.byref(n), vis.ctypes.data_as(_ip) if visit else None)
    return (acc, n.value, vis) if visit else (acc, n.value)


def ngen(nc):
    return L.orc_ngen(nc)


def mshape(nd, nc, dims):
    return (*dims[:nd], nd, ngen(nc))


def refresh_momenta(nd, nc, dims, seed, traj):
    P = np.empty(mshape(nd, nc, dims))
    L.orc_refresh_momenta(nd, nc, _i(dims), seed, traj, _d(P))
    return P


def force(U, nd, nc, dims, beta):
    F = np.empty(mshape(nd, nc, dims))
    L.orc_force(nd, nc, _i(dims), _d(np.ascontiguousarray(U)), beta, _d(F))
    return F


def update_momentum(P, F, nd, nc, dims, dtau):
    L.orc_update_momentum(nd, nc, _i(dims), _d(P), _d(np.ascontiguousarray(F)), dtau)


def update_gauge(U, P, nd, nc, dims, dtau):
    L.orc_update_gauge(nd, nc, _i(dims), _d(U), _d(np.ascontiguousarray(P)), dtau)


def restore_field(U, nd, nc, dims):
    L.orc_restore_field(nd, nc, _i(dims), _d(U))


def kinetic(P, nd, nc, dims):
    return L.orc_kinetic(nd, nc, _i(dims), _d(np.ascontiguousarray(P)))


def gauge_action(U, nd, nc, dims, beta):
    return L.orc_gauge_action(nd, nc, _i(dims), _d(np.ascontiguousarray(U)), beta)


def leapfrog(U, P, nd, nc, dims, beta, tau, nsteps):
    L.orc_leapfrog(nd, nc, _i(dims), _d(U), _d(P), beta, tau, nsteps)


def hmc_trajectory(U, nd, nc, dims, beta, tau, nsteps, seed, traj):
    dH = C.c_double()
    plaq = C.c_double()
    acc = L.orc_hmc_trajectory(nd, nc, _i(dims), _d(U), beta, tau, nsteps, seed, traj, C.byref(dH),
                               C.byref(plaq))
    return bool(acc), dH.value, plaq.value


def matmul(a, b):
    nc = a.shape[0]
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    c = np.empty_like(a)
    L.orc_matmul(nc, _d(a), _d(b), _d(c))
    return c


def restore_sun(a):
    a = np.array(a, dtype=np.complex128, order="C")
    L.orc_restore_sun(a.shape[0], _d(a))
    return a


def rand_sun(nc, delta, seed, sweep, site, mu, hit):
    r = np.empty((nc, nc), dtype=np.complex128)
    n = L.orc_rand_sun(nc, delta, seed, sweep, site, mu, hit, _d(r))
    return r, n


def exp_iq(nc, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty((nc, nc), dtype=np.complex128)
    L.orc_exp_iq(nc, _d(v), _d(out))
    return out


def defect_field(nd, dims, length, pos, cr):
    C_ = np.empty((*dims[:nd], nd))
    L.orc_defect_field(nd, _i(dims), length, _i(pos), cr, _d(C_))
    return C_


def plaquette_defect(U, Cf, nd, nc, dims, normalize=True):
    return L.orc_plaquette_defect(nd, nc, _i(dims), _d(np.ascontiguousarray(U)),
                                  _d(np.ascontiguousarray(Cf)), int(normalize))


L.orc_metropolis_sweep_defect_narrow.argtypes = L.orc_metropolis_sweep_defect.argtypes
L.orc_metropolis_sweep_defect_narrow.restype = C.c_double


def metropolis_sweep_defect(U, Cf, nd, nc, dims, nhits, beta, delta, seed, sweep, chain, narrow=False):
    """narrow: the FP32 links' draws (24-bit uniforms, one stream word each)."""
    n = C.c_long()
    fn = L.orc_metropolis_sweep_defect_narrow if narrow else L.orc_metropolis_sweep_defect
    acc = fn(nd, nc, _i(dims), _d(U), _d(np.ascontiguousarray(Cf)), nhits, beta, delta, seed, sweep, chain,
             C.byref(n))
    return acc, n.value
