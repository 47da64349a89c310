"""ctypes wrapper of the FP64 CPU oracle (qcb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, always as the result the GPU path is compared
with, never as the code measured or shipped.  Fields are numpy float64 arrays of shape (V, 4, 3, 3, 2) or any
contiguous array with V*72 doubles, in the reference's interior order:
site x-fastest, mu, row-major 3x3, (re, im).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libqcboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile the oracle with its Makefile (gcc -O2 -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE] + (["-B"] if force else []))
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_exp_su3.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_proj_lie_su3.argtypes = [_dp, ctypes.c_int, _dp]
        L.orc_proj_su3.argtypes = [_dp, ctypes.c_int, _dp]
        L.orc_staple.argtypes = [_dp, _ip, ctypes.c_long, ctypes.c_int, _dp]
        L.orc_compute_force.argtypes = [_dp, _ip, _dp]
        L.orc_rk3.argtypes = [_dp, _ip, ctypes.c_double, ctypes.c_int]
        L.orc_plaquette.argtypes = [_dp, _ip]
        L.orc_plaquette.restype = ctypes.c_double
        L.orc_clover.argtypes = [_dp, _ip, ctypes.c_long, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_q_e_clover.argtypes = [_dp, _ip, _dp, _dp]
        L.orc_local_q_e.argtypes = [_dp, _ip, ctypes.c_long, _dp]
        L.orc_flow_measure.argtypes = [_dp, _ip, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       _dp, _dp, _dp]
        L.orc_time_flow.argtypes = [_dp, _ip, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_time_flow.restype = ctypes.c_double
        L.orc_hot_start.argtypes = [_dp, _ip, ctypes.c_uint64]
        L.orc_cold_start.argtypes = [_dp, _ip]
        L.orc_heatbath.argtypes = [_dp, _ip, ctypes.c_double, ctypes.c_int, ctypes.c_uint64]
        L.orc_gauge_transform.argtypes = [_dp, _ip, ctypes.c_uint64, _dp]
        L.orc_exp_warnings.restype = ctypes.c_int
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _dims(L):
    if np.isscalar(L):
        L = (L, L, L, L)
    return (ctypes.c_int * 4)(*[int(v) for v in L])


def volume(L):
    if np.isscalar(L):
        return int(L) ** 4
    return int(np.prod(L))


def new_field(L):
    return np.zeros((volume(L), 4, 3, 3, 2), dtype=np.float64)


def to_complex(a):
    """(..., 3, 3, 2) float64 -> (..., 3, 3) complex128"""
    return a[..., 0] + 1j * a[..., 1]


def from_complex(c):
    c = np.ascontiguousarray(c, dtype=np.complex128)
    return np.ascontiguousarray(np.stack([c.real, c.imag], axis=-1))


# --- su(3) ------------------------------------------------------------------
def exp_su3(Z, coeff=1.0):
    """Z: (n, 3, 3) complex -> exp_su3_luscher(Z, coeff), (n, 3, 3) complex"""
    Zr = from_complex(np.asarray(Z).reshape(-1, 3, 3))
    out = np.empty_like(Zr)
    lib().orc_exp_su3(_p(Zr), Zr.shape[0], float(coeff), _p(out))
    return to_complex(out).reshape(np.shape(Z))


def proj_lie_su3(A):
    Ar = from_complex(np.asarray(A).reshape(-1, 3, 3))
    out = np.empty_like(Ar)
    lib().orc_proj_lie_su3(_p(Ar), Ar.shape[0], _p(out))
    return to_complex(out).reshape(np.shape(A))


def proj_su3(A):
    Ar = from_complex(np.asarray(A).reshape(-1, 3, 3))
    out = np.empty_like(Ar)
    lib().orc_proj_su3(_p(Ar), Ar.shape[0], _p(out))
    return to_complex(out).reshape(np.shape(A))


def exp_warnings():
    return lib().orc_exp_warnings()


# --- fields -----------------------------------------------------------------
def hot_start(L, seed):
    f = new_field(L)
    lib().orc_hot_start(_p(f), _dims(L), int(seed))
    return f


def cold_start(L):
    f = new_field(L)
    lib().orc_cold_start(_p(f), _dims(L))
    return f


def heatbath(f, L, beta, n_sweeps, seed):
    lib().orc_heatbath(_p(f), _dims(L), float(beta), int(n_sweeps), int(seed))
    return f


def thermalised(L, beta, seed, n_sweeps):
    """Haar hot start from `seed`, then `n_sweeps` heatbath sweeps."""
    f = hot_start(L, seed)
    return heatbath(f, L, beta, n_sweeps, seed + 0x5EED)


def gauge_transform(f, L, seed):
    V = volume(L)
    g = np.empty((V, 3, 3, 2))
    lib().orc_gauge_transform(_p(f), _dims(L), int(seed), _p(g))
    return to_complex(g)


def staple(f, L, site, mu):
    out = np.empty((3, 3, 2))
    lib().orc_staple(_p(f), _dims(L), int(site), int(mu), _p(out))
    return to_complex(out)


def compute_force(f, L):
    out = np.empty_like(f)
    lib().orc_compute_force(_p(f), _dims(L), _p(out))
    return out


def rk3(f, L, eps, n):
    lib().orc_rk3(_p(f), _dims(L), float(eps), int(n))
    return f


def plaquette(f, L):
    return lib().orc_plaquette(_p(f), _dims(L))


def clover(f, L, site, mu, nu):
    out = np.empty((3, 3, 2))
    lib().orc_clover(_p(f), _dims(L), int(site), int(mu), int(nu), _p(out))
    return to_complex(out)


def local_q_e(f, L, site):
    """local_q_e_clover at one site: (q(x), e(x))."""
    out = np.empty(2)
    lib().orc_local_q_e(_p(f), _dims(L), int(site), _p(out))
    return float(out[0]), float(out[1])


def q_e_clover(f, L):
    Q = ctypes.c_double()
    E = ctypes.c_double()
    lib().orc_q_e_clover(_p(f), _dims(L), ctypes.byref(Q), ctypes.byref(E))
    return Q.value, E.value


def flow_measure(f, L, eps, n_steps_gf, n_rk_steps, want_field=False):
    """topo_charge_flowed cadence.  Returns (tQE (n,3), plaquette (n,), field or None)."""
    tqe = np.empty(3 * n_steps_gf)
    plaq = np.empty(n_steps_gf)
    out = np.empty_like(f) if want_field else None
    lib().orc_flow_measure(_p(f), _dims(L), float(eps), int(n_steps_gf), int(n_rk_steps),
                           _p(tqe), _p(plaq), _p(out) if want_field else None)
    return tqe.reshape(-1, 3), plaq, out


def time_flow(f, L, eps, n_steps, measure=True):
    return lib().orc_time_flow(_p(f), _dims(L), float(eps), int(n_steps), int(bool(measure)))


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return lib().orc_num_threads()
