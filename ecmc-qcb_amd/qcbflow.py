"""Python binding of libqcbflow (ctypes over the C ABI in include/qcbflow.h).

The class `GradientFlow` mirrors the reference's C++ class of the same name
(src/flow/gradient_flow.h:11-27) and the functions below mirror
namespace mpi::observables (src/observables/observables_mpi.h:8-19) and
read_ildg_clime / save_ildg_clime (src/io/ildg.h:17-20).  Every computation
runs in the HIP kernels of libqcbflow.so; there is no CPU fallback: if the
library is missing, `load()` raises.

Link arrays are float64 numpy arrays with V*72 doubles in the reference's
interior order (site x-fastest, mu, 3x3 row-major, re/im), e.g. of shape
(V, 4, 3, 3, 2).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqcbflow.so")

QCBF_OK = 0
QCBF_ERR_ARG = 1
QCBF_ERR_HIP = 2
QCBF_ERR_IO = 3
QCBF_ERR_FORMAT = 4
QCBF_ERR_OOM = 5
QCBF_ERR_STATE = 6

FP64 = 0
FP32_MIXED = 1

K_STAGE1, K_STAGE2, K_STAGE3, K_MEASURE, K_STAGE1_MEASURE = 0, 1, 2, 3, 4
TUNE_MAP_MODE, TUNE_MAP_WIN, TUNE_BLOCKS_PER_CU, TUNE_TILE_SHAPE = 0, 1, 2, 3
TUNE_FUSED_MEASURE = 4
TUNE_THREADS_PER_SITE = 5
TUNE_LDS_STAGE = 6


def tile_code(bx, by, bz, bt):
    """qcbf_tune code of a 4-D tile of extents bx x by x bz x bt (powers of two, product 64)."""
    lg = [int(v).bit_length() - 1 for v in (bx, by, bz, bt)]
    assert all(1 << l == v for l, v in zip(lg, (bx, by, bz, bt))) and sum(lg) == 6
    return lg[0] + 8 * lg[1] + 64 * lg[2] + 512 * lg[3]
K_COUNT = 5

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every exported symbol of include/qcbflow.h
SIGNATURES = {
    "qcbf_last_error": (ctypes.c_char_p, [_vp]),
    "qcbf_create": (ctypes.c_int, [_ip, ctypes.c_int, ctypes.POINTER(_vp)]),
    "qcbf_destroy": (ctypes.c_int, [_vp]),
    "qcbf_get_dims": (ctypes.c_int, [_vp, _ip, _ip]),
    "qcbf_set_links": (ctypes.c_int, [_vp, _dp]),
    "qcbf_get_links": (ctypes.c_int, [_vp, _dp]),
    "qcbf_hot_start": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "qcbf_copy_field": (ctypes.c_int, [_vp, _vp]),
    "qcbf_shift": (ctypes.c_int, [_vp, _vp]),
    "qcbf_batch_copy_from": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "qcbf_cold_start": (ctypes.c_int, [_vp]),
    "qcbf_heatbath": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_uint64]),
    "qcbf_load_ildg": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "qcbf_save_ildg": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "qcbf_rk3_step": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int]),
    "qcbf_compute_force": (ctypes.c_int, [_vp, ctypes.c_int]),
    "qcbf_compute_w1": (ctypes.c_int, [_vp, ctypes.c_double]),
    "qcbf_compute_w2": (ctypes.c_int, [_vp, ctypes.c_double]),
    "qcbf_replace_force_0": (ctypes.c_int, [_vp]),
    "qcbf_compute_w3": (ctypes.c_int, [_vp, ctypes.c_double]),
    "qcbf_get_force": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "qcbf_plaquette": (ctypes.c_int, [_vp, _dp]),
    "qcbf_q_e_clover": (ctypes.c_int, [_vp, _dp, _dp]),
    "qcbf_measure": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "qcbf_clover_site": (ctypes.c_int, [_vp, ctypes.c_long, ctypes.c_int, ctypes.c_int, _dp]),
    "qcbf_local_q_e": (ctypes.c_int, [_vp, ctypes.c_long, _dp, _dp]),
    "qcbf_flow_measure": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp,
                                         _dp]),
    "qcbf_exp_warnings": (ctypes.c_int, [_vp, _ip]),
    "qcbf_unitarity": (ctypes.c_int, [_vp, _dp, _dp]),
    "qcbf_synchronize": (ctypes.c_int, [_vp]),
    "qcbf_su3_exp": (ctypes.c_int, [_dp, ctypes.c_int, ctypes.c_double, _dp, _ip]),
    "qcbf_tune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_long]),
    "qcbf_get_tune": (ctypes.c_int, [_vp, ctypes.c_int, _lp]),
    "qcbf_autotune": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _lp, _lp]),
    "qcbf_batch_last_error": (ctypes.c_char_p, [_vp]),
    "qcbf_batch_create": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, ctypes.POINTER(_vp)]),
    "qcbf_batch_destroy": (ctypes.c_int, [_vp]),
    "qcbf_batch_size": (ctypes.c_int, [_vp, _ip, _ip]),
    "qcbf_batch_set_links": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "qcbf_batch_get_links": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "qcbf_batch_hot_start": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint64]),
    "qcbf_batch_rk3_step": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int]),
    "qcbf_batch_measure": (ctypes.c_int, [_vp, _dp]),
    "qcbf_batch_flow_measure": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                               _dp, _dp]),
    "qcbf_batch_exp_warnings": (ctypes.c_int, [_vp, _ip]),
    "qcbf_batch_synchronize": (ctypes.c_int, [_vp]),
    "qcbf_batch_tune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_long]),
    "qcbf_batch_get_tune": (ctypes.c_int, [_vp, ctypes.c_int, _lp]),
    "qcbf_batch_use_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "qcbf_batch_autotune": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, _lp, _lp, _lp]),
    "qcbf_batch_timing": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _lp]),
    "qcbf_timing_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "qcbf_timing_read": (ctypes.c_int, [_vp, _dp, _lp]),
    "qcbf_copy_bench": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_int, _dp]),
    "qcbf_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _ip]),
    "qcbf_measure_device": (ctypes.c_int, [_vp, _ip, ctypes.c_int, _vp, _vp]),
}

_lib = None


class QcbfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qcbflow error {code}: {msg}")
        self.code = code


def load():
    """Load libqcbflow.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (make -C ecmc-qcb_amd)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _check(rc, handle=None):
    if rc != QCBF_OK:
        msg = load().qcbf_last_error(handle)
        raise QcbfError(rc, msg.decode() if msg else "")


def _p(a):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def _dims(L):
    if np.isscalar(L):
        L = (L, L, L, L)
    if len(L) != 4:
        raise ValueError("lattice extents must be 4 integers (Lx, Ly, Lz, Lt)")
    return tuple(int(v) for v in L)


def device_info():
    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    _check(load().qcbf_device_info(name, 256, ctypes.byref(cus)))
    return name.value.decode(), cus.value


def su3_exp(Z, coeff=1.0):
    """exp_su3_luscher(Z, coeff) on the device for (n, 3, 3) complex Z.
    Returns (exp, flagged)."""
    Z = np.asarray(Z, dtype=np.complex128)
    shape = Z.shape
    Zr = np.ascontiguousarray(np.stack([Z.real, Z.imag], axis=-1).reshape(-1, 18))
    out = np.empty_like(Zr)
    flagged = ctypes.c_int()
    _check(load().qcbf_su3_exp(_p(Zr), Zr.shape[0], float(coeff), _p(out), ctypes.byref(flagged)))
    o = out.reshape(-1, 3, 3, 2)
    return (o[..., 0] + 1j * o[..., 1]).reshape(shape), bool(flagged.value)


class GradientFlow:
    """Device-resident Wilson flow of one SU(3) lattice (the reference's
    GradientFlow, src/flow/gradient_flow.h:11-27, holding its field_c on the
    GPU).  `epsilon` is the flow step used by rk3_step / topo_charge_flowed."""

    def __init__(self, L, epsilon=0.02, precision=FP64):
        self.L = _dims(L)
        self.V = int(np.prod(self.L))
        self.epsilon = float(epsilon)
        self.precision = precision
        h = _vp()
        _check(load().qcbf_create((ctypes.c_int * 4)(*self.L), int(precision), ctypes.byref(h)))
        self._h = h

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None):
            load().qcbf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, rc):
        _check(rc, self._h)

    # -- field in / out
    def copy(self, links):
        """GradientFlow::copy: upload V*72 doubles (reference order)."""
        links = np.ascontiguousarray(links, dtype=np.float64)
        if links.size != self.V * 72:
            raise ValueError(f"expected {self.V * 72} doubles, got {links.size}")
        self._c(load().qcbf_set_links(self._h, _p(links)))

    set_links = copy

    def field(self):
        """field_c read back, shape (V, 4, 3, 3, 2)."""
        out = np.empty((self.V, 4, 3, 3, 2))
        self._c(load().qcbf_get_links(self._h, _p(out)))
        return out

    def hot_start(self, seed):
        self._c(load().qcbf_hot_start(self._h, int(seed)))

    def copy_from(self, other):
        """field_c <- other's field, on the device (precisions may differ)."""
        self._c(load().qcbf_copy_field(self._h, other._h))

    def shift(self, s):
        """U_mu(x) <- U_mu(x - s) on the periodic lattice (mpi::shift::shift_field)."""
        arr = (ctypes.c_int * 4)(*[int(v) for v in s])
        self._c(load().qcbf_shift(self._h, arr))

    def cold_start(self):
        self._c(load().qcbf_cold_start(self._h))

    def heatbath(self, beta, n_sweeps, seed):
        self._c(load().qcbf_heatbath(self._h, float(beta), int(n_sweeps), int(seed)))

    def load_ildg(self, path):
        self._c(load().qcbf_load_ildg(self._h, os.fsencode(path)))

    def save_ildg(self, path):
        self._c(load().qcbf_save_ildg(self._h, os.fsencode(path)))

    # -- flow
    def rk3_step(self, n=1, eps=None):
        self._c(load().qcbf_rk3_step(self._h, self.epsilon if eps is None else float(eps), int(n)))

    def compute_force(self, i):
        self._c(load().qcbf_compute_force(self._h, int(i)))

    def compute_w1(self):
        self._c(load().qcbf_compute_w1(self._h, self.epsilon))

    def compute_w2(self):
        self._c(load().qcbf_compute_w2(self._h, self.epsilon))

    def replace_force_0(self):
        self._c(load().qcbf_replace_force_0(self._h))

    def compute_w3(self):
        self._c(load().qcbf_compute_w3(self._h, self.epsilon))

    def force(self, i):
        out = np.empty((self.V, 4, 3, 3, 2))
        self._c(load().qcbf_get_force(self._h, int(i), _p(out)))
        return out

    # -- observables
    def plaquette(self):
        p = ctypes.c_double()
        self._c(load().qcbf_plaquette(self._h, ctypes.byref(p)))
        return p.value

    def q_e_clover(self):
        q, e = ctypes.c_double(), ctypes.c_double()
        self._c(load().qcbf_q_e_clover(self._h, ctypes.byref(q), ctypes.byref(e)))
        return q.value, e.value

    def measure(self):
        """(P, E, Q) in one fused pass."""
        p, e, q = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._c(load().qcbf_measure(self._h, ctypes.byref(p), ctypes.byref(e), ctypes.byref(q)))
        return p.value, e.value, q.value

    def clover_site(self, site, mu, nu):
        """clover_site: F_munu at one site as a (3, 3) complex array."""
        out = np.empty(18)
        self._c(load().qcbf_clover_site(self._h, int(site), int(mu), int(nu), _p(out)))
        return out[0::2].reshape(3, 3) + 1j * out[1::2].reshape(3, 3)

    def local_q_e(self, site):
        """local_q_e_clover: (q(x), e(x)) at one site."""
        q, e = ctypes.c_double(), ctypes.c_double()
        self._c(load().qcbf_local_q_e(self._h, int(site), ctypes.byref(q), ctypes.byref(e)))
        return q.value, e.value

    def flow_measure(self, n_steps_gf, n_rk_steps, eps=None):
        """topo_charge_flowed cadence: returns (tQE of shape (n, 3), plaquette (n,))."""
        tqe = np.empty(3 * n_steps_gf)
        plaq = np.empty(n_steps_gf)
        self._c(load().qcbf_flow_measure(self._h, self.epsilon if eps is None else float(eps),
                                         int(n_steps_gf), int(n_rk_steps), _p(tqe), _p(plaq)))
        return tqe.reshape(-1, 3), plaq

    def exp_warnings(self):
        f = ctypes.c_int()
        self._c(load().qcbf_exp_warnings(self._h, ctypes.byref(f)))
        return f.value

    def unitarity(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._c(load().qcbf_unitarity(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def synchronize(self):
        self._c(load().qcbf_synchronize(self._h))

    # -- launch tuning
    def tune(self, key, value):
        self._c(load().qcbf_tune(self._h, int(key), int(value)))

    def get_tune(self, key):
        v = ctypes.c_long()
        self._c(load().qcbf_get_tune(self._h, int(key), ctypes.byref(v)))
        return v.value

    def autotune(self, trial_steps=3, eps=None):
        """Pick the fastest launch shape: tile order, tile shape, then threads per
        site (field restored afterwards); returns (mode, win), the rest through
        get_tune."""
        m, w = ctypes.c_long(), ctypes.c_long()
        self._c(load().qcbf_autotune(self._h, self.epsilon if eps is None else float(eps),
                                     int(trial_steps), ctypes.byref(m), ctypes.byref(w)))
        return m.value, w.value

    # -- measurement support
    def timing(self, on=True):
        """Time the kernel launches of this handle: True (every launch), False
        (off), or an int n >= 1 (every n-th launch of each kernel class)."""
        n = 1 if on is True else 0 if on is False else int(on)
        self._c(load().qcbf_timing_enable(self._h, n))

    def timing_read(self):
        ms = (ctypes.c_double * K_COUNT)()
        n = (ctypes.c_long * K_COUNT)()
        self._c(load().qcbf_timing_read(self._h, ms, n))
        return list(ms), list(n)

    def copy_bench(self, nbytes, reps=10):
        ms = ctypes.c_double()
        self._c(load().qcbf_copy_bench(self._h, int(nbytes), int(reps), ctypes.byref(ms)))
        return ms.value


class Ensemble:
    """Configurations of uneven volumes flowed together (one persistent kernel
    per RK3 stage over all of them; optional hipGraph replay of each step)."""

    def __init__(self, dims, epsilon=0.02, precision=FP64):
        self.dims = [_dims(d) for d in dims]
        self.n = len(self.dims)
        self.V = [int(np.prod(d)) for d in self.dims]
        self.epsilon = float(epsilon)
        flat = (ctypes.c_int * (4 * self.n))(*[v for d in self.dims for v in d])
        h = _vp()
        rc = load().qcbf_batch_create(self.n, flat, int(precision), ctypes.byref(h))
        if rc != QCBF_OK:
            raise QcbfError(rc, load().qcbf_batch_last_error(None).decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().qcbf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, rc):
        if rc != QCBF_OK:
            raise QcbfError(rc, load().qcbf_batch_last_error(self._h).decode())

    def copy(self, c, links):
        links = np.ascontiguousarray(links, dtype=np.float64)
        if links.size != self.V[c] * 72:
            raise ValueError("wrong number of doubles for configuration %d" % c)
        self._c(load().qcbf_batch_set_links(self._h, int(c), _p(links)))

    def field(self, c):
        out = np.empty((self.V[c], 4, 3, 3, 2))
        self._c(load().qcbf_batch_get_links(self._h, int(c), _p(out)))
        return out

    def hot_start(self, c, seed):
        self._c(load().qcbf_batch_hot_start(self._h, int(c), int(seed)))

    def copy_from(self, c, flow):
        """configuration c <- the field of a GradientFlow handle, on the device."""
        self._c(load().qcbf_batch_copy_from(self._h, int(c), flow._h))

    def rk3_step(self, n=1, eps=None):
        self._c(load().qcbf_batch_rk3_step(self._h, self.epsilon if eps is None else float(eps),
                                           int(n)))

    def measure(self):
        out = np.empty(3 * self.n)
        self._c(load().qcbf_batch_measure(self._h, _p(out)))
        return out.reshape(self.n, 3)

    def flow_measure(self, n_steps_gf, n_rk_steps, eps=None):
        """Returns (tQE of shape (n_cfg, n_steps_gf, 3), plaquettes (n_cfg, n_steps_gf))."""
        tqe = np.empty(3 * n_steps_gf * self.n)
        plaq = np.empty(n_steps_gf * self.n)
        self._c(load().qcbf_batch_flow_measure(self._h, self.epsilon if eps is None else float(eps),
                                               int(n_steps_gf), int(n_rk_steps), _p(tqe),
                                               _p(plaq)))
        return tqe.reshape(self.n, n_steps_gf, 3), plaq.reshape(self.n, n_steps_gf)

    def exp_warnings(self):
        f = ctypes.c_int()
        self._c(load().qcbf_batch_exp_warnings(self._h, ctypes.byref(f)))
        return f.value

    def synchronize(self):
        self._c(load().qcbf_batch_synchronize(self._h))

    def tune(self, key, value):
        self._c(load().qcbf_batch_tune(self._h, int(key), int(value)))

    def get_tune(self, key):
        v = ctypes.c_long()
        self._c(load().qcbf_batch_get_tune(self._h, int(key), ctypes.byref(v)))
        return v.value

    def use_graph(self, on=True):
        self._c(load().qcbf_batch_use_graph(self._h, 1 if on else 0))

    def autotune(self, trial_steps=2, eps=None):
        m, w, b = ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
        self._c(load().qcbf_batch_autotune(self._h, self.epsilon if eps is None else float(eps),
                                           int(trial_steps), ctypes.byref(m), ctypes.byref(w),
                                           ctypes.byref(b)))
        return m.value, w.value, b.value

    def timing(self, on=True):
        """Returns (ms, rk3 steps) of flow_measure calls since the previous call."""
        ms, n = ctypes.c_double(), ctypes.c_long()
        self._c(load().qcbf_batch_timing(self._h, 1 if on else 0, ctypes.byref(ms),
                                         ctypes.byref(n)))
        return ms.value, n.value


# ---- mpi::observables mirrors (single device; the topology argument of the
# reference is dropped because one GPU holds the whole lattice) -------------
def mean_plaquette_global(flow):
    return flow.plaquette()


def topo_q_e_clover_global(flow):
    return flow.q_e_clover()


def topo_charge_flowed(links, flow, N_steps_gf, N_rk_steps):
    """Copies `links` into the flow (gf.copy) and returns the flat tQE vector
    of 3*N_steps_gf doubles, as observables_mpi.cpp:159-186."""
    flow.copy(links)
    tqe, _ = flow.flow_measure(N_steps_gf, N_rk_steps)
    return tqe.reshape(-1)
