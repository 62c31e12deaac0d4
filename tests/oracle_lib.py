"""ctypes bindings of the test oracles (test infrastructure only).

  Oracle      oracle/liboracle.so                (C restatement of the reference CPU path)
  RefLib      oracle/_ref/libdmfe_ref_strict.so  (the reference's own sources + ref_harness.cpp)

Plus the seeded synthetic history used by the parity tests and the benchmark.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libdmfe_ref_strict.so")
REF_FAST_PATH = os.path.join(ROOT, "oracle", "_ref", "libdmfe_ref_fast.so")

_dp = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def build_oracle():
    """Builds oracle/liboracle.so (gcc) if missing."""
    if not os.path.exists(ORACLE_PATH):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True, capture_output=True)
    return ORACLE_PATH


_SETUP_ARGT = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int64, C.c_int, C.c_int, C.c_double,
               C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]


def _setup_args(grid, p, p2, lam, T0, Gamma, dt_min, delta_max, log_resp):
    return [grid.N, _p(grid.theta), _p(grid.integ), _p(grid.phi2), _p(grid.inds[0]), _p(grid.inds[1]),
            _p(grid.inds[2]), _p(grid.weights[0]), _p(grid.weights[1]), _p(grid.weights[2]), grid.m, p, p2, lam,
            T0, Gamma, dt_min, delta_max, log_resp]


class _Common:
    """Shared methods of the oracle and the reference harness (same call surface)."""

    def _hist(self, T1):
        N = self.N
        a = [np.zeros(T1) for _ in range(4)] + [np.zeros(T1 * N) for _ in range(4)]
        return a

    def state(self):
        T1 = self.T1
        t1, rho, r, dr, QK, QR, dQK, dQR = self._hist(T1)
        self._get_history(t1, rho, r, dr, QK, QR, dQK, dQR)
        return dict(t1grid=t1, dtratio=rho, rvec=r, drvec=dr, QK=QK, QR=QR, dQK=dQK, dQR=dQR)

    def fields(self):
        P = self.N * self.N
        out = np.zeros(15 * P + self.N)
        self._get_fields(out)
        names = ("QKA1", "QRA1", "QKA2", "QRA2", "QKB1", "QRB1", "QKB2", "QRB2",
                 "SigmaKA1", "SigmaRA1", "SigmaRA2", "SigmaKB1", "SigmaRB1", "SigmaKB2", "SigmaRB2")
        d = {n: out[i * P:(i + 1) * P] for i, n in enumerate(names)}
        d["rInt"] = out[15 * P:]
        return d


class Oracle(_Common):
    """C restatement (oracle/dmfe_oracle.c)."""

    def __init__(self, grid, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8, dt_min=1e-5, delta_max=1e-10, log_resp=0):
        L = C.CDLL(build_oracle())
        self.L = L
        L.or_create.argtypes = _SETUP_ARGT
        L.or_create.restype = C.c_void_p
        for f in ("or_rstep", "or_update", "or_loop_iteration", "or_energy", "or_drstep2", "or_flambda"):
            getattr(L, f).restype = C.c_double
        L.or_T1.restype = C.c_int64
        L.or_T1.argtypes = [C.c_void_p]
        L.or_energy.argtypes = [C.c_void_p, C.c_double]
        L.or_flambda.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.or_drstep2.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_double]
        L.or_set_control.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
        L.or_sparsify.argtypes = [C.c_void_p, C.c_double]
        L.or_interpolate.argtypes = [C.c_void_p, C.c_int]
        L.or_set_history.argtypes = [C.c_void_p, C.c_int64] + [_dp] * 8
        L.or_bsearch.argtypes = [C.c_int64, _dp, C.c_int64, _dp, _dp]
        L.or_isearch.argtypes = [C.c_int64, _dp, C.c_int64, C.c_int64, _dp, _dp, _dp]
        L.or_serk2_coeffs.argtypes = [C.c_int, _dp]
        L.or_set_serk2.argtypes = [C.c_void_p, C.c_int]
        L.or_set_maint_every.argtypes = [C.c_void_p, C.c_long]
        L.or_set_rollback.argtypes = [C.c_void_p, C.c_int]
        L.or_set_tail_fit.argtypes = [C.c_void_p, C.c_int]
        L.or_set_gpu_variants.argtypes = [C.c_void_p, C.c_int]
        L.or_tail_fit_applied.argtypes = [C.c_void_p]
        L.or_tail_fit_applied.restype = C.c_long
        L.or_tail_fit_alpha.argtypes = [C.c_void_p]
        L.or_tail_fit_alpha.restype = C.c_double
        for f in ("or_destroy", "or_init_quench", "or_update", "or_loop_iteration", "or_rstep", "or_rhs",
                  "or_get_fields", "or_get_history", "or_get_control", "or_set_pos_old", "or_get_pos_old"):
            if getattr(L, f).argtypes is None:
                getattr(L, f).argtypes = None
        self.N = grid.N
        self.h = C.c_void_p(L.or_create(*_setup_args(grid, p, p2, lam, T0, Gamma, dt_min, delta_max, log_resp)))

    def close(self):
        if self.h:
            self.L.or_destroy(self.h)
            self.h = None

    def set_serk2(self, on):
        self.L.or_set_serk2(self.h, 1 if on else 0)

    def set_maint_every(self, every):
        self.L.or_set_maint_every(self.h, every)

    def set_rollback(self, on):
        self.L.or_set_rollback(self.h, 1 if on else 0)

    def set_tail_fit(self, on):
        self.L.or_set_tail_fit(self.h, 1 if on else 0)

    def set_gpu_variants(self, on):
        self.L.or_set_gpu_variants(self.h, 1 if on else 0)

    def tail_fit_stats(self):
        return self.L.or_tail_fit_applied(self.h), self.L.or_tail_fit_alpha(self.h)

    def serk2_coeffs(self, q):
        out = np.zeros(2 * q + 1)
        self.L.or_serk2_coeffs(q, _p(out))
        return out

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def T1(self):
        return self.L.or_T1(self.h)

    def init_quench(self):
        self.L.or_init_quench(self.h)

    def set_history(self, t1grid, dtratio, rvec, drvec, QK, QR, dQK, dQR):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (t1grid, dtratio, rvec, drvec, QK, QR, dQK, dQR)]
        self.L.or_set_history(self.h, len(arrs[0]), *[_p(a) for a in arrs])

    def _get_history(self, *a):
        self.L.or_get_history(self.h, *[_p(x) for x in a])

    def _get_fields(self, out):
        self.L.or_get_fields(self.h, _p(out))

    def set_control(self, delta_t, delta, delta_old, loop, method):
        self.L.or_set_control(self.h, delta_t, delta, delta_old, loop, method)

    def control(self):
        dt, de, do, lp, me = C.c_double(), C.c_double(), C.c_double(), C.c_int(), C.c_int()
        self.L.or_get_control(self.h, C.byref(dt), C.byref(de), C.byref(do), C.byref(lp), C.byref(me))
        return dict(delta_t=dt.value, delta=de.value, delta_old=do.value, loop=lp.value, method=me.value)

    def set_pos_old(self, b1, b2):
        self.L.or_set_pos_old(self.h, _p(np.ascontiguousarray(b1)), _p(np.ascontiguousarray(b2)))

    def positions(self):
        b1, b2 = np.zeros(self.N), np.zeros(self.N * self.N)
        self.L.or_get_pos_old(self.h, _p(b1), _p(b2))
        return b1, b2

    def interpolate(self, mode=0):
        self.L.or_interpolate(self.h, mode)

    def rhs(self):
        hK, hR = np.zeros(self.N), np.zeros(self.N)
        self.L.or_rhs(self.h, _p(hK), _p(hR))
        return hK, hR

    def rstep(self):
        return self.L.or_rstep(self.h)

    def drstep2(self, qK, qR, dqK, dqR, t):
        return self.L.or_drstep2(self.h, _p(qK), _p(qR), _p(dqK), _p(dqR), t)

    def update(self):
        return self.L.or_update(self.h)

    def loop_iteration(self):
        return self.L.or_loop_iteration(self.h)

    def energy(self, t):
        return self.L.or_energy(self.h, t)

    def sparsify(self, thr):
        self.L.or_sparsify(self.h, thr)

    def flambda(self, q, d):
        return self.L.or_flambda(self.h, q, d)

    def bsearch(self, lst, el):
        lst = np.ascontiguousarray(lst, np.float64)
        el = np.ascontiguousarray(el, np.float64)
        out = np.zeros(len(el))
        self.L.or_bsearch(len(lst), _p(lst), len(el), _p(el), _p(out))
        return out

    def isearch(self, lst, el, inits, rowlen):
        lst = np.ascontiguousarray(lst, np.float64)
        el = np.ascontiguousarray(el, np.float64)
        inits = np.ascontiguousarray(inits, np.float64)
        out = np.zeros(len(el))
        self.L.or_isearch(len(lst), _p(lst), len(el), rowlen, _p(el), _p(inits), _p(out))
        return out


class RefLib(_Common):
    """The reference's own CPU sources, built by oracle/Makefile (process-global state)."""

    def __init__(self, grid, p=3, p2=4, lam=0.5, T0=1e50, Gamma=0.8, dt_min=1e-5, delta_max=1e-10, log_resp=0,
                 path=REF_PATH):
        os.environ.setdefault("OMP_NUM_THREADS", "1")
        L = C.CDLL(path)
        self.L = L
        L.ref_setup.argtypes = _SETUP_ARGT
        for f in ("ref_rstep", "ref_update", "ref_loop_iteration", "ref_energy", "ref_drstep2"):
            getattr(L, f).restype = C.c_double
        L.ref_T1.restype = C.c_int64
        L.ref_energy.argtypes = [C.c_double]
        L.ref_drstep2.argtypes = [_dp, _dp, _dp, _dp, C.c_double]
        L.ref_set_control.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
        L.ref_sparsify.argtypes = [C.c_double]
        L.ref_set_history.argtypes = [C.c_int64] + [_dp] * 8
        L.ref_bsearch.argtypes = [C.c_int64, _dp, C.c_int64, _dp, _dp]
        L.ref_isearch.argtypes = [C.c_int64, _dp, C.c_int64, _dp, _dp, _dp]
        self.N = grid.N
        L.ref_setup(*_setup_args(grid, p, p2, lam, T0, Gamma, dt_min, delta_max, log_resp))

    @property
    def T1(self):
        return self.L.ref_T1()

    def last_qk0(self):
        """QK(t,0) of the last history row."""
        self.L.ref_last_qk0.restype = C.c_double
        return self.L.ref_last_qk0()

    def t1grid(self):
        out = np.zeros(self.T1)
        self.L.ref_get_t1grid.argtypes = [_dp]
        self.L.ref_get_t1grid(_p(out))
        return out

    def init_quench(self):
        self.L.ref_init_quench()

    def set_history(self, t1grid, dtratio, rvec, drvec, QK, QR, dQK, dQR):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (t1grid, dtratio, rvec, drvec, QK, QR, dQK, dQR)]
        self.L.ref_set_history(len(arrs[0]), *[_p(a) for a in arrs])

    def _get_history(self, *a):
        self.L.ref_get_history(*[_p(x) for x in a])

    def _get_fields(self, out):
        self.L.ref_get_fields(_p(out))

    def set_control(self, delta_t, delta, delta_old, loop, method):
        self.L.ref_set_control(delta_t, delta, delta_old, loop, method)

    def control(self):
        dt, de, do, lp, me = C.c_double(), C.c_double(), C.c_double(), C.c_int(), C.c_int()
        self.L.ref_get_control(C.byref(dt), C.byref(de), C.byref(do), C.byref(lp), C.byref(me))
        return dict(delta_t=dt.value, delta=de.value, delta_old=do.value, loop=lp.value, method=me.value)

    def set_pos_old(self, b1, b2):
        self.L.ref_set_pos_old(_p(np.ascontiguousarray(b1)), _p(np.ascontiguousarray(b2)))

    def positions(self):
        b1, b2 = np.zeros(self.N), np.zeros(self.N * self.N)
        self.L.ref_get_pos_old(_p(b1), _p(b2))
        return b1, b2

    def interpolate(self, mode=0):
        self.L.ref_interpolate(mode)

    def rhs(self):
        hK, hR = np.zeros(self.N), np.zeros(self.N)
        self.L.ref_rhs(_p(hK), _p(hR))
        return hK, hR

    def rstep(self):
        return self.L.ref_rstep()

    def drstep2(self, qK, qR, dqK, dqR, t):
        return self.L.ref_drstep2(_p(qK), _p(qR), _p(dqK), _p(dqR), t)

    def update(self):
        return self.L.ref_update()

    def loop_iteration(self):
        return self.L.ref_loop_iteration()

    def energy(self, t):
        return self.L.ref_energy(t)

    def sparsify(self, thr):
        self.L.ref_sparsify(thr)

    def bsearch(self, lst, el):
        lst = np.ascontiguousarray(lst, np.float64)
        el = np.ascontiguousarray(el, np.float64)
        out = np.zeros(len(el))
        self.L.ref_bsearch(len(lst), _p(lst), len(el), _p(el), _p(out))
        return out

    def isearch(self, lst, el, inits, rowlen):
        assert rowlen == self.N
        lst = np.ascontiguousarray(lst, np.float64)
        el = np.ascontiguousarray(el, np.float64)
        inits = np.ascontiguousarray(inits, np.float64)
        out = np.zeros(len(el))
        self.L.ref_isearch(len(lst), _p(lst), len(el), _p(el), _p(inits), _p(out))
        return out


def ref_available():
    return os.path.exists(REF_PATH)


from dynamite_amd.workloads import SEED, synthetic_history  # noqa: E402,F401  (re-export for the tests)
