"""dynamite_amd - MI355X (gfx950) per-RK-step renormalized DMFT update.

Thin ctypes binding of dynamite_amd/libdmfe.so (C ABI in include/dmfe.h) used by
the tests, smoke() and bench.py.  There is no CPU fallback: every compute call
goes to the HIP library, and importing a missing library raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdmfe.so")

DMFE_OK = 0
INTERP_POLY, INTERP_RATIONAL, INTERP_BSPLINE = 0, 1, 2
RK54, SSPRK104 = 1, 2
POS_BSEARCH, POS_ISEARCH, POS_SAME = 0, 1, 2

FIELD_NAMES = ("QKA1", "QRA1", "QKA2", "QRA2", "QKB1", "QRB1", "QKB2", "QRB2",
               "SigmaKA1", "SigmaRA1", "SigmaRA2", "SigmaKB1", "SigmaRB1", "SigmaKB2", "SigmaRB2")

_lib = None


class DmfeError(RuntimeError):
    pass


class GridSpec(C.Structure):
    _fields_ = [("N", C.c_int64), ("Tmax", C.c_double), ("alpha", C.c_double), ("delta", C.c_double),
                ("spline_order", C.c_int32), ("method", C.c_int32), ("order", C.c_int32),
                ("fh_stencil", C.c_int32)]


class Grid(C.Structure):
    _fields_ = [("N", C.c_int64), ("m", C.c_int64), ("theta", C.c_void_p), ("integ", C.c_void_p),
                ("phi2", C.c_void_p), ("inds", C.c_void_p * 3), ("weights", C.c_void_p * 3)]


class Params(C.Structure):
    _fields_ = [("p", C.c_int32), ("p2", C.c_int32), ("lambda_", C.c_double), ("T0", C.c_double),
                ("Gamma", C.c_double), ("delta_t_min", C.c_double), ("delta_max", C.c_double),
                ("rmax", C.c_double * 11), ("log_response", C.c_int32), ("maint_every", C.c_int32)]


class RhsShape(C.Structure):
    """dmfe_rhs_shape: G lanes per B2 entry, split column chunks per row."""
    _fields_ = [("G", C.c_int32), ("split", C.c_int32)]

    def __repr__(self):
        return f"RhsShape(G={self.G}, split={self.split})"

    def astuple(self):
        return (self.G, self.split)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib():
    """Loads libdmfe.so (build it with `python -m dynamite_amd.build`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DmfeError(f"{LIB_PATH} is missing: run `python -m dynamite_amd.build` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        L.dmfe_last_error.restype = C.c_char_p
        L.dmfe_version.restype = C.c_char_p
        L.dmfe_grid_stencil_width.restype = C.c_int64
        L.dmfe_grid_stencil_width.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != DMFE_OK:
        raise DmfeError(f"{what} failed ({rc}): {lib().dmfe_last_error().decode()}")


class GridPackage:
    """In-memory Grid_data/<L> package (theta, int.dat, phi1/phi2, pos*, inds*, weights*)."""

    def __init__(self, N, m, theta, integ, phi2, inds, weights, phi1=None, pos=None, spec=None):
        self.N, self.m = int(N), int(m)
        self.theta = np.ascontiguousarray(theta, np.float64)
        self.integ = np.ascontiguousarray(integ, np.float64)
        self.phi2 = np.ascontiguousarray(phi2, np.float64)
        self.inds = np.ascontiguousarray(inds, np.int64).reshape(3, self.N * self.N)
        self.weights = np.ascontiguousarray(weights, np.float64).reshape(3, self.N * self.N * self.m)
        self.phi1 = phi1
        self.pos = pos
        self.spec = spec

    def as_struct(self):
        g = Grid()
        g.N, g.m = self.N, self.m
        g.theta, g.integ, g.phi2 = _ptr(self.theta), _ptr(self.integ), _ptr(self.phi2)
        for f in range(3):
            g.inds[f] = self.inds[f].ctypes.data
            g.weights[f] = self.weights[f].ctypes.data
        return g

    def params_text(self):
        s = self.spec
        if s is None:
            return None
        meth = ("poly", "rational", "bspline")[s.method]
        txt = ("# DYNAMITE grid generation parameters\n"
               f"len={s.N}\nTmax={s.Tmax:.17e}\nspline_order={s.spline_order}\n"
               f"interp_method={meth}\ninterp_order={s.order}\n")
        if s.method == INTERP_RATIONAL:
            txt += f"fh_stencil={self.m}\n"
        txt += f"alpha={s.alpha:.17e}\ndelta={s.delta:.17e}\nsubdir={s.N}\n"
        return txt

    def write(self, directory):
        os.makedirs(directory, exist_ok=True)
        pos = self.pos if self.pos is not None else (None, None, None)
        txt = self.params_text()
        rc = lib().dmfe_grid_write_dir(directory.encode(), C.c_int64(self.N), C.c_int64(self.m), _ptr(self.theta),
                                       _ptr(self.integ), _ptr(self.phi1), _ptr(self.phi2), _ptr(pos[0]), _ptr(pos[1]),
                                       _ptr(pos[2]), _ptr(self.inds), _ptr(self.weights),
                                       txt.encode() if txt else None)
        _check(rc, "dmfe_grid_write_dir")


def stencil_width(N, method=INTERP_POLY, order=9, fh_stencil=0):
    return int(lib().dmfe_grid_stencil_width(N, method, order, fh_stencil))


def generate_grid(N, Tmax=1e5, spline_order=5, method=INTERP_POLY, order=9, fh_stencil=0, alpha=0.0, delta=0.0,
                  with_pos=False, device=False):
    """Grid package generator (the reference's `grid` subcommand, grid_cli.cpp:35-233).
    device=True forms the dense B-spline weights on the current HIP device
    (dmfe_grid_generate_device)."""
    spec = GridSpec(N, Tmax, alpha, delta, spline_order, method, order, fh_stencil)
    m = stencil_width(N, method, order, fh_stencil)
    P = N * N
    theta, integ = np.zeros(N), np.zeros(N)
    phi1, phi2 = np.zeros(P), np.zeros(P)
    pos = (np.zeros(P), np.zeros(P), np.zeros(P)) if with_pos else (None, None, None)
    inds = np.zeros(3 * P, np.int64)
    weights = np.zeros(3 * P * m)
    gen = lib().dmfe_grid_generate_device if device else lib().dmfe_grid_generate
    rc = gen(C.byref(spec), _ptr(theta), _ptr(integ), _ptr(phi1), _ptr(phi2), _ptr(pos[0]), _ptr(pos[1]),
             _ptr(pos[2]), _ptr(inds), _ptr(weights))
    _check(rc, "dmfe_grid_generate_device" if device else "dmfe_grid_generate")
    return GridPackage(N, m, theta, integ, phi2, inds, weights, phi1=phi1, pos=pos if with_pos else None, spec=spec)


def read_grid(directory):
    """Loads a Grid_data/<L> directory (import(), io_input.cpp:90-110)."""
    N, m = C.c_int64(), C.c_int64()
    _check(lib().dmfe_grid_read_info(directory.encode(), C.byref(N), C.byref(m)), "dmfe_grid_read_info")
    N, m = N.value, m.value
    P = N * N
    theta, integ, phi2 = np.zeros(N), np.zeros(N), np.zeros(P)
    inds, weights = np.zeros(3 * P, np.int64), np.zeros(3 * P * m)
    rc = lib().dmfe_grid_read_dir(directory.encode(), C.c_int64(N), C.c_int64(m), _ptr(theta), _ptr(integ),
                                  _ptr(phi2), _ptr(inds), _ptr(weights))
    _check(rc, "dmfe_grid_read_dir")
    return GridPackage(N, m, theta, integ, phi2, inds, weights)


def default_params(**kw):
    p = Params()
    lib().dmfe_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, "lambda_" if k == "lam" or k == "lambda_" else k, v)
    return p


class Context:
    """One quench instance on the current HIP device (dmfe_ctx)."""

    def __init__(self, grid, params, rows_hint=256):
        self.grid = grid
        self.N = grid.N
        self._h = C.c_void_p()
        self._gs = grid.as_struct()
        _check(lib().dmfe_ctx_create(C.byref(self._gs), C.byref(params), C.c_int64(rows_hint), C.byref(self._h)),
               "dmfe_ctx_create")

    def close(self):
        if self._h:
            lib().dmfe_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state -------------------------------------------------------------------
    def init_quench(self):
        _check(lib().dmfe_init_quench(self._h), "dmfe_init_quench")

    @property
    def T1(self):
        t = C.c_int64()
        _check(lib().dmfe_get_T1(self._h, C.byref(t)), "dmfe_get_T1")
        return t.value

    def set_state(self, t1grid, QK, QR, dQK, dQR, rvec, drvec, dtratio=None):
        arrs = [np.ascontiguousarray(a, np.float64).ravel() for a in (t1grid, QK, QR, dQK, dQR, rvec, drvec)]
        dtr = None if dtratio is None else np.ascontiguousarray(dtratio, np.float64).ravel()
        T1 = len(arrs[0])
        # the C ABI reads T1 * N history values and T1 per-row values from each array
        # (time_steps.cpp:55,136 reject vectors of unequal sizes the same way)
        sizes = [a.size for a in arrs] + ([] if dtr is None else [dtr.size])
        want = [T1] + [T1 * self.N] * 4 + [T1] * 2 + ([] if dtr is None else [T1])
        if sizes != want:
            raise ValueError(f"set_state: array sizes {sizes} do not match T1={T1}, N={self.N} (expected {want})")
        _check(lib().dmfe_set_state(self._h, C.c_int64(T1), *[_ptr(a) for a in arrs], _ptr(dtr)), "dmfe_set_state")

    def get_state(self):
        T1, N = self.T1, self.N
        out = {k: np.zeros(T1) for k in ("t1grid", "rvec", "drvec", "dtratio")}
        for k in ("QK", "QR", "dQK", "dQR"):
            out[k] = np.zeros(T1 * N)
        _check(lib().dmfe_get_state(self._h, _ptr(out["t1grid"]), _ptr(out["QK"]), _ptr(out["QR"]), _ptr(out["dQK"]),
                                    _ptr(out["dQR"]), _ptr(out["rvec"]), _ptr(out["drvec"]), _ptr(out["dtratio"])),
               "dmfe_get_state")
        return out

    def set_control(self, delta_t, delta, delta_old, loop, method):
        _check(lib().dmfe_set_control(self._h, C.c_double(delta_t), C.c_double(delta), C.c_double(delta_old),
                                      C.c_int64(loop), C.c_int(method)), "dmfe_set_control")

    def get_control(self):
        dt, de, do, lp, me = C.c_double(), C.c_double(), C.c_double(), C.c_int64(), C.c_int()
        _check(lib().dmfe_get_control(self._h, C.byref(dt), C.byref(de), C.byref(do), C.byref(lp), C.byref(me)),
               "dmfe_get_control")
        return dict(delta_t=dt.value, delta=de.value, delta_old=do.value, loop=lp.value, method=me.value)

    # -- hot path ----------------------------------------------------------------
    def interpolate(self, mode=POS_BSEARCH):
        _check(lib().dmfe_interpolate(self._h, C.c_int(mode)), "dmfe_interpolate")

    def positions(self):
        b1, b2 = np.zeros(self.N), np.zeros(self.N * self.N)
        _check(lib().dmfe_get_positions(self._h, _ptr(b1), _ptr(b2)), "dmfe_get_positions")
        return b1, b2

    def set_positions(self, b1, b2):
        b1 = np.ascontiguousarray(b1, np.float64)
        b2 = np.ascontiguousarray(b2, np.float64)
        _check(lib().dmfe_set_positions(self._h, _ptr(b1), _ptr(b2)), "dmfe_set_positions")

    def fields(self):
        P = self.N * self.N
        out = np.zeros(15 * P + self.N)
        _check(lib().dmfe_get_fields(self._h, _ptr(out)), "dmfe_get_fields")
        d = {name: out[i * P:(i + 1) * P] for i, name in enumerate(FIELD_NAMES)}
        d["rInt"] = out[15 * P:]
        return d

    def rhs(self):
        hK, hR = np.zeros(self.N), np.zeros(self.N)
        _check(lib().dmfe_rhs(self._h, _ptr(hK), _ptr(hR)), "dmfe_rhs")
        return hK, hR

    def rstep(self):
        r = C.c_double()
        _check(lib().dmfe_rstep(self._h, C.byref(r)), "dmfe_rstep")
        return r.value

    def rk_step(self):
        e = C.c_double()
        _check(lib().dmfe_rk_step(self._h, C.byref(e)), "dmfe_rk_step")
        return e.value

    def loop_iteration(self):
        t = C.c_double()
        _check(lib().dmfe_loop_iteration(self._h, C.byref(t)), "dmfe_loop_iteration")
        return t.value

    def run(self, tmax, max_loop, corr_cap=0):
        corr = np.zeros(3 * max(corr_cap, 1))
        steps = C.c_int64()
        _check(lib().dmfe_run(self._h, C.c_double(tmax), C.c_int64(max_loop), _ptr(corr) if corr_cap else None,
                              C.c_int64(corr_cap), C.byref(steps)), "dmfe_run")
        n = min(steps.value, corr_cap)
        return steps.value, corr[:3 * n].reshape(n, 3)

    def observables(self, t):
        e, q = C.c_double(), C.c_double()
        _check(lib().dmfe_observables(self._h, C.c_double(t), C.byref(e), C.byref(q)), "dmfe_observables")
        return e.value, q.value

    def sparsify(self, threshold):
        _check(lib().dmfe_sparsify(self._h, C.c_double(threshold)), "dmfe_sparsify")

    def save_state(self, path):
        """data.bin version-1 checkpoint (the reference's binary layout)."""
        _check(lib().dmfe_save_state(self._h, os.fsencode(path)), "dmfe_save_state")

    def load_state(self, path):
        """Resume from a data.bin version-1 checkpoint; returns the energy stored in it."""
        e = C.c_double()
        _check(lib().dmfe_load_state(self._h, os.fsencode(path), C.byref(e)), "dmfe_load_state")
        return e.value

    # -- parameter sensitivities (forward mode) ------------------------------------
    def tangent_init(self, dGamma, dLambda):
        g = np.ascontiguousarray(dGamma, np.float64).ravel()
        l = np.ascontiguousarray(dLambda, np.float64).ravel()
        _check(lib().dmfe_tangent_init(self._h, C.c_int(len(g)), _ptr(g), _ptr(l)), "dmfe_tangent_init")
        self._tanK = len(g)

    def tangent_observables(self, t):
        dE, dQ = np.zeros(self._tanK), np.zeros(self._tanK)
        _check(lib().dmfe_tangent_observables(self._h, C.c_double(t), _ptr(dE), _ptr(dQ)), "dmfe_tangent_observables")
        return dE, dQ

    def tangent_state(self, k):
        T1, N = self.T1, self.N
        dQK, dQR, dr = np.zeros(T1 * N), np.zeros(T1 * N), np.zeros(T1)
        _check(lib().dmfe_tangent_state(self._h, C.c_int(k), _ptr(dQK), _ptr(dQR), _ptr(dr)), "dmfe_tangent_state")
        return dict(QK=dQK, QR=dQR, rvec=dr)

    def gradients(self, tmax, max_loop):
        """[dE/dGamma, dE/dlambda, dQK0/dGamma, dQK0/dlambda] of a quench run to tmax / max_loop,
        plus the final proposed time and the number of steps."""
        grad = np.zeros(4)
        t, n = C.c_double(), C.c_int64()
        _check(lib().dmfe_gradients(self._h, C.c_double(tmax), C.c_int64(max_loop), _ptr(grad), C.byref(t),
                                    C.byref(n)), "dmfe_gradients")
        self._tanK = 2
        return grad, t.value, n.value

    def adjoint(self, seeds, tmax, max_loop):
        """dmfe_adjoint: (seeds . dE,QK0/dGamma, seeds . dE,QK0/dlambda) of a quench run."""
        sd = np.ascontiguousarray(seeds, np.float64)
        if sd.size != 2:
            raise ValueError("adjoint: two seeds (energy, QK(t,0))")
        dG, dL, t, n = C.c_double(), C.c_double(), C.c_double(), C.c_int64()
        _check(lib().dmfe_adjoint(self._h, _ptr(sd), C.c_double(tmax), C.c_int64(max_loop), C.byref(dG), C.byref(dL),
                                  C.byref(t), C.byref(n)), "dmfe_adjoint")
        self._tanK = 2
        return dG.value, dL.value, t.value, n.value

    def synchronize(self):
        _check(lib().dmfe_synchronize(self._h), "dmfe_synchronize")


def profile(ctx, enable=True):
    _check(lib().dmfe_profile(ctx._h, C.c_int(1 if enable else 0)), "dmfe_profile")


def profile_read(ctx, kind):
    """(total_ms, launches) of kernel `kind` (0 search, 1 rhs_rows, 2 finish) since profile(ctx)."""
    ms, n = C.c_double(), C.c_int64()
    _check(lib().dmfe_profile_read(ctx._h, C.c_int(kind), C.byref(ms), C.byref(n)), "dmfe_profile_read")
    return ms.value, n.value


def time_rhs(ctx, reps=100):
    """Average ms of `reps` back-to-back RHS kernel launches on the context's current state."""
    ms = C.c_double()
    _check(lib().dmfe_time_rhs(ctx._h, C.c_int(reps), C.byref(ms)), "dmfe_time_rhs")
    return ms.value


def gradients_many(grids, params, tmax, max_loop, rows_hint=1024):
    """dmfe_gradients for several (grid, params) points concurrently (one context, stream
    and host thread each).  Returns (grads [n, 4], t_final [n], steps [n])."""
    n = len(grids)
    if n != len(params) or n == 0:
        raise ValueError("gradients_many: need one grid and one params per point")
    gs = (Grid * n)(*[g.as_struct() for g in grids])
    ps = (Params * n)(*params)
    grads = np.zeros((n, 4))
    tf = np.zeros(n)
    st = np.zeros(n, dtype=np.int64)
    _check(lib().dmfe_gradients_many(C.c_int(n), gs, ps, C.c_double(tmax), C.c_int64(max_loop), C.c_int64(rows_hint),
                                     _ptr(grads), _ptr(tf), _ptr(st)), "dmfe_gradients_many")
    return grads, tf, st


def set_serk2(ctx, on):
    """The reference GPU loop's SERK2 promotion at the maintenance (off by default)."""
    _check(lib().dmfe_set_serk2(ctx._h, C.c_int(1 if on else 0)), "dmfe_set_serk2")


def set_rollback(ctx, on):
    """The reference GPU loop's rollback on error spikes (off by default, as on the CPU path)."""
    _check(lib().dmfe_set_rollback(ctx._h, C.c_int(1 if on else 0)), "dmfe_set_rollback")


def set_tail_fit(ctx, on):
    """The reference main loop's QK tail fit after every step (dmfe_set_tail_fit)."""
    _check(lib().dmfe_set_tail_fit(ctx._h, C.c_int(1 if on else 0)), "dmfe_set_tail_fit")


def set_gpu_variants(ctx, on):
    """The reference GPU path's interpolation: bisection only, per-element B1 clamp."""
    _check(lib().dmfe_set_gpu_variants(ctx._h, C.c_int(1 if on else 0)), "dmfe_set_gpu_variants")


def tail_fit_stats(ctx):
    """(blends applied, smoothed fit parameter) of the context's tail fit."""
    n, a = C.c_int64(), C.c_double()
    _check(lib().dmfe_tail_fit_stats(ctx._h, C.byref(n), C.byref(a)), "dmfe_tail_fit_stats")
    return n.value, a.value


def set_loop_graphs(ctx, on):
    """dmfe_run replays captured graphs (True, default) or launches directly (False)."""
    _check(lib().dmfe_set_loop_graphs(ctx._h, C.c_int(1 if on else 0)), "dmfe_set_loop_graphs")


def rhs_shape(ctx):
    """(G, split) of the context's RHS kernel."""
    sh = RhsShape()
    _check(lib().dmfe_get_rhs_shape(ctx._h, C.byref(sh)), "dmfe_get_rhs_shape")
    return sh.astuple()


def set_rhs_shape(ctx, G, split):
    _check(lib().dmfe_set_rhs_shape(ctx._h, C.byref(RhsShape(G, split))), "dmfe_set_rhs_shape")


def tuning_get(N, m, log_response=0):
    """(G, split) that contexts and batches of grid class (N, m, log_response) are created with."""
    sh = RhsShape()
    _check(lib().dmfe_tuning_get(C.c_int64(N), C.c_int(m), C.c_int(log_response), C.byref(sh)), "dmfe_tuning_get")
    return sh.astuple()


def tuning_set(N, m, log_response, G, split):
    _check(lib().dmfe_tuning_set(C.c_int64(N), C.c_int(m), C.c_int(log_response), C.byref(RhsShape(G, split))),
           "dmfe_tuning_set")


def autotune(ctx, reps=20, install=True):
    """Times every valid RHS shape on the context's current history and keeps the fastest.
    Returns (best (G, split), {(G, split): average launch ms})."""
    best = RhsShape()
    tried = (RhsShape * 16)()
    ms = (C.c_double * 16)()
    n = C.c_int()
    _check(lib().dmfe_autotune(ctx._h, C.c_int(reps), C.c_int(1 if install else 0), C.byref(best), tried, ms,
                               C.c_int(16), C.byref(n)), "dmfe_autotune")
    return best.astuple(), {tried[q].astuple(): ms[q] for q in range(min(n.value, 16))}


def dense_apply(ctx, X, map_=0, reps=1):
    """Y = W_map . X for a dense (B-spline) grid context on MFMA; X is (N, B) with B <= 16.
    Returns (Y of shape (N*N, B), average device ms per contraction)."""
    X = np.ascontiguousarray(X, np.float64)
    if X.ndim == 1:
        X = X[:, None]
    B = X.shape[1]
    Y = np.zeros((ctx.N * ctx.N, B))
    ms = C.c_double()
    _check(lib().dmfe_dense_apply(ctx._h, C.c_int(map_), C.c_int(B), _ptr(X), _ptr(Y), C.c_int(reps), C.byref(ms)),
           "dmfe_dense_apply")
    return Y, ms.value


def copy_probe(nbytes, reps=10):
    """Average ms of a plain HBM copy of nbytes (counter calibration)."""
    ms = C.c_double()
    _check(lib().dmfe_copy_probe(C.c_int64(int(nbytes)), C.c_int(reps), C.byref(ms)), "dmfe_copy_probe")
    return ms.value


BATCH_RUNNING, BATCH_DONE = 0, 1


class Batch:
    """n independent quenches advanced together with device-side step control (dmfe_batch).

    grids: one GridPackage per instance (instances given the same package share its
    device tables); params: one Params per instance (same log_response)."""

    def __init__(self, grids, params, rows_cap=256):
        if len(grids) != len(params) or not grids:
            raise ValueError("Batch: need one grid and one params per instance")
        self.n = len(grids)
        self.grids = list(grids)
        self.N = [g.N for g in grids]
        self._gs = (Grid * self.n)(*[g.as_struct() for g in grids])
        self._ps = (Params * self.n)(*params)
        self._h = C.c_void_p()
        _check(lib().dmfe_batch_create(C.c_int(self.n), self._gs, self._ps, C.c_int64(rows_cap), C.byref(self._h)),
               "dmfe_batch_create")

    def close(self):
        if self._h:
            lib().dmfe_batch_destroy(self._h)
            self._h = C.c_void_p()

    @property
    def lanes(self):
        """Lane-group width G of the batch RHS kernel."""
        G = C.c_int()
        _check(lib().dmfe_batch_get_lanes(self._h, C.byref(G)), "dmfe_batch_get_lanes")
        return G.value

    @lanes.setter
    def lanes(self, G):
        _check(lib().dmfe_batch_set_lanes(self._h, C.c_int(G)), "dmfe_batch_set_lanes")

    def autotune(self, reps=5):
        """Times the batch RHS for every valid G and keeps the fastest: (G, {G: ms})."""
        best = C.c_int()
        tried = (C.c_int32 * 8)()
        ms = (C.c_double * 8)()
        n = C.c_int()
        _check(lib().dmfe_batch_autotune(self._h, C.c_int(reps), C.byref(best), tried, ms, C.c_int(8), C.byref(n)),
               "dmfe_batch_autotune")
        return best.value, {tried[q]: ms[q] for q in range(min(n.value, 8))}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init_quench(self):
        _check(lib().dmfe_batch_init_quench(self._h), "dmfe_batch_init_quench")
        self._tanK = 0

    def set_limits(self, tmax=float("inf"), max_loop=2**62):
        _check(lib().dmfe_batch_set_limits(self._h, C.c_double(tmax), C.c_int64(max_loop)), "dmfe_batch_set_limits")

    def tangent_init(self, dGamma, dLambda):
        """Forward-mode sensitivities of every instance along K = len(dGamma) directions
        (dmfe_batch_tangent_init); right after init_quench."""
        dG = np.ascontiguousarray(dGamma, np.float64)
        dL = np.ascontiguousarray(dLambda, np.float64)
        if dG.size != dL.size or not 1 <= dG.size <= 2:
            raise ValueError("Batch.tangent_init: one or two directions, dGamma and dLambda of equal length")
        self._tanK = int(dG.size)
        _check(lib().dmfe_batch_tangent_init(self._h, C.c_int(self._tanK), _ptr(dG), _ptr(dL)),
               "dmfe_batch_tangent_init")

    def tangent_observables(self):
        """(dE, dQK0), each (n, K): derivatives of E(t) and QK(t,0) at the proposed t of
        every instance's last step along the directions of tangent_init."""
        K = getattr(self, "_tanK", 0)
        if K < 1:
            raise DmfeError("Batch.tangent_observables: call tangent_init first")
        dE, dQ = np.zeros(self.n * K), np.zeros(self.n * K)
        _check(lib().dmfe_batch_tangent_observables(self._h, _ptr(dE), _ptr(dQ)), "dmfe_batch_tangent_observables")
        return dE.reshape(self.n, K), dQ.reshape(self.n, K)

    def step(self):
        """One batch step (direct launches); returns the number of running instances."""
        r = C.c_int()
        _check(lib().dmfe_batch_step(self._h, C.byref(r)), "dmfe_batch_step")
        return r.value

    def run(self, max_steps, use_graph=True, chunk=16):
        n = C.c_int64()
        _check(lib().dmfe_batch_run(self._h, C.c_int64(max_steps), C.c_int(1 if use_graph else 0), C.c_int(chunk),
                                    C.byref(n)), "dmfe_batch_run")
        return n.value

    def T1(self, i):
        v = C.c_int64()
        _check(lib().dmfe_batch_get_T1(self._h, C.c_int(i), C.byref(v)), "dmfe_batch_get_T1")
        return v.value

    def get_control(self, i):
        dt, de, do, lp, me, st = C.c_double(), C.c_double(), C.c_double(), C.c_int64(), C.c_int(), C.c_int()
        _check(lib().dmfe_batch_get_control(self._h, C.c_int(i), C.byref(dt), C.byref(de), C.byref(do), C.byref(lp),
                                            C.byref(me), C.byref(st)), "dmfe_batch_get_control")
        return dict(delta_t=dt.value, delta=de.value, delta_old=do.value, loop=lp.value, method=me.value,
                    status=st.value)

    def set_control(self, i, delta_t, delta, delta_old, loop, method):
        _check(lib().dmfe_batch_set_control(self._h, C.c_int(i), C.c_double(delta_t), C.c_double(delta),
                                            C.c_double(delta_old), C.c_int64(loop), C.c_int(method)),
               "dmfe_batch_set_control")

    def get_state(self, i):
        T1, N = self.T1(i), self.N[i]
        out = {k: np.zeros(T1) for k in ("t1grid", "rvec", "drvec", "dtratio")}
        for k in ("QK", "QR", "dQK", "dQR"):
            out[k] = np.zeros(T1 * N)
        _check(lib().dmfe_batch_get_state(self._h, C.c_int(i), _ptr(out["t1grid"]), _ptr(out["QK"]), _ptr(out["QR"]),
                                          _ptr(out["dQK"]), _ptr(out["dQR"]), _ptr(out["rvec"]), _ptr(out["drvec"]),
                                          _ptr(out["dtratio"])), "dmfe_batch_get_state")
        return out

    def set_state(self, i, t1grid, QK, QR, dQK, dQR, rvec, drvec, dtratio=None):
        """History of instance i (dmfe_set_state layout) + a fresh binary-search interpolation."""
        arr = [np.ascontiguousarray(a, np.float64) for a in (t1grid, QK, QR, dQK, dQR, rvec, drvec)]
        T1, N = len(arr[0]), self.N[i]
        for a in arr[1:5]:
            if a.size != T1 * N:
                raise ValueError(f"Batch.set_state: history arrays must hold T1*N = {T1 * N} values")
        for a in arr[5:]:
            if a.size != T1:
                raise ValueError("Batch.set_state: rvec and drvec must hold T1 values")
        rho = None if dtratio is None else np.ascontiguousarray(dtratio, np.float64)
        if rho is not None and rho.size != T1:
            raise ValueError("Batch.set_state: dtratio must hold T1 values")
        _check(lib().dmfe_batch_set_state(self._h, C.c_int(i), C.c_int64(T1), *[_ptr(a) for a in arr], _ptr(rho)),
               "dmfe_batch_set_state")

    def record(self, cap):
        self._rec_cap = int(cap)
        _check(lib().dmfe_batch_record(self._h, C.c_int(int(cap))), "dmfe_batch_record")

    def records(self, i):
        """(cap, 4) array of (t, E, QK(t,0), loop) rows; row (loop-1) % cap holds loop."""
        out = np.zeros((self._rec_cap, 4))
        _check(lib().dmfe_batch_records(self._h, C.c_int(i), _ptr(out)), "dmfe_batch_records")
        return out
