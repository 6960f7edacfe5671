"""ctypes binding of the device library's C-ABI (include/mpsd.h).

This is plumbing for tests, benchmarks and Python callers: every computation happens in
libmpsd.so.  Loading fails loudly (RuntimeError) if the library has not been built;
there is no CPU fallback.

Arrays cross the boundary as complex128 numpy arrays indexed like the reference's mwArray
(first index first); they are flattened column-major (Fortran order), which is the
library's storage order.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmpsd.so")

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16X3 = 2

OPT_TILE_K = 1
OPT_SPLIT_STAGE3 = 2
OPT_GRAPH = 3
OPT_BATCH_GROUP = 4
OPT_GAUGE_QR = 5
OPT_FUSED_LANCZOS = 6
OPT_ENV_CHAIN = 7
OPT_KRYLOV = 8
OPT_LP_PACKED = 9
OPT_LP_TILE = 10
OPT_PACKED_FP64 = 11
OPT_PACKED_FP64_TILE = 12
OPT_PACKED_FP64_X = 13

_lib = None


class MpsdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mpsd error {code}: {msg}")
        self.code = code


def load():
    """Load libmpsd.so (once).  Raises RuntimeError when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built: run `python -m mpsdyn_amd.build` first")
    lib = C.CDLL(LIB_PATH)
    P, I, I64, D = C.c_void_p, C.c_int, C.c_int64, C.c_double
    sig = {
        "mpsd_ctx_create": (I, [I, C.POINTER(P)]),
        "mpsd_ctx_destroy": (I, [P]),
        "mpsd_ctx_sync": (I, [P]),
        "mpsd_last_error": (C.c_char_p, [P]),
        "mpsd_ctx_stream": (P, [P]),
        "mpsd_ctx_device_bytes": (I64, [P]),
        "mpsd_tensor_create": (I, [P, I, C.POINTER(I), C.POINTER(P)]),
        "mpsd_tensor_destroy": (I, [P]),
        "mpsd_tensor_rank": (I, [P]),
        "mpsd_tensor_dims": (I, [P, C.POINTER(I)]),
        "mpsd_tensor_size": (I64, [P]),
        "mpsd_tensor_reshape": (I, [P, I, C.POINTER(I)]),
        "mpsd_tensor_upload": (I, [P, P, I64]),
        "mpsd_tensor_download": (I, [P, P, I64]),
        "mpsd_tensor_copy": (I, [P, P]),
        "mpsd_tensor_devptr": (P, [P]),
        "mpsd_env_left": (I, [P, P, P, P, P, P]),
        "mpsd_env_right": (I, [P, P, P, P, P, P]),
        "mpsd_heff_create": (I, [P, P, P, P, D, I, C.POINTER(P)]),
        "mpsd_heff_destroy": (I, [P]),
        "mpsd_heff_size": (I64, [P]),
        "mpsd_heff_apply": (I, [P, P, P, I, I64, I64, I]),
        "mpsd_heff_full": (I, [P, P]),
        "mpsd_heff_set_precision": (I, [P, I]),
        "mpsd_heff_set_hermitian": (I, [P, I]),
        "mpsd_heff_eigs_dense": (I, [P, P, C.POINTER(D)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _optional(lib)
    _lib = lib
    return lib


def _optional(lib):
    """Bind entry points that later parts of the library add (absent ones are skipped)."""
    P, I, I64, D = C.c_void_p, C.c_int, C.c_int64, C.c_double
    extra = {
        "mpsd_vec_dots": (I, [P, P, I64, I, P, I64, P]),
        "mpsd_vec_gemv": (I, [P, P, I64, I, P, P, I64]),
        "mpsd_vec_axpby": (I, [P, P, P, P, P, I64]),
        "mpsd_vec_copy": (I, [P, P, P, I64]),
        "mpsd_vec_zero": (I, [P, P, I64]),
        "mpsd_vec_norm2": (I, [P, P, I64, C.POINTER(D)]),
        "mpsd_site_gauge_r": (I, [P, P, P]),
        "mpsd_site_gauge_l": (I, [P, P, P]),
        "mpsd_site_reduce_dim_r": (I, [P, P, P]),
        "mpsd_site_reduce_dim_l": (I, [P, P, P]),
        "mpsd_site_gauge_r_absorb": (I, [P, P, P, P]),
        "mpsd_site_gauge_l_absorb": (I, [P, P, P, P]),
        "mpsd_gs_solve": (I, [P, C.POINTER(GsConfig), C.POINTER(GsResult)]),
        "mpsd_gs_solve_export": (I, [P, C.POINTER(GsConfig), C.c_char_p, C.POINTER(GsResult)]),
        "mpsd_gs_expectation": (I, [P, C.POINTER(GsConfig), C.c_char_p, C.POINTER(D), C.POINTER(D)]),
        "mpsd_mpo_save": (I, [P, C.POINTER(GsConfig), C.c_char_p, I]),
        "mpsd_vec_proj_update": (I, [P, P, I64, I, C.POINTER(D), P, P, I64]),
        "mpsd_vec_dots_dev": (I, [P, P, I64, I, P, I64, P]),
        "mpsd_vec_gemv_dev": (I, [P, P, I64, I, P, D, P, I64]),
        "mpsd_lanczos_coef_update": (I, [P, P, I64, P, P, I, P, P, I64]),
        "mpsd_lanczos_gs1_dev": (I, [P, P, I64, I, P, I64, P]),
        "mpsd_lanczos_gs2_dev": (I, [P, P, I64, I, P, P, P, P, I64]),
        "mpsd_lanczos_gs_step_dev": (I, [P, P, I64, I, P, I64, P, P, P]),
        "mpsd_site_gauge_r_svd": (I, [P, P, P, I]),
        "mpsd_site_gauge_l_svd": (I, [P, P, P, I]),
        "mpsd_mps_truncate": (I, [P, C.c_char_p, I, C.c_char_p]),
        "mpsd_gs_solve_levels": (I, [P, C.POINTER(GsConfig), C.c_char_p, I, I, D, C.POINTER(D), C.POINTER(D),
                                     C.POINTER(GsResult)]),
        "mpsd_gs_sweep_part": (I, [P, C.POINTER(GsConfig), C.c_char_p, I, C.c_char_p, I, C.POINTER(C.c_char_p),
                                   C.POINTER(D), I, I, C.c_char_p, C.POINTER(D)]),
        "mpsd_gs_solve_mpo": (I, [P, C.POINTER(GsConfig), C.c_char_p, I, C.c_char_p, C.POINTER(GsResult)]),
        "mpsd_mpo_expectation": (I, [P, C.c_char_p, I, C.c_char_p, C.POINTER(D), C.POINTER(D)]),
        "mpsd_gs_solve_from": (I, [P, C.POINTER(GsConfig), C.c_char_p, I, C.c_char_p, C.c_char_p,
                                   C.POINTER(GsResult)]),
        "mpsd_tensor_resize": (I, [P, I, C.POINTER(I)]),
        "mpsd_mps_gauge_site": (I, [P, C.c_char_p, I, C.c_char, I, C.c_char_p]),
        "mpsd_heff_apply_noalloc": (I, [P, P, P, P, D, P, P]),
        "mpsd_heff_eigs_lowest": (I, [P, P, D, I, C.POINTER(D), C.POINTER(D), C.POINTER(I)]),
        "mpsd_heff_apply_grouped": (I, [P, I, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
        "mpsd_gs_footprint": (I64, [C.POINTER(GsConfig)]),
        "mpsd_server_create": (I, [C.POINTER(ServeConfig), C.POINTER(P)]),
        "mpsd_server_destroy": (I, [P]),
        "mpsd_server_last_error": (C.c_char_p, [P]),
        "mpsd_server_submit": (I, [P, C.POINTER(GsConfig), C.POINTER(I)]),
        "mpsd_server_wait": (I, [P, I, C.POINTER(ServeResult)]),
        "mpsd_server_run_trace": (I, [P, I, C.POINTER(GsConfig), C.POINTER(D), C.POINTER(ServeResult)]),
        "mpsd_server_stats": (I, [P, C.POINTER(ServeStats)]),
        "mpsd_ctx_set_timing": (I, [P, I]),
        "mpsd_ctx_set_timing_kinds": (I, [P, C.c_uint]),
        "mpsd_ctx_timing": (I, [P, I, C.POINTER(D), C.POINTER(I64)]),
        "mpsd_ctx_set_tile": (I, [P, I, I]),
        "mpsd_ctx_set_option": (I, [P, I, I]),
        "mpsd_ctx_get_option": (I, [P, I]),
        "mpsd_tune_heff": (I, [P, P, P, I, C.POINTER(TuneResult), I, C.POINTER(I)]),
        "mpsd_tune_env": (I, [P, I, P, P, P, P, P, I, C.POINTER(TuneResult), I, C.POINTER(I)]),
        "mpsd_tune_count": (I, [P]),
        "mpsd_tune_clear": (I, [P]),
        "mpsd_tune_save": (I, [P, C.c_char_p]),
        "mpsd_tune_load": (I, [P, C.c_char_p]),
    }
    for name, (res, args) in extra.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args


MODEL_ISING, MODEL_HEISENBERG, MODEL_SCHWINGER = 0, 1, 2


class TuneResult(C.Structure):
    _fields_ = [("M", C.c_int), ("N", C.c_int), ("K", C.c_int), ("NB", C.c_int), ("arith", C.c_int),
                ("kind", C.c_int), ("tm", C.c_int), ("tn", C.c_int), ("tk", C.c_int), ("nbw", C.c_int),
                ("us_builtin", C.c_double), ("us_best", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class GsConfig(C.Structure):
    """mpsd_gs_config (include/mpsd_gs.h)."""
    _fields_ = [("model", C.c_int), ("L", C.c_int), ("D", C.c_int), ("p", C.c_double * 4),
                ("offset", C.c_double), ("convtol", C.c_double), ("eigtol", C.c_double),
                ("max_halfsweeps", C.c_int), ("precision", C.c_int), ("seed", C.c_uint64), ("solver", C.c_int),
                ("want_residual", C.c_int), ("breakdown", C.c_int), ("tol_schedule", C.c_int)]


class GsResult(C.Structure):
    """mpsd_gs_result (include/mpsd_gs.h)."""
    _fields_ = [("energy", C.c_double), ("initial_energy", C.c_double), ("halfsweeps", C.c_int),
                ("matvecs", C.c_int64), ("env_updates", C.c_int64), ("seconds", C.c_double),
                ("residual", C.c_double), ("device_bytes", C.c_int64), ("matvecs_lowp", C.c_int64),
                ("refine_from", C.c_int), ("n_history", C.c_int),
                ("history", C.c_double * 256), ("residual_norm", C.c_double), ("t_env", C.c_double),
                ("t_eig", C.c_double), ("t_gauge", C.c_double), ("t_matvec", C.c_double),
                ("matvec_flops", C.c_double), ("env_flops", C.c_double), ("residual_floor", C.c_double),
                ("hs_seconds", C.c_double * 256), ("hs_tol", C.c_double * 256), ("hs_matvecs", C.c_int64 * 256)]


class ServeConfig(C.Structure):
    """mpsd_serve_config (include/mpsd_serve.h)."""
    _fields_ = [("device", C.c_int), ("slots", C.c_int), ("hbm_budget", C.c_int64), ("batching", C.c_int),
                ("window_us", C.c_int), ("graph", C.c_int)]


class ServeResult(C.Structure):
    """mpsd_serve_result (include/mpsd_serve.h)."""
    _fields_ = [("id", C.c_int), ("status", C.c_int), ("energy", C.c_double), ("arrival_s", C.c_double),
                ("admit_s", C.c_double), ("finish_s", C.c_double), ("halfsweeps", C.c_int),
                ("matvecs", C.c_int64), ("matvecs_lowp", C.c_int64), ("footprint_bytes", C.c_int64),
                ("matvec_flops", C.c_double)]


class ServeStats(C.Structure):
    """mpsd_serve_stats (include/mpsd_serve.h)."""
    _fields_ = [("solves", C.c_int64), ("batches", C.c_int64), ("batched_matvecs", C.c_int64),
                ("max_batch", C.c_int64), ("peak_admitted_bytes", C.c_int64), ("peak_running", C.c_int),
                ("graph_replays", C.c_int64)]


class Server:
    """The serving loop (mpsd_server): concurrent solves admitted by HBM footprint."""

    def __init__(self, slots=8, hbm_budget=0, batching=True, window_us=200, graph=False, device=0):
        self.lib = load()
        h = C.c_void_p()
        cfg = ServeConfig(device, slots, hbm_budget, 1 if batching else 0, window_us, 1 if graph else 0)
        rc = self.lib.mpsd_server_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise MpsdError(rc, "mpsd_server_create failed")
        self._h = h

    def run_trace(self, cfgs, arrivals):
        n = len(cfgs)
        arr = (GsConfig * n)(*cfgs)
        t = (C.c_double * n)(*arrivals)
        out = (ServeResult * n)()
        rc = self.lib.mpsd_server_run_trace(self._h, n, arr, t, out)
        if rc != 0:
            raise MpsdError(rc, self.lib.mpsd_server_last_error(self._h).decode())
        return list(out)

    def stats(self):
        st = ServeStats()
        self.lib.mpsd_server_stats(self._h, C.byref(st))
        return st

    def close(self):
        if getattr(self, "_h", None):
            self.lib.mpsd_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gs_footprint(cfg):
    return int(load().mpsd_gs_footprint(C.byref(cfg)))


SOLVER_LANCZOS, SOLVER_FULLEIG = 0, 1


def gs_config(model, L, D, p, convtol=1e-10, eigtol=0.0, max_halfsweeps=0, precision=PREC_FP64,
              seed=1, offset=0.0, solver=SOLVER_LANCZOS, want_residual=False, breakdown=False,
              tol_schedule=False):
    pp = list(p) + [0.0] * (4 - len(p))
    return GsConfig(model, L, D, (C.c_double * 4)(*pp), offset, convtol, eigtol, max_halfsweeps,
                    precision, seed, solver, 1 if want_residual else 0, 1 if breakdown else 0,
                    1 if tol_schedule else 0)


def gs_solve(ctx, cfg, export_path=None):
    """Run one ground-state solve on the device; returns GsResult."""
    res = GsResult()
    if export_path is None:
        ctx.check(ctx.lib.mpsd_gs_solve(ctx.handle, C.byref(cfg), C.byref(res)))
    else:
        ctx.check(ctx.lib.mpsd_gs_solve_export(ctx.handle, C.byref(cfg), export_path.encode(), C.byref(res)))
    return res


MPO_BINARY, MPO_TEXT = 0, 1


def mpo_save(cfg, path, fmt=MPO_BINARY, ctx=None):
    """Write the Hamiltonian MPO of cfg's model to `path` (MPO::saveMPO or exportMPOtext
    layout).  Host-only: works without a GPU when ctx is None."""
    lib = ctx.lib if ctx is not None else load()
    rc = lib.mpsd_mpo_save(ctx.handle if ctx is not None else None, C.byref(cfg), str(path).encode(), fmt)
    if ctx is not None:
        ctx.check(rc)
    elif rc != 0:
        raise MpsdError(rc, "mpsd_mpo_save failed")


def gs_solve_mpo(ctx, cfg, mpo_path, fmt=MPO_BINARY, export_path=None):
    """Ground state of the Hamiltonian in an MPO file; cfg supplies D, tolerances, precision,
    seed and solver.  Returns GsResult."""
    res = GsResult()
    ctx.check(ctx.lib.mpsd_gs_solve_mpo(ctx.handle, C.byref(cfg), str(mpo_path).encode(), fmt,
                                        None if export_path is None else str(export_path).encode(), C.byref(res)))
    return res


def gs_solve_from(ctx, cfg, mps_in, mpo_path=None, fmt=MPO_BINARY, export_path=None):
    """findGroundState from the exported MPS mps_in, fitted to the Hamiltonian and to cfg.D
    (grown by zero padding or cut; D = 0 keeps its bond).  Returns GsResult."""
    res = GsResult()
    ctx.check(ctx.lib.mpsd_gs_solve_from(ctx.handle, C.byref(cfg), None if mpo_path is None else str(mpo_path).encode(),
                                         fmt, str(mps_in).encode(),
                                         None if export_path is None else str(export_path).encode(), C.byref(res)))
    return res


def mpo_expectation(ctx, mpo_path, fmt, mps_path):
    """<psi|H|psi> of an exported MPS for the MPO in an MPO file."""
    re_, im_ = C.c_double(), C.c_double()
    ctx.check(ctx.lib.mpsd_mpo_expectation(ctx.handle, str(mpo_path).encode(), fmt, str(mps_path).encode(),
                                           C.byref(re_), C.byref(im_)))
    return complex(re_.value, im_.value)


def gs_solve_levels(ctx, cfg, nlevels, mpo_path=None, fmt=MPO_BINARY, exc_orth_tol=0.0):
    """Ground state and the next nlevels-1 excited states (findGroundState, then
    findNextExcitedState).  Returns (energies, |<i|j>| matrix, per-level GsResult list)."""
    e = (C.c_double * nlevels)()
    ov = (C.c_double * (nlevels * nlevels))()
    res = (GsResult * nlevels)()
    ctx.check(ctx.lib.mpsd_gs_solve_levels(ctx.handle, C.byref(cfg), None if mpo_path is None else str(mpo_path).encode(),
                                           fmt, nlevels, exc_orth_tol, e, ov, res))
    return (np.array(e[:]), np.array(ov[:]).reshape(nlevels, nlevels, order="F"), [res[k] for k in range(nlevels)])


def gs_sweep_part(ctx, cfg, mps_in, pos0, lim, mps_out=None, proj=(), weights=(), mpo_path=None, fmt=MPO_BINARY):
    """Contractor::sweepPart (no proj) or sweepPartWithProjectorPenalty on an exported MPS;
    returns <psi|H|psi>/<psi|psi> of the result."""
    n = len(proj)
    paths = (C.c_char_p * max(n, 1))(*[str(p).encode() for p in proj])
    w = (C.c_double * max(n, 1))(*[float(x) for x in weights])
    e = C.c_double()
    ctx.check(ctx.lib.mpsd_gs_sweep_part(ctx.handle, C.byref(cfg), None if mpo_path is None else str(mpo_path).encode(),
                                         fmt, str(mps_in).encode(), n, paths, w, pos0, lim,
                                         None if mps_out is None else str(mps_out).encode(), C.byref(e)))
    return e.value


def mps_gauge_site(ctx, mps_in, pos, direction, nxt, mps_out):
    """MPS::gaugeCond(pos, dir, next) on an exported state, written to mps_out."""
    ctx.check(ctx.lib.mpsd_mps_gauge_site(ctx.handle, str(mps_in).encode(), int(pos), direction.encode(),
                                          1 if nxt else 0, str(mps_out).encode()))


def mps_truncate(ctx, mps_in, D, mps_out):
    """MPS(const MPS&, D) on an exported state, written to mps_out."""
    ctx.check(ctx.lib.mpsd_mps_truncate(ctx.handle, str(mps_in).encode(), D, str(mps_out).encode()))


def exported_symbols():
    return [n for n in _declared_symbols() if hasattr(load(), n)]


def _declared_symbols():
    """Function names declared in include/*.h."""
    import glob
    import re
    names = []
    inc = os.path.join(os.path.dirname(_HERE), "include")
    for h in sorted(glob.glob(os.path.join(inc, "*.h"))):
        txt = open(h).read()
        names += re.findall(r"^\s*(?:const\s+)?[a-zA-Z_][\w\s\*]*?\b(mpsd_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


class Context:
    """One device context (mpsd_ctx): a HIP stream plus scratch memory."""

    def __init__(self, device=0):
        lib = load()
        h = C.c_void_p()
        rc = lib.mpsd_ctx_create(device, C.byref(h))
        if rc != 0:
            raise MpsdError(rc, "mpsd_ctx_create failed (is a HIP device visible?)")
        self._h = h
        self.lib = lib
        # Tensors and Heff handles of this context; closed before the context itself.
        self._children = weakref.WeakSet()

    @property
    def handle(self):
        return self._h

    @property
    def alive(self):
        return getattr(self, "_h", None) is not None

    def check(self, rc):
        if rc != 0:
            raise MpsdError(rc, self.lib.mpsd_last_error(self._h).decode())
        return rc

    def sync(self):
        self.check(self.lib.mpsd_ctx_sync(self._h))

    def stream(self):
        return self.lib.mpsd_ctx_stream(self._h)

    def device_bytes(self):
        return int(self.lib.mpsd_ctx_device_bytes(self._h))

    def tensor(self, arr=None, dims=None):
        return Tensor(self, arr=arr, dims=dims)

    def close(self):
        if getattr(self, "_h", None):
            kids = list(self._children)
            for k in kids:
                if isinstance(k, Heff):
                    k.close()
            for k in kids:
                k.close()
            self.lib.mpsd_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # timing (HIP events around each launch, accumulated per kernel kind)
    def set_timing(self, on=True):
        self.check(self.lib.mpsd_ctx_set_timing(self._h, 1 if on else 0))

    def set_timing_kinds(self, mask):
        """Time only the kernel kinds whose bit (1 << kind) is set (default all)."""
        self.check(self.lib.mpsd_ctx_set_timing_kinds(self._h, mask))

    def timing(self, kind):
        ms = C.c_double()
        cnt = C.c_int64()
        self.check(self.lib.mpsd_ctx_timing(self._h, kind, C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def set_tile(self, tm, tn):
        self.check(self.lib.mpsd_ctx_set_tile(self._h, tm, tn))

    def set_option(self, option, value):
        self.check(self.lib.mpsd_ctx_set_option(self._h, option, value))

    def get_option(self, option):
        return self.lib.mpsd_ctx_get_option(self._h, option)

    # tile autotuner table (mpsd_tune_*)
    def tune_env(self, side, env, ket, W, bra, out, reps=5):
        rows = (TuneResult * 16)()
        n = C.c_int()
        self.check(self.lib.mpsd_tune_env(self._h, int(side), env.handle, ket.handle, W.handle, bra.handle,
                                          out.handle, int(reps), rows, 16, C.byref(n)))
        return [rows[i].as_dict() for i in range(n.value)]

    def _count(self, rc):
        if rc < 0:
            raise MpsdError(rc, self.lib.mpsd_last_error(self._h).decode())
        return rc

    def tune_count(self):
        return self._count(self.lib.mpsd_tune_count(self._h))

    def tune_clear(self):
        self.check(self.lib.mpsd_tune_clear(self._h))

    def tune_save(self, path):
        self.check(self.lib.mpsd_tune_save(self._h, str(path).encode()))

    def tune_load(self, path):
        return self._count(self.lib.mpsd_tune_load(self._h, str(path).encode()))


class Tensor:
    """HBM-resident complex128 tensor (mpsd_tensor)."""

    def __init__(self, ctx, arr=None, dims=None):
        self.ctx = ctx
        if arr is not None:
            arr = np.asarray(arr, dtype=np.complex128)
            dims = arr.shape if arr.ndim else (1,)
        dims = tuple(int(x) for x in dims)
        cd = (C.c_int * max(1, len(dims)))(*dims)
        h = C.c_void_p()
        ctx.check(ctx.lib.mpsd_tensor_create(ctx.handle, len(dims), cd, C.byref(h)))
        self._h = h
        ctx._children.add(self)
        if arr is not None:
            self.upload(arr)

    @property
    def handle(self):
        return self._h

    @property
    def dims(self):
        lib = self.ctx.lib
        r = lib.mpsd_tensor_rank(self._h)
        out = (C.c_int * 8)()
        lib.mpsd_tensor_dims(self._h, out)
        return tuple(out[i] for i in range(r))

    @property
    def size(self):
        return int(self.ctx.lib.mpsd_tensor_size(self._h))

    def devptr(self):
        return self.ctx.lib.mpsd_tensor_devptr(self._h)

    def resize(self, dims):
        """mwArray::resize with zero noise: keep elements by multi-index, zero-pad, cut."""
        d = (C.c_int * len(dims))(*dims)
        self.ctx.check(self.ctx.lib.mpsd_tensor_resize(self._h, len(dims), d))

    def upload(self, arr):
        buf = np.ascontiguousarray(np.reshape(np.asarray(arr, dtype=np.complex128), -1, order="F"))
        self.ctx.check(self.ctx.lib.mpsd_tensor_upload(self._h, buf.ctypes.data, buf.size))

    def download(self):
        buf = np.empty(self.size, dtype=np.complex128)
        self.ctx.check(self.ctx.lib.mpsd_tensor_download(self._h, buf.ctypes.data, buf.size))
        return np.reshape(buf, self.dims, order="F")

    def close(self):
        if getattr(self, "_h", None) and self.ctx.alive:
            self.ctx.lib.mpsd_tensor_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def env_left(ctx, envL, ket, W, bra, out):
    ctx.check(ctx.lib.mpsd_env_left(ctx.handle, envL.handle, ket.handle, W.handle, bra.handle, out.handle))
    return out


def env_right(ctx, envR, ket, W, bra, out):
    ctx.check(ctx.lib.mpsd_env_right(ctx.handle, envR.handle, ket.handle, W.handle, bra.handle, out.handle))
    return out


def heff_apply_grouped(ctx, heffs, xs, ys):
    """y_i = Heff_i x_i for device tensors xs/ys, in one grouped launch sequence."""
    n = len(heffs)
    H = (C.c_void_p * n)(*[h._h.value for h in heffs])
    X = (C.c_void_p * n)(*[x.devptr() for x in xs])
    Y = (C.c_void_p * n)(*[y.devptr() for y in ys])
    ctx.check(ctx.lib.mpsd_heff_apply_grouped(ctx.handle, n, H, X, Y))


class Heff:
    """Effective Hamiltonian handle (mpsd_heff): y = (L W R) x + offset x."""

    def __init__(self, ctx, L, W, R, offset=0.0, precision=PREC_FP64):
        self.ctx = ctx
        self._keep = (L, W, R)
        h = C.c_void_p()
        ctx.check(ctx.lib.mpsd_heff_create(ctx.handle, L.handle, W.handle, R.handle, float(offset),
                                           int(precision), C.byref(h)))
        self._h = h
        ctx._children.add(self)

    @property
    def size(self):
        return int(self.ctx.lib.mpsd_heff_size(self._h))

    def set_precision(self, p):
        self.ctx.check(self.ctx.lib.mpsd_heff_set_precision(self._h, int(p)))

    def apply(self, x):
        """Host-buffer matvec (x flat or [n, nvec] in column-major)."""
        x = np.asarray(x, dtype=np.complex128)
        n = self.size
        nvec = 1 if x.ndim == 1 else x.shape[1]
        xb = np.asfortranarray(x.reshape(n, nvec, order="F"))
        dims = self._out_len()
        yb = np.empty((dims, nvec), dtype=np.complex128, order="F")
        self.ctx.check(self.ctx.lib.mpsd_heff_apply(self._h, xb.ctypes.data, yb.ctypes.data, nvec, n, dims, 0))
        return yb[:, 0] if x.ndim == 1 else yb

    def apply_device(self, x_ptr, y_ptr, nvec=1, ldx=None, ldy=None):
        n = self.size
        self.ctx.check(self.ctx.lib.mpsd_heff_apply(self._h, x_ptr, y_ptr, nvec, ldx or n,
                                                    ldy or self._out_len(), 1))

    def _out_len(self):
        L, W, R = self._keep
        return W.dims[0] * L.dims[0] * R.dims[0]

    def set_hermitian(self, on=True):
        self.ctx.check(self.ctx.lib.mpsd_heff_set_hermitian(self._h, 1 if on else 0))

    def eigs_dense(self, x_tensor):
        """Lowest eigenpair from the dense matrix (rocSOLVER zheevd); x_tensor receives the
        eigenvector.  Returns the eigenvalue."""
        v = C.c_double()
        self.ctx.check(self.ctx.lib.mpsd_heff_eigs_dense(self._h, x_tensor.devptr(), C.byref(v)))
        return v.value

    def tune(self, x_ptr, y_ptr, reps=5):
        """Autotune this matvec's BRGEMM tiles (device vectors x, y); returns one row per
        shape class with the chosen configuration and its timings."""
        rows = (TuneResult * 16)()
        n = C.c_int()
        self.ctx.check(self.ctx.lib.mpsd_tune_heff(self._h, x_ptr, y_ptr, int(reps), rows, 16, C.byref(n)))
        return [rows[i].as_dict() for i in range(n.value)]

    def eigs_lowest(self, x_tensor, tol=1e-10, max_krylov=30):
        """Lowest eigenpair with the device Lanczos solver; x_tensor holds the start vector
        and receives the eigenvector.  Returns (value, residual, matvecs)."""
        v, r, m = C.c_double(), C.c_double(), C.c_int()
        self.ctx.check(self.ctx.lib.mpsd_heff_eigs_lowest(self._h, x_tensor.devptr(), tol, max_krylov,
                                                          C.byref(v), C.byref(r), C.byref(m)))
        return v.value, r.value, m.value

    def full(self):
        nr = self._out_len()
        out = np.empty((nr, self.size), dtype=np.complex128, order="F")
        self.ctx.check(self.ctx.lib.mpsd_heff_full(self._h, out.ctypes.data))
        return out

    def close(self):
        if getattr(self, "_h", None) and self.ctx.alive:
            self.ctx.lib.mpsd_heff_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
