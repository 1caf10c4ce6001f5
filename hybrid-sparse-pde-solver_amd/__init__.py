"""ctypes binding of the gfx950 solve path (lib/libhsps_dev.so, C ABI in
include/hsps_dev.h).

The directory name is not a Python identifier; load the package by path:
    spec = importlib.util.spec_from_file_location("hsps_amd", ".../hybrid-sparse-pde-solver_amd/__init__.py")

This is a thin host-side mirror for Python callers (tests, bench, serving):
every call goes straight to the C ABI.  There is no CPU fallback — if the
library is missing or no GPU is present, the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhsps_dev.so")

HSPS_FP64, HSPS_MIXED = 0, 1
PRECOND = {"none": 0, "jacobi": 1, "ilu": 2, "amg": 3}
STATUS = {0: "OK", -1: "INVALID", -2: "SIZE", -3: "ZERO_DIAG", -4: "PIVOT", -5: "HIP", -6: "NOMEM",
          -7: "UNSUPPORTED"}


class Params(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int), ("restart_size", C.c_int),
                ("drop_tol", C.c_double), ("fill_per_row", C.c_int), ("amg_levels", C.c_int),
                ("amg_strong", C.c_double), ("amg_smooth_pre", C.c_int), ("amg_smooth_post", C.c_int),
                ("amg_omega", C.c_double), ("jacobi_omega", C.c_double), ("precision", C.c_int),
                ("reliable_delta", C.c_double), ("check_every", C.c_int), ("batch_tile_rows", C.c_int),
                ("batch_graphs", C.c_int), ("cg_engine", C.c_int), ("fgmres_inner_rel", C.c_double),
                ("sstep_k", C.c_int), ("sstep_cond_tol", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("restarts", C.c_int),
                ("reliable_updates", C.c_int), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("true_residual", C.c_double),
                ("solve_time_s", C.c_double), ("setup_time_s", C.c_double),
                ("flop_count", C.c_longlong), ("mem_bytes", C.c_longlong), ("energy_joules", C.c_double),
                ("sstep_k", C.c_int), ("comm_events", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class HspsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None

# (name, restype, argtypes) of every C ABI entry point in include/hsps_dev.h
_vp, _ip, _dp, _fp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_float)
_i, _d, _sz = C.c_int, C.c_double, C.c_size_t
_PP, _SP = C.POINTER(Params), C.POINTER(Stats)
# hsps_user_apply_fn: int (*)(void* user, const double* d_r, double* d_z, void* hip_stream)
USER_APPLY = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)
SIGNATURES = {
    "hsps_last_error": (C.c_char_p, []),
    "hsps_params_default": (None, [_PP]),
    "hsps_ctx_create": (_i, [_i, C.POINTER(_vp)]),
    "hsps_ctx_destroy": (_i, [_vp]),
    "hsps_ctx_set_stream": (_i, [_vp, _vp]),
    "hsps_ctx_stream": (_vp, [_vp]),
    "hsps_ctx_sync": (_i, [_vp]),
    "hsps_ctx_bytes_in_use": (_sz, [_vp]),
    "hsps_malloc": (_i, [_vp, _sz, C.POINTER(_vp)]),
    "hsps_free": (_i, [_vp, _vp]),
    "hsps_memcpy_h2d": (_i, [_vp, _vp, _vp, _sz]),
    "hsps_memcpy_d2h": (_i, [_vp, _vp, _vp, _sz]),
    "hsps_memcpy_d2d": (_i, [_vp, _vp, _vp, _sz]),
    "hsps_memset": (_i, [_vp, _vp, _i, _sz]),
    "hsps_csr_create": (_i, [_vp, _i, _i, _i, _ip, _ip, _dp, C.POINTER(_vp)]),
    "hsps_csr_destroy": (_i, [_vp]),
    "hsps_csr_info": (_i, [_vp, _ip, _ip, _ip]),
    "hsps_csr_update_values": (_i, [_vp, _dp]),
    "hsps_spmv_f64": (_i, [_vp, _vp, _vp]),
    "hsps_spmv_f32": (_i, [_vp, _vp, _vp]),
    "hsps_residual_f64": (_i, [_vp, _vp, _vp, _vp]),
    "hsps_dot_f64": (_i, [_vp, _i, _vp, _vp, _dp]),
    "hsps_norm2_f64": (_i, [_vp, _i, _vp, _dp]),
    "hsps_axpy_f64": (_i, [_vp, _i, _d, _vp, _vp]),
    "hsps_axpby_f64": (_i, [_vp, _i, _d, _vp, _d, _vp]),
    "hsps_scale_f64": (_i, [_vp, _i, _d, _vp]),
    "hsps_precond_create": (_i, [_vp, _i, _PP, C.POINTER(_vp), _dp]),
    "hsps_precond_destroy": (_i, [_vp]),
    "hsps_precond_create_user": (_i, [_vp, _i, USER_APPLY, _vp, C.POINTER(_vp)]),
    "hsps_precond_kind": (_i, [_vp]),
    "hsps_precond_apply_f64": (_i, [_vp, _vp, _vp]),
    "hsps_ilu_factor": (_i, [_i, _ip, _ip, _dp, _d, _i, _ip, _ip, _ip, _dp]),
    "hsps_amg_levels": (_i, [_vp]),
    "hsps_amg_level_info": (_i, [_vp, _i, _ip, _ip]),
    "hsps_amg_aggregates": (_i, [_vp, _i, _ip]),
    "hsps_amg_level_matrix": (_i, [_vp, _i, _i, _ip, _ip, _dp]),
    "hsps_amg_setup_profile": (_i, [_vp, _dp]),
    "hsps_ilu_sweep_kinds": (_i, [_vp, _ip, _ip, _ip, _ip]),
    "hsps_cg_solve": (_i, [_vp, _vp, _vp, _vp, _PP, _SP]),
    "hsps_cg_solve_noalloc": (_i, [_vp, _vp, _vp, _vp, _PP, _SP]),
    "hsps_fgmres_solve": (_i, [_vp, _vp, _vp, _vp, _PP, _SP]),
    "hsps_pipelined_fgmres_solve": (_i, [_vp, _vp, _vp, _vp, _PP, _SP]),
    "hsps_sstep_cg_solve": (_i, [_vp, _vp, _vp, _PP, _SP]),
    "hsps_spectral_interval": (_i, [_vp, _i, _dp, _dp]),
    "hsps_energy_open": (_i, [_i, C.POINTER(_vp)]),
    "hsps_energy_read": (_i, [_vp, _dp, _dp, _dp]),
    "hsps_energy_info": (_i, [_vp, _dp, _dp, _dp, _dp, C.c_char_p, _i]),
    "hsps_energy_close": (_i, [_vp]),
    "hsps_features": (_i, [_vp, _dp]),
    "hsps_batch_create": (_i, [_vp, _i, _ip, _ip, _ip, _dp, C.POINTER(_vp)]),
    "hsps_batch_destroy": (_i, [_vp]),
    "hsps_batch_solve": (_i, [_vp, _vp, _vp, _PP, _ip, _ip, _dp]),
    "hsps_batch_storage": (_i, [_vp, _ip, _ip, _ip]),
    "hsps_stream_probe": (_i, [_vp, C.c_longlong, _i, _dp]),
    "hsps_session_create": (_i, [_vp, _i, C.c_longlong, C.c_longlong, _PP, C.POINTER(_vp)]),
    "hsps_session_destroy": (_i, [_vp]),
    "hsps_session_add": (_i, [_vp, _i, _ip, _ip, _dp, _dp, _vp, _ip]),
    "hsps_session_step": (_i, [_vp, _ip, _ip, _i, _ip]),
    "hsps_session_result": (_i, [_vp, _i, _dp, _vp, _ip, _ip, _dp]),
    "hsps_session_live": (_i, [_vp]),
}


def lib():
    """Load libhsps_dev.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() or make -C "
                               "hybrid-sparse-pde-solver_amd")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise HspsError(rc, lib().hsps_last_error().decode())
    return rc


def default_params(**kw):
    p = Params()
    lib().hsps_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = HSPS_MIXED if v == "mixed" else HSPS_FP64
        setattr(p, k, v)
    return p


class Context:
    def __init__(self, device=0):
        h = _vp()
        check(lib().hsps_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().hsps_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().hsps_ctx_sync(self.h))

    def bytes_in_use(self):
        return lib().hsps_ctx_bytes_in_use(self.h)

    def array(self, host=None, n=None, dtype=np.float64):
        return DeviceArray(self, host=host, n=n, dtype=dtype)


class DeviceArray:
    """A device buffer owned through the C ABI (hsps_malloc / hsps_free)."""

    def __init__(self, ctx, host=None, n=None, dtype=np.float64):
        self.ctx = ctx
        # the C ABI keeps the context alive while this allocation exists
        self.ctx_h = ctx.h
        if host is not None:
            host = np.ascontiguousarray(host, dtype=dtype)
            n = host.size
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = _vp()
        check(lib().hsps_malloc(ctx.h, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p
        if host is not None:
            self.upload(host)
        else:
            check(lib().hsps_memset(ctx.h, self.ptr, 0, max(self.nbytes, 1)))

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.size == self.n
        check(lib().hsps_memcpy_h2d(self.ctx_h, self.ptr, host.ctypes.data_as(_vp), self.nbytes))

    def download(self):
        out = np.empty(self.n, self.dtype)
        check(lib().hsps_memcpy_d2h(self.ctx_h, out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().hsps_free(self.ctx_h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Csr:
    """Device CSR matrix (hsps_csr_create): int32 pattern, FP64 + FP32 values."""

    def __init__(self, ctx, rows, cols, rp, ci, v):
        self.ctx = ctx
        rp = np.ascontiguousarray(rp, np.int32)
        ci = np.ascontiguousarray(ci, np.int32)
        v = np.ascontiguousarray(v, np.float64)
        h = _vp()
        check(lib().hsps_csr_create(ctx.h, rows, cols, len(v), rp.ctypes.data_as(_ip), ci.ctypes.data_as(_ip),
                                    v.ctypes.data_as(_dp), C.byref(h)))
        self.h = h
        self.rows, self.cols, self.nnz = rows, cols, len(v)

    def close(self):
        if getattr(self, "h", None):
            lib().hsps_csr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update_values(self, v):
        """New values, same pattern (hsps_csr_update_values)."""
        v = np.ascontiguousarray(v, np.float64)
        if len(v) != self.nnz:
            raise ValueError("update_values: wrong number of values")
        check(lib().hsps_csr_update_values(self.h, v.ctypes.data_as(_dp)))

    def spmv(self, x, y, precision=HSPS_FP64):
        f = lib().hsps_spmv_f32 if precision == HSPS_MIXED else lib().hsps_spmv_f64
        check(f(self.h, x.ptr, y.ptr))

    def residual(self, b, x, r):
        check(lib().hsps_residual_f64(self.h, b.ptr, x.ptr, r.ptr))

    def features(self):
        out = np.zeros(8)
        check(lib().hsps_features(self.h, out.ctypes.data_as(_dp)))
        keys = ["n", "nnz", "dd", "sym", "fro", "cond", "lmax", "spd"]
        return dict(zip(keys, out.tolist()))


class UserPrecond:
    """A caller-written preconditioner (hsps_precond_create_user): apply(d_r,
    d_z, stream) receives device pointers (ints) of n FP64 entries and must
    leave z = M^-1 r there (e.g. through hsps_memcpy_* and host numpy)."""

    def __init__(self, ctx, n, apply):
        self.ctx, self.n = ctx, n

        def tramp(user, r, z, stream):
            try:
                apply(r, z, stream)
                return 0
            except Exception:  # a negative status stops the solve
                return -1
        self._cb = USER_APPLY(tramp)  # kept alive as long as the object
        h = _vp()
        check(lib().hsps_precond_create_user(ctx.h, n, self._cb, None, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().hsps_precond_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Precond:
    def __init__(self, A, kind, params=None):
        self.A = A
        self.kind = PRECOND[kind] if isinstance(kind, str) else kind
        p = params or default_params()
        h = _vp()
        t = C.c_double()
        check(lib().hsps_precond_create(A.h, self.kind, C.byref(p), C.byref(h), C.byref(t)))
        self.h = h
        self.setup_time_s = t.value

    def close(self):
        if getattr(self, "h", None):
            lib().hsps_precond_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, r, z):
        check(lib().hsps_precond_apply_f64(self.h, r.ptr, z.ptr))

    def amg_levels(self):
        return check(lib().hsps_amg_levels(self.h))

    def ilu_sweeps(self):
        """((forward kind, W), (backward kind, W)): kind 1 = band wavefront, 0 = levels,
        2 = nothing to sweep (forward, no entry below the diagonal: y = r)."""
        v = [C.c_int() for _ in range(4)]
        check(lib().hsps_ilu_sweep_kinds(self.h, *[C.byref(x) for x in v]))
        return (v[0].value, v[2].value), (v[1].value, v[3].value)

    def amg_level_info(self, lv):
        r, z = C.c_int(), C.c_int()
        check(lib().hsps_amg_level_info(self.h, lv, C.byref(r), C.byref(z)))
        return r.value, z.value

    def amg_aggregates(self, lv):
        rows, _ = self.amg_level_info(lv)
        out = np.zeros(rows, np.int32)
        check(lib().hsps_amg_aggregates(self.h, lv, out.ctypes.data_as(_ip)))
        return out

    def amg_setup_profile(self):
        """Seconds per AMG setup stage."""
        out = np.zeros(6, np.float64)
        check(lib().hsps_amg_setup_profile(self.h, out.ctypes.data_as(_dp)))
        return dict(zip(["strength", "aggregate", "prolongation", "transpose", "galerkin", "coarse"], out.tolist()))

    def amg_level_matrix(self, lv, which="A"):
        w = {"A": 0, "P": 1, "R": 2}[which]
        rows, nnz = self.amg_level_info(lv)
        if w == 0:
            nr = rows
        else:
            nc, _ = self.amg_level_info(lv + 1)
            nr = rows if w == 1 else nc
        rp = np.zeros(nr + 1, np.int32)
        # row pointer first (NULL column / value arrays): sizes the rest
        check(lib().hsps_amg_level_matrix(self.h, lv, w, rp.ctypes.data_as(_ip), None, None))
        nnz = int(rp[-1])
        ci = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1), np.float64)
        check(lib().hsps_amg_level_matrix(self.h, lv, w, rp.ctypes.data_as(_ip), ci.ctypes.data_as(_ip),
                                          v.ctypes.data_as(_dp)))
        z = rp[-1]
        return rp, ci[:z], v[:z]


def ilu_factor(rp, ci, v, drop_tol=0.0, fill_per_row=0):
    """L\\U factors (row_ptr, col_idx, values) of a host CSR matrix, computed
    on the host (hsps_ilu_factor): ILU(0), or ILUT(drop_tol, fill_per_row)."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    n = len(rp) - 1
    nz = C.c_int()
    args = (n, rp.ctypes.data_as(_ip), ci.ctypes.data_as(_ip), v.ctypes.data_as(_dp), drop_tol, fill_per_row)
    check(lib().hsps_ilu_factor(*args, C.byref(nz), None, None, None))
    lrp = np.zeros(n + 1, np.int32)
    lci = np.zeros(max(nz.value, 1), np.int32)
    lv = np.zeros(max(nz.value, 1), np.float64)
    check(lib().hsps_ilu_factor(*args, C.byref(nz), lrp.ctypes.data_as(_ip), lci.ctypes.data_as(_ip),
                                lv.ctypes.data_as(_dp)))
    return lrp, lci[:nz.value], lv[:nz.value]


def concat_systems(systems):
    """Concatenate per-system CSRs (local column indices) into the batched
    block-diagonal layout: row_off, global row_ptr, global col_idx, values."""
    sizes = [len(rp) - 1 for rp, _, _ in systems]
    row_off = np.zeros(len(systems) + 1, np.int64)
    np.cumsum(sizes, out=row_off[1:])
    nnz = [len(v) for _, _, v in systems]
    nz_off = np.zeros(len(systems) + 1, np.int64)
    np.cumsum(nnz, out=nz_off[1:])
    rp = np.empty(row_off[-1] + 1, np.int64)
    ci = np.empty(nz_off[-1], np.int32)
    v = np.empty(nz_off[-1], np.float64)
    for s, (srp, sci, sv) in enumerate(systems):
        rp[row_off[s]:row_off[s + 1]] = np.asarray(srp[:-1], np.int64) + nz_off[s]
        ci[nz_off[s]:nz_off[s + 1]] = np.asarray(sci, np.int64) + row_off[s]
        v[nz_off[s]:nz_off[s + 1]] = sv
    rp[-1] = nz_off[-1]
    return row_off.astype(np.int32), rp.astype(np.int32), ci, v


class Batch:
    """Many independent systems solved in one launch (hsps_batch_*)."""

    def __init__(self, ctx, row_off, rp, ci, v):
        self.ctx = ctx
        row_off = np.ascontiguousarray(row_off, np.int32)
        rp = np.ascontiguousarray(rp, np.int32)
        ci = np.ascontiguousarray(ci, np.int32)
        v = np.ascontiguousarray(v, np.float64)
        h = _vp()
        check(lib().hsps_batch_create(ctx.h, len(row_off) - 1, row_off.ctypes.data_as(_ip), rp.ctypes.data_as(_ip),
                                      ci.ctypes.data_as(_ip), v.ctypes.data_as(_dp), C.byref(h)))
        self.h = h
        self.n_systems = len(row_off) - 1
        self.n = int(row_off[-1])
        self.row_off = row_off

    @classmethod
    def from_systems(cls, ctx, systems):
        return cls(ctx, *concat_systems(systems))

    def close(self):
        if getattr(self, "h", None):
            lib().hsps_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, b, x, params=None):
        p = params or default_params()
        it = np.zeros(self.n_systems, np.int32)
        cv = np.zeros(self.n_systems, np.int32)
        tr = np.zeros(self.n_systems, np.float64)
        rc = check(lib().hsps_batch_solve(self.h, b.ptr, x.ptr, C.byref(p), it.ctypes.data_as(_ip),
                                          cv.ctypes.data_as(_ip), tr.ctypes.data_as(_dp)))
        return bool(rc), it, cv, tr

    def storage(self):
        """Systems per storage kind chosen at creation: {"csr", "dia", "stencil"}."""
        n = [C.c_int(), C.c_int(), C.c_int()]
        check(lib().hsps_batch_storage(self.h, *[C.byref(c) for c in n]))
        return {"csr": n[0].value, "dia": n[1].value, "stencil": n[2].value}


class Session:
    """Continuous batching (hsps_session_*): systems join a running batched
    PCG + Jacobi at chunk boundaries and leave when they end."""

    def __init__(self, ctx, max_systems, max_rows, max_nnz, params=None):
        self.ctx = ctx
        self.max_systems = max_systems
        p = params or default_params()
        h = _vp()
        check(lib().hsps_session_create(ctx.h, max_systems, int(max_rows), int(max_nnz), C.byref(p), C.byref(h)))
        self.h = h

    def add(self, rp, ci, v, b, x0=None):
        rp = np.ascontiguousarray(rp, np.int32)
        ci = np.ascontiguousarray(ci, np.int32)
        v = np.ascontiguousarray(v, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        slot = C.c_int()
        check(lib().hsps_session_add(self.h, len(rp) - 1, rp.ctypes.data_as(_ip), ci.ctypes.data_as(_ip),
                                     v.ctypes.data_as(_dp), b.ctypes.data_as(_dp), x0.ptr if x0 is not None else None,
                                     C.byref(slot)))
        return slot.value

    def step(self):
        done = np.zeros(self.max_systems, np.int32)
        nd = C.c_int()
        check(lib().hsps_session_step(self.h, done.ctypes.data_as(_ip), None, self.max_systems, C.byref(nd)))
        return done[:nd.value].tolist()

    def result(self, slot, n, d_x=None):
        """(x on the host, iterations, converged, true relative residual);
        d_x (a DeviceArray of n doubles) also receives x on the device."""
        x = np.zeros(n, np.float64)
        it, cv, tr = C.c_int(), C.c_int(), C.c_double()
        check(lib().hsps_session_result(self.h, slot, x.ctypes.data_as(_dp), d_x.ptr if d_x is not None else None,
                                        C.byref(it), C.byref(cv), C.byref(tr)))
        return x, it.value, bool(cv.value), tr.value

    def live(self):
        return check(lib().hsps_session_live(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().hsps_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve(A, b, x, solver="cg", M=None, params=None):
    """Device solve (solver: cg, fgmres, pipelined_fgmres, sstep_cg); b, x
    DeviceArrays.  s-step CG takes no preconditioner."""
    p = params or default_params()
    st = Stats()
    if solver == "sstep_cg":
        rc = check(lib().hsps_sstep_cg_solve(A.h, b.ptr, x.ptr, C.byref(p), C.byref(st)))
        return bool(rc), st
    f = {"cg": lib().hsps_cg_solve, "fgmres": lib().hsps_fgmres_solve,
         "pipelined_fgmres": lib().hsps_pipelined_fgmres_solve}[solver]
    rc = check(f(A.h, M.h if M is not None else None, b.ptr, x.ptr, C.byref(p), C.byref(st)))
    return bool(rc), st


class Energy:
    """GPU energy counter (ROCm SMI).  Raises HspsError (HSPS_ERR_UNSUPPORTED)
    where telemetry is unavailable."""

    def __init__(self, device=0):
        h = _vp()
        check(lib().hsps_energy_open(device, C.byref(h)))
        self.h = h

    def read(self):
        """(cumulative joules, current watts, counter timestamp s)."""
        j, w, t = C.c_double(), C.c_double(), C.c_double()
        check(lib().hsps_energy_read(self.h, C.byref(j), C.byref(w), C.byref(t)))
        return j.value, w.value, t.value

    def info(self):
        v = [C.c_double() for _ in range(4)]
        name = C.create_string_buffer(128)
        check(lib().hsps_energy_info(self.h, *[C.byref(x) for x in v], name, 128))
        return {"power_cap_w": v[0].value, "temperature_c": v[1].value, "busy_pct": v[2].value,
                "vram_used_mb": v[3].value, "name": name.value.decode(errors="replace")}

    def close(self):
        if getattr(self, "h", None):
            lib().hsps_energy_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spectral_interval(A, steps=20):
    lo, hi = C.c_double(), C.c_double()
    check(lib().hsps_spectral_interval(A.h, steps, C.byref(lo), C.byref(hi)))
    return lo.value, hi.value


# ---- workload generators (host, numpy) -----------------------------------------
def stencil_2d(n, diag, west, east, south, north):
    """CSR of a 5-point operator on an n x n grid, DOF (i,j) -> i*n+j with j the
    x index, columns ascending: (i-1,j), (i,j-1), (i,j), (i,j+1), (i+1,j).
    Coefficients may be scalars or length-N arrays.  Matches the layout of
    SparseMatrix::poisson_2d / convection_diffusion_2d (sparse_matrix.cpp:288-350)."""
    N = n * n
    idx = np.arange(N, dtype=np.int64)
    i, j = idx // n, idx % n
    parts = [(i > 0, -n, south), (j > 0, -1, west), (np.ones(N, bool), 0, diag), (j < n - 1, 1, east),
             (i < n - 1, n, north)]
    cnt = sum(m.astype(np.int64) for m, _, _ in parts)
    rp = np.zeros(N + 1, np.int64)
    np.cumsum(cnt, out=rp[1:])
    ci = np.empty(rp[-1], np.int32)
    v = np.empty(rp[-1], np.float64)
    pos = rp[:-1].copy()
    for m, off, coef in parts:
        c = np.broadcast_to(np.asarray(coef, np.float64), (N,))
        k = pos[m]
        ci[k] = (idx[m] + off).astype(np.int32)
        v[k] = c[m]
        pos[m] += 1
    return rp.astype(np.int32), ci, v


def poisson_2d(n):
    """Unscaled 5-point Poisson: 4 on the diagonal, -1 off (sparse_matrix.cpp:288-311)."""
    return stencil_2d(n, 4.0, -1.0, -1.0, -1.0, -1.0)


def convection_diffusion_2d(n, eps, bx, by):
    """Upwind convection-diffusion entries of sparse_matrix.cpp:317-350."""
    h = 1.0 / (n + 1)
    eh, bxh, byh = eps / (h * h), bx / (2 * h), by / (2 * h)
    N = n * n
    idx = np.arange(N)
    i, j = idx // n, idx % n
    diag = np.full(N, 4.0 * eh)
    diag += np.where(j > 0, max(-bxh, 0.0), 0.0) + np.where(j < n - 1, max(bxh, 0.0), 0.0)
    diag += np.where(i > 0, max(-byh, 0.0), 0.0) + np.where(i < n - 1, max(byh, 0.0), 0.0)
    return stencil_2d(n, diag, -eh + max(bxh, 0.0), -eh - min(bxh, 0.0), -eh + max(byh, 0.0),
                      -eh - min(byh, 0.0))


def anisotropic_2d(n, eps):
    """-eps u_xx - u_yy, 5-point, unscaled: 2 eps + 2 on the diagonal."""
    return stencil_2d(n, 2.0 * eps + 2.0, -eps, -eps, -1.0, -1.0)


def uniform_pm1(seed, n):
    """x_i = 2u-1, u = top 53 bits of the i-th splitmix64 draw / 2^53."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53) - 1.0


# ---- solve server (lib/libhsps.so, include/hsps_serve.h) ---------------------------------
SERVE_LIB_PATH = os.path.join(HERE, "lib", "libhsps.so")
STATE_NAMES = {0: "EASY", 1: "MODERATE", 2: "HARD"}


class ServeResult(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int), ("state", C.c_int), ("batched", C.c_int),
                ("escalations", C.c_int), ("true_residual", C.c_double), ("queue_s", C.c_double),
                ("latency_s", C.c_double), ("submit_s", C.c_double), ("done_s", C.c_double),
                ("setup_s", C.c_double), ("solve_s", C.c_double), ("classify_s", C.c_double)]


_llp = C.POINTER(C.c_longlong)
SERVE_SIGNATURES = {
    "hsps_server_create": (_i, [_i, _i, _d, _i, _i, C.c_longlong, C.POINTER(_vp)]),
    "hsps_server_destroy": (_i, [_vp]),
    "hsps_server_submit": (_i, [_vp, _i, _ip, _ip, _dp, _dp, _PP, _llp]),
    "hsps_server_poll": (_i, [_vp, C.c_longlong, _ip]),
    "hsps_server_wait": (_i, [_vp, C.c_longlong, _dp, C.POINTER(ServeResult)]),
    "hsps_server_last_error": (C.c_char_p, []),
}
_serve_lib = None


def serve_lib():
    """Load libhsps.so (the C++ host layer with the solve server)."""
    global _serve_lib
    if _serve_lib is None:
        lib()  # libhsps_dev.so first (same file the host layer links against)
        if not os.path.exists(SERVE_LIB_PATH):
            raise RuntimeError(f"{SERVE_LIB_PATH} is not built: run make -C hybrid-sparse-pde-solver_amd")
        L = C.CDLL(SERVE_LIB_PATH)
        for name, (res, args) in SERVE_SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _serve_lib = L
    return _serve_lib


def serve_check(rc):
    if rc < 0:
        raise HspsError(rc, serve_lib().hsps_server_last_error().decode())
    return rc


class Server:
    """Stream of independent solves through the adaptive ladder (hsps_server_*)."""

    def __init__(self, device=0, workers=4, hbm_budget=64e9, batch_max=64, batch_rows_max=262144,
                 batch_rows_total=0):
        h = _vp()
        serve_check(serve_lib().hsps_server_create(device, workers, float(hbm_budget), batch_max, batch_rows_max,
                                                   int(batch_rows_total), C.byref(h)))
        self.h = h
        self.sizes = {}

    def submit(self, rp, ci, v, b, params=None):
        rp = np.ascontiguousarray(rp, np.int32)
        ci = np.ascontiguousarray(ci, np.int32)
        v = np.ascontiguousarray(v, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        if len(b) != len(rp) - 1:
            raise HspsError(-1, "Server.submit: b must have one entry per row")
        t = C.c_longlong()
        serve_check(serve_lib().hsps_server_submit(self.h, len(rp) - 1, rp.ctypes.data_as(_ip), ci.ctypes.data_as(_ip),
                                                   v.ctypes.data_as(_dp), b.ctypes.data_as(_dp),
                                                   C.byref(params) if params is not None else None, C.byref(t)))
        self.sizes[t.value] = len(rp) - 1
        return t.value

    def poll(self, ticket):
        d = C.c_int()
        serve_check(serve_lib().hsps_server_poll(self.h, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        """(converged, x, result) of a submitted request."""
        x = np.empty(self.sizes.pop(ticket), np.float64)
        r = ServeResult()
        rc = serve_check(serve_lib().hsps_server_wait(self.h, ticket, x.ctypes.data_as(_dp), C.byref(r)))
        return bool(rc), x, r

    def close(self):
        if getattr(self, "h", None):
            serve_lib().hsps_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
