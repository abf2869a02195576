"""Python front-end of the OpenSTC-MI355X C ABI (include/ostc.h).

Loads the in-tree libostc.so with ctypes; there is no Python compute path:
if the library is missing or the GPU is unavailable every call raises.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "libostc.so")

FIELDS = dict(bv=0, sv=1, cv=2, dq=3, vel=4, acoustic=5, mach=6, mul=7, cond=8, rhoD=9, gamma=10, cp=11,
              dt_local=12, qn=13, bv_last=14, jac=15, metric=16, chem_jac=17, mut=18, turb_cond=19,
              wall_distance=20, turb_jac=21)


class OstcError(RuntimeError):
    pass


_lib = None


class TrainCfg(C.Structure):
    """ostc_surrogate_train_cfg"""
    _fields_ = [("epochs", C.c_int64), ("batch", C.c_int64), ("lr", C.c_double), ("lr_min", C.c_double),
                ("beta1", C.c_double), ("beta2", C.c_double), ("eps", C.c_double), ("seed", C.c_uint64)]


class Gate(C.Structure):
    """ostc_surrogate_gate"""
    _fields_ = [("cons_tol", C.c_double), ("dy_max", C.c_double), ("dT_max", C.c_double), ("margin", C.c_double),
                ("sig_min", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OstcError(f"{LIB_PATH} not built; run openstc-dev_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        fp = C.POINTER(C.c_float)
        sig = {
            "ostc_abi_version": ([], C.c_int),
            "ostc_last_error": ([], C.c_char_p),
            "ostc_init": ([C.c_char_p, C.c_int, C.POINTER(vp)], C.c_int),
            "ostc_destroy": ([vp], C.c_int),
            "ostc_get_int": ([vp, C.c_char_p, ip], C.c_int),
            "ostc_get_real": ([vp, C.c_char_p, dp], C.c_int),
            "ostc_get_table": ([vp, C.c_char_p, dp, C.c_int64], C.c_int),
            "ostc_add_block": ([vp, C.c_int, C.c_int, C.c_int, dp, dp, dp, ip], C.c_int),
            "ostc_add_boundary": ([vp] + [C.c_int] * 8, C.c_int),
            "ostc_read_boundary_file": ([vp], C.c_int),
            "ostc_finalize": ([vp], C.c_int),
            "ostc_block_geometry": ([vp, C.c_int, dp, dp], C.c_int),
            "ostc_field_info": ([vp, C.c_int, C.c_int, C.POINTER(C.c_int64), ip, ip], C.c_int),
            "ostc_upload": ([vp, C.c_int, C.c_int, dp], C.c_int),
            "ostc_download": ([vp, C.c_int, C.c_int, dp], C.c_int),
            "ostc_field_device_ptr": ([vp, C.c_int, C.c_int, C.POINTER(vp)], C.c_int),
            "ostc_initialize_uniform": ([vp, C.c_int], C.c_int),
            "ostc_inflow_state": ([vp, C.c_char_p, dp, C.c_int], C.c_int),
            "ostc_initialize": ([vp], C.c_int),
            "ostc_bind_zone": ([vp, C.c_int, C.POINTER(ZoneView)], C.c_int),
            "ostc_local_dt": ([vp, C.c_int], C.c_int),
            "ostc_store_last": ([vp, C.c_int], C.c_int),
            "ostc_residual": ([vp, C.c_int, dp, dp], C.c_int),
            "ostc_rhs": ([vp, C.c_int], C.c_int),
            "ostc_global_dt": ([vp, dp], C.c_int),
            "ostc_rk3_stage": ([vp, C.c_int, C.c_double, C.c_int], C.c_int),
            "ostc_apply_bc": ([vp, C.c_int, C.c_int], C.c_int),
            "ostc_update_properties": ([vp, C.c_int], C.c_int),
            "ostc_step_rk3": ([vp, C.c_int, dp, dp], C.c_int),
            "ostc_step_steady": ([vp, C.c_int, dp], C.c_int),
            "ostc_implicit": ([vp, C.c_int], C.c_int),
            "ostc_time": ([vp, dp, ip], C.c_int),
            "ostc_set_flux_precision": ([vp, C.c_int], C.c_int),
            "ostc_chem_batch": ([vp, C.c_int64, vp, vp, vp, C.c_double, vp, vp, C.c_int,
                                 C.POINTER(C.c_int64)], C.c_int),
            "ostc_chem_batch_noalloc": ([vp, C.c_int64, vp, vp, vp, C.c_double, vp, vp, C.c_int,
                                         C.POINTER(C.c_int64)], C.c_int),
            "ostc_set_chem_split": ([vp, C.c_int], C.c_int),
            "ostc_write_flowfield": ([vp, C.c_char_p], C.c_int),
            "ostc_read_flowfield": ([vp, C.c_char_p], C.c_int),
            "ostc_set_steady_chem": ([vp, C.c_double, C.c_int], C.c_int),
            "ostc_chem_states": ([vp, C.c_int, vp, vp, vp], C.c_int),
            "ostc_steady_chem_stats": ([vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
            "ostc_read_grid": ([vp], C.c_int),
            "ostc_add_inner_face": ([vp, C.c_int, ip, ip], C.c_int),
            "ostc_read_inner_file": ([vp], C.c_int),
            "ostc_read_mesh": ([vp], C.c_int),
            "ostc_inner_faces": ([vp, C.c_int, ip, ip, C.c_int], C.c_int),
            "ostc_exchange": ([vp, C.c_int], C.c_int),
            "ostc_block_count": ([vp, ip], C.c_int),
            "ostc_block_dims": ([vp, C.c_int, ip], C.c_int),
            "ostc_chem_step": ([vp, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int64)], C.c_int),
            "ostc_dev_alloc": ([vp, C.c_int64, C.POINTER(vp)], C.c_int),
            "ostc_dev_free": ([vp, vp], C.c_int),
            "ostc_memcpy_h2d": ([vp, vp, vp, C.c_int64], C.c_int),
            "ostc_memcpy_d2h": ([vp, vp, vp, C.c_int64], C.c_int),
            "ostc_memcpy_d2d": ([vp, vp, vp, C.c_int64], C.c_int),
            "ostc_sync": ([vp], C.c_int),
            "ostc_profile": ([vp, C.c_int], C.c_int),
            "ostc_profile_report": ([vp, C.c_char_p, C.c_int], C.c_int),
            "ostc_surrogate_init": ([vp, C.c_uint64], C.c_int),
            "ostc_surrogate_param_count": ([vp, C.POINTER(C.c_int64)], C.c_int),
            "ostc_surrogate_get_params": ([vp, fp], C.c_int),
            "ostc_surrogate_set_params": ([vp, fp], C.c_int),
            "ostc_surrogate_forward": ([vp, C.c_int64, fp, fp], C.c_int),
            "ostc_surrogate_gradient": ([vp, C.c_int64, fp, fp, fp, dp], C.c_int),
            "ostc_surrogate_train": ([vp, C.c_int64, vp, vp, vp, C.c_double, C.POINTER(TrainCfg), dp], C.c_int),
            "ostc_surrogate_set_gate": ([vp, C.POINTER(Gate)], C.c_int),
            "ostc_surrogate_get_gate": ([vp, C.POINTER(Gate)], C.c_int),
            "ostc_surrogate_norm": ([vp, dp], C.c_int),
            "ostc_surrogate_save": ([vp, C.c_char_p], C.c_int),
            "ostc_surrogate_load": ([vp, C.c_char_p], C.c_int),
            "ostc_surrogate_last_stats": ([vp, dp], C.c_int),
            "ostc_surrogate_gate_reasons": ([vp, C.POINTER(C.c_int64)], C.c_int),
            "ostc_surrogate_set_kernel": ([vp, C.c_int], C.c_int),
            "ostc_surrogate_network_outputs": ([vp, C.c_int64, vp, vp, vp, fp], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def exported_symbols():
    return [n for n in dir(lib()) if n.startswith("ostc_")]


def _chk(rc):
    if rc != 0:
        raise OstcError(lib().ostc_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class ZoneView(C.Structure):
    """ostc_zone_view: a block's state in caller-owned device memory."""
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int), ("ngg", C.c_int), ("n_var", C.c_int),
                ("n_scalar", C.c_int), ("bv", C.c_void_p), ("sv", C.c_void_p), ("cv", C.c_void_p),
                ("qn", C.c_void_p), ("dq", C.c_void_p)]


class DeviceArray:
    """A device buffer owned by a Solver (ostc_dev_alloc)."""

    def __init__(self, solver, nbytes):
        self.s = solver
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _chk(lib().ostc_dev_alloc(solver.h, self.nbytes, C.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, solver, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        d = cls(solver, arr.nbytes)
        _chk(lib().ostc_memcpy_h2d(solver.h, d.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes))
        return d

    def to_host(self, shape):
        out = np.empty(shape, dtype=np.float64)
        _chk(lib().ostc_memcpy_d2h(self.s.h, out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            _chk(lib().ostc_dev_free(self.s.h, self.ptr))
            self.ptr = None


class Solver:
    """One solver instance over a reference-format case directory."""

    def __init__(self, case_dir, device=0):
        h = C.c_void_p()
        _chk(lib().ostc_init(case_dir.encode(), device, C.byref(h)))
        self.h = h
        self.root = case_dir
        self.blocks = []

    def close(self):
        if getattr(self, "h", None):
            lib().ostc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get_int(self, key):
        v = C.c_int()
        _chk(lib().ostc_get_int(self.h, key.encode(), C.byref(v)))
        return v.value

    def get_real(self, key):
        v = C.c_double()
        _chk(lib().ostc_get_real(self.h, key.encode(), C.byref(v)))
        return v.value

    def table(self, name, count):
        out = np.zeros(count)
        _chk(lib().ostc_get_table(self.h, name.encode(), _dp(out), count))
        return out

    def add_block(self, mx, my, mz, x, y, z):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.ascontiguousarray(y, dtype=np.float64).ravel()
        z = np.ascontiguousarray(z, dtype=np.float64).ravel()
        b = C.c_int()
        _chk(lib().ostc_add_block(self.h, mx, my, mz, _dp(x), _dp(y), _dp(z), C.byref(b)))
        self.blocks.append((mx, my, mz))
        return b.value

    def add_boundary(self, blk, raw, label):
        _chk(lib().ostc_add_boundary(self.h, blk, *[int(v) for v in raw], int(label)))

    def read_boundary_file(self):
        _chk(lib().ostc_read_boundary_file(self.h))

    def read_grid(self):
        _chk(lib().ostc_read_grid(self.h))

    def read_inner_file(self):
        _chk(lib().ostc_read_inner_file(self.h))

    def read_mesh(self):
        _chk(lib().ostc_read_mesh(self.h))

    def add_inner_face(self, blk, src7, tar7):
        a = np.ascontiguousarray(src7, dtype=np.int32)
        b = np.ascontiguousarray(tar7, dtype=np.int32)
        _chk(lib().ostc_add_inner_face(self.h, blk, a.ctypes.data_as(C.POINTER(C.c_int)),
                                       b.ctypes.data_as(C.POINTER(C.c_int))))

    def inner_faces(self, blk):
        """Registered inner faces of a block: (n, 31) int array."""
        n = C.c_int()
        _chk(lib().ostc_inner_faces(self.h, blk, C.byref(n), None, 0))
        out = np.zeros((max(n.value, 1), 31), dtype=np.int32)
        _chk(lib().ostc_inner_faces(self.h, blk, C.byref(n), out.ctypes.data_as(C.POINTER(C.c_int)), n.value))
        return out[:n.value]

    def exchange(self, with_cv=False):
        _chk(lib().ostc_exchange(self.h, 1 if with_cv else 0))

    def finalize(self):
        _chk(lib().ostc_finalize(self.h))

    def block_count(self):
        n = C.c_int()
        _chk(lib().ostc_block_count(self.h, C.byref(n)))
        return n.value

    def block_dims(self, blk=0):
        d = (C.c_int * 3)()
        _chk(lib().ostc_block_dims(self.h, blk, d))
        return tuple(d)

    def geometry(self, blk=0):
        mx, my, mz = self.block_dims(blk)
        g = self.get_int("ngg")
        n = (mx + 2 * g) * (my + 2 * g) * (mz + 2 * g)
        jac, met = np.zeros(n), np.zeros(9 * n)
        _chk(lib().ostc_block_geometry(self.h, blk, _dp(jac), _dp(met)))
        return jac, met

    def field_info(self, field, blk=0):
        n, g, c = C.c_int64(), C.c_int(), C.c_int()
        _chk(lib().ostc_field_info(self.h, blk, FIELDS[field], C.byref(n), C.byref(g), C.byref(c)))
        return n.value, g.value, c.value

    def download(self, field, blk=0):
        n, g, c = self.field_info(field, blk)
        out = np.zeros(n)
        if n:
            _chk(lib().ostc_download(self.h, blk, FIELDS[field], _dp(out)))
        return out

    def upload(self, field, arr, blk=0):
        n, g, c = self.field_info(field, blk)
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if a.size != n:
            raise OstcError(f"{field}: expected {n} values, got {a.size}")
        _chk(lib().ostc_upload(self.h, blk, FIELDS[field], _dp(a)))

    def device_ptr(self, field, blk=0):
        p = C.c_void_p()
        _chk(lib().ostc_field_device_ptr(self.h, blk, FIELDS[field], C.byref(p)))
        return p.value

    def initialize_uniform(self, blk=0):
        _chk(lib().ostc_initialize_uniform(self.h, blk))

    def inflow_state(self, name, n=7):
        out = np.zeros(n)
        _chk(lib().ostc_inflow_state(self.h, name.encode(), _dp(out), n))
        return out

    def initialize(self):
        _chk(lib().ostc_initialize(self.h))

    def bind_zone(self, mx, my, mz, ngg, n_var, n_scalar, bv, cv, sv=None, qn=None, dq=None, blk=0):
        """Binds DeviceArray buffers (DZone layout) as block blk's state."""
        v = ZoneView(mx, my, mz, ngg, n_var, n_scalar, bv.ptr, sv.ptr if sv else None, cv.ptr,
                     qn.ptr if qn else None, dq.ptr if dq else None)
        _chk(lib().ostc_bind_zone(self.h, blk, C.byref(v)))

    def local_dt(self, blk=0):
        _chk(lib().ostc_local_dt(self.h, blk))

    def store_last(self, blk=0):
        _chk(lib().ostc_store_last(self.h, blk))

    def residual(self, step):
        res = np.zeros(4)
        e = C.c_double()
        _chk(lib().ostc_residual(self.h, step, _dp(res), C.byref(e)))
        return res, e.value

    def rhs(self, blk=0):
        _chk(lib().ostc_rhs(self.h, blk))

    def global_dt(self):
        v = C.c_double()
        _chk(lib().ostc_global_dt(self.h, C.byref(v)))
        return v.value

    def rk3_stage(self, dt, rk, blk=0):
        _chk(lib().ostc_rk3_stage(self.h, blk, dt, rk))

    def apply_bc(self, with_cv, blk=0):
        _chk(lib().ostc_apply_bc(self.h, blk, 1 if with_cv else 0))

    def update_properties(self, blk=0):
        _chk(lib().ostc_update_properties(self.h, blk))

    def step(self, nsteps):
        dt = np.zeros(3 * nsteps)
        res = np.zeros(5 * nsteps)
        _chk(lib().ostc_step_rk3(self.h, nsteps, _dp(dt), _dp(res)))
        return dt.reshape(nsteps, 3), res.reshape(nsteps, 5)

    def step_steady(self, nsteps):
        res = np.zeros((nsteps, 5))
        _chk(lib().ostc_step_steady(self.h, nsteps, _dp(res)))
        return res

    def implicit(self, blk=0):
        _chk(lib().ostc_implicit(self.h, blk))

    def time(self):
        t, s = C.c_double(), C.c_int()
        _chk(lib().ostc_time(self.h, C.byref(t), C.byref(s)))
        return t.value, s.value

    def set_flux_fp32(self, on):
        _chk(lib().ostc_set_flux_precision(self.h, 1 if on else 0))

    def set_chem_split(self, on):
        _chk(lib().ostc_set_chem_split(self.h, 1 if on else 0))

    def set_steady_chem(self, dt, mode=0):
        """Inline chemistry on the steady path (dt <= 0: off)."""
        _chk(lib().ostc_set_steady_chem(self.h, dt, mode))

    def chem_states(self, rho, T, Y, blk=0):
        """Interior rho, T, Y of a block into DeviceArrays."""
        _chk(lib().ostc_chem_states(self.h, blk, rho.ptr, T.ptr, Y.ptr))

    def steady_chem_stats(self):
        c, f = C.c_int64(), C.c_int64()
        _chk(lib().ostc_steady_chem_stats(self.h, C.byref(c), C.byref(f)))
        return c.value, f.value

    def write_flowfield(self, root):
        """<root>/output/field/flowfield.plt and message/step.txt (+ residual_scale.txt)."""
        _chk(lib().ostc_write_flowfield(self.h, str(root).encode()))

    def read_flowfield(self, root):
        """Restart from <root>/output; call initialize() afterwards."""
        _chk(lib().ostc_read_flowfield(self.h, str(root).encode()))

    def chem_step(self, dt, mode=0, blk=0):
        """One operator-split chemistry step on a block; returns the fallback count."""
        nf = C.c_int64()
        _chk(lib().ostc_chem_step(self.h, blk, dt, mode, C.byref(nf)))
        return nf.value

    def chem_batch(self, n, rho, T, Y, dt, dY, dT, mode=0):
        """rho/T/Y/dY/dT: DeviceArray (Y, dY component-major n_spec x n)."""
        nf = C.c_int64()
        _chk(lib().ostc_chem_batch(self.h, n, rho.ptr, T.ptr, Y.ptr, dt, dY.ptr, dT.ptr, mode, C.byref(nf)))
        return nf.value

    def chem_batch_noalloc(self, n, rho, T, Y, dt, dY, dT, mode=0):
        """As chem_batch, through the allocation-free entry point (n <= 2^20)."""
        nf = C.c_int64()
        _chk(lib().ostc_chem_batch_noalloc(self.h, n, rho.ptr, T.ptr, Y.ptr, dt, dY.ptr, dT.ptr, mode,
                                           C.byref(nf)))
        return nf.value

    def sync(self):
        _chk(lib().ostc_sync(self.h))

    # ---- MLP chemistry surrogate
    def surrogate_init(self, seed):
        _chk(lib().ostc_surrogate_init(self.h, seed))

    def surrogate_param_count(self):
        n = C.c_int64()
        _chk(lib().ostc_surrogate_param_count(self.h, C.byref(n)))
        return n.value

    def surrogate_params(self):
        w = np.empty(self.surrogate_param_count(), dtype=np.float32)
        _chk(lib().ostc_surrogate_get_params(self.h, _fp(w)))
        return w

    def surrogate_set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.surrogate_param_count():
            raise OstcError("parameter count mismatch")
        _chk(lib().ostc_surrogate_set_params(self.h, _fp(w)))

    def surrogate_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((x.shape[0], 8), dtype=np.float32)
        _chk(lib().ostc_surrogate_forward(self.h, x.shape[0], _fp(x), _fp(y)))
        return y

    def surrogate_gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.surrogate_param_count(), dtype=np.float32)
        loss = C.c_double()
        _chk(lib().ostc_surrogate_gradient(self.h, x.shape[0], _fp(x), _fp(t), _fp(g), C.byref(loss)))
        return g, loss.value

    def surrogate_train(self, n, rho, T, Y, dt, epochs=8, batch=8192, lr=2e-3, lr_min=2e-5, seed=2004,
                        beta1=0.9, beta2=0.999, eps=1e-8):
        """Train on n device states (DeviceArray); returns the per-epoch loss."""
        cfg = TrainCfg(epochs, batch, lr, lr_min, beta1, beta2, eps, seed)
        hist = np.zeros(epochs)
        _chk(lib().ostc_surrogate_train(self.h, n, rho.ptr, T.ptr, Y.ptr, dt, C.byref(cfg), _dp(hist)))
        return hist

    def surrogate_gate(self, **kw):
        g = Gate()
        _chk(lib().ostc_surrogate_get_gate(self.h, C.byref(g)))
        for k, v in kw.items():
            setattr(g, k, v)
        if kw:
            _chk(lib().ostc_surrogate_set_gate(self.h, C.byref(g)))
        return {k: getattr(g, k) for k, _ in Gate._fields_}

    def surrogate_norm(self):
        o = np.zeros(123)
        _chk(lib().ostc_surrogate_norm(self.h, _dp(o)))
        return dict(x_mean=o[0:11], x_std=o[11:22], x_min=o[22:33], x_max=o[33:44], u_mean=o[44:52],
                    u_std=o[52:60], dt=o[60], basis=o[61:115].reshape(9, 6), c_scale=o[115:123])

    def surrogate_save(self, path):
        _chk(lib().ostc_surrogate_save(self.h, path.encode()))

    def surrogate_load(self, path):
        _chk(lib().ostc_surrogate_load(self.h, path.encode()))

    def surrogate_set_kernel(self, kernel):
        """3 fused (128 cells per workgroup, pipelined halves; default), 1 the same
        unpipelined, 2 fused (64), 0 one GEMM launch per layer."""
        _chk(lib().ostc_surrogate_set_kernel(self.h, kernel))

    def surrogate_network_outputs(self, n, rho, T, Y):
        out = np.empty((n, 8), dtype=np.float32)
        _chk(lib().ostc_surrogate_network_outputs(self.h, n, rho.ptr, T.ptr, Y.ptr, _fp(out)))
        return out

    def surrogate_last_stats(self):
        o = np.zeros(4)
        _chk(lib().ostc_surrogate_last_stats(self.h, _dp(o)))
        return dict(n_fallback=int(o[0]), net_ms=o[1], gate_ms=o[2], fallback_ms=o[3])

    def surrogate_gate_reasons(self):
        """Fallback count of the last inference and the cells per first failing
        gate condition: dict(total, range, nonfinite, dy, dT, sigma, conservation,
        newton, dt)."""
        o = (C.c_int64 * 9)()
        _chk(lib().ostc_surrogate_gate_reasons(self.h, o))
        keys = ("total", "range", "nonfinite", "dy", "dT", "sigma", "conservation", "newton", "dt")
        return dict(zip(keys, (int(v) for v in o)))

    def profile(self, on):
        _chk(lib().ostc_profile(self.h, 1 if on else 0))

    def profile_report(self):
        buf = C.create_string_buffer(1 << 16)
        _chk(lib().ostc_profile_report(self.h, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, calls, ms = line.split()
            out[name] = (int(calls), float(ms))
        return out


def solver_for_mesh_case(mcase, root, device=0, fields=None):
    """Build a solver for a multi-block case (see ostc.cases.split_case)
    through the reference's grid, boundary and inner-face files; fields:
    per-block (bv, sv) to upload before the initialisation (None: no
    initialisation)."""
    from . import cases as _cases
    _cases.write_mesh_case_dir(mcase, root)
    s = Solver(root, device)
    s.read_mesh()
    s.finalize()
    if fields is not None:
        for b, (bv, sv) in enumerate(fields):
            s.upload("bv", bv, blk=b)
            if sv.shape[0]:
                s.upload("sv", sv, blk=b)
        s.initialize()
    return s


def solver_for_case(case, root, device=0, initial=True):
    """Build a solver for a generated case dict (see ostc.cases)."""
    from . import cases as _cases
    _cases.write_case_dir(case, root)
    s = Solver(root, device)
    s.add_block(case["mx"], case["my"], case["mz"], case["x"], case["y"], case["z"])
    s.read_boundary_file()
    s.finalize()
    if initial:
        bv, sv = _cases.initial_fields(case)
        s.upload("bv", bv)
        if sv.shape[0]:
            s.upload("sv", sv)
        s.initialize()
    return s
