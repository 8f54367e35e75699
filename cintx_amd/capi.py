"""Python mirror of the reference's raw operator interface over libcintx_hip.so.

Three layers, matching include/cintx_hip.h:

* ``int2e_sph(shls, atm, bas, env, dims=None)`` & co. -- libcint raw
  signature (libcint-master/src/cint2e.c:1186), returning (out, not0);
* ``eval_raw(api, out, dims, shls, atm, bas, env)`` -- cintx-compat's
  eval_raw contract (crates/cintx-compat/src/raw.rs:874) through the C shim
  ``cintrs_eval`` (crates/cintx-capi/src/shim.rs:371), raising CintxError
  with the shim's status codes;
* ``Plan`` -- the batched executor boundary (device-resident basis, class
  bins, device output).

The library is loaded from this directory.  There is no fallback: if the HIP
library is missing, importing the symbols raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcintx_hip.so")

# crates/cintx-capi/src/shim.rs:8-49 (+ this library's int3c2e_sph id)
INT2E_SPH = 10
INT3C2E_IP1_SPH = 19
INT2E_IP1_SPH = 36
INT3C2E_SPH = 40
INT2C2E_SPH = 13
INT2E_CART = 9
INT2E_SPINOR = 11
INT3C2E_SPINOR = 41

FAMILY = {"int2e_sph": 0, "int2e_ip1_sph": 1, "int3c2e_sph": 2, "int3c2e_ip1_sph": 3,
          "int2c2e_sph": 4, "int2e_ip2_sph": 5, "int3c2e_ip2_sph": 6, "int2e_cart": 7,
          "int3c2e_cart": 8, "int2e_spinor": 9, "int3c2e_spinor": 10}
API_OF_FAMILY = {0: INT2E_SPH, 1: INT2E_IP1_SPH, 2: INT3C2E_SPH, 3: INT3C2E_IP1_SPH,
                 4: INT2C2E_SPH, 7: INT2E_CART, 9: INT2E_SPINOR, 10: INT3C2E_SPINOR}
PREC_FP64 = 0
# launch-shape variants compiled per class (kVariants in csrc/launch.h)
NVARIANTS = 6
PREC_MIXED = 1

STATUS = {
    0: "Success", 1: "InvalidInput", 2: "UnsupportedApi", 3: "UnsupportedRepresentation",
    4: "BufferTooSmall", 5: "MemoryLimitExceeded", 6: "AllocationFailed",
    7: "ExecutionFailed", 8: "NullPointer", 9: "Panic", 10: "BackendNotCompiled",
}

EXPORTS = [
    "int2e_sph", "int2e_ip1_sph", "int3c2e_sph", "int3c2e_ip1_sph", "int2c2e_sph",
    "int2e_ip2_sph", "int3c2e_ip2_sph", "int2e_cart", "int3c2e_cart", "int2e_spinor",
    "int3c2e_spinor", "cintx_spinor_coeffs", "int2e_optimizer", "int2e_ip1_optimizer", "int3c2e_optimizer", "int3c2e_ip1_optimizer",
    "int2c2e_optimizer", "int2e_ip2_optimizer", "int3c2e_ip2_optimizer",
    "CINTdel_optimizer",
    "cintrs_query_workspace", "cintrs_eval", "cintrs_last_error_code",
    "cintrs_clear_last_error", "cintrs_copy_last_error_message",
    "cintx_plan_create", "cintx_plan_create_screened", "cintx_plan_destroy",
    "cintx_plan_sizes", "cintx_plan_info", "cintx_plan_tuples", "cintx_plan_execute", "cintx_plan_sync",
    "cintx_plan_copy_output", "cintx_plan_copy_tuples", "cintx_plan_execute_graph", "cintx_plan_last_time_ms",
    "cintx_plan_work", "cintx_plan_execute_timed", "cintx_plan_jk", "cintx_plan_bin_stats",
    "cintx_plan_autotune", "cintx_plan_set_variant", "cintx_plan_bin_variants",
    "cintx_plan_fill_output", "cintx_plan_jk_noalloc", "cintx_device_alloc", "cintx_device_free", "cintx_device_count",
    "cintx_device_upload", "cintx_device_download",
]


class CintxError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status
        self.message = message


class WorkspaceQuery(ctypes.Structure):
    _fields_ = [("bytes", ctypes.c_size_t), ("alignment", ctypes.c_size_t),
                ("required_bytes", ctypes.c_size_t), ("chunk_count", ctypes.c_size_t),
                ("work_units", ctypes.c_size_t)]


class EvalSummary(ctypes.Structure):
    _fields_ = [("not0", ctypes.c_int32), ("bytes_written", ctypes.c_size_t),
                ("workspace_bytes", ctypes.c_size_t)]


_lib = None
DP = ctypes.POINTER(ctypes.c_double)
IP = ctypes.POINTER(ctypes.c_int)
I64P = ctypes.POINTER(ctypes.c_int64)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run python -m cintx_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        raw = [DP, IP, IP, IP, ctypes.c_int, IP, ctypes.c_int, DP, ctypes.c_void_p, DP]
        for n in ("int2e_sph", "int2e_ip1_sph", "int3c2e_sph", "int3c2e_ip1_sph", "int2c2e_sph",
                  "int2e_ip2_sph", "int3c2e_ip2_sph", "int2e_cart", "int3c2e_cart",
                  "int2e_spinor", "int3c2e_spinor"):
            getattr(L, n).argtypes = raw
            getattr(L, n).restype = ctypes.c_int
        sz = ctypes.c_size_t
        L.cintrs_eval.argtypes = [ctypes.c_int, DP, sz, IP, sz, IP, sz, IP, sz, IP, sz, DP, sz,
                                  ctypes.c_void_p, DP, sz, ctypes.POINTER(EvalSummary)]
        L.cintrs_query_workspace.argtypes = [ctypes.c_int, IP, sz, IP, sz, IP, sz, IP, sz, DP,
                                             sz, ctypes.c_void_p,
                                             ctypes.POINTER(WorkspaceQuery)]
        L.cintrs_last_error_code.restype = ctypes.c_int
        L.cintrs_copy_last_error_message.argtypes = [ctypes.c_char_p, sz]
        L.cintrs_copy_last_error_message.restype = sz
        vp = ctypes.c_void_p
        L.cintx_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, IP, ctypes.c_int, IP,
                                        ctypes.c_int, DP, IP, ctypes.c_int64, ctypes.c_int,
                                        ctypes.c_double]
        L.cintx_plan_create_screened.argtypes = [ctypes.POINTER(vp), IP, ctypes.c_int, IP,
                                                 ctypes.c_int, DP, ctypes.c_double, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_double]
        L.cintx_plan_destroy.argtypes = [vp]
        L.cintx_plan_sizes.argtypes = [vp, I64P, I64P]
        L.cintx_plan_info.argtypes = [vp, I64P, I64P]
        L.cintx_plan_tuples.argtypes = [vp, IP, I64P]
        L.cintx_plan_execute.argtypes = [vp, DP]
        L.cintx_plan_execute_graph.argtypes = [vp, DP]
        L.cintx_plan_sync.argtypes = [vp]
        L.cintx_plan_fill_output.argtypes = [vp, ctypes.c_double]
        L.cintx_plan_jk_noalloc.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.cintx_plan_copy_output.argtypes = [vp, DP]
        L.cintx_plan_copy_tuples.argtypes = [vp, I64P, ctypes.c_int64, DP, I64P]
        L.cintx_plan_last_time_ms.argtypes = [vp, DP]
        L.cintx_plan_work.argtypes = [vp, DP, DP, DP]
        L.cintx_plan_execute_timed.argtypes = [vp, DP]
        L.cintx_plan_jk.argtypes = [vp, DP, DP, DP]
        L.cintx_plan_bin_stats.argtypes = [vp, ctypes.c_int64, IP, I64P, DP, DP, DP]
        L.cintx_plan_autotune.argtypes = [vp, ctypes.c_int]
        L.cintx_plan_set_variant.argtypes = [vp, ctypes.c_int]
        L.cintx_plan_bin_variants.argtypes = [vp, ctypes.c_int64, IP, IP, IP, IP, IP]
        L.cintx_device_count.argtypes = [IP]
        L.cintx_device_alloc.argtypes = [ctypes.POINTER(DP), ctypes.c_int64]
        L.cintx_device_free.argtypes = [DP]
        L.cintx_device_upload.argtypes = [DP, DP, ctypes.c_int64]
        L.cintx_device_download.argtypes = [DP, DP, ctypes.c_int64]
        L.cintx_spinor_coeffs.argtypes = [ctypes.c_int, ctypes.c_int, DP, DP]
        L.cintx_spinor_coeffs.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(DP)


def _ip(a):
    return None if a is None else a.ctypes.data_as(IP)


def _arrays(atm, bas, env):
    return (np.ascontiguousarray(atm, dtype=np.int32), np.ascontiguousarray(bas, dtype=np.int32),
            np.ascontiguousarray(env, dtype=np.float64))


def last_error():
    L = lib()
    buf = ctypes.create_string_buffer(4096)
    L.cintrs_copy_last_error_message(buf, 4096)
    return L.cintrs_last_error_code(), buf.value.decode()


def _check(st):
    if st != 0:
        code, msg = last_error()
        raise CintxError(st, msg)


def arity(family):
    return 2 if family.startswith("int2c2e") else 3 if family.startswith("int3c2e") else 4


def spinor_len(l, kappa):
    """CINTlen_spinor: 4l+2 (kappa 0), 2l+2 (kappa < 0), 2l (kappa > 0)"""
    return 4 * l + 2 if kappa == 0 else (2 * l + 2 if kappa < 0 else 2 * l)


def shell_functions(family, l, kappa=0, slot=0):
    if family.endswith("_spinor"):
        if family.startswith("int3c2e") and slot == 2:
            return 2 * l + 1
        return spinor_len(l, kappa)
    return (l + 1) * (l + 2) // 2 if family.endswith("_cart") else 2 * l + 1


def natural_dims(family, shls, bas):
    return [shell_functions(family, int(bas[s][1]), int(bas[s][4]), a) * int(bas[s][3])
            for a, s in enumerate(shls[:arity(family)])]


def ncomp(family):
    return 3 if "_ip" in family else 1


def is_complex(family):
    return family.endswith("_spinor")


def spinor_coeffs(l, kappa):
    """The library's Cartesian -> spinor coefficients: complex [row][spin][cart]."""
    nf = (l + 1) * (l + 2) // 2
    nd = spinor_len(l, kappa)
    re = np.zeros(max(1, nd * 2 * nf))
    im = np.zeros_like(re)
    n = lib().cintx_spinor_coeffs(l, kappa, _dp(re), _dp(im))
    if n < 0:
        raise CintxError(1, f"l = {l} outside 0..4")
    return (re + 1j * im)[:nd * 2 * nf].reshape(nd, 2, nf)


class Optimizer:
    """A device-resident CINTOpt of one basis (``<family>_optimizer``): the
    pair data of every shell pair, built once and reused by the raw calls
    that pass it.  Immutable, so threads may share it."""

    def __init__(self, family, atm, bas, env):
        self._atm, self._bas, self._env = _arrays(atm, bas, env)
        name = {"int2e_cart": "int2e", "int3c2e_cart": "int3c2e", "int2e_spinor": "int2e",
                "int3c2e_spinor": "int3c2e"}.get(family, family)
        name = name[:-4] if name.endswith("_sph") else name
        self.h = ctypes.c_void_p()
        fn = getattr(lib(), name + "_optimizer")
        fn.argtypes = [ctypes.POINTER(ctypes.c_void_p), IP, ctypes.c_int, IP, ctypes.c_int, DP]
        fn(ctypes.byref(self.h), _ip(self._atm), len(self._atm), _ip(self._bas), len(self._bas),
           _dp(self._env))
        if not self.h.value:
            raise CintxError(7, "optimizer construction failed")

    def close(self):
        if self.h and self.h.value:
            L = lib()
            L.CINTdel_optimizer.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
            L.CINTdel_optimizer(ctypes.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def raw_call(family, shls, atm, bas, env, dims=None, opt=None):
    """libcint raw signature: returns (out, not0).  opt: an Optimizer or None.
    Spinor families return complex128 (the F-order libcint buffer)."""
    atm, bas, env = _arrays(atm, bas, env)
    d = natural_dims(family, shls, bas) if dims is None else list(dims)
    out = np.zeros(int(np.prod(d)) * ncomp(family) * (2 if is_complex(family) else 1))
    sh = np.ascontiguousarray(list(shls), dtype=np.int32)
    dd = None if dims is None else np.ascontiguousarray(dims, dtype=np.int32)
    fn = getattr(lib(), family)
    not0 = fn(_dp(out), _ip(dd), _ip(sh), _ip(atm), len(atm), _ip(bas), len(bas), _dp(env),
              None if opt is None else opt.h, None)
    if is_complex(family):
        out = out.view(np.complex128)
    return out, not0


def cache_size(family, shls, atm, bas, env):
    atm, bas, env = _arrays(atm, bas, env)
    sh = np.ascontiguousarray(list(shls), dtype=np.int32)
    return getattr(lib(), family)(None, None, _ip(sh), _ip(atm), len(atm), _ip(bas), len(bas),
                                  _dp(env), None, None)


def int2e_sph(shls, atm, bas, env, dims=None):
    return raw_call("int2e_sph", shls, atm, bas, env, dims)


def int2e_ip1_sph(shls, atm, bas, env, dims=None):
    return raw_call("int2e_ip1_sph", shls, atm, bas, env, dims)


def int3c2e_sph(shls, atm, bas, env, dims=None):
    return raw_call("int3c2e_sph", shls, atm, bas, env, dims)


def int3c2e_ip1_sph(shls, atm, bas, env, dims=None):
    return raw_call("int3c2e_ip1_sph", shls, atm, bas, env, dims)


def int2c2e_sph(shls, atm, bas, env, dims=None):
    return raw_call("int2c2e_sph", shls, atm, bas, env, dims)


def int2e_cart(shls, atm, bas, env, dims=None):
    return raw_call("int2e_cart", shls, atm, bas, env, dims)


def eval_raw(api, out, dims, shls, atm, bas, env, cache=None):
    """cintx-compat eval_raw through cintrs_eval; returns the summary dict."""
    atm, bas, env = _arrays(atm, bas, env)
    sh = np.ascontiguousarray(shls, dtype=np.int32)
    dd = None if dims is None else np.ascontiguousarray(dims, dtype=np.int32)
    s = EvalSummary()
    st = lib().cintrs_eval(api, _dp(out), 0 if out is None else out.size, _ip(dd),
                           0 if dd is None else dd.size, _ip(sh), sh.size, _ip(atm), atm.size,
                           _ip(bas), bas.size, _dp(env), env.size, None, _dp(cache),
                           0 if cache is None else cache.size, ctypes.byref(s))
    _check(st)
    return {"not0": s.not0, "bytes_written": s.bytes_written,
            "workspace_bytes": s.workspace_bytes}


def query_workspace_raw(api, dims, shls, atm, bas, env):
    atm, bas, env = _arrays(atm, bas, env)
    sh = np.ascontiguousarray(shls, dtype=np.int32)
    dd = None if dims is None else np.ascontiguousarray(dims, dtype=np.int32)
    q = WorkspaceQuery()
    st = lib().cintrs_query_workspace(api, _ip(dd), 0 if dd is None else dd.size, _ip(sh),
                                      sh.size, _ip(atm), atm.size, _ip(bas), bas.size, _dp(env),
                                      env.size, None, ctypes.byref(q))
    _check(st)
    return {"bytes": q.bytes, "required_bytes": q.required_bytes}


class Plan:
    """Batched executor: device-resident basis/pairs, class-binned tuples."""

    def __init__(self, family, atm, bas, env, tuples=None, precision=PREC_FP64,
                 mixed_threshold=0.0, screened=None):
        self._atm, self._bas, self._env = _arrays(atm, bas, env)
        self.family = family
        h = ctypes.c_void_p()
        L = lib()
        if screened is not None:
            thr, stride = screened
            st = L.cintx_plan_create_screened(ctypes.byref(h), _ip(self._atm), len(self._atm),
                                              _ip(self._bas), len(self._bas), _dp(self._env),
                                              float(thr), int(stride), precision,
                                              float(mixed_threshold))
        else:
            if isinstance(tuples, np.ndarray) and tuples.ndim == 2:
                t = np.zeros((len(tuples), 4), dtype=np.int32)
                t[:, :tuples.shape[1]] = tuples
            else:
                t = np.zeros((len(tuples), 4), dtype=np.int32)
                for i, s in enumerate(tuples):
                    t[i, :len(s)] = s
            st = L.cintx_plan_create(ctypes.byref(h), FAMILY[family], _ip(self._atm),
                                     len(self._atm), _ip(self._bas), len(self._bas),
                                     _dp(self._env), _ip(t), len(t), precision,
                                     float(mixed_threshold))
        _check(st)
        self.h = h
        n = ctypes.c_int64()
        nt = ctypes.c_int64()
        _check(L.cintx_plan_sizes(h, ctypes.byref(n), ctypes.byref(nt)))
        self.nelem = n.value
        self.ntuples = nt.value
        n32 = ctypes.c_int64()
        nb = ctypes.c_int64()
        _check(L.cintx_plan_info(h, ctypes.byref(n32), ctypes.byref(nb)))
        self.ntuples_fp32 = n32.value
        self.nbins = nb.value

    def tuples(self):
        t = np.zeros((self.ntuples, 4), dtype=np.int32)
        off = np.zeros(self.ntuples, dtype=np.int64)
        _check(lib().cintx_plan_tuples(self.h, _ip(t), off.ctypes.data_as(I64P)))
        return t, off

    def execute(self, graph=False):
        L = lib()
        _check((L.cintx_plan_execute_graph if graph else L.cintx_plan_execute)(self.h, None))

    def sync(self):
        _check(lib().cintx_plan_sync(self.h))

    def fill_output(self, value):
        """fill the plan's output buffer with `value` (asynchronous)"""
        _check(lib().cintx_plan_fill_output(self.h, float(value)))

    def last_ms(self):
        v = ctypes.c_double()
        _check(lib().cintx_plan_last_time_ms(self.h, ctypes.byref(v)))
        return v.value

    def execute_timed(self):
        """run every bin with HIP events around it (synchronous)"""
        _check(lib().cintx_plan_execute_timed(self.h, None))

    def bin_stats(self):
        n = self.nbins
        cls = np.zeros((n, 6), dtype=np.int32)
        nt = np.zeros(n, dtype=np.int64)
        ms = np.zeros(n)
        pt = np.zeros(n)
        fl = np.zeros(n)
        _check(lib().cintx_plan_bin_stats(self.h, n, _ip(cls), nt.ctypes.data_as(I64P), _dp(ms),
                                          _dp(pt), _dp(fl)))
        return [{"class": tuple(int(v) for v in c[:4]), "precision": int(c[4]),
                 "generic": bool(c[5]), "tuples": int(t), "ms": float(m),
                 "prim_tuples": float(p), "flops": float(f)}
                for c, t, m, p, f in zip(cls, nt, ms, pt, fl)]

    def autotune(self, reps=3):
        """time every compiled launch-shape variant on each bin, keep the fastest"""
        _check(lib().cintx_plan_autotune(self.h, int(reps)))

    def set_variant(self, v):
        _check(lib().cintx_plan_set_variant(self.h, int(v)))

    def bin_variants(self):
        n = self.nbins
        cols = [np.zeros(n, dtype=np.int32) for _ in range(5)]
        _check(lib().cintx_plan_bin_variants(self.h, n, *[_ip(c) for c in cols]))
        return [{"variant": int(a), "nvariants": int(b), "lanes": int(c), "prims": int(d),
                 "tuples_per_block": int(e)} for a, b, c, d, e in zip(*cols)]

    def jk(self, D):
        """J and K against the symmetric density D (unique-quartet plans)."""
        D = np.ascontiguousarray(D, dtype=np.float64)
        J = np.zeros_like(D)
        K = np.zeros_like(D)
        _check(lib().cintx_plan_jk(self.h, _dp(D), _dp(J), _dp(K)))
        return J, K

    def jk_noalloc(self, D_ptr, J_ptr, K_ptr):
        """J/K into caller-owned device buffers (raw device addresses), no
        allocation; asynchronous on the plan stream (sync() waits)."""
        _check(lib().cintx_plan_jk_noalloc(self.h, ctypes.c_void_p(D_ptr), ctypes.c_void_p(J_ptr),
                                           ctypes.c_void_p(K_ptr)))

    def output(self):
        out = np.zeros(self.nelem)
        _check(lib().cintx_plan_copy_output(self.h, _dp(out)))
        return out

    def output_tuples(self, which):
        """libcint-layout blocks of the tuples `which` (packed) and offsets."""
        which = np.ascontiguousarray(which, dtype=np.int64)
        t, _ = self.tuples_cached()
        sizes = np.array([self.block_size(t[q]) for q in which], dtype=np.int64)
        out = np.zeros(int(sizes.sum()))
        offs = np.zeros(len(which), dtype=np.int64)
        _check(lib().cintx_plan_copy_tuples(self.h, which.ctypes.data_as(I64P), len(which),
                                            _dp(out), offs.ctypes.data_as(I64P)))
        return out, offs, sizes

    def tuples_cached(self):
        if not hasattr(self, "_tuples"):
            self._tuples = self.tuples()
        return self._tuples

    def block_size(self, shls):
        """doubles of a tuple's block (two per complex element)"""
        n = (3 if "_ip" in self.family else 1) * (2 if is_complex(self.family) else 1)
        for a, s in enumerate(shls[:arity(self.family)]):
            n *= shell_functions(self.family, int(self._bas[s][1]), int(self._bas[s][4]),
                                 a) * int(self._bas[s][3])
        return n

    def work(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().cintx_plan_work(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def close(self):
        if getattr(self, "h", None):
            lib().cintx_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count():
    n = ctypes.c_int()
    _check(lib().cintx_device_count(ctypes.byref(n)))
    return n.value


class DeviceArray:
    """A device buffer of doubles owned by the library (cintx_device_alloc),
    for callers without a GPU runtime of their own."""

    def __init__(self, nelem):
        self.nelem = int(nelem)
        self.ptr = DP()
        _check(lib().cintx_device_alloc(ctypes.byref(self.ptr), self.nelem))

    @property
    def address(self):
        return ctypes.cast(self.ptr, ctypes.c_void_p).value

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64).ravel()
        assert host.size == self.nelem
        _check(lib().cintx_device_upload(self.ptr, _dp(host), self.nelem))

    def download(self):
        out = np.zeros(self.nelem)
        _check(lib().cintx_device_download(_dp(out), self.ptr, self.nelem))
        return out

    def free(self):
        if self.ptr:
            _check(lib().cintx_device_free(self.ptr))
            self.ptr = DP()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
