"""ctypes binding of the CPU oracle (oracle/liboracle_cint.so) and of its
extended-precision build (oracle/liboracle_cint_ext.so: the same algorithm
in 80-bit long double, the judge of FP64 rounding on ill-conditioned
quartets; pass ext=True).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product package never
imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_cint.so")
EXT_PATH = os.path.join(HERE, "liboracle_cint_ext.so")
FMA_PATH = os.path.join(HERE, "liboracle_cint_fma.so")
CPU_PATH = os.path.join(HERE, "liboracle_cpu.so")
# build selector of lib()/eval_*(): False the FP64 oracle, True the long
# double judge, "fma" the FP64 oracle with fused multiply-add contraction
_PATHS = {False: LIB_PATH, True: EXT_PATH, "fma": FMA_PATH}

API_INT2E_SPH = 0
API_INT2E_IP1_SPH = 1
API_INT3C2E_SPH = 2
API_INT3C2E_IP1_SPH = 3
API_INT2E_CART = 4

_libs = {}


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib(ext=False):
    path = _PATHS[ext]
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_boys.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.orc_rys_roots.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp]
        L.orc_rys_roots_m.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp, ctypes.c_int]
        ldp = ctypes.POINTER(ctypes.c_longdouble)
        L.orc_rys_rule_ld.argtypes = [ctypes.c_int, ctypes.c_double, ldp, ldp]
        L.orc_c2s_coef.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_c2s_coef.restype = ctypes.c_double
        sig = [dp, ip, ip, ip, ctypes.c_int, ip, ctypes.c_int, dp, ctypes.c_void_p, dp]
        for name in ("orc_int2e_sph", "orc_int2e_cart", "orc_int2e_ip1_sph",
                     "orc_int3c2e_sph", "orc_int3c2e_ip1_sph", "orc_int2c2e_sph",
                     "orc_int2e_ip2_sph", "orc_int3c2e_ip2_sph", "orc_int3c2e_cart"):
            getattr(L, name).argtypes = sig
            getattr(L, name).restype = ctypes.c_int
        L.orc_eval_list.argtypes = [ctypes.c_int, ctypes.c_int, ip,
                                    ctypes.POINTER(ctypes.c_long), dp, ip, ctypes.c_int,
                                    ip, ctypes.c_int, dp, ctypes.c_int]
        L.orc_eval_list.restype = ctypes.c_long
        L.orc_work_model.argtypes = [ctypes.c_int, ip, ip, ctypes.c_int, ip, ctypes.c_int,
                                     dp, dp]
        _libs[path] = L
    return _libs[path]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def boys(mmax, x):
    f = np.zeros(mmax + 1)
    lib().orc_boys(mmax, float(x), _dp(f))
    return f


def pairdata_one(ai, aj, ri, rj):
    """(P[3], exp(-ai aj/(ai+aj) r^2), cce) of one primitive pair."""
    ri = np.ascontiguousarray(ri, dtype=np.float64)
    rj = np.ascontiguousarray(rj, dtype=np.float64)
    out = np.zeros(5)
    lib().orc_pairdata_one(ctypes.c_double(ai), ctypes.c_double(aj), _dp(ri), _dp(rj), _dp(out))
    return out[:3], out[3], out[4]


def rys_roots(n, x, nodes=0):
    u = np.zeros(n)
    w = np.zeros(n)
    if nodes:
        lib().orc_rys_roots_m(n, float(x), _dp(u), _dp(w), nodes)
    else:
        lib().orc_rys_roots(n, float(x), _dp(u), _dp(w))
    return u, w


def rys_rule_ld(n, x):
    """(s, w) of the n-point rule in long double (numpy longdouble arrays)."""
    s = np.zeros(n, dtype=np.longdouble)
    w = np.zeros(n, dtype=np.longdouble)
    ldp = ctypes.POINTER(ctypes.c_longdouble)
    lib().orc_rys_rule_ld(n, float(x), s.ctypes.data_as(ldp), w.ctypes.data_as(ldp))
    return s, w


def c2s_matrix(l):
    nc = (l + 1) * (l + 2) // 2
    ns = 2 * l + 1
    return np.array([[lib().orc_c2s_coef(l, m, c) for c in range(nc)] for m in range(ns)])


_FN = {
    "int2e_sph": "orc_int2e_sph",
    "int2e_cart": "orc_int2e_cart",
    "int2e_ip1_sph": "orc_int2e_ip1_sph",
    "int3c2e_sph": "orc_int3c2e_sph",
    "int3c2e_ip1_sph": "orc_int3c2e_ip1_sph",
    "int2c2e_sph": "orc_int2c2e_sph",
    "int2e_ip2_sph": "orc_int2e_ip2_sph",
    "int3c2e_ip2_sph": "orc_int3c2e_ip2_sph",
    "int3c2e_cart": "orc_int3c2e_cart",
}
_API = {"int2e_sph": 0, "int2e_ip1_sph": 1, "int3c2e_sph": 2, "int3c2e_ip1_sph": 3,
        "int2e_cart": 4, "int2c2e_sph": 5, "int2e_ip2_sph": 6, "int3c2e_ip2_sph": 7,
        "int3c2e_cart": 8}


def ncomp(name):
    return 3 if "_ip" in name else 1


def shell_dim(bas, sh, cart=False):
    l = int(bas[sh, 1])
    n = (l + 1) * (l + 2) // 2 if cart else 2 * l + 1
    return n * int(bas[sh, 3])


def tuple_dims(name, bas, shls):
    nsh = 2 if name.startswith("int2c2e") else 3 if name.startswith("int3c2e") else 4
    cart = name.endswith("cart")
    return [shell_dim(bas, s, cart) for s in shls[:nsh]]


def eval_tuple(name, shls, atm, bas, env, dims=None, out=None, ext=False):
    """One libcint-signature call; returns (out array (F-order block), not0)."""
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    d = tuple_dims(name, bas, shls)
    if out is None:
        n = int(np.prod(dims if dims is not None else d)) * ncomp(name)
        out = np.zeros(n)
    sh = np.ascontiguousarray(list(shls) + [0] * (4 - len(shls)), dtype=np.int32)
    dims_arr = None
    if dims is not None:
        dims_arr = np.ascontiguousarray(dims, dtype=np.int32)
    fn = getattr(lib(ext), _FN[name])
    not0 = fn(_dp(out), _ip(dims_arr) if dims_arr is not None else None, _ip(sh), _ip(atm),
              len(atm), _ip(bas), len(bas), _dp(env), None, None)
    return out, not0


def eval_list(name, shls_list, atm, bas, env, nthreads=0, ext=False):
    """Evaluate many tuples (OpenMP); returns (flat out, offsets)."""
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    sl = np.zeros((len(shls_list), 4), dtype=np.int32)
    for q, s in enumerate(shls_list):
        sl[q, :len(s)] = s
    sizes = np.array([int(np.prod(tuple_dims(name, bas, s))) * ncomp(name)
                      for s in sl], dtype=np.int64)
    offsets = np.zeros(len(sl), dtype=np.int64)
    if len(sl):
        offsets[1:] = np.cumsum(sizes)[:-1]
    out = np.zeros(int(sizes.sum()))
    lib(ext).orc_eval_list(_API[name], len(sl), _ip(sl),
                        offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), _dp(out),
                        _ip(atm), len(atm), _ip(bas), len(bas), _dp(env), nthreads)
    return out, offsets, sizes


def work_model(name, shls, atm, bas, env):
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    sh = np.ascontiguousarray(list(shls) + [0] * (4 - len(shls)), dtype=np.int32)
    c = np.zeros(3)
    lib().orc_work_model(_API[name], _ip(sh), _ip(atm), len(atm), _ip(bas), len(bas),
                         _dp(env), _dp(c))
    return c


def full_tensor_int2e(atm, bas, env, nthreads=0, ext=False):
    """Full (nao,nao,nao,nao) int2e_sph tensor, libcint F-order per block."""
    bas = np.asarray(bas)
    nbas = len(bas)
    dims = [shell_dim(bas, s) for s in range(nbas)]
    offs = np.concatenate([[0], np.cumsum(dims)])
    nao = int(offs[-1])
    shls = [(i, j, k, l) for l in range(nbas) for k in range(nbas)
            for j in range(nbas) for i in range(nbas)]
    out, offsets, sizes = eval_list("int2e_sph", shls, atm, bas, env, nthreads, ext=ext)
    eri = np.zeros((nao, nao, nao, nao))
    for q, (i, j, k, l) in enumerate(shls):
        blk = out[offsets[q]:offsets[q] + sizes[q]].reshape(
            dims[l], dims[k], dims[j], dims[i]).transpose(3, 2, 1, 0)
        eri[offs[i]:offs[i + 1], offs[j]:offs[j + 1], offs[k]:offs[k + 1],
            offs[l]:offs[l + 1]] = blk
    return eri


_cpu = None


def cpu_lib():
    """The CPU baseline (oracle/cpu_simd.cpp: cintx's f64x4 CPU kernel
    restated), with its own copy of the oracle for its tables."""
    global _cpu
    if _cpu is None:
        if not os.path.exists(CPU_PATH):
            build()
        L = ctypes.CDLL(CPU_PATH)
        ip = ctypes.POINTER(ctypes.c_int)
        L.cpus_int2e_sph.argtypes = [ctypes.c_int, ip, ctypes.POINTER(ctypes.c_long),
                                     ctypes.POINTER(ctypes.c_double), ip, ctypes.c_int, ip,
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.cpus_int2e_sph.restype = ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        L.cpus_int2e_jk.argtypes = [ctypes.c_int, ip, ip, ctypes.c_int, ip, ctypes.c_int, dp, dp,
                                    dp, dp, ctypes.c_int]
        _cpu = L
    return _cpu


def cpu_baseline_int2e(shls_list, atm, bas, env, nthreads=0, keep=False):
    """int2e_sph of many quartets on the CPU baseline; returns (out, offsets,
    sizes) when keep, else the checksum alone."""
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    sl = np.ascontiguousarray(np.asarray(shls_list, dtype=np.int32).reshape(-1, 4))
    if keep:
        sizes = np.array([int(np.prod(tuple_dims("int2e_sph", bas, s))) for s in sl],
                         dtype=np.int64)
        offsets = np.zeros(len(sl), dtype=np.int64)
        if len(sl):
            offsets[1:] = np.cumsum(sizes)[:-1]
        out = np.zeros(int(sizes.sum()))
        chk = cpu_lib().cpus_int2e_sph(len(sl), _ip(sl), offsets.ctypes.data_as(
            ctypes.POINTER(ctypes.c_long)), _dp(out), _ip(atm), len(atm), _ip(bas), len(bas),
            _dp(env), nthreads)
        if np.isnan(chk):
            raise ValueError("the CPU baseline takes segmented shells with l <= 4")
        return out, offsets, sizes
    return cpu_lib().cpus_int2e_sph(len(sl), _ip(sl), None, None, _ip(atm), len(atm), _ip(bas),
                                    len(bas), _dp(env), nthreads)


def cpu_baseline_jk(shls_list, atm, bas, env, D, nthreads=0):
    """J and K of a unique-quartet list on the CPU baseline (every distinct
    8-fold image of each quartet added): an independent check of the device
    J/K digestion."""
    atm = np.ascontiguousarray(atm, dtype=np.int32)
    bas = np.ascontiguousarray(bas, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.float64)
    D = np.ascontiguousarray(D, dtype=np.float64)
    sl = np.ascontiguousarray(np.asarray(shls_list, dtype=np.int32).reshape(-1, 4))
    J = np.zeros_like(D)
    K = np.zeros_like(D)
    st = cpu_lib().cpus_int2e_jk(len(sl), _ip(sl), _ip(atm), len(atm), _ip(bas), len(bas),
                                 _dp(env), _dp(D), _dp(J), _dp(K), nthreads)
    if st != 0:
        raise ValueError("the CPU baseline takes segmented shells with l <= 4")
    return J, K
