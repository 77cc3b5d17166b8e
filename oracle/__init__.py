"""ctypes binding of the oracle (liboracle.so) — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, as the checker / CPU baseline.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

FLAG_TRACKED, FLAG_INFINITY, FLAG_REFINED, FLAG_REAL, FLAG_UNIQUE = 0x01, 0x02, 0x04, 0x08, 0x10
FLAG_BUDGET, FLAG_MINSTEP, FLAG_SINGULAR = 0x20, 0x40, 0x80
FLAG_REAL_REF, FLAG_UNIQUE_REF = 0x200, 0x400


class Tables(C.Structure):
    _fields_ = [("n", C.c_int), ("C", C.c_int), ("M", C.c_int), ("hx_terms", C.c_int), ("hx_parts", C.c_int),
                ("ht_terms", C.c_int), ("ht_parts", C.c_int), ("hx_idx", C.c_void_p), ("ht_idx", C.c_void_p)]


class Opts(C.Structure):
    _fields_ = [("max_steps", C.c_int), ("max_corr", C.c_int), ("incr_steps", C.c_int), ("dt0", C.c_double),
                ("end_zone", C.c_double), ("t_tol", C.c_double), ("dt_min", C.c_double), ("corr_tol", C.c_double),
                ("inf_tol", C.c_double), ("refine_iters", C.c_int), ("refine_tol", C.c_double),
                ("real_tol", C.c_double), ("real_ref_tol", C.c_double)]


DEFAULTS = dict(max_steps=200, max_corr=3, incr_steps=3, dt0=0.05, end_zone=0.05, t_tol=1e-7, dt_min=1e-7,
                corr_tol=1e-6, inf_tol=1e14, refine_iters=8, refine_tol=1e-10, real_tol=1e-8, real_ref_tol=1e-4)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.orc_coeffs.argtypes = [C.POINTER(Tables), C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
        _lib.orc_eval.argtypes = [C.POINTER(Tables), C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                  C.c_void_p, C.c_void_p]
        _lib.orc_lu_solve.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_p2c_affine.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                        C.c_void_p]
        _lib.orc_p2c_poly.argtypes = [C.c_int] * 5 + [C.c_void_p] * 4 + [C.c_int, C.c_void_p, C.c_void_p]
        _lib.orc_track.argtypes = [C.POINTER(Tables), C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                   C.POINTER(Opts), C.c_int] + [C.c_void_p] * 8
        _lib.orc_collect.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class OracleProblem:
    """Reference-format tables (dHdx_indx / dHdt_indx layouts)."""

    def __init__(self, n, num_coeffs, hx_idx, ht_idx, max_order_t=1):
        self.hx = np.ascontiguousarray(hx_idx, dtype=np.int32)
        self.ht = np.ascontiguousarray(ht_idx, dtype=np.int32)
        self.n, self.C, self.M = n, num_coeffs, max_order_t
        self.t = Tables(n, num_coeffs, max_order_t, self.hx.shape[-2], self.hx.shape[-1], self.ht.shape[-2],
                        self.ht.shape[-1], self.hx.ctypes.data, self.ht.ctypes.data)

    def coeffs(self, hx, ht, t):
        c = np.empty(self.C + 1, dtype=np.complex128)
        d = np.empty(self.C + 1, dtype=np.complex128)
        hx = np.ascontiguousarray(hx, dtype=np.complex128)
        ht = np.ascontiguousarray(ht, dtype=np.complex128)
        lib().orc_coeffs(C.byref(self.t), _p(hx), _p(ht), t, _p(c), _p(d))
        return c, d

    def eval(self, hx, ht, x, t):
        n = self.n
        H = np.empty(n, dtype=np.complex128)
        J = np.empty((n, n), dtype=np.complex128)
        Ht = np.empty(n, dtype=np.complex128)
        hx = np.ascontiguousarray(hx, dtype=np.complex128)
        ht = np.ascontiguousarray(ht, dtype=np.complex128)
        x = np.ascontiguousarray(x, dtype=np.complex128)
        lib().orc_eval(C.byref(self.t), _p(hx), _p(ht), _p(x), t, _p(H), _p(J), _p(Ht))
        return H, J, Ht

    def track(self, start_sols, coeffs_hx, coeffs_ht, threads=0, **opts):
        o = dict(DEFAULTS)
        o.update(opts)
        oo = Opts(**o)
        start = np.ascontiguousarray(start_sols, dtype=np.complex128).reshape(-1, self.n)
        K = start.shape[0]
        hx = np.ascontiguousarray(coeffs_hx, dtype=np.complex128).reshape(-1, (self.C + 1) * (self.M + 1))
        ht = np.ascontiguousarray(coeffs_ht, dtype=np.complex128).reshape(-1, (self.C + 1) * self.M)
        I = hx.shape[0]
        P = I * K
        out = dict(roots=np.empty((P, self.n), dtype=np.complex128), flags=np.empty(P, dtype=np.uint32),
                   steps=np.empty(P, dtype=np.int32), newton=np.empty(P, dtype=np.int32),
                   refine=np.empty(P, dtype=np.int32), t_final=np.empty(P), dt_final=np.empty(P),
                   residual=np.empty(P))
        lib().orc_track(C.byref(self.t), K, _p(start), I, _p(hx), _p(ht), C.byref(oo), threads,
                        _p(out["roots"]), _p(out["flags"]), _p(out["steps"]), _p(out["newton"]), _p(out["refine"]),
                        _p(out["t_final"]), _p(out["dt_final"]), _p(out["residual"]))
        return out


def lu_solve(A, b):
    A = np.ascontiguousarray(A, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    n = A.shape[0]
    x = np.empty(n, dtype=np.complex128)
    piv = np.empty(n, dtype=np.int32)
    info = lib().orc_lu_solve(n, _p(A), _p(b), _p(x), _p(piv))
    return x, piv, info


def p2c_affine(affine, p0, p1):
    affine = np.ascontiguousarray(affine, dtype=np.complex128)
    Cn, k1 = affine.shape
    p0 = np.ascontiguousarray(p0, dtype=np.complex128)
    p1 = np.ascontiguousarray(p1, dtype=np.complex128).reshape(-1, k1 - 1)
    I = p1.shape[0]
    hx = np.empty((I, Cn + 1, 2), dtype=np.complex128)
    ht = np.empty((I, Cn + 1, 1), dtype=np.complex128)
    lib().orc_p2c_affine(Cn, k1 - 1, _p(affine), _p(p0), _p(p1), I, _p(hx), _p(ht))
    return hx, ht


def p2c_poly(scalar, params, p0, p1, M):
    """Coefficient-polynomial tables of a parameter-polynomial coefficient map
    for each target row of p1: hx [I][C+1][M+1], ht [I][C+1][M]."""
    scalar = np.ascontiguousarray(scalar, dtype=np.complex128)
    params = np.ascontiguousarray(params, dtype=np.int32)
    Cn, T = scalar.shape
    D = params.shape[2]
    p0 = np.ascontiguousarray(p0, dtype=np.complex128)
    k = p0.shape[0]
    p1 = np.ascontiguousarray(p1, dtype=np.complex128).reshape(-1, k)
    I = p1.shape[0]
    hx = np.empty((I, Cn + 1, M + 1), dtype=np.complex128)
    ht = np.empty((I, Cn + 1, M), dtype=np.complex128)
    if lib().orc_p2c_poly(Cn, T, D, k, M, _p(scalar), _p(params), _p(p0), _p(p1), I, _p(hx), _p(ht)) != 0:
        raise ValueError("p2c_poly: parameter index out of range or degree above M")
    return hx, ht


def collect(n, I, K, roots, flags, tol=1e-4):
    roots = np.ascontiguousarray(roots, dtype=np.complex128)
    flags = np.ascontiguousarray(flags, dtype=np.uint32).copy()
    counts = np.zeros(7, dtype=np.int64)
    lib().orc_collect(n, I, K, _p(roots), _p(flags), tol, _p(counts))
    return dict(zip(["tracked", "infinity", "refined", "real", "unique", "unique_ref", "real_ref"],
                    counts.tolist())), flags
