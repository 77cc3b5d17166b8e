"""Python binding of libhcx (the MI355X homotopy-continuation tracker).

Thin ctypes layer over the C-ABI in include/hcx.h.  Device buffers are torch
tensors (torch is plumbing for HBM allocations and streams; all tracking runs
in libhcx's HIP kernels).  Host-side system generation, table compilation and
P2C work without a GPU.
"""
import ctypes as C

import numpy as np

from . import _capi
from ._capi import (BatchDesc, Counts, HcxComplex, ProblemDesc, ProblemInfo, ResultDesc, SynthSpec,
                    SystemInfo, TrackOpts, RefSettings)

FLAG_TRACKED = 0x01
FLAG_INFINITY = 0x02
FLAG_REFINED = 0x04
FLAG_REAL = 0x08
FLAG_UNIQUE = 0x10
FLAG_BUDGET = 0x20
FLAG_MINSTEP = 0x40
FLAG_SINGULAR = 0x80
FLAG_ESCALATED = 0x100
FLAG_REAL_REF = 0x200
FLAG_UNIQUE_REF = 0x400

PREC_FP64 = 0
PREC_MIXED = 1

_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _capi.load()
    return _lib


class HcxError(RuntimeError):
    pass


def _check(st, what):
    if st != 0:
        msg = lib().hcx_last_error()
        raise HcxError(f"{what} failed (status {st}): {msg.decode() if msg else ''}")


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def device_count():
    return lib().hcx_device_count()


def default_opts(**kw):
    o = TrackOpts()
    lib().hcx_track_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def random_cn(seed, count):
    out = np.empty(count, dtype=np.complex128)
    _check(lib().hcx_random_cn(C.c_uint64(seed), count, _ptr(out)), "hcx_random_cn")
    return out


def random_gamma(seed):
    z = lib().hcx_random_gamma(C.c_uint64(seed))
    return complex(z.re, z.im)


class System:
    """Square polynomial system with parameter-affine coefficients (host)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    @classmethod
    def synthetic(cls, n, num_params, num_quadratic=None, monomials=0, seed=0):
        spec = SynthSpec(n, num_params, n if num_quadratic is None else num_quadratic, monomials, seed)
        h = C.c_void_p()
        _check(lib().hcx_system_create_synthetic(C.byref(spec), C.byref(h)), "hcx_system_create_synthetic")
        return cls(h)

    @classmethod
    def from_monomials(cls, exps, eq, affine=None, num_params=0):
        exps = np.ascontiguousarray(exps, dtype=np.int32)
        eq = np.ascontiguousarray(eq, dtype=np.int32)
        Cn, n = exps.shape
        aff = None if affine is None else np.ascontiguousarray(affine, dtype=np.complex128)
        h = C.c_void_p()
        _check(lib().hcx_system_create(n, num_params, Cn, _ptr(exps), _ptr(eq), _ptr(aff), C.byref(h)),
               "hcx_system_create")
        return cls(h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hcx_system_destroy(self._h)
            self._h = None

    @property
    def info(self):
        i = SystemInfo()
        _check(lib().hcx_system_get_info(self._h, C.byref(i)), "hcx_system_get_info")
        return {f: getattr(i, f) for f, _ in SystemInfo._fields_}

    def monomials(self):
        inf = self.info
        exps = np.empty((inf["num_coeffs"], inf["num_vars"]), dtype=np.int32)
        eq = np.empty(inf["num_coeffs"], dtype=np.int32)
        _check(lib().hcx_system_get_monomials(self._h, _ptr(exps), _ptr(eq)), "get_monomials")
        return exps, eq

    def affine(self):
        inf = self.info
        a = np.empty((inf["num_coeffs"], 1 + inf["num_params"]), dtype=np.complex128)
        _check(lib().hcx_system_get_affine(self._h, _ptr(a)), "get_affine")
        return a

    def tables(self):
        inf = self.info
        n = inf["num_vars"]
        hx = np.empty((n, n, inf["hx_terms"], inf["hx_parts"]), dtype=np.int32)
        ht = np.empty((n, inf["ht_terms"], inf["ht_parts"]), dtype=np.int32)
        _check(lib().hcx_system_compile_tables(self._h, _ptr(hx), _ptr(ht)), "compile_tables")
        return hx, ht

    def eval_coeffs(self, params):
        inf = self.info
        p = np.ascontiguousarray(params, dtype=np.complex128) if inf["num_params"] else None
        out = np.empty(inf["num_coeffs"], dtype=np.complex128)
        _check(lib().hcx_system_eval_coeffs(self._h, _ptr(p), _ptr(out)), "eval_coeffs")
        return out

    def param_homotopy_coeffs(self, p0, p1):
        Cn = self.info["num_coeffs"]
        p0 = np.ascontiguousarray(p0, dtype=np.complex128)
        p1 = np.ascontiguousarray(p1, dtype=np.complex128)
        hx = np.empty((Cn + 1, 2), dtype=np.complex128)
        ht = np.empty((Cn + 1, 1), dtype=np.complex128)
        _check(lib().hcx_system_param_homotopy_coeffs(self._h, _ptr(p0), _ptr(p1), _ptr(hx), _ptr(ht)),
               "param_homotopy_coeffs")
        return hx, ht

    def total_degree(self, target_coeffs, gamma):
        """Returns (merged System, hx [C'+1,2], ht [C'+1,1]) of (1-t) gamma G + t F."""
        tc = np.ascontiguousarray(target_coeffs, dtype=np.complex128)
        g = HcxComplex(gamma.real, gamma.imag)
        cnt = C.c_int()
        _check(lib().hcx_total_degree_homotopy(self._h, _ptr(tc), g, None, C.byref(cnt), None, None), "tdh size")
        c2 = cnt.value
        hx = np.empty((c2 + 1, 2), dtype=np.complex128)
        ht = np.empty((c2 + 1, 1), dtype=np.complex128)
        h = C.c_void_p()
        _check(lib().hcx_total_degree_homotopy(self._h, _ptr(tc), g, C.byref(h), C.byref(cnt), _ptr(hx), _ptr(ht)),
               "hcx_total_degree_homotopy")
        return System(h), hx, ht

    def start_roots(self, indices):
        idx = np.ascontiguousarray(indices, dtype=np.int64)
        out = np.empty((len(idx), self.info["num_vars"]), dtype=np.complex128)
        _check(lib().hcx_total_degree_start_roots(self._h, len(idx), _ptr(idx), _ptr(out)), "start_roots")
        return out

    def subset(self, count, seed):
        out = np.empty(count, dtype=np.int64)
        _check(lib().hcx_total_degree_subset(self._h, count, C.c_uint64(seed), _ptr(out)), "subset")
        return out


class Problem:
    """Device-resident problem type: tables + start solutions (+ a coefficient map:
    affine, or parameter-polynomial given as cmap = (scalar [C][T], params [C][T][D]))."""

    def __init__(self, n, num_coeffs, hx_idx, ht_idx, start_sols, max_order_t=1, affine=None, p0=None,
                 cmap=None):
        hx_idx = np.ascontiguousarray(hx_idx, dtype=np.int32)
        ht_idx = np.ascontiguousarray(ht_idx, dtype=np.int32)
        self.start_sols = np.ascontiguousarray(start_sols, dtype=np.complex128).reshape(-1, n)
        self._keep = [hx_idx, ht_idx, self.start_sols]
        d = ProblemDesc()
        d.num_vars = n
        d.num_coeffs = num_coeffs
        d.max_order_t = max_order_t
        d.hx_terms, d.hx_parts = hx_idx.shape[-2], hx_idx.shape[-1]
        d.ht_terms, d.ht_parts = ht_idx.shape[-2], ht_idx.shape[-1]
        d.h_hx_idx = hx_idx.ctypes.data_as(C.POINTER(C.c_int))
        d.h_ht_idx = ht_idx.ctypes.data_as(C.POINTER(C.c_int))
        d.num_tracks = self.start_sols.shape[0]
        d.h_start_sols = self.start_sols.ctypes.data
        if affine is not None:
            affine = np.ascontiguousarray(affine, dtype=np.complex128)
            p0 = np.ascontiguousarray(p0, dtype=np.complex128)
            self._keep += [affine, p0]
            d.num_params = affine.shape[1] - 1
            d.h_coeff_affine = affine.ctypes.data
            d.h_start_params = p0.ctypes.data
        elif cmap is not None:
            scal = np.ascontiguousarray(cmap[0], dtype=np.complex128)
            prm = np.ascontiguousarray(cmap[1], dtype=np.int32)
            p0 = np.ascontiguousarray(p0, dtype=np.complex128)
            self._keep += [scal, prm, p0]
            d.num_params = p0.shape[0]
            d.cmap_terms, d.cmap_deg = prm.shape[1], prm.shape[2]
            d.h_cmap_scalar = scal.ctypes.data
            d.h_cmap_params = prm.ctypes.data_as(C.POINTER(C.c_int))
            d.h_start_params = p0.ctypes.data
        self.n = n
        self.num_coeffs = num_coeffs
        self.num_tracks = d.num_tracks
        self.max_order_t = max_order_t
        self.num_params = d.num_params
        h = C.c_void_p()
        _check(lib().hcx_problem_create(C.byref(d), C.byref(h)), "hcx_problem_create")
        self._h = h

    @classmethod
    def from_system(cls, system, start_sols, p0=None, with_affine=True):
        inf = system.info
        hx, ht = system.tables()
        aff = system.affine() if (with_affine and inf["num_params"] > 0) else None
        return cls(inf["num_vars"], inf["num_coeffs"], hx, ht, start_sols,
                   affine=aff, p0=p0 if aff is not None else None)

    @property
    def handle(self):
        return self._h

    @property
    def info(self):
        i = ProblemInfo()
        _check(lib().hcx_problem_get_info(self._h, C.byref(i)), "hcx_problem_get_info")
        return {f: getattr(i, f) for f, _ in ProblemInfo._fields_}

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hcx_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class _DeviceBuffers:
    """Device tensors allocated (and zero-filled or uploaded) on the stream
    that was current at construction.  A launch on another stream first
    waits for that initialisation (an event recorded at construction, so the
    wait does not order the launch behind later, unrelated work), and the
    tensors are recorded as used by the launch stream, so the caching
    allocator does not hand them out while the launch still runs."""

    def _mark_ready(self):
        import torch
        self._ready = torch.cuda.Event()
        self._ready.record(torch.cuda.current_stream())
        self._streams = set()

    def _use_on(self, s):
        key = s.cuda_stream
        if key in self._streams:
            return
        s.wait_event(self._ready)
        for t in self.tensors():
            t.record_stream(s)
        self._streams.add(key)


class Result(_DeviceBuffers):
    """Per-path device outputs of one segment (torch tensors on cuda)."""

    def __init__(self, P, n, I, device="cuda"):
        import torch
        self.P, self.n, self.I = P, n, I
        self.roots = torch.zeros((max(P, 1), n), dtype=torch.complex128, device=device)
        self.flags = torch.zeros(max(P, 1), dtype=torch.int32, device=device)
        self.steps = torch.zeros(max(P, 1), dtype=torch.int32, device=device)
        self.newton = torch.zeros(max(P, 1), dtype=torch.int32, device=device)
        self.refine = torch.zeros(max(P, 1), dtype=torch.int32, device=device)
        self.t_final = torch.zeros(max(P, 1), dtype=torch.float64, device=device)
        self.dt_final = torch.zeros(max(P, 1), dtype=torch.float64, device=device)
        self.residual = torch.zeros(max(P, 1), dtype=torch.float64, device=device)
        self.inst_done = torch.zeros(max(I, 1), dtype=torch.int64, device=device)
        self.batch_start = torch.zeros(1, dtype=torch.int64, device=device)
        self._mark_ready()

    def tensors(self):
        return (self.roots, self.flags, self.steps, self.newton, self.refine, self.t_final, self.dt_final,
                self.residual, self.inst_done, self.batch_start)

    def desc(self):
        r = ResultDesc()
        r.d_roots = self.roots.data_ptr()
        r.d_flags = self.flags.data_ptr()
        r.d_steps = self.steps.data_ptr()
        r.d_newton = self.newton.data_ptr()
        r.d_refine = self.refine.data_ptr()
        r.d_t_final = self.t_final.data_ptr()
        r.d_dt_final = self.dt_final.data_ptr()
        r.d_residual = self.residual.data_ptr()
        r.d_inst_done = self.inst_done.data_ptr()
        r.d_batch_start = self.batch_start.data_ptr()
        return r

    def to_host(self):
        out = {}
        for k in ("roots", "flags", "steps", "newton", "refine", "t_final", "dt_final", "residual", "inst_done",
                  "batch_start"):
            out[k] = getattr(self, k).cpu().numpy()
        out["roots"] = out["roots"][: self.P]
        for k in ("flags", "steps", "newton", "refine", "t_final", "dt_final", "residual"):
            out[k] = out[k][: self.P]
        out["flags"] = out["flags"].view(np.uint32)
        return out


class Batch(_DeviceBuffers):
    """Instances of one problem: coefficient tables or target parameters on device."""

    def __init__(self, problem, num_instances, coeffs_hx=None, coeffs_ht=None, target_params=None,
                 start_params=None, device="cuda"):
        import torch
        self.problem = problem
        self.I = num_instances
        self._t = []

        def dev(a):
            if a is None:
                return None
            if isinstance(a, np.ndarray):
                a = torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).to(device)
            self._t.append(a)
            return a

        self.coeffs_hx = dev(coeffs_hx)
        self.coeffs_ht = dev(coeffs_ht)
        self.target_params = dev(target_params)
        self.start_params = dev(start_params)
        self.work = None
        if self.coeffs_hx is None:
            self.work = torch.zeros(max(num_instances, 1) * (problem.num_coeffs + 1) * (2 * problem.max_order_t + 1),
                                    dtype=torch.complex128, device=device)
        self._mark_ready()

    def tensors(self):
        return tuple(self._t) + ((self.work,) if self.work is not None else ())

    def desc(self):
        b = BatchDesc()
        b.num_instances = self.I
        b.d_coeffs_hx = self.coeffs_hx.data_ptr() if self.coeffs_hx is not None else None
        b.d_coeffs_ht = self.coeffs_ht.data_ptr() if self.coeffs_ht is not None else None
        b.d_target_params = self.target_params.data_ptr() if self.target_params is not None else None
        b.d_start_params = self.start_params.data_ptr() if self.start_params is not None else None
        b.d_coeff_work = self.work.data_ptr() if self.work is not None else None
        return b


def _stream_handle(stream, uses=()):
    """HIP stream of a launch (torch's current stream when None); the
    buffers in `uses` are made safe for it (see _DeviceBuffers)."""
    import torch
    cur = torch.cuda.current_stream()
    s = stream if stream is not None else cur
    if s != cur:
        for obj in uses:
            obj._use_on(s)
    return C.c_void_p(s.cuda_stream)


def track(problem, batch, opts, result=None, static=False, stream=None):
    """Launches the tracker for all paths of `batch`; returns the Result (async)."""
    if result is None:
        result = Result(batch.I * problem.num_tracks, problem.n, batch.I)
    b, r = batch.desc(), result.desc()
    fn = lib().hcx_track_static if static else lib().hcx_track
    _check(fn(problem.handle, C.byref(b), C.byref(opts), C.byref(r), _stream_handle(stream, (batch, result))),
           "hcx_track")
    return result


def track_multi(problems, batches, opts, results=None, stream=None):
    if results is None:
        results = [Result(b.I * p.num_tracks, p.n, b.I) for p, b in zip(problems, batches)]
    ns = len(problems)
    ph = (C.c_void_p * ns)(*[p.handle for p in problems])
    bd = (BatchDesc * ns)(*[b.desc() for b in batches])
    rd = (ResultDesc * ns)(*[r.desc() for r in results])
    _check(lib().hcx_track_multi(ns, ph, bd, rd, C.byref(opts), _stream_handle(stream, list(batches) + list(results))),
           "hcx_track_multi")
    return results


def autotune(problem, batch, opts, stream=None):
    """Times launch shapes on `batch` and stores the fastest in the problem."""
    r = _capi.TuneResult()
    b = batch.desc()
    _check(lib().hcx_autotune(problem.handle, C.byref(b), C.byref(opts), C.byref(r), _stream_handle(stream)),
           "hcx_autotune")
    return {f: getattr(r, f) for f, _ in _capi.TuneResult._fields_}


class Graph:
    """One multi-type batch launch captured as a HIP graph; replays reuse the
    batch and result buffers given at capture."""

    def __init__(self, problems, batches, opts, results=None):
        if results is None:
            results = [Result(b.I * p.num_tracks, p.n, b.I) for p, b in zip(problems, batches)]
        self.problems, self.batches, self.results, self.opts = problems, batches, results, opts
        ns = len(problems)
        self._ph = (C.c_void_p * ns)(*[p.handle for p in problems])
        self._bd = (BatchDesc * ns)(*[b.desc() for b in batches])
        self._rd = (ResultDesc * ns)(*[r.desc() for r in results])
        h = C.c_void_p()
        _check(lib().hcx_graph_capture_multi(ns, self._ph, self._bd, self._rd, C.byref(opts), C.byref(h)),
               "hcx_graph_capture_multi")
        self._h = h

    def launch(self, stream=None):
        _check(lib().hcx_graph_launch(self._h, _stream_handle(stream, list(self.batches) + list(self.results))),
               "hcx_graph_launch")
        return self.results

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hcx_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Server:
    """Continuous-batching solve server over resident problem types (libhcx)."""

    def __init__(self, track_opts, hbm_budget_bytes=1 << 30, max_batch_requests=64, streams=4,
                 max_queue=1 << 20):
        so = _capi.ServerOpts()
        lib().hcx_server_opts_default(C.byref(so))
        so.hbm_budget_bytes, so.max_batch_requests, so.streams, so.max_queue = (
            hbm_budget_bytes, max_batch_requests, streams, max_queue)
        self._opts = track_opts
        h = C.c_void_p()
        _check(lib().hcx_server_create(C.byref(so), C.byref(track_opts), C.byref(h)), "hcx_server_create")
        self._h = h
        self.types = []

    def add_type(self, problem):
        tid = C.c_int()
        _check(lib().hcx_server_add_type(self._h, problem.handle, C.byref(tid)), "hcx_server_add_type")
        self.types.append(problem)
        return tid.value

    def submit(self, type_id, params):
        p = np.ascontiguousarray(params, dtype=np.complex128)
        rid = C.c_uint64()
        _check(lib().hcx_server_submit(self._h, type_id, _ptr(p), C.byref(rid)), "hcx_server_submit")
        return rid.value

    def done(self, request_id):
        d = C.c_int()
        _check(lib().hcx_server_poll(self._h, C.c_uint64(request_id), C.byref(d), None, None, None),
               "hcx_server_poll")
        return bool(d.value)

    def result(self, request_id, type_id):
        """Blocking-free fetch: returns None until done, else (info dict, roots, flags)."""
        if not self.done(request_id):
            return None
        prob = self.types[type_id]
        roots = np.empty((prob.num_tracks, prob.n), dtype=np.complex128)
        flags = np.empty(prob.num_tracks, dtype=np.uint32)
        done = C.c_int()
        rr = _capi.RequestResult()
        _check(lib().hcx_server_poll(self._h, C.c_uint64(request_id), C.byref(done), C.byref(rr), _ptr(roots),
                                     _ptr(flags)), "hcx_server_poll")
        if not done.value:
            return None
        return {f: getattr(rr, f) for f, _ in _capi.RequestResult._fields_}, roots, flags

    def drain(self, timeout_s=120.0):
        _check(lib().hcx_server_drain(self._h, timeout_s), "hcx_server_drain")

    def stats(self):
        st = _capi.ServerStats()
        _check(lib().hcx_server_get_stats(self._h, C.byref(st)), "hcx_server_get_stats")
        return {f: getattr(st, f) for f, _ in _capi.ServerStats._fields_}

    def reset_stats(self):
        lib().hcx_server_reset_stats(self._h)

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.hcx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def collect(result, opts, stream=None):
    c = Counts()
    _check(lib().hcx_collect(result.n, result.I, result.P // max(result.I, 1), C.c_void_p(result.roots.data_ptr()),
                             C.c_void_p(result.flags.data_ptr()), opts.unique_tol, C.byref(c),
                             _stream_handle(stream, (result,))), "hcx_collect")
    return {f: getattr(c, f) for f, _ in Counts._fields_}


def stream_release(stream):
    """Frees the launch workspace libhcx keeps for `stream` (after waiting for it)."""
    _check(lib().hcx_stream_release(C.c_void_p(stream.cuda_stream)), "hcx_stream_release")


def collect_host(n, num_instances, num_tracks, roots, flags, unique_tol=1e-4):
    roots = np.ascontiguousarray(roots, dtype=np.complex128)
    flags = np.ascontiguousarray(flags, dtype=np.uint32)
    c = Counts()
    _check(lib().hcx_collect_host(n, num_instances, num_tracks, _ptr(roots), _ptr(flags), unique_tol, C.byref(c)),
           "hcx_collect_host")
    return {f: getattr(c, f) for f, _ in Counts._fields_}, flags


def flops_per_path(info, steps, newton, refine):
    """Algorithmic FLOPs of the reference algorithm for per-path counters."""
    steps = np.asarray(steps, dtype=np.float64)
    newton = np.asarray(newton, dtype=np.float64)
    refine = np.asarray(refine, dtype=np.float64)
    return (4.0 * steps * info["flops_stage"] + newton * info["flops_newton"]
            + refine * info["flops_newton"] + (refine > 0) * info["flops_coef"])


# ---------------------------------------------------------------- reference problem folders
def ref_read_settings(problem_dir):
    """gpuhc_settings.yaml of a reference problem folder as a dict."""
    st = RefSettings()
    _check(lib().hcx_ref_read_settings(problem_dir.encode(), C.byref(st)), "hcx_ref_read_settings")
    out = {f: getattr(st, f) for f, _ in RefSettings._fields_}
    for key in ("problem_name", "gpuhc_type"):
        out[key] = out[key].decode()
    return out


def ref_read_params(path, cap):
    out = np.empty(cap, dtype=np.complex128)
    cnt = C.c_int()
    _check(lib().hcx_ref_read_params(path.encode(), _ptr(out), cap, C.byref(cnt)), "hcx_ref_read_params")
    return out[:cnt.value].copy()


def ref_read_start_sols(path, n, cap_tracks):
    out = np.empty((cap_tracks, n), dtype=np.complex128)
    cnt = C.c_int()
    _check(lib().hcx_ref_read_start_sols(path.encode(), n, _ptr(out), cap_tracks, C.byref(cnt)),
           "hcx_ref_read_start_sols")
    return out[:cnt.value].copy()


def ref_read_indices(path, cap):
    out = np.empty(cap, dtype=np.int32)
    cnt = C.c_int()
    _check(lib().hcx_ref_read_indices(path.encode(), _ptr(out), cap, C.byref(cnt)), "hcx_ref_read_indices")
    return out[:cnt.value].copy()


def ref_read_cmap(path):
    """(scalar [C][T], params [C][T][D], k) of a coefficient map file."""
    Cn, T, D, k = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    _check(lib().hcx_ref_read_cmap(path.encode(), C.byref(Cn), C.byref(T), C.byref(D), C.byref(k), None, None, 0),
           "hcx_ref_read_cmap")
    scal = np.empty((Cn.value, T.value), dtype=np.complex128)
    prm = np.empty((Cn.value, T.value, D.value), dtype=np.int32)
    _check(lib().hcx_ref_read_cmap(path.encode(), C.byref(Cn), C.byref(T), C.byref(D), C.byref(k), _ptr(scal),
                                   _ptr(prm), scal.size), "hcx_ref_read_cmap")
    return scal, prm, k.value


def ref_ph_to_p2c(n, P, ph_ht, ph_hx):
    """C-ABI PH -> P2C conversion: (C, ht [n][htT][htP-1], hx [n][n][hxT][hxP-1], pairs [C][2])."""
    ph_ht = np.ascontiguousarray(ph_ht, dtype=np.int32)
    ph_hx = np.ascontiguousarray(ph_hx, dtype=np.int32)
    htT, htP = ph_ht.shape[0], ph_ht.shape[1]
    hxT, hxP = ph_hx.shape[1], ph_hx.shape[2]
    ht = np.empty((n, htT, htP - 1), dtype=np.int32)
    hx = np.empty((n, n, hxT, hxP - 1), dtype=np.int32)
    cap = (P + 1) * (P + 2) // 2
    pairs = np.empty((cap, 2), dtype=np.int32)
    Cn = C.c_int()
    _check(lib().hcx_ref_ph_to_p2c(n, P, htT, htP, _ptr(ph_ht), hxT, hxP, _ptr(ph_hx), C.byref(Cn), _ptr(ht),
                                   _ptr(hx), _ptr(pairs), cap), "hcx_ref_ph_to_p2c")
    return Cn.value, ht, hx, pairs[:Cn.value].copy()
