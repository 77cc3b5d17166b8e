"""hcx_track_noalloc (launch workspace in a static device array, nothing
allocated) against hcx_track through the C-ABI: two calls in a row, the
second with a smaller batch, in mixed precision (re-track list in the
scratch) and FP64, with instance completion times requested (instance
counters in the scratch).  Results must be identical bit for bit.  A batch
beyond the scratch is rejected on the host before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(hcx, prob, batch, opts, result, stream):
    b, r = batch.desc(), result.desc()
    return hcx.lib().hcx_track_noalloc(prob.handle, C.byref(b), C.byref(opts), C.byref(r), stream)


def test_noalloc_equals_track(hcx_mod, gpu):
    import torch
    from hcx import workloads
    hcx = hcx_mod
    ty = workloads.quadratic_types([(6, 64)], seed=5)[0]
    prob = hcx.Problem.from_system(ty["system"], ty["start"], p0=ty["p0"])
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for I, seed in ((48, 11), (7, 12)):  # the second call is smaller
        p1 = workloads.targets(ty["k"], I, seed)
        for prec in (hcx.PREC_MIXED, hcx.PREC_FP64):
            print(f"[noalloc] instances={I} precision={prec}", flush=True)
            opts = hcx.default_opts(precision=prec)
            ref = hcx.track(prob, hcx.Batch(prob, I, target_params=p1), opts).to_host()
            r = hcx.Result(I * prob.num_tracks, prob.n, I)
            b = hcx.Batch(prob, I, target_params=p1)
            st = _noalloc(hcx, prob, b, opts, r, stream)
            assert st == 0, hcx.lib().hcx_last_error()
            torch.cuda.synchronize()
            g = r.to_host()
            for key in ("flags", "roots", "steps", "newton", "refine", "t_final", "dt_final", "residual"):
                assert np.array_equal(g[key], ref[key]), (key, I, prec)
            # every instance completed after the launch started
            assert (g["inst_done"][:I] >= g["batch_start"][0]).all()
            assert ((g["flags"] & hcx.FLAG_REFINED) != 0).mean() > 0.9


def test_noalloc_rejects_batches_beyond_the_scratch(hcx_mod, gpu):
    """65,537 instances with completion times need one instance counter more
    than the scratch holds: refused before any launch.  The buffers are
    sized for the batch, so even a launch would stay in bounds."""
    import torch
    hcx = hcx_mod
    from test_oracle import linear_problem
    import oracle
    a = np.array([1 + 2j, -0.5j])
    b = np.array([-2.0 + 0j, 0.25])
    s, _, hx, ht = linear_problem(hcx, oracle, a, b)
    hxi, hti = s.tables()
    prob = hcx.Problem(2, 4, hxi, hti, a[None])
    I = 65537
    batch = hcx.Batch(prob, I, coeffs_hx=np.repeat(hx[None], I, axis=0), coeffs_ht=np.repeat(ht[None], I, axis=0))
    r = hcx.Result(I, 2, I)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = _noalloc(hcx, prob, batch, hcx.default_opts(), r, stream)
    assert st == 1
    assert b"static scratch" in hcx.lib().hcx_last_error()
    torch.cuda.synchronize()
    assert int(r.flags.abs().sum()) == 0  # nothing ran
