"""Edge cases of the GPU tracker through the C-ABI: an empty batch, the
largest supported system (n = 32, generic instance, 32 lanes per path)
against the FP64 oracle, and a ragged batch whose path count fills no whole
wave of lane groups.  Bar as in test_gpu_parity.py: identical refined and
real counts, identical flags in FP64, refined roots within 1e-10."""
import numpy as np
import pytest

from test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def _quadratic_type(hcx, n, K, seed):
    from hcx import workloads
    ty = workloads.quadratic_types([(n, K)], seed=seed)[0]
    return ty, hcx.Problem.from_system(ty["system"], ty["start"], p0=ty["p0"])


def _oracle_track(oracle, ty, p1, threads=8):
    s = ty["system"]
    hxi, hti = s.tables()
    op = oracle.OracleProblem(ty["n"], s.info["num_coeffs"], hxi, hti, 1)
    ohx, oht = oracle.p2c_affine(s.affine(), ty["p0"], p1)
    return op.track(ty["start"], ohx, oht, threads=threads)


def test_empty_batch(hcx_mod, gpu):
    """Zero instances: the launch is a no-op and the counts are zero."""
    import torch
    hcx = hcx_mod
    from hcx import workloads
    ty, prob = _quadratic_type(hcx, 6, 64, seed=5)
    b = hcx.Batch(prob, 0, target_params=np.zeros((0, ty["k"]), dtype=np.complex128))
    for prec in (0, 1):
        opts = hcx.default_opts(precision=prec)
        r = hcx.track(prob, b, opts)
        torch.cuda.synchronize()
        c = hcx.collect(r, opts)
        print(f"[empty prec={prec}] {c}", flush=True)
        assert c == {"tracked": 0, "infinity": 0, "refined": 0, "real": 0, "unique": 0, "unique_ref": 0, "real_ref": 0}
    # the problem is still usable afterwards
    b1 = hcx.Batch(prob, 1, target_params=workloads.targets(ty["k"], 1, seed=3))
    g = hcx.track(prob, b1, hcx.default_opts()).to_host()
    torch.cuda.synchronize()
    assert ((g["flags"] & hcx.FLAG_REFINED) != 0).sum() > 0
    prob.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_max_size_n32_vs_oracle(hcx_mod, oracle_mod, gpu, prec):
    """n = 32 (the largest size): generic NMAX = 32 instance on 32 lanes."""
    import torch
    hcx, oracle = hcx_mod, oracle_mod
    from hcx import workloads
    ty, prob = _quadratic_type(hcx, 32, 16, seed=9)
    p1 = workloads.targets(ty["k"], 2, seed=41)
    r = hcx.track(prob, hcx.Batch(prob, 2, target_params=p1), hcx.default_opts(precision=prec))
    torch.cuda.synchronize()
    g = r.to_host()
    o = _oracle_track(oracle, ty, p1)
    compare(hcx, oracle, g, o, f"n=32 prec={prec}", exact_flags=(prec == 0))
    assert ((g["flags"] & hcx.FLAG_REFINED) != 0).sum() >= 0.9 * len(g["flags"])
    prob.close()


def test_ragged_batch_vs_oracle(hcx_mod, oracle_mod, gpu):
    """n = 9 (seven 9-lane groups per wave, one lane left over), 3 instances of
    142 paths: 426 paths, a partial last wave."""
    import torch
    hcx, oracle = hcx_mod, oracle_mod
    from hcx import workloads
    ty, prob = _quadratic_type(hcx, 9, 142, seed=5)
    p1 = workloads.targets(ty["k"], 3, seed=43)
    r = hcx.track(prob, hcx.Batch(prob, 3, target_params=p1), hcx.default_opts(precision=0))
    torch.cuda.synchronize()
    g = r.to_host()
    o = _oracle_track(oracle, ty, p1)
    compare(hcx, oracle, g, o, "ragged n=9", exact_flags=True)
    prob.close()
