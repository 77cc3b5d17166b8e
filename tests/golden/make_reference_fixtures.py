"""Writes tests/golden/reference/<problem>.npz from the reference's own problem
data.  Run in a checkout that has the reference tree (default /root/reference):

    python tests/golden/make_reference_fixtures.py [reference_root]

Inputs per problem (all text data of the reference):
  GPU-HC/problems/<p>/{gpuhc_settings.yaml, dHdt_indx.txt, dHdx_indx.txt,
                       start_params.txt, target_params.txt, start_sols.txt}
  auto-data-gen-tools/p2c-problems/sys_<p>.m      (polynomial system)
  GPU-HC/Output_Write_Files/GPU_Converged_HC_tracks.txt (converged roots of
                       the reference's own run of uncalibrated_trifocal_rel_pos_CY)

Each fixture holds data only: the index tables and parameter / start-solution
values as read, the coefficient map recovered by
hcx.compiler.infer_coefficient_map (every table entry checked against the
system), the exact coefficient-polynomial tables Hx / Ht of the start ->
target parameter path (rational arithmetic), and values F(x; p), dF/dx of the
system at seeded random points, evaluated from the system definition directly
(sympy), for the tests to compare the table route against.
"""
import os
import sys
import time
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "homotopy-continuation-tracker-on-gpu_amd"))
from hcx import compiler as cc  # noqa: E402

OUT = os.path.join(HERE, "reference")
SYS_NAME = {"5pt_rel_pos_full_form_quat": "5pt_rel_pos_full_form"}
# start solutions kept per problem (all of them up to this many)
MAX_SOLS = {"uncalibrated_trifocal_rel_pos_CY": 668, "generalized_3views_4pts": 583}
DEFAULT_MAX_SOLS = 128
CHECK_POINTS = 2


def system_values(sd, k, points, seed):
    """F and dF/dx at seeded CN(0,1) points (x, p), p padded to the file's k."""
    import sympy as sp
    rng = np.random.default_rng(seed)
    n = sd.n
    ps = sd.ps[:k]
    f = sp.lambdify(sd.xs + ps, sd.fs, "numpy")
    jac = sp.lambdify(sd.xs + ps, [[sp.diff(fi, x) for x in sd.xs] for fi in sd.fs], "numpy")
    X, P, F, J = [], [], [], []
    for _ in range(points):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
        p = (rng.standard_normal(k) + 1j * rng.standard_normal(k)) / np.sqrt(2)
        X.append(x)
        P.append(p)
        F.append(np.array(f(*x, *p), dtype=np.complex128).reshape(n))
        J.append(np.array(jac(*x, *p), dtype=np.complex128).reshape(n, n))
    return np.array(X), np.array(P), np.array(F), np.array(J)


def read_converged_tracks(path, n):
    roots, cur = [], []
    with open(path) as fh:
        for line in fh:
            s = line.split()
            if len(s) == 2:
                try:
                    cur.append(complex(float(s[0]), float(s[1])))
                except ValueError:
                    continue
                if len(cur) == n:
                    roots.append(cur)
                    cur = []
    return np.array(roots, dtype=np.complex128)


def main(root):
    os.makedirs(OUT, exist_ok=True)
    probs = os.path.join(root, "GPU-HC", "problems")
    for name in sorted(os.listdir(probs)):
        t0 = time.time()
        sysfile = os.path.join(root, "auto-data-gen-tools", "p2c-problems", f"sys_{SYS_NAME.get(name, name)}.m")
        with open(sysfile) as fh:
            sd = cc.parse_system(fh.read(), name)
        d = cc.load_problem_dir(os.path.join(probs, name))
        st = d["settings"]
        n = st["Num_Of_Vars"]
        k = len(d["start_params"])
        if k != st["Num_Of_Params"]:
            raise SystemExit(f"{name}: start_params.txt disagrees with Num_Of_Params")
        # uncalibrated_trifocal_rel_pos_CH ships 18 of its 24 target parameters;
        # such a problem gets seeded CN(0,1) targets in the tests instead
        target_ok = len(d["target_params"]) == k
        K = d["start_sols"].shape[0]
        keep = min(K, MAX_SOLS.get(name, DEFAULT_MAX_SOLS))
        out = dict(name=np.array(name), kind=np.array(st["GPUHC_Type"]), n=n, k=k, num_tracks=K,
                   max_steps=st["GPUHC_Max_Steps"], max_corr=st["GPUHC_Max_Correction_Steps"],
                   incr_steps=st["GPUHC_Num_Of_Steps_to_Increase_Delta_t"], M=st["Max_Order_Of_T"],
                   hx_terms=st["dHdx_Max_Terms"], hx_parts=st["dHdx_Max_Parts"], ht_terms=st["dHdt_Max_Terms"],
                   ht_parts=st["dHdt_Max_Parts"], ht=d["ht"], hx=d["hx"], start_params=d["start_params"],
                   target_params=d["target_params"], target_complete=target_ok,
                   start_sols=d["start_sols"][:keep])
        X, P, F, J = system_values(sd, k, CHECK_POINTS, seed=zlib.crc32(name.encode()))
        out.update(check_x=X, check_p=P, check_f=F, check_j=J)
        if st["GPUHC_Type"] == "P2C":
            C = st["Num_Of_Coeffs_From_Params"]
            ht = d["ht"].reshape(n, st["dHdt_Max_Terms"], st["dHdt_Max_Parts"])
            hx = d["hx"].reshape(n, n, st["dHdx_Max_Terms"], st["dHdx_Max_Parts"])
            coef = cc.infer_coefficient_map(sd, ht, hx, C)
            used = set().union(*[c.free_symbols for c in coef]) & set(sd.ps)
            if any(sd.ps.index(s) >= k for s in used):
                raise SystemExit(f"{name}: a coefficient uses a parameter beyond the {k} in the files")
            scal, prm = cc.coefficient_map_arrays(coef, sd.ps[:k])
            out.update(C=C, cmap_scalar=scal, cmap_params=prm)
            if target_ok:
                hxt, htt = cc.coefficient_tables(coef, sd.ps[:k], d["start_params"], d["target_params"],
                                                 st["Max_Order_Of_T"])
                out.update(coeffs_hx=hxt, coeffs_ht=htt)
        if name == "uncalibrated_trifocal_rel_pos_CY":
            out["ref_converged"] = read_converged_tracks(
                os.path.join(root, "GPU-HC", "Output_Write_Files", "GPU_Converged_HC_tracks.txt"), n)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **out)
        print(f"{name}: n={n} k={k} tracks={K} kept={keep} {time.time() - t0:.1f}s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
