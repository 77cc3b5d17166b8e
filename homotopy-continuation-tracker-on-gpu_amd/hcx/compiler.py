"""Problem compiler: polynomial systems -> index tables + coefficient maps.

Replaces the reference's offline generator and its generated per-problem code:

* ``auto-data-gen-tools/autogen_gpuhc.m:503-699`` (dHdt / dHdx index tables,
  parts = {integer scalar, coefficient index, variable index x (parts-2)},
  variable index n = "times 1", padding term {0, C, n, ...});
* ``createUnivPoly4Coeffs.m`` (coefficients as polynomials in t along the
  straight-line parameter path p(t) = q + t (p - q));
* ``GPU-HC/magmaHC/PHC_Coeffs/p2c-<p>.h`` (coefficients from parameters,
  called from ``Data_Reader.cpp:59-89``), here a data table instead of code:
  c_j(p) = sum_u scalar[j][u] * prod_d p[params[j][u][d]], p[k] = 1.

Systems are read from the reference's definition files
(``auto-data-gen-tools/p2c-problems/sys_<p>.m``: ``syms`` lines and
``f(i) = <polynomial>;`` lines) with sympy.  Two ways to get tables:

* :func:`compile_system` numbers the coefficients itself;
* :func:`infer_coefficient_map` keeps the numbering of an existing
  ``problems/<p>/`` folder (its dHdt/dHdx tables) and recovers, term by term,
  which polynomial in the parameters each coefficient index stands for; every
  table entry is checked against the system, so a folder whose tables do not
  encode the system is rejected.

:func:`ph_to_p2c` turns the PH-type tables of ``trifocal_2op1p_30x30``
(coefficient = s * p_a(t) * p_b(t), tables transposed with the equation index
fastest; ``dev-eval-indxing-trifocal_2op1p_30x30.cuh:44-168``) into P2C tables
with one coefficient per parameter product.
"""
import re
from collections import defaultdict
from fractions import Fraction

import numpy as np

__all__ = ["SystemDef", "parse_system", "compile_system", "infer_coefficient_map", "coefficient_map_arrays",
           "coefficient_tables", "ph_to_p2c", "read_table", "read_settings", "read_complex", "load_problem_dir",
           "write_coefficient_map", "compile_problem_dir"]


class SystemDef:
    """Polynomial system f_i(x; p): sympy expressions, variables, parameters."""

    def __init__(self, fs, xs, ps, name=""):
        self.fs, self.xs, self.ps, self.name = list(fs), list(xs), list(ps), name

    @property
    def n(self):
        return len(self.xs)

    @property
    def k(self):
        return len(self.ps)


def _sympy():
    import sympy
    return sympy


def parse_system(text, name=""):
    """Parses a sys_<p>.m definition: ``syms`` declarations (variables are the
    x<i> symbols, parameters the others, in declaration order) and
    ``f(i) = expr;`` equations (MATLAB ``^`` powers)."""
    sp = _sympy()
    from sympy.parsing.sympy_parser import parse_expr
    names = []
    eqs = {}
    for raw in text.splitlines():
        line = raw.split("%", 1)[0].strip()
        if not line:
            continue
        if line.startswith("syms "):
            names += line[5:].replace(";", " ").split()
            continue
        m = re.match(r"f\((\d+)\)\s*=\s*(.+?);?\s*$", line)
        if m:
            eqs[int(m.group(1))] = m.group(2).rstrip(";")
    if not eqs:
        raise ValueError("no equations f(i) = ...; found")
    syms = {nm: sp.Symbol(nm) for nm in names}
    xs = [syms[nm] for nm in names if re.fullmatch(r"x\d+", nm)]
    ps = [syms[nm] for nm in names if not re.fullmatch(r"x\d+", nm)]
    xs.sort(key=lambda s: int(s.name[1:]))
    order = sorted(eqs)
    if order != list(range(1, len(order) + 1)):
        raise ValueError(f"equations are not numbered 1..{len(order)}")
    fs = [parse_expr(eqs[i].replace("^", "**"), local_dict=syms) for i in order]
    if len(fs) != len(xs):
        raise ValueError(f"{len(fs)} equations in {len(xs)} unknowns: not square")
    return SystemDef(fs, xs, ps, name)


def _monomials(f, xs):
    """{exponent tuple: coefficient expression in the parameters} of f."""
    sp = _sympy()
    P = sp.Poly(sp.expand(f), *xs)
    return {m: sp.expand(c) for m, c in P.terms()}


def _split_content(E, ps):
    """E = s * E' with s the signed integer content of E (E' = 1 for an
    integer constant).  Returns (s, E')."""
    sp = _sympy()
    if not E.free_symbols & set(ps):
        v = sp.nsimplify(E)
        if v.is_Integer and v != 0:
            return int(v), sp.Integer(1)
        return 1, E
    P = sp.Poly(E, *ps)
    cs = [c for _, c in P.terms()]
    if not all(c.is_Integer for c in cs):
        return 1, E
    g = 0
    for c in cs:
        g = int(sp.igcd(g, int(c)))
    s = g if cs[0] > 0 else -g
    return s, sp.expand(E / s)


def _vars_of(mono, n, width):
    v = []
    for j, e in enumerate(mono):
        v += [j] * int(e)
    if len(v) > width:
        raise ValueError("monomial degree exceeds the table width")
    return v + [n] * (width - len(v))


def compile_system(sysdef):
    """Index tables in the reference's layout and a coefficient map, with
    coefficients numbered here (distinct parameter polynomials after taking
    out their integer content; the constant 1 is index C).

    Returns dict(n, C, k, ht [n][htT][htP], hx [n][n][hxT][hxP], coeffs [C]
    sympy expressions, max_order_t)."""
    n, k = sysdef.n, sysdef.k
    ids, coeffs = {}, []
    rows = []
    maxdeg = 1
    for f in sysdef.fs:
        terms = []
        for mono, E in _monomials(f, sysdef.xs).items():
            s, Ep = _split_content(E, sysdef.ps)
            if Ep == 1:
                cid = -1
            else:
                key = Ep
                if key not in ids:
                    ids[key] = len(coeffs)
                    coeffs.append(Ep)
                cid = ids[key]
            terms.append((s, cid, mono))
            maxdeg = max(maxdeg, sum(mono))
        rows.append(terms)
    C = len(coeffs)
    htT = max(len(r) for r in rows)
    htP = max(3, 2 + maxdeg)
    ht = np.zeros((n, htT, htP), dtype=np.int32)
    dx = [[[] for _ in range(n)] for _ in range(n)]
    for i, terms in enumerate(rows):
        for u in range(htT):
            if u < len(terms):
                s, cid, mono = terms[u]
                ht[i, u] = [s, C if cid < 0 else cid] + _vars_of(mono, n, htP - 2)
                for j in range(n):
                    if mono[j] > 0:
                        m2 = list(mono)
                        m2[j] -= 1
                        dx[i][j].append((s * int(mono[j]), C if cid < 0 else cid, tuple(m2)))
            else:
                ht[i, u] = [0, C] + [n] * (htP - 2)
    hxT = max(1, max(len(dx[i][j]) for i in range(n) for j in range(n)))
    hxP = max(3, 2 + maxdeg - 1)
    hx = np.zeros((n, n, hxT, hxP), dtype=np.int32)
    for i in range(n):
        for j in range(n):
            for u in range(hxT):
                if u < len(dx[i][j]):
                    s, cid, mono = dx[i][j][u]
                    hx[i, j, u] = [s, cid] + _vars_of(mono, n, hxP - 2)
                else:
                    hx[i, j, u] = [0, C] + [n] * (hxP - 2)
    order = max([_param_degree(c, sysdef.ps) for c in coeffs] + [1])
    return dict(n=n, C=C, k=k, ht=ht, hx=hx, coeffs=coeffs, max_order_t=order)


def _param_degree(E, ps):
    sp = _sympy()
    if not E.free_symbols & set(ps):
        return 0
    return sp.Poly(E, *ps).total_degree()


def infer_coefficient_map(sysdef, ht, hx, C, check_points=2, seed=0):
    """Coefficient polynomials c_0..c_{C-1} of the parameters under the
    numbering of existing tables ht [n][htT][htP] and hx [n][n][hxT][hxP].

    Every dHdt term s * c[idx] * prod x must match the system's coefficient of
    that monomial in its equation (several terms on one monomial are checked
    as a sum), each index must mean one polynomial wherever it appears, index
    C must be the constant 1, and the dHdx table must give the derivatives of
    the system at random points.  Raises ValueError otherwise."""
    sp = _sympy()
    n = sysdef.n
    ht = np.asarray(ht).reshape(n, -1, np.asarray(ht).shape[-1])
    coef = [None] * C
    deferred = []
    for i, f in enumerate(sysdef.fs):
        mons = _monomials(f, sysdef.xs)
        groups = defaultdict(list)
        for rec in ht[i]:
            s, c = int(rec[0]), int(rec[1])
            if s == 0:
                continue
            mono = [0] * n
            for v in rec[2:]:
                if v < n:
                    mono[int(v)] += 1
            groups[tuple(mono)].append((s, c))
        if set(groups) != set(mons):
            extra = set(groups) ^ set(mons)
            raise ValueError(f"equation {i + 1}: table monomials differ from the system's ({len(extra)} differ)")
        for mono, lst in groups.items():
            E = mons[mono]
            if len(lst) == 1:
                s, c = lst[0]
                e = sp.expand(E / s)
                if c == C:
                    if sp.simplify(e - 1) != 0:
                        raise ValueError(f"equation {i + 1}: constant-index term is not {s} * 1")
                    continue
                if not 0 <= c < C:
                    raise ValueError(f"equation {i + 1}: coefficient index {c} out of range")
                if coef[c] is None:
                    coef[c] = e
                elif sp.expand(coef[c] - e) != 0:
                    raise ValueError(f"coefficient {c} stands for two different polynomials")
            else:
                deferred.append((i, mono, lst, E))
    for i, mono, lst, E in deferred:
        tot = 0
        for s, c in lst:
            if c == C:
                tot += s
            elif coef[c] is None:
                raise ValueError(f"coefficient {c} only appears in shared monomials")
            else:
                tot += s * coef[c]
        if sp.expand(tot - E) != 0:
            raise ValueError(f"equation {i + 1}: terms of one monomial do not sum to its coefficient")
    missing = [c for c in range(C) if coef[c] is None]
    if missing:
        raise ValueError(f"coefficients never used by dHdt: {missing[:8]}")
    _check_jacobian(sysdef, coef, hx, C, check_points, seed)
    return coef


def _lambdas(sysdef, coef):
    sp = _sympy()
    fx = [sp.lambdify(sysdef.xs + sysdef.ps, f, "numpy") for f in sysdef.fs]
    jx = [[sp.lambdify(sysdef.xs + sysdef.ps, sp.diff(f, x), "numpy") for x in sysdef.xs] for f in sysdef.fs]
    cf = [sp.lambdify(sysdef.ps, c, "numpy") for c in coef]
    return fx, jx, cf


def _table_value(recs, cvals, x, n):
    tot = 0j
    for rec in recs:
        s = int(rec[0])
        if s == 0:
            continue
        v = s * cvals[int(rec[1])]
        for q in rec[2:]:
            if q < n:
                v *= x[int(q)]
        tot += v
    return tot


def _check_jacobian(sysdef, coef, hx, C, points, seed):
    n = sysdef.n
    hx = np.asarray(hx).reshape(n, n, -1, np.asarray(hx).shape[-1])
    _, jx, cf = _lambdas(sysdef, coef)
    rng = np.random.default_rng(seed)
    for _ in range(points):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
        p = (rng.standard_normal(sysdef.k) + 1j * rng.standard_normal(sysdef.k)) / np.sqrt(2)
        cv = [complex(f(*p)) for f in cf] + [1.0]
        for i in range(n):
            for j in range(n):
                want = complex(jx[i][j](*x, *p))
                got = _table_value(hx[i, j], cv, x, n)
                if abs(got - want) > 1e-9 * max(1.0, abs(want)):
                    raise ValueError(f"dHdx table entry (eq {i + 1}, var {j + 1}) is not the derivative "
                                     f"({got} vs {want})")


def coefficient_map_arrays(coeffs, ps):
    """(scalar [C][T] complex, params [C][T][D] int32) of the coefficient
    polynomials; padding terms have scalar 0, padding factors index k."""
    sp = _sympy()
    k = len(ps)
    rows = []
    for E in coeffs:
        terms = []
        if not E.free_symbols & set(ps):
            terms.append((complex(E), []))
        else:
            for mono, c in sp.Poly(E, *ps).terms():
                idx = []
                for a, e in enumerate(mono):
                    idx += [a] * int(e)
                terms.append((complex(c), idx))
        rows.append(terms)
    T = max(1, max(len(r) for r in rows))
    D = max(1, max((len(idx) for r in rows for _, idx in r), default=1))
    scal = np.zeros((len(coeffs), T), dtype=np.complex128)
    prm = np.full((len(coeffs), T, D), k, dtype=np.int32)
    for j, r in enumerate(rows):
        for u, (c, idx) in enumerate(r):
            scal[j, u] = c
            prm[j, u, :len(idx)] = idx
    return scal, prm


def coefficient_tables(coeffs, ps, q, p, M):
    """Exact coefficient-polynomial-in-t tables of c_j(q + t (p - q)):
    Hx [C+1][M+1], Ht [C+1][M] (entry C = the constant 1), computed in
    rational arithmetic from the decimal parameter values."""
    sp = _sympy()
    t = sp.Symbol("t")

    def rat(z):
        return sp.Rational(Fraction(repr(float(z.real)))) + sp.I * sp.Rational(Fraction(repr(float(z.imag))))

    sub = {ps[a]: rat(q[a]) + t * (rat(p[a]) - rat(q[a])) for a in range(len(ps))}
    C = len(coeffs)
    hx = np.zeros((C + 1, M + 1), dtype=np.complex128)
    ht = np.zeros((C + 1, M), dtype=np.complex128)
    for j, E in enumerate(coeffs):
        poly = sp.Poly(sp.expand(E.subs(sub)), t)
        cs = poly.all_coeffs()[::-1]
        if len(cs) > M + 1:
            raise ValueError(f"coefficient {j} has degree {len(cs) - 1} > M = {M} in t")
        for m, c in enumerate(cs):
            hx[j, m] = complex(sp.N(c, 30))
        for m in range(1, len(cs)):
            ht[j, m - 1] = m * hx[j, m]
    hx[C, 0] = 1.0
    return hx, ht


def ph_to_p2c(n, P, ht_ph, hx_ph):
    """PH-type tables -> P2C tables + coefficient map.

    ht_ph: [htT][htP][n] (parts: scalar, p_a, p_b, x ...; equation fastest),
    hx_ph: [n var][hxT][hxP][n eq].  Parameter index P is the constant 1.
    Each distinct unordered pair {a, b} other than {P, P} becomes one P2C
    coefficient p_a * p_b (in order of first use); {P, P} and padding terms
    take the constant index C.
    Returns dict(C, ht [n][htT][htP-1], hx [n][n][hxT][hxP-1],
    cmap=(scalar [C][1], params [C][1][2]), max_order_t=2)."""
    ht_ph = np.asarray(ht_ph, dtype=np.int64)
    hx_ph = np.asarray(hx_ph, dtype=np.int64)
    htT, htP = ht_ph.shape[0], ht_ph.shape[1]
    hxT, hxP = hx_ph.shape[1], hx_ph.shape[2]
    pairs = {}

    def pid(a, b):
        a, b = (int(a), int(b)) if a <= b else (int(b), int(a))
        if a == P and b == P:
            return -1
        if (a, b) not in pairs:
            pairs[(a, b)] = len(pairs)
        return pairs[(a, b)]

    ht_l = np.zeros((n, htT, htP - 1), dtype=np.int64)
    hx_l = np.zeros((n, n, hxT, hxP - 1), dtype=np.int64)
    for i in range(n):
        for u in range(htT):
            rec = ht_ph[u, :, i]
            ht_l[i, u] = [rec[0], pid(rec[1], rec[2]) if rec[0] != 0 else -1] + list(rec[3:])
    for j in range(n):
        for u in range(hxT):
            for i in range(n):
                rec = hx_ph[j, u, :, i]
                hx_l[i, j, u] = [rec[0], pid(rec[1], rec[2]) if rec[0] != 0 else -1] + list(rec[3:])
    C = len(pairs)
    for tab in (ht_l, hx_l):
        c = tab[..., 1]
        c[(c < 0) | (tab[..., 0] == 0)] = C
    scal = np.ones((C, 1), dtype=np.complex128)
    prm = np.full((C, 1, 2), P, dtype=np.int32)
    for (a, b), idx in pairs.items():
        prm[idx, 0] = [a, b]
    return dict(C=C, ht=ht_l.astype(np.int32), hx=hx_l.astype(np.int32), cmap=(scal, prm), max_order_t=2)


# ------------------------------------------------------------- file readers
def read_table(path):
    """Whitespace-separated integers of an index table file."""
    with open(path) as fh:
        return np.array(fh.read().split(), dtype=np.int64).astype(np.int32)


def read_complex(path):
    """"re im" pairs of a parameter or start-solution file."""
    with open(path) as fh:
        v = np.array(fh.read().split(), dtype=np.float64)
    return v[0::2] + 1j * v[1::2]


def read_settings(path):
    """Flat ``key: value`` settings (gpuhc_settings.yaml)."""
    out = {}
    with open(path) as fh:
        for line in fh:
            line = line.split("#", 1)[0].strip()
            if not line or line.startswith("%") or ":" not in line:
                continue
            key, val = line.split(":", 1)
            val = val.strip()
            try:
                out[key.strip()] = int(val)
            except ValueError:
                out[key.strip()] = val
    return out


def load_problem_dir(path):
    """Settings, tables, parameters and start solutions of a problems/<p>/
    folder (Data_Reader.cpp:91-225 formats)."""
    import os
    st = read_settings(os.path.join(path, "gpuhc_settings.yaml"))
    n = st["Num_Of_Vars"]
    out = dict(settings=st, n=n)
    out["ht"] = read_table(os.path.join(path, "dHdt_indx.txt"))
    out["hx"] = read_table(os.path.join(path, "dHdx_indx.txt"))
    out["start_params"] = read_complex(os.path.join(path, "start_params.txt"))
    out["target_params"] = read_complex(os.path.join(path, "target_params.txt"))
    out["start_sols"] = read_complex(os.path.join(path, "start_sols.txt")).reshape(-1, n)
    return out


def write_coefficient_map(path, scalar, params, k):
    """Coefficient map file read by hcx_ref_read_cmap: "C T D k", then per
    coefficient j and term u one line "re im i_1 .. i_D"."""
    scalar = np.asarray(scalar, dtype=np.complex128)
    params = np.asarray(params)
    Cn, T = scalar.shape
    D = params.shape[2]
    with open(path, "w") as fh:
        fh.write(f"{Cn} {T} {D} {k}\n")
        for j in range(Cn):
            for u in range(T):
                z = scalar[j, u]
                idx = " ".join(str(int(a)) for a in params[j, u])
                fh.write(f"{float(z.real)!r} {float(z.imag)!r} {idx}\n")


def compile_problem_dir(problem_dir, sys_file, out=None):
    """Coefficient map of a P2C problem folder in its own numbering, from the
    system definition; writes it to `out` (default <problem_dir>/coeff_map.txt
    when writable) and returns (scalar, params, k)."""
    import os
    with open(sys_file) as fh:
        sd = parse_system(fh.read(), os.path.basename(problem_dir.rstrip("/")))
    d = load_problem_dir(problem_dir)
    st = d["settings"]
    n = st["Num_Of_Vars"]
    if st.get("GPUHC_Type") != "P2C":
        raise ValueError("only P2C problems carry a coefficient map (PH problems convert with ph_to_p2c)")
    C = st["Num_Of_Coeffs_From_Params"]
    ht = d["ht"].reshape(n, st["dHdt_Max_Terms"], st["dHdt_Max_Parts"])
    hx = d["hx"].reshape(n, n, st["dHdx_Max_Terms"], st["dHdx_Max_Parts"])
    coef = infer_coefficient_map(sd, ht, hx, C)
    k = st["Num_Of_Params"]
    scal, prm = coefficient_map_arrays(coef, sd.ps[:k])
    if out is None:
        out = os.path.join(problem_dir, "coeff_map.txt")
    write_coefficient_map(out, scal, prm, k)
    return scal, prm, k


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(
        prog="python -m hcx.compiler",
        description="Coefficient map of a reference P2C problem folder from its system definition")
    ap.add_argument("problem_dir", help="problems/<p>/ folder (gpuhc_settings.yaml, dHdt_indx.txt, ...)")
    ap.add_argument("sys_file", help="sys_<p>.m polynomial system definition")
    ap.add_argument("-o", "--out", help="output file (default <problem_dir>/coeff_map.txt)")
    a = ap.parse_args(argv)
    scal, prm, k = compile_problem_dir(a.problem_dir, a.sys_file, a.out)
    print(f"coefficient map: {scal.shape[0]} coefficients, {scal.shape[1]} terms x degree {prm.shape[2]}, k={k}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
