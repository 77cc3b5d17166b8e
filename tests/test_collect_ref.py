"""The reference's own solution counts (Evaluations.cpp:88-176), pinned by
hand-traced cases.

Each case's expected unique_ref is what Find_Unique_Sols (Evaluations.cpp:
127-176) computes, traced by hand through its std::set logic: a converged
path bs that is not in Skip gets Duplicate = every LATER path ds (converged
or not) whose every coordinate agrees to < 1e-4 (the Duplicate set is flushed
only while Skip is non-empty); an empty set counts bs as unique, a non-empty
one REPLACES Skip.  Each case is chosen so that one plausible misreading
fails it: accumulating Skip instead of replacing it, counting one member of a
duplicate pair, ignoring non-converged later paths, comparing with earlier
paths, or testing one coordinate only.  The real_ref count is the
converged paths whose tracked root has every |Im x_i| <= 1e-4
(Evaluate_GPUHC_Sols, :88-110)."""
import numpy as np
import pytest

T, INF = 0x01, 0x02  # tracked ("converged"), infinity
REF = 0x04


def roots_from(labels, n=2):
    """One root per path: paths with the same label share a root exactly."""
    rng = np.random.default_rng(7)
    table = {}
    out = np.zeros((len(labels), n), dtype=np.complex128)
    for j, lab in enumerate(labels):
        if lab not in table:
            table[lab] = rng.normal(size=n) * 3 + 1j * rng.normal(size=n) * 3
        out[j] = table[lab]
    return out


# (name, root labels, flags per path, expected unique_ref)
CASES = [
    # bs=0: D={2}, Skip={2}; bs=1: flush, D={3}, Skip={3} (replaced); bs=2 not in
    # {3}: D={} -> unique; bs=3 in Skip.  Accumulating Skip would give 0.
    ("skip set is replaced", ["A", "B", "A", "B"], [T] * 4, 1),
    # a duplicate pair adds 0 (counting one member would give 1)
    ("pair adds nothing", ["A", "A"], [T] * 2, 0),
    # a later non-converged path at the same point still counts as a duplicate
    ("later failed path counts", ["A", "A"], [T, INF], 0),
    # an EARLIER non-converged duplicate does not (only ds > bs are compared)
    ("earlier failed path ignored", ["A", "A"], [INF, T], 1),
    # triple: bs=0 D={1,2}; 1 and 2 are skipped
    ("triple", ["A", "A", "A"], [T] * 3, 0),
    # bs=0 D={2} Skip={2}; bs=1 unique; bs=2 skipped; bs=3 unique
    ("unique around a skipped path", ["A", "B", "A", "C"], [T] * 4, 2),
    # A B B A: bs=0 D={3} Skip={3}; bs=1 D={2} Skip={2}; bs=2 in Skip;
    # bs=3 not in {2}: D={} -> unique.  Accumulating Skip would give 0.
    ("replaced skip lets the last A count", ["A", "B", "B", "A"], [T] * 4, 1),
    # all distinct
    ("distinct", ["A", "B", "C", "D", "E"], [T] * 5, 5),
    # non-converged paths are never counted
    ("failures", ["A", "B", "C"], [T, INF, 0x20], 1),
]


@pytest.mark.parametrize("name,labels,flags,expect", CASES, ids=[c[0] for c in CASES])
def test_find_unique_sols_traced(hcx_mod, oracle_mod, name, labels, flags, expect):
    roots = roots_from(labels)
    f = np.array([x | (REF if x & T else 0) for x in flags], dtype=np.uint32)
    oc, of = oracle_mod.collect(2, 1, len(labels), roots, f)
    hc, hf = hcx_mod.collect_host(2, 1, len(labels), roots, f)
    assert oc["unique_ref"] == expect, name
    assert hc == oc and (hf == of).all(), name
    assert ((of & oracle_mod.FLAG_UNIQUE_REF) != 0).sum() == expect


def test_tolerance_is_strict_and_per_coordinate(hcx_mod, oracle_mod):
    """|dRe| and |dIm| < 1e-4 in EVERY coordinate (DUPLICATE_SOL_DIFF_TOL,
    definitions.hpp:24; Evaluations.cpp:151-156)."""
    base = np.array([1 + 1j, -2 + 0.5j])
    near = base + np.array([0.5e-4, 0.5e-4j])        # within tol in both coordinates
    far_one = base + np.array([0, 3e-4])              # one coordinate's real part off
    far_im = base + np.array([3e-4j, 0])              # one imaginary part off
    for other, dup in ((near, True), (far_one, False), (far_im, False)):
        roots = np.stack([base, other])
        f = np.array([T, T], dtype=np.uint32)
        oc, _ = oracle_mod.collect(2, 1, 2, roots, f)
        hc, _ = hcx_mod.collect_host(2, 1, 2, roots, f)
        assert oc == hc
        assert oc["unique_ref"] == (0 if dup else 2)


def test_instances_are_independent(hcx_mod, oracle_mod):
    """The same root in two instances is not a duplicate (per-instance sets)."""
    roots = roots_from(["A", "B", "A", "B"])
    f = np.full(4, T, dtype=np.uint32)
    oc, _ = oracle_mod.collect(2, 2, 2, roots, f)
    hc, _ = hcx_mod.collect_host(2, 2, 2, roots, f)
    assert oc == hc and oc["unique_ref"] == 4


def test_real_ref_counts_flags_of_tracked_paths(hcx_mod, oracle_mod):
    roots = roots_from(["A", "B", "C"])
    f = np.array([T | 0x200, T, INF | 0x200], dtype=np.uint32)
    oc, _ = oracle_mod.collect(2, 1, 3, roots, f)
    hc, _ = hcx_mod.collect_host(2, 1, 3, roots, f)
    assert oc == hc and oc["real_ref"] == 1


@pytest.mark.parametrize("im,real", [(0.0, True), (0.5e-4, True), (0.99e-4, True), (1.01e-4, False), (2e-4, False)])
def test_real_ref_on_the_tracked_root(hcx_mod, oracle_mod, im, real):
    """The oracle sets the reference's real flag from the tracked root with
    |Im x_i| <= 1e-4 (Evaluations.cpp:100-104 uses <=), independent of the
    refined-root test at 1e-8.  A linear homotopy lands on b exactly."""
    from test_oracle import linear_problem
    a = np.array([1 + 2j, -0.5j])
    b = np.array([-2.0 + 0j, 0.25 + im * 1j])
    _, op, hx, ht = linear_problem(hcx_mod, oracle_mod, a, b)
    out = op.track(a[None], hx[None], ht[None], threads=1, dt_min=0.0)
    fl = int(out["flags"][0])
    assert fl & oracle_mod.FLAG_TRACKED
    assert bool(fl & oracle_mod.FLAG_REAL_REF) == real
    assert bool(fl & oracle_mod.FLAG_REAL) == (im == 0.0)
