"""Code tables: generators, reference-table loader and the level planner.

The planner's level counts are pinned against independent measurements of
the reference's own tables (26 levels for 4000x2000, 12 for 1944x972, a
32 399-level chain for 64800x32400; 9 and 17 greedy colours) when the
reference tree is present, and against a Python restatement elsewhere."""
from collections import Counter

import numpy as np
import pytest

import fast_ldpc_decoder_for_gpu_fixed_amd as L
from tests.conftest import REFERENCE


def var_checks(code):
    deg, pos = code.export()
    off = np.concatenate([[0], np.cumsum(deg)])
    vc = [[] for _ in range(code.n_vars)]
    for c in range(code.n_checks):
        for v in pos[off[c]:off[c + 1]]:
            vc[v].append(c)
    return deg, pos, vc


def has_4cycle(vc):
    pairs = Counter()
    for l in vc:
        l = sorted(l)
        for i in range(len(l)):
            for j in range(i + 1, len(l)):
                pairs[(l[i], l[j])] += 1
    return max(pairs.values()) > 1


def levels_py(deg, pos, n):
    last = [-1] * n
    e, top = 0, 0
    for d in deg:
        vs = pos[e:e + d]
        e += d
        lv = max([last[v] + 1 for v in vs] + [0])
        for v in vs:
            last[v] = lv
        top = max(top, lv + 1)
    return top


def test_regular_3_6_n4000():
    c = L.Code.regular(4000, 3, 6, 1)
    assert (c.n_vars, c.n_checks, c.n_edges, c.max_deg) == (4000, 2000, 12000, 6)
    deg, pos, vc = var_checks(c)
    assert set(deg) == {6}
    assert Counter(len(x) for x in vc) == {3: 4000}
    assert not has_4cycle(vc)
    off = np.concatenate([[0], np.cumsum(deg)])
    for i in range(c.n_checks):  # no repeated variable inside a check
        assert len(set(pos[off[i]:off[i + 1]])) == 6
    assert c.n_levels == levels_py(list(deg), list(pos), 4000)
    # deterministic
    d2, p2 = L.Code.regular(4000, 3, 6, 1).export()
    np.testing.assert_array_equal(pos, p2)
    assert not np.array_equal(pos, L.Code.regular(4000, 3, 6, 2).export()[1])


def test_colour_reorder_keeps_H_and_cuts_levels():
    c = L.Code.regular(4000, 3, 6, 1)
    r = c.reorder_by_colour()
    assert r.n_levels <= 12 < c.n_levels
    deg, pos = c.export()
    deg2, pos2 = r.export()
    off, off2 = np.concatenate([[0], np.cumsum(deg)]), np.concatenate([[0], np.cumsum(deg2)])
    a = sorted(tuple(pos[off[i]:off[i + 1]]) for i in range(c.n_checks))
    b = sorted(tuple(pos2[off2[i]:off2[i + 1]]) for i in range(r.n_checks))
    assert a == b
    assert r.n_levels == levels_py(list(deg2), list(pos2), 4000)


def test_qc_structure():
    c = L.Code.qc(12, 24, 81, 0.35, 1)
    assert (c.n_vars, c.n_checks) == (1944, 972)
    deg, pos, vc = var_checks(c)
    # info columns have degree >= 2; parity columns have degree 2 or 3
    assert all(len(x) >= 2 for x in vc)
    assert Counter(len(vc[v]) for v in range(12 * 81 + 81, 24 * 81)) == {2: 11 * 81}
    assert {len(vc[v]) for v in range(12 * 81, 13 * 81)} == {3}
    # checks of one base row never share a variable: levels <= 12
    assert c.n_levels <= 12
    c2 = L.Code.qc(4, 24, 96, 0.35, 2)
    assert (c2.n_vars, c2.n_checks) == (2304, 384) and c2.n_info == 1920
    assert all(len(x) >= 2 for x in var_checks(c2)[2])


def test_ira_structure():
    c = L.Code.ira(32400, 12960, 8, 3, 1)
    assert (c.n_vars, c.n_checks, c.n_edges) == (64800, 32400, 226799)
    deg, pos, vc = var_checks(c)
    assert Counter(deg) == {7: 32399, 6: 1}
    vd = Counter(len(vc[v]) for v in range(32400))
    assert vd == {8: 12960, 3: 19440}
    assert Counter(len(vc[v]) for v in range(32400, 64800)) == {2: 32399, 1: 1}
    assert c.n_levels == 32400            # staircase chain in generation order
    assert c.reorder_by_colour().n_levels == 12


def test_balanced_colouring():
    """reorder_by_colour: deterministic, never below the clique bound (the
    checks of one variable need distinct levels), and for the generated codes
    of the benchmarks few, evenly filled levels (first-fit greedy colouring
    gave 9 for the (3,6) code and 17 for the IRA code, DSatur alone 7 and 13)."""
    raw = L.Code.regular(4000, 3, 6, 1)
    a, b = raw.reorder_by_colour(), raw.reorder_by_colour()
    for x, y in zip(a.export(), b.export()):
        np.testing.assert_array_equal(x, y)
    assert a.n_levels == 6
    deg, pos = a.export()
    # same H: every check's variable set appears once
    rows = sorted(tuple(sorted(pos[o:o + d])) for o, d in zip(np.concatenate([[0], np.cumsum(deg)[:-1]]), deg))
    deg0, pos0 = raw.export()
    rows0 = sorted(tuple(sorted(pos0[o:o + d])) for o, d in zip(np.concatenate([[0], np.cumsum(deg0)[:-1]]), deg0))
    assert rows == rows0
    # levels of the new order: 320..350 checks each
    lv = np.full(4000, -1)
    lev = []
    e = 0
    for d in deg:
        vs = pos[e:e + d]
        e += d
        lev.append(int(lv[vs].max()) + 1)
        lv[vs] = lev[-1]
    sizes = np.bincount(lev)
    assert sizes.min() >= 320 and sizes.max() <= 350
    for c in (a, L.Code.ira(3240, 1296, 8, 3, 11).reorder_by_colour()):
        d, p_ = c.export()
        assert c.n_levels >= np.bincount(p_).max()


def _table_levels(deg, pos, n):
    """Dependency levels of the serial table order (numpy restatement)."""
    last = np.full(n, -1)
    out, e = [], 0
    for d in deg:
        vs = pos[e:e + d]
        e += d
        out.append(int(last[vs].max()) + 1)
        last[vs] = out[-1]
    return np.array(out)


def _keeps_table_order(deg, pos, lv):
    """Every variable's checks, in table order, sit on strictly increasing levels."""
    off = np.concatenate([[0], np.cumsum(deg)])
    seen = {}
    for c in range(len(deg)):
        for v in pos[off[c]:off[c + 1]]:
            if v in seen and not lv[seen[v]] < lv[c]:
                return False
            seen[v] = c
    return True


def _lane_waste(sizes, words):
    return int(sum(-(-n * words // 64) * 64 - n * words for n in sizes))


def test_wave_balanced_levels():
    """Levels with sizes in whole waves: the same number of levels, checks that
    share a variable still in table order (so decisions do not change), and for
    the (3,6) N=4000 benchmark code with 2-word blocks the minimum of partly
    filled wave lanes (4000 items per iteration = 62.5 waves: 32 lanes)."""
    code = L.code_3_6(4000)
    deg, pos = code.export()
    asap = _table_levels(deg, pos, 4000)
    np.testing.assert_array_equal(code.levels(1), asap)
    assert list(np.bincount(asap)) == [344, 339, 335, 332, 328, 322]
    lv = code.levels(32)
    assert lv.max() + 1 == asap.max() + 1
    assert _keeps_table_order(deg, pos, lv)
    sizes = np.bincount(lv)
    assert list(sizes) == [320, 352, 320, 346, 320, 342]
    assert _lane_waste(np.bincount(asap), 2) == 224 and _lane_waste(sizes, 2) == 32
    assert not _keeps_table_order(deg, pos, np.zeros_like(lv))  # the check detects a violation
    for c, g in ((L.Code.regular(240, 3, 6, 5), 8), (L.Code.qc(12, 24, 81, 0.35, seed=1), 32),
                 (L.Code.qc(4, 24, 96, 0.35, seed=2), 16), (L.Code.ira(3240, 1296, 8, 3, 11).reorder_by_colour(), 64)):
        d, p_ = c.export()
        a = _table_levels(d, p_, c.n_vars)
        b = c.levels(g)
        assert b.max() == a.max() and _keeps_table_order(d, p_, b)
        assert _lane_waste(np.bincount(b), 64 // g) <= _lane_waste(np.bincount(a), 64 // g)


@pytest.mark.skipif(not (REFERENCE / "src/matrix").exists(), reason="reference tree not present")
@pytest.mark.parametrize("table,n,expect", [
    ("4000x2000", 4000, dict(C=2000, E=12000, levels=26, colours=6, maxdeg=6)),
    ("1944x972", 1944, dict(C=972, E=6966, levels=12, maxdeg=8, q1=162)),
    ("64800x32400", 64800, dict(C=32400, E=226799, levels=32399, colours=12, maxdeg=7, q1=32399)),
])
def test_reference_tables(table, n, expect):
    c = L.Code.load_table(REFERENCE / "src/matrix" / table / "constantes_decoder.h", n)
    assert c.n_checks == expect["C"] and c.n_edges == expect["E"] and c.max_deg == expect["maxdeg"]
    assert c.n_levels == expect["levels"]
    if "colours" in expect:
        assert c.reorder_by_colour().n_levels == expect["colours"]
    if "q1" in expect:
        assert c.q1_split == expect["q1"]
