"""Formula lists (gecco-public_amd/formula.py): the schedule of GeCCo's
formula items into grouped launches, the committed formula files, and the
CCSD driver's levels derived from its flat formula."""
import os

import pytest

import ccsd
import formula as F

H, P = 1, 2
HERE = os.path.dirname(os.path.abspath(__file__))
FORMULAS = os.path.join(HERE, '..', 'gecco-public_amd', 'formulas')


def _names(levels):
    return [[(c['target'], c['op1'], c['op2']) for c in lv['contr']] for lv in levels]


def test_ccsd_levels_are_the_dependency_levels():
    """The flat CCSD formula schedules into three levels of 18, 12 and 7
    contractions in list order: every contraction of level 1 reads a list
    written in level 0, every one of level 2 a list written in level 1."""
    levels = ccsd.SCHEDULE
    assert [len(lv['contr']) for lv in levels] == [18, 12, 7]
    flat = [t[:3] for t in ccsd.TERMS]
    assert _names(levels) == [flat[:18], flat[18:30], flat[30:]]
    written = [{c['target'] for c in lv['contr']} for lv in levels]
    for li in (1, 2):
        for c in levels[li]['contr']:
            assert c['op1'] in written[li - 1] or c['op2'] in written[li - 1], c
    # every initialisation runs first, projections once per list
    assert {it['target'] for it in levels[0]['init']} == \
        {n for n in ccsd.LISTS if n not in ccsd.INTEGRALS and n not in ('T1', 'T2')}
    assert not levels[1]['init'] and not levels[2]['init']
    assert ccsd.SYM_AFTER == [['tau', 'taut', 'Fme', 'Z'],
                              ['Fae', 'F2ae', 'Fmi', 'F2mi', 'Wmnij', 'Zamij', 'Wmbej', 'E'],
                              ['R1', 'R2']]


def test_dependency_rules():
    c = F.contr
    items = [F.init('X'), F.init('Y'), F.init('Z'),
             c('X', 'A', 'B', {}, {}, 1.0),       # level 0
             c('Y', 'X', 'B', {}, {}, 1.0),       # reads X: level 1
             c('X', 'A', 'A', {}, {}, 1.0),       # writes X after a read of X: level 2
             c('Z', 'A', 'B', {}, {}, 1.0),       # independent: level 0
             c('Z', 'Y', 'A', {}, {}, 1.0),       # reads Y: level 2
             c('Y', 'A', 'A', {}, {}, 1.0),       # writes Y after its read: level 3
             F.init('W', 'Z', 2.0),               # after Z's writes: level 3
             c('W', 'A', 'A', {}, {}, 1.0)]       # after W's init: level 3
    lv = F.schedule(items)
    assert _names(lv) == [[('X', 'A', 'B'), ('Z', 'A', 'B')],
                          [('Y', 'X', 'B')],
                          [('X', 'A', 'A'), ('Z', 'Y', 'A')],
                          [('Y', 'A', 'A'), ('W', 'A', 'A')]]
    assert [it['target'] for it in lv[0]['init']] == ['X', 'Y', 'Z']
    assert [it['target'] for it in lv[3]['init']] == ['W']
    # X: written in 0 and 2, read in 1 -> projected after 0 and after 2
    assert lv[0]['sym'] == ['X'] and lv[1]['sym'] == ['Y']
    assert lv[2]['sym'] == ['X', 'Z'] and lv[3]['sym'] == ['Y', 'W']


def _run_sequential(items, lists):
    x = {k: v.copy() for k, v in lists.items()}
    for it in items:
        if it['cmd'] == 'INIT':
            x[it['target']] = it['factor'] * x[it['source']] if 'source' in it else 0.0 * x[it['target']]
        elif it['cmd'] == 'CONTR':
            x[it['target']] = x[it['target']] + it['factor'] * x[it['op1']] * x[it['op2']]
    return x


def _run_levels(levels, lists):
    """A level's INITs run in order, then its contractions as one grouped
    launch: every task reads its operands while the others add into their
    targets, so no list may be both read and written in one level."""
    x = {k: v.copy() for k, v in lists.items()}
    for lv in levels:
        for it in lv['init']:
            x[it['target']] = it['factor'] * x[it['source']] if 'source' in it else 0.0 * x[it['target']]
        reads = {c['op1'] for c in lv['contr']} | {c['op2'] for c in lv['contr']}
        writes = {c['target'] for c in lv['contr']}
        assert not reads & writes, (reads & writes, lv)
        add = {}
        for c in lv['contr']:
            add[c['target']] = add.get(c['target'], 0.0) + c['factor'] * x[c['op1']] * x[c['op2']]
        for k, v in add.items():
            x[k] = x[k] + v
    return x


def test_scheduled_levels_equal_the_sequential_formula():
    """Random formulas over a few lists (bilinear items as elementwise
    products): executing the scheduled levels -- a level's contractions
    grouped, reading the lists as they stand at the level's start -- gives
    the item-by-item result, and no level reads a list it writes.  The
    earlier rule that let a writer share a level with an earlier reader of
    its target fails this test."""
    import numpy as np
    rng = np.random.default_rng(3)
    names = ['A', 'B', 'X', 'Y', 'Z', 'W']
    for trial in range(300):
        items = []
        for _ in range(int(rng.integers(3, 14))):
            t = names[2 + int(rng.integers(0, 4))]
            if rng.random() < 0.2:
                src = names[int(rng.integers(0, 6))]
                items.append(F.init(t, src, float(rng.normal())) if src != t and rng.random() < 0.6
                             else F.init(t))
            else:
                o1, o2 = (names[int(i)] for i in rng.integers(0, 6, 2))
                if t in (o1, o2):
                    continue
                items.append(F.contr(t, o1, o2, {}, {}, float(rng.normal())))
        lists = {n: rng.normal(size=3) for n in names}
        seq = _run_sequential(items, lists)
        lev = _run_levels(F.schedule(items), lists)
        for n in names:
            assert np.allclose(seq[n], lev[n], rtol=1e-12, atol=1e-12), (trial, n, items)


def test_formula_files(tmp_path):
    f = F.load(os.path.join(FORMULAS, 'ccsd.json'))
    assert f['lists'] == ccsd.LISTS and f['constant'] == ccsd.INTEGRALS
    assert f['items'] == ccsd.FORMULA['items']
    F.dump(f, tmp_path / 'x.json')
    assert F.load(tmp_path / 'x.json') == f
    m = F.load(os.path.join(FORMULAS, 'mp2.json'))
    lv = F.schedule(m['items'])
    assert _names(lv) == [[('E', 'Hoovv', 'T2')]]
    assert [it['target'] for it in lv[0]['init']] == ['R1', 'R2', 'E']
    bad = tmp_path / 'bad.json'
    bad.write_text('{"lists": {}, "items": [{"cmd": "CONTR", "target": "Q", "op1": "A", "op2": "B", '
                   '"cnt_c": {}, "cnt_a": {}, "factor": 1}]}')
    with pytest.raises(ValueError):
        F.load(bad)


def test_linearised_formula_is_the_derivative():
    """formula.linearize: with the base lists at t, one pass of the delta
    items gives the directional derivative of every list along dt.  Checked
    on the CCSD formula's item structure and on random formulas, with
    bilinear items as elementwise products, against central differences."""
    import numpy as np
    rng = np.random.default_rng(5)

    def run(items, lists):
        x = {k: v.copy() for k, v in lists.items()}
        return _run_sequential(items, x)

    def check(formula, varying, names):
        try:
            items, dn = F.linearize(formula, varying)
        except ValueError:       # a list written after a read: refused
            return False
        assert all(d.startswith('d') for d in dn)
        base = {n: rng.normal(size=4) for n in names}
        dt = {n: rng.normal(size=4) for n in varying}
        h = 1e-5
        plus = dict(base, **{n: base[n] + h * dt[n] for n in varying})
        minus = dict(base, **{n: base[n] - h * dt[n] for n in varying})
        xp, xm, x0 = run(formula['items'], plus), run(formula['items'], minus), run(formula['items'], base)
        lin = dict(x0)
        for n in dn:
            lin[n] = np.zeros(4)
        for n in varying:
            lin['d' + n] = dt[n]
        lin = _run_sequential(items, lin)
        for n in dn:
            fd = (xp[n[1:]] - xm[n[1:]]) / (2 * h)
            assert np.allclose(lin[n], fd, rtol=1e-6, atol=1e-6), (n, lin[n], fd)
        for n in names:
            if 'd' + n not in dn and n not in varying:
                assert np.allclose(xp[n], xm[n])          # lists that do not depend on t
        return True
    ccsd_names = list(ccsd.LISTS)
    assert check(ccsd.FORMULA, ['T1', 'T2'], ccsd_names)
    checked = 0
    names = ['T', 'A', 'B', 'X', 'Y', 'Z']
    for trial in range(100):
        items = [F.init(n) for n in ('X', 'Y', 'Z')]
        for _ in range(int(rng.integers(2, 10))):
            t = names[3 + int(rng.integers(0, 3))]
            o1, o2 = (names[int(i)] for i in rng.integers(0, 6, 2))
            if t in (o1, o2):
                continue
            items.append(F.contr(t, o1, o2, {}, {}, float(rng.normal())))
        checked += check({'items': items}, ['T'], names)
    assert checked > 30
    with pytest.raises(ValueError):
        F.linearize({'items': [F.contr('Y', 'X', 'T', {}, {}, 1.0), F.contr('X', 'T', 'A', {}, {}, 1.0)]},
                    ['T'])
