"""Formula lists (gecco-public_amd/formula.py): the schedule of GeCCo's
formula items into grouped launches, the committed formula files, and the
CCSD driver's levels derived from its flat formula."""
import os

import numpy as np
import pytest

import ccsd
import formula as F

H, P = 1, 2
HERE = os.path.dirname(os.path.abspath(__file__))
FORMULAS = os.path.join(HERE, '..', 'gecco-public_amd', 'formulas')


def _names(levels):
    return [[(c['target'], c['op1'], c['op2']) for c in lv['contr']] for lv in levels]


def test_ccsd_levels_are_the_dependency_levels():
    """The flat CCSD formula schedules into three levels of 16, 10 and 7
    contractions in list order: every contraction of level 1 reads a list
    written in level 0, every one of level 2 a list written in level 1; the
    ADD items F2ae += Fae, F2mi += Fmi run in level 2's initialisation
    launch, before the contractions that read F2ae and F2mi."""
    levels = ccsd.SCHEDULE
    assert [len(lv['contr']) for lv in levels] == [16, 10, 7]
    flat = [t[:3] for t in ccsd.TERMS]
    assert _names(levels) == [flat[:16], flat[16:26], flat[26:]]
    assert [(a['target'], a['source']) for a in levels[2]['add']] == [('F2ae', 'Fae'), ('F2mi', 'Fmi')]
    assert not levels[0]['add'] and not levels[1]['add']
    written = [{c['target'] for c in lv['contr']} for lv in levels]
    for li in (1, 2):
        for c in levels[li]['contr']:
            assert c['op1'] in written[li - 1] or c['op2'] in written[li - 1], c
    # a list's initialisation runs in the level of its first write (E is
    # first written in level 1), projections once per list
    assert {it['target'] for it in levels[0]['init']} == \
        {n for n in ccsd.LISTS if n not in ccsd.INTEGRALS and n not in ('T1', 'T2', 'E', 'F2ae', 'F2mi')}
    assert [it['target'] for it in levels[1]['init']] == ['F2ae', 'F2mi', 'E'] and not levels[2]['init']
    assert ccsd.SYM_AFTER == [['tau', 'taut', 'Fme', 'Z'],
                              ['Fae', 'Fmi', 'F2ae', 'F2mi', 'Wmnij', 'Zamij', 'Wmbej', 'E'],
                              ['R1', 'R2', 'F2ae', 'F2mi']]


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
    # an INIT to zero runs in the level of its list's first write
    assert [it['target'] for it in lv[0]['init']] == ['X', 'Z']
    assert [it['target'] for it in lv[1]['init']] == ['Y']
    assert [it['target'] for it in lv[3]['init']] == ['W']
    # X: written in 0 and 2, read in 1 -> projected after 0 and after 2
    assert lv[0]['sym'] == ['X'] and lv[1]['sym'] == ['Y']
    assert lv[2]['sym'] == ['X', 'Z'] and lv[3]['sym'] == ['Y', 'W']


def _run_sequential(items, lists):
    x = {k: v.copy() for k, v in lists.items()}
    for it in F.expand(items):
        if it['cmd'] == 'INIT':
            x[it['target']] = it['factor'] * x[it['source']] if 'source' in it else 0.0 * x[it['target']]
        elif it['cmd'] == 'ADD':
            x[it['target']] = x[it['target']] + it['factor'] * x[it['source']]
        elif it['cmd'] == 'CONTR':
            x[it['target']] = x[it['target']] + it['factor'] * x[it['op1']] * x[it['op2']]
        elif it['cmd'] == 'ADD_REO':        # a reorder as the identity map
            x[it['target']] = x[it['target']] + it['factor'] * it['sign'] * x[it['source']]
    return x


class _Store(dict):
    """Lists as views: pooled ones into one shared array."""

    def __init__(self, lists, offsets=None, pool=0):
        super().__init__()
        self.pool = np.full(max(pool, 1), np.nan)
        for k, v in lists.items():
            if offsets and k in offsets:
                self.views = getattr(self, 'views', {})
                dict.__setitem__(self, k, self.pool[offsets[k]:offsets[k] + len(v)])
            else:
                dict.__setitem__(self, k, v.copy())

    def __setitem__(self, k, v):
        self[k][...] = v


def _run_levels(levels, lists, offsets=None, pool=0):
    """A level's INITs and ADDs run as one launch (an INIT with a source
    and an ADD read their sources as they stand at the level's start), then
    its contractions as one grouped launch: every task reads its operands
    while the others add into their targets, so no list may be both read
    and written in one level.  offsets: pooled lists share one array."""
    x = _Store(lists, offsets, pool)
    for lv in levels:
        reads0 = {i['source'] for i in lv['init'] if 'source' in i} | {a['source'] for a in lv['add']}
        writes0 = {i['target'] for i in lv['init']} | {a['target'] for a in lv['add']}
        assert not reads0 & writes0, (reads0 & writes0, lv)
        assert not {i['target'] for i in lv['init']} & {a['target'] for a in lv['add']}, lv
        old = {k: np.array(x[k]) for k in reads0}
        for it in lv['init']:
            x[it['target']] = it['factor'] * old[it['source']] if 'source' in it else 0.0
        for a in lv['add']:
            x[a['target']] = x[a['target']] + a['factor'] * old[a['source']]
        reads = {c['op1'] for c in lv['contr']} | {c['op2'] for c in lv['contr']}
        writes = {c['target'] for c in lv['contr']}
        assert not reads & writes, (reads & writes, lv)
        add = {}
        for c in lv['contr']:
            add[c['target']] = add.get(c['target'], 0.0) + c['factor'] * x[c['op1']] * x[c['op2']]
        for k, v in add.items():
            x[k] = x[k] + v
        # the level's reorders, after its contractions: they read lists that
        # no contraction or reorder of the level writes
        rr = {r['source'] for r in lv.get('reo', [])}
        rw = {r['target'] for r in lv.get('reo', [])}
        assert not rr & (rw | writes), (rr, lv)
        assert not rw & (reads | reads0), (rw, lv)
        for r in lv.get('reo', []):
            x[r['target']] = x[r['target']] + r['factor'] * r['sign'] * x[r['source']]
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
        items = _random_items(rng, names)
        lists = {n: rng.normal(size=3) for n in names}
        seq = _run_sequential(items, lists)
        lev = _run_levels(F.schedule(items), lists)
        for n in names:
            assert np.allclose(seq[n], lev[n], rtol=1e-12, atol=1e-12), (trial, n, items)


def _random_items(rng, names):
    """Random INIT (zero or copy), ADD, BC and CONTR items over the lists
    names[2:] with operands from all of names."""
    items = []
    for _ in range(int(rng.integers(3, 16))):
        t = names[2 + int(rng.integers(0, len(names) - 2))]
        u = rng.random()
        if u < 0.2:
            src = names[int(rng.integers(0, len(names)))]
            items.append(F.init(t, src, float(rng.normal())) if src != t and rng.random() < 0.6
                         else F.init(t))
        elif u < 0.3:
            src = names[int(rng.integers(0, len(names)))]
            if src != t:
                items.append(F.add(t, src, float(rng.normal())))
        elif u < 0.4:                       # REORDER / ADD_REO
            src = names[int(rng.integers(0, len(names)))]
            if src != t:
                items.append(F.reorder(t, src, 0, 0, {}, {}, float(rng.normal()),
                                       int(rng.choice([-1, 1])), add=rng.random() < 0.6))
        else:
            o1, o2 = (names[int(i)] for i in rng.integers(0, len(names), 2))
            if t in (o1, o2):
                continue
            make = F.bc if rng.random() < 0.15 else F.contr
            items.append(make(t, o1, o2, {}, {}, float(rng.normal())))
    return items


def test_pooled_intermediates_equal_the_sequential_formula():
    """formula.lifetimes and formula.pool_layout: with every intermediate
    that has a lifetime placed in one shared array (lists whose lifetimes
    meet never overlap), executing the scheduled levels gives the outputs of
    the item-by-item run.  Random formulas whose outputs are the lists named
    in `keep`; A and B are constant inputs."""
    rng = np.random.default_rng(11)
    names = ['A', 'B', 'X', 'Y', 'Z', 'W', 'V', 'O']
    shared = 0
    for trial in range(400):
        items = _random_items(rng, names)
        form = {'items': items, 'constant': ['A', 'B']}
        keep = ['O']
        spans = F.lifetimes(form, keep=keep)
        assert not set(spans) & {'A', 'B', 'O'}
        sizes = {n: 3 for n in spans}
        off, pool = F.pool_layout(spans, sizes, align=1)
        for a in spans:               # lists alive together never overlap
            for b in spans:
                if a < b and not (spans[a][1] < spans[b][0] or spans[b][1] < spans[a][0]):
                    assert off[a] + 3 <= off[b] or off[b] + 3 <= off[a], (a, b, spans, off)
        shared += pool < 3 * len(spans)
        lists = {n: rng.normal(size=3) for n in names}
        seq = _run_sequential(items, lists)
        lev = _run_levels(F.schedule(items), lists, off, pool)
        for n in keep + ['A', 'B']:
            assert np.allclose(seq[n], lev[n], rtol=1e-12, atol=1e-12), (trial, n, items, spans)
    assert shared >= 3          # memory is actually shared in some of them


def test_new_and_delete_bound_lifetimes():
    c = F.contr
    items = [F.new('X'), F.init('X'), c('X', 'A', 'B', {}, {}, 1.0), c('O', 'X', 'A', {}, {}, 1.0),
             F.delete('X'), F.new('Y'), F.init('Y'), c('Y', 'O', 'A', {}, {}, 1.0), c('Q', 'Y', 'B', {}, {}, 1.0),
             F.delete('Y')]
    spans = F.lifetimes({'items': items, 'constant': ['A', 'B']}, keep=['O', 'Q'])
    assert spans == {'X': (0, 1), 'Y': (2, 3)}
    off, pool = F.pool_layout(spans, {'X': 10, 'Y': 10}, align=1)
    assert off == {'X': 0, 'Y': 0} and pool == 10
    with pytest.raises(ValueError):
        F.lifetimes({'items': items + [c('O', 'X', 'A', {}, {}, 1.0)], 'constant': ['A', 'B']})


def test_formula_files(tmp_path):
    f = F.load(os.path.join(FORMULAS, 'ccsd.json'))
    assert f['lists'] == ccsd.LISTS and f['constant'] == ccsd.INTEGRALS
    assert f['items'] == ccsd.FORMULA['items']
    F.dump(f, tmp_path / 'x.json')
    assert F.load(tmp_path / 'x.json') == f
    m = F.load(os.path.join(FORMULAS, 'mp2.json'))
    lv = F.schedule(m['items'])
    assert _names(lv) == [[('E', 'Hoovv', 'T2')]]
    assert sorted(it['target'] for it in lv[0]['init']) == ['E', 'R1', 'R2']
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
            u = rng.random()
            if u < 0.35:        # ADD and ADD_REO items: dX += f dS (f R(dS))
                src = names[int(rng.integers(0, 6))]
                if src != t:
                    items.append(F.add(t, src, float(rng.normal())) if u < 0.2 else
                                 F.reorder(t, src, 0, 0, {}, {}, float(rng.normal()), -1, add=True))
                continue
            o1, o2 = (names[int(i)] for i in rng.integers(0, 6, 2))
            if t in (o1, o2):
                continue
            items.append(F.contr(t, o1, o2, {}, {}, float(rng.normal())))
        checked += check({'items': items}, ['T'], names)
    assert checked > 30
    with pytest.raises(ValueError):
        F.linearize({'items': [F.contr('Y', 'X', 'T', {}, {}, 1.0), F.contr('X', 'T', 'A', {}, {}, 1.0)]},
                    ['T'])


def test_add_bc_new_delete_formula_gives_the_ccsd_energy():
    """The CCSD formula written with ADD, BC, NEW and DELETE items
    (tests/models.py) gives the hand formula's energy and residuals on the
    oracle's route 3, and its NEW/DELETE spans give the intermediates
    lifetimes."""
    from models import ccsd_with_add_bc_new_delete
    from oracle.ccsd_route3 import Route3CCSD
    from synth import synthetic
    f = ccsd_with_add_bc_new_delete()
    cmds = {it['cmd'] for it in f['items']}
    assert {'ADD', 'BC', 'NEW', 'DELETE'} <= cmds
    ham = synthetic(3, 6, 18, nsym=2, seed=4, sigma=0.1)
    e0, _, _ = Route3CCSD(ham, ccsd.FORMULA, ccsd.SCHEDULE).solve(tol=1e-11)
    e1, _, _ = Route3CCSD(ham, f, F.schedule(f['items'])).solve(tol=1e-11)
    assert abs(e0 - e1) < 1e-12
    spans = F.lifetimes(f, keep=['T1', 'T2', 'R1', 'R2', 'E'])
    assert set(spans) == {n for n in ccsd.LISTS if n not in ccsd.INTEGRALS and
                          n not in ('T1', 'T2', 'R1', 'R2', 'E')}


def test_joined_intermediate_and_reorder_formula_gives_the_ccsd_energy(tmp_path):
    """A CCSD formula whose T1 x T1 products go through the two-vertex
    intermediate X11 and ADD_REO items (tests/models.py) gives the hand
    formula's energy on the oracle's route 3; the formula file round-trips
    its vertex data and reorder items."""
    from models import ccsd_with_joined_tau
    from oracle.ccsd_route3 import Route3CCSD
    from synth import synthetic
    f = ccsd_with_joined_tau()
    assert {'ADD_REO'} <= {it['cmd'] for it in f['items']}
    assert any('vertices' in it for it in f['items'])
    F.dump(f, tmp_path / 'j.json')
    g = F.load(tmp_path / 'j.json')
    assert F.schedule(g['items']) == F.schedule(f['items'])
    lv = F.schedule(f['items'])
    assert any(l['reo'] for l in lv)
    ham = synthetic(3, 6, 18, nsym=2, seed=4, sigma=0.1)
    e0, _, _ = Route3CCSD(ham, ccsd.FORMULA, ccsd.SCHEDULE).solve(tol=1e-11)
    e1, _, _ = Route3CCSD(ham, g, F.schedule(g['items'])).solve(tol=1e-11)
    assert abs(e0 - e1) < 1e-12
