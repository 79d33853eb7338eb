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
             c('X', 'A', 'A', {}, {}, 1.0),       # writes X after a read of X: level 1 (not 0)
             c('Z', 'A', 'B', {}, {}, 1.0),       # independent: level 0
             c('Z', 'Y', 'A', {}, {}, 1.0),       # reads Y: level 2
             c('Y', 'A', 'A', {}, {}, 1.0),       # writes Y after its read: level 2
             F.init('W', 'Z', 2.0),               # after Z's writes: level 3
             c('W', 'A', 'A', {}, {}, 1.0)]       # after W's init: level 3
    lv = F.schedule(items)
    assert _names(lv) == [[('X', 'A', 'B'), ('Z', 'A', 'B')],
                          [('Y', 'X', 'B'), ('X', 'A', 'A')],
                          [('Z', 'Y', 'A'), ('Y', 'A', 'A')],
                          [('W', 'A', 'A')]]
    assert [it['target'] for it in lv[0]['init']] == ['X', 'Y', 'Z']
    assert [it['target'] for it in lv[3]['init']] == ['W']
    # X: written in 0 and 1, read in 1 -> projected after 0 and after 1
    assert lv[0]['sym'] == ['X'] and lv[1]['sym'] == ['Y', 'X']
    assert lv[2]['sym'] == ['Z', 'Y'] and lv[3]['sym'] == ['W']


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
