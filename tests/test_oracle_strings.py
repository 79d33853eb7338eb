"""Oracle pins for strings, graph addressing, maps and ME-list layout (CPU)."""
import itertools

import pytest

from oracle.strings import (Space, strings_of, weight_4sg, idx4sg, iordstr2,
                            n_strings_closed_form, to_spinorbs, from_spinorbs,
                            strmap, flipmap, string_index)
from oracle.melist import Orbitals, MEList

SPACES = [([1] * 5, 1), ([1, 2, 1, 2, 3, 4], 4), ([1, 2, 3, 4, 1, 2, 3], 4),
          ([1, 2, 1, 2, 1, 2, 1], 2)]


@pytest.mark.parametrize('irr,nsym', SPACES)
def test_enumeration_order_is_graph_index(irr, nsym):
    """next_string order == idx4sg over weight_4sg arcs (independent route)."""
    sp = Space(irr, nsym)
    for nel in range(0, 5):
        iy = weight_4sg(sp, nel, nel)
        for ms in range(nel, -nel - 1, -2):
            for g in range(1, nsym + 1):
                ss = strings_of(sp, nel, ms, g)
                assert [idx4sg(s, sp, iy) for s in ss] == list(range(len(ss)))


@pytest.mark.parametrize('irr,nsym', SPACES)
def test_string_counts_closed_form(irr, nsym):
    sp = Space(irr, nsym)
    nb = [sum(1 for x in irr if x == g) for g in range(1, nsym + 1)]
    for nel in range(0, 5):
        for ms in range(nel, -nel - 1, -2):
            for g in range(1, nsym + 1):
                assert len(strings_of(sp, nel, ms, g)) == \
                    n_strings_closed_form(nb, nel, ms, g, nsym)


def test_strings_are_canonical_and_distinct():
    sp = Space([1, 2, 1, 2, 1], 2)
    for nel in range(1, 5):
        seen = set()
        for ms in range(nel, -nel - 1, -2):
            for g in (1, 2):
                for s in strings_of(sp, nel, ms, g):
                    so = to_spinorbs(s)
                    key = [(o, 0 if p == 1 else 1) for o, p in so]
                    assert key == sorted(key) and len(set(key)) == len(key)
                    assert from_spinorbs(so) == s
                    assert s not in seen
                    seen.add(s)


def _parity(seq):
    seq = list(seq)
    n = 0
    for i in range(len(seq)):
        for j in range(i + 1, len(seq)):
            if seq[i] > seq[j]:
                n += 1
    return -1 if n % 2 else 1


def test_merge_sign_is_sort_parity():
    """iordstr2 sign == parity of sorting the concatenated spin-orbitals
    (alpha before beta on one orbital); 0 iff a spin-orbital repeats."""
    sp = Space([1, 1, 1, 1], 1)
    for n1, n2 in [(1, 1), (2, 1), (1, 2), (2, 2), (3, 1)]:
        for ms1 in range(n1, -n1 - 1, -2):
            for ms2 in range(n2, -n2 - 1, -2):
                for s1 in strings_of(sp, n1, ms1, 1):
                    for s2 in strings_of(sp, n2, ms2, 1):
                        so = [(o, 0 if p == 1 else 1) for o, p in to_spinorbs(s1) + to_spinorbs(s2)]
                        sgn, s12 = iordstr2(s1, s2)
                        if len(set(so)) < len(so):
                            assert sgn == 0
                            continue
                        assert sgn == _parity(so)
                        assert [(o, 0 if p == 1 else 1) for o, p in to_spinorbs(s12)] == sorted(so)


def test_strmap_entries_point_at_merged_string():
    sp = Space([1, 2, 1, 2, 1], 2)
    for (n1, m1, g1), (n2, m2, g2) in itertools.product([(1, 1, 1), (1, 1, 2), (1, -1, 1)],
                                                        [(1, 1, 2), (1, -1, 1), (2, 0, 2)]):
        mp = strmap(sp, n1, (n1 - m1) // 2 + 1, g1, n2, (n2 - m2) // 2 + 1, g2)
        S1 = strings_of(sp, n1, m1, g1)
        S2 = strings_of(sp, n2, m2, g2)
        assert len(mp) == len(S1) * len(S2)
        from oracle.strings import multd2h
        for j, s2 in enumerate(S2):
            for i, s1 in enumerate(S1):
                e = mp[i + len(S1) * j]
                sgn, s12 = iordstr2(s1, s2)
                if sgn == 0:
                    assert e == 0
                else:
                    assert e == sgn * (1 + string_index(sp, n1 + n2, m1 + m2, multd2h(g1, g2), s12))


def test_flipmap_is_an_involution():
    sp = Space([1, 2, 1, 2, 1], 2)
    for nel in (1, 2, 3):
        for ms in range(nel, -nel - 1, -2):
            for g in (1, 2):
                f = flipmap(sp, nel, (nel - ms) // 2 + 1, g)
                fb = flipmap(sp, nel, (nel + ms) // 2 + 1, g)
                for i, e in enumerate(f):
                    j = abs(e) - 1
                    assert abs(fb[j]) - 1 == i
                    assert (1 if e > 0 else -1) * (1 if fb[j] > 0 else -1) == 1


def test_melist_lengths_closed_form():
    """T2 and V(PP,PP) of the nocc=8, nvir=32, C1 ladder: C(n,2)/n^2/C(n,2)
    strings per Ms, so T2 = 496*28 + 1024*64 + 496*28."""
    orbs = Orbitals([1] * 8, [1] * 32, 1)
    assert MEList(orbs, {2: 2}, {1: 2}).length == 93312
    assert MEList(orbs, {2: 2}, {2: 2}).length == 1540608
    assert MEList(orbs, {2: 1}, {1: 1}).length == 2 * 32 * 8
    assert MEList(orbs, {}, {}).length == 1
