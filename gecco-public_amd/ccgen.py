"""Coupled-cluster amplitude equations as formula lists of binary
contractions, generated for any excitation level.

The role of GeCCo's symbolic layer for the standard truncated CC methods
(the method targets of cc_spec/ expanded and factorised by formula/,
contract/frm_sched2.f runs the result): the projections

    E     = <0| (H e^T)_c |0>
    R_r   = <r| (H e^T)_c |0>  - D t_r          (r = 1 .. n)

with T = T1 + ... + Tn, for a canonical reference (diagonal Fock operator,
whose contribution D t_r is the solver's diagonal preconditioner, so the
formula carries only the two-electron operator V).  Every connected term
(V T_k1 ... T_km)_c is written as a chain of binary contractions,
V . T_k1 -> I1, I1 . T_k2 -> I2, ..., the last one added into R_r or E:

  * V's creators of holes and annihilators of particles are the lines the
    T's contract with (a T has particle creators and hole annihilators
    only), so every such line of V is contracted and V's other lines stay
    external; each T meets V in at least one line (connectedness);
  * contraction j takes h_j of V's hole creators and p_j of its particle
    annihilators: cnt_c = {H: h_j} on the intermediate's C side, cnt_a =
    {P: p_j} on its A side (the engine's operand roles, contr_op1op2);
  * e^T contributes 1/m! for m factors; the m! orderings of distinct
    (rank, h, p) factors give the same operator, so each multiset of
    factors appears once in a canonical order with the factor
    1 / prod(multiplicity!) of identical factors;
  * a contraction of the engine is the operator product restricted to the
    given contraction pattern (restricted strings, all signs from the
    string merges), so no other factor enters.
Chains that share a prefix share its intermediate list.

Lists: T1..Tn, R1..Rn, E, the nine two-electron integral lists
V<nh C><np C><nh A><np A> = <C||A> (C holes/particles, A holes/particles),
and the intermediates.  The formula runs on ccsd.CCSD (formula=...).
"""
import itertools
import math

import formula as _f

H, P = 1, 2


def vname(hc, pc, ha, pa):
    return f'V{hc}{pc}{ha}{pa}'


def _occ(h, p):
    return {k: v for k, v in ((H, h), (P, p)) if v}


# component order of a string side: particles before holes (the list
# layout's hpvx sequence X, P, V, H)
_RANK = {P: 0, H: 1}


def _perm_parity(seq):
    """Parity (+1/-1) of the permutation that sorts seq."""
    seq = list(seq)
    sgn = 1
    for i in range(len(seq)):
        for j in range(i + 1, len(seq)):
            if seq[i] > seq[j]:
                sgn = -sgn
    return sgn


def _merge_sign(part1, part2):
    """The kernel's sign for merging two strings (lists of (space, label)):
    per space, the parity of sorting part1's labels followed by part2's."""
    sgn = 1
    for sp in (H, P):
        sgn *= _perm_parity([l for s, l in part1 if s == sp] + [l for s, l in part2 if s == sp])
    return sgn


def _canon(ops):
    return sorted(ops, key=lambda o: (_RANK[o[0]], o[1]))


def bc_sign(op1, op2, cnt_c, cnt_a):
    """Sign that turns the engine's binary contraction op1 . op2 over
    (cnt_c, cnt_a) into the operator product: op1/op2 = (C occupation, A
    occupation) dicts over H, P.  A list element with C strings C and A
    strings A is the coefficient of a+(C) a(A reversed), C and A in
    component order (particles, then holes; each ascending).  The engine
    multiplies the elements with the per-space merge signs of its string
    maps (result C = EX1 C + EX2 C, A = EX2 A + EX1 A; op1 C = CNT C + EX1 C,
    op1 A = CNT A + EX1 A, op2 C = CNT A + EX2 C, op2 A = CNT C + EX2 A);
    the operator product a+(C1) a(A1 rev) a+(C2) a(A2 rev) with the
    contracted pairs taken out (+1 each: a particle annihilator of op1
    before its creator in op2, a hole creator of op1 before its annihilator
    in op2) and the rest in the result's order carries the parity of the
    reordering.  The ratio of the two does not depend on the orbital labels,
    so one labelling decides it."""
    (c1, a1), (c2, a2) = op1, op2
    lab = iter(range(1000))
    part = lambda occ: [(sp, next(lab)) for sp in (P, H) for _ in range(occ.get(sp, 0))]
    cc, ca = part(cnt_c), part(cnt_a)
    sub = lambda o, c: {sp: o.get(sp, 0) - c.get(sp, 0) for sp in (H, P)}
    e1c, e1a = part(sub(c1, cnt_c)), part(sub(a1, cnt_a))
    e2c, e2a = part(sub(c2, cnt_a)), part(sub(a2, cnt_c))
    # the engine's sign for this labelling
    eng = (_merge_sign(cc, e1c) * _merge_sign(ca, e1a) * _merge_sign(ca, e2c) * _merge_sign(cc, e2a) *
           _merge_sign(e1c, e2c) * _merge_sign(e2a, e1a))
    # the operator product: (op, label, creator?) in the order of the product
    C1, A1 = _canon(cc + e1c), _canon(ca + e1a)
    C2, A2 = _canon(ca + e2c), _canon(cc + e2a)
    seq = [(o, True) for o in C1] + [(o, False) for o in reversed(A1)] + \
          [(o, True) for o in C2] + [(o, False) for o in reversed(A2)]
    target = []
    for o in ca:                 # particle lines: op1's annihilator, op2's creator
        target += [(o, False), (o, True)]
    for o in cc:                 # hole lines: op1's creator, op2's annihilator
        target += [(o, True), (o, False)]
    target += [(o, True) for o in _canon(e1c + e2c)] + [(o, False) for o in reversed(_canon(e2a + e1a))]
    pos = {x: i for i, x in enumerate(seq)}
    alg = _perm_parity([pos[x] for x in target])
    return alg * eng


def _factors(rank_max, need_h, need_p):
    """Ordered (non-increasing) tuples of T factors (k, h, p) whose h sum to
    need_h and p to need_p, each with h + p >= 1."""
    single = [(k, h, p) for k in range(rank_max, 0, -1) for h in range(min(k, need_h), -1, -1)
              for p in range(min(k, need_p), -1, -1) if h + p >= 1]
    out = []

    def rec(prefix, start, rh, rp):
        if rh == 0 and rp == 0:
            if prefix:
                out.append(tuple(prefix))
            return
        for i in range(start, len(single)):
            k, h, p = single[i]
            if h <= rh and p <= rp:
                rec(prefix + [single[i]], i, rh - h, rp - p)
    rec([], 0, need_h, need_p)
    return out


def generate(rank_max, name=None):
    """The formula ({'lists', 'constant', 'items'}) of CC with T1..T_rank_max
    (2: CCSD, 3: CCSDT)."""
    lists = {}
    for k in range(1, rank_max + 1):
        lists[f'T{k}'] = (_occ(0, k), _occ(k, 0))
    for k in range(1, rank_max + 1):
        lists[f'R{k}'] = (_occ(0, k), _occ(k, 0))
    lists['E'] = ({}, {})
    vl = []
    for hc in range(3):
        for ha in range(3):
            n = vname(hc, 2 - hc, ha, 2 - ha)
            lists[n] = (_occ(hc, 2 - hc), _occ(ha, 2 - ha))
            vl.append((n, hc, 2 - hc, ha, 2 - ha))
    items = []
    inits, contrs = [], []
    made = set()
    for n, hc, pc, ha, pa in vl:
        # hc hole creators and pa particle annihilators contract with the T's
        for facs in _factors(rank_max, hc, pa):
            r_c = pc + sum(k - p for k, h, p in facs)
            r_a = ha + sum(k - h for k, h, p in facs)
            if r_c != r_a or r_c > rank_max:
                continue
            mult = 1
            for _, grp in itertools.groupby(facs):
                mult *= math.factorial(len(list(grp)))
            target = 'E' if r_c == 0 else f'R{r_c}'
            cur, cur_occ = n, ((hc, pc), (ha, pa))     # (C (h, p), A (h, p))
            for j, (k, h, p) in enumerate(facs):
                (ch, cp), (ah, ap) = cur_occ
                new = ((ch - h, cp + k - p), (ah + k - h, ap - p))
                sgn = bc_sign((_occ(ch, cp), _occ(ah, ap)), (_occ(0, k), _occ(k, 0)), _occ(h, 0), _occ(0, p))
                last = j == len(facs) - 1
                if last:
                    contrs.append(_f.contr(target, cur, f'T{k}', _occ(h, 0), _occ(0, p), sgn / mult))
                else:
                    iname = 'I_' + n + '_' + '_'.join(f'{a}{b}{c}' for a, b, c in facs[:j + 1])
                    if iname not in made:
                        made.add(iname)
                        lists[iname] = (_occ(*new[0]), _occ(*new[1]))
                        inits.append(_f.init(iname))
                        contrs.append(_f.contr(iname, cur, f'T{k}', _occ(h, 0), _occ(0, p), float(sgn)))
                    cur = iname
                cur_occ = new
    # V with no T: the driving term of R2, <ab||ij>
    res_init = [_f.init(f'R{k}', vname(0, 2, 2, 0), 1.0) if k == 2 else _f.init(f'R{k}')
                for k in range(1, rank_max + 1)]
    items = res_init + [_f.init('E')] + inits + contrs
    constant = [v[0] for v in vl]
    # order: intermediates before their readers (the generator appends a
    # chain's prefixes before its last step; sort contractions so that every
    # intermediate's writes precede its reads)
    return {'lists': lists, 'constant': constant, 'items': _order(items)}


def _order(items):
    """Inits first, then contractions grouped so that all writes of a list
    come before its first read (the chains are acyclic)."""
    inits = [it for it in items if it['cmd'] == 'INIT']
    contrs = [it for it in items if it['cmd'] == 'CONTR']
    depth = {}

    def d(name):
        if name in depth:
            return depth[name]
        ws = [c for c in contrs if c['target'] == name]
        v = 0
        for c in ws:
            v = max(v, 1 + max(d(c['op1']), d(c['op2'])))
        depth[name] = v
        return v
    contrs.sort(key=lambda c: d(c['target']))
    return inits + contrs
