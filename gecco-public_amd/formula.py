"""Formula lists for the engine: a serialised form of GeCCo's formula items
and the schedule that turns them into grouped plan launches.

The reference keeps a factorised formula as a linked list of items
(include/def_formula_item.h:1-28) with a command code each
(include/def_formula_item.h:28-46) and executes it item by item in
contract/frm_sched2.f:117-365.  The items kept here, with their reference
commands:

  {"cmd": "INIT", "target": L}                  command_set_target_init (0):
                                                L = 0
  {"cmd": "INIT", "target": L, "source": S,     command_cp_intm (14) with a
   "factor": f}                                 factor: L = f * S
  {"cmd": "CONTR", "target": L, "op1": A,       command_add_bc (8): the binary
   "op2": B, "cnt_c": {...}, "cnt_a": {...},    contraction of A and B over the
   "factor": f}                                 strings of occupation cnt_c (on
                                                A's creator side) and cnt_a (on
                                                A's annihilator side), added to
                                                L (contr_op1op2)
  {"cmd": "SYMMETRISE", "target": L}            command_symmetrise (5): the
                                                spin-flip projection of L
                                                (contract/sym_ab_blk.f)
  {"cmd": "ADD", "target": L, "source": S,      command_add_intm (7),
   "factor": f}                                 contract/fs_add_drv.f: L += f * S
                                                (S with L's occupation)
  {"cmd": "BC", ... as CONTR}                   command_bc (10): L = f A . B,
                                                i.e. INIT L = 0 then the CONTR
  {"cmd": "REORDER", "target": L, "source": S,  command_reorder (6),
   "from": v, "to": w, "shift_c": {...},        contract/fs_reo_drv.f: L = f *
   "shift_a": {...}, "factor": f, "sign": s}    s * (S with the lines shift_c /
                                                shift_a moved from its vertex v
                                                to vertex w of L; vertices
                                                0-based; reo_op_wmaps_c)
  {"cmd": "ADD_REO", ... as REORDER}            command_add_reo (13): L += the
                                                same
  {"cmd": "NEW", "target": L}                   command_new_intermediate (2) and
  {"cmd": "DELETE", "target": L}                command_del_intermediate (3)
                                                (contract/fs_newintm_drv.f,
                                                fs_delintm_drv.f): the span of
                                                items in which intermediate L
                                                exists

Occupations are dicts over the spaces {"H": n, "P": n}.  A formula file is
JSON: {"lists": {name: {"C": occ, "A": occ}}, "constant": [names],
"items": [...]}.  A list of several vertices (an intermediate with joined
vertices) gives "C" and "A" as lists of occupations, one per vertex; a
CONTR or BC item between such lists carries "vertices": {"ex1": [[C, A],
...], "ex2": [...], "cnt": [...], "merge": [merge_op1, merge_op2,
merge_op1op2, merge_op2op1]} -- contr_op1op2's iocc_ex1, iocc_ex2, iocc_cnt
and short-form merge maps (contract/contr_op1op2.f:2-17,
contract/get_bc_info3.f:271-330).  A CONTR or BC item may carry "tra":
[op1, op2, target], contr_op1op2's dagger flags tra_op1, tra_op2 and
tra_op1op2: the flagged one-vertex list takes part as its adjoint (C and A
exchanged, the same storage).  Every list is declared with the formula; NEW and DELETE
(or, without them, the items themselves) bound the levels in which an
intermediate holds data (`lifetimes`), so that a driver may let intermediates
whose lifetimes do not meet share device memory (`pool_layout`).

Schedule: items run in levels; the items of a level are independent and
run as one grouped launch.  A CONTR item goes into the first level after
every earlier (in list order) write of its operands -- all contributions to
a list commute, so several writers share a level -- and into a level after
every earlier item that reads its target (the reader must see the target
as it was before this contribution; a level's contributions are added while
its other items read).  An INIT goes after the earlier reads
and writes of its target and before the later ones.  The spin-flip
projection is linear and idempotent and the initial contents of every list
are fixed points of it, so contributions of several levels are projected
together: a list is projected at the end of the last level that writes it
before a later level reads it, and residual lists (never read) after their
last write; explicit SYMMETRISE items are thereby implied.
"""
import json

H, P = 1, 2
_SPACE = {'H': H, 'P': P}
_NAME = {H: 'H', P: 'P'}


def _occ_in(o):
    if isinstance(o, list):
        return [_occ_in(x) for x in o]
    return {_SPACE[k]: int(v) for k, v in o.items() if int(v)}


def _occ_out(o):
    if isinstance(o, list):
        return [_occ_out(x) for x in o]
    return {_NAME[k]: int(v) for k, v in o.items() if int(v)}


def _vertices_io(v, conv):
    return {'ex1': [[conv(c), conv(a)] for c, a in v['ex1']],
            'ex2': [[conv(c), conv(a)] for c, a in v['ex2']],
            'cnt': [[conv(c), conv(a)] for c, a in v['cnt']],
            'merge': [[int(x) for x in m] for m in v['merge']]}


def contr(target, op1, op2, cnt_c, cnt_a, factor):
    """A CONTR item (occupations keyed by the engine's space codes)."""
    return {'cmd': 'CONTR', 'target': target, 'op1': op1, 'op2': op2,
            'cnt_c': dict(cnt_c), 'cnt_a': dict(cnt_a), 'factor': float(factor)}


def contr_n(target, op1, op2, ex1, ex2, cnt, merge, factor):
    """A CONTR item between lists of several vertices: ex1 / ex2 / cnt are
    per-vertex (occ_c, occ_a) of EX1, EX2 and the contraction lines, merge
    the four short-form merge maps (gecco_hip_plan_add_contraction_n)."""
    it = contr(target, op1, op2, {}, {}, factor)
    it['vertices'] = _vertices_io({'ex1': ex1, 'ex2': ex2, 'cnt': cnt, 'merge': merge}, dict)
    return it


def reorder(target, source, frm, to, shift_c, shift_a, factor=1.0, sign=1, add=False):
    """A REORDER item (target = factor * sign * the reordered source; with
    add=True an ADD_REO item, target += ...)."""
    return {'cmd': 'ADD_REO' if add else 'REORDER', 'target': target, 'source': source,
            'from': int(frm), 'to': int(to), 'shift_c': dict(shift_c), 'shift_a': dict(shift_a),
            'factor': float(factor), 'sign': int(sign)}


def init(target, source=None, factor=1.0):
    it = {'cmd': 'INIT', 'target': target}
    if source is not None:
        it.update(source=source, factor=float(factor))
    return it


def add(target, source, factor=1.0):
    """An ADD item: target += factor * source."""
    return {'cmd': 'ADD', 'target': target, 'source': source, 'factor': float(factor)}


def bc(target, op1, op2, cnt_c, cnt_a, factor):
    """A BC item: target = factor * op1 . op2 (no update)."""
    return dict(contr(target, op1, op2, cnt_c, cnt_a, factor), cmd='BC')


def new(target):
    return {'cmd': 'NEW', 'target': target}


def delete(target):
    return {'cmd': 'DELETE', 'target': target}


_COMMANDS = ('INIT', 'CONTR', 'SYMMETRISE', 'ADD', 'BC', 'NEW', 'DELETE', 'REORDER', 'ADD_REO')


def expand(items):
    """The items with every BC written as INIT (zero) then CONTR, and every
    REORDER as INIT (zero) then ADD_REO."""
    out = []
    for it in items:
        if it['cmd'] == 'BC':
            out.append(init(it['target']))
            out.append(dict(it, cmd='CONTR'))
        elif it['cmd'] == 'REORDER':
            out.append(init(it['target']))
            out.append(dict(it, cmd='ADD_REO'))
        else:
            out.append(it)
    return out


def dump(formula, path):
    """Write a formula ({"lists", "constant", "items"}) as JSON."""
    out = {'lists': {n: {'C': _occ_out(c), 'A': _occ_out(a)} for n, (c, a) in formula['lists'].items()},
           'constant': list(formula.get('constant', [])), 'items': []}
    for it in formula['items']:
        it = dict(it)
        if it['cmd'] in ('CONTR', 'BC'):
            it['cnt_c'] = _occ_out(it['cnt_c'])
            it['cnt_a'] = _occ_out(it['cnt_a'])
            if 'vertices' in it:
                it['vertices'] = _vertices_io(it['vertices'], _occ_out)
        elif it['cmd'] in ('REORDER', 'ADD_REO'):
            it['shift_c'] = _occ_out(it['shift_c'])
            it['shift_a'] = _occ_out(it['shift_a'])
        out['items'].append(it)
    with open(path, 'w') as f:
        json.dump(out, f, indent=1)


def load(path):
    """Read a formula written by dump (or by hand in the same form)."""
    with open(path) as f:
        d = json.load(f)
    lists = {n: (_occ_in(v['C']), _occ_in(v['A'])) for n, v in d['lists'].items()}
    items = []
    for it in d['items']:
        if it['cmd'] not in _COMMANDS:
            raise ValueError(f'unknown formula command {it["cmd"]!r}')
        if it['cmd'] in ('CONTR', 'BC'):
            v = it.get('vertices')
            tra = it.get('tra')
            it = dict(contr(it['target'], it['op1'], it['op2'], _occ_in(it['cnt_c']), _occ_in(it['cnt_a']),
                            it['factor']), cmd=it['cmd'])
            if v is not None:
                it['vertices'] = _vertices_io(v, _occ_in)
            if tra is not None:
                it['tra'] = [bool(x) for x in tra]
        elif it['cmd'] in ('REORDER', 'ADD_REO'):
            it = reorder(it['target'], it['source'], it['from'], it['to'], _occ_in(it['shift_c']),
                         _occ_in(it['shift_a']), it.get('factor', 1.0), it.get('sign', 1),
                         add=it['cmd'] == 'ADD_REO')
        elif it['cmd'] == 'ADD':
            it = add(it['target'], it['source'], it.get('factor', 1.0))
        items.append(it)
    for it in items:
        for key in ('target', 'op1', 'op2', 'source'):
            if key in it and it[key] not in lists:
                raise ValueError(f'formula item refers to undefined list {it[key]!r}')
    return {'lists': lists, 'constant': list(d.get('constant', [])), 'items': items}


def schedule(items):
    """Levels of a formula: a list of dicts {"init": [INIT items],
    "add": [ADD items], "contr": [CONTR items], "reo": [ADD_REO items],
    "sym": [list names]} in execution order (the INITs and ADDs of a level
    run in one launch before its contractions, its reorders after them and
    the projections last; BC items count as INIT then CONTR, REORDER items
    as INIT then ADD_REO, NEW and DELETE do no work).

    A contraction reads its operands as they are after its level's
    initialisation launch, so with items taken in list order:
      read after write   an operand's earlier writes by contractions (and
                         reorders) are in earlier levels, its INITs and
                         ADDs in the same or earlier levels;
      write after read   the target's earlier readers are in earlier
                         levels (the writer adds into the target while the
                         readers of its own level run);
      after an INIT      the target's last INIT is in the same or an earlier
                         level.
    Contributions to one list commute and need no order among themselves.
    An ADD_REO is placed like a contraction with the one operand `source`.
    An INIT follows the earlier reads and writes of its target and the
    writes of its source.  An ADD reads its source like a contraction and
    runs in the launch of the level's INITs, so it follows its target's
    last INIT by a level."""
    return _place(items)[0]


def _place(items):
    """schedule() and the level of every item of expand(items) (None for
    SYMMETRISE, NEW and DELETE).

    An INIT to zero is placed late: in the level of the first item that
    writes its list (a CONTR shares that level, the INITs running first;
    an ADD runs in the INIT launch, so the INIT goes one level earlier),
    or, when a read or another INIT of the list comes first, as early as
    its own constraints allow.  A list then holds data only from the level
    where it is first needed (lifetimes)."""
    items = expand(items)
    last_write, last_read, last_init = {}, {}, {}     # list -> level
    last_late = {}        # list -> level of its last write by a contraction or reorder
    inits, contrs, adds, reos = {}, {}, {}, {}
    where = [None] * len(items)
    pending = {}          # list -> (item index, earliest level) of a deferred INIT

    def put_init(t, lv):
        k, _ = pending.pop(t)
        inits.setdefault(lv, []).append(items[k])
        where[k] = lv
        last_write[t] = max(last_write.get(t, -1), lv)
        last_init[t] = lv

    def flush(n):
        if n in pending:
            put_init(n, pending[n][1])

    for k, it in enumerate(items):
        t = it['target']
        if it['cmd'] in ('SYMMETRISE', 'NEW', 'DELETE'):
            continue
        if it['cmd'] == 'ADD':
            src = it['source']
            flush(src)
            lv = max(last_write.get(src, -1) + 1, last_read.get(t, -1) + 1, last_init.get(t, -1) + 1)
            if t in pending:
                put_init(t, max(pending[t][1], lv - 1))
                lv = max(lv, last_init[t] + 1)
            adds.setdefault(lv, []).append(it)
            last_write[t] = max(last_write.get(t, -1), lv)
            last_read[src] = max(last_read.get(src, -1), lv)
            where[k] = lv
            continue
        if it['cmd'] == 'INIT':
            src = it.get('source')
            flush(t)
            if src:
                flush(src)
            lv = max(last_read.get(t, -1) + 1, last_write.get(t, -1) + 1,
                     last_write.get(src, -1) + 1 if src else 0)
            if not src:
                pending[t] = (k, lv)
                continue
            inits.setdefault(lv, []).append(it)
            last_write[t] = max(last_write.get(t, -1), lv)
            last_init[t] = lv
            last_read[src] = max(last_read.get(src, -1), lv)
            where[k] = lv
            continue
        # CONTR (operands op1, op2) and ADD_REO (operand source): a level's
        # reorders run after its contractions, and like them read lists
        # that no item of their level writes
        ops = (it['source'],) if it['cmd'] == 'ADD_REO' else (it['op1'], it['op2'])
        for op in ops:
            flush(op)
        # an operand's writes by contractions and reorders lie in earlier
        # levels; its INIT and ADD writes may share the level (the
        # initialisation launch runs before the contractions)
        lv = max([max(last_late.get(op, -1) + 1, last_write.get(op, -1)) for op in ops] +
                 [last_read.get(t, -1) + 1, last_init.get(t, 0)])
        if t in pending:
            put_init(t, max(pending[t][1], lv))
            lv = max(lv, last_init[t])
        (reos if it['cmd'] == 'ADD_REO' else contrs).setdefault(lv, []).append(it)
        last_write[t] = max(last_write.get(t, -1), lv)
        last_late[t] = max(last_late.get(t, -1), lv)
        for op in ops:
            last_read[op] = max(last_read.get(op, -1), lv)
        where[k] = lv
    for t in list(pending):
        flush(t)
    nlev = max(list(contrs) + list(inits) + list(adds) + list(reos) + [-1]) + 1
    levels = [{'init': inits.get(l, []), 'add': adds.get(l, []), 'contr': contrs.get(l, []),
               'reo': reos.get(l, []), 'sym': []} for l in range(nlev)]
    writes = [{c['target'] for c in lv['contr'] + lv['add'] + lv['reo']} for lv in levels]
    reads = [{c['op1'] for c in lv['contr']} | {c['op2'] for c in lv['contr']} |
             {a['source'] for a in lv['add'] + lv['reo']} |
             {i['source'] for i in lv['init'] if 'source' in i}
             for lv in levels]
    for li, lv in enumerate(levels):
        for res in dict.fromkeys(c['target'] for c in lv['contr'] + lv['add'] + lv['reo']):
            nxt = next((l for l in range(li + 1, nlev) if res in writes[l]), None)
            last = nlev - 1 if nxt is None else nxt
            if nxt is None or any(res in reads[l] for l in range(li + 1, last + 1)):
                lv['sym'].append(res)
    return levels, items, where


def lifetimes(formula, keep=()):
    """The levels [first, last] in which each intermediate holds data: a
    list (not constant, not in `keep`) whose first item, in formula order,
    writes it without reading it -- an INIT, or a BC -- is defined from that
    item's level to the level of its last item (read or write; its
    projection runs at the end of a level that writes it).  Lists read
    before any such write carry data between evaluations and have no
    lifetime.  NEW and DELETE items, where given, must enclose every item
    of their list (ValueError otherwise)."""
    levels, items, where = _place(formula['items'])
    fixed = set(formula.get('constant', ())) | set(keep)
    first, last, born = {}, {}, {}
    opened, closed = {}, {}
    for k, (it, lv) in enumerate(zip(items, where)):
        t = it['target']
        if it['cmd'] == 'NEW':
            opened[t] = k
            continue
        if it['cmd'] == 'DELETE':
            closed[t] = k
            continue
        if lv is None:
            continue
        used = [t] + [it[x] for x in ('op1', 'op2', 'source') if x in it]
        for n in used:
            if n in closed and closed[n] < k:
                raise ValueError(f'formula item {k} uses {n} after its DELETE')
            if n in opened and opened[n] > k:
                raise ValueError(f'formula item {k} uses {n} before its NEW')
            if n not in born:
                born[n] = n == t and it['cmd'] == 'INIT' and it.get('source') != t and \
                    n not in used[1:]
            first[n] = min(first.get(n, lv), lv)
            last[n] = max(last.get(n, lv), lv)
    return {n: (first[n], last[n]) for n in first if born.get(n) and n not in fixed}


def pool_layout(spans, sizes, align=64):
    """Offsets (elements) of lists in one shared pool such that lists
    whose spans (level intervals, from lifetimes) meet never overlap in
    memory: first fit, largest list first.  Returns ({name: offset},
    pool length)."""
    placed = []          # (offset, end, first, last)
    off = {}
    for n in sorted(spans, key=lambda x: (-sizes[x], x)):
        f, l = spans[n]
        need = max(1, int(sizes[n]))
        busy = sorted((o, e) for o, e, f2, l2 in placed if not (l2 < f or l < f2))
        at = 0
        for o, e in busy:
            if at + need <= o:
                break
            at = max(at, (e + align - 1) // align * align)
        off[n] = at
        placed.append((at, at + need, f, l))
    return off, max([e for _, e, _, _ in placed] + [0])


def delta_name(n):
    return 'd' + n


def linearize(formula, varying):
    """The linearisation of a formula about a fixed point of its lists: for
    every list X that depends on the `varying` lists, a list dX with
      INIT X = 0           ->  dX = 0
      INIT X = f * S       ->  dX = f * dS          (0 when S is fixed)
      ADD X += f * S       ->  dX += f * dS         (nothing when S is fixed)
      ADD_REO X += f R(S)  ->  dX += f R(dS)        (R: the same reorder)
      CONTR X += f A . B   ->  dX += f dA . B + f A . dB
    (terms of fixed operands dropped), so that with the base lists holding
    the formula's values at t, the delta lists hold the directional
    derivative of every list along dt (the varying lists' deltas).  Returns
    (items over base and delta lists, the names of the delta lists).  The
    base lists are read only; the items schedule like any formula.

    The base value of a list is its value after the whole formula, so every
    read of a list must see that value: a formula that writes (or
    initialises) a list after an item has read it is refused
    (ValueError)."""
    fitems = [it for it in expand(formula['items']) if it['cmd'] not in ('NEW', 'DELETE', 'SYMMETRISE')]
    read = set()
    for it in fitems:
        if it['target'] in read:
            raise ValueError(f'linearize: {it["target"]} is written after it is read')
        if it['cmd'] == 'CONTR':
            read.update((it['op1'], it['op2']))
        elif it.get('source'):
            read.add(it['source'])
    varies = set(varying)
    for it in fitems:                  # lists written from varying ones
        if it['cmd'] == 'CONTR' and (it['op1'] in varies or it['op2'] in varies):
            varies.add(it['target'])
        elif it['cmd'] in ('INIT', 'ADD', 'ADD_REO') and it.get('source') in varies:
            varies.add(it['target'])
    items = []
    for it in fitems:
        t = it['target']
        if t not in varies or t in varying:
            continue
        if it['cmd'] in ('ADD', 'ADD_REO'):
            if it['source'] in varies:
                items.append(dict(it, target=delta_name(t), source=delta_name(it['source'])))
        elif it['cmd'] == 'INIT':
            if it.get('source') in varies:
                items.append(init(delta_name(t), delta_name(it['source']), it['factor']))
            else:
                items.append(init(delta_name(t)))
        elif it['cmd'] == 'CONTR':
            if it['op1'] in varies:
                items.append(dict(it, target=delta_name(t), op1=delta_name(it['op1'])))
            if it['op2'] in varies:
                items.append(dict(it, target=delta_name(t), op2=delta_name(it['op2'])))
    return items, sorted(delta_name(n) for n in varies)
