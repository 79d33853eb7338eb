"""GeCCo's binary formula lists: reader, writer, and the translation to and
from the engine's formula items (formula.py).

File format, restated from formula/write_form_list.f:33-45,
formula/read_form_list.f:40-60, formula/wr_formula.f:24-45,
formula/rd_formula.f:24-48 and the record kernels
formula/rw_bcontr_kernel.f:119-146, formula/rw_reo_kernel.f:90-108,
formula/rw_opdef_kernel.f:76-85 and formula/rw_contr_kernel.f:96-178.  It is
a Fortran sequential unformatted file as GeCCo writes it when built with
8-byte default INTEGER and LOGICAL (the -i8 / -fdefault-integer-8 flags of
config/*.conf): every record is framed by 4-byte little-endian byte counts
before and after it (a negative leading count marks a record continued in
the next subrecord); INTEGER and LOGICAL fields take 8 bytes (a LOGICAL is
true when its lowest bit is set), REAL(8) 8 bytes, CHARACTER one byte per
character, arrays in column-major order.

  title        len_trim(title), title
  item         command, target                     (one record)
               then, by command (include/def_formula_item.h:29-46):
    0 [INIT TARGET], 1, 5 [SYMMETRISE], 12, 15, 16   nothing more
    2 [NEW INTERMEDIATE]   opdef: nblk, njoined, ngas, nspin /
                           len, name, tra, len, parent1, tra1, len,
                           parent2, tra2, ihpvca_occ(4,2,nblk*nj),
                           igasca_restr(2,ngas,2,2,nspin,nblk*nj), incore
    3 [DEL INTERMEDIATE]   label (32 characters)
    4 [ADD] (unfactorised)  fac, n, buffer(n) (INTEGER(2): iblk_res
                           (negative: dagger), nvtx, nsupvtx, narc, nxarc,
                           nfac, per vertex (op index, negative: dagger;
                           block), supervertex of each vertex, per arc and
                           per external arc (link1, link2, occ(4,2)), per
                           factor 5 entries)
    6 [REORDER]            reo
    7, 8, 10, 14           bcontr
    9, 11, 13              bcontr, reo
  bcontr       n_operands, n_cnt, nj_res, nj_op1, nj_op2, ngas, nspin,
               len_m1, len_m2, len_m12, len_m21, len_idxinf /
               fact, fact_itf, len, label_res, len, label_op1, len,
               label_op2, iblk_res, iblk_op1, iblk_op2, tra_res, tra_op1,
               tra_op2, occ_res(4,2,nj_res), rst_res, occ_op1, rst_op1,
               perm(4,2), itf_index_info(len_idxinf) /
               [2 operands:] occ_op2, rst_op2 /
               [n_cnt > 0:] occ_ex1, occ_cnt, rst_ex1, rst_cnt, merge_op1,
               merge_op1op2, merge_op2op1 /
               [n_cnt > 0 and 2 operands:] occ_ex2, rst_ex2, merge_op2
  reo          nj_out, nj_in, nreo, nreo_i0, ngas, nspin, len_m1, len_m1i,
               len_m2, len_m2i /
               len, label_in, len, label_out, tra_in, tra_out, sign,
               iblk_in, iblk_out, occ_opin, rst_opin, occ_opout, rst_opout,
               occ_op0(4,2,nj_in), occ_shift(4,2,nreo),
               from_to(2,nreo+nreo_i0), merge_stp1, merge_stp1inv,
               merge_stp2, merge_stp2inv
  end          command -1

(rst arrays are (2, ngas, 2, 2, nspin, nj); the occupation index runs over
the spaces H, P, V, X, then C / A.)  The title record is skipped on reading,
as read_form_list does, and reading stops at command -1 or the end of the
file.

Translation (to_formula): every operator block (label, iblk) becomes one
list of the engine, named label when the file uses a single block of that
label and label:iblk otherwise, with the occupation of its vertices as the
items give it; lists used transposed (tra flags) are declared with C and A
exchanged, as stored.  The commands map as the scheduler executes them
(contract/frm_sched2.f:175-300):
  [INIT TARGET] t          INIT (zero) of every list of operator t
  [NEW INTERMEDIATE]       NEW and INIT (zero) of each block's list
  [DEL INTERMEDIATE]       DELETE of each block's list
  [ADD_BC] / [BC]          CONTR / BC: between one-vertex lists with one
                           contraction vertex the one-vertex item (lines of
                           CNT on op1's C and A side), else the vertex data
                           (EX1, EX2, CNT occupations and the four
                           short-form merge maps, contr_op1op2's arguments,
                           contract/fs_contr_drv.f:189-236, 272-288); the
                           dagger flags as "tra"
  [ADD_BC_REO] / [BC_REO]  the contraction into a temporary list of the
                           reorder's input occupation, then ADD_REO into the
                           result (after INIT for [BC_REO]); the temporary is
                           enclosed in NEW / DELETE
  [ADD_INTM] / [CP_INTM]   ADD / INIT with source and factor
  [REORDER] / [ADD_REO]    REORDER / ADD_REO (one vertex move; factor of the
                           bcontr record for [ADD_REO])
  [SYMMETRISE] t           SYMMETRISE of every list of operator t
Unfactorised [ADD] items, inversions and reorders of several steps or with
vertex merges are read but not translated (FormulaListError).  from_formula
writes the engine's formula items as such a file (one block per label).
"""
import struct

import numpy as np

import formula as _formula

END, INIT_TARGET, UPDATE_TARGET, NEW_INTERMEDIATE, DEL_INTERMEDIATE = -1, 0, 1, 2, 3
ADD_CONTRIBUTION, SYMMETRISE, REORDER, ADD_INTM, ADD_BC, ADD_BC_REO = 4, 5, 6, 7, 8, 9
BC, BC_REO, INTERNAL, ADD_REO, CP_INTM, PSINV, INV = 10, 11, 12, 13, 14, 15, 16
COMMAND_NAMES = {END: 'END', INIT_TARGET: 'INIT TARGET', UPDATE_TARGET: 'UPDATE TARGET',
                 NEW_INTERMEDIATE: 'NEW INTERMEDIATE', DEL_INTERMEDIATE: 'DEL INTERMEDIATE',
                 ADD_CONTRIBUTION: 'ADD', SYMMETRISE: 'SYMMETRISE', REORDER: 'REORDER',
                 ADD_INTM: 'ADD_INTM', ADD_BC: 'ADD_BC', ADD_BC_REO: 'ADD_BC_REO', BC: 'BC',
                 BC_REO: 'BC_REO', INTERNAL: 'INTERNAL', ADD_REO: 'ADD_REO', CP_INTM: 'CP_INTM',
                 PSINV: 'PSINV', INV: 'INV'}
NGASTP = 4                    # H, P, V, X (include/opdim.h)
LEN_OPNAME = 32               # include/def_operator.h
_SPACES = ('H', 'P', 'V', 'X')


class FormulaListError(ValueError):
    pass


# ------------------------------------------------------------------ records
def split_records(data):
    """Records of a Fortran sequential unformatted file (bytes)."""
    out, pos, cur = [], 0, b''
    while pos < len(data):
        if pos + 4 > len(data):
            raise FormulaListError(f'truncated record marker at byte {pos}')
        head = struct.unpack_from('<i', data, pos)[0]
        n = abs(head)
        if pos + 8 + n > len(data):
            raise FormulaListError(f'truncated record at byte {pos}')
        cur += data[pos + 4:pos + 4 + n]
        tail = struct.unpack_from('<i', data, pos + 4 + n)[0]
        if abs(tail) != n:
            raise FormulaListError(f'record markers {head} and {tail} disagree at byte {pos}')
        pos += 8 + n
        if head >= 0:
            out.append(cur)
            cur = b''
    if cur:
        raise FormulaListError('file ends inside a continued record')
    return out


def join_records(records):
    return b''.join(struct.pack('<i', len(r)) + r + struct.pack('<i', len(r)) for r in records)


class _Reader:
    def __init__(self, rec, what):
        self.rec, self.pos, self.what = rec, 0, what

    def _take(self, n):
        if self.pos + n > len(self.rec):
            raise FormulaListError(f'{self.what}: record shorter than its fields')
        b = self.rec[self.pos:self.pos + n]
        self.pos += n
        return b

    def i8(self):
        return struct.unpack('<q', self._take(8))[0]

    def f8(self):
        return struct.unpack('<d', self._take(8))[0]

    def l8(self):
        return bool(self.i8() & 1)

    def chars(self, n):
        return self._take(n).decode('latin-1')

    def label(self):
        return self.chars(self.i8())

    def i8s(self, shape):
        """INTEGER(8) array of Fortran shape `shape`, returned with the
        index order reversed (C order, last Fortran index first)."""
        n = int(np.prod(shape)) if len(shape) else 1
        a = np.frombuffer(self._take(8 * n), dtype='<i8').astype(np.int64)
        return a.reshape(tuple(reversed(shape)))

    def i2s(self, n):
        return np.frombuffer(self._take(2 * n), dtype='<i2').astype(np.int64)

    def done(self):
        if self.pos != len(self.rec):
            raise FormulaListError(f'{self.what}: {len(self.rec) - self.pos} bytes left in the record')


class _Writer:
    def __init__(self):
        self.parts = []

    def i8(self, *v):
        for x in v:
            self.parts.append(struct.pack('<q', int(x)))
        return self

    def f8(self, x):
        self.parts.append(struct.pack('<d', float(x)))
        return self

    def l8(self, *v):
        return self.i8(*[1 if x else 0 for x in v])

    def label(self, s):
        b = s.encode('latin-1')
        self.i8(len(b))
        self.parts.append(b)
        return self

    def chars(self, s, n):
        self.parts.append(s.encode('latin-1').ljust(n, b' ')[:n])
        return self

    def arr(self, a):
        """C-ordered array (Fortran index order reversed) as INTEGER(8)."""
        self.parts.append(np.ascontiguousarray(np.asarray(a, dtype='<i8')).tobytes())
        return self

    def bytes(self):
        return b''.join(self.parts)


def _occ_shape(nj):
    return (NGASTP, 2, nj)


def _rst_shape(ngas, nspin, nj):
    return (2, ngas, 2, 2, nspin, nj)


# ------------------------------------------------------------------ kernels
def _read_bcontr(recs):
    r = _Reader(recs.pop(0), 'bcontr dimensions')
    (n_op, n_cnt, nj_res, nj_op1, nj_op2, ngas, nspin, len_m1, len_m2, len_m12, len_m21,
     len_idx) = [r.i8() for _ in range(12)]
    r.done()
    b = dict(n_operands=n_op, n_cnt=n_cnt, nj_res=nj_res, nj_op1=nj_op1, nj_op2=nj_op2,
             ngas=ngas, nspin=nspin)
    r = _Reader(recs.pop(0), 'bcontr')
    b['fact'], b['fact_itf'] = r.f8(), r.f8()
    b['label_res'], b['label_op1'], b['label_op2'] = r.label(), r.label(), r.label()
    b['iblk_res'], b['iblk_op1'], b['iblk_op2'] = r.i8(), r.i8(), r.i8()
    b['tra_res'], b['tra_op1'], b['tra_op2'] = r.l8(), r.l8(), r.l8()
    b['occ_res'] = r.i8s(_occ_shape(nj_res))
    b['rst_res'] = r.i8s(_rst_shape(ngas, nspin, nj_res))
    b['occ_op1'] = r.i8s(_occ_shape(nj_op1))
    b['rst_op1'] = r.i8s(_rst_shape(ngas, nspin, nj_op1))
    b['perm'] = r.i8s((NGASTP, 2)) & 1
    b['itf_index_info'] = r.i8s((len_idx,))
    r.done()
    if n_op == 2:
        r = _Reader(recs.pop(0), 'bcontr operand 2')
        b['occ_op2'] = r.i8s(_occ_shape(nj_op2))
        b['rst_op2'] = r.i8s(_rst_shape(ngas, nspin, nj_op2))
        r.done()
    if n_cnt > 0:
        r = _Reader(recs.pop(0), 'bcontr contraction')
        b['occ_ex1'] = r.i8s(_occ_shape(nj_op1))
        b['occ_cnt'] = r.i8s(_occ_shape(n_cnt))
        b['rst_ex1'] = r.i8s(_rst_shape(ngas, nspin, nj_op1))
        b['rst_cnt'] = r.i8s(_rst_shape(ngas, nspin, n_cnt))
        b['merge_op1'] = r.i8s((len_m1,))
        b['merge_op1op2'] = r.i8s((len_m12,))
        b['merge_op2op1'] = r.i8s((len_m21,))
        r.done()
        if n_op == 2:
            r = _Reader(recs.pop(0), 'bcontr contraction, operand 2')
            b['occ_ex2'] = r.i8s(_occ_shape(nj_op2))
            b['rst_ex2'] = r.i8s(_rst_shape(ngas, nspin, nj_op2))
            b['merge_op2'] = r.i8s((len_m2,))
            r.done()
    return b


def _write_bcontr(b):
    n_op, n_cnt = b['n_operands'], b['n_cnt']
    nj_res, nj_op1, nj_op2 = b['nj_res'], b['nj_op1'], b['nj_op2']
    ngas, nspin = b['ngas'], b['nspin']
    lens = [len(b.get(k, ())) if n_cnt > 0 else 0
            for k in ('merge_op1', 'merge_op2', 'merge_op1op2', 'merge_op2op1')]
    if n_op != 2:
        lens[1] = 0
    idx = np.asarray(b.get('itf_index_info', np.zeros(3, np.int64)))
    recs = [_Writer().i8(n_op, n_cnt, nj_res, nj_op1, nj_op2, ngas, nspin, *lens, len(idx)).bytes()]
    w = _Writer().f8(b['fact']).f8(b.get('fact_itf', b['fact']))
    w.label(b['label_res']).label(b['label_op1']).label(b.get('label_op2', ''))
    w.i8(b['iblk_res'], b['iblk_op1'], b.get('iblk_op2', 0))
    w.l8(b['tra_res'], b['tra_op1'], b.get('tra_op2', False))
    w.arr(b['occ_res']).arr(b['rst_res']).arr(b['occ_op1']).arr(b['rst_op1'])
    w.arr(b.get('perm', np.zeros((2, NGASTP), np.int64))).arr(idx)
    recs.append(w.bytes())
    if n_op == 2:
        recs.append(_Writer().arr(b['occ_op2']).arr(b['rst_op2']).bytes())
    if n_cnt > 0:
        recs.append(_Writer().arr(b['occ_ex1']).arr(b['occ_cnt']).arr(b['rst_ex1']).arr(b['rst_cnt'])
                    .arr(b['merge_op1']).arr(b['merge_op1op2']).arr(b['merge_op2op1']).bytes())
        if n_op == 2:
            recs.append(_Writer().arr(b['occ_ex2']).arr(b['rst_ex2']).arr(b['merge_op2']).bytes())
    return recs


def _read_reo(recs):
    r = _Reader(recs.pop(0), 'reorder dimensions')
    nj_out, nj_in, nreo, nreo_i0, ngas, nspin, len_m1, len_m1i, len_m2, len_m2i = \
        [r.i8() for _ in range(10)]
    r.done()
    o = dict(nj_out=nj_out, nj_in=nj_in, nreo=nreo, nreo_i0=nreo_i0, ngas=ngas, nspin=nspin)
    r = _Reader(recs.pop(0), 'reorder')
    o['label_in'], o['label_out'] = r.label(), r.label()
    o['tra_in'], o['tra_out'] = r.l8(), r.l8()
    o['sign'] = r.i8()
    o['iblk_in'], o['iblk_out'] = r.i8(), r.i8()
    o['occ_opin'] = r.i8s(_occ_shape(nj_in))
    o['rst_opin'] = r.i8s(_rst_shape(ngas, nspin, nj_in))
    o['occ_opout'] = r.i8s(_occ_shape(nj_out))
    o['rst_opout'] = r.i8s(_rst_shape(ngas, nspin, nj_out))
    o['occ_op0'] = r.i8s(_occ_shape(nj_in))
    o['occ_shift'] = r.i8s(_occ_shape(nreo))
    o['from_to'] = r.i8s((2, nreo + nreo_i0))
    o['merge_stp1'] = r.i8s((len_m1,))
    o['merge_stp1inv'] = r.i8s((len_m1i,))
    o['merge_stp2'] = r.i8s((len_m2,))
    o['merge_stp2inv'] = r.i8s((len_m2i,))
    r.done()
    return o


def _write_reo(o):
    ms = [np.asarray(o.get(k, np.zeros(0, np.int64))) for k in
          ('merge_stp1', 'merge_stp1inv', 'merge_stp2', 'merge_stp2inv')]
    recs = [_Writer().i8(o['nj_out'], o['nj_in'], o['nreo'], o['nreo_i0'], o['ngas'], o['nspin'],
                         *[len(m) for m in ms]).bytes()]
    w = _Writer().label(o['label_in']).label(o['label_out']).l8(o['tra_in'], o['tra_out'])
    w.i8(o['sign'], o['iblk_in'], o['iblk_out'])
    for k in ('occ_opin', 'rst_opin', 'occ_opout', 'rst_opout', 'occ_op0', 'occ_shift', 'from_to'):
        w.arr(o[k])
    for m in ms:
        w.arr(m)
    recs.append(w.bytes())
    return recs


def _read_opdef(recs):
    r = _Reader(recs.pop(0), 'operator dimensions')
    nblk, nj, ngas, nspin = [r.i8() for _ in range(4)]
    r.done()
    r = _Reader(recs.pop(0), 'operator')
    o = dict(nblk=nblk, njoined=nj, ngas=ngas, nspin=nspin)
    o['name'], o['tra'] = r.label(), r.l8()
    o['parent1'], o['tra1'] = r.label(), r.l8()
    o['parent2'], o['tra2'] = r.label(), r.l8()
    o['occ'] = r.i8s(_occ_shape(nblk * nj))
    o['rst'] = r.i8s(_rst_shape(ngas, nspin, nblk * nj))
    o['incore'] = r.i8()
    r.done()
    return o


def _write_opdef(o):
    recs = [_Writer().i8(o['nblk'], o['njoined'], o['ngas'], o['nspin']).bytes()]
    w = _Writer().label(o['name']).l8(o.get('tra', False))
    w.label(o.get('parent1', '')).l8(o.get('tra1', False))
    w.label(o.get('parent2', '')).l8(o.get('tra2', False))
    w.arr(o['occ']).arr(o['rst']).i8(o.get('incore', 0))
    recs.append(w.bytes())
    return recs


def _read_contr(rec):
    r = _Reader(rec, 'contraction')
    fac, n = r.f8(), r.i8()
    buf = r.i2s(n)
    r.done()
    if n < 6:
        raise FormulaListError('contraction record shorter than its header')
    c = dict(fac=fac, iblk_res=abs(int(buf[0])), dagger=bool(buf[0] < 0), nvtx=int(buf[1]),
             nsupvtx=int(buf[2]), narc=int(buf[3]), nxarc=int(buf[4]), nfac=int(buf[5]))
    need = 6 + 3 * c['nvtx'] + 10 * (c['narc'] + c['nxarc']) + 5 * c['nfac']
    if n != need:
        raise FormulaListError(f'contraction record holds {n} entries, its header asks for {need}')
    k = 6
    c['vertices'] = []
    for _ in range(c['nvtx']):
        c['vertices'].append(dict(idx_op=abs(int(buf[k])), dagger=bool(buf[k] < 0), iblk_op=int(buf[k + 1])))
        k += 2
    c['svertex'] = [int(x) for x in buf[k:k + c['nvtx']]]
    k += c['nvtx']
    for key, cnt in (('arcs', c['narc']), ('xarcs', c['nxarc'])):
        c[key] = []
        for _ in range(cnt):
            c[key].append(dict(link=(int(buf[k]), int(buf[k + 1])),
                               occ=buf[k + 2:k + 10].reshape(2, NGASTP).copy()))
            k += 10
    c['inffac'] = [buf[k + 5 * i:k + 5 * i + 5].copy() for i in range(c['nfac'])]
    return c


def _write_contr(c):
    buf = [(-1 if c.get('dagger') else 1) * c['iblk_res'], len(c['vertices']), c['nsupvtx'],
           len(c['arcs']), len(c['xarcs']), len(c['inffac'])]
    for v in c['vertices']:
        buf += [(-1 if v['dagger'] else 1) * v['idx_op'], v['iblk_op']]
    buf += list(c['svertex'])
    for a in list(c['arcs']) + list(c['xarcs']):
        buf += list(a['link']) + [int(x) for x in np.asarray(a['occ']).ravel()]
    for f in c['inffac']:
        buf += [int(x) for x in f]
    b = np.asarray(buf, dtype='<i2')
    if not np.array_equal(b.astype(np.int64), np.asarray(buf, dtype=np.int64)):
        raise FormulaListError('contraction entries do not fit INTEGER(2)')
    return [_Writer().f8(c['fac']).i8(len(buf)).bytes() + b.tobytes()]


# ------------------------------------------------------------------ files
def parse(data):
    """GeCCo formula-list bytes -> {'title', 'items'}; each item has
    'command', 'target' and, by command, 'bcontr', 'reo', 'opdef', 'label'
    or 'contr' (arrays C-ordered: occupations [vertex][C/A][space])."""
    recs = split_records(data)
    if not recs:
        raise FormulaListError('empty formula file')
    r = _Reader(recs.pop(0), 'title')
    n = r.i8()
    title = r.chars(len(r.rec) - 8)          # the whole CHARACTER variable
    items = []
    while recs:
        r = _Reader(recs.pop(0), 'command')
        cmd, target = r.i8(), r.i8()
        r.done()
        if cmd == END:
            break
        if cmd not in COMMAND_NAMES:
            raise FormulaListError(f'unknown formula command {cmd}')
        it = dict(command=cmd, target=target)
        try:
            if cmd == ADD_CONTRIBUTION:
                it['contr'] = _read_contr(recs.pop(0))
            elif cmd == DEL_INTERMEDIATE:
                it['label'] = recs.pop(0).decode('latin-1').rstrip()
            elif cmd == NEW_INTERMEDIATE:
                it['opdef'] = _read_opdef(recs)
            elif cmd == REORDER:
                it['reo'] = _read_reo(recs)
            elif cmd in (ADD_BC_REO, BC_REO, ADD_REO):
                it['bcontr'] = _read_bcontr(recs)
                it['reo'] = _read_reo(recs)
            elif cmd in (ADD_INTM, CP_INTM, ADD_BC, BC):
                it['bcontr'] = _read_bcontr(recs)
        except IndexError:
            raise FormulaListError(f'formula file ends inside a {COMMAND_NAMES[cmd]} item') from None
        items.append(it)
    return dict(title=title[:n] if 0 <= n <= len(title) else title, title_field=title, items=items)


def read(path):
    with open(path, 'rb') as f:
        return parse(f.read())


def serialize(title, items):
    """Title and items (as parse returns them) -> formula-list bytes (the
    [END] item is appended).  The title record holds len_trim(title) and
    the whole string, trailing blanks included (parse's title_field)."""
    tb = title.encode('latin-1')
    recs = [struct.pack('<q', len(tb.rstrip(b' '))) + tb]
    for it in items:
        cmd = it['command']
        recs.append(_Writer().i8(cmd, it.get('target', 0)).bytes())
        if cmd == ADD_CONTRIBUTION:
            recs += _write_contr(it['contr'])
        elif cmd == DEL_INTERMEDIATE:
            recs.append(it['label'].encode('latin-1').ljust(LEN_OPNAME, b' ')[:LEN_OPNAME])
        elif cmd == NEW_INTERMEDIATE:
            recs += _write_opdef(it['opdef'])
        elif cmd == REORDER:
            recs += _write_reo(it['reo'])
        elif cmd in (ADD_BC_REO, BC_REO, ADD_REO):
            recs += _write_bcontr(it['bcontr']) + _write_reo(it['reo'])
        elif cmd in (ADD_INTM, CP_INTM, ADD_BC, BC):
            recs += _write_bcontr(it['bcontr'])
    recs.append(_Writer().i8(END, 0).bytes())
    return join_records(recs)


def write(path, title, items):
    with open(path, 'wb') as f:
        f.write(serialize(title, items))


# ------------------------------------------------------------------ translation
def _occ_dict(row):
    """One (C or A) row of an occupation [space] -> the engine's dict over
    its space codes (H 1, P 2)."""
    if any(int(row[k]) for k in (2, 3)):
        raise FormulaListError('valence and external spaces are not supported by the engine')
    return {s: int(row[k]) for k, s in enumerate((_formula.H, _formula.P)) if int(row[k])}


def _vertices(occ):
    """occupation [vertex][C/A][space] -> [(C, A)] per vertex."""
    return [(_occ_dict(v[0]), _occ_dict(v[1])) for v in occ]


def _decl(occ, tra):
    """The list declaration {'C', 'A'} of an operand used with occupation
    occ: stored as the adjoint when the item uses it transposed."""
    vs = _vertices(occ)
    if tra:
        if len(vs) != 1:
            raise FormulaListError('transposed operands of several vertices are not supported')
        vs = [(vs[0][1], vs[0][0])]
    if len(vs) == 1:
        return vs[0]
    return ([c for c, _ in vs], [a for _, a in vs])


def to_formula(parsed, constant=(), names=None):
    """The engine's formula {'lists', 'constant', 'items'} (in memory, as
    formula.load returns it) of a parsed GeCCo formula list.  constant: operator labels whose lists hold fixed
    data (integrals); names(label, iblk, single) -> list name overrides the
    naming rule."""
    items = parsed['items']
    blocks = {}                                    # label -> set of iblk
    index_label = {}                               # operator index -> label

    def see(label, iblk):
        blocks.setdefault(label, set()).add(int(iblk))
    for it in items:
        b, o, d = it.get('bcontr'), it.get('reo'), it.get('opdef')
        if b is not None:
            see(b['label_res'], b['iblk_res'])
            see(b['label_op1'], b['iblk_op1'])
            if b['n_operands'] == 2:
                see(b['label_op2'], b['iblk_op2'])
            if it['command'] not in (ADD_BC_REO, BC_REO, ADD_REO):
                index_label.setdefault(it['target'], b['label_res'])
        if o is not None:
            see(o['label_in'], o['iblk_in'])
            see(o['label_out'], o['iblk_out'])
            index_label.setdefault(it['target'], o['label_out'])
        if d is not None:
            for k in range(1, d['nblk'] + 1):
                see(d['name'], k)

    def lname(label, iblk):
        single = len(blocks.get(label, ())) == 1
        if names is not None:
            return names(label, iblk, single)
        return label if single else f'{label}:{iblk}'

    def label_lists(label):
        return [lname(label, k) for k in sorted(blocks.get(label, ()))]

    def label_of(it):
        t = it['target']
        if t not in index_label:
            raise FormulaListError(f'{COMMAND_NAMES[it["command"]]} of operator {t}, which no '
                                   f'item of the file names')
        return index_label[t]

    lists = {}

    def declare(name, decl):
        if name in lists and lists[name] != decl:
            raise FormulaListError(f'list {name} is used with two occupations: {lists[name]} and {decl}')
        lists[name] = decl

    out = []
    ntmp = 0

    def contraction(b, target, cmd):
        n1 = lname(b['label_op1'], b['iblk_op1'])
        declare(n1, _decl(b['occ_op1'], b['tra_op1']))
        if b['n_operands'] != 2:
            raise FormulaListError('a contraction item with one operand')
        n2 = lname(b['label_op2'], b['iblk_op2'])
        declare(n2, _decl(b['occ_op2'], b['tra_op2']))
        tra = [bool(b['tra_op1']), bool(b['tra_op2']), bool(b['tra_res'])]
        if b['nj_op1'] == b['nj_op2'] == b['nj_res'] == 1 and b['n_cnt'] in (0, 1):
            # one vertex everywhere: the one-vertex form, lines contracted
            # on op1's creator (cnt_c) and annihilator (cnt_a) side; op1 =
            # EX1 + CNT and op2 = EX2 + CNT^+ as the item uses them (no
            # contraction vertex: a product without contraction lines)
            cc, ca = {}, {}
            if b['n_cnt'] == 1:
                (c1, a1), = _vertices(b['occ_op1'])
                (c2, a2), = _vertices(b['occ_op2'])
                (cc, ca), = _vertices(b['occ_cnt'])
                (e1c, e1a), = _vertices(b['occ_ex1'])
                (e2c, e2a), = _vertices(b['occ_ex2'])
                def add(x, y):
                    return {k: x.get(k, 0) + y.get(k, 0) for k in set(x) | set(y) if x.get(k, 0) + y.get(k, 0)}
                if add(e1c, cc) != c1 or add(e1a, ca) != a1 or add(e2c, ca) != c2 or add(e2a, cc) != a2:
                    raise FormulaListError(f'contraction of {n1} and {n2}: EX/CNT occupations do not '
                                           f'add up to the operands')
            it = _formula.contr(target, n1, n2, cc, ca, b['fact'])
        elif b['n_cnt'] <= 0:
            raise FormulaListError('products of lists with several vertices and no contraction vertex '
                                   'carry no merge maps in the file and are not supported')
        else:
            ex1, ex2, cnt = _vertices(b['occ_ex1']), _vertices(b['occ_ex2']), _vertices(b['occ_cnt'])
            if any(tra):
                raise FormulaListError('transposed operands of several vertices are not supported')
            merge = [[int(x) for x in b[k]] for k in ('merge_op1', 'merge_op2', 'merge_op1op2',
                                                        'merge_op2op1')]
            it = _formula.contr_n(target, n1, n2, ex1, ex2, cnt, merge, b['fact'])
        if cmd == 'BC':
            it['cmd'] = 'BC'
        if any(tra):
            it['tra'] = tra
        return it

    def reo_args(o):
        if o['nreo'] != 1 or o['nreo_i0'] != 0:
            raise FormulaListError(f'reorders of {o["nreo"]} steps and {o["nreo_i0"]} vertex merges '
                                   f'are not supported (one vertex move only)')
        if o['tra_in'] or o['tra_out']:
            raise FormulaListError('transposed reorders are not supported')
        frm, to = (int(x) for x in o['from_to'][0])
        sc, sa = _vertices(o['occ_shift'])[0]
        return frm - 1, to - 1, sc, sa, int(o['sign'])

    for it in items:
        cmd = it['command']
        b, o = it.get('bcontr'), it.get('reo')
        if cmd == INIT_TARGET:
            out += [_formula.init(n) for n in label_lists(label_of(it))]
        elif cmd == UPDATE_TARGET:
            pass
        elif cmd == NEW_INTERMEDIATE:
            d = it['opdef']
            nj = d['njoined']
            for k in range(d['nblk']):
                n = lname(d['name'], k + 1)
                declare(n, _decl(d['occ'][k * nj:(k + 1) * nj], False))
                out += [_formula.new(n), _formula.init(n)]
        elif cmd == DEL_INTERMEDIATE:
            out += [_formula.delete(n) for n in label_lists(it['label'])]
        elif cmd in (ADD_BC, BC):
            t = lname(b['label_res'], b['iblk_res'])
            declare(t, _decl(b['occ_res'], b['tra_res']))
            out.append(contraction(b, t, 'CONTR' if cmd == ADD_BC else 'BC'))
        elif cmd in (ADD_BC_REO, BC_REO):
            t = lname(b['label_res'], b['iblk_res'])
            declare(t, _decl(o['occ_opout'], b['tra_res']))
            ntmp += 1
            tmp = f'_reo{ntmp}'
            declare(tmp, _decl(o['occ_opin'], False))
            frm, to, sc, sa, sign = reo_args(o)
            out += [_formula.new(tmp), _formula.init(tmp), contraction(b, tmp, 'CONTR')]
            if cmd == BC_REO:
                out.append(_formula.init(t))
            out += [_formula.reorder(t, tmp, frm, to, sc, sa, 1.0, sign, add=True), _formula.delete(tmp)]
        elif cmd in (ADD_INTM, CP_INTM):
            t = lname(b['label_res'], b['iblk_res'])
            s = lname(b['label_op1'], b['iblk_op1'])
            if b['tra_res'] != b['tra_op1']:
                raise FormulaListError('transposed additions are not supported')
            declare(t, _decl(b['occ_res'], b['tra_res']))
            declare(s, _decl(b['occ_op1'], b['tra_op1']))
            out.append(_formula.add(t, s, b['fact']) if cmd == ADD_INTM else _formula.init(t, s, b['fact']))
        elif cmd in (REORDER, ADD_REO):
            t, s = lname(o['label_out'], o['iblk_out']), lname(o['label_in'], o['iblk_in'])
            declare(t, _decl(o['occ_opout'], False))
            declare(s, _decl(o['occ_opin'], False))
            frm, to, sc, sa, sign = reo_args(o)
            fact = b['fact'] if cmd == ADD_REO else 1.0
            out.append(_formula.reorder(t, s, frm, to, sc, sa, fact, sign, add=cmd == ADD_REO))
        elif cmd == SYMMETRISE:
            out += [{'cmd': 'SYMMETRISE', 'target': n} for n in label_lists(label_of(it))]
        else:
            raise FormulaListError(f'{COMMAND_NAMES[cmd]} items are not translated to the engine')
    const = [n for lab in constant for n in label_lists(lab)]
    missing = [n for n in const if n not in lists]
    if missing:
        raise FormulaListError(f'constant lists {missing} are never used')
    return {'lists': lists, 'constant': const, 'items': out}


# ------------------------------------------------------------------ from the engine
def _occ_arr(vs):
    """[(C, A)] per vertex (dicts over 'H'/'P' or the space codes 1/2) ->
    occupation [vertex][C/A][space]."""
    a = np.zeros((len(vs), 2, NGASTP), dtype=np.int64)
    for v, (c, an) in enumerate(vs):
        for side, occ in ((0, c), (1, an)):
            for s, n in occ.items():
                k = {'H': 0, 'P': 1, 1: 0, 2: 1}[s]
                a[v, side, k] = int(n)
    return a


def _list_vertices(decl):
    c, a = decl
    if isinstance(c, list):
        return list(zip(c, a))
    return [(c, a)]


def _sub(x, y):
    out = dict(x)
    for k, n in y.items():
        out[k] = out.get(k, 0) - n
        if out[k] < 0:
            raise FormulaListError('contraction lines exceed the operand occupation')
        if out[k] == 0:
            del out[k]
    return out


def from_formula(formula, title='formula of the engine', ngas=1, nspin=1):
    """The engine's formula items as GeCCo formula-list items (parse's
    form; serialize writes them): every list is an operator of one block
    (operator index = order of first use, from 1).  One-vertex CONTR items
    get their EX1 / EX2 / CNT occupations and one-vertex merge maps."""
    lists = formula['lists']
    index = {}

    def idx(n):
        return index.setdefault(n, len(index) + 1)
    for n in lists:
        idx(n)

    def occ(n):
        return _occ_arr(_list_vertices(lists[n]))

    def rst(nj):
        return np.zeros(tuple(reversed(_rst_shape(ngas, nspin, nj))), dtype=np.int64)

    def used(n, tra):
        """Occupation of list n as an item uses it (the adjoint: C and A
        exchanged)."""
        vs = _list_vertices(lists[n])
        if tra:
            if len(vs) != 1:
                raise FormulaListError('transposed operands of several vertices are not supported')
            vs = [(vs[0][1], vs[0][0])]
        return vs

    def bcontr(res, op1, op2, fact, extra=None, tra=(False, False, False)):
        o1, o2, o_res = _occ_arr(used(op1, tra[0])), _occ_arr(used(op2, tra[1])) if op2 else None, \
            _occ_arr(used(res, tra[2]))
        b = dict(n_operands=2 if op2 else 1, n_cnt=0, nj_res=len(o_res), nj_op1=len(o1),
                 nj_op2=len(o2) if op2 else 0, ngas=ngas, nspin=nspin, fact=fact, fact_itf=fact,
                 label_res=res, label_op1=op1, label_op2=op2 or '', iblk_res=1, iblk_op1=1,
                 iblk_op2=1 if op2 else 0, tra_res=bool(tra[2]), tra_op1=bool(tra[0]),
                 tra_op2=bool(tra[1]), occ_res=o_res, rst_res=rst(len(o_res)), occ_op1=o1,
                 rst_op1=rst(len(o1)), perm=np.zeros((2, NGASTP), np.int64),
                 itf_index_info=np.zeros(3, np.int64))
        if op2:
            b.update(occ_op2=o2, rst_op2=rst(len(o2)))
        if extra:
            b.update(extra)
        return b

    def contraction_extra(it, tra):
        v = it.get('vertices')
        if v is None:
            (c1, a1), = used(it['op1'], tra[0])
            (c2, a2), = used(it['op2'], tra[1])
            if len(used(it['target'], tra[2])) != 1:
                raise FormulaListError('one-vertex CONTR item into a list of several vertices')
            cc, ca = it['cnt_c'], it['cnt_a']
            ex1 = [(_sub(c1, cc), _sub(a1, ca))]
            ex2 = [(_sub(c2, ca), _sub(a2, cc))]
            cnt = [(cc, ca)]
            merge = [[1, 1, 1, 1]] * 4
        else:
            if any(tra):
                raise FormulaListError('transposed operands of several vertices are not supported')
            ex1, ex2, cnt, merge = v['ex1'], v['ex2'], v['cnt'], v['merge']
            if not cnt:
                raise FormulaListError('a product of lists with several vertices without a contraction '
                                       'vertex: the file keeps no merge maps for it')
        return dict(n_cnt=len(cnt), occ_ex1=_occ_arr(ex1), occ_ex2=_occ_arr(ex2), occ_cnt=_occ_arr(cnt),
                    rst_ex1=rst(len(ex1)), rst_ex2=rst(len(ex2)), rst_cnt=rst(len(cnt)),
                    merge_op1=np.asarray(merge[0], np.int64), merge_op2=np.asarray(merge[1], np.int64),
                    merge_op1op2=np.asarray(merge[2], np.int64), merge_op2op1=np.asarray(merge[3], np.int64))

    def reo(it):
        nj_in, nj_out = len(occ(it['source'])), len(occ(it['target']))
        return dict(nj_out=nj_out, nj_in=nj_in, nreo=1, nreo_i0=0, ngas=ngas, nspin=nspin,
                    label_in=it['source'], label_out=it['target'], tra_in=False, tra_out=False,
                    sign=int(it.get('sign', 1)), iblk_in=1, iblk_out=1, occ_opin=occ(it['source']),
                    rst_opin=rst(nj_in), occ_opout=occ(it['target']), rst_opout=rst(nj_out),
                    occ_op0=np.zeros((nj_in, 2, NGASTP), np.int64),
                    occ_shift=_occ_arr([(it['shift_c'], it['shift_a'])]),
                    from_to=np.asarray([[it['from'] + 1, it['to'] + 1]], np.int64))

    out = []
    for it in formula['items']:
        c, t = it['cmd'], it['target']
        if c == 'INIT' and 'source' not in it:
            out.append(dict(command=INIT_TARGET, target=idx(t)))
        elif c == 'INIT':
            out.append(dict(command=CP_INTM, target=idx(t),
                            bcontr=bcontr(t, it['source'], None, it['factor'])))
        elif c == 'ADD':
            out.append(dict(command=ADD_INTM, target=idx(t),
                            bcontr=bcontr(t, it['source'], None, it['factor'])))
        elif c in ('CONTR', 'BC'):
            tra = tuple(bool(x) for x in (it.get('tra') or (False, False, False)))
            out.append(dict(command=ADD_BC if c == 'CONTR' else BC, target=idx(t),
                            bcontr=bcontr(t, it['op1'], it['op2'], it['factor'],
                                          contraction_extra(it, tra), tra)))
        elif c == 'REORDER':
            if float(it.get('factor', 1.0)) != 1.0:
                raise FormulaListError('REORDER items with a factor have no GeCCo command')
            out.append(dict(command=REORDER, target=idx(t), reo=reo(it)))
        elif c == 'ADD_REO':
            out.append(dict(command=ADD_REO, target=idx(t),
                            bcontr=bcontr(t, it['source'], None, it['factor']), reo=reo(it)))
        elif c == 'SYMMETRISE':
            out.append(dict(command=SYMMETRISE, target=idx(t)))
        elif c == 'NEW':
            d = lists[t]
            vs = _list_vertices(d)
            out.append(dict(command=NEW_INTERMEDIATE, target=idx(t),
                            opdef=dict(nblk=1, njoined=len(vs), ngas=ngas, nspin=nspin, name=t,
                                       tra=False, parent1='', tra1=False, parent2='', tra2=False,
                                       occ=_occ_arr(vs), rst=rst(len(vs)), incore=0)))
        elif c == 'DELETE':
            out.append(dict(command=DEL_INTERMEDIATE, target=idx(t), label=t))
        else:
            raise FormulaListError(f'formula item {c} has no GeCCo command')
    return dict(title=title, items=out)
