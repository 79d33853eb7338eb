"""Oracle: binary contraction of two ME-lists, GeCCo route 3 and route 2.

TEST INFRASTRUCTURE ONLY (see oracle/strings.py header).

Restated routines (reference file:line):
  * condensed contraction info     contract/condense_bc_info.f:242-365 (one
                                   vertex per operator: one component per
                                   space type and side)
  * block enumerator               contract/contr_op1op2_wmaps_c.f:964-2070
      Ms loop and use_tr skips     :984-1019 (next_dist, baserout/next_dist.f)
      Ms of contraction            :1071-1101
      IRREP loops                  :1117-1255
      EX1/EX2/CNT distributions    :1255-1517 (merge_msgmdis)
      norm screening 1d-28         :1652-1655, :1767-1770
      kernel call, xfac*casign*fac_scal  :1892-1949
  * map blocks                     strings/get_strmap_blk_c.f:68-200
  * spin-flip symmetrisation       contract/sym_ab_blk.f:2-405, called with
                                   prefac 1/2 and relfac absym at
                                   contract/contr_op1op2_wmaps_c.f:2080-2103
  * kernels (C library)            oracle/gecco_oracle.c

Parity status: pinned; see tests/test_oracle_contract.py (dense
antisymmetrised einsum, route 2 vs route 3, blocking invariance,
use_tr on/off).
"""
import ctypes
import itertools
import os
import numpy as np

from .strings import multd2h, strmap, flipmap, strings_of, to_spinorbs, std_spsign_msdis
from .melist import condense, condense_n, msgamdists, idxms_of, msa2idxms4op, set_op_ldim_c, HPVXSEQ

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


_BLAS = None


def blas():
    """numpy's bundled OpenBLAS (ILP64, symbols with a 64_ suffix), or None:
    (library handle, address of its Fortran-interface dgemm)."""
    global _BLAS
    if _BLAS is None:
        import glob
        libs = os.path.join(os.path.dirname(os.path.dirname(np.__file__)), 'numpy.libs')
        found = sorted(glob.glob(os.path.join(libs, 'libscipy_openblas64_*.so')))
        _BLAS = (None, None)
        if found:
            h = ctypes.CDLL(found[0])
            if hasattr(h, 'scipy_dgemm_64_'):
                _BLAS = (h, ctypes.cast(h.scipy_dgemm_64_, ctypes.c_void_p).value)
    return _BLAS if _BLAS[0] is not None else None


def blas_threads(n):
    """Thread count of the bundled OpenBLAS (affects numpy too)."""
    b = blas()
    if b is not None and hasattr(b[0], 'scipy_openblas_set_num_threads64_'):
        b[0].scipy_openblas_set_num_threads64_(ctypes.c_int(int(n)))


_USE_BLAS = True


def use_blas(on):
    """Route 3's dgemm from the bundled ILP64 OpenBLAS (on, the default when
    present; the reference links a BLAS dgemm) or the plain loop (off)."""
    global _USE_BLAS
    _USE_BLAS = bool(on)
    if _LIB is not None:
        b = blas() if _USE_BLAS else None
        _LIB.orc_set_dgemm(b[1] if b is not None else None)


def lib():
    """The C restatement (see use_blas for its dgemm)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, '_build', 'libgecco_oracle.so')
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(['make', '-s', '-C', _HERE])
        _LIB = ctypes.CDLL(path)
        _LIB.orc_contr_blk1blk2_blocked_mm.restype = ctypes.c_int
        _LIB.orc_set_dgemm.argtypes = [ctypes.c_void_p]
        _LIB.orc_set_collect_scatter.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        use_blas(_USE_BLAS)
    return _LIB


def use_collect_scatter(collect=None, scatter=None):
    """Route 3's collect and scatter steps through external functions with
    the reference's Fortran argument lists (addresses; None: the
    restatements) -- the test hook for the product's collect_block_impr_
    and scatter_block_ entries."""
    lib().orc_set_collect_scatter(collect, scatter)


def _sub(a, b):
    out = {}
    for h in HPVXSEQ:
        v = a.get(h, 0) - b.get(h, 0)
        if v < 0:
            raise ValueError('contraction removes more than the operator holds')
        if v:
            out[h] = v
    return out


def _add(a, b):
    out = {}
    for h in HPVXSEQ:
        v = a.get(h, 0) + b.get(h, 0)
        if v:
            out[h] = v
    return out


def _cinfo(occ):
    oc, hc, _, _ = condense(occ, {})
    return oc, hc


def _map_info(tgt_h, s1_h, s2_h):
    mi = []
    for h in tgt_h:
        i1 = [i + 1 for i, x in enumerate(s1_h) if x == h]
        i2 = [i + 1 for i, x in enumerate(s2_h) if x == h]
        mi += [len(i1)] + i1 + [len(i2)] + i2
    return mi


def _merge_lists(mmap, ntgt):
    """contract/condense_merge_map.f short form -> per target vertex the
    (source-1 vertices, source-2 vertices), 1-based."""
    out, k = [], 0
    for _ in range(ntgt):
        n = mmap[k]; v1 = list(mmap[k + 1:k + 1 + n]); k += 1 + n
        n = mmap[k]; v2 = list(mmap[k + 1:k + 1 + n]); k += 1 + n
        out.append((v1, v2))
    return out


def _idx_seq(vocc, side):
    """contract/set_idx_seq.f (sep_ca): condensed index (1-based) of every
    (space, vertex) of one side, X,P,V,H outer and vertex inner."""
    seq, k = {}, 0
    for h in HPVXSEQ:
        for v, occ in enumerate(vocc):
            if occ[side].get(h, 0) > 0:
                k += 1
                seq[(h, v + 1)] = k
    return seq


def _set_mapping_info(tgt, merge, src1, dag1, src2, dag2, side):
    """contract/set_mapping_info.f:84-170 for one side (0: C, 1: A).  tgt,
    src1, src2: per vertex (occ_c, occ_a) dicts; merge: short-form merge map
    of the target vertices; dagger swaps a source's C and A sides."""
    s1, s2 = (1 - side if dag1 else side), (1 - side if dag2 else side)
    seq1, seq2 = _idx_seq(src1, s1), _idx_seq(src2, s2)
    lists = _merge_lists(merge, len(tgt))
    mi = []
    for h in HPVXSEQ:
        for tv, occ in enumerate(tgt):
            if occ[side].get(h, 0) == 0:
                continue
            v1, v2 = lists[tv]
            i1 = [seq1[(h, v)] for v in v1 if src1[v - 1][s1].get(h, 0) > 0]
            i2 = [seq2[(h, v)] for v in v2 if src2[v - 1][s2].get(h, 0) > 0]
            mi += [len(i1)] + i1 + [len(i2)] + i2
    return mi


def _vocc(mel):
    return [(dict(c), dict(a)) for c, a in zip(mel.vocc_c, mel.vocc_a)]


class Contraction:
    """Binary contraction  Res += fact * Op1 . Op2  over CNT.

    cnt_c: occupation (hpvx -> n) of the contracted part on Op1's C side
           (it meets Op2's A side); cnt_a: on Op1's A side (meets Op2's C).
    An operand may be an oracle.melist.Adjoint (the vertex's dagger flag):
    its kernel calls then carry tra_op1 / tra_op2 / tra_op1op2.
    Operand roles follow contract/contr_op1op2_wmaps_c.f:1541-1551,1657-1668:
      op1 C = CNT C + EX1 C, op1 A = CNT A + EX1 A,
      op2 C = CNT A + EX2 C, op2 A = CNT C + EX2 A,
      res C = EX1 C + EX2 C, res A = EX2 A + EX1 A.
    """

    def __init__(self, op1, op2, res, cnt_c, cnt_a, fact=1.0):
        if max(op1.njoined, op2.njoined, res.njoined) > 1:
            raise ValueError('Contraction: one-vertex operators only')
        self.op1, self.op2, self.res = op1, op2, res
        self.orbs = op1.orbs
        self.fact = fact
        self.cnt_c = {h: n for h, n in cnt_c.items() if n}
        self.cnt_a = {h: n for h, n in cnt_a.items() if n}
        self.ex1_c = _sub(op1.occ_c, self.cnt_c)
        self.ex1_a = _sub(op1.occ_a, self.cnt_a)
        self.ex2_c = _sub(op2.occ_c, self.cnt_a)
        self.ex2_a = _sub(op2.occ_a, self.cnt_c)
        if _add(self.ex1_c, self.ex2_c) != {h: n for h, n in res.occ_c.items() if n} or \
           _add(self.ex1_a, self.ex2_a) != {h: n for h, n in res.occ_a.items() if n}:
            raise ValueError('result occupation does not match EX1+EX2')
        self.ex1c_o, self.ex1c_h = _cinfo(self.ex1_c)
        self.ex1a_o, self.ex1a_h = _cinfo(self.ex1_a)
        self.ex2c_o, self.ex2c_h = _cinfo(self.ex2_c)
        self.ex2a_o, self.ex2a_h = _cinfo(self.ex2_a)
        self.cntc_o, self.cntc_h = _cinfo(self.cnt_c)
        self.cnta_o, self.cnta_h = _cinfo(self.cnt_a)
        self.mi_12c = _map_info(res.hc, self.ex1c_h, self.ex2c_h)
        self.mi_12a = _map_info(res.ha, self.ex2a_h, self.ex1a_h)
        self.mi_1c = _map_info(op1.hc, self.cntc_h, self.ex1c_h)
        self.mi_1a = _map_info(op1.ha, self.cnta_h, self.ex1a_h)
        self.mi_2c = _map_info(op2.hc, self.cnta_h, self.ex2c_h)
        self.mi_2a = _map_info(op2.ha, self.cntc_h, self.ex2a_h)

    @classmethod
    def joined(cls, op1, op2, res, ex1, ex2, cnt, merge1, merge2, merge12, merge21, fact=1.0):
        """Contraction of operators with several vertices (njoined >= 1), as
        contr_op1op2 receives it (contract/contr_op1op2.f:2-17): ex1 / ex2 /
        cnt are per-vertex (occ_c, occ_a) dicts of EX1 (one per op1 vertex),
        EX2 (one per op2 vertex) and the contraction lines CNT (one per
        contracted vertex pair); merge1 / merge2 / merge12 / merge21 are the
        short-form merge maps of contract/get_bc_info3.f:271-330
        (condense_merge_map): per op1 vertex its (CNT vertices, EX1
        vertices), per op2 vertex its (CNT vertices, EX2 vertices), per result
        vertex its (EX1 vertices, EX2 vertices), and that map with the two
        lists swapped (the A-side order).  The condensed description follows
        contract/condense_bc_info.f:242-365 and the map_info arrays
        contract/set_mapping_info.f (OP2 takes CNT daggered).  res is the
        list the kernel calls write (op1op2tmp); a different target vertex
        grouping is reached by a reorder of res afterwards
        (contract/contr_op1op2_wmaps_c.f:1962-1995).

        A target component may take at most one component from each source
        list: the route-3 collect and scatter steps carry one string per
        target component and source (contract/collect_block_impr.f:129-194,
        ireo from contract/set_strmapdim_c2.f), so anything else is
        refused."""
        self = cls.__new__(cls)
        if any(getattr(x, 'transposed', False) for x in (op1, op2, res)):
            raise ValueError('Contraction.joined: stored lists only')
        self.op1, self.op2, self.res = op1, op2, res
        self.orbs = op1.orbs
        self.fact = fact
        ex1 = [(dict(c), dict(a)) for c, a in ex1]
        ex2 = [(dict(c), dict(a)) for c, a in ex2]
        cnt = [(dict(c), dict(a)) for c, a in cnt]
        if len(ex1) != op1.njoined or len(ex2) != op2.njoined:
            raise ValueError('Contraction.joined: EX1/EX2 need one occupation per operand vertex')
        o1, o2, o12 = _vocc(op1), _vocc(op2), _vocc(res)
        dag = lambda x: (x[1], x[0])  # noqa: E731

        def check(tgt, mmap, s1, s2, what, sides=(0, 1)):
            for tv, (v1, v2) in enumerate(_merge_lists(mmap, len(tgt))):
                for side in sides:
                    tot = {}
                    for v in v1:
                        for h, n in s1[v - 1][side].items():
                            tot[h] = tot.get(h, 0) + n
                    for v in v2:
                        for h, n in s2[v - 1][side].items():
                            tot[h] = tot.get(h, 0) + n
                    if {h: n for h, n in tot.items() if n} != \
                       {h: n for h, n in tgt[tv][side].items() if n}:
                        raise ValueError(f'Contraction.joined: {what} vertex {tv + 1} does not '
                                         'equal the sum of its merged parts')
        check(o1, merge1, cnt, ex1, 'op1')
        check(o2, merge2, [dag(x) for x in cnt], ex2, 'op2')
        check(o12, merge12, ex1, ex2, 'result', sides=(0,))
        check(o12, merge21, ex2, ex1, 'result', sides=(1,))
        self.cnt_vocc, self.ex1_vocc, self.ex2_vocc = cnt, ex1, ex2
        self.merge = (list(merge1), list(merge2), list(merge12), list(merge21))

        def cond(v):
            oc, hc, _, oa, ha, _ = condense_n([c for c, _ in v], [a for _, a in v])
            return oc, hc, oa, ha
        self.ex1c_o, self.ex1c_h, self.ex1a_o, self.ex1a_h = cond(ex1)
        self.ex2c_o, self.ex2c_h, self.ex2a_o, self.ex2a_h = cond(ex2)
        self.cntc_o, self.cntc_h, self.cnta_o, self.cnta_h = cond(cnt)
        self.mi_12c = _set_mapping_info(o12, merge12, ex1, False, ex2, False, 0)
        self.mi_12a = _set_mapping_info(o12, merge21, ex2, False, ex1, False, 1)
        self.mi_1c = _set_mapping_info(o1, merge1, cnt, False, ex1, False, 0)
        self.mi_1a = _set_mapping_info(o1, merge1, cnt, False, ex1, False, 1)
        self.mi_2c = _set_mapping_info(o2, merge2, cnt, True, ex2, False, 0)
        self.mi_2a = _set_mapping_info(o2, merge2, cnt, True, ex2, False, 1)
        for mi, n in ((self.mi_12c, len(res.hc)), (self.mi_12a, len(res.ha)),
                      (self.mi_1c, len(op1.hc)), (self.mi_1a, len(op1.ha)),
                      (self.mi_2c, len(op2.hc)), (self.mi_2a, len(op2.ha))):
            k = 0
            for _ in range(n):
                n1 = mi[k]; k += 1 + n1
                n2 = mi[k]; k += 1 + n2
                if n1 > 1 or n2 > 1:
                    raise ValueError('Contraction.joined: multi map needed (two components of '
                                     'one source merge into one target component)')
        return self

    # ---------------------------------------------------------------
    def _lens(self, occs, hs, idxms, gams):
        return [self.orbs.lenstr(h, o, i, g) for o, h, i, g in zip(occs, hs, idxms, gams)]

    def _strmap_blk(self, tgt_h, mi, s1, s2):
        """strings/get_strmap_blk_c.f:68-200.  s1/s2 = (occs, hpvxs, idxms,
        gams, lens) of source 1 and 2."""
        out = []
        k = 0
        for h in tgt_h:
            n1 = mi[k]; i1 = mi[k + 1:k + 1 + n1]; k += 1 + n1
            n2 = mi[k]; i2 = mi[k + 1:k + 1 + n2]; k += 1 + n2
            if n1 and n2:
                a, b = i1[0] - 1, i2[0] - 1
                out += strmap(self.orbs.spaces[h], s1[0][a], s1[2][a], s1[3][a],
                              s2[0][b], s2[2][b], s2[3][b])
            elif n1:
                out += list(range(1, s1[4][i1[0] - 1] + 1))
            else:
                out += list(range(1, s2[4][i2[0] - 1] + 1))
        return out

    @staticmethod
    def _merge(mi, n_tgt, ms1, g1, ms2, g2):
        """contract/merge_msgmdis.f:2-64."""
        msr, gr = [], []
        k = 0
        for _ in range(n_tgt):
            n = mi[k]; k += 1
            m, g = 0, 1
            for _ in range(n):
                m += ms1[mi[k] - 1]; g = multd2h(g, g1[mi[k] - 1]); k += 1
            n = mi[k]; k += 1
            for _ in range(n):
                m += ms2[mi[k] - 1]; g = multd2h(g, g2[mi[k] - 1]); k += 1
            msr.append(m); gr.append(g)
        return msr, gr

    def blocks(self, use_tr=True):
        """Enumerate kernel calls as contr_op1op2_wmaps_c.f:984-1950 does."""
        op1, op2, res, nsym = self.op1, self.op2, self.res, self.orbs.nsym
        na = (op1.na, op2.na, res.na)
        nc = (op1.nc, op2.nc, res.nc)
        nc_cnt, na_cnt = sum(self.cntc_o), sum(self.cnta_o)
        nc_ex1, na_ex1 = sum(self.ex1c_o), sum(self.ex1a_o)
        nc_ex2, na_ex2 = sum(self.ex2c_o), sum(self.ex2a_o)
        use_tr_here = use_tr and op1.absym != 0 and op2.absym != 0 and \
            res.absym != 0 and op1.absym * op2.absym == res.absym
        idx_restr = 1 if na[1] <= na[0] else 2
        out = []
        # next_dist over (ms1,ms2,ms3) from max, step -2, first index fastest
        ranges = [list(range(n, -n - 1, -2)) for n in na]
        for ms3 in ranges[2]:
            for ms2 in ranges[1]:
                for ms1 in ranges[0]:
                    msa = (ms1, ms2, ms3)
                    msc12 = msa
                    if any(abs(msc12[k]) > nc[k] for k in range(3)):
                        continue
                    msc_ac = msa[0] + msa[1] - msa[2]
                    if na_cnt + nc_cnt < abs(msc_ac):
                        continue
                    mr = msa[idx_restr - 1]
                    if use_tr_here and (msa[2] < 0 or (msa[2] == 0 and mr < 0)):
                        continue
                    fac0 = 1.0
                    if use_tr_here and msa[2] > 0:
                        fac0 *= 2.0
                    if use_tr_here and msa[2] == 0 and mr > 0:
                        fac0 *= 2.0
                    for msc_a in range(na_cnt, -na_cnt - 1, -2):
                        msc_c = msc_ac - msc_a
                        if abs(msc_c) > nc_cnt:
                            continue
                        msex1_c = msc12[0] - msc_c
                        msex1_a = msa[0] - msc_a
                        msex2_a = msa[1] - msc_c
                        msex2_c = msc12[1] - msc_a
                        if abs(msex1_c) > nc_ex1 or abs(msex1_a) > na_ex1 or \
                           abs(msex2_c) > nc_ex2 or abs(msex2_a) > na_ex2:
                            continue
                        if use_tr_here and mr == 0 and msa[2] == 0 and msc_c - msc_a > 0:
                            continue
                        fac = fac0
                        if use_tr_here and mr == 0 and msa[2] == 0 and msc_c - msc_a < 0:
                            fac *= 2.0
                        self._gam_loops(out, msa, msc_c, msc_a, msex1_c, msex1_a,
                                        msex2_c, msex2_a, fac, nsym)
        return out

    def _gam_loops(self, out, msa, msc_c, msc_a, msex1_c, msex1_a, msex2_c,
                   msex2_a, fac, nsym):
        op1, op2, res = self.op1, self.op2, self.res
        for g3 in range(1, nsym + 1):
            for g2 in range(1, nsym + 1):
                for g1 in range(1, nsym + 1):
                    if op1.block_len(msa[0], g1) == 0 or op2.block_len(msa[1], g2) == 0 \
                       or res.block_len(msa[2], g3) == 0:
                        continue
                    gc_ac = multd2h(multd2h(g1, g2), g3)
                    for gca in range(1, nsym + 1):
                        gcc = multd2h(gca, gc_ac)
                        if not self.cntc_o and gcc > 1:
                            continue
                        if not self.cnta_o and gca > 1:
                            continue
                        gex1_a = multd2h(g1, gca)
                        gex1_c = multd2h(g1, gcc)
                        gex2_a = multd2h(g2, gcc)
                        gex2_c = multd2h(g2, gca)
                        if (not self.ex1c_o and gex1_c > 1) or (not self.ex1a_o and gex1_a > 1) or \
                           (not self.ex2c_o and gex2_c > 1) or (not self.ex2a_o and gex2_a > 1):
                            continue
                        self._dist_loops(out, msa, (g1, g2, g3), msc_c, msc_a, gcc, gca,
                                         msex1_c, msex1_a, gex1_c, gex1_a,
                                         msex2_c, msex2_a, gex2_c, gex2_a, fac, nsym)

    def _dist_loops(self, out, msa, gams, msc_c, msc_a, gcc, gca, msex1_c, msex1_a,
                    gex1_c, gex1_a, msex2_c, msex2_a, gex2_c, gex2_a, fac, nsym):
        op1, op2, res = self.op1, self.op2, self.res
        for m1c, m1a, g1c, g1a in msgamdists(self.ex1c_o, self.ex1a_o, msex1_c, msex1_a,
                                             gex1_c, gex1_a, nsym):
            i1c = [idxms_of(o, m) for o, m in zip(self.ex1c_o, m1c)]
            i1a = [idxms_of(o, m) for o, m in zip(self.ex1a_o, m1a)]
            l1c = self._lens(self.ex1c_o, self.ex1c_h, i1c, g1c)
            l1a = self._lens(self.ex1a_o, self.ex1a_h, i1a, g1a)
            if any(x == 0 for x in l1c + l1a):
                continue
            for m2c, m2a, g2c, g2a in msgamdists(self.ex2c_o, self.ex2a_o, msex2_c, msex2_a,
                                                 gex2_c, gex2_a, nsym):
                i2c = [idxms_of(o, m) for o, m in zip(self.ex2c_o, m2c)]
                i2a = [idxms_of(o, m) for o, m in zip(self.ex2a_o, m2a)]
                l2c = self._lens(self.ex2c_o, self.ex2c_h, i2c, g2c)
                l2a = self._lens(self.ex2a_o, self.ex2a_h, i2a, g2a)
                if any(x == 0 for x in l2c + l2a):
                    continue
                mrc, grc = self._merge(self.mi_12c, len(res.hc), m1c, g1c, m2c, g2c)
                mra, gra = self._merge(self.mi_12a, len(res.ha), m2a, g2a, m1a, g1a)
                irc = [idxms_of(o, m) for o, m in zip(res.oc, mrc)]
                ira = [idxms_of(o, m) for o, m in zip(res.oa, mra)]
                dres = res.find(msa[2], gams[2], irc, grc, ira, gra)
                if dres is None:
                    continue
                ex1 = (self.ex1c_o + self.ex1a_o, self.ex1c_h + self.ex1a_h,
                       i1c + i1a, list(g1c) + list(g1a), l1c + l1a)
                ex2 = (self.ex2c_o + self.ex2a_o, self.ex2c_h + self.ex2a_h,
                       i2c + i2a, list(g2c) + list(g2a), l2c + l2a)
                nc1, nc2 = len(self.ex1c_o), len(self.ex2c_o)
                s_ex1c = tuple(x[:nc1] for x in ex1)
                s_ex1a = tuple(x[nc1:] for x in ex1)
                s_ex2c = tuple(x[:nc2] for x in ex2)
                s_ex2a = tuple(x[nc2:] for x in ex2)
                map12c = self._strmap_blk(res.hc, self.mi_12c, s_ex1c, s_ex2c)
                map12a = self._strmap_blk(res.ha, self.mi_12a, s_ex2a, s_ex1a)
                for mcc, mca, gcc_d, gca_d in msgamdists(self.cntc_o, self.cnta_o, msc_c, msc_a,
                                                         gcc, gca, nsym):
                    icc = [idxms_of(o, m) for o, m in zip(self.cntc_o, mcc)]
                    ica = [idxms_of(o, m) for o, m in zip(self.cnta_o, mca)]
                    lcc = self._lens(self.cntc_o, self.cntc_h, icc, gcc_d)
                    lca = self._lens(self.cnta_o, self.cnta_h, ica, gca_d)
                    if any(x == 0 for x in lcc + lca):
                        continue
                    m1oc, g1oc = self._merge(self.mi_1c, len(op1.hc), mcc, gcc_d, m1c, g1c)
                    m1oa, g1oa = self._merge(self.mi_1a, len(op1.ha), mca, gca_d, m1a, g1a)
                    d1 = op1.find(msa[0], gams[0],
                                  [idxms_of(o, m) for o, m in zip(op1.oc, m1oc)], g1oc,
                                  [idxms_of(o, m) for o, m in zip(op1.oa, m1oa)], g1oa)
                    m2oc, g2oc = self._merge(self.mi_2c, len(op2.hc), mca, gca_d, m2c, g2c)
                    m2oa, g2oa = self._merge(self.mi_2a, len(op2.ha), mcc, gcc_d, m2a, g2a)
                    d2 = op2.find(msa[1], gams[1],
                                  [idxms_of(o, m) for o, m in zip(op2.oc, m2oc)], g2oc,
                                  [idxms_of(o, m) for o, m in zip(op2.oa, m2oa)], g2oa)
                    if d1 is None or d2 is None:
                        continue
                    s_cc = (self.cntc_o, self.cntc_h, icc, list(gcc_d), lcc)
                    s_ca = (self.cnta_o, self.cnta_h, ica, list(gca_d), lca)
                    map1c = self._strmap_blk(op1.hc, self.mi_1c, s_cc, s_ex1c)
                    map1a = self._strmap_blk(op1.ha, self.mi_1a, s_ca, s_ex1a)
                    map2c = self._strmap_blk(op2.hc, self.mi_2c, s_ca, s_ex2c)
                    map2a = self._strmap_blk(op2.ha, self.mi_2a, s_cc, s_ex2a)
                    out.append(dict(
                        xfac=self.fact * fac, fac_scal=fac,
                        tra_op1=int(getattr(op1, 'transposed', False)),
                        tra_op2=int(getattr(op2, 'transposed', False)),
                        tra_op1op2=int(getattr(res, 'transposed', False)),
                        off1=d1['off'], len1=d1['len'], off2=d2['off'], len2=d2['len'],
                        off12=dres['off'], len12=dres['len'],
                        ncblk_op1=len(op1.hc), nablk_op1=len(op1.ha),
                        ncblk_ex1=len(self.ex1c_o), nablk_ex1=len(self.ex1a_o),
                        ncblk_op2=len(op2.hc), nablk_op2=len(op2.ha),
                        ncblk_ex2=len(self.ex2c_o), nablk_ex2=len(self.ex2a_o),
                        ncblk_cnt=len(self.cntc_o), nablk_cnt=len(self.cnta_o),
                        ncblk_op1op2=len(res.hc), nablk_op1op2=len(res.ha),
                        hpvxop1c=op1.hc, hpvxop1a=op1.ha, hpvxop2c=op2.hc, hpvxop2a=op2.ha,
                        hpvxop1op2c=res.hc, hpvxop1op2a=res.ha,
                        lstrop1=d1['lens'], lstrop2=d2['lens'], lstrop1op2=dres['lens'],
                        lstr_ex1=l1c + l1a, lstr_ex2=l2c + l2a, lstr_cnt=lcc + lca,
                        map_info_12c=self.mi_12c, map_info_12a=self.mi_12a,
                        map_info_1c=self.mi_1c, map_info_1a=self.mi_1a,
                        map_info_2c=self.mi_2c, map_info_2a=self.mi_2a,
                        map_ex1ex2c=map12c, map_ex1ex2a=map12a,
                        map_ex1cntc=map1c, map_ex1cnta=map1a,
                        map_ex2cntc=map2c, map_ex2cnta=map2a))

    # ---------------------------------------------------------------
    def run(self, x1, x2, xr, route=3, use_tr=True, lenblock=200, lenblock_cnt=-1,
            lenscr=None, screen=True, blocks=None):
        """Res list xr (in place) += contraction; returns the number of kernel
        calls.  route 3: collect/dgemm/scatter; route 2: scalar kernel."""
        L = lib()
        if blocks is None:
            blocks = self.blocks(use_tr)
        ncall = 0
        for b in blocks:
            a1 = x1[b['off1']:b['off1'] + b['len1']]
            a2 = x2[b['off2']:b['off2'] + b['len2']]
            if screen and (float(np.dot(a1, a1)) < 1e-28 or float(np.dot(a2, a2)) < 1e-28):
                continue
            ncall += 1
            call_kernel(L, b, x1, x2, xr, route, lenblock, lenblock_cnt, lenscr)
        if use_tr:
            use_tr_here = self.op1.absym * self.op2.absym == self.res.absym != 0
            if use_tr_here:     # the projection of the stored list
                sym_ab(getattr(self.res, 'base', self.res), xr, 0.5, float(self.res.absym))
        return ncall

    def flops(self, use_tr=True):
        """Algorithmic FLOP: sum over kernel calls of 2*n_ex1*n_ex2*n_cnt."""
        f = 0
        for b in self.blocks(use_tr):
            f += 2 * int(np.prod(b['lstr_ex1'])) * int(np.prod(b['lstr_ex2'])) * \
                int(np.prod(b['lstr_cnt']))
        return f


def _i64(seq):
    a = np.asarray(list(seq) if len(seq) else [0], dtype=np.int64)
    return a


def kernel_args(b):
    """Integer argument arrays of one kernel call (kept alive by caller)."""
    keys = ['hpvxop1c', 'hpvxop1a', 'hpvxop2c', 'hpvxop2a', 'hpvxop1op2c',
            'hpvxop1op2a', 'lstrop1', 'lstrop2', 'lstrop1op2', 'lstr_ex1',
            'lstr_ex2', 'lstr_cnt', 'map_info_12c', 'map_info_12a', 'map_info_1c',
            'map_info_1a', 'map_info_2c', 'map_info_2a', 'map_ex1ex2c',
            'map_ex1ex2a', 'map_ex1cntc', 'map_ex1cnta', 'map_ex2cntc', 'map_ex2cnta']
    return [_i64(b[k]) for k in keys]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def call_kernel(L, b, x1, x2, xr, route, lenblock=200, lenblock_cnt=-1, lenscr=None):
    arrs = kernel_args(b)
    dims = [b[k] for k in ('ncblk_op1', 'nablk_op1', 'ncblk_ex1', 'nablk_ex1',
                           'ncblk_op2', 'nablk_op2', 'ncblk_ex2', 'nablk_ex2',
                           'ncblk_cnt', 'nablk_cnt', 'ncblk_op1op2', 'nablk_op1op2')]
    p1 = ctypes.c_void_p(x1.ctypes.data + 8 * b['off1'])
    p2 = ctypes.c_void_p(x2.ctypes.data + 8 * b['off2'])
    p12 = ctypes.c_void_p(xr.ctypes.data + 8 * b['off12'])
    I = ctypes.c_int64
    tra = [I(int(b.get(k, 0))) for k in ('tra_op1', 'tra_op2', 'tra_op1op2')]
    if route == 3:
        nex1 = int(np.prod(b['lstr_ex1']))
        nex2 = int(np.prod(b['lstr_ex2']))
        ncnt = int(np.prod(b['lstr_cnt']))
        if lenscr is None:
            lenscr = 3 * lenblock * lenblock + (nex1 + nex2 + 2) * ncnt + \
                lenblock * ncnt + 16
        scr = np.zeros(lenscr)
        rc = L.orc_contr_blk1blk2_blocked_mm(
            ctypes.c_double(b['xfac']), p12, p1, p2, *tra,
            _ptr(scr), I(lenscr), I(lenblock), I(lenblock_cnt),
            *[I(d) for d in dims], *[_ptr(a) for a in arrs])
        if rc != 0:
            raise RuntimeError('contr_blk1blk2_blocked_mm: insufficient scratch')
    else:
        L.orc_contr_blk1blk2_wmaps_c(
            ctypes.c_double(b['xfac']), p12, p1, p2, *tra,
            *[I(d) for d in dims], *[_ptr(a) for a in arrs])


# ------------------------------------------------------------------ sym_ab
def sym_ab_pairs(mel, relfac):
    """contract/sym_ab_blk.f:214-405 for a one-vertex list, as the list of
    element pairs (i1 <= i2) and factors f = csign*asign*relfac2 that the
    reference updates as v = prefac*(x1 + f*x2); x1 = v; x2 = f*v."""
    orbs = mel.orbs
    nsym = orbs.nsym
    msc_max, msa_max = mel.nc, mel.na
    if (((msc_max + msa_max) // 2 + 1) * (msc_max - msa_max) // 2) % 2:
        relfac2 = -relfac
    else:
        relfac2 = relfac
    I1, I2, F = [], [], []
    for msa in range(msa_max, -1, -2):
        msc = msa
        if abs(msc) > msc_max:
            continue
        for igama in range(1, nsym + 1):
            if mel.block_len(msa, igama) <= 0:
                continue
            for mc, ma, gc, ga in msgamdists(mel.oc, mel.oa, msc, msa, igama, igama, nsym):
                mc2 = [-m for m in mc]
                ma2 = [-m for m in ma]
                gsign = std_spsign_msdis(mc, mel.oc) * std_spsign_msdis(ma, mel.oa) * \
                    std_spsign_msdis(mc2, mel.oc) * std_spsign_msdis(ma2, mel.oa)
                imc = [idxms_of(o, m) for o, m in zip(mel.oc, mc)]
                ima = [idxms_of(o, m) for o, m in zip(mel.oa, ma)]
                imc2 = [idxms_of(o, m) for o, m in zip(mel.oc, mc2)]
                ima2 = [idxms_of(o, m) for o, m in zip(mel.oa, ma2)]
                d1 = mel.find(msa, igama, imc, gc, ima, ga)
                if d1 is None:
                    continue
                d2 = mel.find(-msa, igama, imc2, gc, ima2, ga)
                if msa == 0 and d1['off'] > d2['off']:
                    continue
                lens = d1['lens']
                ncb = len(mel.oc)
                fl = []
                for k in range(ncb):
                    fl.append(np.asarray(flipmap(orbs.spaces[mel.hc[k]], mel.oc[k], imc[k], gc[k])))
                for k in range(len(mel.oa)):
                    fl.append(np.asarray(flipmap(orbs.spaces[mel.ha[k]], mel.oa[k], ima[k], ga[k])))
                ld = d1['ldc'] + d1['lda']
                # all elements of the block: component string indices, first fastest
                if lens:
                    grids = np.meshgrid(*[np.arange(n) for n in lens], indexing='ij')
                    i1 = d1['off'] + sum(g * l for g, l in zip(grids, ld))
                    i2 = d2['off'] + sum((np.abs(f[g]) - 1) * l for f, g, l in zip(fl, grids, ld))
                    sg = gsign * np.prod([np.sign(f[g]) for f, g in zip(fl, grids)], axis=0)
                else:   # operator without strings (a scalar)
                    i1 = np.array([d1['off']])
                    i2 = np.array([d2['off']])
                    sg = np.array([gsign])
                i1, i2, sg = i1.ravel(), i2.ravel(), np.asarray(sg).ravel()
                keep = i1 <= i2
                I1.append(i1[keep])
                I2.append(i2[keep])
                F.append(sg[keep] * relfac2)
    if not I1:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0)
    return np.concatenate(I1), np.concatenate(I2), np.concatenate(F).astype(float)


def apply_sym_pairs(x, pairs, prefac):
    """The reference's pair update, vectorised (pairs are disjoint)."""
    i1, i2, f = pairs
    v = prefac * (x[i1] + f * x[i2])
    x[i1] = v
    x[i2] = f * v


def sym_ab(mel, x, prefac, relfac):
    """contract/sym_ab_blk.f:214-405 applied to list x in place."""
    apply_sym_pairs(x, sym_ab_pairs(mel, relfac), prefac)


# ------------------------------------------------------------ dense maps
def _so_index(orb, spin):
    return 2 * (orb - 1) + (0 if spin == 1 else 1)


def element_spinorbs(mel, d):
    """For a distribution block d of list mel, yield (offset, groups) where
    groups is the tuple of spin-orbital index tuples per component (C comps
    then A comps), with elements addressed as in contract/set_op_ldim_c.f."""
    orbs = mel.orbs
    comps = []
    for k in range(len(mel.oc)):
        ms = d['msc'][k]
        comps.append(strings_of(orbs.spaces[mel.hc[k]], mel.oc[k], ms, d['gamc'][k]))
    for k in range(len(mel.oa)):
        ms = d['msa'][k]
        comps.append(strings_of(orbs.spaces[mel.ha[k]], mel.oa[k], ms, d['gama'][k]))
    ld = d['ldc'] + d['lda']
    for idx in itertools.product(*[range(len(c)) for c in comps]):
        off = d['off'] + sum(i * l for i, l in zip(idx, ld))
        groups = tuple(tuple(_so_index(o, s) for o, s in to_spinorbs(c[i]))
                       for c, i in zip(comps, idx))
        yield off, groups


def _perm_sign(p):
    p = list(p)
    s = 1
    for i in range(len(p)):
        while p[i] != i:
            j = p[i]
            p[i], p[j] = p[j], p[i]
            s = -s
    return s


def to_dense(mel, x):
    """Expand list x into a dense tensor antisymmetric within each component;
    axes: C components (hpvxblkseq order, each expanded to its occupation),
    then A components."""
    orbs = mel.orbs
    dims = []
    for h, o in zip(mel.hc + mel.ha, mel.oc + mel.oa):
        dims += [2 * orbs.spaces[h].norb] * o
    D = np.zeros(dims)
    occs = mel.oc + mel.oa
    perms = [list(itertools.permutations(range(o))) for o in occs]
    for dl in mel.blocks.values():
        for d in dl:
            for off, groups in element_spinorbs(mel, d):
                v = x[off]
                for combo in itertools.product(*perms):
                    sgn = 1
                    idx = []
                    for g, p in zip(groups, combo):
                        sgn *= _perm_sign(p)
                        idx += [g[i] for i in p]
                    D[tuple(idx)] = sgn * v
    return D


def from_dense(mel, D):
    x = np.zeros(mel.length)
    for dl in mel.blocks.values():
        for d in dl:
            for off, groups in element_spinorbs(mel, d):
                x[off] = D[tuple(i for g in groups for i in g)]
    return x
