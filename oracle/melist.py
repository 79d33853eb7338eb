"""Oracle: ME-list storage layout.

TEST INFRASTRUCTURE ONLY (see oracle/strings.py header).

Restated routines (reference file:line):
  * condensed occupations      operators/condense_occ.f (hpvxblkseq order,
                               include/hpvxseq.h:4 = X,P,V,H)
  * Ms/IRREP distributions     operators/next_msgamdist2.f:2-263
  * string lengths             operators/set_len_str2.f (resort=.false.)
  * list layout and offsets    operators/set_op_dim2.f:146-459
  * Ms(A) -> block index       operators/msa2idxms4op.f:24-27
  * in-block leading dims      contract/set_op_ldim_c.f:237-308
  * distribution id            operators/msgmdid2.f:2-34

Scope: operators with one or more vertices (njoined >= 1; the condensed
components of all vertices, X,P,V,H outer and vertex inner), spaces H and P
with one orbital subspace each, mst=0, gamt=1 (closed-shell, totally
symmetric operators), no diagonal restriction.

Parity status: pinned by closed-form block lengths (e.g. T2 at nocc=8,
nvir=32 has 93 312 elements) and by the dense-tensor round trip in
tests/test_oracle_contract.py.
"""
from .strings import multd2h, lenstr

IHOLE, IPART, IVALE, IEXTR = 1, 2, 3, 4
HPVXSEQ = (4, 2, 3, 1)          # include/hpvxseq.h:2-4


def condense(occ_c, occ_a):
    """occ_c/occ_a: dict hpvx->count.  Returns (occ_csub, hpvx_csub, occ_asub,
    hpvx_asub) in hpvxblkseq order (operators/condense_occ.f)."""
    oc, hc, oa, ha = [], [], [], []
    for h in HPVXSEQ:
        if occ_c.get(h, 0) > 0:
            oc.append(occ_c[h])
            hc.append(h)
        if occ_a.get(h, 0) > 0:
            oa.append(occ_a[h])
            ha.append(h)
    return oc, hc, oa, ha


def condense_n(vocc_c, vocc_a):
    """Several vertices (lists of dicts hpvx->count): operators/condense_occ.f
    with the vertex loop inside the hpvxblkseq loop.  Returns the condensed
    lists plus the vertex (0-based) of every C and A component."""
    oc, hc, vc, oa, ha, va = [], [], [], [], [], []
    for h in HPVXSEQ:
        for v in range(len(vocc_c)):
            if vocc_c[v].get(h, 0) > 0:
                oc.append(vocc_c[v][h]); hc.append(h); vc.append(v)
            if vocc_a[v].get(h, 0) > 0:
                oa.append(vocc_a[v][h]); ha.append(h); va.append(v)
    return oc, hc, vc, oa, ha, va


def first_msdistn(occ, mst):
    n = len(occ)
    if n == 0:
        return (mst == 0), []
    if n > 1:
        d = list(occ[:-1]) + [mst - sum(occ[:-1])]
    else:
        d = [mst]
    if abs(d[-1]) <= occ[-1]:
        return True, d
    ok = next_msdistn(d, mst, occ)
    return ok, d


def next_msdistn(d, mst, occ):
    n = len(occ)
    if n <= 1:
        return False
    while True:
        ok = False
        for i in range(n - 1):
            if d[i] > -occ[i]:
                d[i] -= 2
                ok = True
                break
            else:
                d[i] = occ[i]
        if not ok:
            return False
        d[n - 1] = mst - sum(d[:n - 1])
        if abs(d[n - 1]) <= occ[n - 1]:
            return True


def first_gamdistn(gamt, nsym, n):
    if n == 0:
        return (gamt == 1), []
    return True, [1] * (n - 1) + [gamt]


def next_gamdistn(g, gamt, nsym):
    n = len(g)
    if n <= 1:
        return False
    ok = False
    ilast = 0
    for i in range(n - 1):
        if g[i] < nsym:
            g[i] += 1
            ilast = i
            ok = True
            break
        else:
            g[i] = 1
    if ok:
        acc = gamt
        for i in range(ilast, n - 1):
            acc = multd2h(acc, g[i])
        g[n - 1] = acc
    return ok


def msgamdists(occ_c, occ_a, ms_c, ms_a, gam_c, gam_a, nsym):
    """Generator over next_msgamdist2 (operators/next_msgamdist2.f:39-84):
    innermost C IRREP, then C Ms, then A IRREP, then A Ms."""
    ok1, mc = first_msdistn(occ_c, ms_c)
    ok2, ma = first_msdistn(occ_a, ms_a)
    ok3, gc = first_gamdistn(gam_c, nsym, len(occ_c))
    ok4, ga = first_gamdistn(gam_a, nsym, len(occ_a))
    if not (ok1 and ok2 and ok3 and ok4):
        return
    while True:
        yield list(mc), list(ma), list(gc), list(ga)
        if next_gamdistn(gc, gam_c, nsym):
            continue
        if next_msdistn(mc, ms_c, occ_c):
            _, gc[:] = first_gamdistn(gam_c, nsym, len(occ_c))
            continue
        if next_gamdistn(ga, gam_a, nsym):
            _, mc[:] = first_msdistn(occ_c, ms_c)
            _, gc[:] = first_gamdistn(gam_c, nsym, len(occ_c))
            continue
        if next_msdistn(ma, ms_a, occ_a):
            _, ga[:] = first_gamdistn(gam_a, nsym, len(occ_a))
            _, mc[:] = first_msdistn(occ_c, ms_c)
            _, gc[:] = first_gamdistn(gam_c, nsym, len(occ_c))
            continue
        return


def idxms_of(occ, ms):
    return (occ - ms) // 2 + 1


def msa2idxms4op(msa, mst, occ_a, occ_c):
    """operators/msa2idxms4op.f:24-27."""
    idx = (occ_a - msa) // 2 + 1
    off = occ_a - occ_c + mst
    if off > 0:
        idx -= off // 2
    return idx


def set_op_ldim_c(hpvx_c, hpvx_a, lens, transpose=False):
    """contract/set_op_ldim_c.f:237-308."""
    nc, na = len(hpvx_c), len(hpvx_a)
    ldc = [0] * nc
    lda = [0] * na
    ld = 1
    for h in HPVXSEQ:
        if not transpose:
            for i in range(nc):
                if hpvx_c[i] == h:
                    ldc[i] = ld
                    ld *= lens[i]
            for i in range(na):
                if hpvx_a[i] == h:
                    lda[i] = ld
                    ld *= lens[nc + i]
        else:
            for i in range(na - 1, -1, -1):
                if hpvx_a[i] == h:
                    lda[i] = ld
                    ld *= lens[nc + i]
            for i in range(nc - 1, -1, -1):
                if hpvx_c[i] == h:
                    ldc[i] = ld
                    ld *= lens[i]
    return ldc, lda


class Orbitals:
    """Orbital spaces of the model: H (occupied) and P (virtual)."""

    def __init__(self, h_irreps, p_irreps, nsym):
        from .strings import Space
        self.nsym = nsym
        self.spaces = {IHOLE: Space(h_irreps, nsym), IPART: Space(p_irreps, nsym)}

    def lenstr(self, hpvx, nel, idxms, igam):
        return lenstr(self.spaces[hpvx], nel, idxms, igam)


class MEList:
    """ME-list layout of a one-vertex operator (operators/set_op_dim2.f)."""

    def __init__(self, orbs, occ_c, occ_a, absym=1):
        """occ_c/occ_a: dict hpvx->count (one vertex) or a list of such
        dicts, one per vertex."""
        self.orbs = orbs
        if isinstance(occ_c, dict):
            occ_c, occ_a = [occ_c], [occ_a]
        self.vocc_c = [dict(x) for x in occ_c]
        self.vocc_a = [dict(x) for x in occ_a]
        self.njoined = len(self.vocc_c)
        self.occ_c, self.occ_a = {}, {}
        for x in self.vocc_c:
            for h, n in x.items():
                self.occ_c[h] = self.occ_c.get(h, 0) + n
        for x in self.vocc_a:
            for h, n in x.items():
                self.occ_a[h] = self.occ_a.get(h, 0) + n
        self.absym = absym
        self.oc, self.hc, self.vc, self.oa, self.ha, self.va = condense_n(self.vocc_c, self.vocc_a)
        if (sum(self.oc) - sum(self.oa)) % 2:
            raise ValueError('MEList: Ms(total) = 0 needs creator and annihilator counts of equal parity')
        self.nc = sum(self.oc)
        self.na = sum(self.oa)
        nsym = orbs.nsym
        self.blocks = {}        # (idxmsa, igama) -> list of distribution dicts
        self.gam_ms = {}        # (igama, idxmsa) -> (offset, length)
        self.dis_index = {}     # (idxmsa, igama, key) -> distribution dict
        idx = 0
        idxmsa = 0
        for msa in range(self.na, -self.na - 1, -2):
            msc = msa
            if abs(msc) > self.nc:
                continue
            idxmsa += 1
            for igama in range(1, nsym + 1):
                igamc = igama
                start = idx
                dlist = []
                for mc, ma, gc, ga in msgamdists(self.oc, self.oa, msc, msa,
                                                 igamc, igama, nsym):
                    imc = [idxms_of(o, m) for o, m in zip(self.oc, mc)]
                    ima = [idxms_of(o, m) for o, m in zip(self.oa, ma)]
                    lens = ([orbs.lenstr(h, o, i, g) for h, o, i, g in
                             zip(self.hc, self.oc, imc, gc)] +
                            [orbs.lenstr(h, o, i, g) for h, o, i, g in
                             zip(self.ha, self.oa, ima, ga)])
                    ln = 1
                    for x in lens:
                        ln *= x
                    if ln <= 0:
                        continue
                    ldc, lda = set_op_ldim_c(self.hc, self.ha, lens)
                    d = dict(off=idx, len=ln, msc=mc, msa=ma, gamc=gc, gama=ga,
                             idxmsc=imc, idxmsa=ima, lens=lens, ldc=ldc, lda=lda,
                             msa_tot=msa, igama=igama)
                    dlist.append(d)
                    key = (tuple(imc), tuple(gc), tuple(ima), tuple(ga))
                    self.dis_index[(idxmsa, igama, key)] = d
                    idx += ln
                self.blocks[(idxmsa, igama)] = dlist
                self.gam_ms[(igama, idxmsa)] = (start, idx - start)
        self.length = idx

    def find(self, msa_tot, igama, imc, gc, ima, ga):
        idxmsa = msa2idxms4op(msa_tot, 0, self.na, self.nc)
        key = (tuple(imc), tuple(gc), tuple(ima), tuple(ga))
        return self.dis_index.get((idxmsa, igama, key))

    def block_len(self, msa_tot, igama):
        idxmsa = msa2idxms4op(msa_tot, 0, self.na, self.nc)
        return self.gam_ms.get((igama, idxmsa), (0, 0))[1]


class Adjoint:
    """The adjoint X^+ of a one-vertex list X, as a contraction uses an
    operand whose vertex carries the dagger flag (contract/get_bc_info3.f:173,
    191: tra_op = vertex%dagger; the result with contr%dagger, :391).

    Its C side is X's A side and its A side X's C side.  Element (c; a) of X^+
    is the stored element (a; c) of X: the reference reaches it through the
    transposed leading dimensions of contract/set_op_ldim_c.f:51-70 and the
    C/A-swapped distribution lookup of operators/idx_msgmdst2.f (called with
    the transpose flag, contract/contr_op1op2_wmaps_c.f:1644,1759), with no
    sign.  With Ms(total) = 0 and a totally symmetric X, Ms(C) = Ms(A) and
    IRREP(C) = IRREP(A), so block (Ms(A), IRREP(A)) of X^+ is X's block with
    the same labels (contract/contr_op1op2_wmaps_c.f:427-429,1127-1146 flip
    mst and take IRREP(C) in that case)."""

    transposed = True

    def __init__(self, mel):
        if mel.njoined != 1:
            raise ValueError('Adjoint: one-vertex lists only')
        self.base = mel
        self.orbs = mel.orbs
        self.njoined = 1
        self.occ_c, self.occ_a = dict(mel.occ_a), dict(mel.occ_c)
        self.oc, self.hc, self.oa, self.ha = list(mel.oa), list(mel.ha), list(mel.oc), list(mel.hc)
        self.nc, self.na = mel.na, mel.nc
        self.absym = mel.absym
        self.length = mel.length

    def find(self, msa_tot, igama, imc, gc, ima, ga):
        d = self.base.find(msa_tot, igama, ima, ga, imc, gc)
        if d is None:
            return None
        nc = len(self.base.oc)
        return dict(off=d['off'], len=d['len'], lens=d['lens'][nc:] + d['lens'][:nc])

    def block_len(self, msa_tot, igama):
        return self.base.block_len(msa_tot, igama)
