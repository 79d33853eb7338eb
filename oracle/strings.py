"""Oracle: spin-orbital strings, graph addressing and string maps.

TEST INFRASTRUCTURE ONLY.  This module is a CPU restatement of GeCCo's
string machinery, used by tests/, the smoke check and the CPU-baseline leg
of bench.py to check the HIP product path.  Nothing in the product imports it.

Restated routines (reference file:line):
  * lexically lowest string   strings/lexlstr.f:2-145      -> lexlstr()
  * next string               strings/nextstr.f:2-333      -> nextstr()
  * string iterator           strings/next_string.f:2-201  -> strings_of()
    (single subspace per space type: the subspace-distribution loop of
    next_string.f reduces to one pass)
  * string IRREP              strings/igamstr2.f           -> igamstr()
  * merge with sign           strings/iordstr2.f:2-119     -> iordstr2()
  * 4-slope graph weights     strings/weight_4sg.f:2-231   -> weight_4sg()
  * graph string index        strings/idx4sg_sb.f:2-68     -> idx4sg()
  * (str1,str2)->str12 map    strings/set_strmap_kernel.f:55-118 -> strmap()
  * spin-flip map             strings/set_flipmap_kernel.f, std_spsign
                              (strings/std_spsign_msdis.f)  -> flipmap()
Spin labels follow GeCCo: +1 alpha, -1 beta, 2 = both electrons of a doubly
occupied orbital (which appears twice in the orbital list).  Ms is counted in
units of 1/2: ms = nel - 2*(idxms-1)  (strings/set_strmap_kernel.f:55-56).

Parity status: pinned.  tests/test_oracle_strings.py checks the enumeration
order against the independent graph index idx4sg (weight_4sg restated), string
counts against closed-form binomials, and the merge sign against the parity of
an explicit sort of the concatenated spin-orbital list.
"""
from functools import lru_cache
from math import comb


def multd2h(a, b):
    """D2h IRREP product, 1-based labels (include/multd2h.h: XOR table)."""
    return ((a - 1) ^ (b - 1)) + 1


class Space:
    """One orbital space (H or P) with orbitals sorted by IRREP.

    GeCCo numbers the orbitals of a space contiguously with IRREP 1 first
    (mostnd(1:2,igam,ispc), used by strings/lexlstr.f:60,80 and
    strings/nextstr.f:227).  Orbital numbers here are 1-based within the space.
    """

    def __init__(self, irreps, nsym):
        self.nsym = nsym
        self.irreps = sorted(irreps)          # irrep of orbital k is irreps[k-1]
        self.norb = len(self.irreps)

    @property
    def first(self):
        return 1

    @property
    def last(self):
        return self.norb

    def gam(self, orb):
        return self.irreps[orb - 1]


def lexlstr(nel, itms, space):
    """strings/lexlstr.f:2-145 for one subspace (nspc=1, ioss=(nel,)).

    Returns (succ, idorb, idspn) with 1-based lists padded at index 0."""
    idorb = [0] * (nel + 2)
    idspn = [0] * (nel + 2)
    succ = True
    npair = (nel - abs(itms)) // 2
    nalpht = npair
    nbetat = npair
    if itms > 0:
        nalpht += itms
    if itms < 0:
        nbetat -= itms
    npair = 0
    maxnum = space.last - space.first + 1
    lenstr = nel
    npair_res = npair
    if maxnum < lenstr:
        npair4ss = lenstr - maxnum
        if npair4ss > maxnum:
            succ = False
        npair += npair4ss
    if succ and nel > 0:
        inum = space.first
        ipos = 1
        npair = min(nalpht, nbetat) - npair_res
        if npair < 0:
            return False, idorb, idspn
        npair4ss = min(npair, nel // 2)
        nalpht -= npair4ss
        nbetat -= npair4ss
        for _ in range(npair4ss):
            idorb[ipos] = idorb[ipos + 1] = inum
            idspn[ipos] = idspn[ipos + 1] = 2
            ipos += 2
            inum += 1
        nlone = nel - 2 * npair4ss
        nalph4ss = min(nalpht, nlone)
        nbeta4ss = nlone - nalph4ss
        if min(nalpht - nalph4ss, nbetat - nbeta4ss) < npair_res:
            npair = npair_res - min(nalpht - nalph4ss, nbetat - nbeta4ss)
            nalph4ss -= npair
            nbeta4ss = nalph4ss + npair
        nalpht -= nalph4ss
        nbetat -= nbeta4ss
        for _ in range(nbeta4ss):
            idorb[ipos] = inum
            idspn[ipos] = -1
            inum += 1
            ipos += 1
        for _ in range(nalph4ss):
            idorb[ipos] = inum
            idspn[ipos] = +1
            inum += 1
            ipos += 1
        if inum - 1 > space.last:
            succ = False
    return succ, idorb, idspn


def _lexlstr_into(nelsb, ms_sub, space, idorb, idspn):
    succ, o, s = lexlstr(nelsb, ms_sub, space)
    if succ:
        for i in range(1, nelsb + 1):
            idorb[i] = o[i]
            idspn[i] = s[i]
    return succ


def nextstr(nel, idorb, idspn, space):
    """strings/nextstr.f:147-333 (in place on 1-based lists)."""
    ipos = 1
    ms_pos = 0
    ipairc = 0
    lnext = True
    lsame = False
    idspnn = 0
    while True:
        idorbc = idorb[ipos]
        idspnc = idspn[ipos]
        if not lsame:
            if idspnc == 2 and ipairc == 0:
                ipos += 1
                ipairc = 1
                continue
            elif idspnc == 2 and ipairc == 1:
                ipairc = 0
                idspnc = -1
            else:
                ms_pos += idspnc
        else:
            lsame = False
        nalph = (ipos + ms_pos) // 2
        nbeta = (ipos - ms_pos) // 2
        iorbmax = space.last
        idorbn = idorb[ipos + 1] if ipos < nel else 100000
        if ipos < nel:
            idspnn = idspn[ipos + 1]
        if idspnc == +1 and nbeta > 0:
            idspn[ipos] = -1
            ms_sub = ms_pos + 1
        elif (ipos == nel or idorbc < idorbn - 1) and idorbc < iorbmax:
            idorb[ipos] += 1
            if idspnc == -1 and nalph > 0:
                idspn[ipos] = +1
                ms_sub = ms_pos - 1
            else:
                ms_sub = ms_pos - idspnc
        elif (idorbc < idorbn and idorbc < iorbmax and idspnn == -1
              and nalph > 0):
            idorb[ipos] += 1
            idspn[ipos] = 2
            idspn[ipos + 1] = 2
            ms_sub = ms_pos - 1
        elif ipos == nel:
            lnext = False
            break
        else:
            ipos += 1
            continue
        if ipos == 1:
            break
        nelsb = ipos - 1
        lnext = _lexlstr_into(nelsb, ms_sub, space, idorb, idspn)
        if lnext and idspn[nelsb + 1] == -1 and idorb[nelsb] == idorb[nelsb + 1]:
            idspn[nelsb] = idspn[nelsb + 1] = 2
        if lnext:
            break
        lsame = True
    return lnext


def igamstr(orbs, space):
    """String IRREP: product over all occupied spin-orbitals (igamstr2)."""
    g = 1
    for o in orbs:
        g = multd2h(g, space.gam(o))
    return g


@lru_cache(maxsize=None)
def _all_strings_ms(nel, ms, space_key):
    space = _SPACES[space_key]
    if nel == 0:
        return [((), ())] if ms == 0 else []
    succ, idorb, idspn = lexlstr(nel, ms, space)
    out = []
    while succ:
        out.append((tuple(idorb[1:nel + 1]), tuple(idspn[1:nel + 1])))
        succ = nextstr(nel, idorb, idspn, space)
    return out


_SPACES = {}


def _key(space):
    k = (tuple(space.irreps), space.nsym)
    _SPACES[k] = space
    return k


def strings_of(space, nel, ms, igam):
    """All strings of (nel, ms, igam) in GeCCo order (next_string.f:254-384)."""
    if nel == 0:
        return [((), ())] if (ms == 0 and igam == 1) else []
    return [s for s in _all_strings_ms(nel, ms, _key(space))
            if igamstr(s[0], space) == igam]


@lru_cache(maxsize=None)
def _index_table(nel, ms, igam, space_key):
    space = _SPACES[space_key]
    return {s: i for i, s in enumerate(strings_of(space, nel, ms, igam))}


def string_index(space, nel, ms, igam, s):
    return _index_table(nel, ms, igam, _key(space))[s]


def lenstr(space, nel, idxms, igam):
    """Number of strings, graph%lenstr_gm(igam,idxms) (set_lenstr_array)."""
    ms = nel - 2 * (idxms - 1)
    return len(strings_of(space, nel, ms, igam))


def to_spinorbs(s):
    """(idorb, idspn) -> tuple of (orb, spin) with pairs expanded alpha,beta."""
    orbs, spns = s
    out = []
    i = 0
    while i < len(orbs):
        if spns[i] == 2:
            out.append((orbs[i], +1))
            out.append((orbs[i], -1))
            i += 2
        else:
            out.append((orbs[i], spns[i]))
            i += 1
    return tuple(out)


def from_spinorbs(so):
    """Inverse of to_spinorbs for a canonically ordered list."""
    orbs, spns = [], []
    i = 0
    while i < len(so):
        if i + 1 < len(so) and so[i][0] == so[i + 1][0]:
            orbs += [so[i][0]] * 2
            spns += [2, 2]
            i += 2
        else:
            orbs.append(so[i][0])
            spns.append(so[i][1])
            i += 1
    return tuple(orbs), tuple(spns)


def iordstr2(s1, s2):
    """strings/iordstr2.f:2-119.  Returns (sign, merged) ; sign 0 = Pauli."""
    o1, p1 = s1
    o2, p2 = s2
    n1, n2 = len(o1), len(o2)
    if n1 + n2 == 0:
        return 1, ((), ())
    if n2 == 0:
        return 1, (tuple(o1), tuple(p1))
    if n1 == 0:
        return 1, (tuple(o2), tuple(p2))
    o12, p12 = [], []
    i1 = i2 = 0
    iperm = 0
    for _ in range(n1 + n2):
        if i1 < n1 and i2 < n2:
            if o1[i1] == o2[i2] and p1[i1] * p2[i2] != -1:
                return 0, None
        if i2 >= n2 or (i1 < n1 and (o1[i1] < o2[i2] or
                                     (o1[i1] == o2[i2] and p1[i1] == +1))):
            o12.append(o1[i1])
            p12.append(p1[i1])
            i1 += 1
        else:
            o12.append(o2[i2])
            p12.append(p2[i2])
            i2 += 1
            iperm += n1 - i1
    for i in range(n1 + n2 - 1):
        if o12[i] == o12[i + 1]:
            p12[i] = p12[i + 1] = 2
    return (-1 if iperm % 2 else 1), (tuple(o12), tuple(p12))


# ---------------------------------------------------------------- graphs
def weight_4sg(space, nel, maxms):
    """strings/weight_4sg.f:2-231: arc weights iy4sg(ia,n,ms,gam,orb).

    Returned as dict keyed (ianum, nel_before, ms_before, gam_before, orb) with
    ianum 1:+1 arc, 2:-1 arc, 3:pair arc (ianum(-1:2)=(/2,0,1,3/) in
    strings/idx4sg_sb.f:28)."""
    nsym = space.nsym
    norb = space.norb
    iw = {}
    iy = {}

    def W(iel, ims, igam, iorb):
        if iorb == 0:
            return 1 if (iel == 0 and ims == 0 and igam == 1) else 0
        return iw.get((iel, ims, igam, iorb), 0)

    for iorb in range(1, norb + 1):
        g_orb = space.gam(iorb)
        for igam in range(1, nsym + 1):
            for ims in range(-maxms, maxms + 1):
                for iel in range(0, nel + 1):
                    prev = W(iel, ims, igam, iorb - 1)
                    if prev == 0:
                        continue
                    for ia in (-1, 0, 1, 2):
                        eln = iel + abs(ia)
                        if eln > nel:
                            continue
                        if abs(ia) == 1:
                            msn = ims + ia
                            gn = multd2h(igam, g_orb)
                        else:
                            msn = ims
                            gn = igam
                        if abs(msn) > maxms:
                            continue
                        key = (eln, msn, gn, iorb)
                        iw[key] = iw.get(key, 0) + prev
        for igam in range(1, nsym + 1):
            for ims in range(-maxms, maxms + 1):
                for iel in range(0, nel + 1):
                    if W(iel, ims, igam, iorb) == 0:
                        continue
                    acc = W(iel, ims, igam, iorb - 1)
                    for ianum, ia in ((1, +1), (2, -1), (3, 2)):
                        elp = iel - abs(ia)
                        if elp < 0:
                            continue
                        if abs(ia) == 1:
                            msp = ims - ia
                            gp = multd2h(igam, g_orb)
                        else:
                            msp = ims
                            gp = igam
                        if abs(msp) > maxms:
                            continue
                        iy[(ianum, elp, msp, gp, iorb)] = acc
                        acc += W(elp, msp, gp, iorb - 1)
    return iy


def idx4sg(s, space, iy):
    """strings/idx4sg_sb.f:41-64: zero-based graph index of a string."""
    orbs, spns = s
    ianum = {-1: 2, 1: 1, 2: 3}
    idx = 0
    nn = 0
    nm = 0
    ig = 1
    i = 0
    while i < len(orbs):
        sp = spns[i]
        idx += iy.get((ianum[sp], nn, nm, ig, orbs[i]), 0)
        nn += abs(sp)
        if abs(sp) == 1:
            nm += sp
            ig = multd2h(ig, space.gam(orbs[i]))
        i += 2 if sp == 2 else 1
    return idx


# ------------------------------------------------------------------ maps
def strmap(space, nel1, idxms1, igam1, nel2, idxms2, igam2):
    """strings/set_strmap_kernel.f:55-118.

    Entry (str1 fastest) = sign*(1+index of merged string in its own
    (ms,IRREP) block of graph (space, nel1+nel2)), 0 if Pauli-forbidden."""
    ms1 = nel1 - 2 * (idxms1 - 1)
    ms2 = nel2 - 2 * (idxms2 - 1)
    nel12 = nel1 + nel2
    out = []
    for s2 in strings_of(space, nel2, ms2, igam2):
        for s1 in strings_of(space, nel1, ms1, igam1):
            sgn, s12 = iordstr2(s1, s2)
            if sgn == 0:
                out.append(0)
                continue
            g12 = multd2h(igam1, igam2)
            idx = string_index(space, nel12, ms1 + ms2, g12, s12)
            out.append(sgn * (idx + 1))
    return out


def std_spsign(spns):
    """strings/std_spsign_msdis.f: sign to bring a spin string to a..ab..b."""
    trp = 0
    nb = 0
    i = 0
    while i < len(spns):
        sig = spns[i]
        if sig in (1, 2):
            trp += nb
            i += 1
        if sig in (-1, 2):
            nb += 1
            i += 1
    return -1 if trp % 2 else 1


def std_spsign_msdis(msdis, occ):
    """strings/std_spsign_msdis.f: standard sign of an Ms distribution."""
    if len(msdis) <= 1:
        return 1
    scr = []
    for ms, n in zip(msdis, occ):
        scr += [1] * ((n + ms) // 2) + [-1] * ((n - ms) // 2)
    return std_spsign(scr)


@lru_cache(maxsize=None)
def _flipmap_cached(space_key, nel, idxms, igam):
    return tuple(_flipmap(_SPACES[space_key], nel, idxms, igam))


def flipmap(space, nel, idxms, igam):
    """strings/set_flipmap_kernel.f (cached)."""
    return list(_flipmap_cached(_key(space), nel, idxms, igam))


def _flipmap(space, nel, idxms, igam):
    """strings/set_flipmap_kernel.f: string -> spin-flipped string, signed."""
    ms = nel - 2 * (idxms - 1)
    out = []
    for s in strings_of(space, nel, ms, igam):
        orbs, spns = s
        fl = tuple(p if p == 2 else -p for p in spns)
        sgn = std_spsign(spns) * std_spsign(fl)
        idx = string_index(space, nel, -ms, igam, (orbs, fl))
        out.append(sgn * (idx + 1))
    return out


def n_strings_closed_form(norb_by_irrep, nel, ms, igam, nsym):
    """Independent count: number of (alpha-set, beta-set) pairs with the right
    sizes and IRREP, for checking lenstr()."""
    na = (nel + ms) // 2
    nb = (nel - ms) // 2
    orbs = []
    for g, n in enumerate(norb_by_irrep, start=1):
        orbs += [g] * n
    from itertools import combinations

    def irrep_counts(k):
        cnt = {g: 0 for g in range(1, nsym + 1)}
        for c in combinations(range(len(orbs)), k):
            gg = 1
            for i in c:
                gg = multd2h(gg, orbs[i])
            cnt[gg] += 1
        return cnt
    ca = irrep_counts(na)
    cb = irrep_counts(nb)
    tot = 0
    for ga in range(1, nsym + 1):
        tot += ca[ga] * cb[multd2h(ga, igam)]
    return tot


__all__ = [n for n in dir() if not n.startswith('__')] + ['comb']
