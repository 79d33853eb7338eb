"""Integral import from a MOLPRO FCIDUMP file into the CCSD driver.

Restates the reference's FCIDUMP path:
  input/read_env_molpro_dump.f:37-80      namelist &FCI (NORB, NELEC, MS2,
                                          ORBSYM, ISYM), IRREP count from
                                          ORBSYM (3 -> 4, > 4 -> 8)
  input/import_h2_molpro_dump.f           records "value i j k l": (ij|kl)
                                          for k, l > 0, h_ij for k = l = 0,
                                          the core energy for i = j = 0
  input/read_fock_from_molpro_dump.f:30-40
  input/import_fock_molpro_dump.f         Fock operator = h + the closed-shell
                                          two-electron terms

Orbitals are numbered as in the file, grouped by IRREP (ORBSYM), and the
closed, frozen and occupied shells of an IRREP are its first orbitals (the
orb_space shell definitions of a GeCCo input, e.g. test/python-interface.s.ref:
41-44).  MOLPRO numbers the IRREPs of D2h and its subgroups so that the
direct product is the XOR of (label - 1), as the driver's IRREP table
assumes.
"""
import gzip
import re

import numpy as np

from synth import DenseHamiltonian


def _open(path):
    return gzip.open(path, 'rt') if str(path).endswith('.gz') else open(path)


def read_fcidump(path):
    """Header and integrals of an FCIDUMP file (plain or gzip).  Returns a
    dict with norb, nelec, ms2, isym, orbsym (array of IRREP labels), nsym,
    h (norb x norb), eri ((pq|rs), norb^4, 8-fold symmetric) and ecore."""
    with _open(path) as f:
        text = f.read()
    m = re.search(r'&FCI(.*?)(?:&END|/)', text, re.S | re.I)
    if not m or not text.lstrip().upper().startswith('&FCI'):
        raise ValueError(f'{path}: not an FCIDUMP file (no leading &FCI namelist)')
    head, body = m.group(1), text[m.end():]
    # "KEY=v1,v2,..., KEY2=..." : split at the keys
    parts = re.split(r'([A-Za-z_][A-Za-z0-9_]*)\s*=', head)
    fields = {}
    for key, val in zip(parts[1::2], parts[2::2]):
        fields[key.upper()] = [int(x) for x in re.split(r'[,\s]+', val.strip()) if x]
    norb = fields['NORB'][0]
    nelec = fields['NELEC'][0]
    ms2 = fields.get('MS2', [0])[0]
    isym = fields.get('ISYM', [1])[0]
    orbsym = np.array(fields.get('ORBSYM', [1] * norb)[:norb], dtype=np.int64)
    if len(orbsym) != norb:
        raise ValueError(f'{path}: ORBSYM has {len(orbsym)} entries for NORB = {norb}')
    nsym = int(orbsym.max())
    nsym = 4 if nsym == 3 else (8 if nsym > 4 else nsym)
    h = np.zeros((norb, norb))
    eri = np.zeros((norb, norb, norb, norb))
    ecore = 0.0
    rec = np.array(body.split(), dtype=object)
    if len(rec) % 5:
        raise ValueError(f'{path}: integral records are not "value i j k l"')
    rec = rec.reshape(-1, 5)
    vals = rec[:, 0].astype(float)
    idx = rec[:, 1:].astype(np.int64)
    for v, (i, j, k, l) in zip(vals, idx):
        if k or l:
            i, j, k, l = i - 1, j - 1, k - 1, l - 1
            for a, b, c, d in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                               (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
                eri[a, b, c, d] = v
        elif i and j:
            h[i - 1, j - 1] = h[j - 1, i - 1] = v
        elif i:
            continue                  # orbital energy record: the Fock matrix is built below
        else:
            ecore += v
    return dict(norb=norb, nelec=nelec, ms2=ms2, isym=isym, orbsym=orbsym, nsym=nsym, h=h, eri=eri,
                ecore=ecore)


def _first_of_irrep(orbsym, counts):
    """Indices of the first counts[s] orbitals of each IRREP s+1."""
    out = []
    for s, n in enumerate(counts):
        idx = np.flatnonzero(orbsym == s + 1)
        if len(idx) < n:
            raise ValueError(f'IRREP {s + 1} has {len(idx)} orbitals, {n} requested')
        out += list(idx[:n])
    return sorted(out)


def hamiltonian_from_fcidump(path, closed, frozen=None, canonical_tol=1e-6):
    """Closed-shell DenseHamiltonian for the CCSD driver.

    closed[s] / frozen[s]: closed-shell / frozen-core orbitals of IRREP s+1
    (the first ones of that IRREP).  The Fock operator is h plus the
    closed-shell two-electron terms, f_pq = h_pq + sum_i [2 (pq|ii) - (pi|iq)];
    the reference energy is ecore + sum_i (h_ii + f_ii).  The correlated
    orbitals are the closed non-frozen ones (occupied) and all others
    (virtual), each ordered by IRREP; their orbital energies are the Fock
    diagonal, and off-diagonal Fock elements among them above canonical_tol
    are an error (the driver assumes canonical orbitals)."""
    d = read_fcidump(path)
    orbsym, h, eri = d['orbsym'], d['h'], d['eri']
    nsym = d['nsym']
    closed = list(closed) + [0] * (nsym - len(closed))
    frozen = list(frozen or [0] * nsym) + [0] * (nsym - len(frozen or []))
    if 2 * sum(closed) != d['nelec'] or d['ms2'] != 0:
        raise ValueError(f'closed shells {closed} do not hold the {d["nelec"]} electrons of a '
                         'closed-shell (MS2 = 0) state')
    occ = _first_of_irrep(orbsym, closed)
    core = _first_of_irrep(orbsym, frozen)
    if not set(core) <= set(occ):
        raise ValueError('frozen orbitals must be closed shells')
    fock = h.copy()
    for i in occ:
        fock += 2.0 * eri[:, :, i, i] - eri[:, i, i, :]
    e_ref = d['ecore'] + sum(h[i, i] + fock[i, i] for i in occ)
    act_o = [i for i in occ if i not in core]
    act_v = [p for p in range(d['norb']) if p not in occ]
    act = act_o + act_v
    off = fock[np.ix_(act, act)] - np.diag(np.diag(fock)[act])
    if np.abs(off).max() > canonical_tol:
        raise ValueError(f'orbitals are not canonical: max |f_pq| (p != q) = {np.abs(off).max():.2e}')
    return DenseHamiltonian(np.diag(fock)[act_o], np.diag(fock)[act_v], orbsym[act_o], orbsym[act_v],
                            eri[np.ix_(act, act, act, act)], nsym, e_ref)
