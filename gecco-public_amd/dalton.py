"""Integral import from DALTON's SIRIFC and MOTWOINT files into the CCSD
driver.

File formats, restated from the reference's readers:
  * Fortran sequential unformatted records with 4-byte length markers;
    label records are four character*8 words '********', date, time, label
    (input/mollab.f).
  * SIRIFC: after label 'SIR IPH ': potnuc, emy, eactiv, emcscf (real*8),
    istate, ispin, nactel, lsym (int*4); the next record holds nsym and the
    orbital counts per IRREP (input/read_env_dalton.f:57-62); the inactive
    Fock matrix, IRREP-blocked packed triangles, is the 6th record after the
    label (input/read_fock_from_sirifc.f:30-40, nskip_in_sirifc = 6).  After
    label 'TRCCINT ' the second record holds the orbital energies
    (input/read_env_dalton.f:66-68).
  * MOTWOINT: record 2 = (lbuf, itrlevel); after label 'MOLTWOEL' records of
    xbuf(lbuf) real*8, ibuf(lbuf) int*4, len int*4 (len < 0 ends); ibuf packs
    (rs) in the high and (pq) in the low 16 bits, each pair as 256*p + q with
    DALTON's IRREP-ordered orbital numbers (input/import_h2_dalton.f:90-140).
Frozen core: orbitals whose energy is below -2.5 Eh, as the automatic frozen
core of input/read_env_dalton.f:12 (thr_auto_freeze).  Files may be gzipped.
"""
import gzip
import struct

import numpy as np

from synth import DenseHamiltonian

FREEZE_BELOW = -2.5


def _read(path):
    opener = gzip.open if str(path).endswith('.gz') else open
    with opener(path, 'rb') as f:
        return f.read()


def records(path):
    """Records of a Fortran sequential unformatted file."""
    data = _read(path)
    pos, out = 0, []
    while pos < len(data):
        n = struct.unpack_from('<i', data, pos)[0]
        if n < 0 or pos + 8 + n > len(data) or struct.unpack_from('<i', data, pos + 4 + n)[0] != n:
            raise ValueError(f'{path}: bad record marker at byte {pos}')
        out.append(data[pos + 4:pos + 4 + n])
        pos += 8 + n
    return out


def _after_label(recs, label):
    for i, r in enumerate(recs):
        if len(r) == 32 and r[:8] == b'********' and r[24:32] == label:
            return i + 1
    raise KeyError(label)


def read_sirifc(path):
    """Energies, orbital counts, packed inactive Fock matrix and orbital
    energies from SIRIFC."""
    recs = records(path)
    i = _after_label(recs, b'SIR IPH ')
    potnuc, emy, eactiv, emcscf = struct.unpack_from('<4d', recs[i], 0)
    istate, ispin, nactel, lsym = struct.unpack_from('<4i', recs[i], 32)
    ints = np.frombuffer(recs[i + 1], dtype='<i4')
    nsym = int(ints[14])
    tail = ints[15 + 64:]
    nrhf, norb = tail[0:8][:nsym].copy(), tail[32:40][:nsym].copy()
    nfock = int(sum(n * (n + 1) // 2 for n in norb))
    fock_packed = np.frombuffer(recs[i + 6][:8 * nfock], dtype='<f8').copy()
    j = _after_label(recs, b'TRCCINT ')
    norbt = int(norb.sum())
    eorb = np.frombuffer(recs[j + 1][:8 * norbt], dtype='<f8').copy()
    return dict(potnuc=potnuc, emy=emy, emcscf=emcscf, ispin=ispin, nactel=nactel, nsym=nsym,
                nrhf=nrhf, norb=norb, fock_packed=fock_packed, eorb=eorb)


def unpack_fock(packed, norb):
    """IRREP blocks, each a packed triangle (row-wise lower = column-wise
    upper), into the full matrix."""
    n = int(norb.sum())
    F = np.zeros((n, n))
    k = off = 0
    for nb in norb:
        for p in range(nb):
            for q in range(p + 1):
                F[off + p, off + q] = F[off + q, off + p] = packed[k]
                k += 1
        off += nb
    return F


def read_motwoint(path, n):
    """(pq|rs) over the n orbitals from MOTWOINT (8-fold symmetry filled);
    returns the array and the number of records read."""
    recs = records(path)
    lbuf = int(np.frombuffer(recs[1][:4], dtype='<i4')[0])
    i = _after_label(recs, b'MOLTWOEL')
    eri = np.zeros((n, n, n, n))
    count = 0
    for r in recs[i:]:
        x = np.frombuffer(r[:8 * lbuf], dtype='<f8')
        ib = np.frombuffer(r[8 * lbuf:12 * lbuf], dtype='<i4')
        ln = int(np.frombuffer(r[12 * lbuf:12 * lbuf + 4], dtype='<i4')[0])
        if ln < 0:
            break
        for v, lab in zip(x[:ln], ib[:ln]):
            rs = (int(lab) >> 16) & 0xFFFF
            pq = int(lab) & 0xFFFF
            p, q = (pq >> 8) - 1, (pq & 0xFF) - 1
            r_, s = (rs >> 8) - 1, (rs & 0xFF) - 1
            for a, b, c, d in ((p, q, r_, s), (q, p, r_, s), (p, q, s, r_), (q, p, s, r_),
                               (r_, s, p, q), (s, r_, p, q), (r_, s, q, p), (s, r_, q, p)):
                eri[a, b, c, d] = v
            count += 1
    return eri, count


def hamiltonian_from_dalton(sirifc, motwoint, freeze_below=FREEZE_BELOW, check_tol=1e-9, frozen=None,
                            deleted=None):
    """Closed-shell DenseHamiltonian for the CCSD driver from SIRIFC and
    MOTWOINT.  The SCF energy rebuilt from the Fock diagonal and the
    integrals, E = E_nuc + sum_i (2 f_ii - sum_j (2 (ii|jj) - (ij|ji))), and
    the Fock diagonal against SIRIFC's orbital energies check the decoding
    (ValueError beyond check_tol).  Occupied orbitals of energy below
    freeze_below are frozen, or, with frozen = counts per IRREP, the lowest
    frozen[s] orbitals of IRREP s (the orb_space shell type=frozen,def=(...)
    of the reference's input); deleted = counts per IRREP drops the highest
    deleted[s] orbitals of IRREP s (shell type=deleted).  The correlated
    orbitals are the other occupied ones and the remaining virtual ones, in
    file (IRREP) order."""
    sir = read_sirifc(sirifc)
    norb, nrhf, nsym = sir['norb'], sir['nrhf'], sir['nsym']
    n = int(norb.sum())
    irrep = np.concatenate([[s + 1] * int(norb[s]) for s in range(nsym)])
    F = unpack_fock(sir['fock_packed'], norb)
    eri, _ = read_motwoint(motwoint, n)
    occ = np.concatenate([np.arange(int(norb[:s].sum()), int(norb[:s].sum()) + int(nrhf[s]))
                          for s in range(nsym)]).astype(int)
    e_scf = sir['potnuc']
    for i in occ:
        e_scf += 2 * F[i, i] - sum(2 * eri[i, i, j, j] - eri[i, j, j, i] for j in occ)
    if abs(e_scf - sir['emcscf']) > check_tol:
        raise ValueError(f'integral decoding gives E(SCF) = {e_scf:.12f}, SIRIFC holds '
                         f'{sir["emcscf"]:.12f}')
    if np.abs(np.diag(F) - sir['eorb']).max() > check_tol:
        raise ValueError('Fock diagonal differs from the orbital energies')
    start = [int(norb[:s].sum()) for s in range(nsym)]
    if frozen is None:
        frozen = [i for i in occ if F[i, i] < freeze_below]
    else:
        frozen = [start[s] + k for s in range(nsym) for k in range(int(frozen[s]) if s < len(frozen) else 0)]
    gone = set()
    if deleted is not None:
        gone = {start[s] + int(norb[s]) - 1 - k for s in range(min(nsym, len(deleted)))
                for k in range(int(deleted[s]))}
        if gone & set(occ):
            raise ValueError('deleted orbitals overlap the occupied ones')
    act_o = [int(i) for i in occ if i not in frozen]
    act_v = [p for p in range(n) if p not in set(occ) and p not in gone]
    act = act_o + act_v
    return DenseHamiltonian(np.diag(F)[act_o], np.diag(F)[act_v], irrep[act_o], irrep[act_v],
                            eri[np.ix_(act, act, act, act)], nsym, sir['emcscf'])
