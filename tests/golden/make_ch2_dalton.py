"""Writes tests/golden/ch2_dz_dalton.npz: the CH2/DZ Hamiltonian of the
reference's test molecule (test/molecules/CH2-DZ-DALTON/{SIRIFC,MOTWOINT},
DALTON files that the reference's own tests read), reduced to what the CCSD
engine needs, plus the reference's CCSD energy for it
(test/CCPT-energy.e.ref:690, 'CC energy : -38.907618798245', CCSD with the
automatic frozen core of input/read_env_dalton.f: orbitals below -2.5 Eh).

Run in the development container only (it reads the reference tree); the
fixture it writes is what the tests and the GPU box use.

File formats, restated from the reference's readers:
  * Fortran sequential unformatted records with 4-byte length markers;
    label records are four character*8 words '********', date, time, label
    (input/mollab.f).
  * SIRIFC: after label 'SIR IPH ': potnuc, emy, eactiv, emcscf (real*8),
    istate, ispin, nactel, lsym (int*4); the next record holds nsym, the
    orbital counts per IRREP (input/read_env_dalton.f:57-62); the inactive
    Fock matrix, IRREP-blocked packed triangles, is the 6th record after the
    label (input/read_fock_from_sirifc.f:30-40, nskip_in_sirifc = 6).
    After label 'TRCCINT ' the second record holds the orbital energies
    (input/read_env_dalton.f:66-68).
  * MOTWOINT: record 2 = (lbuf, itrlevel); after label 'MOLTWOEL' records of
    xbuf(lbuf) real*8, ibuf(lbuf) int*4, len int*4 (len < 0 ends); ibuf
    packs (rs) in the high and (pq) in the low 16 bits, each pair as
    256*p + q with DALTON's IRREP-ordered orbital numbers
    (input/import_h2_dalton.f:90-140).
"""
import os
import struct
import sys

import numpy as np

REF = '/root/reference/test/molecules/CH2-DZ-DALTON'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'ch2_dz_dalton.npz')
E_CCSD_REF = -38.907618798245        # test/CCPT-energy.e.ref:690
FREEZE_BELOW = -2.5                   # input/read_env_dalton.f:12 (thr_auto_freeze)


def records(path):
    data = open(path, 'rb').read()
    pos, out = 0, []
    while pos < len(data):
        n = struct.unpack_from('<i', data, pos)[0]
        if struct.unpack_from('<i', data, pos + 4 + n)[0] != n:
            raise ValueError(f'{path}: bad record marker at byte {pos}')
        out.append(data[pos + 4:pos + 4 + n])
        pos += 8 + n
    return out


def after_label(recs, label):
    for i, r in enumerate(recs):
        if len(r) == 32 and r[:8] == b'********' and r[24:32] == label:
            return i + 1
    raise KeyError(label)


def read_sirifc(path):
    recs = records(path)
    i = after_label(recs, b'SIR IPH ')
    potnuc, emy, eactiv, emcscf = struct.unpack_from('<4d', recs[i], 0)
    istate, ispin, nactel, lsym = struct.unpack_from('<4i', recs[i], 32)
    ints = np.frombuffer(recs[i + 1], dtype='<i4')
    nsym = int(ints[14])
    tail = ints[15 + 64:]
    nrhf, norb = tail[0:8][:nsym].copy(), tail[32:40][:nsym].copy()
    nfock = int(sum(n * (n + 1) // 2 for n in norb))
    fock_packed = np.frombuffer(recs[i + 6][:8 * nfock], dtype='<f8').copy()
    j = after_label(recs, b'TRCCINT ')
    norbt = int(norb.sum())
    eorb = np.frombuffer(recs[j + 1][:8 * norbt], dtype='<f8').copy()
    return dict(potnuc=potnuc, emy=emy, emcscf=emcscf, ispin=ispin, nactel=nactel, nsym=nsym,
                nrhf=nrhf, norb=norb, fock_packed=fock_packed, eorb=eorb)


def unpack_fock(packed, norb):
    """IRREP blocks, each a packed triangle (row-wise lower = column-wise
    upper)."""
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
    recs = records(path)
    lbuf = int(np.frombuffer(recs[1][:4], dtype='<i4')[0])
    i = after_label(recs, b'MOLTWOEL')
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


def main():
    sir = read_sirifc(os.path.join(REF, 'SIRIFC'))
    norb, nrhf, nsym = sir['norb'], sir['nrhf'], sir['nsym']
    n = int(norb.sum())
    irrep = np.concatenate([[s + 1] * int(norb[s]) for s in range(nsym)])
    F = unpack_fock(sir['fock_packed'], norb)
    eri, nint = read_motwoint(os.path.join(REF, 'MOTWOINT'), n)
    occ = np.concatenate([np.arange(int(norb[:s].sum()), int(norb[:s].sum()) + int(nrhf[s]))
                          for s in range(nsym)])
    # closed-shell energy from the Fock diagonal and the integrals:
    # E = E_nuc + sum_i (2 f_ii - sum_j (2 (ii|jj) - (ij|ji)))
    e_scf = sir['potnuc']
    for i in occ:
        e_scf += 2 * F[i, i] - sum(2 * eri[i, i, j, j] - eri[i, j, j, i] for j in occ)
    print(f'{nint} integrals; E(SCF) from Fock+integrals {e_scf:.12f}, SIRIFC emcscf '
          f'{sir["emcscf"]:.12f}, diff {e_scf - sir["emcscf"]:.2e}')
    if abs(e_scf - sir['emcscf']) > 1e-9:
        sys.exit('integral decoding does not reproduce the SCF energy')
    if np.abs(np.diag(F) - sir['eorb']).max() > 1e-9:
        sys.exit('Fock diagonal differs from the orbital energies')
    off = F - np.diag(np.diag(F))
    frozen = [i for i in occ if F[i, i] < FREEZE_BELOW]
    act_o = [i for i in occ if i not in frozen]
    act_v = [i for i in range(n) if i not in occ]
    act = act_o + act_v
    print(f'frozen {frozen}, active occupied {act_o}, virtual {act_v}; '
          f'max |off-diagonal Fock| {np.abs(off[np.ix_(act, act)]).max():.2e}')
    np.savez_compressed(
        OUT, eps_o=np.diag(F)[act_o], eps_v=np.diag(F)[act_v],
        irrep_o=irrep[act_o], irrep_v=irrep[act_v], nsym=nsym,
        eri=eri[np.ix_(act, act, act, act)], fock_offdiag=off[np.ix_(act, act)],
        e_scf=sir['emcscf'], e_ccsd_ref=E_CCSD_REF)
    print('wrote', OUT)


if __name__ == '__main__':
    main()
