"""Writes tests/golden/ch2_dz_SIRIFC.gz, ch2_dz_MOTWOINT.gz and ch2_dz.json:
the DALTON files of the reference's test molecule CH2/DZ
(test/molecules/CH2-DZ-DALTON/{SIRIFC,MOTWOINT}, read by the reference's
CC tests) and the reference's CCSD energy for it (test/CCPT-energy.e.ref:690,
'CC energy : -38.907618798245', CCSD with the automatic frozen core of
input/read_env_dalton.f: orbitals below -2.5 Eh).  The files are read by the
product reader gecco-public_amd/dalton.py, which checks its decoding against
the SCF energy and orbital energies the files hold.

Run in the development container only (it reads the reference tree); the
fixtures it writes are what the tests and the GPU box use.
"""
import gzip
import json
import os

REF = '/root/reference/test/molecules/CH2-DZ-DALTON'
HERE = os.path.dirname(os.path.abspath(__file__))
E_CCSD_REF = -38.907618798245        # test/CCPT-energy.e.ref:690
E_CCSDT_REF = -38.859570692641       # test/CC-energy-I.e.ref:606


def main():
    for name in ('SIRIFC', 'MOTWOINT'):
        data = open(os.path.join(REF, name), 'rb').read()
        with gzip.GzipFile(os.path.join(HERE, f'ch2_dz_{name}.gz'), 'wb', mtime=0) as f:
            f.write(data)
    json.dump({'e_ccsd_ref': E_CCSD_REF, 'source': 'test/CCPT-energy.e.ref:690 (CC energy)',
               'e_ccsdt_ref': E_CCSDT_REF,
               'e_ccsdt_source': 'test/CC-energy-I.e.ref:606 (CC maxexc=3, frozen (1,0,0,0), '
                                 'deleted (2,1,0,0))'},
              open(os.path.join(HERE, 'ch2_dz.json'), 'w'), indent=1)
    print('wrote', HERE)


if __name__ == '__main__':
    main()
