"""Writes tests/golden/hf_hf_molpro_FCIDUMP.gz and hf_hf_molpro.json: the
MOLPRO FCIDUMP of the reference's test molecule HF (R = 1.5, cc-pVDZ,
test/molecules/HF-HF-MOLPRO/FCIDUMP, read by the reference test
test/python-interface.s.chk) and that test's known answer, the frozen-core
MP2 correlation energy -0.196151291382 Eh (test/python-interface.s.ref:629),
with its orbital spaces (closed (3,1,1,0), frozen (1,0,0,0);
test/python-interface.s.ref:41-44).

Run in the development container only (it reads the reference tree); the
fixture it writes is what the tests and the GPU box use.
"""
import gzip
import json
import os

REF = '/root/reference/test/molecules/HF-HF-MOLPRO/FCIDUMP'
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    data = open(REF, 'rb').read()
    with gzip.GzipFile(os.path.join(HERE, 'hf_hf_molpro_FCIDUMP.gz'), 'wb', mtime=0) as f:
        f.write(data)
    json.dump({'closed': [3, 1, 1, 0], 'frozen': [1, 0, 0, 0], 'e_mp2_ref': -0.196151291382,
               'source': 'test/python-interface.s.ref:629 (final energy, MP2 target)'},
              open(os.path.join(HERE, 'hf_hf_molpro.json'), 'w'), indent=1)
    print('wrote', len(data), 'bytes (gzip)')


if __name__ == '__main__':
    main()
