"""Build libgecco_hip.so in-tree for gfx950 with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'libgecco_hip.so')
SOURCES = ['contract_kernels.hip', 'plan.cpp', 'enumerate.cpp', 'solver.hip', 'mlp.hip', 'op1op2.cpp',
           'collect_scatter.hip']
HIPFLAGS = ['--offload-arch=gfx950', '-O3', '-fPIC', '-shared', '-std=c++17',
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-munsafe-fp-atomics', '-Wall']


def build(verbose=False):
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    tmp = OUT + '.tmp'
    cmd = [hipcc] + HIPFLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ['-o', tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError('hipcc failed for libgecco_hip.so')
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose=True))
