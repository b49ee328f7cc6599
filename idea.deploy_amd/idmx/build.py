"""Builds libidmx.so in-tree with hipcc for gfx950.

    python idea.deploy_amd/idmx/build.py

Each .hip translation unit is compiled to an object in parallel, then
linked into idea.deploy_amd/idmx/libidmx.so. The lattice tables header is
regenerated from idmx/LBM/stencils.py first.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
CSRC = os.path.join(PKG_ROOT, 'csrc')
OBJ = os.path.join(PKG_ROOT, '_build')
LIB = os.path.join(HERE, 'libidmx.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('IDMX_ARCH', 'gfx950')

FLAGS = ['-std=c++17', '-O3', '-fPIC', '--offload-arch=' + ARCH,
         # packed-FP32 instructions are kept out of the generated code
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-result'] + os.environ.get('IDMX_HIPCC_FLAGS', '').split()

# (source, extra flags). The FP64 lattice code (and the batch and runtime
# units, whose FP64 results must agree with it bit for bit) is compiled
# without floating-point contraction, as the reference's CTypes backend
# computes: with fused multiply-adds the FP64 trajectory drifts from it by
# 2e-11 over 10^4 steps, without by 1e-13. The FP32-shifted lattice units
# (single-domain and batched) and the network kernels keep contraction.
NOFMA = ['-ffp-contract=off']
SOURCES = [('runtime.hip', NOFMA), ('sc_d2q9.hip', NOFMA), ('sc_d2q9_f32.hip', []),
           ('sc_d3q19_e4.hip', NOFMA), ('sc_d3q19_e4_f32.hip', []),
           ('sc_d3q19_e6.hip', NOFMA), ('sc_d3q19_e6_f32.hip', []),
           ('sc_d3q19_e8.hip', NOFMA), ('sc_d3q19_e8_f32.hip', []),
           ('batch.hip', NOFMA), ('batch_f32.hip', []), ('nn.hip', []), ('noalloc.hip', NOFMA),
           ('init_fields.hip', NOFMA)]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError('command failed: %s\n%s%s' % (' '.join(cmd), p.stdout, p.stderr))
    return p


def build(verbose=True, jobs=None):
    subprocess.run([sys.executable, os.path.join(CSRC, 'gen_lattice.py')], check=True)
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src, extra in SOURCES:
        o = os.path.join(OBJ, src.replace('.hip', '.o'))
        objs.append(o)
        cmds.append([HIPCC] + FLAGS + extra + ['-c', os.path.join(CSRC, src), '-o', o])
    jobs = jobs or min(len(cmds), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for c in ex.map(_run, cmds):
            pass
    tmp = LIB + '.tmp'
    _run([HIPCC, '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build()
