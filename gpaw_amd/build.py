"""Build libgpaw_amd.so (all HIP sources under gpaw_amd/csrc) for gfx950.

    python -m gpaw_amd.build          # or gpaw_amd.build.build()

The library is built in-tree next to this file so that it travels with the
repository snapshot.  FP64 kernels must stay bit-identical to the
reference C, which is compiled without FMA contraction: the whole library
is built with -ffp-contract=off (FP32 kernels that may contract say so with
a pragma).  Packed-FP32 instructions are disabled for the device pass.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libgpaw_amd.so')
ARCH = os.environ.get('GPAW_AMD_ARCH', 'gfx950')

SOURCES = ['stencil.hip', 'stencil_tile_f64.hip', 'stencil_tile_c128.hip',
           'stencil_tile_f32.hip', 'stencil_tile_c64.hip', 'stencil_box_f64.hip',
           'stencil_box_c128.hip', 'stencil_box_f32.hip', 'stencil_box_c64.hip', 'transfer.hip',
           'transfer_f64.hip', 'transfer_c128.hip', 'transfer_f32.hip',
           'transfer_c64.hip',
           'misc.hip', 'capi.hip', 'poisson.hip', 'coarse.hip', 'mixed.hip', 'gs.hip',
           'precond.hip', 'batch.hip', 'batch_f64.hip', 'batch_c128.hip',
           'batch_f32.hip', 'batch_c64.hip', 'jacobi_fused.hip',
           'resid_restrict.hip', 'jacobi_corr.hip']

FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH,
         '-ffp-contract=off', '-Xclang', '-target-feature', '-Xclang',
         '-packed-fp32-ops', '-Wall', '-Wno-unused-function',
         '-Wno-unused-variable', '-Wno-unused-parameter', '-Wno-unused-value',
         '-Wno-unused-result']


def hipcc():
    for p in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc'):
        if p and os.path.exists(p):
            return p
    return 'hipcc'


def _compile(args):
    cmd, verbose = args
    if verbose:
        print(' '.join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = '\n'.join(l for l in (r.stderr or '').splitlines()
                    if 'packed-fp32-ops' not in l)
    return r.returncode, err


def build(verbose=False, jobs=None):
    """Compile every source to an object in parallel, then link the .so."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for s in SOURCES:
        o = os.path.join(objdir, s.replace('.hip', '.o'))
        objs.append(o)
        cmds.append(([hipcc()] + FLAGS + ['-c', '-I' + CSRC,
                                          os.path.join(CSRC, s), '-o', o],
                     verbose))
    jobs = jobs or min(len(cmds), os.cpu_count() or 4, 16)
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, cmds))
    for (rc, err), s in zip(results, SOURCES):
        if rc != 0:
            raise RuntimeError('hipcc failed on %s:\n%s' % (s, err))
        if verbose and err.strip():
            print(err)
    tmp = LIB + '.tmp'
    rc, err = _compile(([hipcc(), '-shared', '-fPIC', '--offload-arch=' + ARCH]
                        + objs + ['-o', tmp], verbose))
    if rc != 0:
        raise RuntimeError('link failed:\n' + err)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
