"""Build the gx HIP library (libgx.so) for gfx950, in-tree.

``python -m gaussium_amd.build`` or ``gaussium_amd.build.build()``.
The code generators write csrc/_gen/eri_gen.h and csrc/_gen/rys_tab.h; hipcc
compiles the single translation unit csrc/gx_api.hip into
gaussium_amd/libgx.so.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
GEN = os.path.join(CSRC, '_gen')
LIB = os.path.join(HERE, 'libgx.so')
ROOT = os.path.dirname(HERE)

# Angular momentum handled by the generated class kernels.
ERI_LMAX = int(os.environ.get('GX_ERI_LMAX', '2'))

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
    '-munsafe-fp-atomics',
    # packed FP32 (v_pk_*_f32) is kept out of the generated code
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-parameter',
]


def generate_rys(out_dir, lmax=ERI_LMAX):
    """Rys tables and 2D-integral code (gen_rys.py, about 30 s of 60-digit
    arithmetic); kept while newer than the generator."""
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, 'rys_tab.h')
    gen = os.path.join(CSRC, 'gen_rys.py')
    if not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(gen):
        subprocess.run([sys.executable, gen, out + '.tmp', str(lmax)], check=True)
        os.replace(out + '.tmp', out)
    return out


def build(verbose=False):
    os.makedirs(GEN, exist_ok=True)
    subprocess.run([sys.executable, os.path.join(CSRC, 'gen_eri.py'),
                    os.path.join(GEN, 'eri_gen.h'), str(ERI_LMAX)], check=True)
    generate_rys(GEN)
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    tmp = LIB + '.tmp'
    cmd = [hipcc] + HIPCC_FLAGS + ['-I', os.path.join(ROOT, 'include'), '-I', CSRC, '-I', GEN,
                                   os.path.join(CSRC, 'gx_api.hip'), '-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
