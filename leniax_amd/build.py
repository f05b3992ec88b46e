"""Build libleniax_amd.so (HIP, gfx950) in-tree with hipcc.

    python -m leniax_amd.build          # or leniax_amd.build.build()
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'libleniax_amd.so')
SOURCES = ['lx_api.hip', 'lx_fused2d.hip', 'lx_fused2d_n32.hip', 'lx_fused2d_n64.hip', 'lx_fused2d_n128.hip', 'lx_plain.hip', 'lx_fft.hip', 'lx_direct.hip', 'lx_kernels.hip', 'lx_init.hip', 'lx_static.hip', 'lx_adjoint.hip', 'lx_tune.hip']
ARCH = os.environ.get('LENIAX_AMD_ARCH', 'gfx950')

FLAGS = [
    f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC',
    # no packed FP32 (v_pk_*_f32) generation
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result',
]


def _hipcc():
    p = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(p):
        raise RuntimeError('hipcc not found: the HIP toolchain is required to build leniax_amd')
    return p


def build(verbose=False):
    hipcc = _hipcc()
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        obj = os.path.join(objdir, src.replace('.hip', '.o'))
        cmd = [hipcc] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f'hipcc failed for {src}:\n{r.stderr}')
        if verbose:
            sys.stderr.write(r.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = OUT + '.tmp'
    cmd = [hipcc, f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{r.stderr}')
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
