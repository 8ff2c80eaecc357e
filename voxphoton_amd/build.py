"""Builds libvoxphoton.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m voxphoton_amd.build  (or voxphoton_amd.build.build_library()).
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')
OUT = os.path.join(HERE, 'libvoxphoton.so')
OBJDIR = os.path.join(HERE, '_build')

SOURCES = ['plan.cpp', 'setup.hip', 'matvec.hip', 'capi_op.cpp', 'precond.hip', 'krylov.hip',
           'capi_solve.cpp', 'capi_serve.cpp', 'prof.cpp', 'fields.hip', 'mor.hip', 'bicgstab.hip']

ARCH = os.environ.get('VP_OFFLOAD_ARCH', 'gfx950')


def _hipcc():
    h = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(h):
        raise RuntimeError('hipcc not found')
    return h


def _flags():
    return ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
            # packed FP32 instructions stay disabled for all device code
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-I', CSRC, '-I', INCLUDE, '-Wno-unused-result']


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + '.o')
    cmd = [_hipcc()] + _flags() + ['-x', 'hip', '-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'compile failed: {src}\n{" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    return obj, r.stderr


def _stale(src):
    """True when the object of src is older than src or any header."""
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + '.o')
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(d, f) for d in (CSRC, INCLUDE) for f in os.listdir(d)
                                        if f.endswith(('.hpp', '.h'))]
    return any(os.path.getmtime(d) > t for d in deps)


def build_library(verbose=False, incremental=False):
    """Compiles every source (incremental=True: only those whose object is
    older than the source or a header) and links the library."""
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    todo = [s for s in srcs if _stale(s)] if incremental else srcs
    results = []
    if todo:
        with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(todo))) as ex:
            results = list(ex.map(_compile, todo))
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + '.o') for s in srcs]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err)
    tmp = OUT + '.tmp'
    cmd = [_hipcc(), '-shared', '-fPIC', f'--offload-arch={ARCH}', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build_library(verbose='-v' in sys.argv, incremental='--incremental' in sys.argv))
