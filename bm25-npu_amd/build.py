"""Builds libbm25mi.so for gfx950 in-tree (bm25-npu_amd/lib/) with hipcc, plus the
copy_calibration program used to scale rocprofv3 byte counters.

Usage: python bm25-npu_amd/build.py [--jobs N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / 'csrc'
LIB = HERE / 'lib'
OBJ = HERE / 'lib' / 'obj'
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
# FP32 contract of the reference build (-ffp-contract=off -fno-fast-math); packed FP32
# VALU ops are disabled for all device code.
COMMON = ['-O3', '-fPIC', '-std=c++17', '-ffp-contract=off', '-fno-fast-math', '-pthread', '-Wall',
          '-Wno-unused-command-line-argument']
DEVICE = ['--offload-arch=gfx950', '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']
EXTRA = os.environ.get('BM25MI_CFLAGS', '').split()
SOURCES = ['search.hip', 'session.hip', 'server.hip', 'mlp.hip', 'corpus.cpp', 'index.cpp']


def compile_one(src):
    out = OBJ / (src + '.o')
    cmd = [HIPCC] + COMMON + EXTRA + (DEVICE + ['-x', 'hip'] if src.endswith('.hip') else []) + \
        ['-c', str(CSRC / src), '-o', str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'compile {src} failed:\n{r.stdout}\n{r.stderr}')
    return out


def build(jobs=8, lib_dir=None):
    global LIB, OBJ
    if lib_dir is not None:
        LIB = Path(lib_dir)
        OBJ = LIB / 'obj'
    OBJ.mkdir(parents=True, exist_ok=True)
    sources = [s for s in SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, sources))
    target = LIB / 'libbm25mi.so'
    tmp = LIB / 'libbm25mi.so.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=gfx950', '-pthread', '-o', str(tmp)] + [str(o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'link failed:\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, target)
    # Counter-calibration copy program (a plain copy kernel of known size).
    exe = LIB / 'copy_calibration'
    cmd = [HIPCC] + COMMON + DEVICE + ['-o', str(exe) + '.tmp', str(CSRC / 'copy_calibration.hip')]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f'compile copy_calibration failed:\n{r.stdout}\n{r.stderr}')
    os.replace(str(exe) + '.tmp', exe)
    # Check program of the C++ facade (include/bm25mi_index.hpp), run by tests/test_facade.py.
    check_src = HERE.parent / 'tests' / 'cpp' / 'facade_check.cpp'
    if check_src.exists():
        exe = LIB / 'facade_check'
        cmd = ['g++', '-std=c++17', '-O2', '-Wall', '-I', str(HERE.parent / 'include'), str(check_src), '-o',
               str(exe) + '.tmp', '-L', str(LIB), '-lbm25mi', '-Wl,-rpath,$ORIGIN']
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(f'compile facade_check failed:\n{r.stdout}\n{r.stderr}')
        os.replace(str(exe) + '.tmp', exe)
    return target


if __name__ == '__main__':
    ap = argparse.ArgumentParser()
    ap.add_argument('--jobs', type=int, default=8)
    ap.add_argument('--lib-dir', default=None, help='output directory (default bm25-npu_amd/lib)')
    a = ap.parse_args()
    print(build(a.jobs, a.lib_dir))
