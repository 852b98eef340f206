"""The built C-ABI library against include/lcodepy_amd.h, without a GPU:
every function the header declares is exported by liblcodepy_amd.so and
bound by the ctypes layer, and the ctypes structures have the C layout
(offsets and sizes measured by compiling the header with gcc)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from lcodepy_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'lcodepy_amd.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(lc3d_[a-z_0-9]+)\s*\(',
                                 text, flags=re.M)))


def test_header_and_binding_name_the_same_functions():
    assert declared_functions() == sorted(_native.SYMBOLS)


def _exported(path):
    out = subprocess.run(['nm', '-D', '--defined-only', path], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason='library not built')
def test_library_exports_every_declared_function():
    exported = _exported(_native.LIB_PATH)
    missing = [s for s in declared_functions() if s not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason='library not built')
def test_library_loads_and_binds_without_a_gpu():
    L = _native.load_library()
    for name in _native.SYMBOLS:
        assert getattr(L, name) is not None
    with pytest.raises(RuntimeError):
        _native.load_library(os.path.join(ROOT, 'no-such-library.so'))


LAYOUT_C = r'''
#include <stdio.h>
#include <stddef.h>
#include "lcodepy_amd.h"
#define OFF(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
    printf("lc3d_params %zu\n", sizeof(lc3d_params));
    @PARAMS@
    printf("lc3d_sort %zu\n", sizeof(lc3d_sort));
    @SORT@
    printf("lc3d_fields %zu\nlc3d_currents %zu\nlc3d_fields6 %zu\n",
           sizeof(lc3d_fields), sizeof(lc3d_currents), sizeof(lc3d_fields6));
    return 0;
}
'''


@pytest.mark.skipif(shutil.which('gcc') is None, reason='needs gcc')
def test_ctypes_structures_match_the_c_layout(tmp_path):
    P, S = _native.Params, _native.Sort
    src = (LAYOUT_C.replace('@PARAMS@', ' '.join(f'OFF(lc3d_params, {f})' for f, _ in P._fields_))
           .replace('@SORT@', ' '.join(f'OFF(lc3d_sort, {f})' for f, _ in S._fields_)))
    c = tmp_path / 'layout.c'
    c.write_text(src)
    exe = tmp_path / 'layout'
    subprocess.run(['gcc', '-std=c11', '-I', os.path.dirname(HEADER), str(c), '-o', str(exe)],
                   check=True)
    got = dict(line.rsplit(' ', 1) for line in
               subprocess.run([str(exe)], capture_output=True, text=True, check=True)
               .stdout.splitlines())
    assert int(got['lc3d_params']) == ctypes.sizeof(P)
    assert int(got['lc3d_sort']) == ctypes.sizeof(S)
    for f, _ in P._fields_:
        assert int(got[f'lc3d_params.{f}']) == getattr(P, f).offset, f
    for f, _ in S._fields_:
        assert int(got[f'lc3d_sort.{f}']) == getattr(S, f).offset, f
    assert int(got['lc3d_fields']) == ctypes.sizeof(_native.Fields)
    assert int(got['lc3d_currents']) == ctypes.sizeof(_native.Currents)
    assert int(got['lc3d_fields6']) == ctypes.sizeof(_native.Fields6)


@pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason='library not built')
def test_gpu_library_refuses_the_fft_solver_before_touching_the_device():
    L = _native.load_library()
    p = _native.Params()
    p.grid_steps, p.grid_step_size, p.xi_step_size = 33, 0.1, 0.05
    p.precision, p.max_particles = _native.LC3D_FP64, 100
    p.field_solver = _native.SOLVERS['fft']
    h = ctypes.c_void_p()
    assert L.lc3d_create(ctypes.byref(p), ctypes.byref(h)) == 1   # LC3D_EINVAL
    assert not h.value
    p.field_solver = 7
    assert L.lc3d_create(ctypes.byref(p), ctypes.byref(h)) == 1
