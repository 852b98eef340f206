"""The HIP kernels and the C ABI built for the CPU (tests/cpu_emu) and run
under AddressSanitizer + UBSan: every global and LDS access of every entry
point is bounds-checked on the host, and the GPU suite's own parity tests
run against that build, before anything runs on a GPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, 'cpu_emu')
BUILD = os.path.join(EMU, '_build')


def _libasan():
    out = subprocess.run(['g++', '-print-file-name=libasan.so'], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope='module')
def emu_build():
    if shutil.which('g++') is None or _libasan() is None:
        pytest.skip('needs g++ with AddressSanitizer')
    subprocess.run(['make', '-C', EMU, '-j8', '-s'], check=True)
    return BUILD


def _env():
    env = dict(os.environ)
    env['ASAN_OPTIONS'] = 'detect_leaks=0:verify_asan_link_order=0:abort_on_error=1'
    env['UBSAN_OPTIONS'] = 'halt_on_error=1:print_stacktrace=1'
    return env


@pytest.mark.parametrize('N', [33, 65, 129])
def test_abi_entry_points_in_bounds(emu_build, N):
    """Every entry point in both precision forms and every field solver
    (matrix, the switched-off length-(N-1) FFT, and the paired LDS FFT with
    its index checks), one and two sorts, at N = 33 and at the two sizes
    whose length-(N-1) FFT solve faulted on the GPU (65, 129).  N = 513 is
    beyond what the fiber emulation runs in minutes here (the paired FFT's
    solve alone is checked there through run_parity.py)."""
    r = subprocess.run([os.path.join(emu_build, 'abi_check'), str(N)], capture_output=True,
                       text=True, env=_env(), timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert 'ALL OK' in r.stdout
    assert 'ERROR: AddressSanitizer' not in r.stderr and 'runtime error' not in r.stderr


def test_gpu_parity_suite_on_cpu_build(emu_build):
    ids = ['tests/test_gpu_kernels.py::test_compute_fields_matches_oracle[33-fp64-matrix]',
           'tests/test_gpu_kernels.py::test_compute_fields_matches_oracle[97-fp64-matrix]',
           'tests/test_gpu_kernels.py::test_compute_fields_matches_oracle[65-fp32-bulk-matrix]',
           'tests/test_gpu_kernels.py::test_compute_fields_matches_oracle[97-fp64-rfft]',
           'tests/test_gpu_kernels.py::test_compute_fields_matches_oracle[113-fp64-rfft]',
           'tests/test_gpu_kernels.py::test_compute_fields_matches_oracle[65-fp32-bulk-rfft]',
           'tests/test_gpu_kernels.py::test_rfft_plans_and_launch_shapes[65-fp64-plan4-1024-3]',
           'tests/test_gpu_kernels.py::test_rfft_plans_and_launch_shapes[65-fp32-bulk-plan5-2048-7]',
           'tests/test_gpu_kernels.py::test_rfft_plans_and_launch_shapes[97-fp64-plan6-4096-5]',
           'tests/test_gpu_kernels.py::test_rfft_by_waves[257-fp64]',
           'tests/test_gpu_kernels.py::test_rfft_by_waves[257-fp32-bulk]',
           'tests/test_gpu_kernels.py::test_rfft_kernels_mixed[257-fp32-bulk-3-plan2-2048-1]',
           'tests/test_gpu_kernels.py::test_rfft_kernels_mixed[257-fp64-6-plan3-4096-3]',
           'tests/test_gpu_kernels.py::test_deposit_matches_oracle_and_is_deterministic[65-fp64]',
           'tests/test_gpu_kernels.py::test_deposit_matches_oracle_and_is_deterministic[65-fp32-bulk]',
           'tests/test_gpu_kernels.py::test_push_matches_oracle[65-fp64]',
           'tests/test_gpu_kernels.py::test_deposit_of_collapsed_plasma[65-0.005]',
           'tests/test_gpu_kernels.py::test_deposit_of_collapsed_plasma[65-0.05]',
           'tests/test_gpu_kernels.py::test_deposit_of_collapsed_plasma[65-0.3]',
           'tests/test_gpu_kernels.py::test_push_matches_oracle[65-fp32-bulk]',
           'tests/test_gpu_kernels.py::test_push_reflects_and_absorbs_at_the_wall',
           'tests/test_gpu_kernels.py::test_plasma_step_matches_oracle[33-mobile]',
           'tests/test_gpu_kernels.py::test_step_currents_do_not_depend_on_the_visit_order[fp64]',
           'tests/test_gpu_kernels.py::test_step_currents_do_not_depend_on_the_visit_order[fp32-bulk]',
           'tests/test_gpu_declustering.py::test_noise_filter_matches_oracle[fp64-0.6]',
           'tests/test_gpu_beam.py::test_beam_split_keeps_layer_order[700]',
           'tests/test_gpu_beam.py::test_beam_split_keeps_layer_order[5000]',
           'tests/test_gpu_beam.py::test_beam_split_keeps_layer_order[40000]',
           'tests/test_gpu_beam.py::test_beam_deposit_is_reproducible[3000]',
           'tests/test_gpu_declustering.py::test_noise_filter_matches_oracle[fp32-bulk-0.6]']
    env = _env()
    env['LD_PRELOAD'] = _libasan()
    r = subprocess.run(['python3', os.path.join(EMU, 'run_parity.py')] + ids,
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert 'ERROR: AddressSanitizer' not in r.stderr
