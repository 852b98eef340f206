"""Runs GPU test functions against the CPU emulation of the kernels.

The product's Python layer is pointed at _build/libemu_lcodepy.so (the same
kernel and C-ABI sources built for the host with ASan/UBSan) and its device
tensors become host tensors; then pytest runs the given test ids of the GPU
suite.  Started by tests/test_cpu_emulation.py in a child process with
libasan preloaded.  Usage: run_parity.py <pytest node ids...>"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import torch  # noqa: E402

import lcodepy_amd._native as nat  # noqa: E402
import lcodepy_amd.runtime as rt  # noqa: E402

EMU_LIB = os.path.join(HERE, '_build', 'libemu_lcodepy.so')
_load = nat.load_library
nat.load_library = lambda path=None: _load(EMU_LIB)


def _ptr(t):
    if t is None:
        return None
    assert t.is_contiguous()
    return t.data_ptr()


def _stream(stream=None):
    return None


nat.ptr, nat.stream_ptr = _ptr, _stream
for mod in ('lcodepy_amd.plasma3d._bind', 'lcodepy_amd.plasma3d.move',
            'lcodepy_amd.plasma3d.weights', 'lcodepy_amd.plasma3d.fields',
            'lcodepy_amd.plasma3d.solver', 'lcodepy_amd.beam3d.weights',
            'lcodepy_amd.beam3d.move', 'lcodepy_amd.beam3d.beam_calculator'):
    m = __import__(mod, fromlist=['x'])
    if hasattr(m, 'ptr'):
        m.ptr = _ptr
    if hasattr(m, 'stream_ptr'):
        m.stream_ptr = _stream

rt.device = lambda: torch.device('cpu')
torch.cuda.is_available = lambda: True
torch.cuda.synchronize = lambda *a, **k: None
torch.cuda.current_device = lambda: 0
torch.cuda.set_device = lambda *a: None


class _NoDevice:
    def __init__(self, *a):
        pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


torch.cuda.device = _NoDevice


class _DoneEvent:
    def record(self, *a):
        pass

    def synchronize(self):
        pass


torch.cuda.Event = _DoneEvent


class _Stream:
    """Every launch of the host build completes before it returns: one
    stream, and waits between streams are no-ops."""
    cuda_stream = 0

    def __init__(self, *a, **k):
        pass

    def wait_stream(self, other):
        pass

    def wait_event(self, event):
        pass


torch.cuda.Stream = _Stream
torch.cuda.current_stream = lambda *a, **k: _Stream()
torch.cuda.stream = lambda s: _NoDevice()
torch.Tensor.record_stream = lambda self, stream: None
_empty = torch.empty


def _empty_unpinned(*a, **k):
    k.pop('pin_memory', None)
    return _empty(*a, **k)


torch.empty = _empty_unpinned


def _host_factory(fn):
    def make(*a, **k):
        dev = k.get('device')
        if dev == 'cuda' or (isinstance(dev, torch.device) and dev.type == 'cuda'):
            k['device'] = 'cpu'
        return fn(*a, **k)
    return make


for _name in ('tensor', 'full', 'zeros', 'ones', 'as_tensor', 'arange'):
    setattr(torch, _name, _host_factory(getattr(torch, _name)))
torch.empty = _host_factory(torch.empty)
_to = torch.Tensor.to


def _to_host(self, *a, **k):
    dev = k.get('device')
    if dev == 'cuda' or (isinstance(dev, torch.device) and dev.type == 'cuda'):
        k['device'] = 'cpu'
    return _to(self, *tuple('cpu' if x == 'cuda' else x for x in a), **k)


torch.Tensor.to = _to_host
torch.Tensor.is_cuda = property(lambda self: True)

import pytest  # noqa: E402

os.chdir(ROOT)
sys.exit(pytest.main(['-x', '-q', '-p', 'no:cacheprovider', '-s'] + sys.argv[1:]))
