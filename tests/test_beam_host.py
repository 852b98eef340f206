"""Host-side beam bookkeeping on CPU tensors: the drain's periodic joining
keeps the order of the layers pushed (beam3d/beam_io.py:50-73, whose drain
appends every layer to its buffer as it arrives), and buffers rounded to a power of
two are ordinary contiguous tensors of the asked shape."""
import pytest

torch = pytest.importorskip('torch')

from lcodepy_amd.beam3d.beam_io import BeamDrain  # noqa: E402
from lcodepy_amd.beam3d.data import BeamParticles, empty_rounded  # noqa: E402


def _layer(first, n):
    f = torch.arange(first, first + n, dtype=torch.float64).repeat(9, 1)
    f[1] *= -1
    i = torch.arange(first, first + n, dtype=torch.int64).repeat(2, 1)
    return BeamParticles(f=f, i=i)


def test_drain_joins_in_push_order():
    drain = BeamDrain(None)
    sizes = [3, 0, 5, 1, 7] * 20          # 100 layers, empty ones among them
    first = 0
    for n in sizes:
        drain.push_beam_layer(_layer(first, n))
        first += n
        assert len(drain._layers) <= BeamDrain.JOIN_EVERY + 1
    out = drain.beam_slice()
    assert out.size == first == sum(sizes)
    assert torch.equal(out.id, torch.arange(first))
    assert torch.equal(out.x, -torch.arange(first, dtype=torch.float64))
    assert out.f.is_contiguous() and out.i.is_contiguous()


@pytest.mark.parametrize('shape', [(3, 0), (3, 1), (9, 5), (3, 9, 1000), (2, 7)])
def test_rounded_buffers(shape):
    t = empty_rounded(shape, torch.float64, 'cpu')
    n = 1
    for s in shape:
        n *= s
    assert tuple(t.shape) == shape and t.is_contiguous()
    cap = t.untyped_storage().nbytes() // 8
    assert cap >= n and (cap & (cap - 1)) == 0


def test_append_into_rounded_buffer_keeps_values():
    a = _layer(0, 6)
    a = BeamParticles(f=a.f[:, 1:4], i=a.i[:, 1:4])    # a strided column slice
    b = _layer(10, 2)
    a.append(b)
    assert a.size == 5
    assert torch.equal(a.id, torch.tensor([1, 2, 3, 10, 11]))
    assert a.f.is_contiguous()
