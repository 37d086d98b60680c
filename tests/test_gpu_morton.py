"""ZenoParticles::orderByMortonCode on the GPU (zmpm_order_by_morton_code) against the oracle.

Bit-exact: the codes, both maps (a stable sort: particles of equal code keep
their order) and every channel of the permuted buffer, for the three buffer
layouts (SoA, AoS, and the AoSoA32 TileVector with the reference's 20-channel
J-fluid map, x at channel 1), n = 1 .. 300,000 including a cloud where most
particles share a code, and positions outside the box (clamped codes).
Argument errors come back as status codes without a launch.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from zeno_amd.zmpm import LAYOUT_AOS, LAYOUT_AOSOA32, LAYOUT_SOA, Context, ZmpmError

pytestmark = pytest.mark.gpu


def pack(planes, layout):
    """(nch, n) channel planes -> the flat buffer of the layout."""
    nch, n = planes.shape
    if layout == LAYOUT_SOA:
        return planes.reshape(-1).copy()
    if layout == LAYOUT_AOS:
        return planes.T.reshape(-1).copy()
    nt = (n + 31) // 32
    buf = np.zeros((nt, nch, 32), np.float32)
    pad = np.zeros((nch, nt * 32), np.float32)
    pad[:, :n] = planes
    buf[:] = pad.reshape(nch, nt, 32).transpose(1, 0, 2)
    return buf.reshape(-1)


def unpack(buf, layout, nch, n):
    if layout == LAYOUT_SOA:
        return buf.reshape(nch, n)
    if layout == LAYOUT_AOS:
        return buf.reshape(n, nch).T
    nt = (n + 31) // 32
    return buf.reshape(nt, nch, 32).transpose(1, 0, 2).reshape(nch, nt * 32)[:, :n]


@pytest.mark.parametrize("layout,nch,x_chn", [(LAYOUT_SOA, 7, 0), (LAYOUT_AOS, 5, 2), (LAYOUT_AOSOA32, 20, 1)],
                         ids=["soa", "aos", "tilevector20"])
@pytest.mark.parametrize("n", [1, 33, 4097, 300000])
def test_morton_order_matches_oracle(layout, nch, x_chn, n):
    rng = np.random.default_rng(n + nch)
    planes = rng.standard_normal((nch, n)).astype(np.float32)
    x = rng.uniform(-0.05, 1.05, size=(3, n)).astype(np.float32)  # some outside the box: clamped codes
    if n > 1000:
        x[:, : n // 2] = (0.5 + 0.001 * rng.integers(0, 4, size=(3, n // 2))).astype(np.float32)  # many ties
    planes[x_chn:x_chn + 3] = x
    lo, hi = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    codes, o2o, inv = O.order_by_morton_code(x, lo, hi)
    ctx = Context(1.0 / 64, 64, max(n, 1))
    src = torch.from_numpy(pack(planes, layout)).cuda()
    dst = torch.full_like(src, float("nan"))
    c = torch.empty(n, dtype=torch.int32, device="cuda")
    a = torch.empty(n, dtype=torch.int32, device="cuda")
    b = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx.order_by_morton_code(layout, nch, n, src, dst, x_chn, lo, hi, c, a, b)
    torch.cuda.synchronize()
    assert np.array_equal(c.cpu().numpy().view(np.uint32), codes[o2o])
    assert np.array_equal(a.cpu().numpy(), o2o)
    assert np.array_equal(b.cpu().numpy(), inv)
    out = unpack(dst.cpu().numpy(), layout, nch, n)
    assert np.array_equal(out, planes[:, o2o])
    ctx.close()


def test_morton_order_errors():
    ctx = Context(1.0 / 64, 64, 100)
    buf = torch.zeros(300, device="cuda")
    out = torch.zeros(300, device="cuda")
    i32 = torch.zeros(100, dtype=torch.int32, device="cuda")
    with pytest.raises(ZmpmError):  # more particles than the capacity
        ctx.order_by_morton_code(LAYOUT_SOA, 3, 101, buf, out, 0, (0, 0, 0), (1, 1, 1), None, i32, i32)
    with pytest.raises(ZmpmError):  # x channels outside the buffer
        ctx.order_by_morton_code(LAYOUT_SOA, 3, 100, buf, out, 1, (0, 0, 0), (1, 1, 1), None, i32, i32)
    with pytest.raises(ZmpmError):  # empty box
        ctx.order_by_morton_code(LAYOUT_SOA, 3, 100, buf, out, 0, (0, 0, 0), (1, 0, 1), None, i32, i32)
    with pytest.raises(ZmpmError):  # dst aliases src
        ctx.order_by_morton_code(LAYOUT_SOA, 3, 100, buf, buf, 0, (0, 0, 0), (1, 1, 1), None, i32, i32)
    ctx.close()
