"""matpl_tall_xtz (per-type X^T Z over tall, pair-interleaved row sets; the training step's fitting
weight-gradient sums) against float64 torch products of the same rows. Tolerance: the FP32 summation
bound, |Y - Y64| <= 1e-5 * (|X|^T |Z|) entrywise (K up to ~5,000 rows per type here)."""
import numpy as np
import pytest
import torch

from matpl_amd import _lib
from matpl_amd.calcops import _ptr, _stream

pytestmark = pytest.mark.gpu


def _rows(buf, a, e, ld2, ld1, width):
    """The [2 (e - a), width] matrix of rows 2a .. 2e of a pair-interleaved buffer."""
    out = []
    for p in range(a, e):
        for h in range(2):
            off = p * ld2 + h * ld1
            out.append(buf[off:off + width])
    return torch.stack(out) if out else buf.new_zeros((0, width))


@pytest.mark.parametrize("M,N,counts", [(400, 64, [2500, 0, 731, 1]), (64, 64, [3, 2049, 17]), (37, 50, [1000])])
def test_tall_xtz_matches_float64(dev, M, N, counts):
    rng = np.random.default_rng(M + N)
    npos = sum(counts)
    # X rows: a pair per position with an extra gap (ld2 > 2 ld1) so that the strides are exercised
    ldx1, ldx2 = M + 3, 2 * (M + 3) + 5
    ldz1, ldz2 = 64, 3 * 64
    X = torch.tensor(rng.normal(size=npos * ldx2 + ldx1 + M), dtype=torch.float32, device=dev)
    Z = torch.tensor(rng.normal(size=npos * ldz2 + ldz1 + 64), dtype=torch.float32, device=dev)
    start = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    T = len(counts)
    work = torch.empty(_lib.query("matpl_tall_xtz_work", max(counts), T, M, N), dtype=torch.float32, device=dev)
    Y = torch.empty((T, M, N), dtype=torch.float32, device=dev)
    print(f"case tall_xtz M={M} N={N} counts={counts}", flush=True)
    _lib.call("matpl_tall_xtz", _ptr(X), ldx2, ldx1, _ptr(Z), ldz2, ldz1, start.ctypes.data, T, M, N, _ptr(Y),
              _ptr(work), work.numel(), _stream())
    torch.cuda.synchronize()
    Xd, Zd = X.double(), Z.double()
    for t in range(T):
        a, e = int(start[t]), int(start[t + 1])
        Xt = _rows(Xd, a, e, ldx2, ldx1, M)
        Zt = _rows(Zd, a, e, ldz2, ldz1, N)
        want = Xt.t() @ Zt
        bound = Xt.abs().t() @ Zt.abs()
        err = (Y[t].double() - want).abs()
        assert bool((err <= 1e-5 * bound + 1e-30).all()), (t, float((err / (bound + 1e-30)).max()))


@pytest.mark.parametrize("counts", [[2500, 0, 731, 1], [3000], [0, 5, 1030]])
def test_seg_rowsum_matches_float64(dev, counts):
    """matpl_seg_rowsum: per-type column sums of several strided buffers in one launch."""
    rng = np.random.default_rng(len(counts))
    npos = sum(counts)
    shapes = [(128, 50), (256, 50), (64, 64), (1, 1), (70, 3)]       # (ld, cols) per buffer
    bufs = [torch.tensor(rng.normal(size=max(npos, 1) * ld), dtype=torch.float32, device=dev) for ld, _ in shapes]
    start = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    T = len(counts)
    ptrs = np.asarray([b.data_ptr() for b in bufs], np.uint64)
    lds = np.asarray([ld for ld, _ in shapes], np.int64)
    cols = np.asarray([c for _, c in shapes], np.int32)
    totc = int(cols.sum())
    work = torch.empty(_lib.query("matpl_seg_rowsum_work", max(counts), T, totc), dtype=torch.float32, device=dev)
    out = torch.full((T, totc), float("nan"), dtype=torch.float32, device=dev)
    print(f"case seg_rowsum counts={counts}", flush=True)
    _lib.call("matpl_seg_rowsum", len(bufs), ptrs.ctypes.data, lds.ctypes.data, cols.ctypes.data, start.ctypes.data,
              T, _ptr(out), _ptr(work), work.numel(), _stream())
    torch.cuda.synchronize()
    off = 0
    for b, (ld, c) in zip(bufs, shapes):
        rows = b[:max(npos, 1) * ld].double().reshape(-1, ld)[:, :c]
        for t in range(T):
            a, e = int(start[t]), int(start[t + 1])
            want = rows[a:e].sum(0)
            bound = rows[a:e].abs().sum(0)
            err = (out[t, off:off + c].double() - want).abs()
            assert bool((err <= 1e-5 * bound + 1e-30).all()), (t, off, float((err / (bound + 1e-30)).max()))
        off += c
