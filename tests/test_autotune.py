"""Host logic of the batched-plan autotuner (CPU): batch classes and the
tile-shape candidates it times."""

import torch

from serinv_amd.autotune import batch_class, shape_candidates


def _sys(n, b, a):
    return (torch.zeros(n, b, b), torch.zeros(max(n - 1, 0), b, b), torch.zeros(n, a, b), torch.zeros(a, a))


def test_batch_class_counts_b_and_a_buckets():
    key = batch_class([_sys(3, 64, 5), _sys(2, 64, 16), _sys(4, 128, 17)], rhs=[torch.zeros(10, 4)] * 3)
    assert key == "r4:64x16*2,128x32*1"


def test_shape_candidates():
    # one column: every width is a candidate only if not wider than needed
    assert shape_candidates([torch.zeros(5)], selinv=False) == [(1, None)]
    # four columns: widths 1 and 4, both GEMM workgroup sizes
    assert shape_candidates([torch.zeros(5, 4)], selinv=True) == [(1, 256), (1, 512), (4, 256), (4, 512)]
    # 100 columns: width 1 would need 100 chunks (> 64)
    widths = {w for w, _ in shape_candidates([torch.zeros(5, 100)], selinv=False)}
    assert widths == {4, 8}
    # no right-hand sides: the default width only
    assert shape_candidates(None, selinv=True) == [(None, 256), (None, 512)]
