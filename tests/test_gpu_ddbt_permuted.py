"""GPU parity of the non-SPD permuted ("middle partition") forms.

rocbta_dddbtasc_permuted / rocbta_dddbtasci_permuted through
serinv_amd.algs with buffers= against the oracle's restatement of
src/serinv/algs/ddbtasc.py:244-332 and ddbtasci.py:259-341 (with a = 0,
ddbtsc.py:224-269 and ddbtsci.py:240-297), and through serinv's
reduced-system procedure (tests/_ddbt_permuted.py, the reduced system
inverted by the natural routines) against the dense inverse.  Tolerance
1e-12 relative (max-norm per output array); the GPU inverts the diagonal
blocks by Gauss-Jordan without pivoting, the oracle by numpy.linalg.inv.
Shapes: the reference's grid (b, a in {2, 3}, n 3, 4), n = 2 (no middle
block), the arrowless forms, 64-wide tile boundaries (b 63, 64, 65, 100,
130; a 1, 16, 64, 65) and n = 6, b = 256.
"""

import numpy as np
import pytest

from oracle import dense, generators, serinv_np
from tests._ddbt_permuted import selinv_ddbt_permuted
from tests._helpers import relerr, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12

SHAPES = [(2, 2, 2), (3, 3, 4), (2, 3, 3), (63, 1, 3), (64, 16, 4), (65, 64, 5), (100, 3, 4),
          (130, 65, 3), (256, 32, 6), (64, 0, 5), (3, 0, 4)]


@pytest.fixture(scope="module")
def algs(gpu_device):
    import serinv_amd.algs as A

    return A


def bufdict(Lbuf, Ubuf):
    return {"A_lower_buffer_blocks": Lbuf, "A_upper_buffer_blocks": Ubuf}


class GpuRT:
    """serinv_amd.algs under the procedure's signatures (the BT routines
    when the arrow is empty)."""

    def __init__(self, algs, a):
        self.A, self.bt = algs, a == 0

    def sc(self, arrs, Lbuf, Ubuf):
        if self.bt:
            self.A.ddbtsc(*arrs[:3], buffers=bufdict(Lbuf, Ubuf))
        else:
            self.A.ddbtasc(*arrs, buffers=bufdict(Lbuf, Ubuf))

    def sci(self, arrs, Lbuf, Ubuf):
        if self.bt:
            self.A.ddbtsci(*arrs[:3], buffers=bufdict(Lbuf, Ubuf))
        else:
            self.A.ddbtasci(*arrs, buffers=bufdict(Lbuf, Ubuf))

    def rsc(self, r):
        self.A.ddbtasc(*r)

    def rsci(self, r):
        self.A.ddbtasci(*r)


class OracleRT:
    sc = staticmethod(lambda arrs, L, U: serinv_np.ddbtasc_permuted(*arrs, L, U))
    sci = staticmethod(lambda arrs, L, U: serinv_np.ddbtasci_permuted(*arrs, L, U))
    rsc = staticmethod(lambda r: serinv_np.ddbtasc(*r))
    rsci = staticmethod(lambda r: serinv_np.ddbtasci(*r))


def gpu_zeros(dev):
    import torch

    return lambda shape: torch.zeros(shape, dtype=torch.float64, device=dev)


@pytest.mark.parametrize("b,a,n", SHAPES)
def test_ddbtasc_permuted_matches_oracle(algs, gpu_device, b, a, n):
    print(f"ddbtasc_permuted b={b} a={a} n={n}", flush=True)
    A = generators.dd_bta(b, a, n, np.random.default_rng(55000 + b + 7 * a + 13 * n))
    arrs = list(dense.bta_dense_to_arrays_full(A, b, a, n))
    Lbuf, Ubuf = np.zeros_like(arrs[1]), np.zeros_like(arrs[1])
    g = to_dev(gpu_device, *arrs, Lbuf, Ubuf)
    GpuRT(algs, a).sc(list(g[:6]), g[6], g[7])
    serinv_np.ddbtasc_permuted(*arrs, Lbuf, Ubuf)
    for x, y in zip(to_host(*g), arrs + [Lbuf, Ubuf]):
        assert relerr(x, y) < TOL


@pytest.mark.parametrize("b,a,n", SHAPES)
def test_ddbtasci_permuted_matches_oracle_and_dense(algs, gpu_device, b, a, n):
    print(f"ddbtasci_permuted b={b} a={a} n={n}", flush=True)
    A = generators.dd_bta(b, a, n, np.random.default_rng(56000 + b + 7 * a + 13 * n))
    arrs = list(dense.bta_dense_to_arrays_full(A, b, a, n))
    Lbuf, Ubuf = np.zeros_like(arrs[1]), np.zeros_like(arrs[1])
    g = list(to_dev(gpu_device, *arrs, Lbuf, Ubuf))
    selinv_ddbt_permuted(g[:6], g[6], g[7], GpuRT(algs, a), gpu_zeros(gpu_device))
    selinv_ddbt_permuted(arrs, Lbuf, Ubuf, OracleRT, np.zeros)
    h = to_host(*g)
    for x, y in zip(h, arrs + [Lbuf, Ubuf]):
        assert relerr(x, y) < TOL
    X = np.linalg.inv(A)
    for x, y in zip(h, dense.bta_dense_to_arrays_full(X, b, a, n)):
        assert relerr(x, y) < TOL
    for k in range(n - 1):  # buffers: X_{0,k+1}, X_{k+1,0}
        assert relerr(h[6][k], X[:b, (k + 1) * b:(k + 2) * b]) < TOL
        assert relerr(h[7][k], X[(k + 1) * b:(k + 2) * b, :b]) < TOL


def test_ddbt_permuted_argument_errors(algs, gpu_device):
    from serinv_amd.utils import allocate_ddbtax_permutation_buffers, allocate_ddbtx_permutation_buffers

    b, a, n = 3, 2, 3
    A = generators.dd_bta(b, a, n, np.random.default_rng(2))
    g = to_dev(gpu_device, *dense.bta_dense_to_arrays_full(A, b, a, n))
    bufs = allocate_ddbtax_permutation_buffers(g[1])
    assert set(bufs) == {"A_lower_buffer_blocks", "A_upper_buffer_blocks"}
    assert all(x.shape == g[1].shape and float(x.abs().max()) == 0 for x in bufs.values())
    assert len(allocate_ddbtx_permutation_buffers(g[1], quadratic=True)) == 4
    with pytest.raises(ValueError, match="correct arrays for A"):
        algs.ddbtasc(*g, buffers={"A_lower_buffer_blocks": bufs["A_lower_buffer_blocks"]})
    with pytest.raises(ValueError, match="A_upper_buffer_blocks"):
        algs.ddbtasci(*g, buffers={"A_lower_buffer_blocks": bufs["A_lower_buffer_blocks"],
                                   "A_upper_buffer_blocks": g[0]})
    rhs = dict(zip(("B_diagonal_blocks", "B_lower_diagonal_blocks", "B_upper_diagonal_blocks",
                    "B_lower_arrow_blocks", "B_upper_arrow_blocks", "B_arrow_tip_block"),
                   [x.clone() for x in g]))
    with pytest.raises(ValueError, match="correct arrays for B"):
        algs.ddbtasc(*g, rhs=rhs, quadratic=True, buffers=bufs)
    one = [g[0][:1].clone(), g[1][:0].clone(), g[2][:0].clone(), g[3][:1].clone(),
           g[4][:1].clone(), g[5].clone()]
    with pytest.raises(IndexError):
        algs.ddbtasc(*one, buffers={k: v[:0].clone() for k, v in bufs.items()})


_BK = ("B_diagonal_blocks", "B_lower_diagonal_blocks", "B_upper_diagonal_blocks",
       "B_lower_arrow_blocks", "B_upper_arrow_blocks", "B_arrow_tip_block")


class GpuQRT:
    """The A X A^T = B forms under the procedure's signatures."""

    def __init__(self, algs, a):
        self.A, self.bt = algs, a == 0

    def _call(self, fa, fbt, arrs, brr, bufs=None):
        kw = {"quadratic": True}
        if bufs is not None:
            kw["buffers"] = bufs
        if self.bt:
            fbt(*arrs[:3], rhs=dict(zip(_BK[:3], brr[:3])), **kw)
        else:
            fa(*arrs, rhs=dict(zip(_BK, brr)), **kw)

    def sc(self, arrs, L, U, brr, BL, BU):
        bufs = dict(bufdict(L, U), B_lower_buffer_blocks=BL, B_upper_buffer_blocks=BU)
        self._call(self.A.ddbtasc, self.A.ddbtsc, arrs, brr, bufs)

    def sci(self, arrs, L, U, brr, BL, BU):
        bufs = dict(bufdict(L, U), B_lower_buffer_blocks=BL, B_upper_buffer_blocks=BU)
        self._call(self.A.ddbtasci, self.A.ddbtsci, arrs, brr, bufs)

    def rsc(self, r, rb):
        self._call(self.A.ddbtasc, self.A.ddbtsc, r, rb)

    def rsci(self, r, rb):
        self._call(self.A.ddbtasci, self.A.ddbtsci, r, rb)


class OracleQRT:
    sc = staticmethod(lambda arrs, L, U, brr, BL, BU:
                      serinv_np.ddbtasc_quadratic_permuted(*arrs, L, U, *brr, BL, BU))
    sci = staticmethod(lambda arrs, L, U, brr, BL, BU:
                       serinv_np.ddbtasci_quadratic_permuted(*arrs, L, U, *brr, BL, BU))
    rsc = staticmethod(lambda r, rb: serinv_np.ddbtasc_quadratic(*r, *rb))
    rsci = staticmethod(lambda r, rb: serinv_np.ddbtasci_quadratic(*r, *rb))


def quad_case(b, a, n, seed):
    rng = np.random.default_rng(seed)
    A = generators.dd_bta(b, a, n, rng)
    Bm = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(Bm)
    return A, Bm


@pytest.mark.parametrize("b,a,n", SHAPES)
def test_ddbtasc_quadratic_permuted_matches_oracle(algs, gpu_device, b, a, n):
    print(f"ddbtasc_quadratic_permuted b={b} a={a} n={n}", flush=True)
    A, Bm = quad_case(b, a, n, 59000 + b + 7 * a + 13 * n)
    arrs = list(dense.bta_dense_to_arrays_full(A, b, a, n))
    brr = list(dense.bta_dense_to_arrays_full(Bm, b, a, n))
    z = [np.zeros_like(arrs[1]) for _ in range(4)]
    g = list(to_dev(gpu_device, *arrs, *brr, *z))
    GpuQRT(algs, a).sc(g[:6], g[12], g[13], g[6:12], g[14], g[15])
    serinv_np.ddbtasc_quadratic_permuted(*arrs, z[0], z[1], *brr, z[2], z[3])
    for x, y in zip(to_host(*g), arrs + brr + z):
        assert relerr(x, y) < TOL


@pytest.mark.parametrize("b,a,n", SHAPES)
def test_ddbtasci_quadratic_permuted_matches_oracle_and_dense(algs, gpu_device, b, a, n):
    print(f"ddbtasci_quadratic_permuted b={b} a={a} n={n}", flush=True)
    A, Bm = quad_case(b, a, n, 60000 + b + 7 * a + 13 * n)
    arrs = list(dense.bta_dense_to_arrays_full(A, b, a, n))
    brr = list(dense.bta_dense_to_arrays_full(Bm, b, a, n))
    z = [np.zeros_like(arrs[1]) for _ in range(4)]
    g = list(to_dev(gpu_device, *arrs, *brr, *z))
    selinv_ddbt_permuted(g[:6], g[12], g[13], GpuQRT(algs, a), gpu_zeros(gpu_device), g[6:12],
                         (g[14], g[15]))
    selinv_ddbt_permuted(arrs, z[0], z[1], OracleQRT, np.zeros, brr, (z[2], z[3]))
    h = to_host(*g)
    for x, y in zip(h, arrs + brr + z):
        assert relerr(x, y) < TOL
    X = np.linalg.inv(A)
    ref = list(dense.bta_dense_to_arrays_full(X, b, a, n)) + list(
        dense.bta_dense_to_arrays_full(X @ Bm @ X.T, b, a, n))
    for x, y in zip(h, ref):
        assert relerr(x, y) < TOL
