"""GPU parity of the non-SPD family (ddbtasc / ddbtasci, ddbtsc / ddbtsci).

rocbta_dddbtasc / rocbta_dddbtasci through serinv_amd.algs against the
oracle's restatement of src/serinv/algs/ddbtasc.py:187-241,
ddbtasci.py:187-256, ddbtsc.py:180-221, ddbtsci.py:176-237, and the
selected inverse against the dense inverse, on the reference test suite's
diagonally dominant non-symmetric matrices.  Tolerance 1e-12 relative
(max-norm per output array): the GPU inverts the diagonal blocks by
Gauss-Jordan without pivoting, the oracle by numpy.linalg.inv (LU with
partial pivoting); on these matrices both are accurate to ~1e-15.
Shapes: the reference grid (b, a in {2, 3}, n 1..4), 64-wide tile
boundaries of the blocked inverse (b 63, 64, 65, 100, 130, 256; a 1, 16,
64, 65), the arrowless forms in both directions; the A X A^T = B forms
(symmetric B, as the reference's tests) against the oracle and against
A^{-1} B A^{-T} from dense linear algebra, the BT ones in both directions
(ddbtsc.py:271-356, ddbtsci.py:299-445).
"""

import numpy as np
import pytest

from oracle import dense, generators, serinv_np
from tests._helpers import relerr, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def algs(gpu_device):
    import serinv_amd.algs as A

    return A


SHAPES = [(2, 2, 1), (3, 3, 4), (2, 3, 2), (63, 1, 3), (64, 16, 4), (65, 64, 3), (100, 3, 5),
          (130, 65, 2), (256, 32, 6)]


@pytest.mark.parametrize("b,a,n", SHAPES)
def test_ddbtasc_ddbtasci_match_oracle_and_dense(algs, gpu_device, b, a, n):
    print(f"ddbtasc b={b} a={a} n={n}", flush=True)
    A = generators.dd_bta(b, a, n, np.random.default_rng(41000 + b + 7 * a + 13 * n))
    arrs = dense.bta_dense_to_arrays_full(A, b, a, n)
    g = to_dev(gpu_device, *arrs)
    algs.ddbtasc(*g)
    o = [x.copy() for x in arrs]
    serinv_np.ddbtasc(*o)
    for x, y in zip(to_host(*g), o):
        assert relerr(x, y) < TOL
    algs.ddbtasci(*g)
    serinv_np.ddbtasci(*o)
    h = to_host(*g)
    for x, y in zip(h, o):
        assert relerr(x, y) < TOL
    for x, y in zip(h, dense.bta_dense_to_arrays_full(np.linalg.inv(A), b, a, n)):
        assert relerr(x, y) < TOL


@pytest.mark.parametrize("b,a,n", [(64, 8, 4), (100, 3, 3)])
def test_ddbt_partial_last_block(algs, gpu_device, b, a, n):
    print(f"ddbt partial b={b} a={a} n={n}", flush=True)
    A = generators.dd_bta(b, a, n, np.random.default_rng(42000 + b + a + n))
    arrs = dense.bta_dense_to_arrays_full(A, b, a, n)
    g = to_dev(gpu_device, *arrs)
    algs.ddbtasc(*g, invert_last_block=False)
    o = [x.copy() for x in arrs]
    serinv_np.ddbtasc(*o, invert_last_block=False)
    for x, y in zip(to_host(*g), o):
        assert relerr(x, y) < TOL
    # selected inversion with the last block row given as inverse blocks
    X = np.linalg.inv(A)
    Xs = dense.bta_dense_to_arrays_full(X, b, a, n)
    o = [x.copy() for x in arrs]
    serinv_np.ddbtasc(*o)
    o[0][-1], o[3][-1], o[4][-1], o[5][:] = Xs[0][-1], Xs[3][-1], Xs[4][-1], Xs[5]
    g = to_dev(gpu_device, *o)
    algs.ddbtasci(*g, invert_last_block=False)
    for x, y in zip(to_host(*g), Xs):
        assert relerr(x, y) < TOL


@pytest.mark.parametrize("direction", ["downward", "upward"])
@pytest.mark.parametrize("b,n", [(3, 4), (64, 5), (130, 3), (2, 1)])
def test_ddbtsc_ddbtsci(algs, gpu_device, b, n, direction):
    print(f"ddbtsc b={b} n={n} {direction}", flush=True)
    A = generators.dd_bt(b, n, np.random.default_rng(43000 + b + n))
    D, Lo, Up = dense.bta_dense_to_arrays_full(A, b, 0, n)[:3]
    g = to_dev(gpu_device, D, Lo, Up)
    algs.ddbtsc(*g, direction=direction)
    o = [x.copy() for x in (D, Lo, Up)]
    serinv_np.ddbtsc(*o, direction=direction)
    for x, y in zip(to_host(*g), o):
        assert relerr(x, y) < TOL
    algs.ddbtsci(*g, direction=direction)
    for x, y in zip(to_host(*g), dense.bta_dense_to_arrays_full(np.linalg.inv(A), b, 0, n)[:3]):
        assert relerr(x, y) < TOL


def test_ddbt_argument_errors(algs, gpu_device):
    b, a, n = 3, 2, 3
    A = generators.dd_bta(b, a, n, np.random.default_rng(1))
    g = to_dev(gpu_device, *dense.bta_dense_to_arrays_full(A, b, a, n))
    with pytest.raises(TypeError, match="Unexpected keyword"):
        algs.ddbtasc(*g, direction="upward")
    with pytest.raises(ValueError, match="rhs does not contain"):
        algs.ddbtasc(*g, rhs={"B_diagonal_blocks": g[0]}, quadratic=True)
    rhs = dict(zip(("B_diagonal_blocks", "B_lower_diagonal_blocks", "B_upper_diagonal_blocks",
                    "B_lower_arrow_blocks", "B_upper_arrow_blocks", "B_arrow_tip_block"),
                   [x.clone() for x in g]))
    with pytest.raises(NotImplementedError, match="AX=B"):
        algs.ddbtasc(*g, rhs=rhs)
    bt = {k: rhs[k] for k in list(rhs)[:3]}
    with pytest.raises(ValueError, match="direction"):
        algs.ddbtsc(g[0], g[1], g[2], rhs=bt, quadratic=True, direction="sideways")
    with pytest.raises(ValueError, match="buffers does not contain"):
        algs.ddbtasci(*g, buffers={})
    with pytest.raises(ValueError, match="direction"):
        algs.ddbtsc(g[0], g[1], g[2], direction="sideways")
    with pytest.raises(ValueError, match="upper arrow"):
        algs.ddbtasc(g[0], g[1], g[2], g[3], g[4][:, :, :1].contiguous(), g[5])


QSHAPES = [(2, 2, 1), (3, 3, 4), (64, 16, 4), (65, 64, 3), (100, 3, 5), (130, 65, 2), (64, 0, 4)]


@pytest.mark.parametrize("b,a,n", QSHAPES)
def test_ddbt_quadratic_matches_oracle_and_dense(algs, gpu_device, b, a, n):
    print(f"ddbt quadratic b={b} a={a} n={n}", flush=True)
    rng = np.random.default_rng(44000 + b + 7 * a + 13 * n)
    A = generators.dd_bta(b, a, n, rng)
    Bm = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(Bm)
    arrs = dense.bta_dense_to_arrays_full(A, b, a, n)
    brr = dense.bta_dense_to_arrays_full(Bm, b, a, n)
    g, gb = to_dev(gpu_device, *arrs), to_dev(gpu_device, *brr)
    keys = ("B_diagonal_blocks", "B_lower_diagonal_blocks", "B_upper_diagonal_blocks",
            "B_lower_arrow_blocks", "B_upper_arrow_blocks", "B_arrow_tip_block")
    if a:
        rhs = dict(zip(keys, gb))
        algs.ddbtasc(*g, rhs=rhs, quadratic=True)
    else:
        rhs = dict(zip(keys[:3], gb[:3]))
        algs.ddbtsc(g[0], g[1], g[2], rhs=rhs, quadratic=True)
    o, ob = [x.copy() for x in arrs], [x.copy() for x in brr]
    serinv_np.ddbtasc_quadratic(*o, *ob)
    for x, y in zip(to_host(*g, *gb), o + ob):
        assert relerr(x, y) < TOL
    if a:
        algs.ddbtasci(*g, rhs=rhs, quadratic=True)
    else:
        algs.ddbtsci(g[0], g[1], g[2], rhs=rhs, quadratic=True)
    serinv_np.ddbtasci_quadratic(*o, *ob)
    h = to_host(*g, *gb)
    for x, y in zip(h, o + ob):
        assert relerr(x, y) < TOL
    X = np.linalg.inv(A)
    ref = list(dense.bta_dense_to_arrays_full(X, b, a, n)) + list(
        dense.bta_dense_to_arrays_full(X @ Bm @ X.T, b, a, n))
    for x, y in zip(h, ref):
        assert relerr(x, y) < TOL


@pytest.mark.parametrize("direction", ["downward", "upward"])
@pytest.mark.parametrize("b,n", [(2, 1), (3, 4), (63, 3), (64, 5), (100, 4), (130, 2)])
def test_ddbtsc_quadratic_both_directions(algs, gpu_device, b, n, direction):
    print(f"ddbtsc quadratic {direction} b={b} n={n}", flush=True)
    rng = np.random.default_rng(45000 + b + 13 * n + (7 if direction == "upward" else 0))
    A = generators.dd_bt(b, n, rng)
    Bm = generators.dd_bt(b, n, rng)
    generators.symmetrize(Bm)
    arrs = dense.bta_dense_to_arrays_full(A, b, 0, n)[:3]
    brr = dense.bta_dense_to_arrays_full(Bm, b, 0, n)[:3]
    g, gb = to_dev(gpu_device, *arrs), to_dev(gpu_device, *brr)
    rhs = dict(zip(("B_diagonal_blocks", "B_lower_diagonal_blocks", "B_upper_diagonal_blocks"), gb))
    algs.ddbtsc(*g, rhs=rhs, quadratic=True, direction=direction)
    o, ob = [x.copy() for x in arrs], [x.copy() for x in brr]
    serinv_np.ddbtsc_quadratic(*o, *ob, direction=direction)
    for x, y in zip(to_host(*g, *gb), o + ob):
        assert relerr(x, y) < TOL
    algs.ddbtsci(*g, rhs=rhs, quadratic=True, direction=direction)
    serinv_np.ddbtsci_quadratic(*o, *ob, direction=direction)
    h = to_host(*g, *gb)
    for x, y in zip(h, o + ob):
        assert relerr(x, y) < TOL
    X = np.linalg.inv(A)
    ref = list(dense.bta_dense_to_arrays_full(X, b, 0, n)[:3]) + list(
        dense.bta_dense_to_arrays_full(X @ Bm @ X.T, b, 0, n)[:3])
    for x, y in zip(h, ref):
        assert relerr(x, y) < TOL
