"""The reduced-system procedure of serinv's non-SPD permuted tests.

serinv checks ddbtasc / ddbtasci with buffers= by eliminating the middle
blocks, inverting the reduced system formed by block 0, block n-1 and the
tip, mapping it back and running the permuted selected inversion
(tests/tests_algs/permuted/test_bta/test_ddbtasci_permuted.py:90-190; the BT
version under test_bt/).  The reduced system here is inverted with the
routines passed in (the 2-block ddbtasc + ddbtasci), so that the same steps
drive the oracle (numpy arrays) and the GPU routines (torch tensors).
"""


def reduced_from(D, Rl, Ru, T, Lbuf, Ubuf, zeros):
    """The reduced system as a 2-block BTA (test_ddbtasci_permuted.py:
    98-133): diagonal D[0], D[-1]; lower block A~_{n-1,0} = Ubuf[-1]; upper
    block A~_{0,n-1} = Lbuf[-1]; arrow blocks 0 and n-1; the tip."""
    b, a = D.shape[1], Rl.shape[1]
    r = [zeros((2, b, b)), zeros((1, b, b)), zeros((1, b, b)), zeros((2, a, b)),
         zeros((2, b, a)), zeros((a, a))]
    r[0][0] = D[0]
    r[0][1] = D[-1]
    r[1][0] = Ubuf[-1]
    r[2][0] = Lbuf[-1]
    r[3][0] = Rl[0]
    r[3][1] = Rl[-1]
    r[4][0] = Ru[0]
    r[4][1] = Ru[-1]
    r[5][:, :] = T
    return r


def map_back(D, Rl, Ru, T, Lbuf, Ubuf, r):
    """The reduced system's inverse back into the full arrays
    (test_ddbtasci_permuted.py:135-160)."""
    D[0] = r[0][0]
    D[-1] = r[0][1]
    Ubuf[-1] = r[1][0]
    Lbuf[-1] = r[2][0]
    Rl[0] = r[3][0]
    Rl[-1] = r[3][1]
    Ru[0] = r[4][0]
    Ru[-1] = r[4][1]
    T[:, :] = r[5]


def selinv_ddbt_permuted(arrs, Lbuf, Ubuf, rt, zeros, brr=None, Bbuf=None):
    """sc(buffers) -> reduced sc + sci -> map back -> sci(buffers).
    rt: object with sc(arrs, Lbuf, Ubuf[, brr, BL, BU]), sci(same),
    rsc(arrs[, brr]), rsci(arrs[, brr]) on the 2-block reduced system."""
    D, Lo, Up, Rl, Ru, T = arrs
    if brr is None:
        rt.sc(arrs, Lbuf, Ubuf)
        r = reduced_from(D, Rl, Ru, T, Lbuf, Ubuf, zeros)
        rt.rsc(r)
        rt.rsci(r)
        map_back(D, Rl, Ru, T, Lbuf, Ubuf, r)
        rt.sci(arrs, Lbuf, Ubuf)
        return
    BL, BU = Bbuf
    rt.sc(arrs, Lbuf, Ubuf, brr, BL, BU)
    r = reduced_from(D, Rl, Ru, T, Lbuf, Ubuf, zeros)
    rb = reduced_from(brr[0], brr[3], brr[4], brr[5], BL, BU, zeros)
    rt.rsc(r, rb)
    rt.rsci(r, rb)
    map_back(D, Rl, Ru, T, Lbuf, Ubuf, r)
    map_back(brr[0], brr[3], brr[4], brr[5], BL, BU, rb)
    rt.sci(arrs, Lbuf, Ubuf, brr, BL, BU)
