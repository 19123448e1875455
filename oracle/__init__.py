"""CPU oracle for the SPD block-tridiagonal-with-arrowhead (BTA) path.

TEST INFRASTRUCTURE ONLY. Nothing in the product package (`serinv_amd`) may
import this package; only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` use it, and only as the checker / the timed
CPU baseline, never as the computation that is shipped or measured on the GPU.

Contents
--------
* `serinv_np`  : NumPy/SciPy restatement of serinv's sequential routines
                 `pobtaf`, `pobtasi`, `pobtas`, `pobtf`, `pobtsi`, `pobts`
                 (natural ordering, in-memory), op for op, with the reference
                 file:line each step follows.
* `dense`      : dense assembly of BTA/BT block arrays and extraction of the
                 BTA/BT pattern from a dense matrix (restating the semantics of
                 serinv's tests/testing_utils.py converters).
* `generators` : the reference test-suite generator `dd_bta`/`dd_bt`/`rhs`
                 restated with an explicit numpy Generator instead of a global
                 seed (tests/testing_utils.py:346-540).

How the oracle is pinned (tests/test_oracle_*.py):
* pobtaf/pobtf  -> dense Cholesky of the assembled matrix (unique for SPD);
* pobtasi/pobtsi-> dense inverse restricted to the BTA/BT pattern;
* pobtas/pobts  -> dense solve, and the forward half against a dense
                   triangular solve with the dense Cholesky factor;
* partial flags -> dense Schur complements;
* logdet        -> numpy.linalg.slogdet.
Mutation tests plant mistakes in copies of the oracle and check that the
pinning tests catch them.
"""
