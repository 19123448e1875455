/*
 * rocbta: MI355X (gfx950) SPD block-tridiagonal-with-arrowhead (BTA) kernels.
 *
 * C ABI for the drop-in replacement of serinv's sequential SPD BTA/BT path.
 * All matrix pointers are DEVICE pointers to C-contiguous (row-major) float64
 * arrays in serinv's block layout:
 *   D  diagonal blocks        (n, b, b)
 *   L  lower-diagonal blocks  (n-1, b, b)
 *   R  lower arrow blocks     (n, a, b)
 *   T  arrow tip              (a, a)
 *   B  right-hand sides       (n*b + a, r), row stride ldb (>= r)
 * a == 0 selects the arrowless BT routines (pobtf / pobtsi / pobts); R and T
 * may then be NULL.  `stream` is a hipStream_t; every call is stream-ordered
 * and returns without synchronizing.  Return value: ROCBTA_OK or a negative
 * ROCBTA_ERR_* code (argument errors are detected on the host before any
 * launch).  Numerical failure of a Cholesky pivot is reported through the
 * device int `*info` (0 = success, else 1 + global row index of the first
 * non-positive pivot, as LAPACK potrf).
 *
 * Workspace: the caller provides device memory of at least
 * rocbta_workspace_bytes(op, n, b, a, r) bytes, 256-byte aligned.
 */
#ifndef ROCBTA_H
#define ROCBTA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ROCBTA_OK 0
#define ROCBTA_ERR_ARG (-1)
#define ROCBTA_ERR_HIP (-2)
#define ROCBTA_ERR_WORKSPACE (-3)

#define ROCBTA_OP_POBTAF 0
#define ROCBTA_OP_POBTASI 1
#define ROCBTA_OP_POBTAS 2
#define ROCBTA_OP_RESIDUAL 3
#define ROCBTA_OP_LOGDET 4
#define ROCBTA_OP_POBTAF_PERMUTED 5
#define ROCBTA_OP_POBTASI_PERMUTED 6
#define ROCBTA_OP_POBTAS_PERMUTED 7
#define ROCBTA_OP_POBTAF_STREAMING 8
#define ROCBTA_OP_POBTASI_STREAMING 9
#define ROCBTA_OP_DDBTASC 10
#define ROCBTA_OP_DDBTASCI 11

/* Library version string. */
const char* rocbta_version(void);

/* Bytes of device workspace an operation needs for the given shape. */
size_t rocbta_workspace_bytes(int op, int64_t n, int64_t b, int64_t a, int64_t r);

/* Block Cholesky factorization, in place.
 * Replaces serinv.algs.pobtaf / _pobtaf (src/serinv/algs/pobtaf.py:11-94,
 * 97-183) and, with a == 0, serinv.algs.pobtf / _pobtf (pobtf.py:11-122).
 * factorize_last_block = 0 leaves the Schur complements in D[n-1], R[n-1], T
 * (pobtaf.py:161). */
int rocbta_dpobtaf(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R, double* T,
                   int factorize_last_block, int* info, void* workspace, size_t workspace_bytes,
                   void* stream);

/* pobtaf without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as rocbta_dpobtaf, but the workspace is a
 * static device array of ROCBTA_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends (the limit: rocbta_workspace_bytes(ROCBTA_OP_POBTAF, n,
 * b, a, 0) must not exceed it, about 32 KiB per 64 x 64 diagonal tile; e.g.
 * n = 256, b = 256, a = 32 fits; larger shapes return ROCBTA_ERR_WORKSPACE
 * with a message on stderr).  The array is shared by every call on one GPU: calls that overlap
 * on one GPU (two streams, two host threads) race. */
#define ROCBTA_NOALLOC_SCRATCH_BYTES ((size_t)40 << 20)
size_t rocbta_noalloc_scratch_bytes(void);
int rocbta_dpobtaf_noalloc(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R,
                           double* T, int factorize_last_block, int* info, void* stream);

/* Selected inversion from the block Cholesky factor, in place.
 * Replaces serinv.algs.pobtasi / _pobtasi (src/serinv/algs/pobtasi.py:10-89,
 * 92-172) and, with a == 0, serinv.algs.pobtsi / _pobtsi (pobtsi.py:10-125).
 * invert_last_block = 0 takes D[n-1], R[n-1], T as already-inverted blocks
 * (pobtasi.py:114). */
int rocbta_dpobtasi(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R, double* T,
                    int invert_last_block, void* workspace, size_t workspace_bytes, void* stream);

/* Triangular solve with the block Cholesky factor, in place on B.
 * trans 'N': B <- L^{-1} B ; 'T' or 'C': B <- L^{-T} B.
 * Replaces serinv.algs.pobtas / _pobtas (src/serinv/algs/pobtas.py:10-62,
 * 65-161) and, with a == 0, serinv.algs.pobts / _pobts (pobts.py:10-118).
 * partial = 1 skips the last-block and tip solves (pobtas.py:99,129).
 * With partial = 0 the solve runs as one persistent kernel per call; the
 * first int32 of the workspace then holds 0, or -1 if that kernel hit its
 * internal scheduling time-out (the contents of B are then undefined). */
int rocbta_dpobtas(int64_t n, int64_t b, int64_t a, const double* D, const double* L,
                   const double* R, const double* T, double* B, int64_t r, int64_t ldb,
                   char trans, int partial, void* workspace, size_t workspace_bytes,
                   void* stream);

/* log det A = 2 sum log diag(L_ii) + 2 sum log diag(L_T) from the factor
 * written by rocbta_dpobtaf; result stored to the device double *out.
 * Workspace: rocbta_workspace_bytes(ROCBTA_OP_LOGDET, n, b, a, 0) bytes of
 * partial sums (the sum is deterministic: partials are added in a fixed
 * order).  No counterpart in the reference (extension for likelihood
 * evaluation). */
int rocbta_dlogdet(int64_t n, int64_t b, int64_t a, const double* D, const double* T,
                   double* out, void* workspace, size_t workspace_bytes, void* stream);

/* Mixed-precision factorization / selected inversion (no counterpart in the
 * reference).  Same arguments, layout, workspace and results as
 * rocbta_dpobtaf / rocbta_dpobtasi; the trsm, syrk and gemm tile products
 * round their FP64 operands to FP32 and accumulate in FP32 MFMA, while the
 * blocks stay FP64 in memory and the diagonal-tile potrf / trtri, the
 * arrow-tip factorization and rocbta_dlogdet run in FP64.  Accuracy is that
 * of FP32 products (relative error ~ cond(A) * 1e-7); solves regain FP64
 * accuracy by iterative refinement with rocbta_dresidual. */
int rocbta_mpobtaf(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R, double* T,
                   int factorize_last_block, int* info, void* workspace, size_t workspace_bytes,
                   void* stream);
int rocbta_mpobtasi(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R, double* T,
                    int invert_last_block, void* workspace, size_t workspace_bytes, void* stream);

/* FP64 residual Out = B - A X for the SPD BTA matrix A given by its stored
 * blocks (lower triangles of D and T are used; L = A_{i+1,i}, R = A_{t,i}).
 * X, B, Out: (n*b + a, r) with row stride ld; Out may alias B, not X.
 * Workspace: rocbta_workspace_bytes(ROCBTA_OP_RESIDUAL, n, b, a, r).  No
 * counterpart in the reference (iterative refinement of mixed solves). */
int rocbta_dresidual(int64_t n, int64_t b, int64_t a, const double* D, const double* L,
                     const double* R, const double* T, const double* X, const double* B,
                     double* Out, int64_t r, int64_t ld, void* workspace, size_t workspace_bytes,
                     void* stream);

/* Gradient of f = gl * log det A + gq * sum_k y_k^T A^{-1} y_k with respect
 * to the stored blocks of A (no counterpart in the reference; the backward of
 * serinv_amd.autograd).  XD, XL, XR, XT: the selected inverse written by
 * rocbta_dpobtasi; x = A^{-1} y, (n*b + a, r) with row stride ldx; gl, gq:
 * DEVICE pointers to the two upstream scalars (NULL = 0).  Writes
 *   GD, GT: w * (gl X - gq x x^T) with w = 2 below the diagonal, 1 on it,
 *           0 above (only the lower triangle of D and T is read by pobtaf);
 *   GL, GR: 2 * (gl X - gq x x^T) on the sub-diagonal and arrow blocks;
 *   gy (optional, NULL to skip; row stride ldx): 2 gq x.
 * G* may alias X* (in place); gy must not alias x. */
int rocbta_dgrad(int64_t n, int64_t b, int64_t a, const double* XD, const double* XL,
                 const double* XR, const double* XT, const double* x, int64_t r, int64_t ldx,
                 const double* gl, const double* gq, double* GD, double* GL, double* GR,
                 double* GT, double* gy, void* stream);

/* *out = sum_{i < rows, k < r} y[i*ld + k] * x[i*ld + k] (device scalar;
 * deterministic).  No counterpart in the reference. */
int rocbta_dquad(int64_t rows, int64_t r, const double* y, const double* x, int64_t ld,
                 double* out, void* stream);

/* ---------------------------------------------------------------- batched
 * Many independent SPD BTA systems (uneven n, b and a <= 64) as one plan,
 * built once for fixed device arrays and replayed (no counterpart in the
 * reference; for hyperparameter sweeps that factor many matrices at once).
 * A run factors every matrix (pobtaf, in one persistent launch), computes
 * each log-determinant, and with ROCBTA_BATCHED_SOLVE overwrites B[m]
 * (n*b + a, r[m], row stride r[m]) with A^{-1} B (one persistent launch per
 * sweep direction), and with ROCBTA_BATCHED_SELINV overwrites the factor
 * with the selected inverse (pobtasi) as in rocbta_dpobtasi.  Arrays are
 * overwritten in place exactly like the single-matrix calls. */
#define ROCBTA_BATCHED_FACTOR 1
#define ROCBTA_BATCHED_SOLVE 2
#define ROCBTA_BATCHED_SELINV 4
/* with ROCBTA_BATCHED_MIXED the factorization and selected inversion run in
 * the mixed form of rocbta_mpobtaf / rocbta_mpobtasi (FP32 tile products,
 * FP64 tile potrf / trtri, tip and log-determinant); the solves stay FP64
 * with that factor. */
#define ROCBTA_BATCHED_MIXED 8
/* Launch shapes (no counterpart in the reference; set by the autotuner):
 * ROCBTA_BATCHED_RC1 / _RC4 / _RC8 force the solve sweeps' column-chunk
 * width (right-hand-side columns per chain workgroup; default: 1 for r = 1,
 * 4 for r <= 4, else 8), ROCBTA_BATCHED_GEMM512 runs the selected-inversion
 * GEMM phases on 512-thread workgroups whose two halves split each 64x64
 * tile's K range (default: 256 threads, two workgroups per CU). */
#define ROCBTA_BATCHED_RC1 16
#define ROCBTA_BATCHED_RC4 32
#define ROCBTA_BATCHED_RC8 64
#define ROCBTA_BATCHED_GEMM512 128
typedef void* rocbta_batched_plan;

/* Build a plan (allocates its own device workspace; not stream-ordered).
 * Systems run in waves of at most a quarter of the CU count (factorization)
 * and 64 column chunks (solves), so any number of systems is accepted;
 * r[m] <= 512 columns per system. */
int rocbta_batched_create(int nmat, const int64_t* n, const int64_t* b, const int64_t* a,
                          double* const* D, double* const* L, double* const* R,
                          double* const* T, double* const* B, const int64_t* r, int ops,
                          rocbta_batched_plan* plan);
/* Workgroup counts of the persistent factorization / solve launches
 * (<= 0 keeps the current value); used by the autotuner. */
int rocbta_batched_set_grid(rocbta_batched_plan plan, int factor_grid, int solve_grid);
/* Enqueue one run on `stream`.  use_graph = 1 captures the launch sequence
 * into a hipGraph on the first call and replays it afterwards (the stream
 * must then be capturable, i.e. not the legacy null stream). */
int rocbta_batched_run(rocbta_batched_plan plan, int use_graph, void* stream);
/* Synchronise the device and copy back per-matrix info (LAPACK style, -1 =
 * scheduling time-out), log-determinants, and the solve status (0 or -1);
 * any pointer may be NULL. */
int rocbta_batched_results(rocbta_batched_plan plan, int* info, double* logdet, int* status);
/* Number of kernel launches in one run. */
int rocbta_batched_launch_count(rocbta_batched_plan plan);
int rocbta_batched_destroy(rocbta_batched_plan plan);

/* --------------------------------------------------------------- permuted
 * The "middle partition" forms used by serinv's distributed solvers: the
 * first and last diagonal blocks and the tip are left to a reduced system,
 * and the top buffer (n, b, b), buffer[i] = the (0, i) block, carries the
 * fill-in between block 0 and the middle blocks.  buffer[0] is not used.
 * a == 0 gives the BT forms (R and T may then be NULL).
 *
 * Factorization of the middle blocks 1 .. n-2, in place: D[i], L[i], R[i]
 * and buffer[i] become the factor blocks L_ii, L_{i+1,i}, L_{t,i}, L_{0,i};
 * D[0], D[n-1], R[0], R[n-1], T and buffer[n-1] receive the Schur
 * complements of the reduced system (full blocks, both triangles).  The input
 * buffer is overwritten (buffer[1] = L[0]^T is set here).
 * Replaces serinv.algs.pobtaf(..., buffer=) / _pobtaf_permuted
 * (src/serinv/algs/pobtaf.py:186-281) and, with a == 0, pobtf(..., buffer=)
 * / _pobtf_permuted (pobtf.py:125-185).
 * Workspace: rocbta_workspace_bytes(ROCBTA_OP_POBTAF_PERMUTED, n, b, a, 0). */
int rocbta_dpobtaf_permuted(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R,
                            double* T, double* buffer, int* info, void* workspace,
                            size_t workspace_bytes, void* stream);
/* Selected inversion of the middle blocks, in place, given the factor from
 * rocbta_dpobtaf_permuted with D[0], D[n-1], R[0], R[n-1], T and buffer[n-1]
 * already replaced by the inverse's blocks from the reduced system
 * (buffer[n-1] = X_{0,n-1}).  On exit D, L, R, T hold the selected inverse's
 * blocks as rocbta_dpobtasi writes them (L[0] = buffer[1]^T = X_{1,0}) and
 * buffer[i] = X_{0,i}.
 * Replaces serinv.algs.pobtasi(..., buffer=) / _pobtasi_permuted
 * (src/serinv/algs/pobtasi.py:175-244) and, with a == 0, pobtsi(...,
 * buffer=) / _pobtsi_permuted (pobtsi.py:128-179).
 * Workspace: rocbta_workspace_bytes(ROCBTA_OP_POBTASI_PERMUTED, n, b, a, 0). */
int rocbta_dpobtasi_permuted(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R,
                             double* T, double* buffer, void* workspace, size_t workspace_bytes,
                             void* stream);
/* Triangular solves over the middle blocks with that factor, in place on B:
 * trans 'N' runs the forward elimination of blocks 1 .. n-2 (updating B's
 * first block, block n-1 and tip rows), 'T' / 'C' the backward solve of
 * blocks n-2 .. 1 (reading the reduced system's solution in B's first block,
 * block n-1 and tip rows).
 * Replaces serinv.algs.pobtas(..., buffer=) / _pobtas_permuted
 * (src/serinv/algs/pobtas.py:164-218) and, with a == 0, pobts(..., buffer=)
 * / _pobts_permuted (pobts.py:121-165).
 * Workspace: rocbta_workspace_bytes(ROCBTA_OP_POBTAS_PERMUTED, n, b, a, r)
 * (its first int32 is set to 0, the status word of rocbta_dpobtas; T is not
 * read and may be NULL). */
int rocbta_dpobtas_permuted(int64_t n, int64_t b, int64_t a, const double* D, const double* L,
                            const double* R, const double* T, const double* buffer, double* B,
                            int64_t r, int64_t ldb, char trans, void* workspace,
                            size_t workspace_bytes, void* stream);

/* --------------------------------------------------------- non-SPD ddbt*
 * Diagonally dominant BTA / BT matrices without symmetry.  Arrays (device):
 *   D (n, b, b), Lo (n-1, b, b) = A_{i+1,i}, Up (n-1, b, b) = A_{i,i+1},
 *   Rl (n, a, b) = A_{t,i}, Ru (n, b, a) = A_{i,t}, T (a, a).
 * Block inverses are blocked Gauss-Jordan without pivoting (stable for
 * diagonally dominant blocks; the reference calls numpy/cupy linalg.inv).
 * direction: 0 = downward; 1 = upward (a == 0 only, as the reference's BT
 * routines).  Only the "AX=I" forms (no rhs / quadratic / buffers).
 * Workspace: rocbta_workspace_bytes(ROCBTA_OP_DDBTASC / _DDBTASCI, n, b, a, 0).
 *
 * Schur-complement sweep, in place: D[i] <- inverse of the Schur complement
 * of block i, Rl / Ru / T Schur-updated, T <- the tip's inverse (with
 * invert_last_block).  Replaces serinv.algs.ddbtasc / _ddbtasc
 * (src/serinv/algs/ddbtasc.py:9-185, 187-241) and, with a == 0,
 * ddbtsc / _ddbtsc / _ddbtsc_upward (ddbtsc.py:9-221). */
int rocbta_dddbtasc(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                    double* Rl, double* Ru, double* T, int invert_last_block, int direction,
                    void* workspace, size_t workspace_bytes, void* stream);
/* Selected inversion from the sweep's output: the blocks of A^{-1} on the
 * pattern.  Replaces serinv.algs.ddbtasci / _ddbtasci (ddbtasci.py:9-185,
 * 187-256) and, with a == 0, ddbtsci / _ddbtsci / _ddbtsci_upward
 * (ddbtsci.py:9-237). */
int rocbta_dddbtasci(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                     double* Rl, double* Ru, double* T, int invert_last_block, int direction,
                     void* workspace, size_t workspace_bytes, void* stream);

/* The A X A^T = B forms ("quadratic"): the same sweeps with the blocks of a
 * symmetric B (BD, BLo, BUp, BRl, BRu, BT, same shapes as A's) carried
 * along; after both sweeps A's arrays hold A^{-1} and B's A^{-1} B A^{-T}
 * on the pattern.  direction 0 = downward; 1 = upward, BT only (a == 0),
 * and the selected inversion must use the sweep's direction.  Replace
 * _ddbtasc_quadratic (ddbtasc.py:335-475) / _ddbtsc_quadratic (ddbtsc.py:
 * 271-311) / _ddbtsc_upward_quadratic (ddbtsc.py:314-356) and
 * _ddbtasci_quadratic (ddbtasci.py:343-576) / _ddbtsci_quadratic
 * (ddbtsci.py:299-369) / _ddbtsci_upward_quadratic (ddbtsci.py:372-445).
 * Workspace: ROCBTA_OP_DDBTASC. */
int rocbta_dddbtasc_quadratic(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                              double* Rl, double* Ru, double* T, double* BD, double* BLo,
                              double* BUp, double* BRl, double* BRu, double* BT,
                              int invert_last_block, int direction, void* workspace,
                              size_t workspace_bytes, void* stream);
int rocbta_dddbtasci_quadratic(int64_t n, int64_t b, int64_t a, double* D, double* Lo,
                               double* Up, double* Rl, double* Ru, double* T, double* BD,
                               double* BLo, double* BUp, double* BRl, double* BRu, double* BT,
                               int invert_last_block, int direction, void* workspace,
                               size_t workspace_bytes, void* stream);

/* Permuted ("middle partition") forms: the sweep eliminates the middle
 * blocks 1 .. n-2 top-down and leaves blocks 0, n-1 and the tip as the
 * reduced system; Lbuf and Ubuf ((n-1) x b x b, like Lo) carry the fill-ins
 * A~_{0,k+1} / A~_{k+1,0} of row and column 0.  The selected inversion
 * expects the reduced system's inverse in blocks 0, n-1, the tip, their
 * arrow blocks, Lbuf[n-2] and Ubuf[n-2]; on exit Lbuf[k] = X_{0,k+1},
 * Ubuf[k] = X_{k+1,0}.  n >= 2.  Replace _ddbtasc_permuted (ddbtasc.py:
 * 244-332) / _ddbtsc_permuted (ddbtsc.py:224-269) and _ddbtasci_permuted
 * (ddbtasci.py:259-341) / _ddbtsci_permuted (ddbtsci.py:240-297).
 * Workspace: ROCBTA_OP_DDBTASC. */
int rocbta_dddbtasc_permuted(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                             double* Rl, double* Ru, double* T, double* Lbuf, double* Ubuf,
                             void* workspace, size_t workspace_bytes, void* stream);
int rocbta_dddbtasci_permuted(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                              double* Rl, double* Ru, double* T, double* Lbuf, double* Ubuf,
                              void* workspace, size_t workspace_bytes, void* stream);

/* The permuted A X A^T = B forms: the same sweeps with B's blocks and B's
 * fill-in buffers (BLbuf, BUbuf, like Lbuf, Ubuf) carried along; BLbuf[k],
 * BUbuf[k] for k >= 1 accumulate and must be zero on entry (as the
 * reference's allocator leaves them).  For a symmetric B.  Replace
 * _ddbtasc_quadratic_permuted (ddbtasc.py:478-710) / _ddbtsc_quadratic_
 * permuted (ddbtsc.py:357-466) and _ddbtasci_quadratic_permuted
 * (ddbtasci.py:578-845) / _ddbtsci_quadratic_permuted (ddbtsci.py:446-639).
 * Workspace: ROCBTA_OP_DDBTASC. */
int rocbta_dddbtasc_quadratic_permuted(int64_t n, int64_t b, int64_t a, double* D, double* Lo,
                                       double* Up, double* Rl, double* Ru, double* T,
                                       double* Lbuf, double* Ubuf, double* BD, double* BLo,
                                       double* BUp, double* BRl, double* BRu, double* BT,
                                       double* BLbuf, double* BUbuf, void* workspace,
                                       size_t workspace_bytes, void* stream);
int rocbta_dddbtasci_quadratic_permuted(int64_t n, int64_t b, int64_t a, double* D, double* Lo,
                                        double* Up, double* Rl, double* Ru, double* T,
                                        double* Lbuf, double* Ubuf, double* BD, double* BLo,
                                        double* BUp, double* BRl, double* BRu, double* BT,
                                        double* BLbuf, double* BUbuf, void* workspace,
                                        size_t workspace_bytes, void* stream);

/* -------------------------------------------------------------- streaming
 * Host <-> device streaming forms for systems whose blocks do not fit in
 * device memory: D, L, R, T are HOST pointers (page-locked memory, e.g.
 * hipHostMalloc, for asynchronous copies), overwritten in place with the
 * same results as rocbta_dpobtaf / rocbta_dpobtasi.  Two device slots per
 * block kind (workspace) carry the recurrence; the copies run on two streams
 * of the call's own, overlapping the compute on `stream`.  Unlike the other
 * entry points these calls are synchronous: they return when the host
 * arrays hold the results (as the reference's streaming routines end with a
 * device synchronisation).  `info` is a DEVICE int as for rocbta_dpobtaf.
 * Replace serinv.algs.pobtaf(..., device_streaming=True) / _pobtaf_streaming
 * (src/serinv/algs/pobtaf.py:284-524) and, with a == 0, _pobtf_streaming
 * (pobtf.py:188-329); serinv.algs.pobtasi(..., device_streaming=True) /
 * _pobtasi_streaming (pobtasi.py:247-496) and, with a == 0,
 * _pobtsi_streaming (pobtsi.py:182-330).
 * Workspace: rocbta_workspace_bytes(ROCBTA_OP_POBTAF_STREAMING /
 * ROCBTA_OP_POBTASI_STREAMING, n, b, a, 0), about 6 b^2 / 12 b^2 doubles,
 * independent of n. */
int rocbta_dpobtaf_streaming(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R,
                             double* T, int factorize_last_block, int* info, void* workspace,
                             size_t workspace_bytes, void* stream);
int rocbta_dpobtasi_streaming(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R,
                              double* T, int invert_last_block, void* workspace,
                              size_t workspace_bytes, void* stream);

/* Diagnostic (no counterpart in the reference): runs the task decoding and
 * dependency enumeration of the persistent pobtaf kernel on the HOST for the
 * given shape and returns the number of violations (a task waiting on a task
 * that is not scheduled strictly before it, a task index that does not
 * decode back to itself, a tile outside its array); 0 means the schedule is
 * deadlock-free and in bounds.  -1: shape not handled by that kernel. */
int64_t rocbta_dag_selfcheck(int64_t n, int64_t b, int64_t a);

#ifdef __cplusplus
}
#endif

#endif /* ROCBTA_H */
