"""serinv_amd: MI355X-native SPD BTA factorization, selected inversion and solve.

serinv_amd.algs holds the drop-in counterparts of serinv.algs' SPD routines
(pobtaf, pobtasi, pobtas, pobtf, pobtsi, pobts) running as hand-written
gfx950 HIP kernels; serinv_amd.logdet adds the log-determinant.
"""

__version__ = "0.1.0"
