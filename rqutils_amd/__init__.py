"""MI355X-native SQD hot path: Pauli-sum subspace matvec + single-vector LOBPCG.

Drop-in counterparts of ``rqutils.sqd`` (``sqd``, ``hproj``, ``get_xsource``,
``get_diagonal``), ``rqutils.ground_locg`` (``ground_locg``) and
``rqutils.paulis.symplectic.PauliSumXZ``, computed by hand-written HIP kernels
for gfx950 behind the C ABI of ``include/sqdk.h``.
"""
__version__ = "0.1.0"

__all__ = ["__version__"]
