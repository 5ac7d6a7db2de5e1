"""CPU oracle for the SQD hot path: test infrastructure, not product code.

This package restates, in NumPy FP64 / complex128, the algorithms of the
reference's sample-based quantum diagonalization path:

- ``sqd_oracle``: Pauli-sum packing and grouping (rqutils/paulis/symplectic.py),
  state packing, uniquification, the sort-join source-index search, the
  per-group diagonal and the cached-xsource matvec (rqutils/sqd.py);
- ``locg_oracle``: the single-vector LOBPCG with its projection and the
  closed-form 2x2 / 3x3 Hermitian eigenpairs (rqutils/ground_locg.py);
- ``dense``: an independent dense Kronecker-product construction of Pauli
  sums, used to pin the restatement itself;
- ``mlp_oracle``: FP64 forward/backward/Adam of the start network (no
  reference counterpart), pinned by finite differences and torch;
- ``svsim_oracle``: the gate-list conversion and the per-gate update of the
  state-vector simulator (rqutils/svsim.py), pinned by dense gate matrices.

Only the test suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg
of ``bench.py`` may import this package, and only as the checker or as the
timed CPU baseline. The product (``rqutils_amd``) never imports it.
"""
