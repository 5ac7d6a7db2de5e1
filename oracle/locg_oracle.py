"""NumPy restatement of the reference's single-vector LOBPCG (test oracle).

Test infrastructure only: the product never imports this module. It follows
rqutils/ground_locg.py:182-471 step by step -- four matvecs per iteration
(three in the Rayleigh-Ritz step, one for the new residual), two normalising
plus two plain rounds of classical Gram-Schmidt in ``project_out``, the Kopp
closed-form 3x3 eigenpair and the self-consistency convergence test -- so that
iteration counts of the device implementation can be compared against it.
"""
from __future__ import annotations

import numpy as np

__all__ = ["ground_locg", "project_out", "eigenpair_2x2", "eigenpair_3x3",
           "compute_sas"]


def _vdot(a, b):
    # sum(conj(a) * b) -- the reference's inner product (ground_locg.py:243)
    return np.sum(np.conjugate(a) * b)


def _norm(a):
    return float(np.sqrt(np.sum(np.abs(a) ** 2)))


def compute_sas(matvec, *vectors):
    """``S^H A S`` built from one matvec per vector -- ground_locg.py:235-248."""
    nv = len(vectors)
    dtype = np.result_type(vectors[0].dtype, np.complex128) \
        if np.iscomplexobj(vectors[0]) else np.float64
    sas = np.zeros((nv, nv), dtype=dtype)
    mvs = []
    for i1, v1 in enumerate(vectors):
        mv1 = matvec(v1)
        mvs.append(mv1)
        for i2 in range(i1 + 1, nv):
            sas[i2, i1] = _vdot(vectors[i2], mv1)
    sas = sas + np.conjugate(sas).T
    for i1, (v1, mv1) in enumerate(zip(vectors, mvs)):
        sas[i1, i1] = _vdot(v1, mv1)
    return sas


def eigenpair_2x2(mat):
    """Lowest eigenpair of a 2x2 Hermitian matrix -- ground_locg.py:424-435."""
    d = np.real(np.diagonal(mat))
    det = d[0] * d[1] - np.abs(mat[1, 0]) ** 2
    tr = d[0] + d[1]
    eigval = (tr - np.sqrt(tr * tr - 4.0 * det)) * 0.5
    e0 = (d[1] - eigval + np.conjugate(mat[1, 0])) / (d[0] - eigval + mat[1, 0])
    eigvec = np.array([e0, -1.0], dtype=np.result_type(mat.dtype, np.float64))
    eigvec = eigvec / np.sqrt(np.real(e0) ** 2 + np.imag(e0) ** 2 + 1.0)
    return eigval, eigvec


def eigenpair_3x3(mat):
    """Lowest eigenpair of a 3x3 Hermitian matrix by Cardano's method.

    J. Kopp, Int. J. Mod. Phys. C 19, 523 (2008); follows
    ground_locg.py:438-471: characteristic-polynomial coefficients, the
    trigonometric root, and the eigenvector as the conjugated cross product of
    columns 1 and 2 of ``A - lambda I``.
    """
    with np.errstate(invalid="ignore", divide="ignore"):
        d = np.real(np.diagonal(mat))
        modod = np.abs(np.array([mat[1, 0], mat[2, 0], mat[2, 1]])) ** 2
        c2 = -np.sum(d)
        c1 = np.sum(d * np.roll(d, 1)) - np.sum(modod)
        c0 = np.sum(d * modod[::-1])
        c0 -= np.prod(d)
        c0 -= 2.0 * np.real(mat[0, 2] * mat[1, 0] * mat[2, 1])
        p = c2 * c2 - 3.0 * c1
        q = -13.5 * c0 - c2 * c2 * c2 + 4.5 * c2 * c1
        phi = np.arctan2(np.sqrt(27.0 * (0.25 * c1 * c1 * (p - c1) + c0 * (q + 6.75 * c0))), q) / 3.0
        cphi, sphi = np.cos(phi), np.sin(phi)
        xmin = min(2.0 * cphi, -cphi - np.sqrt(3.0) * sphi, -cphi + np.sqrt(3.0) * sphi)
        eigval = np.sqrt(p) / 3.0 * xmin - c2 / 3.0
        v0 = mat[:, 1].copy()
        v0[1] -= eigval
        v1 = mat[:, 2].copy()
        v1[2] -= eigval
        eigvec = np.conjugate(np.cross(v0, v1))
        eigvec = eigvec / np.sqrt(np.sum(np.real(eigvec) ** 2 + np.imag(eigvec) ** 2))
    return eigval, eigvec


def project_out(basis, vector):
    """Remove ``basis`` components from ``vector`` -- ground_locg.py:380-421.

    Two rounds of (classical Gram-Schmidt, normalise), two rounds of plain
    subtraction, then zero the result if its norm fell below 0.99.
    """
    vector = vector.copy()
    for _ in range(2):
        ips = [_vdot(vb, vector) for vb in basis]
        for vb, ip in zip(basis, ips):
            vector = vector - vb * ip
        nrm = _norm(vector)
        vector = vector / (1.0 if nrm == 0.0 else nrm)
    for _ in range(2):
        ips = [_vdot(vb, vector) for vb in basis]
        for vb, ip in zip(basis, ips):
            vector = vector - vb * ip
    return vector * (1.0 if _norm(vector) >= 0.99 else 0.0)


def _diagnostics(matvec, x, y, r, theta, kappa=None, reltol=0.0, converged=False):
    """One diagnostics record -- ground_locg.py:211-233: S^H A S of (x, y, r),
    rho = Re <x, A x>, kappa (zeros when absent), reltol, converged."""
    sas = compute_sas(matvec, x, y, r)
    rho = float(np.real(_vdot(x, matvec(x))))
    return {"x": x, "y": y, "r": r, "theta": float(np.real(theta)), "rho": rho,
            "kappa": np.zeros(3, complex) if kappa is None else np.asarray(kappa, complex),
            "sas": sas, "reltol": float(reltol), "converged": bool(converged)}


def ground_locg(matvec, xinit, maxiter=1000, tol=None, n_for_tol=None, history=None,
                rel_tol=None, debug=False):
    """Single-vector LOBPCG -- ground_locg.py:191-377 (callable path).

    ``n_for_tol`` is the vector length used in the convergence threshold
    (``xcurr.shape[0]`` in the reference, i.e. the padded length in sqd).
    If ``history`` is a list, ``(niter, theta, |r|, reltol)`` tuples are
    appended once per loop iteration. ``rel_tol`` replaces the reference's
    stopping rule by ``|r| <= rel_tol |theta|`` (the benchmark's residual).
    Returns ``(theta, x, niter)``; with ``debug`` the reference's diagnostics
    mode (ground_locg.py:358-376): every one of ``maxiter`` iterations runs
    and a list of per-step records (``_diagnostics``) is returned fourth.
    """
    x = np.array(xinit, dtype=np.result_type(xinit.dtype, np.float64))
    if tol is None:
        tol = float(np.finfo(x.dtype).eps)
    nlen = x.shape[0] if n_for_tol is None else n_for_tol
    x = x / _norm(x)

    # body_iter0 -- ground_locg.py:264-272
    ax = matvec(x)
    rho = np.real(_vdot(x, ax))
    r = ax - rho * x
    records = [_diagnostics(matvec, x, np.zeros_like(x), r, rho)] if debug else None

    # body_iter1 -- ground_locg.py:274-286
    nr = _norm(r)
    p = r / (1.0 if nr == 0.0 else nr)
    theta, kappa = eigenpair_2x2(compute_sas(matvec, x, p))
    t = p * kappa[0] - x * kappa[1]
    u = x * kappa[0] + p * kappa[1]
    xn = u / _norm(u)
    yn = t / _norm(t)
    rn = matvec(xn) - theta * xn
    x, y, r = xn, yn, rn
    if debug:
        records.append(_diagnostics(matvec, x, y, r, theta, np.insert(kappa, 1, 0.0)))

    # while_loop over body -- ground_locg.py:288-341, 366-370 (a scan over
    # maxiter steps in the diagnostics mode, :358-364)
    niter, converged = 0, False
    while niter < maxiter and (debug or not converged):
        p = project_out((x, y), r)
        theta, kappa = eigenpair_3x3(compute_sas(matvec, x, y, p))
        s = y * kappa[1] + p * kappa[2]
        ns = _norm(s)
        t = s * (kappa[0] / ns) - x * ns
        u = x * kappa[0] + s
        xn = u / _norm(u)
        yn = t / _norm(t)
        axn = matvec(xn)
        rn = axn - xn * theta
        reltol = (_norm(axn) - theta) * nlen * 10
        nrn = _norm(rn)
        converged = bool(nrn < tol * reltol) if rel_tol is None else \
            bool(nrn <= rel_tol * abs(theta))
        niter += 1
        x, y, r = xn, yn, rn
        if history is not None:
            history.append((niter, float(np.real(theta)), nrn, float(reltol)))
        if debug:
            records.append(_diagnostics(matvec, x, y, r, theta, kappa, reltol, converged))
    if debug:
        return float(np.real(theta)), x, niter, records
    return float(np.real(theta)), x, niter
