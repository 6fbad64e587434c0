"""CPU oracle for the FWI steps of the 3D Helmholtz path (TEST INFRASTRUCTURE).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module; the product never does.

Restates, on small grids and with exact (sparse direct) solves:
  * ``assemble``       the matrix of ``do_Hmvp`` (any MEX variant), recovered by
                       27-colour probing of the C oracle stencil
                       (Helm3d_27pt.h:827-875 -- a 27-point stencil couples
                       only points that differ by at most 1 per axis, so
                       points of equal (i mod 3, j mod 3, k mod 3) have
                       disjoint column supports);
  * ``sinc_weights``   opSincInterp(xin, xout, 4) with kaiser_window
                       (opSincInterp.m:26-51, kaiser_window.m:17-19), a scalar
                       loop per entry;
  * ``interp3``        opKron(Iz, Iy, Ix) of opInterp.m:51-57;
  * ``extension`` / ``padding``   Pext and Ppad: opKron of opExtension modes
                       1 and 0 (opExtension.m:85-101); their transposes are
                       comp2phys;
  * ``Problem``        PDEfunc.m:252-394 (OBJ), :421-441 (FORW_MODEL,
                       JACOB_FORW, JACOB_ADJ, HESS_GN) for the operto27 scheme,
                       with T(U) dm = dH (dm .* U) and
                       T(U)' z = conj(U) .* (dH' z) (discrete_helmholtz.m:286-289),
                       comp2phys = Ppad' (cut_pml) or Pext' (:231-235), the
                       source term of PDEfunc.m:324-339 and LS_misfit.m:21-26.

Pinning (tests/test_oracle_fwi.py): the gradient against central finite
differences of the objective (exact solves), the Jacobian pair by the
adjoint identity, HESS_GN against J'J, the interpolation against the
closed-form on-grid impulse and band-limited reconstruction, and the
extension pair by its dot identity.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O

KAISER_B4 = 6.31   # opSincInterp.m:42, r_b(4)


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------

def assemble(variant: str, wn, h, n, npml):
    """Sparse matrix of the stencil variant on the grid n (x fastest)."""
    n = tuple(int(a) for a in n)
    N = int(np.prod(n))
    q = np.arange(N)
    ci = q % n[0]
    cj = (q // n[0]) % n[1]
    ck = q // (n[0] * n[1])
    color = (ci % 3) + 3 * (cj % 3) + 9 * (ck % 3)
    rows, cols, vals = [], [], []
    for c in range(27):
        sel = np.nonzero(color == c)[0]
        y = O.hmvp(variant, wn, h, n, npml, (color == c).astype(np.complex128))
        for dk in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    a, b, e = ci[sel] + di, cj[sel] + dj, ck[sel] + dk
                    ok = (a >= 0) & (a < n[0]) & (b >= 0) & (b < n[1]) & (e >= 0) & (e < n[2])
                    p = a[ok] + n[0] * (b[ok] + n[1] * e[ok])
                    rows.append(p)
                    cols.append(sel[ok])
                    vals.append(y[p])
    rows, cols, vals = (np.concatenate(a) for a in (rows, cols, vals))
    keep = vals != 0
    return sp.csc_matrix((vals[keep], (rows[keep], cols[keep])), shape=(N, N))


def sinc_weights(xin, xout, r: int = 4, b: float = KAISER_B4) -> np.ndarray:
    """opSincInterp(xin, xout, r): W[o, i] = kaiser(d) sinc(d), d = (xout[o] -
    xin[i]) / (xout[1] - xout[0]); kaiser(d) = I0(b sqrt(1 - (d/r)^2)) / I0(b)
    for |d| <= r, else 0."""
    xin = [float(a) for a in np.asarray(xin).reshape(-1)]
    xout = [float(a) for a in np.asarray(xout).reshape(-1)]
    dx = xout[1] - xout[0] if len(xout) > 1 else 1.0
    W = np.zeros((len(xout), len(xin)))
    i0b = float(np.i0(b))
    for o, xo in enumerate(xout):
        for i, xi in enumerate(xin):
            d = xo / dx - xi / dx
            if abs(d) > r:
                continue
            win = float(np.i0(b * math.sqrt(max(0.0, 1.0 - (d / r) ** 2)))) / i0b
            s = 1.0 if d == 0 else math.sin(math.pi * d) / (math.pi * d)
            W[o, i] = win * s
    return W


def interp3(px, py, pz, cx, cy, cz) -> sp.csr_matrix:
    """ncomp x npts interpolation from the tensor grid (px, py, pz) onto
    (cx, cy, cz), x fastest."""
    Wx, Wy, Wz = (sp.csr_matrix(sinc_weights(p, c)) for p, c in ((px, cx), (py, cy), (pz, cz)))
    return sp.kron(Wz, sp.kron(Wy, Wx)).tocsr()


def _ext1(n: int, lo: int, hi: int) -> sp.csr_matrix:
    rows = np.arange(n + lo + hi)
    cols = np.clip(rows - lo, 0, n - 1)
    return sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(n + lo + hi, n))


def extension(n_phys, npml) -> sp.csr_matrix:
    """Pext: opKron(Ez, Ey, Ex) with edge replication (opExtension.m:94-96)."""
    npml = np.asarray(npml).reshape(2, 3)
    E = [_ext1(int(n_phys[a]), int(npml[0, a]), int(npml[1, a])) for a in range(3)]
    return sp.kron(E[2], sp.kron(E[1], E[0])).tocsr()


def padding(n_phys, npml) -> sp.csr_matrix:
    """Ppad: zero padding (opExtension mode 0); Ppad' is the truncation."""
    npml = np.asarray(npml).reshape(2, 3)
    P = []
    for a in range(3):
        n, lo, hi = int(n_phys[a]), int(npml[0, a]), int(npml[1, a])
        P.append(sp.csr_matrix((np.ones(n), (np.arange(n) + lo, np.arange(n))), shape=(n + lo + hi, n)))
    return sp.kron(P[2], sp.kron(P[1], P[0])).tocsr()


# ---------------------------------------------------------------------------
# PDEfunc restatement
# ---------------------------------------------------------------------------

class Problem:
    """Small 3D FWI problem: one model grid, tensor-grid sources (Q = I) and
    receivers, a list of frequencies, npml on every face."""

    def __init__(self, o, d, n, freq, src, rec, npml, unit="s2/m2", f0=10.0, t0=0.0, cut_pml=True):
        self.o = tuple(float(a) for a in o)
        self.d = tuple(float(a) for a in d)
        self.n = tuple(int(a) for a in n)
        self.freq = list(freq)
        self.src = src      # (xs, ys, zs)
        self.rec = rec      # (xr, yr, zr)
        self.npml = np.asarray(npml).reshape(2, 3)
        self.unit = unit
        self.f0 = f0
        self.t0 = t0
        self.cut_pml = cut_pml
        self.nt = tuple(self.n[a] + int(self.npml[:, a].sum()) for a in range(3))
        ot = [self.o[a] - self.npml[0, a] * self.d[a] for a in range(3)]
        comp = [ot[a] + self.d[a] * np.arange(self.nt[a]) for a in range(3)]
        self.Ps = interp3(*src, *comp)
        self.Pr = interp3(*rec, *comp).T.tocsr()
        self.Pext = extension(self.n, self.npml)
        self.comp2phys = (padding(self.n, self.npml) if cut_pml else self.Pext).T.tocsr()
        self.nsrc = self.Ps.shape[1]

    # per frequency: operator, factorisations, dwn on the extended grid
    def _freq(self, m, freq):
        vp = self.Pext @ np.asarray(m, dtype=np.float64).reshape(-1)
        wn, dwn, _ = O.param2wavenum(vp, freq, self.unit)
        h = self.d
        Hf = assemble("forw", wn, h, self.nt, self.npml)
        Ha = assemble("adj", wn, h, self.nt, self.npml)
        return dict(wn=wn, dwn=dwn, Hf=spla.splu(Hf), Ha=spla.splu(Ha),
                    T=lambda U, dm: O.hmvp("forw_deriv", dwn * (self.Pext @ dm), h, self.nt, self.npml, U),
                    Tadj=lambda U, z: np.conj(U) * O.hmvp("adj_deriv", dwn, h, self.nt, self.npml, z))

    def _source(self, k, j):
        w = O.fwi_wavelet([self.freq[k]], self.t0, self.f0)[0]
        q = np.zeros(self.nsrc, complex)
        q[j] = w
        return self.Ps @ q

    def forward(self, m):
        """FORW_MODEL: nrec x (nsrc nfreq), frequencies outer."""
        cols = []
        for k, f in enumerate(self.freq):
            F = self._freq(m, f)
            for j in range(self.nsrc):
                cols.append(self.Pr @ F["Hf"].solve(self._source(k, j)))
        return np.stack(cols, axis=1)

    def objective(self, m, D):
        """OBJ: (f, g) with g = comp2phys sum Re(T(U)' V)."""
        f = 0.0
        g = np.zeros(int(np.prod(self.n)))
        for k, fr in enumerate(self.freq):
            F = self._freq(m, fr)
            for j in range(self.nsrc):
                U = F["Hf"].solve(self._source(k, j))
                phi, r = O.ls_misfit(self.Pr @ U, D[:, k * self.nsrc + j])
                f += phi
                V = F["Ha"].solve(-(self.Pr.T @ r))
                g += self.comp2phys @ np.real(F["Tadj"](U, V))
        return f, g

    def jacob_forw(self, m, dm):
        cols = []
        for k, fr in enumerate(self.freq):
            F = self._freq(m, fr)
            for j in range(self.nsrc):
                U = F["Hf"].solve(self._source(k, j))
                cols.append(self.Pr @ F["Hf"].solve(-F["T"](U, dm)))
        return np.stack(cols, axis=1)

    def jacob_adj(self, m, dd):
        g = np.zeros(int(np.prod(self.n)))
        for k, fr in enumerate(self.freq):
            F = self._freq(m, fr)
            for j in range(self.nsrc):
                U = F["Hf"].solve(self._source(k, j))
                V = F["Ha"].solve(-(self.Pr.T @ dd[:, k * self.nsrc + j]))
                g += self.comp2phys @ np.real(F["Tadj"](U, V))
        return g

    def hess(self, m, dm, D):
        r"""HESS (PDEfunc.m:444-450) with the LS misfit (d2phi = I):
        dU = Hf \ (-T(U) dm), V = Ha \ (-Pr' r), dV = Ha \ (-T(V) dm - Pr' Pr dU),
        out += comp2phys Re(conj(U) .* (dm .* ddH' V) + conj(dU) .* dH' V + conj(U) .* dH' dV)
        (discrete_helmholtz.m:286-293: T(u) dm = dH (dm .* u); DT_adj)."""
        g = np.zeros(int(np.prod(self.n)))
        h = self.d
        for k, fr in enumerate(self.freq):
            F = self._freq(m, fr)
            vp = self.Pext @ np.asarray(m, dtype=np.float64).reshape(-1)
            _, _, ddwn = O.param2wavenum(vp, fr, self.unit)
            dmp = self.Pext @ np.asarray(dm, dtype=np.float64).reshape(-1)
            for j in range(self.nsrc):
                U = F["Hf"].solve(self._source(k, j))
                _, r = O.ls_misfit(self.Pr @ U, D[:, k * self.nsrc + j])
                dU = F["Hf"].solve(-F["T"](U, dm))
                V = F["Ha"].solve(-(self.Pr.T @ r))
                dV = F["Ha"].solve(-F["T"](V, dm) - self.Pr.T @ (self.Pr @ dU))
                t1 = np.conj(U) * (dmp * O.hmvp("adj_deriv", ddwn, h, self.nt, self.npml, V))
                g += self.comp2phys @ np.real(t1 + F["Tadj"](dU, V) + F["Tadj"](U, dV))
        return g

    def hess_gn(self, m, dm):
        g = np.zeros(int(np.prod(self.n)))
        for k, fr in enumerate(self.freq):
            F = self._freq(m, fr)
            for j in range(self.nsrc):
                U = F["Hf"].solve(self._source(k, j))
                dU = F["Hf"].solve(-F["T"](U, dm))
                dU = F["Ha"].solve(-(self.Pr.T @ (self.Pr @ dU)))
                g += self.comp2phys @ np.real(F["Tadj"](U, dU))
        return g
