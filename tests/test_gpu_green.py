"""The reference's analytic check of the 3D Helmholtz discretisation, on the
device path: tools/algorithms/3DFreqModeling/test_ToolsAlgorithms3DFreqModeling.m:50-73
(testAnalytic) with its set-up (:10-47).

A homogeneous medium v = 2000 m/s on a 100^3 physical grid at 10 m, 4 Hz,
30-point PML on all six faces (opts.pml = 30, no pml_max, no free surface),
one source at (450, 450, 100) m -- linspace(4 dx, max(x) - 4 dx, 1) over the
50-point set-up grid is its end point -- injected by the sinc interpolation,
which on a grid node is a unit impulse.  The solve (FGMRES preconditioned by
the ML_GMRES V-cycle, tol 1e-10, :31-36) must match

    G(R) = prod(d) exp(i omega R / v) / (4 pi R)

to better than 1 % in the physical domain (comp2phys, cut_pml), with the
source node itself excluded.  This pins together the wavenumber
wn = (2 pi f / v)^2 of param2wavenum, the sign convention of the operator
(outgoing exp(+i k R)), the unit-impulse source scaling and the PML.

The wavenumber is formed by the device kernel (param2wavenum + edge
extension into the PML) and the solve is the library's multigrid FGMRES.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_PHYS = (100, 100, 100)
H = 10.0
V0 = 2000.0
FREQ = 4.0
NPML = 30
SRC_M = (450.0, 450.0, 100.0)


def _green(nt, src_idx):
    """prod(d) exp(i w R) / (4 pi R) on the computational grid (x fastest)."""
    ax = [(np.arange(nt[a]) - src_idx[a]) * H for a in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    R = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    R[tuple(src_idx)] = 1.0  # replaced below; avoids a division by zero
    w = 2 * np.pi * FREQ / V0
    G = H ** 3 * np.exp(1j * w * R) / (4 * np.pi * R)
    return G


@pytest.mark.parametrize("precision", ["c128", "mixed"])
def test_analytic_green_function(precision):
    import torch
    from helm3d.mg import MLGMRES
    dev = torch.device("cuda:0")
    nt = tuple(n + 2 * NPML for n in N_PHYS)
    src = tuple(int(round(SRC_M[a] / H)) + NPML for a in range(3))  # computational index
    print(f"Green's function check: nt={nt}, source at {src}, precision={precision}", flush=True)
    v = torch.full((int(np.prod(N_PHYS)),), V0, dtype=torch.float64, device=dev)
    M = MLGMRES(v, N_PHYS, H, NPML, FREQ, unit="m/s", nlevels=3, device=dev, autotune=False)
    try:
        nx, ny, _ = nt
        b = torch.zeros(M.H.N, dtype=torch.complex128, device=dev)
        b[src[0] + src[1] * nx + src[2] * nx * ny] = 1.0
        x, info = M.solve(b, tol=1e-10, maxinnerit=20, precision=precision)
        print(f"  outer iterations {info.iters}, relres {info.relres:.2e}", flush=True)
        assert info.converged and info.relres < 1e-10
        u = x.cpu().numpy().reshape(nt, order="F")
    finally:
        M.close()
    G = _green(nt, src)
    r = u - G
    r[src] = 0.0
    core = tuple(slice(NPML, nt[a] - NPML) for a in range(3))
    err = float(np.linalg.norm(r[core]) / np.linalg.norm(u[core]))
    print(f"  ||comp2phys (u - G)|| / ||comp2phys u|| = {err:.4e}", flush=True)
    assert err < 0.01
    # the plausible mistakes fail the same bar: the incoming wave, or the
    # field of a source scaled by 1/prod(d)
    G_in = np.conj(G)
    r_in = u - G_in
    r_in[src] = 0.0
    assert float(np.linalg.norm(r_in[core]) / np.linalg.norm(u[core])) > 0.1
    r_sc = u - G / H ** 3
    r_sc[src] = 0.0
    assert float(np.linalg.norm(r_sc[core]) / np.linalg.norm(u[core])) > 0.1
