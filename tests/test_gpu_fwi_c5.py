"""Gradient check of the FWI objective at full size: 256x256x128 + 20-point
PML at 12.5 m, 16 sources, 64x64 receivers, 4 and 6 Hz, squared slowness,
every solve to 1e-11 (FGMRES + 4-level complex64 multigrid V-cycle, eight
hierarchies solving in turn).

Central differences of the device objective against the device gradient on
20 seeded directions per setting (four tests of five directions each share
one forward model, observed data and gradient per setting):
  * cut_pml = True (PDEfunc's default, comp2phys = truncation): the gradient
    of PDEfunc.m:383-394 omits the model's extension into the PML, so it is
    the derivative only along directions that vanish on the faces of the
    physical grid (their edge-replicated extension is zero); such directions
    are used;
  * cut_pml = False: the gradient is folded back with the extension adjoint
    and is the full derivative; unrestricted random directions are used.
Bar: relative difference <= 1e-4 per direction.  The central difference
divides the objective's solver noise by the step, so that noise, not the
gradient, decides directions whose derivative is small: with solves to
1e-10 and a step of 1e-4 the absolute difference stayed near 0.02-0.04 on
every direction, i.e. 1.04e-4 relative on one whose derivative was 20x below
the typical one.  The solves here go to 1e-11 and the step is 3e-4 of a
direction scaled to max(m0), which puts that noise near 1e-3 absolute;
central differences are exact for the quadratic part of the objective and
the next term is O(eps^2)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NDIR = 20    # directions per setting
CHUNK = 5    # directions per test
_STATE = {}  # cut_pml -> (problem, observed data, start model, gradient)


def _state(cut_pml):
    if cut_pml not in _STATE:
        import torch
        from helm3d.fwi import FWIOpts, FWIProblem
        from helm3d.models import fwi_c5
        _STATE.clear()  # one problem on the device at a time
        import gc
        gc.collect()
        torch.cuda.empty_cache()
        model, m_true, m0 = fwi_c5()
        P = FWIProblem(model, FWIOpts(pml=20, pml_max=20, tol=1e-11, solver="mg", mg_levels=4,
                                      mg_precision="mixed", mg_workers=8, mg_k=(1, 2, 3, 5), cut_pml=cut_pml),
                       device=torch.device("cuda:0"))
        print(f"C5 gradient check, cut_pml={cut_pml}: forward data and gradient", flush=True)
        D = P.forw_model(m_true)
        _, g = P.objective(m0, D)
        _STATE[cut_pml] = (P, D, m0, g.cpu().numpy(), tuple(model.n))
    return _STATE[cut_pml]


@pytest.fixture(scope="module", autouse=True)
def _release_device_state():
    yield
    import gc
    import torch
    _STATE.clear()
    gc.collect()
    torch.cuda.empty_cache()


@pytest.mark.timeout(300)
@pytest.mark.parametrize("chunk", range(NDIR // CHUNK))
@pytest.mark.parametrize("cut_pml", [True, False])
def test_c5_gradient_vs_central_differences(cut_pml, chunk):
    P, D, m0, g, n = _state(cut_pml)
    rng = np.random.default_rng(20261205 + 1000 * int(cut_pml) + chunk)
    eps = 3e-4
    for it in range(CHUNK):
        dm = rng.standard_normal(n[::-1]).transpose(2, 1, 0)  # (nx, ny, nz)
        if cut_pml:
            dm[0], dm[-1] = 0.0, 0.0
            dm[:, 0], dm[:, -1] = 0.0, 0.0
            dm[:, :, 0], dm[:, :, -1] = 0.0, 0.0
        dm = dm.reshape(-1, order="F") * m0.max()
        fp, _ = P.objective(m0 + eps * dm, D, need_grad=False)
        fm, _ = P.objective(m0 - eps * dm, D, need_grad=False)
        fd = (fp - fm) / (2 * eps)
        an = float(np.dot(g, dm))
        rel = abs(fd - an) / abs(an)
        print(f"  direction {chunk * CHUNK + it}: central difference {fd:.10e}, gradient {an:.10e}, rel {rel:.2e}",
              flush=True)
        assert rel <= 1e-4
