"""Write tests/golden/stencil_golden.npz from the reference stencil.

The reference's own tests hold no vectors for this path, so these fixtures are
outputs of the reference header itself (oracle/_ref, compiled in place from
tools/algorithms/3DFreqModeling/Helm3d_27pt.h by oracle/Makefile) on seeded
inputs: 3 grids (symmetric, asymmetric and free-surface PML, anisotropic h)
x the six MEX variants.  Run where the reference tree exists:
    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

CASES = [
    ((12, 10, 8), [[3, 2, 2], [3, 2, 2]], (10.0, 10.0, 10.0)),
    ((9, 11, 7), [[2, 3, 1], [4, 2, 3]], (10.0, 12.0, 9.0)),
    ((10, 10, 10), [[0, 2, 2], [3, 2, 2]], (10.0, 10.0, 10.0)),
    ((7, 13, 9), [[1, 4, 2], [1, 4, 2]], (12.5, 12.5, 12.5)),
]


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref not built (needs the reference tree)")
    out = {"ncase": np.array(len(CASES))}
    for c, (n, npml, h) in enumerate(CASES):
        rng = np.random.default_rng(777 + c)
        N = int(np.prod(n))
        wn = rng.uniform(2e-4, 2e-3, N)
        wnc = wn + 1j * rng.uniform(-2e-4, 2e-4, N)
        x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
        out.update({f"n{c}": np.array(n), f"npml{c}": np.array(npml), f"h{c}": np.array(h),
                    f"wn{c}": wn, f"wnc{c}": wnc, f"x{c}": x})
        for v in O.VARIANTS:
            w = wnc if v.endswith("wnc") else wn
            out[f"y_{v}{c}"] = O.hmvp(v, w, h, n, np.array(npml), x, use_ref=True)
    np.savez_compressed(os.path.join(HERE, "stencil_golden.npz"), **out)


if __name__ == "__main__":
    main()
