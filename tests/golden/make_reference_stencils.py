"""Write tests/golden/reference_stencils.npz: input/output vectors of the
reference's own stencil functions (tools/finite_differences.f90 of
FlavioRosales/KG, compiled from its source into oracle/_ref by
oracle/Makefile), so the oracle's restatement stays pinned where the
reference is not available.

Each case holds the seeded input f (complex for the live 2nd-order set, real
for the 4th-order set), the upwinding coefficient beta (values -0.7, 0 and
1.3, so every branch of advec_x / advec_x_4 is taken), the spacing h and the
reference's output.  Run from the repository root after `oracle.build()`:

    python tests/golden/make_reference_stencils.py
"""
import os
import sys
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

SHAPES = [(13, 9, 10), (9, 7, 5)]  # (ni, nj, nk) with the padding g = 2
NAMES2 = ["dx1", "dx2", "dx3", "advec_x"]
NAMES4 = ["dx1_4", "dx2_4", "dx3_4", "advec_x_4"]


def main():
    if not oracle.have_reference():
        raise SystemExit("oracle/_ref is not built: the reference source is needed")
    out = {}
    for shape in SHAPES:
        ni, nj, nk = shape
        for name in NAMES2 + NAMES4:
            key = "%s_%dx%dx%d" % (name, ni, nj, nk)
            rng = np.random.default_rng(zlib.crc32(("golden", shape, name).__repr__().encode()))
            beta = rng.choice([-0.7, 0.0, 1.3], size=(nk, nj, ni))
            h = float(rng.uniform(0.05, 0.5))
            if name in NAMES2:
                f = rng.standard_normal((nk, nj, ni)) + 1j * rng.standard_normal((nk, nj, ni))
            else:
                f = rng.standard_normal((nk, nj, ni))
            ref = oracle.ref_stencil(name, f, h, 2, beta=beta)
            out[key + "_f"] = f
            out[key + "_beta"] = beta
            out[key + "_h"] = np.array(h)
            out[key + "_out"] = ref
    path = os.path.join(ROOT, "tests", "golden", "reference_stencils.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out) // 4, "cases")


if __name__ == "__main__":
    main()
