"""Write colour_golden.json: the reference's own column colouring of a set of
Jacobian patterns.

The exact colour vector (modax/_sparsity.py:182-210) is pinned by none of the
reference's tests -- they check colour counts and orthogonality -- so this
script loads that one module by file path (its dependencies, numpy and
networkx, are installed here; the package's __init__ needs JAX and is not
run), colours each pattern with it and stores the inputs and outputs as data.
The reference never travels: tests/test_colour_golden.py compares this build's
colouring with the committed fixture.

Run from the repository root:  python tests/golden/make_colour_golden.py
"""

import importlib.util
import json
import os
import sys

import networkx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests import systems as S  # noqa: E402
from menodax_amd._sparsity import normalize_sparsity  # noqa: E402

REFERENCE_SPARSITY = "/root/reference/modax/_sparsity.py"


def _patterns():
    out = []
    for n_osc in (2, 4, 8, 16, 32):
        for wrap in (True, False):
            out.append((f"vdp_{'ring' if wrap else 'chain'}{n_osc}", 2 * n_osc,
                        S.vdp_sparsity(n_osc, wrap=wrap)))
    for n in (12, 24, 50):
        out.append((f"arrow{n}", n, S.arrow_system(n)[1]))
    for n in (6, 9, 16):
        out.append((f"ring{n}", n, S.ring_mask(n)))
    tri = np.eye(6, dtype=bool) | np.eye(6, k=1, dtype=bool) | np.eye(6, k=-1, dtype=bool)
    out.append(("tridiagonal6", 6, tri))
    rng = np.random.default_rng(7)
    for k, (n, density) in enumerate(((10, 0.2), (20, 0.1), (40, 0.05), (64, 0.04))):
        mask = (rng.random((n, n)) < density) | np.eye(n, dtype=bool)
        out.append((f"random{k}_{n}", n, mask))
    return out


def main():
    spec = importlib.util.spec_from_file_location("reference_sparsity", REFERENCE_SPARSITY)
    ref = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = ref  # its dataclasses look their module up
    spec.loader.exec_module(ref)
    cases = []
    for name, n, mask in _patterns():
        pattern = normalize_sparsity(mask, n)
        got = ref.colour_sparsity(ref.normalize_sparsity(np.asarray(mask), n))
        cases.append({"name": name, "n_vars": n, "pattern": [list(r) for r in pattern],
                      "colour": list(got.colour), "n_colours": int(got.n_colours)})
    path = os.path.join(HERE, "colour_golden.json")
    with open(path, "w") as f:
        json.dump({"source": "modax/_sparsity.py colour_sparsity, loaded by path",
                   "networkx": networkx.__version__, "cases": cases}, f, indent=0)
    print(f"wrote {len(cases)} cases to {path}")


if __name__ == "__main__":
    main()
