"""CPU-only checks of the HIP library: it loads, exports every entry point
declared in include/exwave_amd.h, and its host-side mesh generator reproduces
the oracle's mesh (no GPU calls)."""
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "exwave_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(exw_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import exwave_amd.capi as capi
    lib = capi.load()
    names = _declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    # and the ctypes signature table covers all of them
    assert sorted(s[0] for s in capi.SIGNATURES) == names


@pytest.mark.parametrize("dim,nint,nref,tr,f", [(2, 5, 2, 1, 0.1), (3, 3, 1, 1, 0.05),
                                               (3, 4, 0, 2, 0.05), (2, 6, 0, 3, 0.05),
                                               (3, 2, 0, 0, 0.0)])
def test_mesh_generator_matches_oracle(dim, nint, nref, tr, f):
    import exwave_amd as X
    V, nb = X.mesh_hypercube(dim, nint, nref, tr, f, 2)
    m = O.Mesh(dim, nint, nref, tr, f, 2)
    assert np.array_equal(nb, m.neighbor)
    assert np.max(np.abs(V - m.vertices)) < 1e-15


def test_mesh_generator_rejects_bad_input():
    import exwave_amd as X
    with pytest.raises(X.ExwError):
        X.mesh_hypercube(4, 2)
    with pytest.raises(X.ExwError):
        X.mesh_hypercube(2, 2, 0, 0, 0.0, 7)
