"""Term tables of the fused output-layer lifts (host mirror of the kernel
plans in csrc/jet_rowgemm.hpp).

3D curl lift (experiments/metaball/_backend/flow/cases/metaball/
native_cupy_core.py:97-99, velocity_trace_jet2.py:93-101): with the vector
potential A stored as output rows 3k+c,
    u_a     = d_{a+1} A_{a+2} - d_{a+2} A_{a+1}
    d_j u_a = d_j d_{a+1} A_{a+2} - d_j d_{a+2} A_{a+1}
Each field is a signed sum of (last-hidden jet channel) x (weight slice W_c).
"""
from __future__ import annotations

PAIRS3 = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def hid(a: int, b: int) -> int:
    return PAIRS3.index(tuple(sorted((a, b))))


def curl_terms():
    """[(field a, jet channel c, weight slice s, sign)]: field a in 0..2 is u_a,
    3+3a+j is d_j u_a; channel c in 0..2 is grad_c, 3+p is Hessian pair p."""
    terms = []
    for a in range(3):
        terms.append((a, (a + 1) % 3, (a + 2) % 3, 1.0))
        terms.append((a, (a + 2) % 3, (a + 1) % 3, -1.0))
    for a in range(3):
        for j in range(3):
            terms.append((3 + 3 * a + j, 3 + hid(j, (a + 1) % 3), (a + 2) % 3, 1.0))
            terms.append((3 + 3 * a + j, 3 + hid(j, (a + 2) % 3), (a + 1) % 3, -1.0))
    return terms
