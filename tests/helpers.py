"""Shared helpers for the parity tests."""
import numpy as np


def bits(a: np.ndarray) -> np.ndarray:
    """Bit pattern view of a float32 array for exact comparison."""
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_nodes_equal(ex: dict, nd, what: str = "") -> None:
    """GPU export vs oracle Nodes, bit for bit."""
    for k in ("left", "right", "parent", "prim"):
        a, b = ex[k], getattr(nd, k)
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"{what}{k} differs at {bad[:10]} gpu={a[bad[:10]]} ref={b[bad[:10]]}"
    for k in ("box", "sa", "cost"):
        a, b = bits(ex[k]), bits(getattr(nd, k))
        bad = np.nonzero((a != b).reshape(len(a), -1).any(axis=1))[0]
        assert bad.size == 0, (f"{what}{k} differs at nodes {bad[:10]}: "
                               f"gpu={ex[k][bad[:3]]} ref={getattr(nd, k)[bad[:3]]}")
