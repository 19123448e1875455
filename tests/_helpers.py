"""Shared helpers for the GPU parity tests."""

import numpy as np

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "bta_golden.npz")


def relerr(x, ref):
    """max |x - ref| / max |ref| (0 for empty arrays)."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    if ref.size == 0:
        return 0.0
    den = np.max(np.abs(ref))
    return float(np.max(np.abs(x - ref)) / (den if den > 0 else 1.0))


def golden_cases():
    z = np.load(GOLDEN)
    names = sorted({k.split("/")[0] for k in z.files})
    out = {}
    for nm in names:
        out[nm] = {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(nm + "/")}
    return out


def to_dev(dev, *arrays):
    import torch

    return [None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in arrays]


def to_host(*tensors):
    return [None if t is None else t.detach().cpu().numpy() for t in tensors]
