"""Parameter layout and file format of the learned three-body stand-in (include/d3x_sur.h).

Vector of NPARAMS float32: mean[64] | istd[64] | per element class c = 0, 1, 2:
W1[64][128] | b1[128] | W2[128][128] | b2[128] | w3[128] | b3, scale, shift, pad.
File: b"D3SUR001", uint32 count, count float32 (little endian).
"""

from __future__ import annotations

import numpy as np

NFEAT, NHID, NRBF, NCLASS = 64, 128, 21, 3
HEAD = 2 * NFEAT
O_W1, O_B1 = 0, NFEAT * NHID
O_W2 = O_B1 + NHID
O_B2 = O_W2 + NHID * NHID
O_W3 = O_B2 + NHID
O_TAIL = O_W3 + NHID
CLASS_SIZE = O_TAIL + 4
NPARAMS = HEAD + NCLASS * CLASS_SIZE
MAGIC = b"D3SUR001"


def element_class(Z) -> np.ndarray:
    """0 = H, He; 1 = Li..Ne; 2 = heavier elements."""
    Z = np.asarray(Z)
    return np.where(Z <= 2, 0, np.where(Z <= 10, 1, 2)).astype(np.int32)


def init_params(seed: int, mean, istd, scale, shift) -> np.ndarray:
    """Random weights with variance 1/fan-in, zero biases; normalisation and output scale given."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float64)
    p[:NFEAT] = mean
    p[NFEAT:HEAD] = istd
    for c in range(NCLASS):
        o = HEAD + c * CLASS_SIZE
        p[o + O_W1:o + O_B1] = rng.normal(0, np.sqrt(1.0 / NFEAT), NFEAT * NHID)
        p[o + O_W2:o + O_B2] = rng.normal(0, np.sqrt(1.0 / NHID), NHID * NHID)
        p[o + O_W3:o + O_TAIL] = rng.normal(0, np.sqrt(1.0 / NHID), NHID)
        p[o + O_TAIL + 1] = scale[c]
        p[o + O_TAIL + 2] = shift[c]
    return p.astype(np.float32)


def save(path: str, params) -> None:
    p = np.ascontiguousarray(params, dtype="<f4")
    if p.size != NPARAMS:
        raise ValueError(f"expected {NPARAMS} parameters")
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.uint32(NPARAMS).astype("<u4").tobytes())
        f.write(p.tobytes())


def load(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not a network parameter file")
        n = int(np.frombuffer(f.read(4), "<u4")[0])
        if n != NPARAMS:
            raise ValueError(f"{path}: {n} parameters, expected {NPARAMS}")
        return np.frombuffer(f.read(4 * n), "<f4").astype(np.float32)
