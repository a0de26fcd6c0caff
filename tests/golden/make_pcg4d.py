"""Writes tests/golden/pcg4d.json: PCG4D golden vectors.

PCG4D is the 4-D counter hash of Jarzynski & Olano, "Hash Functions for GPU
Rendering" (JCGT 9(3), 2020), used by the reference in shaders/common/rng.wgsl
(pcg4d, rng_u4, rng_f4).  This is an independent numpy restatement of the
published algorithm (LCG step v*1664525+1013904223 on each lane, the
cross-lane multiply-add round x+=y*w, y+=z*x, z+=x*y, w+=y*z, the xorshift
v ^= v>>16, and a second multiply-add round), written separately from the
oracle's C++ so that the two pin each other.  It also records rng_f4
((u >> 8) * 2^-24) and the reference's hash (rng.wgsl:46-50).
"""
import json
import os

import numpy as np

M32 = np.uint64(0xFFFFFFFF)


def pcg4d_np(v):
    v = v.astype(np.uint64)
    v = (v * np.uint64(1664525) + np.uint64(1013904223)) & M32
    x, y, z, w = (v[:, i].copy() for i in range(4))
    x = (x + y * w) & M32
    y = (y + z * x) & M32
    z = (z + x * y) & M32
    w = (w + y * z) & M32
    x ^= x >> np.uint64(16)
    y ^= y >> np.uint64(16)
    z ^= z >> np.uint64(16)
    w ^= w >> np.uint64(16)
    x = (x + y * w) & M32
    y = (y + z * x) & M32
    z = (z + x * y) & M32
    w = (w + y * z) & M32
    return np.stack([x, y, z, w], 1).astype(np.uint32)


def hash_np(v):
    x = v.astype(np.uint64)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & M32
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & M32
    x ^= x >> np.uint64(16)
    return x.astype(np.uint32)


def main():
    rng = np.random.default_rng(20200315)
    special = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1],
                        [0xFFFFFFFF] * 4, [12345, 7, (1 << 16) + 3, 1], [4096 * 4 + 2, 999, 1, 7]], dtype=np.uint32)
    rand = rng.integers(0, 2 ** 32, size=(56, 4), dtype=np.uint64).astype(np.uint32)
    ins = np.concatenate([special, rand])
    outs = pcg4d_np(ins)
    f4 = (outs >> np.uint32(8)).astype(np.float64) / 16777216.0
    hin = np.concatenate([np.array([0, 1, 0xFFFFFFFF, 0x85EBCA6B], dtype=np.uint32), rand[:, 0]])
    doc = {
        "source": "Jarzynski & Olano, JCGT 2020; reference shaders/common/rng.wgsl:10-50",
        "pcg4d_in": ins.tolist(),
        "pcg4d_out": outs.tolist(),
        "rng_f4_out": f4.tolist(),
        "hash_in": hin.tolist(),
        "hash_out": hash_np(hin).tolist(),
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcg4d.json")
    with open(path, "w") as f:
        json.dump(doc, f, indent=0)
    print(path)


if __name__ == "__main__":
    main()
