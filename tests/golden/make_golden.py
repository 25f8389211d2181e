"""Writes the committed index-table fixtures from the CPU oracle.

radial_table_257x512.npz: the FP32 radial sample table of GRadialAvg.cu:17-31
(counts per radius and (x, y) per sample, y taken mod 512), for a 512 tile.
ring_pix1.2_t512.npy / ring_pix1.35_t512.npy: half-spectrum indices y*257+x
of the 2D correlation ring (GCC2D.cu:28-33) at 1.2 A and 1.35 A.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402


def ring_bounds(pix, tile=512):
    p = np.float32(pix)
    res0 = np.float32(25.0) * p
    res1 = np.float32(2.0) * p / np.float32(0.8)
    if res1 < np.float32(3.5):
        res1 = np.float32(3.5)
    r2 = np.float32(tile) * p
    return float(r2 / res0), float(r2 / res1)


def main():
    counts, xy = O.radial_table()
    np.savez_compressed(os.path.join(HERE, "radial_table_257x512.npz"), counts=counts.astype(np.int16),
                        xy=xy.astype(np.int16))
    for pix in (1.2, 1.35):
        lo, hi = ring_bounds(pix)
        np.save(os.path.join(HERE, f"ring_pix{pix}_t512.npy"), O.ring_pixels(lo, hi).astype(np.int32))


if __name__ == "__main__":
    main()
