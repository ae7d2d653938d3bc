"""Writes the committed golden fixtures under tests/golden/.

  fft_vectors.npz   numpy.fft (numpy 2.2) forward/inverse transforms of seeded
                    inputs: 1-D lengths and small uneven 3-D shapes.  They pin
                    the oracle's own FFT to an independent library routine.
  c1_oracle.npz     the CPU oracle's 300-iteration run of the 32^3 lamellar
                    workload (chiN=15, f=0.5, L=4.0, ds=0.01, seeded fields):
                    the five free-energy terms every 10 iterations and phiA,
                    wA after iteration 300.  The GPU test compares the HIP
                    solver with it without re-running the 75 s oracle there.

Run from the repository root:  python tests/golden/gen_golden.py [--fft-only]
"""
import argparse
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

FFT_1D = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 15, 16, 24, 30, 32, 48, 64, 100, 128]
FFT_3D = [(4, 6, 8), (3, 5, 6), (1, 1, 48), (1, 6, 6), (8, 12, 16), (2, 3, 10)]  # [nz][ny][nx]


def fft_vectors():
    rng = np.random.default_rng(20260101)
    out = {}
    for n in FFT_1D:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"x1_{n}"] = x
        out[f"F1_{n}"] = np.fft.fft(x)
        out[f"I1_{n}"] = np.fft.ifft(x) * n  # unnormalised inverse
    for shp in FFT_3D:
        key = "x".join(map(str, shp))
        x = rng.standard_normal(shp)  # real input, as the propagator's D2Z sees
        out[f"x3_{key}"] = x
        out[f"F3_{key}"] = np.fft.fftn(x)
        out[f"R3_{key}"] = np.fft.rfftn(x)  # half spectrum of cuFFT's D2Z layout
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **out)


def c1_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    spec = importlib.util.spec_from_file_location(
        "scft_fields", os.path.join(ROOT, "scft_ver1.01_amd", "fields.py"))
    fields = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fields)
    wA, wB = fields.workload_fields("c1_lamellar_32")
    head = np.concatenate([wA[:16], wB[:16]])  # start of the seeded input, to detect generator drift
    p = orc.params(32, 32, 32, 4.0, 4.0, 4.0, 15.0, 0.5, 0.01)
    phiA, phiB, hist = orc.run(p, wA, wB, 300)  # updates wA, wB in place
    np.savez_compressed(os.path.join(HERE, "c1_oracle.npz"), fe_every10=hist[9::10], phiA=phiA, wA=wA,
                        input_head=head, iters=300)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--fft-only", action="store_true")
    a = ap.parse_args()
    fft_vectors()
    if not a.fft_only:
        c1_oracle()
