"""Writes tests/golden/ref_utils_vectors.json: sample streams and log-pmf
values produced by the reference's own src/utils.cc (poisson_sample,
binomial_sample after sran(seed); normal_probability, poisson_probability,
negative_binomial_probability, gamma_probability), compiled unmodified into oracle/_ref by
oracle/Makefile.  Run in a container that has /root/reference:

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import ctypes as C  # noqa: E402

import numpy as np  # noqa: E402

from oracle import oracle as O  # noqa: E402

POIS = [(1, 0.05), (2, 0.7), (3, 4.2), (4, 11.5), (5, 12.0), (6, 37.9), (7, 812.3), (8, 20000.0)]
BINOM = [(11, 0.3, 5), (12, 0.02, 1000), (13, 0.5, 100), (14, 0.93, 4000), (15, 0.0001, 30000), (16, 0.7, 17)]
NORMAL = [(3.0, 2.5, 4.0), (0.0, 100.0, 25.0), (1e5, 99990.0, 1e4)]
POISP = [(0, 0.0), (3, 0.0), (0, 2.5), (7, 2.5), (120, 98.4), (-1, 3.0)]
NEGBIN = [(0, 1.0, 2.0), (5, 3.7, 0.5), (40, 55.2, 10.0), (1000, 1234.5, 3.3)]
GAMMA = [(0.37, 1.0, 1.0), (1.4, 3.2, 1.0), (0.05, 0.45, 1.0), (7.5, 12.0, 2.5), (2e-3, 0.11, 1.0)]


def main(n=64):
    L = O.ref_lib()
    if L is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    out = {"poisson_streams": [], "binomial_streams": [], "normal_probability": [], "poisson_probability": [],
           "negative_binomial_probability": [], "gamma_probability": []}
    for seed, lam in POIS:
        a = np.zeros(n, dtype=np.int32)
        L.ref_poisson_stream(seed, lam, n, a.ctypes.data_as(C.POINTER(C.c_int)))
        out["poisson_streams"].append({"seed": seed, "lam": lam, "samples": a.tolist()})
    for seed, p, nn in BINOM:
        a = np.zeros(n, dtype=np.int32)
        L.ref_binomial_stream(seed, p, nn, n, a.ctypes.data_as(C.POINTER(C.c_int)))
        out["binomial_streams"].append({"seed": seed, "p": p, "n": nn, "samples": a.tolist()})
    for x, m, v in NORMAL:
        out["normal_probability"].append({"x": x, "mean": m, "var": v, "value": L.ref_normal_probability(x, m, v)})
    for i, lam in POISP:
        out["poisson_probability"].append({"i": i, "lam": lam, "value": L.ref_poisson_probability(i, lam)})
    for v, m, shape in NEGBIN:
        out["negative_binomial_probability"].append(
            {"val": v, "m": m, "shape": shape, "value": L.ref_negative_binomial_probability(v, m, shape)})
    for x, a, b in GAMMA:
        out["gamma_probability"].append({"x": x, "a": a, "b": b, "value": L.ref_gamma_probability(x, a, b)})
    with open(os.path.join(HERE, "ref_utils_vectors.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
