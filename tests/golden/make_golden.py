"""Regenerate the golden fixtures of tests/golden/.

* reference_normal_quantiles.json -- the N(0,1) quantile table and expected
  CDF values that the reference's RNG tests hold
  (src/tests/rng_matrix_test/mod.rs:5-115), extracted as data from
  /root/reference when it is present.
* stream_normals.json -- bit patterns of the first normals of a few walks of
  the build's Philox4x32-10 stream (pins the stream definition).
* walk_eigenvalues.json -- eigenvalues of small walks on that stream from the
  oracle's numpy restatement of the reference algorithm + LAPACK dggev.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

REF_RNG_TEST = "/root/reference/src/tests/rng_matrix_test/mod.rs"


def reference_quantiles():
    out = os.path.join(HERE, "reference_normal_quantiles.json")
    if not os.path.exists(REF_RNG_TEST):
        print("reference not mounted; keeping", out)
        return
    text = open(REF_RNG_TEST).read()

    def block(name):
        m = re.search(name + r": &\[f64\] = &\[(.*?)\];", text, re.S)
        return [float(v) for v in re.findall(r"-?\d+\.\d+(?:e-?\d+)?", m.group(1))]

    q = block("STANDARD_NORMAL_QUANTILES")
    cdf = block("STANDARD_NORMAL_EXPECTED_CDF")
    assert len(q) == len(cdf) == 99
    json.dump({"source": "src/tests/rng_matrix_test/mod.rs:5-115", "quantiles": q, "expected_cdf": cdf},
              open(out, "w"), indent=1)


def stream_normals():
    cases = []
    for seed, stream, d, T in [(1, 0, 3, 8), (2, 0, 3, 8), (4294967295, 0, 5, 4), (7, 12345, 2, 6)]:
        z = oracle.walk_normals(seed, stream, d, T)
        cases.append({"seed": seed, "rng_stream": stream, "dim": d, "steps": T,
                      "bits": [int(v) for v in z.ravel().view(np.uint32)]})
    json.dump({"layout": "z[t*dim + i] as float32 bit patterns", "cases": cases},
              open(os.path.join(HERE, "stream_normals.json"), "w"), indent=1)


def walk_eigenvalues():
    cases = []
    for model in range(5):
        for d, T in [(1, 50), (2, 100), (3, 64), (5, 200), (8, 150), (12, 300)]:
            for seed in (1, 2, 3):
                ev = oracle.walk_eigenvalues_by_seed(model, d, T, seed)
                cases.append({"model": model, "dim": d, "steps": T, "seed": seed, "eigenvalues": ev.tolist()})
    json.dump({"source": "oracle.walk_eigenvalues_by_seed (reference algorithm + LAPACK dggev), rng_stream 0",
               "cases": cases}, open(os.path.join(HERE, "walk_eigenvalues.json"), "w"), indent=1)


if __name__ == "__main__":
    reference_quantiles()
    stream_normals()
    walk_eigenvalues()
    print("fixtures written to", HERE)
