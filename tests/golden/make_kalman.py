"""Writes kalman_lg.json: the Kalman FP64 log-likelihood and its gradient
(oracle/kalman.py) for the linear-Gaussian settings the gradient bench leg
reports against, with the observations themselves: the T = 100
observations lg_y at (0.9, 1, 0.5), and their first 10 at (0.7, 1.3, 0.7).

Run from the repository root: python tests/golden/make_kalman.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import kalman  # noqa: E402

y = np.load(os.path.join(HERE, "observations.npz"))["lg_y"]
out = {}
for name, T, th in (("lg_T100", 100, (0.9, 1.0, 0.5)),
                    ("lg_T10_offset", 10, (0.7, 1.3, 0.7))):
    out[name] = {"T": T, "theta": list(th), "y": y[:T].tolist(),
                 "loglik": kalman.loglik(y[:T], *th),
                 "grad": kalman.grad(y[:T], *th).tolist()}
with open(os.path.join(HERE, "kalman_lg.json"), "w") as f:
    json.dump(out, f, indent=1)
