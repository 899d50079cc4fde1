"""Derivation of the exp_blend_f32 polynomial (oracle/gsr_oracle.cpp) and of the
GPU's exp_blend: p(f) = 1 + c1 f + ... + c5 f^5 ~ 2^f on [-1/2, 1/2], fitted by
iteratively reweighted least squares on the relative error (weights pushed
towards the largest errors, which approaches the minimax fit), coefficients
rounded to FP32.  Prints the coefficients as hex floats.  Not imported by
anything; run with python oracle/exp_blend_fit.py."""
import numpy as np


def fit(deg=5, iters=200):
    f = np.cos(np.linspace(0, np.pi, 40001)) * 0.5
    y = np.exp2(f)
    A = np.stack([f ** k for k in range(1, deg + 1)], 1)
    w = 1 / y
    for _ in range(iters):
        c, *_ = np.linalg.lstsq(A * w[:, None], (y - 1) * w, rcond=None)
        err = (1 + A @ c - y) / y
        w = w * (np.abs(err) / np.abs(err).max()) ** 0.2 + 1e-30
    return [float(np.float32(v)) for v in c], float(np.abs(err).max())


if __name__ == "__main__":
    c, e = fit()
    print("max relative error of the fit:", e)
    print("c1..c5:", [v.hex() for v in c])
