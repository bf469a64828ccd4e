"""Synthetic denoising workloads (seeded, numpy.random.default_rng).

rects:    background 0.5, `nrect` axis-aligned rectangles with uniform integer
          corners and U[0,1] values (one per channel, shared geometry), plus
          N(0, noise^2) noise -- optionally correlated across channels.
batch_uneven: many instances with side lengths U{lo..hi}, aspect ratio in
          [0.5, 2] (square shapes rejected), L1 or SSD data term by coin flip
          and lambda log-uniform in [lam_lo, lam_hi].
"""
import numpy as np


def rects(H, W, C=1, seed=0, nrect=20, noise=0.1, corr=0.0, rng=None):
    rng = np.random.default_rng(seed) if rng is None else rng
    img = np.full((H, W, C), 0.5)
    for _ in range(nrect):
        y0, y1 = np.sort(rng.integers(0, H, 2))
        x0, x1 = np.sort(rng.integers(0, W, 2))
        img[y0:y1 + 1, x0:x1 + 1, :] = rng.random(C)
    if C > 1 and corr > 0:
        cov = noise ** 2 * ((1 - corr) * np.eye(C) + corr * np.ones((C, C)))
        L = np.linalg.cholesky(cov)
        img += rng.standard_normal((H, W, C)) @ L.T
    else:
        img += noise * rng.standard_normal((H, W, C))
    return img if C > 1 else img[:, :, 0]


def log_uniform(rng, lo, hi, size=None):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size))


def batch_uneven(n, seed=3, lo=64, hi=1024, lam_lo=1.0, lam_hi=20.0, l1_fraction=0.5):
    """Shapes, data terms and weights of an uneven batch (images generated lazily by `images`)."""
    rng = np.random.default_rng(seed)
    specs = []
    while len(specs) < n:
        H, W = int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1))
        if W == H or not 0.5 <= W / H <= 2.0:
            continue
        data = "l1" if rng.random() < l1_fraction else "l2"
        lam = float(log_uniform(rng, lam_lo, lam_hi))
        specs.append({"H": H, "W": W, "data": data, "lam": lam, "seed": int(rng.integers(2 ** 31))})
    return specs


def batch_square(n, side=1024, seed=5, lam_lo=1.0, lam_hi=20.0, data="l2"):
    rng = np.random.default_rng(seed)
    return [{"H": side, "W": side, "data": data, "lam": float(log_uniform(rng, lam_lo, lam_hi)),
             "seed": int(rng.integers(2 ** 31))} for _ in range(n)]


def images(specs, C=1):
    return [rects(s["H"], s["W"], C, seed=s["seed"]) for s in specs]
