"""Seeded synthetic point clouds shared by the tests."""
import glob
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def sphere(n, seed):
    x = np.random.default_rng(seed).normal(size=(n, 3))
    return x / np.linalg.norm(x, axis=1, keepdims=True)


def torus(n, seed, R=2.0, r=0.7):
    rng = np.random.default_rng(seed)
    th, ph = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    return np.stack([(R + r * np.cos(ph)) * np.cos(th), (R + r * np.cos(ph)) * np.sin(th),
                     r * np.sin(ph)], 1)


def flat_torus3(n, seed, D=32, noise=0.01):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 2 * np.pi, (n, 3))
    E = np.concatenate([np.cos(a), np.sin(a)], 1)
    Q, Rq = np.linalg.qr(rng.normal(size=(D, D)))
    Q = Q * np.sign(np.diag(Rq))
    return E @ Q[:6] + noise * rng.normal(size=(n, D))


def golden_cases():
    return sorted(glob.glob(os.path.join(HERE, "golden", "heap_*.npz")))


def flat_torus3_with_field(n, seed, D=32, noise=0.01):
    """flat_torus3's points (same stream) and a smooth tangent field
    sum_i f_i(a) dE/da_i of the angles a, mapped into R^D like the points."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 2 * np.pi, (n, 3))
    E = np.concatenate([np.cos(a), np.sin(a)], 1)
    Q, Rq = np.linalg.qr(rng.normal(size=(D, D)))
    Q = Q * np.sign(np.diag(Rq))
    X = E @ Q[:6] + noise * rng.normal(size=(n, D))
    f = np.stack([np.sin(a[:, 1]) + 0.5 * np.cos(a[:, 2]), np.cos(a[:, 0] + a[:, 2]),
                  np.sin(a[:, 0]) * np.cos(a[:, 1])], 1)
    v6 = np.zeros((n, 6))
    for i in range(3):
        v6[:, i] -= np.sin(a[:, i]) * f[:, i]
        v6[:, 3 + i] += np.cos(a[:, i]) * f[:, i]
    return X, v6 @ Q[:6]
