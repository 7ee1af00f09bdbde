"""Writes svd_golden.npz: 3x3 matrices and their singular values from LAPACK
(numpy.linalg.svd), used to pin the oracle's QR/Givens SVD (QRSVD.hpp).

Cases: random well-conditioned, near-identity, rotations, diagonal with
repeated entries, rank-deficient, inverted (det < 0), large stretch.
Signed singular values follow the reference convention: the smallest carries
the sign of det F.
"""
import os

import numpy as np

rng = np.random.default_rng(20201)
mats = []
mats += [np.eye(3) + 0.3 * rng.standard_normal((3, 3)) for _ in range(40)]
mats += [np.eye(3) + 1e-6 * rng.standard_normal((3, 3)) for _ in range(10)]


def rot(axis, ang):
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


mats += [rot(rng.standard_normal(3), rng.uniform(0, np.pi)) for _ in range(10)]
mats += [np.diag([2.0, 2.0, 0.5]), np.diag([1.0, 1.0, 1.0]), np.diag([3.0, 1.0, 1.0]),
         np.diag([0.5, 0.5, 0.5]), np.diag([1.5, 0.7, 0.7])]
for _ in range(10):  # rank-deficient
    a = rng.standard_normal((3, 2))
    mats.append(a @ rng.standard_normal((2, 3)))
for _ in range(10):  # inverted
    F = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    F[:, 0] *= -1
    mats.append(F)
mats += [np.diag([5.0, 0.2, 1.0]) @ rot(rng.standard_normal(3), 0.7) for _ in range(5)]
mats = np.array(mats)
sv = np.linalg.svd(mats, compute_uv=False)
sgn = np.sign(np.linalg.det(mats))
signed = sv.copy()
signed[:, 2] *= np.where(sgn == 0, 1.0, sgn)
np.savez(os.path.join(os.path.dirname(os.path.abspath(__file__)), "svd_golden.npz"),
         F=mats, sv=sv, signed=signed)
print("wrote", len(mats), "cases")
