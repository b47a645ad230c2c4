import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libkrans_hip.so)")


def principal_sine(U, K):
    """Largest principal-angle sine between span(U) and span(K) (orthonormal
    columns, same dimension): ||(I - K K^T) U||_2."""
    if U.shape[1] == 0 and K.shape[1] == 0:
        return 0.0
    R = U - K @ (K.T @ U)
    return float(np.linalg.norm(R, 2))


def orthonormal(B):
    Q, _ = np.linalg.qr(B)
    return Q
