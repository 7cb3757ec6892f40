import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU and the built liblinqt_amd.so (runs through the C ABI)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    if not os.path.isfile(O._LIB_PATH):
        O.build()
    return O


@pytest.fixture(scope="session")
def ctx():
    import linqt_amd as lq

    c = lq.Context(0)
    yield c
    c.close()


def random_upper_csr(n, density, seed, diag=True, complex_diag=False):
    """Random Hermitian-upper CSR (entries col >= row) as (rowptr, col, val)."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        js = [j for j in range(i + 1, n) if rng.random() < density]
        if diag:
            js = [i] + js
        for j in js:
            v = rng.normal() + 1j * rng.normal()
            if j == i:
                v = complex(v.real, v.imag if complex_diag else 0.0)
            rows.append(i)
            cols.append(j)
            vals.append(v)
    rowptr = np.zeros(n + 1, dtype=np.int32)
    for r in rows:
        rowptr[r + 1] += 1
    rowptr = np.cumsum(rowptr).astype(np.int32)
    return rowptr, np.array(cols, dtype=np.int32), np.array(vals, dtype=np.complex128)


def dense_herm(n, rowptr, col, val):
    d = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        for q in range(rowptr[i], rowptr[i + 1]):
            j = col[q]
            if j < i:
                continue
            d[i, j] += val[q]
            if j != i:
                d[j, i] += np.conj(val[q])
    return d


def rel_inf(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))
