"""GPU parity of the device supercell generator (lq_csr_from_hoppings) against
the oracle's restatement of wannier2sparse's wrap_in_supercell +
save_hopping_list_as_csr: the stored entries bitwise, and the completed
operator through SpMV bitwise against the same CSR loaded the usual way."""
import sys

import numpy as np
import pytest

import linqt_amd as lq

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _hoppings(rng, W, nh, reach, hermitian=True):
    cell = rng.integers(-reach, reach + 1, size=(nh, 3))
    edge = rng.integers(0, W, size=(nh, 2))
    val = rng.normal(size=nh) + 1j * rng.normal(size=nh)
    if hermitian:
        cell = np.concatenate([cell, -cell])
        edge = np.concatenate([edge, edge[:, ::-1]])
        val = np.concatenate([val, val.conj()])
    return cell, edge, val


def _check(ctx, oracle, W, cell, edge, val, dims, spmv=True):
    n, rowptr, col, v = oracle.supercell_csr(W, cell, edge, val, dims)
    A = lq.SparseMatrix.from_hoppings(ctx, W, cell, edge, val, dims)
    e = A.export()
    assert e.n == n and np.array_equal(e.rowptr, rowptr) and np.array_equal(e.col, col)
    assert np.array_equal(e.val.view(np.int64), v.view(np.int64))
    if spmv and n > 0:
        B = lq.SparseMatrix(ctx, n, rowptr, col, v)
        assert A.info() == B.info()
        rng = np.random.default_rng(n)
        x = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
        assert np.array_equal(A.Multiply(1.0, x), B.Multiply(1.0, x))
    return A


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 1, 1), (2, 2, 1), (3, 2, 2), (5, 4, 1), (7, 1, 3), (16, 16, 1)])
@pytest.mark.parametrize("W,nh,reach", [(1, 5, 1), (3, 14, 2), (4, 40, 3)])
def test_from_hoppings_matches_oracle(ctx, oracle, dims, W, nh, reach):
    say("supercell", dims, W, nh, reach)
    rng = np.random.default_rng(hash((dims, W, nh)) % 2**32)
    cell, edge, val = _hoppings(rng, W, nh, reach, hermitian=(nh % 2 == 0))
    _check(ctx, oracle, W, cell, edge, val, dims)


def test_from_hoppings_reference_wrap_fixture_and_edge_cases(ctx, oracle):
    say("supercell edge cases")
    hv = 1.0 + 1.0j
    cell = [(0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0)]
    edge = [(0, 1), (0, 1), (0, 1), (1, 0), (1, 0), (1, 0)]
    _check(ctx, oracle, 2, cell, edge, [hv] * 3 + [np.conj(hv)] * 3, (2, 2, 1))
    # summation order of wrapped duplicates follows the tag strings
    A = _check(ctx, oracle, 1, [(9, 0, 0), (10, 0, 0), (2, 0, 0)], [(0, 0)] * 3, [1e16, 1.0, -1e16], (1, 1, 1))
    assert A.export().val[0] == 0.0
    # all-zero list (the reference's empty "VY" files) and an empty list
    A = lq.SparseMatrix.from_hoppings(ctx, 2, [(1, 0, 0)], [(0, 1)], [0.0], (3, 1, 1))
    assert A.export().nnz == 0 and A.n == 6
    A = lq.SparseMatrix.from_hoppings(ctx, 2, np.zeros((0, 3)), np.zeros((0, 2)), [], (3, 1, 1))
    assert A.export().nnz == 0
    with pytest.raises(lq.LinqtError, match="outside"):
        lq.SparseMatrix.from_hoppings(ctx, 2, [(0, 0, 0)], [(0, 2)], [1.0], (1, 1, 1))
    with pytest.raises(lq.LinqtError, match="positive"):
        lq.SparseMatrix.from_hoppings(ctx, 2, [(0, 0, 0)], [(0, 1)], [1.0], (0, 1, 1))


def test_from_hoppings_large_supercell(ctx, oracle):
    """A 512 x 512 honeycomb-like model with second neighbours (N = 524288,
    the size of the Kubo benchmark) against the oracle, and rescaled."""
    say("supercell 512x512")
    W = 2
    cell, edge, val = [], [], []
    for (c, e, v) in [((0, 0, 0), (0, 1), -1.0), ((-1, 0, 0), (0, 1), -1.0), ((0, -1, 0), (0, 1), -1.0),
                      ((1, 0, 0), (0, 0), 0.1j), ((0, 1, 0), (0, 0), -0.1j), ((-1, 1, 0), (0, 0), 0.1j),
                      ((1, 0, 0), (1, 1), -0.1j), ((0, 1, 0), (1, 1), 0.1j), ((-1, 1, 0), (1, 1), -0.1j),
                      ((0, 0, 0), (0, 0), 0.3), ((0, 0, 0), (1, 1), -0.3)]:
        cell.append(c)
        edge.append(e)
        val.append(v)
        if c != (0, 0, 0) or e[0] != e[1]:
            cell.append(tuple(-x for x in c))
            edge.append(e[::-1])
            val.append(np.conj(v))
    A = _check(ctx, oracle, W, cell, edge, val, (512, 512, 1), spmv=False)
    assert A.info()["n"] == 524288
    x = np.random.default_rng(0).normal(size=524288) + 0j
    y = A.Multiply(1.0, x)
    n, rowptr, col, v = oracle.supercell_csr(W, cell, edge, val, (512, 512, 1))
    B = lq.SparseMatrix(ctx, n, rowptr, col, v)
    assert np.array_equal(y, B.Multiply(1.0, x))
    A.Rescale(1 / 4.6, 0.0)
    B.Rescale(1 / 4.6, 0.0)
    assert np.array_equal(A.Multiply(1.0, x), B.Multiply(1.0, x))
