"""Test-side numpy model of the split-precision (INT8 digit) contraction.

Mirrors the algorithm of linqt_amd/csrc/lq_ozaki.hip step by step so the GPU
result can be checked against it to FP64 rounding of the final fold, and
gives the a-priori error bound of the method:

* every row is cut into groups of GROUP elements; e = 1 + the exponent
  with max(|re|, |im|) < 2^(e-1) over the group (frexp), a = x / 2^e, |a| < 1/2;
* S digits by rounding to nearest: y = 128 a, d = rint(y), a = y - d, so
  |d| <= 64 and every remainder is within 1/2;
* products of digit planes t, u (1-based) with t + u <= S + 1 are summed
  exactly (int64 here, int32 on the device) and weighted 2^(eA+eB-7(t+u)).

Error of one real product of normalised values (|a|, |b| < 1/2): the rounding
remainders (<= 2^-7S / 2 each) give <= 2^-7S / 2, the dropped levels
t + u >= S + 2 give <= (S - 1) 2^-7S / 4, so a complex component (two real
products) is within (S + 1) 2^-7S / 2; summed over a group this is scaled by
2^(eA+eB) * len(group).  Rounding, unlike truncation, leaves remainders whose
sign does not follow x, so correlated rows (L ~ R) do not add them coherently.
"""
from __future__ import annotations

import numpy as np

GROUP = 4096


def slice_rows(X: np.ndarray, S: int):
    """X: rows x K complex128 -> (Dre, Dim) int64 [rows][S][Kp] and E int [rows][ngroups]."""
    rows, K = X.shape
    Kp = -(-K // GROUP) * GROUP
    ng = Kp // GROUP
    Xp = np.zeros((rows, Kp), dtype=np.complex128)
    Xp[:, :K] = X
    G = Xp.reshape(rows, ng, GROUP)
    mx = np.maximum(np.abs(G.real), np.abs(G.imag)).max(axis=2)
    _, E = np.frexp(mx)
    E = np.where(mx > 0, E + 1, 0).astype(np.int64)
    scale = np.ldexp(1.0, -E)[:, :, None]
    ar = (G.real * scale).reshape(rows, Kp)
    ai = (G.imag * scale).reshape(rows, Kp)
    Dre = np.zeros((rows, S, Kp), dtype=np.int64)
    Dim = np.zeros((rows, S, Kp), dtype=np.int64)
    for t in range(S):
        yr, yi = ar * 128.0, ai * 128.0
        dr, di = np.rint(yr), np.rint(yi)
        Dre[:, t], Dim[:, t] = dr.astype(np.int64), di.astype(np.int64)
        ar, ai = yr - dr, yi - di
    return Dre, Dim, E


def contract(A: np.ndarray, B: np.ndarray, S: int) -> np.ndarray:
    """Model of sum_k conj(A[m0][k]) B[m1][k] with S digit planes."""
    Ar, Ai, EA = slice_rows(A, S)
    Br, Bi, EB = slice_rows(B, S)
    M0, M1 = A.shape[0], B.shape[0]
    ng = EA.shape[1]
    out = np.zeros((M0, M1), dtype=np.complex128)
    for g in range(ng):
        sl = slice(g * GROUP, (g + 1) * GROUP)
        sr = np.zeros((M0, M1))
        si = np.zeros((M0, M1))
        for lev in range(S - 1, -1, -1):  # 0-based t + u = lev
            lr = np.zeros((M0, M1), dtype=np.int64)
            li = np.zeros((M0, M1), dtype=np.int64)
            for t in range(lev + 1):
                u = lev - t
                a_r, a_i = Ar[:, t, sl], Ai[:, t, sl]
                b_r, b_i = Br[:, u, sl], Bi[:, u, sl]
                lr += a_r @ b_r.T + a_i @ b_i.T
                li += a_r @ b_i.T - a_i @ b_r.T
            sr += np.ldexp(lr.astype(np.float64), -7 * (lev + 2))
            si += np.ldexp(li.astype(np.float64), -7 * (lev + 2))
        w = np.ldexp(1.0, EA[:, g][:, None] + EB[:, g][None, :])
        out += (sr + 1j * si) * w
    return out


def error_bound(A: np.ndarray, B: np.ndarray, S: int) -> np.ndarray:
    """Per-element bound on |model - exact| for each of re and im."""
    _, _, EA = slice_rows(A, 1)
    _, _, EB = slice_rows(B, 1)
    K = A.shape[1]
    ng = EA.shape[1]
    cnt = np.array([min(GROUP, K - g * GROUP) for g in range(ng)], dtype=np.float64)
    scale = np.ldexp(1.0, EA[:, None, :] + EB[None, :, :])  # M0 x M1 x ng
    return 0.5 * (S + 1) * 2.0 ** (-7 * S) * (scale * cnt).sum(axis=2)


def magnitude(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """sum_g 2^(eA+eB) len(g): the scale the model's fold rounding is relative to."""
    return error_bound(A, B, 1) / (2.0 ** -7)
