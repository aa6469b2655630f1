"""Pins the oracle's Gauss-Legendre rules (LegendreMatrix.h:9-10): textbook routine, exactness, and the
reference's own table text when the reference tree is present."""
import os
import re

import numpy as np
import pytest

import oracle as O

REF_TABLE = "/root/reference/src/LegendreMatrix/LegendreMatrix.cxx"


@pytest.mark.parametrize("n", list(range(1, 101)))
def test_matches_numpy_leggauss(n):
    x, w = O.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xr)) < 1e-14
    assert np.max(np.abs(w - wr)) < 1e-14


@pytest.mark.parametrize("n", [1, 2, 3, 7, 20, 59, 80, 100])
def test_exact_on_polynomials(n):
    x, w = O.gauss_legendre(n)
    assert abs(w.sum() - 2.0) < 1e-13
    for deg in range(0, 2 * n):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 1e-12, (n, deg)


def test_symmetry():
    for n in range(1, 101):
        x, w = O.gauss_legendre(n)
        assert np.all(x == -x[::-1])
        assert np.all(w == w[::-1])
        assert np.all(np.diff(x) > 0)


def _parse_reference_table(name):
    text = open(REF_TABLE).read()
    start = text.index(name)
    body = text[start:]
    body = body[body.index("{") + 1:]
    rows, depth, cur = [], 0, ""
    for ch in body:
        if ch == "{":
            depth += 1
            cur = ""
        elif ch == "}":
            if depth == 0:
                break
            depth -= 1
            rows.append([float(v) for v in re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?", cur)])
        elif depth == 1:
            cur += ch
    return rows


@pytest.mark.skipif(not os.path.exists(REF_TABLE), reason="reference tree not present")
def test_matches_reference_table_text():
    xs = _parse_reference_table("positionMatrix")
    ws = _parse_reference_table("weightsMatrix")
    assert len(xs) == 100 and len(ws) == 100
    for n in range(1, 101):
        x, w = O.gauss_legendre(n)
        assert np.max(np.abs(x - np.array(xs[n - 1]))) < 1e-14, n
        assert np.max(np.abs(w - np.array(ws[n - 1]))) < 1e-14, n
