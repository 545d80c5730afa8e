"""The numeric literals of Hunger's voigt_hjerting and the Harris series
(src/line_absorption.jl:251-310) in every restatement: the C oracle, the hydrogen oracle's Python
copy and the device header.  The fixture tests/golden/voigt_literals.json holds the reference's own
literals (written by tests/golden/make_voigt_literals.py): per function the signed fitted
coefficients (six or more significant digits) in order and the branch thresholds.

Most of these coefficients are pinned otherwise only to the approximation's own accuracy (the
Faddeeva bounds are 4e-2 to 0.13 near a = 1-1.4), so a slipped digit or sign in one of them would
pass those tests; here each copy must carry exactly the reference's values:
  * forward forms (oracle_harris_series, oracle_voigt_hjerting, harris_series64, voigt_hjerting64,
    voigt_core_f32, the Python copy): the same signed coefficients in the same order, and the same
    branch thresholds;
  * the Float32 Harris piece selector (harris_h1_f32), whose coefficients are interleaved: the same
    multiset;
  * the derivative forms (harris_series_dv, harris_series64_dv, the *_grad functions), whose
    coefficients appear scaled by their powers: every magnitude a coefficient of the reference's,
    and each piece's derivative carrying all of that piece's non-constant coefficients;
  * the precomputed constants beside them (sqrt(pi), 1/sqrt(pi), 1/sqrt(2), sqrt(2/pi)) equal to
    the Float64 values of those expressions.
"""
import json
import math
import os
import re
from collections import Counter

import pytest

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")
FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "voigt_literals.json")))
HARRIS = FIX["harris_series"]["fitted"]
PSI = FIX["voigt_hjerting"]["fitted"]
NUM = re.compile(r"(?<![\w.])(\d+\.\d+)f?(?![\w.])")
CMP = re.compile(r"(<=|>=|<|>)\s*\(?\s*(\d+(?:\.\d+)?)f?(?![\w.])")
# constants precomputed beside the coefficients: literals within 1e-9 of one of these are taken as
# that constant and must equal it to an ulp
KNOWN = [math.sqrt(math.pi), 1 / math.sqrt(math.pi), 1 / math.sqrt(2), math.sqrt(2 / math.pi),
         math.sqrt(2)]


def known(x):
    return any(abs(x - e) <= 1e-9 * e for e in KNOWN)


def exact_constant(x):
    return any(abs(x - e) <= 2.5e-16 * e for e in KNOWN)


def read(rel):
    return open(os.path.join(ROOT, rel)).read()


def c_body(src, name):
    """The body of the C/C++ function `name` (brace-matched), comments removed."""
    m = re.search(r"\b" + re.escape(name) + r"\s*\([^;{]*\)\s*\{", src)
    assert m, name
    i = m.end() - 1
    depth = 0
    for j in range(i, len(src)):
        depth += {"{": 1, "}": -1}.get(src[j], 0)
        if depth == 0:
            break
    text = re.sub(r"/\*.*?\*/", "", src[i:j + 1], flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def py_body(src, name):
    i = src.index("def " + name + "(")
    j = src.find("\ndef ", i + 1)
    text = src[i:j if j > 0 else None]
    return re.sub(r'""".*?"""', "", re.sub(r"#[^\n]*", "", text), flags=re.S)


def fitted(text):
    """Signed literals of six or more significant digits, in order (a literal after a minus sign,
    unary or binary, is negative); precomputed constants are returned separately."""
    out, consts = [], []
    for m in NUM.finditer(text):
        if len(m.group(1).replace(".", "").lstrip("0")) < 6:
            continue
        x = float(m.group(1))
        if known(x):
            consts.append(x)
            continue
        k = m.start() - 1
        while k >= 0 and text[k] in " (\n":
            k -= 1
        out.append(-x if k >= 0 and text[k] == "-" else x)
    return out, consts


def thresholds(text):
    return [float(m.group(2)) for m in CMP.finditer(text)]


ORACLE_C = "oracle/line_oracle.c"
DEVICE_H = "korg.jl_amd/csrc/korg_line_device.h"
HYDRO_PY = "oracle/hydrogen_oracle.py"


def test_fixture_shape():
    assert len(HARRIS) == 15 and len(PSI) == 4
    assert FIX["harris_series"]["thresholds"] == [1.3, 2.4]
    assert FIX["voigt_hjerting"]["thresholds"] == [0.2, 5.0, 0.2, 1.4, 3.2]


@pytest.mark.parametrize("path,name", [(ORACLE_C, "oracle_harris_series"),
                                       (DEVICE_H, "harris_series64")])
def test_harris_forward(path, name):
    text = c_body(read(path), name)
    assert fitted(text)[0] == HARRIS
    assert thresholds(text) == FIX["harris_series"]["thresholds"]


@pytest.mark.parametrize("path,name", [(ORACLE_C, "oracle_voigt_hjerting"),
                                       (DEVICE_H, "voigt_hjerting64"),
                                       (DEVICE_H, "voigt_core_f32")])
def test_voigt_forward(path, name):
    text = c_body(read(path), name)
    vals, consts = fitted(text)
    assert vals == PSI
    assert all(exact_constant(c) for c in consts)
    if name != "voigt_core_f32":  # the Float32 core takes its branch from the caller
        assert thresholds(text) == FIX["voigt_hjerting"]["thresholds"]


def test_python_copy():
    text = py_body(read(HYDRO_PY), "voigt_hjerting")
    assert fitted(text)[0] == HARRIS + PSI
    assert thresholds(text) == (FIX["harris_series"]["thresholds"] +
                                FIX["voigt_hjerting"]["thresholds"])


def test_f32_piece_selector():
    text = c_body(read(DEVICE_H), "harris_h1_f32")
    assert Counter(fitted(text)[0]) == Counter(HARRIS)


def test_branch_classifier():
    text = c_body(read("korg.jl_amd/csrc/korg_line.hip"), "hunger_branch")
    assert thresholds(text) == FIX["voigt_hjerting"]["thresholds"]


PIECES = [HARRIS[0:5], HARRIS[5:10], HARRIS[10:15]]


@pytest.mark.parametrize("path,name", [(ORACLE_C, "harris_series_dv"),
                                       (DEVICE_H, "harris_series64_dv"),
                                       (DEVICE_H, "voigt_hjerting_grad_f32")])
def test_harris_derivatives(path, name):
    text = c_body(read(path), name)
    mags = {abs(x) for x in fitted(text)[0]}
    ref = {abs(x) for x in HARRIS + PSI}
    assert mags <= ref, mags - ref
    # pieces 0 and 1: the derivative drops the constant term; piece 2 (a quotient) keeps it
    for k, piece in enumerate(PIECES):
        need = {abs(x) for x in (piece if k == 2 else piece[1:])}
        assert need <= mags, (k, need - mags)
    assert thresholds(text)[:2] in ([1.3, 2.4], [0.2, 5.0])


@pytest.mark.parametrize("path,name", [(ORACLE_C, "oracle_voigt_hjerting_grad"),
                                       (DEVICE_H, "voigt_hjerting64_grad"),
                                       (DEVICE_H, "voigt_hjerting_grad_f32")])
def test_voigt_derivatives(path, name):
    text = c_body(read(path), name)
    vals, consts = fitted(text)
    mags = {abs(x) for x in vals}
    assert {abs(x) for x in PSI} <= mags
    assert mags <= {abs(x) for x in HARRIS + PSI}
    assert all(exact_constant(c) for c in consts)
    assert thresholds(text)[:5] == FIX["voigt_hjerting"]["thresholds"]
