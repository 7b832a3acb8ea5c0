"""The oracle's ASCII frame writer and gauge sampler against the reference's
(fc2d_geoclaw_output_ascii_fort.f, geoclaw_gauges_fort.f90): byte-exact on
the golden frame (tests/golden/output_vectors.npz), exact on the golden gauge
samples, against the reference build on fresh random inputs when it is
present, and against what the format and the interpolation fix by
themselves."""
import os
import re
import tempfile

import numpy as np
import pytest

from oracle.oracle import CLib, Params, ref_lib

GOLD = os.path.join(os.path.dirname(__file__), "golden", "output_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.fixture(scope="module")
def clib():
    c = CLib()
    c.set_params(Params())
    return c


def frame_patches(gold):
    k = 0
    out = []
    while f"frame_q{k}" in gold:
        out.append((gold[f"frame_q{k}"].copy(), gold[f"frame_aux{k}"].copy(), gold["frame_geom"][k],
                    gold["frame_ints"][k]))
        k += 1
    return out


def test_frame_bytes_match_reference(gold, clib):
    hdr = clib.frame_header(float(gold["frame_time"][0]), 3, 1, int(gold["frame_hdr"][1]), 2)
    assert hdr == gold["frame_t_bytes"].tobytes()
    body = b"".join(clib.frame_patch(q, a, *[float(x) for x in g], *[int(x) for x in i])
                    for q, a, g, i in frame_patches(gold))
    assert body == gold["frame_q_bytes"].tobytes()


def test_frame_flushes_tiny_values_in_place(gold, clib):
    q, a, g, i = frame_patches(gold)[0]
    before = q.copy()
    clib.frame_patch(q, a, *[float(x) for x in g], *[int(x) for x in i])
    inner = (slice(2, -2), slice(2, -2))
    tiny = np.abs(before[inner]) < 1e-99
    assert tiny.any()
    assert np.all(q[inner][tiny] == 0.0) and not np.any(np.signbit(q[inner][tiny]))
    assert np.array_equal(q[inner][~tiny], before[inner][~tiny])
    ghost = np.ones(q.shape[:2], bool)
    ghost[2:-2, 2:-2] = False
    assert np.array_equal(q[ghost], before[ghost], equal_nan=True)


def _parse_E(field: str) -> float:
    s = field.strip()
    m = re.fullmatch(r"(-?)0\.(\d+)(?:E([+-]\d\d)|([+-]\d\d\d))", s)
    assert m, field
    exp = int(m.group(3) or m.group(4))
    return float(f"{m.group(1)}0.{m.group(2)}e{exp}")


@pytest.mark.parametrize("w,d", [(26, 16), (24, 16), (30, 20)])
def test_format_E_reads_back_rounded(clib, w, d):
    """Every field reads back as the value rounded to d significant digits,
    with the mantissa in [0.1, 1) and the width w."""
    rng = np.random.default_rng(5)
    vals = rng.normal(0, 1, 3000) * 10.0 ** rng.integers(-300, 300, 3000)
    for v in vals:
        s = clib.format_E(float(v), w, d)
        assert len(s) == w
        got = _parse_E(s)
        assert got == float(f"{v:.{d - 1}e}")
        assert s.strip().lstrip("-").startswith("0.") and s.strip().lstrip("-")[2] != "0"


def test_format_E_special_cases(clib):
    assert clib.format_E(0.0, 26, 16) == "    0.0000000000000000E+00"
    assert clib.format_E(-0.0, 24, 16) == " -0.0000000000000000E+00"
    assert clib.format_E(1.5e100, 26, 16) == "    0.1500000000000000+101"
    assert clib.format_E(1e-99, 26, 16) == "    0.1000000000000000E-98"
    assert clib.format_E(1234567890123456.5, 26, 16) == "    0.1234567890123456E+16"  # tie to even
    assert clib.format_E(1234567890123457.5, 26, 16) == "    0.1234567890123458E+16"
    assert clib.format_E(float("inf"), 24, 16).strip() == "Inf"
    assert clib.format_E(float("nan"), 24, 16).strip() == "NaN"


def test_gauge_samples_match_reference(gold, clib):
    for k in range(len(gold["gauge_out"])):
        qv, av = clib.update_gauge(gold["gauge_q"][k].copy(), gold["gauge_aux"][k].copy(),
                                   *gold["gauge_geom"][k], *gold["gauge_xy"][k])
        assert np.array_equal(qv, gold["gauge_out"][k, :3]), k
        assert av[0] == gold["gauge_out"][k, 3], k


def test_gauge_bilinear_is_exact_for_linear_fields(clib):
    """A wet linear field is reproduced at the gauge (to rounding); with a dry
    corner the containing cell's value is returned."""
    rng = np.random.default_rng(9)
    mx = 8
    for _ in range(200):
        xl, yl, dx, dy = rng.normal(), rng.normal(), rng.uniform(0.2, 2), rng.uniform(0.2, 2)
        c = rng.normal(0, 1, (4, 3))
        jj, ii = np.meshgrid(np.arange(mx + 4), np.arange(mx + 4), indexing="ij")
        xc_c = xl + (ii - 2 + 0.5) * dx
        yc_c = yl + (jj - 2 + 0.5) * dy
        q = np.stack([5.0 + c[0, m] * xc_c + c[1, m] * yc_c for m in range(3)], -1)
        q[..., 0] = 5.0 + 0.1 * xc_c + 0.2 * yc_c + 10
        aux = (c[2, 0] * xc_c + c[3, 0] * yc_c)[..., None]
        xg = xl + rng.uniform(0.5, mx - 0.5) * dx
        yg = yl + rng.uniform(0.5, mx - 0.5) * dy
        qv, av = clib.update_gauge(q, aux, xl, yl, dx, dy, xg, yg)
        exact = [5.0 + 0.1 * xg + 0.2 * yg + 10] + [5.0 + c[0, m] * xg + c[1, m] * yg for m in (1, 2)]
        np.testing.assert_allclose(qv, exact, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(av[0], c[2, 0] * xg + c[3, 0] * yg, atol=1e-11)
        ic = int((xg - xl) / dx) + 1 + 1
        jc = int((yg - yl) / dy) + 1 + 1
        q[jc + 1, ic + 1, 0] = 0.0           # one of the four cells dry
        qv, av = clib.update_gauge(q, aux, xl, yl, dx, dy, xg, yg)
        assert np.array_equal(qv, q[jc, ic]) and av[0] == aux[jc, ic, 0]


def test_gauge_file_lines_known_answers(clib):
    assert clib.gauge_header(7, 86.001, 0.001) == (
        b"# gauge_id=     7 location=(  8.6001000000e+01  1.0000000000e-03 ) num_eqn=  4\n"
        b"# Columns: level time h    hu    hv    eta\n")
    assert clib.gauge_line(2, 0.5, [1.25, -0.5, 0.0], [-1.25]) == (
        b"    2   5.0000000e-01   1.2500000e+00  -5.0000000e-01   0.0000000e+00   0.0000000e+00\n")
    assert clib.gauge_line(1, 1e3, [2.0, 3e-7, -4e120], [1e-100 - 2.0]) == (
        b"    1   1.0000000e+03   2.0000000e+00   3.0000000e-07 -4.0000000e+120   0.0000000e+00\n")


def test_frame_and_gauges_match_reference_build_fresh(clib):
    ref = ref_lib()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    ref.set_params(Params())
    rng = np.random.default_rng(11)
    pats = []
    for k in range(4):
        mx, my = int(rng.integers(1, 12)), int(rng.integers(1, 12))
        q = rng.normal(0, 3, (my + 4, mx + 4, 3)) * 10.0 ** rng.integers(-110, 110, (my + 4, mx + 4, 3))
        a = rng.normal(0, 3, (my + 4, mx + 4, 1))
        pats.append((q, a, float(rng.normal()), float(rng.normal()), float(rng.uniform(0.1, 1)),
                     float(rng.uniform(0.1, 1)), k, k, 0))
    tb, qb = ref.write_frame(tempfile.mkdtemp(), 42, 17.25, [(p[0].copy(), p[1].copy(), *p[2:]) for p in pats])
    assert tb == clib.frame_header(17.25, 3, 1, 4, 2)
    assert qb == b"".join(clib.frame_patch(p[0].copy(), p[1], *p[2:8], p[8], 0) for p in pats)
    for _ in range(300):
        q = np.abs(rng.normal(0, 1, (12, 12, 3)))
        q[..., 0][rng.random((12, 12)) < 0.1] = 1e-5
        a = rng.normal(0, 2, (12, 12, 1))
        g = (rng.normal(), rng.normal(), rng.uniform(0.1, 2), rng.uniform(0.1, 2))
        xy = (g[0] + rng.uniform(0, 8) * g[2], g[1] + rng.uniform(0, 8) * g[3])
        r1 = ref.update_gauge(q, a, *g, *xy)
        r2 = clib.update_gauge(q, a, *g, *xy)
        assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])
