"""The library's output writers (fl_write_*, the reference's src/output.rs
formats) against the oracle's pure-Python restatement (oracle/output.py),
byte for byte, on random fields with special values; the number formats
pinned by hand on Rust's std::fmt rules; error paths.  Host-only calls: no
GPU needed."""

import math
import os

import numpy as np
import pytest

from flursys_amd import _lib, output as W
from oracle import output as OO


def test_number_formats_follow_rust_fmt():
    """{:.12e}: mantissa, 'e', exponent with no '+' or padding; {:.12}:
    fixed; NaN / inf spellings of Rust's Display and LowerExp."""
    assert OO.e12(1234.5) == "1.234500000000e3"
    assert OO.e12(-2.5e-7) == "-2.500000000000e-7"
    assert OO.e12(0.0) == "0.000000000000e0"
    assert OO.e12(-0.0) == "-0.000000000000e0"
    assert OO.e12(1.0) == "1.000000000000e0"
    assert OO.e12(1e100) == "1.000000000000e100"
    assert OO.e12(float("nan")) == "NaN" and OO.e12(float("-inf")) == "-inf"
    assert OO.f12(0.5) == "0.500000000000" and OO.f12(0.0078125) == "0.007812500000"
    assert OO.f12(float("inf")) == "inf"
    # ties at the 13th decimal round to even (2^-13 = 0.0001220703125)
    assert OO.f12(2.0 ** -13) == "0.000122070312"


def test_rust_display_of_f64():
    """Rust's `{}` of f64 (case_summary.txt): shortest digits, positional."""
    cases = {1.0: "1", 0.5: "0.5", 100.0: "100", 1e-7: "0.0000001", 1e21: "1" + "0" * 21,
             -0.0: "-0", 0.1: "0.1", 2.5e-5: "0.000025"}
    for x, s in cases.items():
        assert W.rust_display(x) == s


def fields(nx, ny, seed):
    rng = np.random.default_rng(seed)
    f = {k: rng.standard_normal((ny, nx)) * 10.0 ** rng.integers(-8, 8, (ny, nx))
         for k in ("p", "u", "v", "w", "t")}
    f["p"][0, 0] = -0.0
    f["u"][0, 1] = 0.0
    f["v"][1, 0] = 1e-310  # subnormal
    f["w"][1, 1] = 5e307
    f["t"][0, 2] = -1.0
    solid = np.zeros((ny, nx), np.uint8)
    solid[ny // 2, nx // 3] = 1
    return f, solid


@pytest.mark.parametrize("with_temp", [False, True])
def test_field_csv_and_vtk_match_the_oracle(tmp_path, with_temp):
    nx, ny, dx, dy = 7, 5, 1.0 / 7, 0.2
    f, solid = fields(nx, ny, 1)
    t = f["t"] if with_temp else None
    p = tmp_path / "field.csv"
    W.write_field_csv(str(p), nx, ny, dx, dy, solid, f["p"], f["u"], f["v"], f["w"], t)
    assert p.read_text() == OO.field_csv(nx, ny, dx, dy, solid, f["p"], f["u"], f["v"], f["w"], t)
    p = tmp_path / "field.vtk"
    W.write_legacy_vtk(str(p), "Lid-Driven Cavity", nx, ny, dx, dy, solid, f["p"], f["u"], f["v"],
                       f["w"], t)
    assert p.read_text() == OO.vtk_2d("Lid-Driven Cavity", nx, ny, dx, dy, solid, f["p"], f["u"],
                                      f["v"], f["w"], t)


def test_vtk_3d_matches_the_oracle(tmp_path):
    nx, ny, nz = 4, 3, 5
    rng = np.random.default_rng(2)
    u = rng.standard_normal((nz, ny, nx + 1))
    v = rng.standard_normal((nz, ny + 1, nx))
    w = rng.standard_normal((nz + 1, ny, nx))
    p = rng.standard_normal((nz, ny, nx))
    uc = 0.5 * (u[:, :, :-1] + u[:, :, 1:])
    vc = 0.5 * (v[:, :-1, :] + v[:, 1:, :])
    wc = 0.5 * (w[:-1] + w[1:])
    path = tmp_path / "field.vtk"
    W.write_legacy_vtk_3d(str(path), "FLURSYS 3D lid-driven cavity", nx, ny, nz, 0.25, 1 / 3, 0.2,
                          p, uc, vc, wc)
    assert path.read_text() == OO.vtk_3d("FLURSYS 3D lid-driven cavity", nx, ny, nz, 0.25, 1 / 3,
                                         0.2, p, u, v, w)


@pytest.mark.parametrize("symmetric", [False, True])
@pytest.mark.parametrize("kind", ["random", "constant", "nan", "all_solid"])
def test_ppm_matches_the_oracle(tmp_path, symmetric, kind):
    nx, ny = 9, 6
    f, solid = fields(nx, ny, 3)
    field = np.tanh(f["u"])
    if kind == "constant":
        field[:] = 0.25
    elif kind == "nan":
        field[2, 3] = np.nan
        field[1, 1] = np.inf
    elif kind == "all_solid":
        solid[:] = 1
    path = tmp_path / "f.ppm"
    W.write_ppm_frame(str(path), nx, ny, solid, field, symmetric)
    assert path.read_bytes() == OO.ppm(nx, ny, solid, field, symmetric)


def test_ppm_colours_by_hand():
    """t = 0, 1/2, 1 of both colour maps (output.rs:257-279)."""
    assert OO.colour(0.0, False) == (0, 0, 255)
    assert OO.colour(1.0, False) == (255, 255, 0)
    assert OO.colour(0.0, True) == (0, 0, 130)
    assert OO.colour(0.5, True) == (255, 245, 255)
    assert OO.colour(1.0, True) == (255, 0, 0)


def test_history_rows_and_columns(tmp_path):
    path = tmp_path / "history.csv"
    vals = [(5, 0.005, 1.5e-9, 37, 2e-12, 3e-4, 1.25, 0.0, 0.0, float("nan"), 0.3, 0.02, 1e-3),
            (10, 0.01, 2.5e-9, 41, 3e-12, 2e-4, 1.5, 1.2, -0.1, 0.5, 0.31, 0.02, 1e-3)]
    for v in vals:
        row = _lib.HistoryRow(*v, v[3], 0.0, 0.0, 0)
        W.append_history(str(path), row, extended=False)
    expect = OO.HISTORY_HEADER + "\n" + "\n".join(OO.history_row(*v) for v in vals) + "\n"
    assert path.read_text() == expect
    path = tmp_path / "history_ext.csv"
    W.append_history(str(path), _lib.HistoryRow(*vals[0], 1234, 5123.5, 0.64, 2), extended=True)
    lines = path.read_text().splitlines()
    assert lines[0] == OO.HISTORY_HEADER + ",cg_iters,hbm_GBps,roofline_frac,fallback_count"
    assert lines[1] == OO.history_row(*vals[0]) + ",1234," + OO.e12(5123.5) + "," + \
        OO.e12(0.64) + ",2"
    path = tmp_path / "cols.csv"
    x, y = np.linspace(0, 1, 4), np.array([1e-3, -2.0, np.nan, 7.0])
    W.write_columns(str(path), "y,u", [x, y], "fe")
    assert path.read_text() == "y,u\n" + "".join(f"{OO.f12(a)},{OO.e12(b)}\n" for a, b in zip(x, y))


def test_write_errors_carry_the_reference_text(tmp_path):
    bad = tmp_path / "missing" / "field.csv"
    z = np.zeros((2, 2))
    with pytest.raises(W.OutputError, match="Cannot create"):
        W.write_field_csv(str(bad), 2, 2, 0.5, 0.5, None, z, z, z, z)
    with pytest.raises(W.OutputError, match="Cannot create"):
        W.write_ppm_frame(str(bad), 2, 2, None, z, False)


def test_cylinder_forces_and_reattachment_by_hand():
    """Cd/Cl: one fluid cell next to one solid cell on each side picks up
    the pressure and the wall-shear terms of incompressible.rs:1634-1676;
    reattachment: the linear zero crossing of the wall shear (:1678-1705)."""
    p = np.array([[0.0, 2.0, 0.0]])
    u = np.array([[0.0, 0.5, 0.0]])
    v = np.array([[0.0, 0.25, 0.0]])
    solid = np.array([[1, 0, 1]])
    # east and west neighbours solid: fx = p dy - p dy = 0, fy = 2 mu v / (dx/2) dy
    cd, cl = W.cylinder_force_coefficients(p, u, v, solid, 1.0, 1.0, 0.1, 1.0, 1.0, 1.0)
    assert cd == 0.0 and cl == pytest.approx(2 * 0.1 * 0.25 / 0.5 / 0.5)
    ub = np.array([[0.0, -1.0, -1.0, 1.0, 2.0, 0.0]])
    xr = W.reattachment_length_ratio(ub, None, 1.0, 1.0, 0.5, 1.0, 1.0)
    # tau changes sign between x = 2.5 and 3.5 at the midpoint, 3.0
    assert xr == pytest.approx((3.0 - 0.5) / 1.0)
    assert math.isnan(W.reattachment_length_ratio(-np.abs(ub) - 1, None, 1.0, 1.0, 0.5, 1.0, 1.0))
