"""Oracle: the reference's on-disk output formats (src/output.rs and the
history row of src/solver/incompressible.rs:1744-1767), restated in pure
Python.  TEST INFRASTRUCTURE ONLY -- the library's C++ writers are checked
byte for byte against these; nothing here is part of the shipped path.

Number formats of Rust's std::fmt used by the reference:
  {:.12}   fixed with 12 decimals (Python's '.12f': both round the exact
           binary value to nearest, ties to even);
  {:.12e}  12-decimal mantissa, 'e', then the exponent as a plain integer
           (no '+', no zero padding): 1234.5 -> "1.234500000000e3";
  NaN -> "NaN", infinities -> "inf" / "-inf" in both.
"""

import math

import numpy as np


def f12(x):
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    return format(x, ".12f")


def e12(x):
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    mant, exp = format(x, ".12e").split("e")
    return f"{mant}e{int(exp)}"


def field_csv(nx, ny, dx, dy, solid, p, u, v, vort, temperature=None):
    """write_field_csv (output.rs:31-67) as text."""
    head = "i,j,x,y,solid,p,u,v,speed,vorticity" + (",temperature" if temperature is not None else "")
    out = [head]
    for j in range(ny):
        for i in range(nx):
            speed = math.sqrt(u[j, i] ** 2 + v[j, i] ** 2)
            row = [str(i), str(j), f12((i + 0.5) * dx), f12((j + 0.5) * dy),
                   "1" if solid is not None and solid[j, i] else "0"]
            row += [e12(a) for a in (p[j, i], u[j, i], v[j, i], speed, vort[j, i])]
            if temperature is not None:
                row.append(e12(temperature[j, i]))
            out.append(",".join(row))
    return "\n".join(out) + "\n"


def vtk_2d(title, nx, ny, dx, dy, solid, p, u, v, vort, temperature=None):
    """write_legacy_vtk (output.rs:69-125) as text."""
    out = ["# vtk DataFile Version 3.0", title, "ASCII", "DATASET STRUCTURED_POINTS",
           f"DIMENSIONS {nx} {ny} 1", f"ORIGIN {f12(0.5 * dx)} {f12(0.5 * dy)} 0",
           f"SPACING {f12(dx)} {f12(dy)} 1", f"POINT_DATA {nx * ny}"]

    def scalars(name, vals):
        out.append(f"SCALARS {name} double 1")
        out.append("LOOKUP_TABLE default")
        out.extend(e12(a) for a in vals)

    cells = [(j, i) for j in range(ny) for i in range(nx)]
    scalars("pressure", [p[c] for c in cells])
    scalars("speed", [math.sqrt(u[c] ** 2 + v[c] ** 2) for c in cells])
    scalars("vorticity", [vort[c] for c in cells])
    if temperature is not None:
        scalars("temperature", [temperature[c] for c in cells])
    out += ["SCALARS solid int 1", "LOOKUP_TABLE default"]
    out.extend("1" if solid is not None and solid[c] else "0" for c in cells)
    out.append("VECTORS velocity double")
    out.extend(f"{e12(u[c])} {e12(v[c])} 0" for c in cells)
    return "\n".join(out) + "\n"


def vtk_3d(title, nx, ny, nz, dx, dy, dz, p, u, v, w):
    """write_legacy_vtk_3d (output.rs:127-183) from the staggered u, v, w
    (shapes (nz, ny, nx+1), (nz, ny+1, nx), (nz+1, ny, nx)): the face means
    are formed here, as the reference does."""
    out = ["# vtk DataFile Version 3.0", title, "ASCII", "DATASET STRUCTURED_POINTS",
           f"DIMENSIONS {nx} {ny} {nz}", f"ORIGIN {f12(0.5 * dx)} {f12(0.5 * dy)} {f12(0.5 * dz)}",
           f"SPACING {f12(dx)} {f12(dy)} {f12(dz)}", f"POINT_DATA {nx * ny * nz}",
           "SCALARS pressure double 1", "LOOKUP_TABLE default"]
    out.extend(e12(a) for a in p.ravel())
    out.append("VECTORS velocity double")
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                uc = 0.5 * (u[k, j, i] + u[k, j, i + 1])
                vc = 0.5 * (v[k, j, i] + v[k, j + 1, i])
                wc = 0.5 * (w[k, j, i] + w[k + 1, j, i])
                out.append(f"{e12(uc)} {e12(vc)} {e12(wc)}")
    return "\n".join(out) + "\n"


def _as_u8(x):
    """Rust's `f64 as u8`: NaN -> 0, saturating, truncating toward zero."""
    if math.isnan(x) or x <= 0.0:
        return 0
    return 255 if x >= 255.0 else int(x)


def _clamp01(x):
    return x if math.isnan(x) else min(max(x, 0.0), 1.0)


def _smooth(e0, e1, x):
    t = _clamp01((x - e0) / (e1 - e0))
    return t * t * (3.0 - 2.0 * t)


def colour(t, symmetric):
    """sequential_color / diverging_color (output.rs:257-279)."""
    if symmetric:
        if t < 0.5:
            a = t / 0.5
            return (_as_u8(240.0 * a), _as_u8(245.0 * a), _as_u8(130.0 + 125.0 * a))
        a = (t - 0.5) / 0.5
        return (255, _as_u8(245.0 * (1.0 - a)), _as_u8(255.0 * (1.0 - a)))
    return (_as_u8(255.0 * _smooth(0.45, 1.0, t)), _as_u8(255.0 * _smooth(0.05, 0.85, t)),
            _as_u8(255.0 * (1.0 - _smooth(0.0, 0.7, t))))


def ppm(nx, ny, solid, field, symmetric):
    """write_ppm_frame (output.rs:201-255) as bytes."""
    lo, hi = math.inf, -math.inf
    for j in range(ny):
        for i in range(nx):
            if not (solid is not None and solid[j, i]):
                val = float(field[j, i])
                if math.isfinite(val):
                    lo, hi = min(lo, val), max(hi, val)
    if not (math.isfinite(lo) and math.isfinite(hi)):
        lo, hi = 0.0, 1.0
    elif symmetric:
        b = max(abs(lo), abs(hi), 1.0e-14)
        lo, hi = -b, b
    elif hi <= lo:
        hi = lo + 1.0
    body = bytearray()
    for jr in range(ny):
        j = ny - 1 - jr
        for i in range(nx):
            if solid is not None and solid[j, i]:
                body += bytes((20, 20, 20))
            else:
                body += bytes(colour(_clamp01((float(field[j, i]) - lo) / (hi - lo)), symmetric))
    return f"P6\n{nx} {ny}\n255\n".encode() + bytes(body)


HISTORY_HEADER = ("step,time,pressure_residual,pressure_iterations,max_divergence,velocity_change,"
                  "max_speed,cd,cl,reattachment_x_over_h,momentum_cfl,viscous_diffusion_number,"
                  "time_step")


def history_row(step, time, pres, pits, div, dv, umax, cd, cl, xr, cfl, dnu, dt):
    """append_histories' row format (incompressible.rs:1744-1767)."""
    return ",".join([str(step), f12(time), e12(pres), str(pits)] +
                    [e12(a) for a in (div, dv, umax, cd, cl, xr, cfl, dnu, dt)])


def to_grid(a, nx, ny):
    return np.asarray(a, dtype=np.float64).reshape(ny, nx)
