// On-disk outputs of the solvers in the reference's formats (src/output.rs),
// written from host arrays that the device fills (cell fields computed on
// the GPU and copied out once per snapshot).  Host code only.
//
// Number formats follow Rust's std::fmt, which the reference uses:
//   {:.12}   fixed, 12 decimals (printf %.12f: both round the exact binary
//            value to nearest, ties to even);
//   {:.12e}  12-decimal mantissa, exponent with no '+' and no leading zeros
//            ("1.234500000000e3", "-2.500000000000e-7");
//   NaN prints "NaN", infinities "inf" / "-inf" in both forms.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "common.hpp"

namespace fl {
namespace out {

void fmt_e12(std::string &s, double v) {
    if (std::isnan(v)) {
        s += "NaN";
        return;
    }
    if (std::isinf(v)) {
        s += v < 0 ? "-inf" : "inf";
        return;
    }
    char buf[64];
    snprintf(buf, sizeof buf, "%.12e", v);
    const char *e = strchr(buf, 'e');
    s.append(buf, (size_t)(e - buf));
    s += 'e';
    const char *p = e + 1;
    const bool neg = *p == '-';
    ++p;
    while (*p == '0' && p[1] != '\0') ++p;
    if (neg) s += '-';
    s += p;
}

void fmt_f12(std::string &s, double v) {
    if (std::isnan(v)) {
        s += "NaN";
        return;
    }
    if (std::isinf(v)) {
        s += v < 0 ? "-inf" : "inf";
        return;
    }
    char buf[400];
    snprintf(buf, sizeof buf, "%.12f", v);
    s += buf;
}

void fmt_i(std::string &s, long long v) { s += std::to_string(v); }

// Writes `data` to `path` (create/truncate), or appends.
fl_status put(const char *path, const std::string &data, bool append, char *err, int cap) {
    FILE *f = fopen(path, append ? "ab" : "wb");
    if (!f) {
        if (err && cap > 0) snprintf(err, cap, "Cannot create %s: %s", path, strerror(errno));
        return FL_ERR_IO;
    }
    const size_t n = fwrite(data.data(), 1, data.size(), f);
    const int rc = fclose(f);
    if (n != data.size() || rc != 0) {
        if (err && cap > 0) snprintf(err, cap, "Cannot write %s: %s", path, strerror(errno));
        return FL_ERR_IO;
    }
    return FL_OK;
}

// Rust's `as u8` of an f64: NaN -> 0, saturating, truncating.
inline unsigned char as_u8(double x) {
    if (std::isnan(x) || x <= 0.0) return 0;
    if (x >= 255.0) return 255;
    return (unsigned char)x;
}
// f64::clamp (NaN passes through)
inline double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }
inline double smoothstep(double e0, double e1, double x) {
    const double t = clamp01((x - e0) / (e1 - e0));
    return t * t * (3.0 - 2.0 * t);
}
// sequential_color / diverging_color (output.rs:257-279)
inline void sequential(double t, unsigned char *rgb) {
    rgb[0] = as_u8(255.0 * smoothstep(0.45, 1.0, t));
    rgb[1] = as_u8(255.0 * smoothstep(0.05, 0.85, t));
    rgb[2] = as_u8(255.0 * (1.0 - smoothstep(0.0, 0.7, t)));
}
inline void diverging(double t, unsigned char *rgb) {
    if (t < 0.5) {
        const double a = t / 0.5;
        rgb[0] = as_u8(240.0 * a);
        rgb[1] = as_u8(245.0 * a);
        rgb[2] = as_u8(130.0 + 125.0 * a);
    } else {
        const double a = (t - 0.5) / 0.5;
        rgb[0] = 255;
        rgb[1] = as_u8(245.0 * (1.0 - a));
        rgb[2] = as_u8(255.0 * (1.0 - a));
    }
}

}  // namespace out
}  // namespace fl

using namespace fl::out;

extern "C" {

fl_status fl_write_field_csv(const char *path, int64_t nx, int64_t ny, double dx, double dy,
                             const uint8_t *solid, const double *p, const double *u, const double *v,
                             const double *vorticity, const double *temperature, char *err,
                             int32_t errcap) {
    if (!path || nx <= 0 || ny <= 0 || !p || !u || !v || !vorticity) return FL_ERR_INVALID_ARGUMENT;
    std::string s;
    s.reserve((size_t)(nx * ny) * (temperature ? 130 : 110) + 64);
    s += temperature ? "i,j,x,y,solid,p,u,v,speed,vorticity,temperature\n" : "i,j,x,y,solid,p,u,v,speed,vorticity\n";
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            const int64_t c = i + nx * j;
            const double speed = std::sqrt(u[c] * u[c] + v[c] * v[c]);  // powi(2) is an exact square
            fmt_i(s, i);
            s += ',';
            fmt_i(s, j);
            s += ',';
            fmt_f12(s, ((double)i + 0.5) * dx);  // grid.cell_x (src/grid.rs:30-32)
            s += ',';
            fmt_f12(s, ((double)j + 0.5) * dy);
            s += ',';
            s += (solid && solid[c]) ? '1' : '0';
            for (double x : {p[c], u[c], v[c], speed, vorticity[c]}) {
                s += ',';
                fmt_e12(s, x);
            }
            if (temperature) {
                s += ',';
                fmt_e12(s, temperature[c]);
            }
            s += '\n';
        }
    return put(path, s, false, err, errcap);
}

fl_status fl_write_vtk_2d(const char *path, const char *title, int64_t nx, int64_t ny, double dx, double dy,
                          const uint8_t *solid, const double *p, const double *u, const double *v,
                          const double *vorticity, const double *temperature, char *err, int32_t errcap) {
    if (!path || !title || nx <= 0 || ny <= 0 || !p || !u || !v || !vorticity) return FL_ERR_INVALID_ARGUMENT;
    const int64_t n = nx * ny;
    std::string s;
    s.reserve((size_t)n * 120 + 512);
    s += "# vtk DataFile Version 3.0\n";
    s += title;
    s += "\nASCII\nDATASET STRUCTURED_POINTS\nDIMENSIONS ";
    fmt_i(s, nx);
    s += ' ';
    fmt_i(s, ny);
    s += " 1\nORIGIN ";
    fmt_f12(s, 0.5 * dx);
    s += ' ';
    fmt_f12(s, 0.5 * dy);
    s += " 0\nSPACING ";
    fmt_f12(s, dx);
    s += ' ';
    fmt_f12(s, dy);
    s += " 1\nPOINT_DATA ";
    fmt_i(s, n);
    s += '\n';
    auto scalars = [&](const char *name, auto &&value) {
        s += "SCALARS ";
        s += name;
        s += " double 1\nLOOKUP_TABLE default\n";
        for (int64_t c = 0; c < n; ++c) {
            fmt_e12(s, value(c));
            s += '\n';
        }
    };
    scalars("pressure", [&](int64_t c) { return p[c]; });
    scalars("speed", [&](int64_t c) { return std::sqrt(u[c] * u[c] + v[c] * v[c]); });
    scalars("vorticity", [&](int64_t c) { return vorticity[c]; });
    if (temperature) scalars("temperature", [&](int64_t c) { return temperature[c]; });
    s += "SCALARS solid int 1\nLOOKUP_TABLE default\n";
    for (int64_t c = 0; c < n; ++c) s += (solid && solid[c]) ? "1\n" : "0\n";
    s += "VECTORS velocity double\n";
    for (int64_t c = 0; c < n; ++c) {
        fmt_e12(s, u[c]);
        s += ' ';
        fmt_e12(s, v[c]);
        s += " 0\n";
    }
    return put(path, s, false, err, errcap);
}

fl_status fl_write_vtk_3d(const char *path, const char *title, int64_t nx, int64_t ny, int64_t nz, double dx,
                          double dy, double dz, const double *p, const double *uc, const double *vc,
                          const double *wc, char *err, int32_t errcap) {
    if (!path || !title || nx <= 0 || ny <= 0 || nz <= 0 || !p || !uc || !vc || !wc)
        return FL_ERR_INVALID_ARGUMENT;
    const int64_t n = nx * ny * nz;
    std::string s;
    s.reserve((size_t)n * 80 + 512);
    s += "# vtk DataFile Version 3.0\n";
    s += title;
    s += "\nASCII\nDATASET STRUCTURED_POINTS\nDIMENSIONS ";
    fmt_i(s, nx);
    s += ' ';
    fmt_i(s, ny);
    s += ' ';
    fmt_i(s, nz);
    s += "\nORIGIN ";
    fmt_f12(s, 0.5 * dx);
    s += ' ';
    fmt_f12(s, 0.5 * dy);
    s += ' ';
    fmt_f12(s, 0.5 * dz);
    s += "\nSPACING ";
    fmt_f12(s, dx);
    s += ' ';
    fmt_f12(s, dy);
    s += ' ';
    fmt_f12(s, dz);
    s += "\nPOINT_DATA ";
    fmt_i(s, n);
    s += "\nSCALARS pressure double 1\nLOOKUP_TABLE default\n";
    for (int64_t c = 0; c < n; ++c) {
        fmt_e12(s, p[c]);
        s += '\n';
    }
    s += "VECTORS velocity double\n";
    for (int64_t c = 0; c < n; ++c) {
        fmt_e12(s, uc[c]);
        s += ' ';
        fmt_e12(s, vc[c]);
        s += ' ';
        fmt_e12(s, wc[c]);
        s += '\n';
    }
    return put(path, s, false, err, errcap);
}

fl_status fl_write_ppm(const char *path, int64_t nx, int64_t ny, const uint8_t *solid, const double *field,
                       int32_t symmetric, char *err, int32_t errcap) {
    if (!path || nx <= 0 || ny <= 0 || !field) return FL_ERR_INVALID_ARGUMENT;
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t c = 0; c < nx * ny; ++c)
        if (!(solid && solid[c]) && std::isfinite(field[c])) {
            lo = std::fmin(lo, field[c]);
            hi = std::fmax(hi, field[c]);
        }
    if (!std::isfinite(lo) || !std::isfinite(hi)) {
        lo = 0.0;
        hi = 1.0;
    } else if (symmetric) {
        const double bound = std::fmax(std::fmax(std::fabs(lo), std::fabs(hi)), 1.0e-14);
        lo = -bound;
        hi = bound;
    } else if (hi <= lo) {
        hi = lo + 1.0;
    }
    std::string s = "P6\n" + std::to_string(nx) + " " + std::to_string(ny) + "\n255\n";
    const size_t head = s.size();
    s.resize(head + (size_t)(3 * nx * ny));
    unsigned char *px = (unsigned char *)&s[head];
    for (int64_t jr = 0; jr < ny; ++jr) {  // image rows top to bottom
        const int64_t j = ny - 1 - jr;
        for (int64_t i = 0; i < nx; ++i, px += 3) {
            const int64_t c = i + nx * j;
            if (solid && solid[c]) {
                px[0] = px[1] = px[2] = 20;
                continue;
            }
            const double t = clamp01((field[c] - lo) / (hi - lo));
            if (symmetric) diverging(t, px);
            else sequential(t, px);
        }
    }
    return put(path, s, false, err, errcap);
}

fl_status fl_write_csv_columns(const char *path, const char *header, int32_t ncols, int64_t nrows,
                               const double *const *cols, const char *fmt, char *err, int32_t errcap) {
    if (!path || !header || ncols <= 0 || nrows < 0 || !cols || !fmt || (int32_t)strlen(fmt) != ncols)
        return FL_ERR_INVALID_ARGUMENT;
    for (int32_t c = 0; c < ncols; ++c)
        if (!cols[c] || (fmt[c] != 'f' && fmt[c] != 'e')) return FL_ERR_INVALID_ARGUMENT;
    std::string s = header;
    s += '\n';
    for (int64_t r = 0; r < nrows; ++r) {
        for (int32_t c = 0; c < ncols; ++c) {
            if (c) s += ',';
            if (fmt[c] == 'f') fmt_f12(s, cols[c][r]);
            else fmt_e12(s, cols[c][r]);
        }
        s += '\n';
    }
    return put(path, s, false, err, errcap);
}

fl_status fl_append_history(const char *path, const fl_history_row *r, int32_t extended, char *err,
                            int32_t errcap) {
    if (!path || !r) return FL_ERR_INVALID_ARGUMENT;
    std::string s;
    FILE *probe = fopen(path, "rb");
    if (probe) fclose(probe);
    else {
        s += "step,time,pressure_residual,pressure_iterations,max_divergence,velocity_change,max_speed,cd,cl,"
             "reattachment_x_over_h,momentum_cfl,viscous_diffusion_number,time_step";
        if (extended) s += ",cg_iters,hbm_GBps,roofline_frac,fallback_count";
        s += '\n';
    }
    fmt_i(s, r->step);
    s += ',';
    fmt_f12(s, r->time);
    s += ',';
    fmt_e12(s, r->pressure_residual);
    s += ',';
    fmt_i(s, r->pressure_iterations);
    for (double x : {r->max_divergence, r->velocity_change, r->max_speed, r->cd, r->cl, r->reattachment,
                     r->momentum_cfl, r->viscous_diffusion_number, r->time_step}) {
        s += ',';
        fmt_e12(s, x);
    }
    if (extended) {
        s += ',';
        fmt_i(s, r->cg_iters);
        s += ',';
        fmt_e12(s, r->hbm_GBps);
        s += ',';
        fmt_e12(s, r->roofline_frac);
        s += ',';
        fmt_i(s, r->fallback_count);
    }
    s += '\n';
    return put(path, s, true, err, errcap);
}

}  // extern "C"
