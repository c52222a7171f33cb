"""Golden text of the driver's output files, from the reference's format strings.

Reads /root/reference/LBM.cu and main.cu as text (nothing is compiled,
imported or evaluated), takes the fprintf format strings of save_data_tecplot,
save_data_end, save_data_dmd and record_umax and the sprintf of the DMD file
name, and applies them with Python's printf-style % operator (C semantics for
%g, %f, %d) to a small synthetic field set.  The wall extrapolation applied
before formatting is restated from LBM.cu:1457-1472, 1582-1585 and 1690-1694.
Writes tests/golden/writers_golden.json: the inputs as hex floats, the format
strings found, and the expected file texts.  tests/test_writers.py compares
the library's writers (include/ehd_io.h) against it byte for byte.

Run once in the build container:  python tests/golden/gen_writer_golden.py
"""
from __future__ import annotations

import json
import os
import re

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "writers_golden.json")
NX, NY, NZ = 4, 3, 5
DX = DY = DZ = 0.25
T = 0.0123456789


def read(name):
    with open(os.path.join(REF, name)) as f:
        return f.read()


def section(text, start, end):
    i = text.index(start)
    return text[i:text.index(end, i + len(start))]


def fprintf_formats(body):
    """Every fprintf format string of a function body, in order."""
    return [bytes(m, "utf-8").decode("unicode_escape") for m in re.findall(r'fprintf\(\w+,\s*"((?:[^"\\]|\\.)*)"', body)]


def c_string(body, name):
    m = re.search(r'char\s+' + name + r'\[\]\s*=\s*"((?:[^"\\]|\\.)*)"', body)
    return bytes(m.group(1), "utf-8").decode("unicode_escape")


def synthetic_fields():
    """Nine fields with mixed signs and magnitudes (tiny values exercise the
    exponent form of %g, values near rounding boundaries exercise %10.6f)."""
    S = NX * NY * NZ
    i = np.arange(S, dtype=np.float64)
    f = np.empty((9, S))
    scales = [1600.0, 10.0, 1.0e4, 1e-3, 1e-4, 2e-3, 5e3, 5e3, 1e4]
    for k in range(9):
        f[k] = scales[k] * np.sin(0.37 * i + 1.1 * k)
    f[3, 7] = 1.234e-9           # exponent form of %g
    f[4, 8] = -0.0               # negative zero
    f[8, 9] = 0.0000005          # %10.6f rounding of 5e-7
    return f


def extrap(v, bottom=True, top=True):
    v = v.reshape(NZ, NY, NX).copy()
    if bottom:
        v[0] = 2.0 * v[1] - v[2]
    if top:
        v[NZ - 1] = 2.0 * v[NZ - 2] - v[NZ - 3]
    return v.reshape(-1)


def main():
    cu, mainc = read("LBM.cu"), read("main.cu")
    tec = section(cu, "void save_data_tecplot(FILE *fout", "void save_data_end(")
    end = section(cu, "void save_data_end(FILE *fend", "void save_data_dmd(")
    dmd = section(cu, "void save_data_dmd(FILE *fend", "void read_data(")
    umx = section(cu, "void record_umax(FILE *fend", "\n}\n")
    tec_fmt = fprintf_formats(tec)          # "%s\n", "\n", zone, row
    variables = c_string(tec, "str")
    end_fmt = fprintf_formats(end)[0]
    dmd_fmt = fprintf_formats(dmd)[0]
    umax_fmt = fprintf_formats(umx)[0]
    name_fmt = re.search(r'sprintf\(filename,\s*"([^"]*)"', mainc).group(1)
    assert len(tec_fmt) == 4, tec_fmt

    f = synthetic_fields()
    RHO, Q, PHI, UX, UY, UZ, EX, EY, EZ = range(9)
    c = f.copy()
    for k in (RHO, Q, UX, UY, UZ):
        c[k] = extrap(f[k])
    S = NX * NY * NZ

    def tecplot(first):
        out = (tec_fmt[0] % variables) if first else ""
        out += tec_fmt[1]
        out += tec_fmt[2] % (T, NX, NY, NZ)
        for z in range(NZ):
            for y in range(NY):
                for x in range(NX):
                    i = NX * (NY * z + y) + x
                    out += tec_fmt[3] % (DX * x, DY * y, DZ * z, c[UX][i], c[UY][i], c[UZ][i], c[RHO][i], c[Q][i],
                                         c[PHI][i], c[EX][i], c[EY][i], c[EZ][i])
        return out

    data_end = "".join(end_fmt % (T, c[UX][i], c[UY][i], c[UZ][i], c[RHO][i], c[Q][i], c[PHI][i], c[EX][i],
                                  c[EY][i], c[EZ][i]) for i in range(S))
    w = extrap(f[UZ])
    dmd_text = "".join(dmd_fmt % (T, w[i]) for i in range(S))
    wt = extrap(f[UZ], bottom=False)
    umax = 0.0
    for v in wt:
        umax = umax if umax > v else v
    umax_text = umax_fmt % (T, umax)

    golden = {
        "grid": [NX, NY, NZ], "spacing": [DX, DY, DZ], "t": T.hex(),
        "fields9_hex": [[float(v).hex() for v in row] for row in f],
        "formats": {"tecplot": tec_fmt, "variables": variables, "data_end": end_fmt, "dmd": dmd_fmt,
                    "umax": umax_fmt, "dmd_name": name_fmt},
        "tecplot_first": tecplot(True), "tecplot_next": tecplot(False), "data_end": data_end,
        "dmd": dmd_text, "umax": umax_text, "dmd_name": name_fmt % T,
    }
    with open(OUT, "w") as fo:
        json.dump(golden, fo, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
