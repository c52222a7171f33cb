"""Parse the index and sign tables of the reference step out of its source text.

Reads /root/reference/LBM.h and LBM.cu as text (nothing is compiled, imported
or evaluated) and writes tests/golden/reference_tables.json: integer tables
only -- direction vectors of the pull stream, bounce-back and injection maps,
the moving-wall signs, the summation lists of the moments, the signed term
order of every c.u in the equilibria and the term structure of every Guo
force population.  tests/test_oracle.py checks the oracle against it.

Run once in the build container:  python tests/golden/gen_reference_tables.py
"""
from __future__ import annotations

import json
import os
import re
import sys

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")

OFF = {"xm1": -1, "x": 0, "xp1": 1, "ym1": -1, "y": 0, "yp1": 1, "zm1": -1, "z": 0, "zp1": 1}
AX = {"x": 1, "y": 2, "z": 3}
CLS = {"s": 1, "a": 2, "d": 3, "0": 0}


def read(name):
    with open(os.path.join(REF, name)) as f:
        return f.read()


def section(text, start, end):
    i = text.index(start)
    j = text.index(end, i + len(start))
    return text[i:j]


def weights(h):
    out = {}
    for name, a, b in re.findall(r"double\s+(w0|ws|wa|wd)\s*=\s*([0-9.]+)\s*/\s*([0-9.]+)\s*;", h):
        out[name] = [float(a), float(b)]
    return out


def stream_vectors(cu):
    body = section(cu, "__global__ void gpu_stream(double *f0", "__global__ void gpu_bc_charge")
    res = {}
    for dst, src in (("f1", "f2"), ("h1", "h2")):
        pat = (dst + r"\[gpu_fieldn_index\(x, y, z, (\d+)\)\] = " + src +
               r"\[gpu_fieldn_index\((\w+), (\w+), (\w+), (\d+)\)\];")
        vec = {}
        for d, xs, ys, zs, d2 in re.findall(pat, body):
            assert d == d2
            vec[int(d)] = [-OFF[xs], -OFF[ys], -OFF[zs]]
        res[dst] = vec
    return res


def boundary(cu):
    body = section(cu, "__global__ void gpu_boundary(double *f0", "__global__ void gpu_stream(double *f0")
    bottom = {}
    for a, b in re.findall(r"f2\[gpu_fieldn_index\(x, y, 0, (\d+)\)\] = f1\[gpu_fieldn_index\(x, y, 0, (\d+)\)\];", body):
        bottom[int(a)] = int(b)
    top = {}
    pat = (r"f2\[gpu_fieldn_index\(x, y, NZ - 1, (\d+)\)\] = f1\[gpu_fieldn_index\(x, y, NZ - 1, (\d+)\)\]"
           r"\s*(?:([+-])\s*multi([sad]))?\s*;")
    for a, b, sg, cl in re.findall(pat, body):
        top[int(a)] = [int(b), (1 if sg == "+" else -1) if sg else 0, CLS[cl] if cl else 0]
    hcopy = {}
    for a, b in re.findall(r"h2\[gpu_fieldn_index\(x, y, NZ - 1, (\d+)\)\] = h2\[gpu_fieldn_index\(x, y, NZ - 2, (\d+)\)\];", body):
        hcopy[int(a)] = int(b)
    return bottom, top, hcopy


def bc_charge(cu):
    body = section(cu, "__global__ void gpu_bc_charge", "__host__ void compute_parameters")
    out = {}
    for a, b, cl in re.findall(r"h1\[gpu_fieldn_index\(x, y, 0, (\d+)\)\] = -ht(\d+) \+ multi([sad])c;", body):
        out[int(a)] = [int(b), CLS[cl]]
    rest = re.findall(r"h0\[gpu_field0_index\(x, y, 0\)\] = -h0\[gpu_field0_index\(x, y, 0\)\] \+ multi0c;", body)
    return out, len(rest)


def moments(cu):
    body = section(cu, "__global__ void gpu_collide_save(double *f0", "__global__ void gpu_boundary(double *f0")
    out = {}
    for ax in "xyz":
        m = re.search(r"double u" + ax + r" = rhoinv\*\(\((.*?)\)\) / CFL", body, re.S)
        txt = m.group(1)
        pos_txt, neg_txt = txt.split("- (")
        out[ax] = [[int(v) for v in re.findall(r"ft(\d+)", pos_txt)], [int(v) for v in re.findall(r"ft(\d+)", neg_txt)]]
    return out


def cdot_terms(expr):
    terms = []
    for sg, ax in re.findall(r"([+-]?)\s*tu([xyz])(?:_c)?", expr):
        terms.append((-1 if sg == "-" else 1) * AX[ax])
    return terms


def equilibria(text, start, end, fe_pat, he_pat):
    body = section(text, start, end)
    lines = body.split("\n")
    ef, eh = {}, {}
    cur = None
    for ln in lines:
        m = re.search(r"cidot3u = (.*?);", ln)
        if m:
            cur = cdot_terms(m.group(1))
            continue
        mf = re.search(fe_pat, ln)
        mh = re.search(he_pat, ln)
        if mf and cur is not None:
            ef[int(mf.group(1))] = cur
            cur = None
        elif mh and cur is not None:
            eh[int(mh.group(1))] = cur
            cur = None
    return ef, eh


A_TERM = re.compile(r"\(\((-?)cflinv\s*-\s*u([xyz])\)\s*\+\s*(?:\(cflinv2\*u([xyz])\)|\(([^()]*)\)\*cflinv2)\)\*force([xyz])")
U_TERM = re.compile(r"u([xyz])\*force([xyz])")


def fpops(cu):
    body = section(cu, "__global__ void gpu_collide_save(double *f0", "__global__ void gpu_boundary(double *f0")
    out = {}
    for d, coe, expr in re.findall(r"double fpop(\d+)\s*=\s*(-?coe[0sad])\s*\*\s*\((.*?)\);\n", body):
        d = int(d)
        toks = []
        i = 0
        sign = 1
        s = expr.replace(" ", "")
        while i < len(s):
            ch = s[i]
            if ch in "+-":
                sign = 1 if ch == "+" else -1
                i += 1
                if i < len(s) and s[i] == "-":   # unary minus after a binary operator
                    sign = -sign
                    i += 1
                continue
            m = A_TERM.match(expr.replace(" ", ""), i)
            if m:
                head_neg, ua, paxis, pexpr, fax = m.groups()
                if paxis:
                    pterms = [AX[paxis]]
                else:
                    pterms = [(-1 if sg == "-" else 1) * AX[a] for sg, a in re.findall(r"([+-]?)u([xyz])", pexpr)]
                assert ua == fax, (d, ua, fax)
                toks.append({"kind": "A", "sign": sign, "axis": AX[fax], "head": -1 if head_neg else 1, "P": pterms})
                i = m.end()
                sign = 1
                continue
            m = U_TERM.match(s, i)
            if m:
                assert m.group(1) == m.group(2)
                toks.append({"kind": "U", "sign": sign, "axis": AX[m.group(1)]})
                i = m.end()
                sign = 1
                continue
            raise ValueError(f"fpop{d}: cannot parse at {s[i:i+30]!r}")
        cls = coe.lstrip("-")[-1]
        out[d] = {"coe": CLS[cls], "neg": coe.startswith("-"), "terms": toks}
    return out


def main():
    if not os.path.isdir(REF):
        sys.exit("reference sources not present")
    h = read("LBM.h")
    cu = read("LBM.cu")
    st = stream_vectors(cu)
    bottom, top, hcopy = boundary(cu)
    inj, rest_inj = bc_charge(cu)
    eqf_c, eqh_c = equilibria(cu, "__global__ void gpu_collide_save(double *f0", "// calculate force population",
                              r"double fe(\d+) =", r"double he(\d+) =")
    eqf_i, eqh_i = equilibria(cu, "__global__ void gpu_init_equilibrium(double *f0",
                              "__host__ void stream_collide_save",
                              r"f1\[gpu_fieldn_index\(x,\s*y,\s*z,\s*(\d+)\)\]",
                              r"h1\[gpu_fieldn_index\(x,\s*y,\s*z,\s*(\d+)\)\]")
    tables = {
        "source": "LBM.h, LBM.cu (parsed as text)",
        "weights": weights(h),
        "stream_f": st["f1"], "stream_h": st["h1"],
        "bounce_bottom": bottom, "bounce_top": top, "h_top_copy": hcopy,
        "inject": inj, "inject_rest_lines": rest_inj,
        "moments": moments(cu),
        "eq_collide_f": eqf_c, "eq_collide_h": eqh_c,
        "eq_init_f": eqf_i, "eq_init_h": eqh_i,
        "fpop": fpops(cu),
    }
    for key in ("stream_f", "stream_h", "bounce_bottom", "bounce_top", "h_top_copy", "inject",
                "eq_collide_f", "eq_collide_h", "eq_init_f", "eq_init_h"):
        assert len(tables[key]) == 26, (key, len(tables[key]))
    assert len(tables["fpop"]) == 27
    with open(OUT, "w") as f:
        json.dump(tables, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
