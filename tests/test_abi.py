"""The C-ABI library loads and exports every entry point include/ehd.h declares;
host-only entry points (parameter derivation) work without a GPU."""
import os
import re
import subprocess

import pytest

import ehd_amd as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("ehd.h", "ehd_io.h", "ehd_fno.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(ehd_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


def test_header_declares_the_binding_exports():
    from ehd_amd import fno
    assert declared_functions() == sorted(E.EXPORTS + fno.EXPORTS)


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", E.LIB_PATH], check=True, capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (ehd_[a-z_0-9]+)$", out, flags=re.M))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    L = E.lib()
    for f in declared_functions():
        assert hasattr(L, f)


def test_abi_version():
    assert E.lib().ehd_abi_version() == 2


def test_struct_layouts_match_header(tmp_path):
    """The ctypes structures have the C structures' sizes and field offsets
    (a C program compiled against include/ehd.h prints them)."""
    from ehd_amd import fno
    structs = {"ehd_params": E.Params, "ehd_derived": E.Derived, "ehd_diag": E.Diag, "ehd_steady": E.Steady,
               "ehd_profile": E.Profile, "ehd_tuning": E.Tuning, "ehd_fno_config": fno.Config}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ehd.h"', '#include "ehd_fno.h"',
             'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == __import__("ctypes").sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, f"{cname}.{fname}"


def test_params_default_match_reference_header():
    p = E.default_params()
    assert (p.nx, p.ny, p.nz) == (122, 122, 101)          # LBM.h:34-36
    assert abs(p.T - 996.0000637) < 1e-6 and abs(p.M - 10) < 1e-12 and abs(p.C - 10) < 1e-12
    assert abs(p.Fe - 4000) < 1e-9
    assert p.lambda_f == 1.0 / 12.0 and p.lambda_c == 1e-6   # LBM.h:79-80
    d = E.derive(p)
    assert abs(d.K - 2.5e-5) < 1e-18 and abs(d.nu / 0.0251004 - 1) < 1e-14
    assert abs(d.diffu - 6.25e-5) < 1e-18 and abs(d.charge0 - 10.0) < 1e-13
    assert d.dz == 0.01 and abs(d.dt - 1e-4) < 1e-19 and abs(d.lx - 1.22) < 1e-15


def test_derive_matches_oracle():
    from oracle import pyoracle as O
    for (nx, ny, nz, T, C) in ((64, 64, 32, 200.0, 10.0), (256, 256, 128, 400.0, 10.0), (48, 48, 32, 150.0, 5.0)):
        p = E.make_params(nx, ny, nz, T=T, C_=C)
        d = E.derive(p)
        o = O.make_params(nx, ny, nz, T=T, M=10.0, C_=C, Fe=4000.0, lam=0.25)
        for a, b in (("dx", "dx"), ("dz", "dz"), ("dt", "dt"), ("lx", "lx"), ("cs_square", "cs_square"),
                     ("K", "K"), ("nu", "nu"), ("diffu", "diffu"), ("charge0", "charge0")):
            assert getattr(d, a) == getattr(o, b), a
        T2, M2, C2, Fe2 = E.compute_parameters(d, p)
        assert abs(T2 / T - 1) < 1e-14 and abs(C2 / C - 1) < 1e-14 and abs(M2 / 10 - 1) < 1e-14
        assert abs((d.tau_plus - 0.5) * (d.tau_minus - 0.5) - 0.25) < 1e-12


@pytest.mark.parametrize("bad", [dict(nx=7), dict(ny=0), dict(nz=3), dict(T=-1.0), dict(precision=5),
                                 dict(init_mode=9), dict(lam=0.0)])
def test_invalid_params_rejected_on_host(bad):
    kw = dict(nx=8, ny=8, nz=8)
    nx, ny, nz = bad.pop("nx", 8), bad.pop("ny", 8), bad.pop("nz", 8)
    p = E.make_params(nx, ny, nz, **bad)
    with pytest.raises(E.EhdError) as ei:
        E.derive(p)
    assert ei.value.code == -1
    assert E.lib().ehd_last_error()
    del kw


@pytest.mark.parametrize("C", [5.0, 10.0, 20.0])
def test_hydrostatic_current_matches_oracle(C):
    """ehd_hydrostatic_current = the oracle's current() (LBM.cu:1646-1671) of
    the hydrostatic initial state, so Nu = 1 there by construction."""
    from oracle import pyoracle as O
    nx, ny, nz = 12, 10, 17
    ob = O.Box(O.make_params(nx, ny, nz, T=200.0, M=10.0, C_=C, Fe=4000.0, lam=0.25))
    ob.init_hydrostatic(1, 0.0, 0)
    ref = ob.current()
    got = E.hydrostatic_current(E.make_params(nx, ny, nz, T=200.0, C_=C, lam=0.25))
    assert ref > 0
    assert abs(got - ref) <= 1e-13 * ref, (got, ref)


def test_strided_gemm_rejects_bad_layouts_on_host():
    """ehd_fno_gemm_strided checks its arguments before any device work:
    strides that reach outside the host arrays, negative strides and empty
    shapes are refused with EHD_EINVAL and a message."""
    import ctypes as C
    import numpy as np
    from ehd_amd import fno as G

    L = G._lib()
    A = np.zeros(64, np.float32)
    B = np.zeros(64, np.float32)
    Cm = np.zeros(64, np.float32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    arr = lambda v: (C.c_longlong * len(v))(*v)
    # 4 x 4 x 4 fits 16-element operands with unit strides
    ok = (4, 1, 16)
    # A reaches element 4*5 + 3 = 23 > 15 of a 16-element array
    rc = L.ehd_fno_gemm_strided(4, 4, 4, 1, 0, fp(A), arr((5, 1, 0)), 16, fp(B), arr(ok), 64, fp(Cm),
                                arr((4, 1, 16, 0)), 64, 1.0, 0.0)
    assert rc == -1 and b"outside" in L.ehd_last_error()
    rc = L.ehd_fno_gemm_strided(4, 4, 4, 1, 0, fp(A), arr((4, -1, 0)), 64, fp(B), arr(ok), 64, fp(Cm),
                                arr((4, 1, 16, 0)), 64, 1.0, 0.0)
    assert rc == -1 and b"negative" in L.ehd_last_error()
    rc = L.ehd_fno_gemm_strided(0, 4, 4, 1, 0, fp(A), arr(ok), 64, fp(B), arr(ok), 64, fp(Cm),
                                arr((4, 1, 16, 0)), 64, 1.0, 0.0)
    assert rc == -1
    # split-K slices: 8 slices of k = 32 / 4 need C slices up to 7 * 16 + 15 = 127 > 63
    rc = L.ehd_fno_gemm_strided(4, 4, 32, 1, 4, fp(A), arr((0, 1, 0)), 64, fp(B), arr((1, 0, 0)), 64, fp(Cm),
                                arr((4, 1, 0, 16)), 64, 1.0, 0.0)
    assert rc == -1 and b"outside" in L.ehd_last_error()


def test_complex_gemm_rejects_bad_arguments_on_host():
    """ehd_fno_cgemm_strided checks its arguments before any device work: signs
    other than +-1, strides that reach outside the host arrays, negative
    strides and a missing real plane are refused with EHD_EINVAL."""
    import ctypes as C
    import numpy as np
    from ehd_amd import fno as G

    L = G._lib()
    P = [np.zeros(64, np.float32) for _ in range(6)]
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    arr = lambda v: (C.c_longlong * len(v))(*v)
    ok = (4, 1, 16)

    def call(sa=1, sb=1, a_str=ok, a_len=64, ar=True):
        return L.ehd_fno_cgemm_strided(4, 4, 4, 1, fp(P[0]) if ar else None, fp(P[1]), arr(a_str), a_len, fp(P[2]),
                                       fp(P[3]), arr(ok), 64, fp(P[4]), fp(P[5]), arr(ok), 64, sa, sb, 1.0, 0.0)

    assert call(sa=2) == -1 and b"complex GEMM" in L.ehd_last_error()
    assert call(sb=0) == -1
    assert call(a_str=(5, 1, 0), a_len=16) == -1 and b"outside" in L.ehd_last_error()
    assert call(a_str=(4, -1, 0)) == -1 and b"negative" in L.ehd_last_error()
    assert call(ar=False) == -1


def test_resident_dataset_entry_points_reject_bad_arguments_on_host():
    """ehd_fno_set_dataset and ehd_fno_train_step_indices refuse a missing
    network, an empty set and missing arrays before any device work."""
    import ctypes as C
    import numpy as np
    from ehd_amd import fno as G

    L = G._lib()
    a = np.zeros(8, np.float32)
    fp = a.ctypes.data_as(C.POINTER(C.c_float))
    assert L.ehd_fno_set_dataset(None, 1, fp, fp) == -1
    assert L.ehd_fno_set_dataset(None, 0, fp, fp) == -1
    idx = (C.c_int * 2)(0, 1)
    loss = C.c_double()
    assert L.ehd_fno_train_step_indices(None, 2, idx, 1e-3, C.byref(loss)) == -1
    assert L.ehd_last_error()
