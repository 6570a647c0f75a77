"""Host-side checks of the C ABI (no GPU needed): the library loads, exports
every function the header declares, the Python binding matches the header's
constants, and the host-only entry point agrees with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "diffmpm_amd.h")


def header_text():
    with open(HEADER) as f:
        return f.read()


def declared_functions():
    txt = re.sub(r"/\*.*?\*/", "", header_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|const char\*)\s+(dmpm_\w+)\s*\(", txt)))


def header_enums():
    txt = re.sub(r"/\*.*?\*/", "", header_text(), flags=re.S)
    out = {}
    for body in re.findall(r"enum\s*\{(.*?)\}", txt, flags=re.S):
        for name, val in re.findall(r"(DMPM_\w+)\s*=\s*(-?\d+)", body):
            out[name] = int(val)
    return out


def test_library_exports_every_declared_symbol():
    import diffmpm_amd._lib as L
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 35
    for n in names:
        assert hasattr(lib, n), n
        assert n in L.SIGNATURES or n == "dmpm_last_error", n


def test_python_constants_match_header():
    import diffmpm_amd._lib as L
    e = header_enums()
    rows = {"X": L.P_X, "V": L.P_V, "MASS": L.P_MASS, "VOLUME": L.P_VOLUME,
            "DENSITY": L.P_DENSITY, "SIZE": L.P_SIZE, "MOMENTUM": L.P_MOMENTUM,
            "ACCEL": L.P_ACCEL, "STRAIN": L.P_STRAIN, "STRESS": L.P_STRESS,
            "STRAIN_RATE": L.P_STRAIN_RATE, "DSTRAIN": L.P_DSTRAIN, "FEXT": L.P_FEXT,
            "TRACTION": L.P_TRACTION, "XI": L.P_XI, "DVOLSTRAIN": L.P_DVOLSTRAIN,
            "VOLSTRAIN_C": L.P_VOLSTRAIN_C, "YOUNGS": L.P_YOUNGS, "ROWS": L.P_ROWS}
    for k, v in rows.items():
        assert e["DMPM_P_" + k] == v, k
    nrows = {"LOC": L.N_LOC, "V": L.N_V, "A": L.N_A, "MASS": L.N_MASS, "P": L.N_P,
             "FINT": L.N_FINT, "FEXT": L.N_FEXT, "FDAMP": L.N_FDAMP, "ROWS": L.N_ROWS}
    for k, v in nrows.items():
        assert e["DMPM_N_" + k] == v, k
    assert e["DMPM_F32"] == L.DMPM_F32 and e["DMPM_F64"] == L.DMPM_F64
    assert e["DMPM_QUAD4"] == L.DMPM_QUAD4 and e["DMPM_LINEAR1D"] == L.DMPM_LINEAR1D
    for k, v in L.KC.items():
        assert e["DMPM_KC_" + k.upper()] == v, k
    assert e["DMPM_OPT_USF_SPLIT"] == L.DMPM_OPT_USF_SPLIT
    assert e["DMPM_OPT_STEP_KERNEL"] == L.DMPM_OPT_STEP_KERNEL
    assert e["DMPM_RUN_CLOCK"] == L.DMPM_RUN_CLOCK


def test_plan_desc_layout():
    import diffmpm_amd._lib as L
    assert ctypes.sizeof(L.PlanDesc) == 4 * 4 + 8 * 2 + 8 * 3 + 4 * 2 + 8 * 3


@pytest.mark.parametrize("h", [(1.0, 1.0), (0.1, 0.1), (2.0 ** -10, 2.0 ** -10), (0.3, 0.7)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_element_volume_matches_oracle(h, dtype):
    import diffmpm_amd._lib as L
    from oracle import mpm as om
    el = om.Elements("quad4", (2, 2), h, dtype=dtype)
    el.compute_volume()
    code = L.DMPM_F32 if dtype == np.float32 else L.DMPM_F64
    got = L.element_volume(L.DMPM_QUAD4, code, h[0], h[1])
    assert got == float(el.volume[0])


def test_abi_version_and_errors():
    import diffmpm_amd._lib as L
    assert L.lib().dmpm_abi_version() == 101
    with pytest.raises(ValueError, match="element"):
        L.element_volume(7, 0, 1.0, 1.0)


def header_declarations():
    """{name: [parameter type, ...]} of every function the header declares."""
    txt = re.sub(r"/\*.*?\*/", "", header_text(), flags=re.S)
    out = {}
    for name, params in re.findall(r"\b(?:int|const char\*)\s+(dmpm_\w+)\s*\(([^)]*)\)\s*;",
                                   txt):
        params = " ".join(params.split())
        if params in ("", "void"):
            out[name] = []
            continue
        # each parameter's type: the declaration without its trailing name
        out[name] = [re.sub(r"\s*\b\w+$", "", p.strip()) for p in params.split(",")]
    return out


def test_binding_signatures_match_header():
    """The ctypes argument list of every entry point has the header's arity,
    and each argument the header's kind: a pointer, a 32- or 64-bit integer,
    or a double (a mismatch would pass garbage through the ABI)."""
    import ctypes as C
    import diffmpm_amd._lib as L
    decls = header_declarations()
    assert set(L.SIGNATURES) <= set(decls)
    for name, argt in L.SIGNATURES.items():
        types = decls[name]
        assert len(types) == len(argt), (name, types, argt)
        for ty, ct in zip(types, argt):
            if "*" in ty:
                assert ct is C.c_void_p or issubclass(ct, C._Pointer), (name, ty, ct)
            elif ty.endswith("int32_t"):
                assert ct is C.c_int32, (name, ty, ct)
            elif ty.endswith("int64_t"):
                assert ct is C.c_int64, (name, ty, ct)
            elif ty == "double":
                assert ct is C.c_double, (name, ty, ct)
            else:
                raise AssertionError((name, ty))
