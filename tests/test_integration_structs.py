"""The Julia structs of INTEGRATION.md (the reference-side binding) against include/oal.h, field for field:
same names, same order, same types (C int32_t / int64_t / float / double / pointers / fixed arrays as
Julia Int32 / Int64 / Float32 / Float64 / Ptr / NTuple), so that Julia's isbits layout is the C one."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

C_TO_JL = {"int32_t": "Int32", "int64_t": "Int64", "float": "Float32", "double": "Float64", "uint8_t": "UInt8",
           "uint16_t": "UInt16"}


def c_structs() -> dict:
    text = open(os.path.join(ROOT, "include", "oal.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for body, name in re.findall(r"typedef struct\s*\w*\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"(const\s+)?(\w+)\s*(\*?)\s*(.*)$", decl)
            base, ptr, names = m.group(2), m.group(3), m.group(4)
            for nm in names.split(","):
                nm = nm.strip()
                ptr2 = ptr
                if nm.startswith("*"):
                    ptr2, nm = "*", nm[1:].strip()
                arr = re.match(r"(\w+)\[(\d+)\]", nm)
                t = C_TO_JL.get(base, base)
                if ptr2:
                    t = f"Ptr{{{t}}}"
                if arr:
                    nm, t = arr.group(1), f"NTuple{{{arr.group(2)},{t}}}"
                fields.append((nm, t))
        out[name] = fields
    return out


def julia_structs() -> dict:
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    code = "\n".join(re.findall(r"```julia\n(.*?)```", text, flags=re.S))
    out = {}
    for name, body in re.findall(r"^struct (\w+)\n(.*?)^end", code, flags=re.S | re.M):
        fields = []
        for part in re.split(r"[;\n]", body):
            part = part.strip()
            if part:
                nm, t = part.split("::")
                fields.append((nm.strip(), t.strip().replace(" ", "")))
        out[name] = fields
    return out


def test_every_julia_struct_matches_the_header():
    c, jl = c_structs(), julia_structs()
    assert set(jl) >= {"OalCollision", "OalDomainDesc", "OalLevelDesc", "OalPlanDesc", "OalLevelInfo", "OalTuning",
                       "OalSurfaceParams", "OalSurfaceForces"}
    for name, fields in jl.items():
        assert name in c, f"{name} is not a struct of oal.h"
        assert fields == c[name], f"{name}: Julia {fields} != C {c[name]}"


def test_every_header_struct_has_a_julia_twin():
    c, jl = c_structs(), julia_structs()
    assert sorted(c) == sorted(jl)


@pytest.mark.parametrize("bad", ["n_blocks::Int64", "tau::Float64"])
def test_the_comparison_catches_a_wrong_type(bad):
    c = c_structs()
    jl = julia_structs()
    name = "OalLevelDesc" if bad.startswith("n_blocks") else "OalDomainDesc"
    field = bad.split("::")[0]
    jl[name] = [(n, bad.split("::")[1] if n == field else t) for n, t in jl[name]]
    assert jl[name] != c[name]


def test_julia_constants_match_the_header():
    """the ABI version and the enum values the Julia shim passes are the header's"""
    h = open(os.path.join(ROOT, "include", "oal.h")).read()
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    code = "\n".join(re.findall(r"```julia\n(.*?)```", text, flags=re.S))
    m = re.search(r"const OAL_ABI_VERSION = Int32\((\d+)\)", code)
    assert m, "the Julia shim defines OAL_ABI_VERSION"
    assert int(m.group(1)) == int(re.search(r"#define OAL_ABI_VERSION (\d+)", h).group(1))
    # every plan descriptor the shim builds passes that constant
    assert re.findall(r"OalPlanDesc\(([^,]+),", code) and all(
        a.strip() == "OAL_ABI_VERSION" for a in re.findall(r"OalPlanDesc\(([^,]+),", code))
    for name in ("OAL_F32", "OAL_F64", "OAL_F16S", "OAL_BF16S", "OAL_STREAM_AB", "OAL_BB_HALFWAY", "OAL_STEP_FORCE",
                 "OAL_ROUND_STOCHASTIC", "OAL_ROUND_NEAREST"):
        cv = re.search(rf"\b{name}\s*=\s*(\d+)", h)
        assert cv, name
        names = re.search(rf"const ([\w, ]*\b{name}\b[\w, ]*) = ([^\n#]+)", code)
        assert names, name
        lhs = [x.strip() for x in names.group(1).split(",")]
        rhs = re.findall(r"\((\d+)\)", names.group(2))
        assert int(rhs[lhs.index(name)]) == int(cv.group(1)), name
