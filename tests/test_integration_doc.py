"""The Julia binding documented in INTEGRATION.md against the C header.

A Julia caller passes descriptor structs (and arrays of them) by pointer, so
every documented Julia struct must list the C struct's fields in the same
order with the same widths; otherwise the library reads the wrong stride.
This test parses both texts and also checks the C layout (size, field
offsets) against the ctypes mirrors the Python harness uses.
"""

import ctypes
import os
import re

import pytest

from gricoamd import _devabi, learned

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = {
    "BucketDesc": ("gmf_bucket_desc", _devabi.BucketDesc),
    "PlanDesc": ("gmf_plan_desc", _devabi.PlanDesc),
    "SolveInfo": ("gmf_solve_info", _devabi.SolveInfo),
    "TransferDesc": ("gmf_transfer_desc", _devabi.TransferDesc),
    "MgDesc": ("gmf_mg_desc", _devabi.MgDesc),
    "MlpDesc": ("gmf_mlp_desc", learned.MlpDesc),
    "BucketTuning": ("gmf_bucket_tuning", _devabi.BucketTuning),
    "BlockDesc": ("gmf_block_desc", _devabi.BlockDesc),
}

C_SCALARS = {"int32_t": "Int32", "int64_t": "Int64", "double": "Float64", "float": "Float32"}
C_STRUCTS = {"gmf_bucket_desc": "BucketDesc", "gmf_transfer_desc": "TransferDesc", "gmf_block_desc": "BlockDesc"}
SIZES = {"Int32": 4, "Int64": 8, "Float64": 8, "Float32": 4}


def _strip_c_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def c_structs(text):
    """{c_name: [(field, julia_type)]} from `typedef struct name { ... } name;`."""
    out = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} \1;", _strip_c_comments(text), flags=re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            decl = " ".join(re.sub(r"\bconst\b", " ", decl).split())
            if "*" in decl:
                base, rest = decl.split("*", 1)
                base = base.strip()
                stars = 1 + rest.count("*")
                name = rest.replace("*", "").strip()
                if base in C_SCALARS:
                    inner = C_SCALARS[base]
                elif base in C_STRUCTS:
                    inner = C_STRUCTS[base]
                else:  # opaque handle
                    inner = "Cvoid"
                jt = inner
                for _ in range(stars):
                    jt = f"Ptr{{{jt}}}"
                fields.append((name, jt))
                continue
            base, names = decl.split(" ", 1)
            for name in names.split(","):
                name = name.strip()
                arr = re.match(r"(\w+)\[(\d+)\]", name)
                if arr:
                    fields.append((arr.group(1), f"NTuple{{{arr.group(2)},{C_SCALARS[base]}}}"))
                else:
                    fields.append((name, C_SCALARS[base]))
        out[m.group(1)] = fields
    return out


def julia_structs(text):
    """{name: [(field, type)]} from the Julia code block of INTEGRATION.md."""
    block = text.split("```julia", 1)[1].split("```", 1)[0]
    block = re.sub(r"#[^\n]*", "", block)
    out = {}
    for m in re.finditer(r"^struct (\w+)\n(.*?)^end", block, flags=re.S | re.M):
        fields = []
        for part in re.split(r"[;\n]", m.group(2)):
            part = part.strip()
            if not part:
                continue
            name, jt = part.split("::")
            fields.append((name.strip(), jt.replace(" ", "")))
        out[m.group(1)] = fields
    return out


def julia_layout(fields):
    """(size, offsets) of a Julia isbits struct, C alignment rules."""
    off, align, offsets = 0, 1, []
    for _, jt in fields:
        tup = re.match(r"NTuple\{(\d+),(\w+)\}", jt)
        if tup:
            a = SIZES[tup.group(2)]
            size = a * int(tup.group(1))
        elif jt.startswith("Ptr{"):
            a = size = 8
        else:
            a = size = SIZES[jt]
        off = (off + a - 1) // a * a
        offsets.append(off)
        off += size
        align = max(align, a)
    return (off + align - 1) // align * align, offsets


@pytest.fixture(scope="module")
def texts():
    with open(os.path.join(ROOT, "include", "grico_mf.h")) as f:
        header = f.read()
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        doc = f.read()
    return c_structs(header), julia_structs(doc)


@pytest.mark.parametrize("jname", sorted(PAIRS))
def test_julia_struct_matches_header(texts, jname):
    cs, js = texts
    cname, _ = PAIRS[jname]
    assert jname in js, f"INTEGRATION.md documents no Julia struct {jname}"
    assert js[jname] == cs[cname], f"{jname} differs from {cname}"


@pytest.mark.parametrize("jname", sorted(PAIRS))
def test_julia_layout_matches_ctypes(texts, jname):
    _, js = texts
    _, ctype = PAIRS[jname]
    size, offsets = julia_layout(js[jname])
    assert size == ctypes.sizeof(ctype)
    assert [getattr(ctype, f[0]).offset for f in ctype._fields_] == offsets
    assert [f[0] for f in ctype._fields_] == [f[0] for f in js[jname]]


def test_parser_sees_every_descriptor(texts):
    cs, _ = texts
    assert {c for c, _ in PAIRS.values()} <= set(cs)
    # a planted omission is caught: drop the last field of BucketDesc
    _, js = texts
    assert js["BucketDesc"][:-1] != cs["gmf_bucket_desc"]
