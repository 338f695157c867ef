"""Writes the struct-layout fixture of the reference's SW2 path (ref_struct_layout.txt) from its own headers.

Run in a container that has the reference tree (/root/reference); the GPU box and the tests only read the
committed fixture.  Steps:
 1. the field names of the structs that cross the drop-in boundary are read, in declaration order, from the
    reference's headers, in the single-rank configuration (no _MESSG, _PETSC, _HDF5, _ADH_GROUNDWATER) --
    into ref_layout_fields.h, an X-macro list of (struct, field) names;
 2. oracle/ref_layout.c, compiled against the reference headers in adh.h's order by `make -C oracle
    _ref/ref_layout`, prints sizeof of every struct and the offset and size of every listed field;
 3. its output becomes the fixture.
tests/test_ref_layout.py compiles the same program against include/adhk_ref.h and compares.
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("ADHK_REFERENCE", "/root/reference")
SKIP = {"_MESSG", "_PETSC", "_HDF5", "_ADH_GROUNDWATER"}

# struct -> header file (under src/) that declares it
STRUCTS = [
    ("SDT", "structs/sdt/sdt.h"), ("SVECT2D", "structs/svect/svect2d.h"), ("SVECT", "structs/svect/svect.h"),
    ("SNODE", "structs/snode/snode.h"), ("FLUX_ELEM", "structs/selem/selem_1d.h"),
    ("SELEM_1D", "structs/selem/selem_1d.h"), ("SELEM_2D", "structs/selem/selem_2d.h"),
    ("SFLAGS", "structs/sflags/sflags.h"), ("SIVAR_POSITION", "structs/sivar_position/sivar_position.h"),
    ("SDVAR_POSITION", "structs/sdvar/sdvar.h"), ("SDVAR", "structs/sdvar/sdvar.h"),
    ("SPDE", "structs/smat/smat_physics.h"), ("SMAT_PHYSICS", "structs/smat/smat_physics.h"),
    ("SSERIES_ENTRY", "structs/sseries/sseries.h"), ("SSERIES", "structs/sseries/sseries.h"),
    ("SLIN_SYS", "structs/slin_sys/slin_sys.h"), ("NODAL_ATTRIBUTES", "structs/sgrid/sgrid.h"),
    ("SGRID", "structs/sgrid/sgrid.h"), ("SMODEL_SUPER", "structs/smodel_super/smodel_super.h"),
    ("INPUT_DATA", "structs/sstr_value/sstr_value.h"), ("FRICTION", "structs/sstr_value/sstr_value.h"),
    ("STR_VALUE", "structs/sstr_value/sstr_value.h"), ("SMODEL_COUPLE", "structs/smodel_design/smodel_design.h"),
    ("SMODEL_DESIGN", "structs/smodel_design/smodel_design.h"),
]


def preprocess(text):
    """drop comments and the #ifdef blocks of the configurations not built"""
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    out, stack = [], []
    for line in text.split("\n"):
        s = line.strip()
        m = re.match(r"#\s*(ifdef|ifndef|if)\b\s*(\w*)", s)
        if m:
            skip = m.group(1) == "ifdef" and m.group(2) in SKIP
            stack.append(skip)
            continue
        if re.match(r"#\s*else\b", s):
            stack[-1] = not stack[-1] if stack else False
            continue
        if re.match(r"#\s*endif\b", s):
            stack.pop()
            continue
        if s.startswith("#"):
            continue
        if not any(stack):
            out.append(line)
    return "\n".join(out)


def fields_of(name, text):
    m = re.search(r"typedef\s+struct\s*\w*\s*\{([^{}]*)\}\s*" + name + r"\s*;", text)
    if not m:
        raise SystemExit(f"{name} not found")
    names = []
    for decl in m.group(1).split(";"):
        d = " ".join(decl.split())
        if not d:
            continue
        fp = re.search(r"\(\s*\*\s*(\w+)\s*\)", d)
        if fp:
            names.append(fp.group(1))
            continue
        parts = d.split(",")
        first = re.sub(r"\[.*?\]", "", parts[0]).strip()
        names.append(re.findall(r"\w+", first)[-1])
        for p in parts[1:]:
            names.append(re.findall(r"\w+", re.sub(r"\[.*?\]", "", p))[-1])
    return names


def main():
    src = os.path.join(REF, "src")
    if not os.path.isdir(src):
        sys.exit(f"{src} not found: the fixture is generated where the reference tree is present")
    lines = ["/* generated by tests/golden/make_ref_layout.py from the reference's headers: (struct, field) in",
             " * declaration order, single-rank configuration */"]
    for name, hdr in STRUCTS:
        text = preprocess(open(os.path.join(src, hdr)).read())
        lines.append(f"STRUCT({name})")
        lines += [f"FIELD({name}, {f})" for f in fields_of(name, text)]
    open(os.path.join(HERE, "ref_layout_fields.h"), "w").write("\n".join(lines) + "\n")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_ref/ref_layout"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "ref_layout")], check=True, capture_output=True,
                         text=True).stdout
    open(os.path.join(HERE, "ref_struct_layout.txt"), "w").write(out)
    print(f"{out.count(chr(10))} lines written")


if __name__ == "__main__":
    main()
