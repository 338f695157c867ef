"""include/adhk_ref.h lays out the reference's structures exactly as its headers do (no GPU).

tests/golden/ref_struct_layout.txt is the sizeof of every structure that crosses the drop-in boundary and the
offset and size of each of its fields, printed by oracle/ref_layout.c compiled against the reference's own
headers in adh.h's order (tests/golden/make_ref_layout.py).  The same program compiled against adhk_ref.h
must print the same table."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "ref_struct_layout.txt")


def table(binary):
    return subprocess.run([binary], check=True, capture_output=True, text=True).stdout


@pytest.fixture(scope="module")
def ours(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("layout") / "ref_layout_adhk")
    subprocess.run(["gcc", "-std=gnu99", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "tests", "golden"), "-o", out,
                    os.path.join(ROOT, "oracle", "ref_layout.c")], check=True)
    return table(out)


def test_struct_layout_matches_reference_headers(ours):
    want = open(FIXTURE).read()
    assert ours == want


def test_fixture_covers_the_path_structs():
    want = open(FIXTURE).read()
    structs = {l.split()[1] for l in want.splitlines() if l.startswith("S ")}
    assert {"SMODEL_SUPER", "SGRID", "SLIN_SYS", "SELEM_2D", "SELEM_1D", "SNODE", "SDT", "SMAT_PHYSICS", "SPDE",
            "SIVAR_POSITION", "SDVAR", "SSERIES", "SFLAGS"} <= structs
    nfields = sum(1 for l in want.splitlines() if l.startswith("F "))
    assert nfields > 350


def test_fixture_is_current_where_the_reference_was_built():
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_layout")
    if not os.path.exists(ref):
        pytest.skip("oracle/_ref/ref_layout is built only where the reference tree is present")
    assert table(ref) == open(FIXTURE).read()
