"""Build step: compiles libadhk.so for gfx950 (hipcc + gcc) in-tree, and the CPU oracle.

The oracle (oracle/liboracle.so) is test infrastructure; building it here only
builds the checker, the product library never links it.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _make(path, *args):
    jobs = str(min(16, os.cpu_count() or 4))
    r = subprocess.run(["make", "-j" + jobs, "-C", path, *args], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed in {path}:\n{r.stdout[-4000:]}\n{r.stderr[-8000:]}")


def build(oracle=True, force=True):
    """Compile libadhk.so (always, when force) and the oracle library."""
    csrc = os.path.join(HERE, "csrc")
    if force:
        _make(csrc, "clean")
    _make(csrc)
    if oracle:
        o = os.path.join(ROOT, "oracle")
        if force:
            _make(o, "clean")
        _make(o)
        # the struct-layout table from the reference's own headers, where the reference tree is present
        # (the checker of include/adhk_ref.h; tests compare against the committed fixture either way)
        if os.path.isdir("/root/reference/src"):
            _make(o, "_ref/ref_layout")
    return os.path.join(HERE, "libadhk.so")


if __name__ == "__main__":
    print(build())
