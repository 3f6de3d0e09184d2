"""Writes tests/golden/rafttypes_symbols.json: every function of libraft.so
that the reference's ctypes loader binds (sscRaft/rafttypes.py), with the
argtypes / restype it sets, as ctypes type names.

Reads the reference's rafttypes.py as text (it is not imported: it needs
libcudart.so, h5py and skimage, none of which this image has).  Run from the
repository root:  python tests/golden/list_rafttypes_symbols.py [path]
"""
import json
import os
import re
import sys

DEFAULT = "/root/reference/sscRaft/rafttypes.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rafttypes_symbols.json")


def parse(text):
    """{name: {"argtypes": [...], "restype": str}} from `libraft.<name>.argtypes
    = [...]` and `libraft.<name>.restype = ...` statements."""
    out = {}
    for m in re.finditer(r"libraft\.(\w+)\.argtypes\s*=\s*\[(.*?)\]", text, flags=re.S):
        body = re.sub(r"#[^\n]*", "", m.group(2))
        args = [a.strip().replace("ctypes.", "") for a in body.split(",") if a.strip()]
        out.setdefault(m.group(1), {})["argtypes"] = args
    for m in re.finditer(r"libraft\.(\w+)\.restype\s*=\s*([\w.]+)", text):
        if m.group(1) in out:
            out[m.group(1)]["restype"] = m.group(2).replace("ctypes.", "")
    for v in out.values():
        v.setdefault("restype", "None")
    return out


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else DEFAULT
    with open(path) as fh:
        syms = parse(fh.read())
    with open(OUT, "w") as fh:
        json.dump(syms, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print(f"{len(syms)} symbols -> {OUT}")


if __name__ == "__main__":
    main()
