"""Writes the reference's precomputed SHE coupling tables as a data fixture.

The reference ships the coupling coefficients a_{x,y,z} and b_{x,y,z} for
expansion orders L = 1, 3, 5, 7 as literal constants in
viennashe/she/harmonics_coupling.hpp:281-1046 (``fill_coupling_matrices_1`` ...
``fill_coupling_matrices_7``).  They are golden vectors of the hot path: every
streaming entry of the assembled system is a multiple of one of them.

This script reads that file as text (it runs only where the reference tree is
mounted) and writes ``coupling_tables.json`` next to itself:
``{"1": {"a_x": [[row, col, value], ...], ...}, "3": {...}, ...}``
with the entries in the order the reference assigns them.
"""
import json
import os
import re
import sys

SRC = "/root/reference/viennashe/she/harmonics_coupling.hpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "coupling_tables.json")


def main(src=SRC, out=OUT):
    text = open(src).read()
    starts = [(m.start(), int(m.group(1))) for m in
              re.finditer(r"void\s+fill_coupling_matrices_(\d)\s*\(", text)]
    tables = {}
    entry = re.compile(r"\b([ab]_[xyz])\(\s*(\d+)\s*,\s*(\d+)\s*\)\s*=\s*([-+0-9.eE]+)\s*;")
    for k, (pos, L) in enumerate(starts):
        end = starts[k + 1][0] if k + 1 < len(starts) else text.index("void fill_coupling_matrices(", pos)
        body = text[pos:end]
        t = {name: [] for name in ("a_x", "a_y", "a_z", "b_x", "b_y", "b_z")}
        for m in entry.finditer(body):
            t[m.group(1)].append([int(m.group(2)), int(m.group(3)), float(m.group(4))])
        tables[str(L)] = t
    with open(out, "w") as f:
        json.dump(tables, f, indent=0)
    return tables


if __name__ == "__main__":
    t = main(*sys.argv[1:])
    for L, d in t.items():
        print(L, {k: len(v) for k, v in d.items()})
