"""Generates coupling_tables.inc from the coupling-table fixture.

The SHE coupling coefficients a_{x,y,z}, b_{x,y,z} for L = 1, 3, 5, 7 are the
reference's precomputed constants (viennashe/she/harmonics_coupling.hpp:281-1046).
Their decimal values fix the assembled matrix to the last digit, so the fill
uses the same numbers; tests/golden/coupling_tables.json holds them as data.
Run: python gen_coupling_tables.py  (rewrites coupling_tables.inc next to this file)
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "..", "tests", "golden", "coupling_tables.json")


def main():
    tab = json.load(open(SRC))
    out = ["// Generated by gen_coupling_tables.py from tests/golden/coupling_tables.json.",
           "// Entries {row, col, value} of a_x, a_y, a_z, b_x, b_y, b_z per expansion order.",
           "#pragma once", "namespace she_tables {",
           "struct Entry { short r, c; double v; };"]
    idx = []
    for L in ("1", "3", "5", "7"):
        for name in ("a_x", "a_y", "a_z", "b_x", "b_y", "b_z"):
            ent = tab[L][name]
            var = f"L{L}_{name}"
            body = ", ".join("{%d, %d, %r}" % (r, c, v) for r, c, v in ent)
            out.append(f"static const Entry {var}[] = {{{body}}};")
            idx.append((L, name, var, len(ent)))
    out.append("struct Table { int L; int which; const Entry *e; int n; };")
    out.append("static const Table tables[] = {")
    for L, name, var, n in idx:
        which = ("a_x", "a_y", "a_z", "b_x", "b_y", "b_z").index(name)
        out.append(f"  {{{L}, {which}, {var}, {n}}},")
    out.append("};")
    out.append("}  // namespace she_tables")
    open(os.path.join(HERE, "coupling_tables.inc"), "w").write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
