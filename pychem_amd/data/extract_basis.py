"""Regenerate ``basis_sets.json`` from the reference's basis library.

The reference keeps every basis set in one Python dict literal
(``Data/basis.py``: ``get[name][ELEMENT] = [[l, [exp, coef], ...], ...]``).
This script reads that literal as *data* with ``ast.literal_eval`` (nothing in
the file is executed) and writes the subset this package needs as JSON.  The
numbers are basis-set data published with the reference under the MIT licence
(``Documentation/LICENSE``); the layout is kept identical so that shell order,
split sp shells and coefficients match the reference exactly.

Run from the repository root:  ``python pychem_amd/data/extract_basis.py``
"""
import ast
import json
import os
import sys

SETS = ["STO3G", "631G", "631GS", "631GSS", "6311G", "6311GS"]
ELEMENTS = ["H", "HE", "LI", "BE", "B", "C", "N", "O", "F", "NE"]


def main(ref_root="/root/reference"):
    src = open(os.path.join(ref_root, "Data", "basis.py")).read()
    table = ast.literal_eval(src[src.index("{"):])
    out = {"_source": "dlc62/pychem Data/basis.py (MIT), subset",
           "_format": "set -> element -> [[l, [exponent, coefficient], ...], ...]"}
    for name in SETS:
        out[name] = {el: table[name][el] for el in ELEMENTS if el in table[name]}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_sets.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, separators=(",", ":"))
    print("wrote", dst)


if __name__ == "__main__":
    main(*sys.argv[1:])
