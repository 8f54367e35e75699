"""Write tests/golden/c2spinor_coeffs.json: the Cartesian -> spinor coupling
coefficients (libcint's g_trans_cart2jR / g_trans_cart2jI tables, l = 0..4)
as the reference's own port of them holds them
(crates/cintx-cubecl/src/transform/c2spinor_coeffs.rs), as numbers only.

Run from the repository root in a container that has the reference at
/root/reference:  python tests/golden/make_c2spinor_fixture.py
"""
import json
import os
import re

SRC = "/root/reference/crates/cintx-cubecl/src/transform/c2spinor_coeffs.rs"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2spinor_coeffs.json")


def main():
    text = open(SRC).read()
    tables = {}
    pat = re.compile(r"pub const (CJ_(GT|LT)_L(\d)_([RI])): \[\[f64; (\d+)\]; (\d+)\] = (\[.*?\]);",
                     re.S)
    for m in pat.finditer(text):
        name, block, l, part, ncol, nrow, body = m.groups()
        nums = [float(v) for v in re.findall(r"-?\d+\.\d+(?:e-?\d+)?|-?\d+(?=[,\]])", body)]
        nrow, ncol = int(nrow), int(ncol)
        assert len(nums) == nrow * ncol, (name, len(nums), nrow, ncol)
        tables[f"{block}_{l}_{part}"] = [nums[r * ncol:(r + 1) * ncol] for r in range(nrow)]
    assert len(tables) == 20, sorted(tables)
    with open(OUT, "w") as f:
        json.dump({"source": "libcint g_trans_cart2jR/I via crates/cintx-cubecl/src/transform/"
                             "c2spinor_coeffs.rs; rows: spinor components m_j = -j..j, "
                             "columns: alpha then beta coefficients of the ncart(l) "
                             "Cartesian functions",
                   "tables": tables}, f, indent=0)
    print("wrote", OUT, len(tables), "tables")


if __name__ == "__main__":
    main()
