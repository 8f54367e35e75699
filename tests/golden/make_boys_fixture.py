"""Write tests/golden/boys_golden.json: the Boys function values F_0..F_4 at
x = 1.0 that the reference's own test holds as golden data
(crates/cintx-cubecl/tests/boys_tests.rs, boys_host_generic_f64_byte_identity:
values of libcint's gamma_inc_like power-series branch, fmt.c:186-226), as
numbers only.

Run from the repository root in a container that has the reference at
/root/reference:  python tests/golden/make_boys_fixture.py
"""
import json
import os
import re

SRC = "/root/reference/crates/cintx-cubecl/tests/boys_tests.rs"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "boys_golden.json")


def main():
    text = open(SRC).read()
    m = re.search(r"let golden: \[f64; (\d+)\] = \[(.*?)\];", text, re.S)
    n = int(m.group(1))
    vals = re.findall(r"(-?\d+\.\d+(?:e-?\d+)?),\s*//\s*f\[(\d+)\]", m.group(2))
    assert len(vals) == n and [int(k) for _, k in vals] == list(range(n)), vals
    # the x and the order of the call that produced them
    call = re.search(r"boys_gamma_inc_host::<f64>\(([\d.]+)_f64,\s*(\d+)\)", text)
    x, mmax = float(call.group(1)), int(call.group(2))
    assert mmax + 1 == n
    with open(OUT, "w") as f:
        json.dump({"source": "crates/cintx-cubecl/tests/boys_tests.rs "
                             "(boys_host_generic_f64_byte_identity)",
                   "x": x, "F": [v for v, _ in vals], "tolerance": 1e-14}, f, indent=1)
    print("wrote", OUT, x, len(vals), "values")


if __name__ == "__main__":
    main()
