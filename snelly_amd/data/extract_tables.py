"""Write spectral_tables.json from the reference's data tables.

The reference keeps its spectral data as literals in JavaScript:
  * js/color.js:24-283  wavelengthToXYZTable(): 1024 x (X, Y, Z, 0) CIE-2006 2-degree
    colour-matching values at 390 + i*360/1023 nm (made by scripts/color/cie.py);
  * js/materials.js:536-627 tabulated_<metal>(): 64 samples of n and k over 390-750 nm
    for 22 metals, registered under display names at js/materials.js:1150-1171.
This script parses those literals (data only, no code) so the runtime and the JS
host can ship them as JSON.  It reads /root/reference and is run by hand:
  python3 snelly_amd/data/extract_tables.py
"""
import json
import os
import re

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def floats(s):
    return [float(x) for x in re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?", s)]


def main():
    color = open(os.path.join(REF, "js/color.js")).read()
    body = color[color.index("function wavelengthToXYZTable"):]
    body = body[body.index("Float32Array([") + len("Float32Array(["):body.index("]);")]
    xyz = floats(body)
    assert len(xyz) == 4096, len(xyz)
    mats = open(os.path.join(REF, "js/materials.js")).read()
    tables = {}
    for m in re.finditer(r"function tabulated_(\w+)\(\)\s*\{.*?n: new Float32Array\(\[(.*?)\]\),\s*k: new Float32Array\(\[(.*?)\]\)", mats, re.S):
        n, k = floats(m.group(2)), floats(m.group(3))
        assert len(n) == 64 and len(k) == 64, m.group(1)
        tables[m.group(1)] = {"n": n, "k": k}
    names = {}
    for m in re.finditer(r'new TabulatedMetal\("([^"]+)",\s*tabulated_(\w+)\(\)', mats):
        names[m.group(1)] = m.group(2)
    metals = {disp: tables[fn] for disp, fn in names.items()}
    assert len(metals) == 22, len(metals)
    out = {"xyz_lut": [xyz[4 * i:4 * i + 3] for i in range(1024)],
           "xyz_lut_nm": [390.0, 750.0],
           "metals": metals,
           "metals_nm": [390.0, 750.0]}
    with open(os.path.join(HERE, "spectral_tables.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", os.path.join(HERE, "spectral_tables.json"))


if __name__ == "__main__":
    main()
