"""Extract the reference's own golden DATA into compact fixtures.

Reads (as text, never compiles or imports) the row tables of
  tests/random/golden_random.h          -> random_core_rows.json, random_key_rows.json
  tests/bandmath/golden_{sin,cos,exp,log,sqrt,tan,tanh,atan,asin,acos,rsqrt,pow}.h -> bandmath_<fn>.npz
of the reference source tree and writes them next to this script. Only the
numbers are kept (inputs and expected outputs); no source text is copied.
Run: python tests/golden/extract_reference_fixtures.py [/path/to/reference]
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _int_literal(s):
    """C integer literal -> int (drops a u/l suffix; hex digits are kept)."""
    return int(re.sub(r"(?i)(ull|llu|ul|lu|ll|u|l)$", "", s), 0)


def extract_random(ref):
    path = os.path.join(ref, "tests", "random", "golden_random.h")
    text = open(path).read()
    core_block = text.split("kCoreRows[] = {", 1)[1].split("};", 1)[0]
    key_block = text.split("kKeyRows[] = {", 1)[1].split("};", 1)[0]
    num = r"([-+0-9A-Za-z.]+)"
    core = []
    for m in re.finditer(r"\{\s*" + r"\s*,\s*".join([num] * 8) + r"\s*\}", core_block):
        g = [_int_literal(s) for s in m.groups()[:6]]
        core.append({
            "seed": g[0], "global": g[1], "counter": g[2],
            "bits64": g[3], "u01d_bits": g[4], "u01f_bits": g[5],
            "normal_d": float(m.group(7)), "normal_f": float(m.group(8).rstrip("fF")),
        })
    keys = []
    for m in re.finditer(r"\{\s*" + r"\s*,\s*".join([num] * 5) + r"\s*\}", key_block):
        g = [_int_literal(s) for s in m.groups()]
        keys.append({"seed": g[0], "stream": g[1], "k": g[2], "key": g[3], "sub_key": g[4]})
    tol = {
        "normal_d": float(re.search(r"kNormalDoubleTol = ([0-9.e+-]+);", text).group(1)),
        "normal_f": float(re.search(r"kNormalFloatTol = ([0-9.e+-]+)f;", text).group(1)),
    }
    n_core = int(re.search(r"kCoreRowCount = (\d+);", text).group(1))
    n_key = int(re.search(r"kKeyRowCount = (\d+);", text).group(1))
    assert len(core) == n_core, (len(core), n_core)
    assert len(keys) == n_key, (len(keys), n_key)
    with open(os.path.join(HERE, "random_core_rows.json"), "w") as f:
        json.dump({"source": "tests/random/golden_random.h", "tolerance": tol, "rows": core}, f)
    with open(os.path.join(HERE, "random_key_rows.json"), "w") as f:
        json.dump({"source": "tests/random/golden_random.h", "rows": keys}, f)
    print("random:", len(core), "core rows,", len(keys), "key rows")


def extract_bandmath(ref, fn):
    path = os.path.join(ref, "tests", "bandmath", f"golden_{fn}.h")
    text = open(path).read()
    block = text.split("kRows[] = {", 1)[1].split("};", 1)[0]
    rows = re.findall(r"\{\s*\{\s*0x([0-9A-Fa-f]+)ULL\s*\}\s*,\s*0x([0-9A-Fa-f]+)ULL\s*,\s*(\d)\s*\}", block)
    x = np.array([int(a, 16) for a, _, _ in rows], np.uint64)
    y = np.array([int(b, 16) for _, b, _ in rows], np.uint64)
    lab = np.array([int(c) for _, _, c in rows], np.uint8)
    np.savez_compressed(os.path.join(HERE, f"bandmath_{fn}.npz"), x_bits=x, ref_bits=y, label=lab,
                        source=np.array(f"tests/bandmath/golden_{fn}.h"))
    print(f"bandmath {fn}: {len(rows)} rows")


def extract_bandmath2(ref, fn):
    """Two-argument rows ({ {x, y}, ref, label }): x_bits, y_bits, ref_bits."""
    path = os.path.join(ref, "tests", "bandmath", f"golden_{fn}.h")
    text = open(path).read()
    block = text.split("kRows[] = {", 1)[1].split("};", 1)[0]
    rows = re.findall(r"\{\s*\{\s*0x([0-9A-Fa-f]+)ULL\s*,\s*0x([0-9A-Fa-f]+)ULL\s*\}\s*,\s*0x([0-9A-Fa-f]+)ULL"
                      r"\s*,\s*(\d)\s*\}", block)
    x = np.array([int(r[0], 16) for r in rows], np.uint64)
    y = np.array([int(r[1], 16) for r in rows], np.uint64)
    z = np.array([int(r[2], 16) for r in rows], np.uint64)
    lab = np.array([int(r[3]) for r in rows], np.uint8)
    np.savez_compressed(os.path.join(HERE, f"bandmath_{fn}.npz"), x_bits=x, y_bits=y, ref_bits=z, label=lab,
                        source=np.array(f"tests/bandmath/golden_{fn}.h"))
    print(f"bandmath {fn}: {len(rows)} rows")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    extract_random(ref)
    for fn in ("sin", "cos", "exp", "log", "sqrt", "tan", "tanh", "atan", "asin", "acos", "rsqrt"):
        extract_bandmath(ref, fn)
    extract_bandmath2(ref, "pow")


if __name__ == "__main__":
    main()
