"""Extract the reference's evaluation constants, Zobrist keys and castle masks
as numbers (no source text) into tests/golden/ref_tables.json.

Sources (read as text from the reference checkout):
  src/zeta.cl:272      EvalPieceValues[7]
  src/zeta.cl:275-285  EvalControl[64]
  src/zeta.cl:288-359  EvalTable[7*64]
  src/zeta.cl:258-265  Zobrist[18] (device copy)
  src/zobrist.h:28-35  Zobrist[18] (host copy)
  src/types.h:114-124  SMCRALL, SMCRWHITEQ/K, SMCRBLACKQ/K

The fixture is data: integer arrays.  tests/test_tables_pinned.py compares it
entry by entry with the oracle's copy (oracle/refcpu.c) and the device code's
copy (zeta_amd/csrc/evaltab.h, chess.h, via the host build in tests/hostcore).

Usage: python tests/golden/extract_tables.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def array_numbers(text, decl):
    """Numbers inside the brace initialiser of the first live declaration matching decl."""
    m = re.search(decl + r"\s*=\s*\{(.*?)\}", text, flags=re.S)
    if not m:
        raise SystemExit("declaration not found: %s" % decl)
    body = m.group(1)
    return [int(tok, 0) for tok in re.findall(r"-?0[xX][0-9A-Fa-f]+|-?\d+", body)]


def define_value(text, name):
    m = re.search(r"#define\s+%s\s+(0[xX][0-9A-Fa-f]+)ULL" % name, text)
    if not m:
        raise SystemExit("define not found: %s" % name)
    return int(m.group(1), 16)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, "src", "zeta.cl")) as f:
        cl = strip_comments(f.read())
    with open(os.path.join(ref, "src", "zobrist.h")) as f:
        zh = strip_comments(f.read())
    with open(os.path.join(ref, "src", "types.h")) as f:
        ty = f.read()
    values = array_numbers(cl, r"__constant\s+Score\s+EvalPieceValues\[7\]")
    control = array_numbers(cl, r"__constant\s+Score\s+EvalControl\[64\]")
    pst = array_numbers(cl, r"__constant\s+Score\s+EvalTable\[7\*64\]")
    zob_cl = array_numbers(cl, r"__constant\s+Hash\s+Zobrist\[18\]")
    zob_h = array_numbers(zh, r"const\s+Hash\s+Zobrist\[18\]")
    assert len(values) == 7 and len(control) == 64 and len(pst) == 7 * 64, (len(values), len(control), len(pst))
    assert len(zob_cl) == 18 and zob_cl == zob_h, "device and host Zobrist keys differ"
    masks = {k: define_value(ty, k) for k in ("SMCRALL", "SMCRWHITEQ", "SMCRWHITEK", "SMCRBLACKQ", "SMCRBLACKK")}
    out = {
        "source": "shintakezou/Zeta 099m: src/zeta.cl:258-359, src/zobrist.h:28-35, src/types.h:114-124",
        "piece_values": values,
        "control": control,
        "pst": pst,
        "zobrist": ["0x%016X" % z for z in zob_cl],
        "castle_masks": {k: "0x%016X" % v for k, v in masks.items()},
    }
    path = os.path.join(HERE, "ref_tables.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print("wrote %s: %d values, %d control, %d pst, %d keys" % (path, len(values), len(control), len(pst), len(zob_cl)))


if __name__ == "__main__":
    main()
