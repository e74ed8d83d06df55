"""Extract the reference's 10x10 small-matrix test fixture, in all six
formats, into ``small_matrix.json`` (data only).

The numbers live in the reference test helpers ``reset_small_container``
(core/tests/utils/Macros_{Coo,Csr,Dia,Ell,Hyb,Hdc}Matrix.hpp, sizes in
core/tests/utils/Macros_Definitions.hpp:29-43).  This script reads those
files as text, takes the body of the first ``reset_small_container`` in each
and records every ``<array>(<idx>[, <idx>]) = <number>`` assignment.  It is
run in the build container only (the reference is not present elsewhere);
its output is committed.
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

REF = Path("/root/reference/core/tests/utils")
OUT = Path(__file__).resolve().parent / "small_matrix.json"

ASSIGN = re.compile(
    r"c\.(?:(ell|coo|dia|csr)\(\)\.)?"
    r"(row_offsets|row_indices|column_indices|values|diagonal_offsets)"
    r"\(\s*(\d+)\s*(?:,\s*(\d+)\s*)?\)\s*=\s*(?:\(value_type\))?\s*(-?[0-9.]+)\s*;")


def reset_body(text: str) -> str:
    start = text.index("void reset_small_container(")
    brace = text.index("{", text.index("nullptr)", start))
    depth, i = 0, brace
    while True:
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return text[brace:i]
        i += 1


def extract(fmt: str) -> dict:
    body = reset_body((REF / f"Macros_{fmt}Matrix.hpp").read_text())
    out: dict = {}
    for part, arr, i, j, num in ASSIGN.findall(body):
        key = f"{part}.{arr}" if part else arr
        idx = [int(i)] + ([int(j)] if j else [])
        val = float(num) if arr == "values" else int(num)
        out.setdefault(key, []).append([idx, val])
    return out


def main() -> int:
    if not REF.exists():
        print("reference not present; the committed fixture is used as is")
        return 0
    defs = (REF / "Macros_Definitions.hpp").read_text()
    consts = {k: int(v) for k, v in re.findall(r"const size_t (\w+)\s*=\s*(\d+);", defs)}
    fixture = {"constants": consts,
               "formats": {f: extract(f) for f in ("Coo", "Csr", "Dia", "Ell", "Hyb", "Hdc")}}
    OUT.write_text(json.dumps(fixture, indent=1, sort_keys=True) + "\n")
    print(f"wrote {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
