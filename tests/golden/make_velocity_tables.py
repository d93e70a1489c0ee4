"""Writes velocity_sets.json: the velocity tables of the reference, read as text.

Source: src/velocities.rs:66-191 (weights ws and velocities qs of D1Q3, D2Q9, D2Q5, D3Q7,
D3Q27) and the qs tables at the top of each update kernel in src/lib.cl (lib.cl:103,
192-193, 304-305, 410-417, 623).  Nothing is imported or executed from the reference: the
files are parsed as text.  Run from the repository root:

    python tests/golden/make_velocity_tables.py /root/reference
"""

import json
import os
import re
import sys
from fractions import Fraction


def parse_velocities_rs(text: str) -> dict:
    out = {}
    for m in re.finditer(r"pub fn (D\dQ\d+)\(\) -> VelocitySet \{(.*?)\n    \}\n", text, re.S):
        name, body = m.group(1), m.group(2)
        ws_txt = re.search(r"ws: arr1\(&\[(.*?)\]\)", body, re.S).group(1)
        qs_txt = re.search(r"qs: arr2\(&\[(.*?)\]\),", body, re.S).group(1)
        ws = []
        for tok in ws_txt.replace("\n", " ").split(","):
            tok = tok.strip()
            if not tok:
                continue
            a, b = tok.split("/")
            ws.append(str(Fraction(int(float(a.strip().rstrip("."))),
                                   int(float(b.strip().rstrip("."))))))
        qs = [[int(x) for x in row.split(",") if x.strip()]
              for row in re.findall(r"\[([-\d,\s]+)\]", qs_txt)]
        out[name] = {"w": ws, "c": qs}
    return out


def parse_lib_cl(text: str) -> dict:
    out = {}
    for m in re.finditer(r"kernel void (update_d(\d)q(\d+)_(bgk|trt))\((.*?)const int qs\[(\d+)\]\[(\d)\] = \{(.*?)\};", text, re.S):
        name = m.group(1)
        rows = re.findall(r"\{([-\d,\s]+)\}", m.group(8))
        out[name] = [[int(x) for x in r.split(",") if x.strip()] for r in rows]
    return out


def main(ref_root: str) -> None:
    with open(os.path.join(ref_root, "src", "velocities.rs")) as fh:
        sets = parse_velocities_rs(fh.read())
    with open(os.path.join(ref_root, "src", "lib.cl")) as fh:
        kernels = parse_lib_cl(fh.read())
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "velocity_sets.json"), "w") as fh:
        json.dump({"velocities_rs": sets, "lib_cl_qs": kernels}, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
