"""Write tests/golden/ref_bond_tables.json: the reference's literal cluster tables.

Reads S_one_half/Hybrid_TPS_params.f90 as TEXT (nothing is compiled or run) and
collects, for the hand-built clusters N = 8, 12, 18, 24, the ``simple==0`` branch's
``ipair`` and ``ibond_type`` array constructors.  The output is data (integers), the
fixture against which pned_amd.lattice is checked.  Run from the repo root:

    python tests/golden/extract_reference_tables.py /root/reference
"""
import json
import os
import re
import sys


def _array_after(text, name, start):
    """Integers of the first ``name= (/ ... /)`` constructor at or after ``start``."""
    i = text.index(name, start)
    j = text.index("(/", i)
    k = text.index("/)", j)
    body = text[j + 2:k]
    # drop Fortran comments and continuation marks
    lines = [ln.split("!")[0] for ln in body.splitlines()]
    nums = re.findall(r"-?\d+", " ".join(lines).replace("&", " "))
    return [int(x) for x in nums], k


def main(ref_root):
    path = os.path.join(ref_root, "S_one_half", "Hybrid_TPS_params.f90")
    text = open(path).read()
    out = {}
    for n in (24, 8, 12, 18):
        m = re.search(r"n\s*==\s*%d\s*\)\s*then" % n, text)
        start = m.end()
        s0 = text.index("if(simple==0)then", start)
        ip, k = _array_after(text, "ipair=", s0)
        bt, _ = _array_after(text, "ibond_type=", k)
        out[str(n)] = {"ipair": ip, "ibond_type": bt}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_bond_tables.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", dst, {k: len(v["ibond_type"]) for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
