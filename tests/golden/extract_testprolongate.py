"""Writes testprolongate_max_error.json from the reference's TestProlongate
golden outputs (TestProlongate/test/<case>/testprolongate-gf*_max_error.*.tsv)
and the parameters of each case (<case>.par).  Data only: the per-case
operator parameters and, per grid function (gfXYZ: X, Y, Z = 1 for a
cell-centred x, y, z direction), the max_error scalar at iterations 0..2.

    python tests/golden/extract_testprolongate.py [/path/to/reference]
"""
import glob
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
TEST = os.path.join(REF, "TestProlongate", "test")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "testprolongate_max_error.json")


def par_values(path):
    vals = {}
    for line in open(path):
        m = re.match(r"\s*CarpetX::(\w+)\s*=\s*(.+?)\s*$", line)
        if m:
            vals[m.group(1)] = m.group(2).strip('"')
        m = re.match(r"\s*\$(\w+)\s*=\s*(.+?)\s*$", line)
        if m:
            vals["$" + m.group(1)] = m.group(2)
    return vals


def main():
    cases = {}
    for par in sorted(glob.glob(os.path.join(TEST, "*.par"))):
        name = os.path.basename(par)[:-4]
        d = os.path.join(TEST, name)
        if not os.path.isdir(d):
            continue
        v = par_values(par)
        case = {
            "prolongation_type": v.get("prolongation_type"),
            "prolongation_order": int(v.get("prolongation_order", "1")),
            "prolongate_per_group": v.get("prolongate_per_group", "no"),
            "ghost_size": int(v.get("ghost_size", "0")),
            "ncells": int(v.get("$ncells", "0")),
            "xmin": [v.get("xmin"), v.get("ymin"), v.get("zmin")],
            "max_error": {},
        }
        for gf in ["gf%d%d%d" % (a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]:
            errs = []
            for it in range(3):
                f = os.path.join(d, "testprolongate-%s_max_error.it%06d.tsv" % (gf, it))
                if not os.path.exists(f):
                    errs.append(None)
                    continue
                rows = [l.split() for l in open(f) if l.strip() and not l.startswith("#")]
                errs.append(float(rows[0][2]))
            case["max_error"][gf] = errs
        cases[name] = case
    with open(OUT, "w") as f:
        json.dump(cases, f, indent=1, sort_keys=True)
    print("wrote %s: %d cases" % (OUT, len(cases)))


if __name__ == "__main__":
    main()
