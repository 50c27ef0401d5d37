"""Extract the expected outputs of PISM's regression scripts into JSON fixtures.

Reads the heredoc blocks of test/regression/test_{14,15,20,21,22,24,25,26}.sh
under the reference tree (text only, nothing is executed) and writes
tests/golden/regression_outputs.json. Run from the repository root:

    python tests/golden/make_golden.py /root/reference
"""
import json
import os
import re
import sys


def heredoc(path):
    text = open(path).read()
    m = re.search(r"<<END-OF-OUTPUT\n(.*?)END-OF-OUTPUT", text, re.S)
    return m.group(1).splitlines()


def flow_table(lines):
    name = re.match(r'flow law:\s+"(.*)"', lines[0]).group(1)
    rows = []
    for ln in lines[4:]:
        m = re.match(r"\s*([-+0-9.e]+)\s+([-+0-9.e]+)\s+([-+0-9.e]+)\s*=\s*([-+0-9.e]+)", ln)
        if m:
            rows.append([float(x) for x in m.groups()] + [m.group(4)])
    return {"name": name, "pressure_line": lines[1], "rows": rows}


def main(ref):
    reg = os.path.join(ref, "test", "regression")
    out = {"flow_tables": {}, "source": "PISM v0.7.3 test/regression heredocs"}
    for num, law in ((20, "gpbld"), (21, "pb"), (24, "hooke"), (25, "gk"), (26, "arr")):
        out["flow_tables"][law] = flow_table(heredoc(os.path.join(reg, "test_%d.sh" % num)))
        out["flow_tables"][law]["script"] = "test/regression/test_%d.sh" % num
    br = heredoc(os.path.join(reg, "test_22.sh"))
    nums = [re.findall(r"[-+]?\d+\.?\d*(?:e[-+]?\d+)?", ln) for ln in br]
    out["bedrough"] = {
        "script": "test/regression/test_22.sh",
        "Nx": int(nums[1][0]), "Ny": int(nums[1][1]),
        "bed": [nums[2][0], nums[2][1]],
        "smoothed": [nums[3][0], nums[3][1]],
        "theta": [nums[4][0], nums[4][1]],
    }
    tc = heredoc(os.path.join(reg, "test_15.sh"))
    rows = [ln.split() for ln in tc if re.match(r"\s+[0-9]", ln)]
    out["test_C"] = {"script": "test/regression/test_15.sh",
                     "columns": ["prcntVOL", "maxH", "avH", "relmaxETA"],
                     "Mx31": rows[0], "Mx41": rows[1]}
    te = heredoc(os.path.join(reg, "test_14.sh"))
    rows = [ln.split() for ln in te if re.match(r"\s+[0-9]", ln)]
    out["test_E"] = {"script": "test/regression/test_14.sh",
                     "geometry": ["prcntVOL", "maxH", "avH", "relmaxETA"],
                     "base_vels": ["maxvector", "avvector", "prcntavvec", "maxub", "maxvb"],
                     "Mx21": {"geometry": rows[0], "base_vels": rows[1]},
                     "Mx41": {"geometry": rows[2], "base_vels": rows[3]}}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "regression_outputs.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
