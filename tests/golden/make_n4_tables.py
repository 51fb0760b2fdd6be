"""Extract the order-4 DG operator tables held by the reference as a JSON fixture.

The reference solver hardcodes its N=4 operators as literal arrays in
``src/constants/constant_{r,s,Dr,Ds,Drw,Dsw,LIFT}.h`` and the face-node mask
``FMASK`` in ``src/constants/constant_util.h:15``.  These literal values are
the only known-answer data the reference holds for the DG path; the operator
generators of the oracle and of the library are checked against them.

The script parses the headers as text (numbers only) and writes
``n4_tables.json`` next to itself.  Run it in a container where the reference
tree is mounted at /root/reference:

    python tests/golden/make_n4_tables.py
"""
import json
import os
import re
import sys

REF = os.environ.get("DG_REFERENCE_ROOT", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))

NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def parse_array(path, name):
    text = open(path).read()
    m = re.search(r"\b" + name + r"\s*\[[^\]]*\]\s*=\s*\{([^}]*)\}", text)
    if m is None:
        raise RuntimeError(f"array {name} not found in {path}")
    return [float(v) if "." in v or "e" in v.lower() else int(v)
            for v in re.findall(NUM, m.group(1))]


def main():
    cdir = os.path.join(REF, "src", "constants")
    out = {"N": 4, "Np": 15, "Nfp": 5, "source": "src/constants/*.h"}
    for nm, fn in [("r", "constant_r.h"), ("s", "constant_s.h"),
                   ("Dr", "constant_Dr.h"), ("Ds", "constant_Ds.h"),
                   ("Drw", "constant_Drw.h"), ("Dsw", "constant_Dsw.h"),
                   ("LIFT", "constant_LIFT.h")]:
        vals = parse_array(os.path.join(cdir, fn), nm)
        out[nm] = [float(v) for v in vals]
    out["FMASK"] = [int(v) for v in parse_array(os.path.join(cdir, "constant_util.h"), "FMASK")]
    sizes = {"r": 15, "s": 15, "Dr": 225, "Ds": 225, "Drw": 225, "Dsw": 225,
             "LIFT": 225, "FMASK": 15}
    for k, n in sizes.items():
        if len(out[k]) != n:
            sys.exit(f"{k}: expected {n} values, parsed {len(out[k])}")
    with open(os.path.join(HERE, "n4_tables.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", os.path.join(HERE, "n4_tables.json"))


if __name__ == "__main__":
    main()
