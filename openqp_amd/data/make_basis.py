"""Extract the H/C/N/O entries of the basis sets this engine ships from the
GAMESS ``$DATA`` files of the OpenQP source tree into small JSON files.

Basis-set data are the published Basis Set Exchange tables (v0.8.13); the
JSON files written here are data, not code.  The harmonic flag per basis
follows the BSE function type: Pople sets are Cartesian (6d), Dunning
cc-pVXZ sets are spherical (5d/7f) -- the ``ispher=auto`` convention of
pyoqp/oqp/library/set_basis.py:123-140.

Run once in a tree that has the reference checked out::

    python openqp_amd/data/make_basis.py /root/reference/basis_sets
"""
import json
import os
import sys

ELEMENTS = {"HYDROGEN": 1, "CARBON": 6, "NITROGEN": 7, "OXYGEN": 8}
SETS = {  # file stem -> (output name, harmonic flag)
    "6-31g": ("6-31g", 0),
    "6-31g(d)": ("6-31g(d)", 0),
    "cc-pvdz": ("cc-pvdz", 1),
    "cc-pvtz": ("cc-pvtz", 1),
}
LETTERS = "SPDFGHI"


def parse(path):
    out = {}
    with open(path) as fh:
        lines = [ln.rstrip("\n") for ln in fh]
    i = 0
    cur = None
    while i < len(lines):
        ln = lines[i].strip()
        if ln.upper() in ELEMENTS:
            cur = ELEMENTS[ln.upper()]
            out[cur] = []
            i += 1
            continue
        if cur is not None and ln and ln[0] in LETTERS + "L" and len(ln.split()) == 2:
            kind, n = ln.split()
            n = int(n)
            rows = [lines[i + 1 + k].split() for k in range(n)]
            exps = [float(r[1].replace("D", "E")) for r in rows]
            if kind == "L":
                cs = [float(r[2].replace("D", "E")) for r in rows]
                cp = [float(r[3].replace("D", "E")) for r in rows]
                out[cur].append({"l": 0, "exps": exps, "coefs": cs})
                out[cur].append({"l": 1, "exps": exps, "coefs": cp})
            else:
                cs = [float(r[2].replace("D", "E")) for r in rows]
                out[cur].append({"l": LETTERS.index(kind), "exps": exps, "coefs": cs})
            i += 1 + n
            continue
        if not ln:
            cur = None
        i += 1
    return out


def main(src):
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")
    os.makedirs(here, exist_ok=True)
    for stem, (name, harm) in SETS.items():
        data = parse(os.path.join(src, stem + ".basis"))
        doc = {"name": name, "harmonic": harm, "source": "Basis Set Exchange v0.8.13 via OpenQP basis_sets/" + stem + ".basis",
               "elements": {str(z): data[z] for z in sorted(data)}}
        with open(os.path.join(here, name + ".json"), "w") as fh:
            json.dump(doc, fh, indent=1)
        print(name, {z: len(v) for z, v in data.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/basis_sets")
