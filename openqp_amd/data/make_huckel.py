"""Extract the extended-Hueckel guess data for H..Ne from the OpenQP source
tree into openqp_amd/data/huckel.json (data, not code):

* the MINI (Huzinaga) minimal basis of basis_sets/MINI_huckel.basis for H, C,
  N, O (written as openqp_amd/data/basis/mini_huckel.json, the format of the
  other basis files);
* the atomic orbital energies (Clementi-Roetti tables) and the core/valence
  shell counts of source/huckel_lut.F90 (huckel_eneg, huckel_ncore,
  huckel_nval, huckel_ndval, huckel_lneg), elements 1..10.

Run once in a tree that has the reference checked out::

    python openqp_amd/data/make_huckel.py /root/reference
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_basis  # noqa: E402


def fortran_array(txt, name):
    m = re.search(r"parameter\s*::\s*" + name + r"\([^)]*\)\s*=\s*reshape\(\[(.*?)\]", txt, flags=re.S)
    body = m.group(1).replace("&", " ").replace("D", "E")
    return [float(x) for x in body.replace(",", " ").split()]


def main(ref):
    lut = open(os.path.join(ref, "source", "huckel_lut.F90")).read()
    eneg = fortran_array(lut, "huckel_eneg")          # (18, 103), column-major
    ncore = fortran_array(lut, "huckel_ncore")
    nval = fortran_array(lut, "huckel_nval")
    ndval = fortran_array(lut, "huckel_ndval")        # (4, 103)
    lneg = fortran_array(lut, "huckel_lneg")          # (56, 6)
    elems = {}
    for z in range(1, 11):
        elems[str(z)] = {"eneg": eneg[18 * (z - 1):18 * z], "ncore": int(ncore[z - 1]), "nval": int(nval[z - 1]),
                         "ndval": [int(x) for x in ndval[4 * (z - 1):4 * z]]}
    doc = {"source": "OpenQP source/huckel_lut.F90 (Clementi-Roetti orbital energies, hartree)",
           "wh_k": 1.75, "bitsy": 0.05, "lneg_type1": [int(x) for x in lneg[:56]], "elements": elems}
    with open(os.path.join(HERE, "huckel.json"), "w") as fh:
        json.dump(doc, fh, indent=1)
    data = make_basis.parse(os.path.join(ref, "basis_sets", "MINI_huckel.basis"))
    mini = {"name": "mini_huckel", "harmonic": 0,
            "source": "Huzinaga MINI basis via OpenQP basis_sets/MINI_huckel.basis",
            "elements": {str(z): data[z] for z in sorted(data)}}
    with open(os.path.join(HERE, "basis", "mini_huckel.json"), "w") as fh:
        json.dump(mini, fh, indent=1)
    print({z: len(v) for z, v in data.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
