"""Copy OpenQP's own RHF gradient known-answer data (the JSON its CI checks)
into tests/golden/h2o_rhf_gradient_kat.json.  Data only: atoms, coordinates
(bohr), energy and gradient (Eh/bohr) of examples/HF/H2O_RHF-HF_GRADIENT.json."""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(REF, "examples", "HF", "H2O_RHF-HF_GRADIENT.json")
d = json.load(open(src))
out = {"source": "examples/HF/H2O_RHF-HF_GRADIENT.json (OpenQP known-answer test: H2O RHF/6-31G*, Cartesian d)",
       "atoms": d["atoms"], "coord_bohr": d["coord"], "energy": d["energy"], "grad": d["grad"]}
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "h2o_rhf_gradient_kat.json"), "w"),
          indent=1)
