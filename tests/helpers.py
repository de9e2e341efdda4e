"""Shared test helpers: workloads, densities, header parsing."""
import os
import re

import numpy as np

from openqp_amd import basis as B, molecules as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def workload(name, seed=0):
    bld, bname, nocc = M.WORKLOADS[name]
    Z, xyz = bld(seed)
    return B.build_basis(bname, Z, xyz), nocc


def small_system(basis_name, nw=1, seed=0):
    if nw == 1:
        Z, xyz = M.h2o_kat()
    else:
        Z, xyz = M.water_grid(nw, 1, 1, seed=seed)
    return B.build_basis(basis_name, Z, xyz)


def density_for(oracle, bas, nocc, seed=0):
    S, _, _ = oracle.int1e(bas)
    D = M.random_density(S, nocc, seed=seed)
    return D, B.pack_tri(D)


def header_functions():
    txt = open(os.path.join(ROOT, "include", "oqph.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:int|void|const char \*)\s*\**\s*(\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))
