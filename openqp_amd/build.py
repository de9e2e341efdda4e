"""Build liboqph.so (HIP, gfx950) in-tree: openqp_amd/_lib/liboqph.so.

Translation units: session.hip (host session + C ABI), int1e.hip (host
one-electron integrals), dispatch.hip (class dispatch, small kernels, Rys
table data) and jk_class.hip compiled once per quartet class with
-DXA..-DXD.  Objects compile in parallel, the slowest classes first, and are
cached by a hash of each unit and the headers it includes, so an unchanged
class is not rebuilt; the shared library is always relinked.
"""
from __future__ import annotations

import hashlib
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
OBJ = os.path.join(OUT, "obj")
LIB = os.path.join(OUT, "liboqph.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("OQPH_ARCH", "gfx950")
FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-variable", "-Wno-pass-failed"]
CLASSES = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 0), (1, 1, 1, 1),
           (2, 0, 0, 0), (2, 0, 1, 0), (2, 0, 1, 1), (2, 0, 2, 0), (2, 1, 0, 0), (2, 1, 1, 0),
           (2, 1, 1, 1), (2, 1, 2, 0), (2, 1, 2, 1), (2, 2, 0, 0), (2, 2, 1, 0), (2, 2, 1, 1),
           (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2)]
# classes with an f shell that get class-templated kernels (the rest run on
# the run-time-class kernels of hl.hip); the lists match jk_template() and
# grad_template() in dispatch.hip
F_JK = [(3, 0, 0, 0), (3, 0, 1, 0), (3, 0, 1, 1), (3, 0, 2, 0), (3, 0, 2, 1), (3, 0, 3, 0), (3, 1, 0, 0),
        (3, 1, 1, 0), (3, 1, 2, 0), (3, 2, 0, 0), (3, 2, 1, 0), (3, 3, 0, 0)]
# the gradient classes whose templates would spill 9-17 KB per lane run on
# the run-time-class kernel instead
GRAD_HL = [(2, 2, 2, 2), (2, 2, 2, 1)]
F_GRAD = [(3, 0, 0, 0), (3, 0, 1, 0), (3, 0, 1, 1), (3, 0, 2, 0), (3, 0, 3, 0), (3, 1, 0, 0), (3, 1, 1, 0),
          (3, 2, 0, 0), (3, 3, 0, 0)]


def _cost(c):
    a, b, cc, d = c
    return (a + 1) * (a + 2) * (b + 1) * (b + 2) * (cc + 1) * (cc + 2) * (d + 1) * (d + 2)


# headers each translation unit depends on (cache key = unit + these + flags)
DEPS = {
    "jk_class.hip": ["jk_device.h", "jkg_device.h", "eri_device.h", "oqph_common.h", "rys_table.inc", "rys_table_f32.inc"],
    "dispatch.hip": ["eri_device.h", "grad_device.h", "grad_params.h", "oqph_common.h", "rys_table.inc", "rys_table_f32.inc"],
    "session.hip": ["oqph_common.h", "grad_params.h", "ao_xform.h", "session_view.h", "../../include/oqph.h"],
    "grad_class.hip": ["grad_device.h", "grad_params.h", "eri_device.h", "oqph_common.h", "rys_table.inc", "rys_table_f32.inc"],
    "int1e.hip": ["../../include/oqph.h", "ao_xform.h", "rys_table.inc", "rys_table_f32.inc"],
    "hl.hip": ["hl_device.h", "jk_device.h", "grad_params.h", "eri_device.h", "oqph_common.h", "rys_table.inc",
               "rys_table_f32.inc"],
    "sigma.hip": ["../../include/oqph.h"],
    "xc.hip": ["../../include/oqph.h", "ao_xform.h", "session_view.h"],
}


EXTRA = {"int1e.hip": ["-fopenmp"], "sigma.hip": ["-fopenmp"]}


def _key(unit: str, defs) -> str:
    h = hashlib.sha1()
    for f in [unit] + DEPS.get(unit, []):
        with open(os.path.join(SRC, f), "rb") as fh:
            h.update(f.encode() + fh.read())
    h.update(" ".join(FLAGS + EXTRA.get(unit, []) + list(defs)).encode())
    return h.hexdigest()[:16]


def _jobs():
    units = [("session.hip", [], "session.o"), ("dispatch.hip", [], "dispatch.o"), ("int1e.hip", [], "int1e.o"),
             ("hl.hip", [], "hl.o"), ("sigma.hip", [], "sigma.o"), ("xc.hip", [], "xc.o")]
    heavy, light = [], []
    per_class = [(c, "jk_class.hip", "jk") for c in CLASSES + F_JK] + \
                [(c, "grad_class.hip", "gr") for c in CLASSES + F_GRAD if c not in GRAD_HL]
    for c, unit, tag in sorted(per_class, key=lambda t: _cost(t[0]), reverse=True):
        a, b, cc, d = c
        defs = [f"-DXA={a}", f"-DXB={b}", f"-DXC={cc}", f"-DXD={d}"]
        (heavy if _cost(c) > 1000 else light).append((unit, defs, f"{tag}_{a}{b}{cc}{d}.o"))
    return heavy + units + light


def _compile(job, verbose):
    unit, defs, obj = job
    out = os.path.join(OBJ, _key(unit, defs) + "_" + obj)
    if os.path.exists(out):
        return out
    cmd = [HIPCC] + FLAGS + EXTRA.get(unit, []) + defs + ["-c", os.path.join(SRC, unit), "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {unit} {defs}:\n{r.stdout}\n{r.stderr}")
    shutil.move(out + ".tmp", out)
    return out


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJ, exist_ok=True)
    wanted = {_key(u, d) + "_" + o for u, d, o in _jobs()}
    for f in os.listdir(OBJ):  # drop objects of other source states
        if f not in wanted:
            os.remove(os.path.join(OBJ, f))
    jobs = jobs or int(os.environ.get("MAX_JOBS", min(16, os.cpu_count() or 8)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda j: _compile(j, verbose), _jobs()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "-fopenmp", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lrocblas"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
