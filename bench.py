#!/usr/bin/env python
"""J+K Fock-build benchmark on one MI355X.

Workload (default): 64 H2O on a 4x4x4 grid (3.0 A spacing, seeded random
orientations), cc-pVDZ (spherical d), 1536 basis functions, a seeded
S-orthonormal density of 320 occupied orbitals (D = 2 C C^T), int2e_cutoff
5e-11.  One step = one J+K build (G = J - K/2 for that density) with the
density and result resident in HBM; the class launches of a build are spread
over the session's four streams.  Per-class times and the roofline of the
dominant kernel come from further builds whose launches are serialised and
timed one by one with HIP events.  The plain FP64 build and the mixed
build (FP32 for quartets whose Schwarz x density bound is below the FP32
cutoff, FP64 accumulation) are both timed; ``value`` is the mixed rate when
its energy error |dE2| = |0.5 Tr D (G_mixed - G_fp64)| is within 1 uEh,
otherwise the FP64 rate.

Further legs, each reported in its own field of the same JSON line and each
switched off by its --no-* flag:
- the 2e gradient on C20H42/6-31G(d) (FP64 and mixed);
- a full RHF SCF of (H2O)32/6-31G(d);
- the MRSF sigma and MRSF 2e gradient seams on (H2O)32/6-31G(d);
- the semilocal Vxc build (B3LYP5) on the (H2O)64/cc-pVDZ workload above, with
  ``--xc-only`` running that leg alone;
- the f-shell leg on (H2O)16/cc-pVTZ.

Prints one JSON line (rank 0, N = 1).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "J+K Fock builds/s, mixed vs FP64 GPU, at |ΔE|≤1 µEh; gradient evals/s"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (= FP64 matrix) dense peak
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector dense peak (packed)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def trace_energy(dp, fp, nbf):
    """0.5 Tr(D G) for packed symmetric D, G: 2*dot - diagonal (mathlib.F90:69-85)."""
    idx = np.arange(1, nbf + 1)
    diag = idx * (idx + 1) // 2 - 1
    return 0.5 * (2.0 * np.dot(dp, fp) - np.dot(dp[diag], fp[diag]))


def cpu_cores():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name"):
                return ln.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def cpu_baseline(bas, dp, gpu_quartets, target_s=15.0):
    """Time the CPU FP64 restatement (oracle) on OpenQP's round-robin 1/size
    share of shell pairs (int2.F90:759-761) and scale to builds/s."""
    from oracle import oracle as O
    Q = O.schwarz(bas)
    size = 4096
    t0 = time.perf_counter()
    _, st = O.fock_jk(bas, dp, Q=Q, rank=0, size=size)
    dt = time.perf_counter() - t0
    # the time per share is not proportional at small shares (fixed costs): refine
    # the share until the sample takes at least half the target
    for _ in range(3):
        if dt >= target_s / 2 or size == 1:
            break
        size = max(1, int(size * dt / target_s))
        t0 = time.perf_counter()
        _, st = O.fock_jk(bas, dp, Q=Q, rank=0, size=size)
        dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", cpu_cores()))
    return {"value": round(1.0 / (dt * size), 6), "unit": "builds/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle FP64 J+K on shell-pair share rank 0 of {size} (int2.F90:759-761 round robin), "
                      f"{int(st[1])} quartets in {dt:.2f} s on {threads} threads of {cpu_model()}; "
                      f"x{size} for one build",
            "quartets_per_s": round(int(st[1]) / dt, 1)}


# mixed-precision 2e gradient: template quartets with Q_ij Q_kl below this run
# in FP32 (max |dgrad| against FP64 below 1e-7 Eh/bohr on both gradient workloads)
GRAD_FP32_CUT = 1e-2


def mixed_gradient_rate(fb, dp, steps, warmup):
    """Seconds per 2e-gradient evaluation in mixed precision, and its gradient;
    leaves the builder's gradient in FP64."""
    fb.set_grad_precision(1, GRAD_FP32_CUT)
    for _ in range(max(warmup, 1)):
        fb.grad2(dp)
    t0 = time.perf_counter()
    for _ in range(steps):
        g = fb.grad2(dp)
    dt = (time.perf_counter() - t0) / steps
    fb.set_grad_precision(0, 0.0)
    return dt, g


def gradient_leg(steps, warmup):
    """2e nuclear gradient evals/s on C20H42 / 6-31G(d) (384 bf, Cartesian d) at
    the converged FP64 density of an SCF run on the device Fock builder."""
    from openqp_amd import basis as B, molecules as M
    from openqp_amd.lib import FockBuilder
    from openqp_amd.scf import RHF
    bld, bname, nocc = M.WORKLOADS["c20h42_631gd"]
    Z, xyz = bld(0)
    bas = B.build_basis(bname, Z, xyz)
    fb = FockBuilder(bas)
    t0 = time.perf_counter()
    res = RHF(bas, fb, conv=1e-6, econv=1e-8).run()
    t_scf = time.perf_counter() - t0
    dp = B.pack_tri(res.density)
    for _ in range(warmup):
        fb.grad2(dp)
    t0 = time.perf_counter()
    for _ in range(steps):
        g = fb.grad2(dp)
    dt = (time.perf_counter() - t0) / steps
    st = fb.last_stats
    st = (st.kernel_ms, st.quartets_computed, st.model_flops)
    dtm, gm = mixed_gradient_rate(fb, dp, steps, warmup)
    out = {"workload": "c20h42_631gd", "nbf": bas.nbf, "natm": bas.natm, "scf_energy": res.energy,
           "scf_iterations": res.iterations, "scf_converged": res.converged, "scf_wall_s": round(t_scf, 3),
           "scf_fock_ms_mean": round(float(np.mean(res.fock_ms)), 3),
           "evals_per_s": round(1.0 / dt, 4), "ms_per_eval": round(dt * 1e3, 3),
           "kernel_ms": round(st[0], 3), "quartets": st[1],
           "model_tflops": round(st[2] / (st[0] * 1e-3) / 1e12, 3) if st[0] else None,
           "mixed_evals_per_s": round(1.0 / dtm, 4), "mixed_fp32_cutoff": GRAD_FP32_CUT,
           "mixed_max_dgrad_vs_fp64": float(np.abs(gm - g).max()),
           "max_abs_grad": float(np.abs(g).max()), "translational_sum": float(np.abs(g.sum(axis=0)).max())}
    fb.close()
    return out


def scf_leg():
    """Full direct RHF SCF on the device Fock builder: 32 H2O (4x4x2 grid) /
    6-31G(d), 608 bf, incremental Fock builds, DIIS, converged to 1e-8 Eh.
    Reports SCF wall time and the J+K builds/s over the SCF's builds."""
    from openqp_amd import basis as B, molecules as M
    from openqp_amd.lib import FockBuilder
    from openqp_amd.scf import RHF
    bld, bname, nocc = M.WORKLOADS["h2o32_631gd"]
    Z, xyz = bld(0)
    bas = B.build_basis(bname, Z, xyz)
    fb = FockBuilder(bas)
    t0 = time.perf_counter()
    res = RHF(bas, fb, nocc=nocc, conv=1e-6, econv=1e-8).run()
    wall = time.perf_counter() - t0
    fms = np.array(res.fock_ms)
    fb.close()
    return {"workload": "h2o32_631gd", "nbf": bas.nbf, "nocc": nocc, "energy": res.energy,
            "converged": res.converged, "iterations": res.iterations, "scf_wall_s": round(wall, 3),
            "fock_builds_per_s": round(1e3 / float(fms.mean()), 3), "fock_ms_mean": round(float(fms.mean()), 3),
            "fock_ms_total": round(float(fms.sum()), 3), "incremental": True, "guess": "extended Hueckel"}


def xc_leg(steps, name="b3lyp5"):
    """Semilocal Vxc build (routec_vxc's work) on the (H2O)64/cc-pVDZ workload:
    B3LYP5 on the default grid (96 radial shells x 450 angular points per
    atom, SSF partition), closed shell, S-orthonormal random density; the grid
    build (first call, per geometry) is timed apart from the builds."""
    from openqp_amd import basis as B, molecules as M
    from openqp_amd import lib as L
    bld, bname, nocc = M.WORKLOADS["h2o64_ccpvdz"]
    Z, xyz = bld(0)
    bas = B.build_basis(bname, Z, xyz)
    S, _, _ = L.int1e(bas)
    dp = B.pack_tri(M.random_density(S, nocc, seed=3))
    fb = L.FockBuilder(bas)
    fb.set_xc(name)
    t0 = time.perf_counter()
    npts = len(fb.xc_grid())
    tg = time.perf_counter() - t0
    fb.vxc(dp, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        V, exc, nel = fb.vxc(dp, 1)
    dt = (time.perf_counter() - t0) / steps
    st = fb.xc_grid_stats()
    gemm_flops = 4.0 * st["sum_np_nb2"]
    log(f"[bench] Vxc {name}: {dt * 1e3:.1f} ms/build over {npts} points (grid build {tg * 1e3:.0f} ms), "
        f"N_elec {nel:.6f}, {st['batches']} batches, GEMM model {gemm_flops / 1e12:.2f} TFLOP")
    return {"workload": "h2o64_ccpvdz", "functional": name, "nbf": bas.nbf, "grid_points": npts,
            "grid": "96 x (15 x 30), SSF", "ms_per_build": round(dt * 1e3, 2), "grid_build_ms": round(tg * 1e3, 1),
            "points_per_s": round(npts / dt, 1), "E_xc": exc, "N_elec": nel, "N_expected": 2 * nocc,
            "batches": st["batches"], "mean_basis_functions_per_point": round(st["sum_np_nb"] / npts, 1),
            "gemm_model_tflop": round(gemm_flops / 1e12, 3),
            "gemm_model_tflops_over_build": round(gemm_flops / dt / 1e12, 2)}


def sigma_leg(steps):
    """MRSF sigma seam (routec_sig_*) on (H2O)32 / 6-31G(d): two trial vectors
    per iteration (14 non-symmetric transition densities through the device
    response pass), synthetic S-orthonormal MOs and a diagonal MO Fock matrix
    (two open shells over the 160 doubly occupied orbitals)."""
    from openqp_amd import basis as B, molecules as M
    from openqp_amd import lib as L
    bld, bname, nocc = M.WORKLOADS["h2o32_631gd"]
    Z, xyz = bld(0)
    bas = B.build_basis(bname, Z, xyz)
    n = bas.nbf
    S, _, _ = L.int1e(bas)
    rng = np.random.default_rng(1)
    w, U = np.linalg.eigh(S)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    va = U @ np.diag(w ** -0.5) @ U.T @ Q
    fmo = np.diag(np.sort(rng.standard_normal(n)))
    nocca, noccb, nv = nocc + 1, nocc - 1, 2
    fb = L.FockBuilder(bas)
    fb.make_default()
    sg = L.SigmaSession(va, va, fmo, fmo, nocca, noccb, 1)
    bv = rng.standard_normal((nocca * (n - noccb), nv)) * 0.01
    sg.iter(bv)
    t0 = time.perf_counter()
    for _ in range(steps):
        sg.iter(bv)
    dt = (time.perf_counter() - t0) / steps
    sg.close()
    # MRSF 2e gradient (routec_grad2_mrsf's work) at synthetic densities of the same system
    A = va[:, :nocca]
    d = np.array([(A @ A.T).T, (A[:, :noccb] @ A[:, :noccb].T).T])
    p = 0.01 * d
    spc = rng.standard_normal((n, n, 7)) * 1e-3
    fb.grad2_mrsf(d, p, spc)
    t0 = time.perf_counter()
    for _ in range(steps):
        fb.grad2_mrsf(d, p, spc)
    dtg = (time.perf_counter() - t0) / steps
    fb.close()
    return {"workload": "h2o32_631gd", "nbf": n, "trial_vectors": nv, "densities": 7 * nv,
            "ms_per_iteration": round(dt * 1e3, 2), "mrsf_grad2_ms": round(dtg * 1e3, 2)}


def fshell_leg(fp32_cutoff, steps):
    """(H2O)16 / cc-pVTZ (928 bf, pure d and f shells): full direct RHF SCF to
    1e-8 Eh on the device builder in FP64 and in mixed precision, the total
    RHF gradient at each converged density, and at the FP64 density the J+K
    build rate (FP64 and mixed) and the 2e-gradient rate."""
    from openqp_amd import basis as B, molecules as M
    from openqp_amd.lib import FockBuilder
    from openqp_amd.scf import RHF
    from openqp_amd.gradient import rhf_gradient
    bld, bname, nocc = M.WORKLOADS["h2o16_ccpvtz"]
    Z, xyz = bld(0)
    bas = B.build_basis(bname, Z, xyz)
    fb = FockBuilder(bas)
    t0 = time.perf_counter()
    r64 = RHF(bas, fb, nocc=nocc, conv=1e-6, econv=1e-8).run()
    w64 = time.perf_counter() - t0
    g64 = rhf_gradient(bas, r64.orbitals, r64.orbital_energies, nocc, fb)["total"]
    dp = B.pack_tri(r64.density)
    fb.grad2(dp)
    t0 = time.perf_counter()
    for _ in range(steps):
        g2 = fb.grad2(dp)
    dtg = (time.perf_counter() - t0) / steps
    # last_stats is overwritten by every call: keep the values
    g_kernel_ms, g_quartets = fb.last_stats.kernel_ms, fb.last_stats.quartets_computed
    dtgm, g2m = mixed_gradient_rate(fb, dp, steps, 1)

    def builds(precision):
        fb.set_precision(precision, fp32_cutoff)
        fb.fock_jk(dp)
        t = time.perf_counter()
        for _ in range(steps):
            fb.fock_jk(dp)
        st = fb.last_stats
        return (time.perf_counter() - t) / steps, st.quartets_computed, st.quartets_fp32
    dt64, q64, _ = builds(0)
    dtmx, _, q32 = builds(1)
    # mixed-precision SCF and gradient against the FP64 ones
    t0 = time.perf_counter()
    rmx = RHF(bas, fb, nocc=nocc, conv=1e-6, econv=1e-8).run()
    wmx = time.perf_counter() - t0
    gmx = rhf_gradient(bas, rmx.orbitals, rmx.orbital_energies, nocc, fb)["total"]
    fb.close()
    return {"workload": "h2o16_ccpvtz", "nbf": bas.nbf, "nshell": bas.nsh, "nocc": nocc,
            "scf_energy_fp64": r64.energy, "scf_iterations_fp64": r64.iterations, "scf_converged": r64.converged,
            "scf_wall_s_fp64": round(w64, 3), "scf_fock_ms_mean_fp64": round(float(np.mean(r64.fock_ms)), 3),
            "fp64_builds_per_s": round(1.0 / dt64, 4), "mixed_builds_per_s": round(1.0 / dtmx, 4),
            "mixed_speedup": round(dt64 / dtmx, 3), "fp32_cutoff": fp32_cutoff,
            "quartets_per_build": q64, "fp32_quartets": q32,
            "gradient_evals_per_s": round(1.0 / dtg, 4), "gradient_kernel_ms": round(g_kernel_ms, 3),
            "gradient_quartets": g_quartets,
            "mixed_gradient_evals_per_s": round(1.0 / dtgm, 4),
            "mixed_gradient_max_dgrad_vs_fp64": float(np.abs(g2m - g2).max()),
            "scf_energy_mixed": rmx.energy, "scf_iterations_mixed": rmx.iterations,
            "scf_wall_s_mixed": round(wmx, 3),
            "mixed_dE_scf_hartree": abs(rmx.energy - r64.energy),
            "mixed_max_dgrad": float(np.abs(gmx - g64).max()),
            "max_abs_grad": float(np.abs(g64).max()), "translational_sum": float(np.abs(g64.sum(axis=0)).max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="h2o64_ccpvdz")
    ap.add_argument("--fp32-cutoff", type=float, default=None,
                    help="FP32 cutoff of the mixed build; default: swept over 1e-3..1e-6, the fastest "
                         "with |dE2| <= 1 uEh is timed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-scf", action="store_true")
    ap.add_argument("--no-fshell", action="store_true")
    ap.add_argument("--no-sigma", action="store_true")
    ap.add_argument("--no-xc", action="store_true")
    ap.add_argument("--xc-only", action="store_true",
                    help="run only the Vxc leg (for profiling it) and print its record")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--profile-builds", type=int, default=3,
                    help="builds after the timed ones with per-launch HIP events (launches serialised): the "
                         "per-class times and the roofline of the dominant kernel")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: measures one GPU only (--gpus {a.gpus} requested); across GPUs the bra "
              "shell pairs would be split by cost and G all-reduced once per build, which is not "
              "implemented here", file=sys.stderr)
        sys.exit(2)
    if a.xc_only:
        print(json.dumps({"vxc": xc_leg(a.steps)}))
        return

    import torch
    from openqp_amd import basis as B, molecules as M
    from openqp_amd import lib as L

    bld, bname, nocc = M.WORKLOADS[a.workload]
    Z, xyz = bld(0)
    bas = B.build_basis(bname, Z, xyz)
    log(f"[bench] workload {a.workload}: nbf {bas.nbf}, nsh {bas.nsh}, natm {bas.natm}")
    t0 = time.perf_counter()
    S, _, _ = L.int1e(bas)
    D = M.random_density(S, nocc, seed=0)
    dp = B.pack_tri(D)
    log(f"[bench] density ready ({time.perf_counter() - t0:.1f} s)")

    dev = torch.device("cuda:0")
    d_dev = torch.tensor(dp, dtype=torch.float64, device=dev)
    f64 = torch.zeros_like(d_dev)
    fmx = torch.zeros_like(d_dev)
    stream = torch.cuda.current_stream().cuda_stream

    fb = L.FockBuilder(bas)
    t0 = time.perf_counter()
    fb.schwarz()
    log(f"[bench] geometry setup (pairs + Schwarz) {time.perf_counter() - t0:.2f} s")

    def run(fout, steps, warm, precision, cut=0.0, profile_builds=1):
        """Time `steps` builds (the class launches spread over the session's
        streams), then run `profile_builds` more with per-launch HIP events,
        which serialise the launches so that each is timed alone."""
        fb.set_precision(precision, cut)
        for _ in range(warm):
            fb.fock_jk_device(d_dev.data_ptr(), fout.data_ptr(), stream=stream)
        torch.cuda.synchronize()
        st_last = None
        t = time.perf_counter()
        for _ in range(steps):
            st_last = fb.fock_jk_device(d_dev.data_ptr(), fout.data_ptr(), stream=stream)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        prof = {}
        launches = {}
        fb.set_profiling(True)
        for _ in range(profile_builds):
            fb.fock_jk_device(d_dev.data_ptr(), fout.data_ptr(), stream=stream)
            # a class may run as two launches (grouped and segmented bra-row
            # kernels): per-class figures are per build
            seen = set()
            for p in fb.last_profile():
                r = launches.setdefault((p["cls"], p["kernel"]), {"ms": 0.0, "flops": 0.0, "n": 0})
                r["ms"] += p["ms"]; r["flops"] += p["flops"]; r["n"] += 1
                q = prof.setdefault(p["cls"], {"ms": 0.0, "flops": 0.0, "quartets": 0, "prims": 0, "prims_eval": 0,
                                               "screened": 0, "slots": 0, "n": 0, "kernel": p["kernel"]})
                q["ms"] += p["ms"]; q["flops"] += p["flops"]; q["quartets"] += p["quartets"]
                q["prims"] += p["prims"]; q["prims_eval"] += p["prims_eval"]
                q["screened"] += p["screened"]; q["slots"] += p["slots"]
                if p["cls"] not in seen:
                    q["n"] += 1
                    seen.add(p["cls"])
                if p["kernel"] not in q["kernel"].split("+"):
                    q["kernel"] += "+" + p["kernel"]
        torch.cuda.synchronize()
        fb.set_profiling(False)
        prof["_launches"] = launches
        return dt, st_last, prof

    log("[bench] FP64 builds")
    dt64, st64, prof64 = run(f64, a.steps, a.warmup, 0, profile_builds=a.profile_builds)
    log(f"[bench] FP64 {dt64 * 1e3:.1f} ms/build, quartets {st64.quartets_computed}, "
        f"model {st64.model_flops / 1e12:.3f} TFLOP")
    e64_ref = trace_energy(dp, f64.cpu().numpy(), bas.nbf)
    # FP32 cutoff: the fastest of the swept values whose energy error is within 1 uEh
    sweep = []
    cuts = [a.fp32_cutoff] if a.fp32_cutoff is not None else [1e-3, 1e-4, 1e-5, 1e-6]
    for cut in cuts:
        dts, sts, _ = run(fmx, 1, 1, 1, cut, profile_builds=0)
        de = abs(trace_energy(dp, fmx.cpu().numpy(), bas.nbf) - e64_ref)
        sweep.append({"fp32_cutoff": cut, "ms": round(dts * 1e3, 3), "dE2_hartree": de,
                      "fp32_quartets": sts.quartets_fp32})
        log(f"[bench] sweep fp32 cutoff {cut:g}: {dts * 1e3:.1f} ms, |dE2| {de:.2e}")
    ok = [x for x in sweep if x["dE2_hartree"] <= 1e-6]
    best = min(ok, key=lambda x: x["ms"]) if ok else min(sweep, key=lambda x: x["dE2_hartree"])
    a.fp32_cutoff = best["fp32_cutoff"]
    log(f"[bench] mixed builds (fp32 cutoff {a.fp32_cutoff:g})")
    dtmx, stmx, profmx = run(fmx, a.steps, a.warmup, 1, a.fp32_cutoff, profile_builds=a.profile_builds)
    log(f"[bench] mixed {dtmx * 1e3:.1f} ms/build, fp32 quartets {stmx.quartets_fp32}")

    g64 = f64.cpu().numpy()
    gmx = fmx.cpu().numpy()
    e64 = trace_energy(dp, g64, bas.nbf)
    emx = trace_energy(dp, gmx, bas.nbf)
    dE = abs(emx - e64)
    dG = float(np.abs(gmx - g64).max())
    mixed_ok = dE <= 1e-6
    dt = dtmx if mixed_ok else dt64
    st = stmx if mixed_ok else st64
    prof = profmx if mixed_ok else prof64

    # dominant kernel: the single launch with the largest average device time
    (dom_cls, dom_kern), dom = max(prof["_launches"].items(), key=lambda kv: kv[1]["ms"] / kv[1]["n"])
    dom_ms = dom["ms"] / dom["n"]
    dom_flops = dom["flops"] / dom["n"]
    achieved = dom_flops / (dom_ms * 1e-3) / 1e12
    build_tflops = st.model_flops / (st.kernel_ms * 1e-3) / 1e12 if st.kernel_ms > 0 else 0.0
    kname = f"{dom_kern}<{dom_cls // 1000},{dom_cls // 100 % 10},{dom_cls // 10 % 10},{dom_cls % 10}>"
    roof = {"bound": "valu", "kernel": kname,
            "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None,
            "launch_ms": round(dom_ms, 3), "flops_per_launch": dom_flops,
            "build_model_tflops": round(build_tflops, 3)}
    # HBM traffic of that kernel from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE record of this workload (the counters cannot be read here)
    tfile = os.path.join(ROOT, "openqp_amd", "data", "traffic_h2o64_ccpvdz.json")
    if a.workload == "h2o64_ccpvdz" and os.path.exists(tfile):
        tr = json.load(open(tfile))
        rec = tr.get("kernels", {}).get(kname, {}).get("mixed" if mixed_ok else "fp64")
        if rec:
            roof["traffic"] = rec["fetch_bytes"] + rec["write_bytes"]
            roof["traffic_source"] = tr["source"]

    grad = None
    if not a.no_grad:
        log("[bench] gradient: RHF on the C20H42/6-31G(d) workload, then the 2e gradient")
        grad = gradient_leg(a.steps, a.warmup)

    scf = None
    if not a.no_scf:
        log("[bench] SCF: RHF on the (H2O)32/6-31G(d) workload, incremental builds")
        scf = scf_leg()

    sig = None
    if not a.no_sigma:
        log("[bench] MRSF sigma: routec_sig_iter on the (H2O)32/6-31G(d) workload, two trial vectors")
        sig = sigma_leg(2)

    xcl = None
    if not a.no_xc:
        log("[bench] Vxc: B3LYP5 on the (H2O)64/cc-pVDZ workload")
        xcl = xc_leg(3)

    fsh = None
    if not a.no_fshell:
        log("[bench] f shells: RHF SCF, builds and gradient on the (H2O)16/cc-pVTZ workload")
        fsh = fshell_leg(a.fp32_cutoff, 2)
        log(f"[bench] f shells: FP64 {fsh['fp64_builds_per_s']} builds/s, gradient {fsh['gradient_evals_per_s']} "
            f"evals/s, mixed |dE_SCF| {fsh['mixed_dE_scf_hartree']:.2e}")

    cpu = None
    if not a.no_cpu_baseline:
        log("[bench] CPU baseline (oracle restatement, sampled)")
        try:
            cpu = cpu_baseline(bas, dp, st64.quartets_computed, a.cpu_seconds)
        except Exception as e:  # reported baseline only
            cpu = {"error": str(e)}

    out = {
        "metric": METRIC, "value": round(1.0 / dt, 4), "unit": "builds/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if mixed_ok else "f64",
        "data": "synthetic (seeded geometry and S-orthonormal random density)",
        "config": {"workload": a.workload, "basis": bname, "nbf": bas.nbf, "nshell": bas.nsh, "nocc": nocc,
                   "int2e_cutoff": 5e-11, "fp32_cutoff": a.fp32_cutoff,
                   "precision": "mixed" if mixed_ok else "fp64"},
        "fp64_builds_per_s": round(1.0 / dt64, 4), "mixed_builds_per_s": round(1.0 / dtmx, 4),
        "mixed_speedup": round(dt64 / dtmx, 3), "mixed_dE2_hartree": dE, "mixed_max_dG": dG,
        "quartets_per_build": st.quartets_computed, "quartets_per_s": round(st.quartets_computed / dt, 1),
        "prim_quartets_per_build": st.prim_quartets, "fp32_quartets": stmx.quartets_fp32,
        "E2_fp64_hartree": e64,
        "fp32_cutoff_sweep": sweep,
        "gradient_evals_per_s": grad["evals_per_s"] if grad else None,
        "mixed_gradient_evals_per_s": grad["mixed_evals_per_s"] if grad else None,
        "gradient": grad,
        "scf": scf,
        "fshell": fsh,
        "mrsf_sigma": sig,
        "vxc": xcl,
        "roofline": roof, "cpu_baseline": cpu,
        "per_class": {str(k): {"kernel": v["kernel"], "ms_fp64": round(v["ms"] / v["n"], 3),
                               "ms_mixed": round(profmx[k]["ms"] / profmx[k]["n"], 3) if k in profmx else None,
                               "quartets": v["quartets"] // v["n"], "prim_quartets": v["prims"] // v["n"],
                               "prim_quartets_evaluated": (profmx.get(k, v)["prims_eval"] // profmx.get(k, v)["n"])
                               or None,
                               "screened": v["screened"] // v["n"], "ket_lanes": v["slots"] // v["n"]}
                      for k, v in sorted((k, v) for k, v in prof64.items() if k != "_launches")},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
