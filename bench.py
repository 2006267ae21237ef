"""FMM matvec benchmark on one MI355X.

Workload: Laplace G FMM matvec, N = 1,000,000 points uniform in the unit
cube (seed 0), charges uniform in [-1, 1], p = 10 (degrees 0..9), unsoftened
kernel with the self pair excluded.  One step = one full matvec with the
inputs already resident in HBM: tree build (Morton keys, radix sort, box
lists, interaction lists) + P2P + P2M + M2M + M2L + L2L + L2P.  The metric
is N / t_step in Mpoints/s for the mixed-precision engine; the FP64 engine
is timed in the same run, and the error of both against an FP64 direct sum
on 1,000 sampled targets is reported (the mixed number is valid only if its
rel-L2 potential error is <= 1e-5).

Prints one JSON line (rank 0).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks: HBM GB/s; dense MFMA TFLOP/s (FP16 without sparsity);
# the FP32 vector peak is the published (packed) figure
PEAK = {"hbm": 8000.0, "fp32_mfma": 157.3, "fp64_mfma": 78.6, "fp16_mfma": 2500.0, "fp32_vector": 157.3,
        "fp64_vector": 78.6}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--p", type=int, default=10)
    ap.add_argument("--lmax", type=int, default=0, help="0: pick the faster of 4 and 5")
    ap.add_argument("--targets", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run one 256 MiB device copy first (reference kernel for HBM counter scaling)")
    ap.add_argument("--workload", default="cube", choices=["cube", "helmholtz-sphere", "bem-sphere", "bem-batch"],
                    help="cube: the headline metric; the others time the sphere workloads (one JSON line each)")
    return ap.parse_args()


def traffic_from_profile(n, p, lmax, all_levels, split):
    """HBM bytes per M2L launch (every level in one launch, or the leaf
    level's) measured with rocprofv3 PMC counters (FETCH_SIZE + WRITE_SIZE,
    scaled by a plain copy of known size), kept as a text summary under
    profiles/; None when no summary matches this workload."""
    if split:
        name = "k_m2l_split_all_levels_n1m_p10_pmc.txt"
    else:
        name = "k_m2l_all_levels_n1m_p10_pmc.txt" if all_levels else "k_m2l_leaf_n1m_p10_pmc.txt"
    path = os.path.join(ROOT, "profiles", name)
    if not (os.path.exists(path) and (n, p, lmax) == (1_000_000, 10, 5)):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1])
    return None


def run_engine(F, torch, x, q, p, lmax, precision, steps, warmup):
    # coordinates resident as three device columns (the reference's xi, yi,
    # zi; the engine's layout), so that a step copies nothing
    x = tuple(x[:, k].contiguous() for k in range(3))
    f = F.FMM(p=p, lmax=lmax, precision=precision, convention="exact", box=(0, 0, 0, 1))
    for _ in range(warmup):
        f.set_points(x)
        f.evaluate(q)
    # M2L launch shapes and offset splits per level, timed by whole
    # evaluations on this tree (untimed set-up)
    f.set_points(x)
    f.tuning = f.autotune(q, reps=6)
    # headline: one event pair per call, near field overlapped with far field
    f.set_timing(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.set_points(x)
        v, g = f.evaluate(q)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    # per-phase device times (phases serialized by their events), same steps
    f.set_timing(True)
    phases = {k: 0.0 for k in ("ms_tree", "ms_p2p", "ms_p2m", "ms_m2m", "ms_m2l", "ms_l2p", "ms_total",
                               "ms_m2l_leaf")}
    for _ in range(steps):
        f.set_points(x)
        f.evaluate(q)
        st = f.stats()
        for k in phases:
            phases[k] += st[k]
    phases = {k: v_ / steps for k, v_ in phases.items()}
    # evaluate-only (tree reused, as in a GMRES iteration)
    f.set_timing(False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        f.evaluate(q)
    torch.cuda.synchronize()
    ms_eval = (time.perf_counter() - t1) * 1e3 / steps
    return f, v, g, ms, ms_eval, phases, st


def timed(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, out


def sphere_workload(a, F, torch):
    """The two sphere workloads, each step a full matvec (tree + all phases)
    with inputs in HBM, mixed and FP64, error against the FP64 direct sum:
      helmholtz-sphere: N points uniform on the unit sphere, k = 5 (k D = 10),
        complex charges with parts uniform in [-1, 1], p = 12, 2,000 targets;
      bem-sphere: the 3 Gauss points of each triangle of an icosphere with
        81,920 triangles (245,760 points), unit normals, density a random
        degree <= 8 spherical-harmonic field times the quadrature weights,
        single layer (G) and double layer (Gnj), p = 10, 1,000 targets."""
    rng = np.random.default_rng(0)
    res = {}
    if a.workload == "helmholtz-sphere":
        n, p, k, nt = (a.n, 12, 5.0, 2000)
        xh = rng.standard_normal((n, 3))
        xh = (xh / np.linalg.norm(xh, axis=1)[:, None]).astype(np.float32)
        qh = ((2 * rng.random(n) - 1) + 1j * (2 * rng.random(n) - 1)).astype(np.complex64)
        x = torch.from_numpy(xh).cuda()
        q = torch.from_numpy(qh).cuda()
        idx = torch.from_numpy(np.random.default_rng(1).choice(n, nt, replace=False)).cuda()
        kernels = [("helmholtz", F.HELMHOLTZ, None)]
        metric = "Helmholtz FMM matvec Mpoints/s, N=1M on the unit sphere, k=5, p=12, complex charges"
        workload = (f"oscillatory Helmholtz exp(ikr)/(4 pi r) FMM matvec, N={n} uniform on the unit sphere, k=5, "
                    f"complex q (parts ~U[-1,1]), p=12; step = tree + lists + all phases")
        lmaxes = [a.lmax] if a.lmax else [5, 6]
        mk = lambda prec, lm: F.FMM(p=p, lmax=lm, precision=prec, convention="exact", k=k)  # noqa: E731
        nrm = None
    else:
        from old_fmm_bem_amd import bem
        from scipy.special import sph_harm_y

        m = bem.icosphere(6)
        xh = m.qx.astype(np.float32)
        n, p, nt = len(xh), 10, 1000
        nrmh = (m.qx / np.linalg.norm(m.qx, axis=1)[:, None]).astype(np.float32)
        th = np.arccos(np.clip(m.qx[:, 2] / np.linalg.norm(m.qx, axis=1), -1, 1))
        ph = np.arctan2(m.qx[:, 1], m.qx[:, 0])
        sig = np.zeros(n)
        for deg in range(9):
            for o in range(deg + 1):
                y = sph_harm_y(deg, o, th, ph)
                sig += rng.standard_normal() * y.real + (rng.standard_normal() * y.imag if o else 0)
        qh = (m.qw * sig).astype(np.float32)
        x = torch.from_numpy(xh).cuda()
        q = torch.from_numpy(qh).cuda()
        nrm = torch.from_numpy(nrmh).cuda()
        idx = torch.from_numpy(np.random.default_rng(1).choice(n, nt, replace=False)).cuda()
        kernels = [("single_layer", F.G, None), ("double_layer", F.GNJ, None)]
        metric = "BEM single+double layer matvec Mpoints/s, 245,760 Gauss points on an icosphere, p=10"
        workload = ("Laplace G and Gnj matvecs on the 3 Gauss points of an 81,920-triangle icosphere (245,760 points,"
                    " unit normals), density = degree<=8 harmonic field x weights, p=10; step = tree + lists + "
                    "one G and one Gnj evaluation")
        lmaxes = [a.lmax] if a.lmax else [4, 5]
        mk = lambda prec, lm: F.FMM(p=p, lmax=lm, precision=prec, convention="exact")  # noqa: E731
    for prec in ("mixed", "fp64"):
        best = None
        for lm in lmaxes:
            f = mk(prec, lm)

            def step():
                f.set_points(x, ni_normals=nrm, nj_normals=nrm)
                return [f.evaluate(q, neq=kk, gradient=False)[0] for _, kk, _ in kernels]

            ms, outs = timed(torch, step, max(2, a.steps // 2), 1)
            if best is None or ms < best[0]:
                if best is not None:
                    best[2].close()
                best = (ms, lm, f, outs)
            else:
                f.close()
        ms, lm, f, outs = best
        errs = {}
        for (name, kk, _), v in zip(kernels, outs):
            xs = x[idx]
            dv, _ = f.direct(xs, x, q, neq=kk, ni_normals=None if nrm is None else nrm[idx], nj_normals=nrm)
            errs[name] = float(torch.linalg.norm(v[idx] - dv) / torch.linalg.norm(dv))
        res[prec] = {"ms_per_step": ms, "value": n / (ms * 1e3), "lmax": lm, "rel_l2_err": errs}
        f.close()
    out = {"metric": metric, "value": res["mixed"]["value"], "unit": "Mpoints/s", "n_gpus": 1,
           "steps": max(2, a.steps // 2), "warmup": 1, "ms_per_step": res["mixed"]["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic", "config": {"workload": workload, "n": n, "p": p},
           "mixed": res["mixed"], "fp64": res["fp64"],
           "speedup_vs_fp64": res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]}
    print(json.dumps(out))


def bem_batch_workload(a, F, torch):
    """Batched BEM solves: 256 ellipsoids (semi-axes uniform in [0.5, 2], 20
    Gaussian bumps of amplitude <= 0.1 each), panel counts log-uniform from
    2,000 to 200,000, three unit charges inside each; the induced-charge
    equation of every surface solved by GMRES to rtol 1e-6 with p = 8, all
    surfaces in one tree (one FMM launch sequence per iteration).  Timed: the
    whole batched solve (tree, right-hand sides, GMRES, energies) once cold
    (first use of the order's operators in the process) and once warm, and
    the matvec alone."""
    from old_fmm_bem_amd import bem, workloads as W

    ns = 256 if a.n == 1_000_000 else max(2, a.n)
    surfaces, charges = W.ellipsoids(ns, 0)
    npan = sum(len(f) for _, f in surfaces)
    print(f"bem-batch: {ns} surfaces, {npan} panels generated", file=sys.stderr, flush=True)
    t_solves = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = bem.batched_solvation(surfaces, charges, p=8, precision="mixed", rtol=1e-6, extra_levels=5)
        torch.cuda.synchronize()
        t_solves.append(time.perf_counter() - t0)
        print(f"bem-batch: solved in {t_solves[-1]:.2f} s, {r['iterations']} iterations", file=sys.stderr,
              flush=True)
    t_solve = t_solves[-1]
    # the matvec alone on the same batch: one replay of the captured
    # evaluation plus the diagonal term
    scale = r["scale"]
    sc = [((np.asarray(v, np.float64) * s_).astype(np.float32), f) for (v, f), s_ in zip(surfaces, scale)]
    op = bem.BatchedInducedCharge(sc, p=8, precision="mixed", extra_levels=5)
    x = torch.ones(op.n, dtype=torch.float64, device=op.dev)
    ms_mv, _ = timed(torch, lambda: op.matvec(x), 10, 2)
    op.close()
    out = {"metric": "batched BEM solve: 256 surfaces, ~10M panels total, GMRES rtol 1e-6, p=8 (panels/s per matvec)",
           "value": npan / (ms_mv * 1e3), "unit": "Mpanels/s", "n_gpus": 1, "steps": 10, "warmup": 2,
           "ms_per_step": ms_mv, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{ns} ellipsoids with 20 bumps, {npan} panels (log-uniform 2k..200k per "
                                  "surface), induced-charge BEM, p=8, mixed precision; step = one batched matvec "
                                  "(replay of the captured evaluation)",
                      "surfaces": ns, "panels": npan},
           "solve_s": t_solve, "solve_first_s": t_solves[0], "solve_breakdown_s": r["times_s"],
           "gmres_iterations": r["iterations"], "max_residual": float(r["residuals"].max()),
           "energy_range_kcal": [float(r["energies"].min()), float(r["energies"].max())]}
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the multi-GPU path is "
              "design-only in this build", file=sys.stderr)
        sys.exit(2)
    import torch

    import old_fmm_bem_amd as F

    assert torch.cuda.is_available(), "needs a HIP device"
    if a.workload == "bem-batch":
        bem_batch_workload(a, F, torch)
        return
    if a.workload != "cube":
        sphere_workload(a, F, torch)
        return
    rng = np.random.default_rng(0)
    xh = rng.random((a.n, 3), dtype=np.float32)
    qh = (2 * rng.random(a.n, dtype=np.float32) - 1).astype(np.float32)
    x = torch.from_numpy(xh).cuda()
    q = torch.from_numpy(qh).cuda()
    if a.copy_calib:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    idx = np.random.default_rng(1).choice(a.n, a.targets, replace=False)

    lmaxes = [a.lmax] if a.lmax else [4, 5]
    best = None
    for lm in lmaxes:
        r = run_engine(F, torch, x, q, a.p, lm, "mixed", a.steps, a.warmup)
        if best is None or r[3] < best[1][3]:
            best = (lm, r)
    lmax, (f, v, g, ms, ms_eval, phases, st) = best

    # error vs FP64 direct sum on sampled targets (GPU direct kernel, FP64)
    xs = torch.from_numpy(xh[idx]).cuda()
    dv, dg = f.direct(xs, x, q)
    vi = v[torch.from_numpy(idx).cuda()]
    gi = g[torch.from_numpy(idx).cuda()]
    err_v = float(torch.linalg.norm(vi - dv) / torch.linalg.norm(dv))
    err_g = float(torch.linalg.norm(gi - dg) / torch.linalg.norm(dg))

    fp64 = None
    if not a.no_fp64:
        r64 = run_engine(F, torch, x, q, a.p, lmax, "fp64", max(3, a.steps // 2), 1)
        f64, v64, g64, ms64, ms64_eval, ph64, st64 = r64
        e64 = float(torch.linalg.norm(v64[torch.from_numpy(idx).cuda()] - dv) / torch.linalg.norm(dv))
        fp64 = {"ms_per_step": ms64, "value": a.n / (ms64 * 1e3), "ms_eval_only": ms64_eval, "err_v": e64,
                "phases_ms": {k: round(v_, 4) for k, v_ in ph64.items()}}
        # the ratio's denominator against its own roof: the FP64 leaf M2L
        # (vector FMA) on the FP64 peak
        fl64 = 2.0 * a.p ** 4 * st64["m2l_pairs_leaf"]
        ach64 = fl64 / (ph64["ms_m2l_leaf"] * 1e-3) / 1e12
        fp64["roofline"] = {"bound": "valu", "achieved": ach64, "peak": PEAK["fp64_vector"], "unit": "TFLOP/s",
                            "frac": ach64 / PEAK["fp64_vector"], "kernel": f"k_m2l_vec<{a.p}, double>, leaf level",
                            "ms_per_launch": ph64["ms_m2l_leaf"]}
        f64.close()

    # dominant kernel and its roofline position.  The M2L launch (one launch
    # over every level, or the leaf level's) carries most of the matvec's
    # flops: algorithmic work 2 p^4 flop per M2L pair (dense p^2 x p^2 real
    # operator), pairs counted by the kernel itself; duration from HIP events
    # recorded on the engine's stream around that launch alone.  The mixed
    # engine's M2L runs on the FP16 matrix cores (error-compensated split:
    # three FP16 products per FP32-accurate product) or on FP32 MFMA.
    p2 = a.p * a.p
    split = f.tuning[lmax][1] >= 3
    launch_pairs = st["m2l_pairs"] if st["m2l_all_levels"] else st["m2l_pairs_leaf"]
    peak_m2l = PEAK["fp16_mfma"] if split else PEAK["fp32_mfma"]
    if phases["ms_m2l"] >= phases["ms_p2p"]:
        flops = 2.0 * p2 * p2 * launch_pairs
        achieved = flops / (phases["ms_m2l_leaf"] * 1e-3) / 1e12
        tr = traffic_from_profile(a.n, a.p, lmax, st["m2l_all_levels"], split)
        if split and f.tuning[lmax][1] == 5:
            # executed MFMA work per pair: 32 MT rows x 16 KH k-positions per target
            mt, ks = (p2 + 31) // 32, (p2 + 1) // 2
            kh = (3 * (ks + (ks & 1)) + 7) // 8
            executed = 2.0 * (32 * mt) * (16 * kh)
            kern = f"k_m2l_w (error-compensated FP16 split, v_mfma_f32_32x32x16_f16), levels 2-{lmax} in one launch"
        elif split:
            # executed MFMA work per pair: 16 MT rows x 32 KH k-positions per target
            mt, ks = (p2 + 15) // 16, (p2 + 3) // 4
            kh = (3 * (ks + (ks & 1)) + 7) // 8
            executed = 2.0 * (16 * mt) * (32 * kh)
            kern = f"k_m2l_hd (error-compensated FP16 split, v_mfma_f32_16x16x32_f16), levels 2-{lmax} in one launch"
        else:
            executed = None
            kern = f"k_m2l (FP32 MFMA 16x16x4), levels 2-{lmax} in one launch"
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak_m2l, "unit": "TFLOP/s",
                "frac": achieved / peak_m2l, "traffic": tr, "kernel": kern,
                "ms_per_launch": phases["ms_m2l_leaf"],
                "work": f"2*p^4 = {2 * p2 * p2} flop x {launch_pairs} M2L pairs per launch (algorithmic)",
                "executed_flop_per_algorithmic": (executed / (2 * p2 * p2)) if executed else None,
                "frac_vs_fp32_mfma_peak": achieved / PEAK["fp32_mfma"]}
    else:
        flops = 20.0 * st["p2p_pairs"]
        achieved = flops / (phases["ms_p2p"] * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": achieved, "peak": PEAK["fp32_vector"], "unit": "TFLOP/s",
                "frac": achieved / PEAK["fp32_vector"], "traffic": None, "kernel": "k_p2p",
                "ms_per_launch": phases["ms_p2p"],
                "work": f"20 flop x {st['p2p_pairs']} near pairs (FP32 vector ALU)"}

    # every phase against its own bound: t_ideal = max(flop / peak, bytes /
    # HBM), fraction = t_ideal / t_measured (serialized per-phase device times)
    nbi = {lv: f.nboxes("i", lv) for lv in range(2, lmax + 1)}
    nbj = {lv: f.nboxes("j", lv) for lv in range(2, lmax + 1)}
    ks = (p2 + 3) // 4
    vs = 4 * ((ks + 3) // 4 * 4)
    passes = (3 * lmax + 7) // 8
    trans = 2.0 * p2 * p2
    work = {
        # keys (12 B in, 4 B out), per radix pass 4 B (counts) + 8 B in + 8 B
        # out, coordinate gather 16 B in + 12 B out, per point
        "ms_tree": (0.0, a.n * (16.0 + 20.0 * passes + 28.0), None),
        "ms_p2p": (20.0 * st["p2p_pairs"], 0.0, PEAK["fp32_vector"]),
        "ms_p2m": (0.0, a.n * 16.0 + nbj[lmax] * vs * 4.0, None),
        "ms_m2m": (trans * sum(nbj[lv] for lv in range(3, lmax + 1)), 0.0, PEAK["fp32_mfma"]),
        "ms_m2l": (trans * st["m2l_pairs"], 0.0, peak_m2l),  # M2L (+ L2L below)
        "ms_l2p": (0.0, a.n * 44.0 + nbi[lmax] * p2 * 4.0, None),
    }
    per_phase, t_ideal_sum = {}, 0.0
    for k_, (fl, by, pk) in work.items():
        t_ideal = max(fl / (pk * 1e12) if pk else 0.0, by / (PEAK["hbm"] * 1e9)) * 1e3
        if k_ == "ms_m2l":
            t_ideal += trans * sum(nbi[lv] for lv in range(3, lmax + 1)) / (PEAK["fp32_mfma"] * 1e12) * 1e3
        per_phase[k_[3:]] = {"ms": round(phases[k_], 4), "ms_ideal": round(t_ideal, 4),
                             "frac": round(t_ideal / phases[k_], 4) if phases[k_] > 0 else None}
        t_ideal_sum += t_ideal
    roofline_phases = {"phases": per_phase, "ms_ideal_step": round(t_ideal_sum, 4),
                       "frac_step": t_ideal_sum / ms,
                       "model": "t_ideal = max(flop/peak, bytes/8 TB/s) per phase; P2P on the FP32 vector peak, "
                                "M2M/L2L on FP32 MFMA, M2L on the peak of the matrix unit it runs on; bytes of "
                                "tree, P2M, L2P are compulsory HBM traffic"}

    cpu = None
    if not a.no_cpu:
        from oracle import oracle as O

        def cpu_model():
            try:
                for line in open("/proc/cpuinfo"):
                    if line.startswith("model name"):
                        return line.split(":", 1)[1].strip()
            except OSError:
                pass
            return "unknown"

        def oracle_run(threads):
            t0 = time.perf_counter()
            o = O.OracleFMM(xh, xh, mp=a.p, lmax=lmax, conv=O.EXACT, box=(0, 0, 0, 1), nthreads=threads)
            o.run(qh)
            tc = time.perf_counter() - t0
            o.close()
            return tc

        # the host's share for one GPU (16 threads), and one thread (the
        # reference's CPU path has no threading)
        cores = min(16, os.cpu_count() or 1)
        tc = oracle_run(cores)
        t1 = oracle_run(1)
        cpu = {"value": a.n / tc / 1e6, "unit": "Mpoints/s", "cores": cores, "kind": "port",
               "sample": f"one full matvec of the same workload (N={a.n}, p={a.p}, lmax={lmax}), FP64, "
                         f"{tc:.2f} s wall on {cores} threads",
               "single_thread": {"value": a.n / t1 / 1e6, "unit": "Mpoints/s", "cores": 1, "s": t1},
               "nproc": os.cpu_count(), "cpu_model": cpu_model()}

    out = {
        "metric": "FMM matvec Mpoints/s, N=1M p=10, err<=1e-5 vs FP64 direct: mixed vs FP64",
        "value": a.n / (ms * 1e3),
        "unit": "Mpoints/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"Laplace G FMM matvec, N={a.n} uniform unit cube, q~U[-1,1], p={a.p}, "
                               f"lmax={lmax}, self excluded, coordinates resident as x, y, z device columns; "
                               "step = tree + lists + all phases (the translation "
                               "operators depend only on p, lmax and the root cube: computed at the first "
                               "set_points of the handle, before timing, and reused)",
                   "n": a.n, "p": a.p, "lmax": lmax, "precision": "mixed", "convention": "exact"},
        "err_v": err_v,
        "err_g": err_g,
        # the metric's bound is on the potential; the gradient carries the
        # p = 10 truncation at about 2e-5 and is held to 3e-5
        "valid": err_v <= 1e-5 and err_g <= 3e-5,
        "ms_eval_only": ms_eval,
        "value_eval_only": a.n / (ms_eval * 1e3),
        "phases_ms": {k: round(v_, 4) for k, v_ in phases.items()},
        "p2p_pairs": st["p2p_pairs"],
        "m2l_pairs": st["m2l_pairs"],
        "fp64": fp64,
        "speedup_vs_fp64": (fp64["ms_per_step"] / ms) if fp64 else None,
        "roofline": roof,
        "roofline_phases": roofline_phases,
        "cpu_baseline": cpu,
        "m2l_tuning": {str(k): v for k, v in f.tuning.items()},
    }
    f.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
