#!/usr/bin/env python3
"""pGM energy+force evaluations per second at 96,000 atoms on one MI355X (mixed precision vs the FP64 path).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload (synthetic, seeded): 32,000 rigid water-like molecules (96,000 atoms) in a 6.84 nm cubic box (3x
liquid density), Gaussian charges O -0.8 / H +0.4 e, covalent dipoles 0.01 e nm along each O-H, alpha
1.0e-3 / 0.4e-3 nm^3, radii 0.075 / 0.055 nm; cutoff 0.9 nm, beta 3.5/nm, PME 72^3 order 6, CG tau 1e-4,
peek 0.65.  The initial guess is the cubic (mu4) extrapolation of 4 seeded prior dipole sets: FP64 solutions
at tau 1e-10 of 4 earlier geometries 0.002 nm apart along a seeded velocity field.

A step is one complete evaluation (positions already in HBM -> energy terms + forces in HBM): list check,
sorted gather, PME binning, pair setup, PME of the permanent multipoles, CG from the predicted guess, final
pair pass + PME energy/forces, covalent chain rule, reductions.  The predictor history is frozen so every
step solves the same problem from the same guess.  The pair list is built once (positions do not move).

Prints one JSON line: value = mixed-precision evals/s; the FP64 path of the same engine, the force errors,
the per-kernel-group timing, the roofline of the dominant kernel group and the CPU baseline ride along, and
(unless --no-large) the 999,000-atom box: both paths' evals/s, device memory and the mixed path's error against
the FP64 oracle's evaluation of that box.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pgm-md_amd"))

METRIC = "pGM E+F evals/s @96k atoms, mixed vs FP64 GPU at ≤5e-5 rel-RMS force error"
PEAK_TFLOPS = {"mixed": 157.3, "double": 78.6}  # MI355X published FP32 / FP64 rates
HBM_PEAK_GBS = 8000.0
CAL_COPY_BYTES = 512 << 20  # the calibration copy of the profiler pass
TRAFFIC_FILE = os.path.join("profiles", "water96k_traffic.md")

# Algorithmic HBM bytes per launch of the pair passes: per entry of the full list within the cutoff (E = 2P
# entries) and per atom (N), s = 4 (mixed) or 8 (FP64) bytes per real.  Entry: the compacted index (4) and
# the stored kernels (G1, G2: 2s; FP64 also G0, G3: 2s more) read or written once; compaction also reads the
# candidate indices (4 per candidate).  Atom: its own and (once, as cache-resident partners) its neighbours'
# position (4s), dipole (4s), radius parameter (s), list offsets and counts, and the FP64 outputs.
PAIR_BYTES = {
    # phase: (per entry, per atom, per candidate) for s = 4 / s = 8
    "pair_compact": {"mixed": (4, 28, 4), "double": (4, 44, 4)},
    "pair_setup": {"mixed": (12, 112, 0), "double": (36, 164, 0)},
    "pair_field": {"mixed": (12, 68, 0), "double": (20, 100, 0)},
    "pair_final": {"mixed": (12, 176, 0), "double": (36, 212, 0)},
}
# compaction fused into the setup pass: candidates in, compacted entries and kernels out, no index round trip
PAIR_BYTES_FUSED_SETUP = {"mixed": (12, 112, 4), "double": (36, 164, 4)}
KERNEL_OF_PHASE = {"pair_compact": "kCompact", "pair_setup": "kPairSetup", "pair_field": "kPairField",
                   "pair_final": "kPairFinal"}  # fmt: skip

# Algorithmic flops (add/mul/sub 1, FMA 2; exp/erfc/sqrt/rsqrt/rint not counted) per half pair within the
# cutoff (P), per atom (N) and per grid: geometry 26, G1+G2 32, G0..G3 44, dual field body 76, dipole
# matrix-vector body 36, final body 108; splines 300, spread 1750, field gather 1300, final gather 3950.
# Per launch of a pair pass (each pair visited from both atoms counts once): setup (compaction fused or not)
# 26 + 32 + 76 = 134 P, field 26 + 36 = 62 P, final 26 + 44 + 108 = 178 P; compaction alone is the geometry of
# the setup's count.
PAIR_FLOPS_PER_P = {"pair_setup": 134, "pair_field": 62, "pair_final": 178, "pair_compact": 0}
VALU_NONPACKED_TFLOPS = 78.6  # FP32 without packed instructions (this build disables them) = the FP64 rate
# Compulsory HBM bytes per evaluation: N (196 + 280 n) + (n + 3)(20 K + 32 Kc) (positions, parameters, forces,
# dipoles, 40 B of dipole/field traffic per field pass, 10 FP64 3-vector streams per CG iteration, grids).


def fft_flops(K):
    Kt = K[0] * K[1] * K[2]
    Kc = K[0] * K[1] * (K[2] // 2 + 1)
    return 5.0 * Kt * math.log2(Kt) + 8.0 * Kc


def flops_per_eval(P, N, K, n):
    """F(n) = P (312 + 62 n) + N (12,100 + 3,050 n) + (n + 3)(5 K log2 K + 8 Kc)."""
    return P * (312 + 62 * n) + N * (12100 + 3050 * n) + (n + 3) * fft_flops(K)


def compulsory_bytes(N, K, n):
    Kt = K[0] * K[1] * K[2]
    Kc = K[0] * K[1] * (K[2] // 2 + 1)
    return N * (196 + 280 * n) + (n + 3) * (20 * Kt + 32 * Kc)


def adjoint_flops(P, N, K, n, m):
    """F_adj = F(n) + m (62 P + 3,050 N + 5 K log2 K + 8 Kc) + 2 (178 P + 3,950 N)."""
    return flops_per_eval(P, N, K, n) + m * (62 * P + 3050 * N + fft_flops(K)) + 2 * (178 * P + 3950 * N)


def prior_dipole_sets(sy, E, nset=4, step=0.002, seed=22):
    """The FP64 (tau 1e-10, no predictor) dipoles of nset earlier geometries, most recent first."""
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, 1.0, sy.pos.shape)
    v /= np.sqrt((v * v).sum(1)).mean()
    eng = E.Engine.for_system(sy, precision="double", tol=1e-10, predictor=0, peek=0.0, max_iter=500)
    sets = []
    for k in range(nset):
        eng.compute(sy.pos - (k + 1) * step * v, sy.box, forces=False)
        sets.append(eng.induced_dipoles())
    return np.stack(sets)


def phase_stats(tab, steps):
    return {k: {"ms_per_eval": v[0] / steps, "launch_avg_us": 1000.0 * v[0] / max(v[1], 1), "launches": v[1] / steps}
            for k, v in tab.items()}  # fmt: skip


def phase_table(eng, pos_d, f_d, box, stream, torch, reps):
    """Every kernel group bracketed with HIP events on the engine stream (each bracket idles the GPU for a few
    microseconds, so this pass is separate from the timed steps)."""
    eng.set_timing(True)
    for _ in range(reps):
        eng.compute_device_f64(pos_d.data_ptr(), box, True, f_d.data_ptr(), stream)
    torch.cuda.synchronize()
    tab = eng.timing()
    eng.set_timing(False)
    return phase_stats(tab, reps)


def dominant_pair_phase(phases):
    live = [ph for ph in PAIR_BYTES if phases.get(ph, {}).get("launches", 0) > 0]
    return max(live, key=lambda k: phases[k]["ms_per_eval"])


def timed_steps(torch, eng, pos_d, f_d, box, steps, warmup, stream, phase):
    """Wall time of `steps` evaluations; during them the engine brackets the launches of `phase` (only) with
    HIP events on its own stream, the stream those kernels run on: their live average duration."""
    for _ in range(warmup):
        eng.compute_device_f64(pos_d.data_ptr(), box, True, f_d.data_ptr(), stream)
    torch.cuda.synchronize()
    eng.set_timing(True, phases=[phase])
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.compute_device_f64(pos_d.data_ptr(), box, True, f_d.data_ptr(), stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tab = eng.timing()
    eng.set_timing(False)
    return dt, phase_stats(tab, steps)[phase]


def traffic_table(path=TRAFFIC_FILE):
    """HBM bytes per launch by kernel from the committed PMC summary (rows `| kernel | ... | bytes |`)."""
    out = {}
    try:
        with open(os.path.join(ROOT, path)) as fh:
            for line in fh:
                cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
                if len(cells) >= 2 and cells[0].startswith("k") and cells[-1].replace(",", "").isdigit():
                    out[cells[0]] = int(cells[-1].replace(",", ""))
    except OSError:
        pass
    return out


def roofline(dom, live, P, C, N, precision, lanes, fused, setup_lanes=None):
    """The pair pass `dom` (the one with the most time per evaluation) against the VALU rate, the bound of the
    path: its algorithmic flops per launch (PAIR_FLOPS_PER_P) over its live launch average `live` in the timed
    steps, against the MI355X FP32 (packed, 157.3 TFLOP/s) or FP64 (78.6) peak.  `traffic` is the HBM bytes per
    launch measured by the counters; `stored_list_bytes` the part of them that is this implementation's stored
    list and kernels (per entry of the full list and per candidate), which is no algorithmic traffic."""
    E = 2 * P
    table = dict(PAIR_BYTES, pair_setup=PAIR_BYTES_FUSED_SETUP) if fused else PAIR_BYTES
    kernels = dict(KERNEL_OF_PHASE, pair_setup="kCompactSetup") if fused else KERNEL_OF_PHASE
    per_launch = {ph: b[precision][0] * E + b[precision][1] * N + b[precision][2] * C for ph, b in table.items()}
    stored = {ph: b[precision][0] * E + b[precision][2] * C for ph, b in table.items()}
    t = live["launch_avg_us"] * 1e-6
    flops = PAIR_FLOPS_PER_P[dom] * P
    achieved = flops / t / 1e12
    peak = PEAK_TFLOPS[precision]
    real = "float" if precision == "mixed" else "double"
    if dom in ("pair_setup", "pair_compact") and setup_lanes:
        lanes = setup_lanes
    prefix = f"{kernels[dom]}<{real}, {lanes},"
    traffic = next((v for k, v in traffic_table().items() if k.startswith(prefix)), None)
    return {
        "bound": "valu", "kernel": f"{kernels[dom]}<{real}, {lanes}>", "phase": dom,
        "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
        "frac_of_nonpacked_fp32": round(achieved / VALU_NONPACKED_TFLOPS, 4) if precision == "mixed" else None,
        "flops_per_launch": flops, "traffic": traffic, "stored_list_bytes": stored[dom],
        "hbm_achieved_gbs": round(per_launch[dom] / t / 1e9, 1), "launch_avg_us": round(live["launch_avg_us"], 2),
        "launches_per_eval": live["launches"], "traffic_source": TRAFFIC_FILE if traffic else None,
    }  # fmt: skip


def eval_roofline(P, N, K, n, evals_per_s, precision):
    """The whole evaluation against the VALU peak (F(n) flops per evaluation) and its compulsory HBM bytes."""
    f = flops_per_eval(P, N, K, n)
    peak = PEAK_TFLOPS[precision]
    b = compulsory_bytes(N, K, n)
    return {
        "flops_per_eval": f, "achieved_tflops": round(f * evals_per_s / 1e12, 3), "peak_tflops": peak,
        "frac": round(f * evals_per_s / 1e12 / peak, 4), "compulsory_bytes_per_eval": b,
        "compulsory_hbm_frac": round(b * evals_per_s / 1e9 / HBM_PEAK_GBS, 4),
    }  # fmt: skip


def gvdw_cost(torch, E, sy, sets, pos_d, f_d, stream, steps, warmup, log):
    """The same workload with GVDW on (pgm_set_gvdw; Gaussian repulsion and dispersion on O, dispersion on H,
    the intramolecular pairs skipped, rvdw = the cutoff): evals/s, the final pair pass (live HIP events), the GVDW
    energy and the mixed-vs-FP64 force difference."""
    n = sy.n
    o = np.arange(0, n, 3)
    sa, sc = np.zeros(n), np.zeros(n)
    sa[o] = math.sqrt(422.0 * 4.184)
    sc[o] = math.sqrt(594.825035 * 4.184e-6)
    sc[o + 1] = sc[o + 2] = 0.3 * sc[o]
    b = np.full(n, 0.9453)
    si, sj = np.concatenate([o, o, o + 1]), np.concatenate([o + 1, o + 2, o + 2])
    out, forces = {}, {}
    for prec in ("mixed", "double"):
        eng = E.Engine.for_system(sy, precision=prec)
        eng.set_history(sets)
        eng.freeze_history(True)
        eng.set_gvdw(0, sy.settings["cutoff"], sa, sc, b, si, sj)
        t, forces[prec], _ = eng.compute(sy.pos, sy.box)
        (lanes, fused), _ = eng.autotune(3)
        dt, live = timed_steps(torch, eng, pos_d, f_d, sy.box, steps, warmup, stream, "pair_final")
        out[prec] = {"evals_per_s": round(steps / dt, 3), "pair_final_ms": round(live["ms_per_eval"], 4),
                     "gvdw_kj_mol": t["gvdw"], "lanes": lanes, "fused_setup": fused}  # fmt: skip
        log(f"gvdw {prec}: {steps / dt:.2f} evals/s")
    out["force_relrms_mixed_vs_fp64"] = E.rel_rms(forces["mixed"], forces["double"])
    out["energy_rel_mixed_vs_fp64"] = out["mixed"]["gvdw_kj_mol"] / out["double"]["gvdw_kj_mol"] - 1
    return out


def large_box(torch, E, S, steps, warmup, log):
    """The 999,000-atom box (333,000 waters, 21.54 nm, 216^3 PME, tau 1e-5) timed like the main workload (mu4
    guess from 4 prior sets, history frozen) in both precisions, with the mixed path's error against the FP64
    oracle's evaluation of the same box (tests/golden/water999k_oracle.npz, tau 1e-10: its 3,000-atom sample)."""
    sy = S.config4()
    sets = prior_dipole_sets(sy, E)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    pos_d = torch.from_numpy(sy.pos.copy()).to(dev)
    f_d = torch.zeros((sy.n, 3), dtype=torch.float64, device=dev)
    g = np.load(os.path.join(ROOT, "tests", "golden", "water999k_oracle.npz"))
    out = {"workload": "water999k: 333,000 rigid waters (999,000 atoms), 21.54 nm cubic box (liquid density), rc "
                       "0.9 nm, PME 216^3 order 6, tau 1e-5, mu4 guess from 4 seeded prior dipole sets, peek 0.65",
           "atoms": sy.n}  # fmt: skip
    for prec in ("mixed", "double"):
        free0 = torch.cuda.mem_get_info()[0]
        eng = E.Engine.for_system(sy, precision=prec)
        eng.set_history(sets)
        eng.freeze_history(True)
        t, f, _ = eng.compute(sy.pos, sy.box)
        used = (free0 - torch.cuda.mem_get_info()[0]) / 2**30
        (lanes, fused), _ = eng.autotune(3)
        beside = eng.pme_overlap()

        def timed():
            for _ in range(warmup):
                eng.compute_device_f64(pos_d.data_ptr(), sy.box, True, f_d.data_ptr(), stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                eng.compute_device_f64(pos_d.data_ptr(), sy.box, True, f_d.data_ptr(), stream)
            torch.cuda.synchronize()
            return time.perf_counter() - t0

        dt = timed()
        eng.set_pme_overlap(not beside)   # the autotuner's other choice, for comparison
        dt_other = timed()
        eng.set_pme_overlap(beside)
        fs = g["forces_sample"]
        err = float(np.sqrt(((f[g["sample"]] - fs) ** 2).sum() / (fs**2).sum()))
        out[prec] = {"evals_per_s": round(steps / dt, 3), "cg_iterations": eng.iterations(), "device_gib": round(used, 2),
                     "lanes": lanes, "setup_lanes": eng.setup_lanes(), "fused_setup": fused,
                     "pme_beside": beside,
                     "cg_product_ms_pme_after_beside": [round(v, 4) for v in eng.pme_overlap_times()],
                     "evals_per_s_other_pme_choice": round(steps / dt_other, 3),
                     "force_relrms_vs_oracle_tau1e-10": err,
                     "energy_rel_vs_oracle_tau1e-10": E.total_energy(t) / float(g["terms"].sum()) - 1}  # fmt: skip
        log(f"999k {prec}: {steps / dt:.2f} evals/s, {used:.1f} GiB, force error {err:.2e}")
        del eng
    out["mixed_over_fp64"] = round(out["mixed"]["evals_per_s"] / out["double"]["evals_per_s"], 3)
    return out


def md_like(torch, E, sy, steps, log, precision="mixed", step_nm=0.002, seed=31, tune=False):
    """MD-like steps of the 96k workload through the asynchronous device boundary (pgm_launch_device +
    pgm_finish_device, float4 positions, fixed-point forces, as an OpenMM caller): the atoms move every step along
    a seeded velocity field (step_nm per step for the mean atom), every step is a new step (the mu4 history
    starts from the FP64 solutions of the 4 steps before along the same field and then follows the trajectory),
    the pair list is rebuilt when the displacement check asks.
    Reports evals/s with the rebuilds in the timed region, the builds, repeats (valid = 0) and the mean CG
    iterations, and what one list build costs (the list phase of a separate pass with every rebuild bracketed).
    tune: autotune the engine after step 0 instead of taking the main workload's choice."""
    n = sy.n
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, 1.0, sy.pos.shape)
    v /= np.sqrt((v * v).sum(1)).mean()
    dev = torch.device("cuda:0")
    posq = torch.zeros((steps + 1, n, 4), dtype=torch.float32, device=dev)
    for k in range(steps + 1):
        posq[k, :, :3] = torch.from_numpy((sy.pos + k * step_nm * v).astype(np.float32)).to(dev)
    sets = prior_dipole_sets(sy, E, step=step_nm, seed=seed)   # the same field: its 4 steps before
    stride, scale = n, 2.0**32
    fixed = torch.zeros(3 * stride, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    out = {}
    for timed in (False, True):
        eng = E.Engine.for_system(sy, precision=precision)
        eng.set_history(sets)
        if not tune:
            apply_variant(eng, precision)
        if timed:
            eng.set_timing(True, phases=["list"])
        its, repeats, stale_repeats = [], 0, 0
        builds0 = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps + 1):
            if k == 1:   # step 0 builds the list and takes the first (host-checked) solve: outside the timed region
                if tune:
                    eng.autotune(3)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                builds0 = eng.list_builds()
            while True:
                fixed.zero_()
                b0 = eng.list_builds()
                eng.launch_device(posq[k].data_ptr(), sy.box, True, fixed.data_ptr(), stride, scale, stream)
                _, valid = eng.finish_device()
                if valid:
                    break
                repeats += 1
                stale_repeats += eng.list_builds() > b0   # else the solve needed more iterations than enqueued
            its.append(eng.iterations())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if timed:
            tab = eng.timing()
            ms, cnt = tab["list"]
            out["list_build_ms"] = round(ms / max(cnt, 1), 3)
        else:
            builds = eng.list_builds() - builds0
            out.update({"evals_per_s": round(steps / dt, 3), "steps": steps, "step_nm": step_nm,
                        "list_builds": builds, "repeated_launches": repeats,
                        "repeats_stale_list": stale_repeats, "repeats_short_solve": repeats - stale_repeats,
                        "late_list_checks": eng.late_list_checks(), "mean_cg_iterations": round(float(np.mean(its[1:])), 3)})  # fmt: skip
            log(f"md-like {precision}: {steps / dt:.2f} evals/s, {builds} list builds in {steps} steps, {repeats} repeats")
        del eng
    return out


def liquid98k(torch, E, S, steps, log):
    """The reference's closest published configuration: 32,768 waters (98,304 atoms) at liquid density, a 9.945 nm
    box, PME 112^3 order 6, rc 0.9 nm, tau 1e-4, mu4 and peek 0.65 (docs/openmm.md:177: 24.0 ns/day of MD at
    1 fs with the reference's CUDA engine under OpenMM on one RTX PRO 6000, about 278 steps/s, integration,
    constraints and LJ included).  Here: MD-like steps of the electrostatics alone (md_like, autotuned), mixed."""
    L = 21.54 * (32768 / 333000) ** (1.0 / 3.0)
    sy = S.water_box(32768, L, 32, 5, 0.03, "water98k-liquid", S.base_settings(grid=(112, 112, 112), tol=1e-4))
    out = md_like(torch, E, sy, steps, log, tune=True)
    return {"workload": f"32,768 rigid waters (98,304 atoms), {L:.3f} nm cubic box (liquid density), rc 0.9 nm, "
                        "PME 112^3 order 6, tau 1e-4, mu4 (history along the trajectory), peek 0.65; MD-like steps "
                        "through launch/finish, mixed precision",
            "mixed": out,
            "reference_published": {"steps_per_s": 278, "ns_per_day_at_1fs": 24.0, "gpu": "NVIDIA RTX PRO 6000 Blackwell",
                                    "what": "MD steps of OpenMM with the reference's CUDA pGM engine (integration, "
                                            "constraints and LJ included)", "source": "docs/openmm.md:177"}}  # fmt: skip


VARIANTS = {}   # the autotuner's choice per precision: (lanes, fused setup, on-the-fly field, PME beside, setup lanes)


def apply_variant(eng, precision):
    lanes, fused, otf, beside, setup_lanes = VARIANTS[precision]
    eng.set_lanes(lanes, fused)
    eng.set_setup_lanes(setup_lanes)
    eng.set_pme_overlap(beside)
    if precision == "mixed":
        eng.set_field_otf(otf)


def mu3_variant(torch, E, sy, sets, pos_d, f_d, stream, steps, warmup, log):
    """The main workload with the quadratic mu3 guess from 3 prior sets (predictor 1), both precisions."""
    out = {}
    for prec in ("mixed", "double"):
        eng = E.Engine.for_system(sy, precision=prec, predictor=1)
        eng.set_history(sets[:3])
        eng.freeze_history(True)
        apply_variant(eng, prec)
        eng.compute(sy.pos, sy.box)
        for _ in range(warmup):
            eng.compute_device_f64(pos_d.data_ptr(), sy.box, True, f_d.data_ptr(), stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.compute_device_f64(pos_d.data_ptr(), sy.box, True, f_d.data_ptr(), stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[prec] = {"evals_per_s": round(steps / dt, 3), "cg_iterations": eng.iterations()}
        log(f"mu3 {prec}: {steps / dt:.2f} evals/s, {eng.iterations()} CG iterations")
        del eng
    out["mixed_over_fp64"] = round(out["mixed"]["evals_per_s"] / out["double"]["evals_per_s"], 3)
    return out


def snapshot_batch(S, nsnap=64):
    """The uneven snapshot batch as one engine's arrays: offsets, grids, boxes, per-atom parameters, covalent."""
    systems = [S.config3_snapshot(k) for k in range(nsnap)]
    off = np.concatenate([[0], np.cumsum([s.n for s in systems])]).astype(np.int32)
    cat = lambda key: np.concatenate([getattr(s, key) for s in systems])  # noqa: E731
    return {
        "systems": systems, "off": off, "pos": cat("pos"), "boxes": np.stack([s.box for s in systems]),
        "grids": np.array([s.settings["grid"] for s in systems], np.int32), "q": cat("q"), "radius": cat("radius"),
        "alpha": cat("alpha"), "cov_strength": cat("cov_strength"),
        "cov_atom": np.concatenate([s.cov_atom + off[i] for i, s in enumerate(systems)]),
        "cov_partner": np.concatenate([s.cov_partner + off[i] for i, s in enumerate(systems)]),
    }  # fmt: skip


def batch_flops(b, P, n, m):
    """Sum over the snapshots of F_adj (adjoint_flops), with the batch's total half pairs P split over
    the snapshots in proportion to their atoms (all snapshots have the same density)."""
    N = int(b["off"][-1])
    total = 0.0
    for s in b["systems"]:
        Ps = P * s.n / N
        total += adjoint_flops(Ps, s.n, s.settings["grid"], n, m)
    return total


def batch_engine(E, b, precision, tol):
    st = b["systems"][0].settings
    keys = ("cutoff", "beta", "order", "max_iter", "ke", "skin", "polarizable")
    eng = E.Engine(precision=precision, grid=tuple(st["grid"]), tol=tol, predictor=0, peek=0.0,
                   **{k: st[k] for k in keys})  # fmt: skip
    eng.set_atoms(b["q"], b["radius"], b["alpha"])
    eng.set_covalent(b["cov_atom"], b["cov_partner"], b["cov_strength"])
    eng.set_batch(b["off"], b["grids"])
    return eng


def gap_report(path, pause_ns=50_000_000):
    """Kernel trace (rocprofv3 --kernel-trace, *_kernel_trace.csv) of a --batch-only run, split at every pause
    longer than 50 ms: per segment, the union of the kernels' [start, end) intervals (both streams) is GPU-busy
    time, the rest of its span is launch gaps and host waits.  In a --batch-only run the segments that follow
    each precision's autotune are its timed steps (then one phase-bracketed step)."""
    import csv

    rows = []
    with open(path) as fh:
        for r in csv.DictReader(fh):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    rows.sort()
    segments, cur = [], [rows[0]]
    reach = rows[0][1]
    for a, b in rows[1:]:
        if a - reach > pause_ns:
            segments.append(cur)
            cur = []
        cur.append((a, b))
        reach = max(reach, b)
    segments.append(cur)
    out = []
    for seg in segments:
        busy, c0, c1 = 0, seg[0][0], seg[0][1]
        for a, b in seg[1:]:
            if a > c1:
                busy += c1 - c0
                c0, c1 = a, b
            else:
                c1 = max(c1, b)
        busy += c1 - c0
        span = max(b for _, b in seg) - seg[0][0]
        out.append({"dispatches": len(seg), "span_ms": round(span / 1e6, 3), "busy_ms": round(busy / 1e6, 3),
                    "idle_share": round(1.0 - busy / span, 4) if span > 0 else 0.0})  # fmt: skip
    return out


def adjoint_batch(torch, E, S, reps, log, pause=False):
    """Force-matching gradients over the 64-snapshot batch in one engine: per step a forward solve from the
    alpha E guess, the same-step re-evaluation and the adjoint solve, FP64 and mixed, tau 1e-8, the pair-pass
    lanes picked by the autotuner.  Unit: snapshot-gradients/s."""
    b = snapshot_batch(S)
    n = int(b["off"][-1])
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    pos_d = torch.from_numpy(b["pos"].copy()).to(dev)
    f_d = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    g_d = torch.from_numpy(np.random.default_rng(9).normal(0.0, 1e-3, (n, 3))).to(dev)
    grads = [torch.empty(m, dtype=torch.float64, device=dev) for m in (n, n, n, len(b["cov_atom"]))]
    # the parameters a fit would set at each step: per-type scalings of q, R, alpha and the covalent strengths
    # (two sets, alternated), through set_atoms + set_covalent (the pair list is kept: same atoms, same order)
    prng = np.random.default_rng(12)
    params = []
    for _ in range(2):
        f = 1.0 + 0.01 * prng.standard_normal(4)
        params.append((b["q"] * f[0], b["radius"] * f[1], b["alpha"] * f[2], b["cov_strength"] * f[3]))
    out, ref = {}, None
    for prec in ("double", "mixed"):
        eng = batch_engine(E, b, prec, 1e-8)
        _, f = eng.compute_batch(b["pos"], b["boxes"])
        P_batch = eng.num_pairs()
        (lanes, fused), tune = eng.autotune(3)
        fwd_it = eng.iterations()
        g = eng.backward_batch(b["pos"], b["boxes"], g_d.cpu().numpy(), None)
        if ref is None:
            ref = (f, g)
        adj_it = eng.adjoint_iterations()

        counter = [0]

        def step():
            # a fitting step: new parameters, so the forward solve starts from the alpha E guess (no history)
            q, r, al, cs = params[counter[0] % 2]
            counter[0] += 1
            eng.set_atoms(q, r, al)
            eng.set_covalent(b["cov_atom"], b["cov_partner"], cs)
            eng.reset_history()
            eng.compute_batch_device(pos_d.data_ptr(), b["boxes"], True, f_d.data_ptr(), stream)
            eng.backward_batch_device(pos_d.data_ptr(), b["boxes"], g_d.data_ptr(), None,
                                      *[t.data_ptr() for t in grads], stream)  # fmt: skip

        step()
        torch.cuda.synchronize()
        if pause:   # a gap a kernel trace can find: the timed steps follow it
            time.sleep(0.1)
        t0 = time.perf_counter()
        for _ in range(reps):
            step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        if pause:   # and after them, so the phase-bracketed step below is a segment of its own
            time.sleep(0.1)
        # one more step with every kernel group bracketed (in line, so not the timed figure): where a batch goes
        eng.set_timing(True)
        step()
        torch.cuda.synchronize()
        tab = eng.timing()
        eng.set_timing(False)
        phases = {k: round(v[0], 3) for k, v in tab.items() if v[1] > 0}
        gq = np.concatenate([g[k] for k in ("q", "radius", "alpha")])
        gq0 = np.concatenate([ref[1][k] for k in ("q", "radius", "alpha")])
        fl = batch_flops(b, P_batch, fwd_it, adj_it)
        peak = PEAK_TFLOPS[prec]
        out[prec] = {
            "snapshot_gradients_per_s": round(64 / dt, 2), "ms_per_batch": round(1000 * dt, 3),
            "roofline": {"bound": "valu", "flops_per_batch": fl, "achieved": round(fl / dt / 1e12, 3), "peak": peak,
                         "unit": "TFLOP/s", "frac": round(fl / dt / 1e12 / peak, 4)},
            "list_builds": eng.list_builds(),
            "cg_iterations_forward": fwd_it, "cg_iterations_adjoint": adj_it, "lanes": lanes,
            "setup_lanes": eng.setup_lanes(), "fused_setup": fused,
            "pme_beside": eng.pme_overlap(),
            "cg_product_ms_pme_after_beside": [round(v, 4) for v in eng.pme_overlap_times()],
            "phases_ms_per_batch_in_line": phases,
            "autotune_ms": {f"{k[0]}{'f' if k[1] else 's'}": round(v, 4) for k, v in tune.items()},
            "force_relrms_vs_fp64": E.rel_rms(f, ref[0]),
            "gradient_maxrel_vs_fp64": float(np.abs(gq - gq0).max() / np.abs(gq0).max()),
        }  # fmt: skip
        log(f"adjoint batch {prec}: {64 / dt:.1f} snapshot-gradients/s, lanes {lanes}")
    return {"snapshots": 64, "atoms": n, "tol": 1e-8, **out}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            return next(ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name"))
    except (OSError, StopIteration):
        return "unknown"


def sub_box(sy, S):
    """A 1/8 sample of the workload: the waters whose oxygen lies in [0, L/2)^3, as a periodic box of edge L/2
    at the same density and settings (PME grid at the same spacing)."""
    L = sy.box[0, 0]
    o = np.arange(0, sy.n, 3)
    keep = o[np.all((sy.pos[o] % L) < 0.5 * L, axis=1)]
    idx = np.sort(np.concatenate([keep, keep + 1, keep + 2]))
    remap = -np.ones(sy.n, np.int64)
    remap[idx] = np.arange(len(idx))
    cov = np.isin(sy.cov_atom, idx)
    st = dict(sy.settings, grid=tuple(k // 2 for k in sy.settings["grid"]))
    return S.System(f"{sy.name}-eighth", sy.pos[idx] % L, sy.box * 0.5, sy.q[idx], sy.radius[idx], sy.alpha[idx],
                    remap[sy.cov_atom[cov]], remap[sy.cov_partner[cov]], sy.cov_strength[cov], st)  # fmt: skip


def cpu_baseline(sy, sets, S):
    """The reference's own FP64 engine, core/pgm_core.cpp (oracle/_ref, built from the reference sources),
    single-threaded as it is, on a bounded sample of the workload: one evaluation of its 1/8 sub-box (12,000
    atoms) from the reference's default guess (alpha E, no history), scaled by the atom ratio to the full box.
    That scaling favours the reference (its pair search is O(N^2)).  Falls back to this build's FP64 oracle
    (oracle/pgm_oracle.cpp) where the reference library was not built.  Returns (baseline, oracle energy and
    forces of the full workload from the GPU's guess, their CG iterations)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    sub = sub_box(sy, S)
    try:
        r, kind = O.Oracle(sub, kind="reference"), "reference"
    except (OSError, RuntimeError):
        r, kind = O.Oracle(sub), "port"
    t0 = time.perf_counter()
    r.compute(sub.pos, sub.box)
    dt = time.perf_counter() - t0
    src = "core/pgm_core.cpp (oracle/_ref)" if kind == "reference" else "oracle/pgm_oracle.cpp"
    cb = {
        "value": round(sub.n / sy.n / dt, 6), "unit": "evals/s", "cores": 1, "kind": kind,
        "sample": f"one FP64 evaluation of the workload's 1/8 sub-box ({sub.n} atoms, {sub.box[0, 0]:.2f} nm, "
                  f"PME {sub.settings['grid'][0]}^3, tau 1e-4, the reference's default alpha E guess: "
                  f"{r.iterations()} CG iterations) by {src} on 1 host thread in {dt:.2f} s, scaled by "
                  f"{sub.n}/{sy.n} to the full box",
        "cpu": cpu_model(), "nproc": os.cpu_count(),
    }  # fmt: skip
    del r
    # the checker: this build's oracle on the full workload from the GPU engines' guess
    o = O.Oracle(sy)
    for mu in sets[::-1]:
        o.push_history(mu)
    e, f = o.compute(sy.pos, sy.box)
    return cb, e, f, o.iterations()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-snapshot adjoint batch")
    ap.add_argument("--no-gvdw", action="store_true", help="skip the GVDW-on timing of the same workload")
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--no-large", action="store_true", help="skip the 999,000-atom box")
    ap.add_argument("--no-md", action="store_true", help="skip the MD-like steps and the mu3 variant")
    ap.add_argument("--md-steps", type=int, default=100)
    ap.add_argument("--large-only", action="store_true",
                    help="profiler pass over the 999,000-atom box alone (both precisions, --steps timed "
                         "evaluations each); prints its figures, not the JSON line")  # fmt: skip
    ap.add_argument("--batch-only", action="store_true",
                    help="the adjoint batch alone, with 0.1 s pauses around each precision's timed steps so a "
                         "kernel trace splits at them; prints its figures, not the JSON line")  # fmt: skip
    ap.add_argument("--gap-report", metavar="KERNEL_TRACE_CSV",
                    help="read a kernel trace of a --batch-only run and print, per segment between pauses, "
                         "dispatches, span, GPU-busy time and the idle (launch-gap) share; no GPU")  # fmt: skip
    ap.add_argument("--profile", type=int, default=0, metavar="N",
                    help="profiler pass: N evaluations of the workload per precision (each engine autotuned "
                         "first) and one calibration copy of "
                         f"{CAL_COPY_BYTES >> 20} MiB, no timing, no JSON line")  # fmt: skip
    a = ap.parse_args()
    if a.gap_report:
        print(json.dumps(gap_report(a.gap_report)))
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs one system on one GPU; --gpus must be 1 (multi-GPU sharding "
                                   "is a design note, not a code path)"}))  # fmt: skip
        sys.exit(2)

    import torch

    from pgmhip import engine as E
    from pgmhip import systems as S

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if a.large_only:
        print(json.dumps(large_box(torch, E, S, a.steps, a.warmup, log)))
        return
    if a.batch_only:
        print(json.dumps(adjoint_batch(torch, E, S, a.batch_reps, log, pause=True)))
        return

    sy = S.config2()
    log(f"workload {sy.name}: {sy.n} atoms")
    sets = prior_dipole_sets(sy, E)
    log("prior dipole sets ready")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    pos_d = torch.from_numpy(sy.pos.copy()).to(dev)
    f_d = torch.zeros((sy.n, 3), dtype=torch.float64, device=dev)
    box = sy.box
    if a.profile:
        for prec in ("mixed", "double"):
            eng = E.Engine.for_system(sy, precision=prec)
            eng.set_history(sets)
            eng.freeze_history(True)
            eng.compute(sy.pos, box)
            eng.autotune(3)
            for _ in range(a.profile):
                eng.compute_device_f64(pos_d.data_ptr(), box, True, f_d.data_ptr(), stream)
            torch.cuda.synchronize()
            del eng
        # a plain device copy of known size: the scale of the HBM byte counters
        src = torch.ones(CAL_COPY_BYTES // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)   # one elementwise kernel: reads and writes CAL_COPY_BYTES
        torch.cuda.synchronize()
        log(f"profile pass: {a.profile} evaluations per precision, "
            f"calibration copy {CAL_COPY_BYTES} bytes read + {CAL_COPY_BYTES} bytes written")  # fmt: skip
        return

    engines, results, tuned = {}, {}, {}
    for prec in ("mixed", "double"):
        eng = E.Engine.for_system(sy, precision=prec)
        eng.set_history(sets)
        eng.freeze_history(True)
        engines[prec] = eng
        t, f, _ = eng.compute(sy.pos, box)  # checked evaluation (host arrays)
        results[prec] = (t, f, eng.iterations())
        tuned[prec] = eng.autotune(3)
        VARIANTS[prec] = (*tuned[prec][0], eng.field_otf(), eng.pme_overlap(), eng.setup_lanes())
        log(f"{prec}: {eng.iterations()} CG iterations, E = {E.total_energy(t):.6f}, pair passes {tuned[prec][0]}")
    P = engines["double"].num_pairs()
    C = engines["double"].num_candidates()
    tight = E.Engine.for_system(sy, precision="double", tol=1e-10, max_iter=500, peek=0.0)
    tight.set_history(sets)
    tt, ft, _ = tight.compute(sy.pos, box)

    secs, phases, live, dom = {}, {}, {}, {}
    for prec in ("mixed", "double"):
        phases[prec] = phase_table(engines[prec], pos_d, f_d, box, stream, torch, 5)
        dom[prec] = dominant_pair_phase(phases[prec])
        secs[prec], live[prec] = timed_steps(torch, engines[prec], pos_d, f_d, box, a.steps, a.warmup, stream,
                                             dom[prec])  # fmt: skip
        log(f"{prec}: {a.steps / secs[prec]:.2f} evals/s")
    n_it = results["mixed"][2]
    roof = roofline(dom["mixed"], live["mixed"], P, C, sy.n, "mixed", *tuned["mixed"][0],
                    engines["mixed"].setup_lanes())
    roof64 = roofline(dom["double"], live["double"], P, C, sy.n, "double", *tuned["double"][0],
                      engines["double"].setup_lanes())

    value = a.steps / secs["mixed"]
    v64 = a.steps / secs["double"]
    tm, fm, _ = results["mixed"]
    td, fd, _ = results["double"]
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "evals/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1000.0 * secs["mixed"] / a.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {
            "workload": "water96k: 32,000 rigid waters (96,000 atoms), 6.84 nm cubic box (3x liquid density), "
                        "rc 0.9 nm, beta 3.5/nm, PME 72^3 order 6, tau 1e-4, mu4 guess from 4 seeded prior "
                        "dipole sets, peek 0.65; FP32 pair/PME arithmetic, FP64 accumulation and CG",
            "atoms": sy.n, "pairs_within_cutoff": P, "list_candidates": C, "cg_iterations": n_it,
            "grid": list(sy.settings["grid"]),
            "pair_passes": {"lanes": tuned["mixed"][0][0], "setup_lanes": engines["mixed"].setup_lanes(),
                            "fused_setup": tuned["mixed"][0][1],
                            "field_on_the_fly": engines["mixed"].field_otf(), "pme_beside": engines["mixed"].pme_overlap(),
                            "cg_product_ms_pme_after_beside": [round(v, 4) for v in engines["mixed"].pme_overlap_times()]},
        },
        "fp64_path": {"value": round(v64, 3), "ms_per_step": round(1000.0 * secs["double"] / a.steps, 4),
                      "cg_iterations": results["double"][2], "roofline": roof64,
                      "pair_passes": {"lanes": tuned["double"][0][0], "setup_lanes": engines["double"].setup_lanes(),
                                      "fused_setup": tuned["double"][0][1],
                                      "pme_beside": engines["double"].pme_overlap(),
                                      "cg_product_ms_pme_after_beside": [round(v, 4) for v in
                                                                         engines["double"].pme_overlap_times()]}},
        "mixed_over_fp64": round(value / v64, 3),
        "accuracy": {
            "force_relrms_mixed_vs_fp64_gpu": E.rel_rms(fm, fd),
            "force_relrms_mixed_vs_fp64_tau1e-10": E.rel_rms(fm, ft),
            "force_relrms_fp64_vs_fp64_tau1e-10": E.rel_rms(fd, ft),
            "energy_rel_mixed_vs_fp64_gpu": E.total_energy(tm) / E.total_energy(td) - 1,
            "energy_rel_mixed_vs_fp64_tau1e-10": E.total_energy(tm) / E.total_energy(tt) - 1,
        },
        "algorithmic_tflops": round(flops_per_eval(P, sy.n, sy.settings["grid"], n_it) * value / 1e12, 3),
        "phases_ms_per_eval": {p: {k: round(v["ms_per_eval"], 4) for k, v in phases[p].items()} for p in phases},
        "roofline": roof,
        "eval_roofline": {"mixed": eval_roofline(P, sy.n, sy.settings["grid"], n_it, value, "mixed"),
                          "double": eval_roofline(P, sy.n, sy.settings["grid"], results["double"][2], v64, "double")},
    }  # fmt: skip
    if not a.no_cpu_baseline:
        log("CPU baseline (the reference's pgm_core.cpp, 1 thread) ...")
        cb, eo, fo, it_o = cpu_baseline(sy, sets, S)
        out["cpu_baseline"] = cb
        out["accuracy"]["force_relrms_fp64_gpu_vs_oracle"] = E.rel_rms(fd, fo)
        out["accuracy"]["force_relrms_mixed_vs_oracle"] = E.rel_rms(fm, fo)
        out["accuracy"]["energy_rel_mixed_vs_oracle"] = E.total_energy(tm) / eo - 1
        out["accuracy"]["oracle_cg_iterations"] = it_o
    else:
        out["cpu_baseline"] = None
    if not a.no_md:
        out["md_like"] = md_like(torch, E, sy, a.md_steps, log)
        out["liquid98k_vs_reference"] = liquid98k(torch, E, S, a.md_steps, log)
        out["mu3"] = mu3_variant(torch, E, sy, sets, pos_d, f_d, stream, max(10, a.steps // 2), a.warmup, log)
    if not a.no_gvdw:
        out["gvdw"] = gvdw_cost(torch, E, sy, sets, pos_d, f_d, stream, max(10, a.steps // 2), a.warmup, log)
        out["gvdw"]["pair_final_ms_without"] = {p: round(phases[p]["pair_final"]["ms_per_eval"], 4) for p in phases}
    if not a.no_batch:
        out["adjoint_batch"] = adjoint_batch(torch, E, S, a.batch_reps, log)
    if not a.no_large:
        for prec in list(engines):
            del engines[prec]
        del tight
        out["large_box"] = large_box(torch, E, S, max(5, a.steps // 5), 2, log)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
