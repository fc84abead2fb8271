"""Benchmark of one VMC step of the continuum spinful Psiformer on one MI355X.

Workload (the headline configuration): N = 16 electrons, rs = 3 (L = rs sqrt(pi N)), Rashba
lambda = 0.8, Zeeman h_z = 0.05, d_model = 128, 4 layers, 8 heads, K = 8 determinants, 8
Fourier harmonics (P = 567,696 real parameters), 8192 walkers; synthetic seeded weights
(backflow ~ (1+i) N(0, 0.05^2)), walkers uniform in the cell then burnt in.  One step =
10 Metropolis sweeps (1 joint move + 2 spin flips each) + E_loc (forward-Laplacian kinetic,
Rashba SOC, Zeeman) + Ewald Coulomb + median/MAD clip + the per-walker gradient O (hand-written
backward) + MinSR (Gram on MFMA, float64 Cholesky, one refinement step, push-through) + the
parameter update.

--precision mixed (default): float32 backbone activations whose dense layers run as bf16x3 on MFMA
(float32 operands split into bf16 hi/lo pairs, float32 accumulation), bf16 MFMA Gram with float32
accumulation, float64 determinants / Ewald / Cholesky.  --precision fp64: everything float64 (the
plain FP64 GPU path).  Prints one JSON line; --gpus > 1 is refused (one GPU only).

The default run times the mixed path ("value"), then --fp64-steps steps of the plain float64 path
from the same parameters and burnt-in walkers ("fp64_value", "ratio_mixed_over_fp64").  Outside the
timed regions it records "accuracy": the mixed path against the float64 GPU path on every walker
(per-step energy, log psi, E_loc, MinSR direction) and both against the float64 oracle on a subset
(log psi, E_loc, Ewald), and "cpu_baseline": the oracle's step measured on a few walkers and
extrapolated to the full walker count.  "roofline" is the Gram (the dominant kernel), timed with
hipEvents around its dispatches; "step_roofline" is the whole step's algorithmic floor over its
measured time.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "continuum-nqs_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np
import torch

METRIC = "walker-steps/s incl. MinSR at N=16, mixed vs FP64 GPU; |dE| vs FP64 oracle"
N_E, RS, LAM, HZ = 16, 3.0, 0.8, 0.05
ARCH = dict(embed_dim=128, depth=4, n_heads=8, n_freq=8, n_dets=8)
PEAK = {"bf16": 2.5e15, "fp32": 157.3e12, "fp64": 78.6e12}


def build_system(precision, W, seed=0, gemm="bf16x3"):
    from nqs.rashba_reference import noninteracting_gs, occupied_orbital_matrix, planewave_envelope
    from nqs.psiformer import ContinuumSpinorPsiformer, ravel_params
    L = RS * math.sqrt(math.pi * N_E)
    g = noninteracting_gs(N_E, L, 1.0, LAM, HZ, n_cut=10, twist=(0.1, 0.17))
    Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
    model = ContinuumSpinorPsiformer(n_elec=N_E, Gs=Gs, W=Wenv, L=L, precision=precision, gemm=gemm, **ARCH)
    params = model.init(seed)
    rng = np.random.default_rng(seed + 1)
    kr = params["params"]["backflow"]["kernel"]
    z = torch.tensor(rng.standard_normal(tuple(kr.shape)) * 0.05)
    params["params"]["backflow"]["kernel"] = (z + 1j * z).to(torch.complex128)
    theta = ravel_params(params)
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed + 2)
    R = torch.rand(W, N_E, 2, generator=gen, dtype=torch.float64, device=dev) * L
    S = torch.randint(0, 2, (W, N_E), generator=gen, device=dev).to(torch.int32)
    return dict(model=model, theta=theta, R=R, S=S, L=L, Gs=Gs, Wenv=Wenv)


def cpu_baseline(sysd, n_walkers=8, seed=0, W_full=8192):
    """The oracle's step (torch float64 on the host cores): one full step measured on a bounded
    sample of walkers, and the headline step extrapolated from measured pieces.  Walker-parallel
    stages (sweeps, dense-Hessian E_loc, reverse-mode O) scale linearly in W; the oracle's MinSR
    (Otil Otil^T, LU solve, push-through) is timed on its own at 2 W_g = 256 rows of random data with
    the real P and scaled as W^2 P for the Gram, W P for the push-through and W^3 for an LU timed at
    4096^2."""
    from oracle import psiformer as opf, ewald as oew, cpu_step
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    a = opf.Arch(N_E, sysd["Gs"], sysd["Wenv"], sysd["L"], **ARCH)
    th = sysd["theta"].cpu()
    _, _, ce = oew.make_ewald(sysd["L"], 1.0, n_real=4, n_recip=8)
    R = sysd["R"][:n_walkers].cpu().numpy()
    S = sysd["S"][:n_walkers].cpu().numpy().astype(np.int64)
    tm = {}
    t0 = time.perf_counter()
    cpu_step.vmc_step(a, th, R, S, lambda_r=LAM, h_z=HZ, coulomb=ce, n_sweeps=10, step=0.3 * RS, lam=1e-4,
                      lr=0.05, seed=seed, timings=tm)
    dt = time.perf_counter() - t0
    # MinSR pieces at scale
    P = th.shape[0]
    Wg = 128
    gen = torch.Generator().manual_seed(3)
    Ot = torch.randn(2 * Wg, P, generator=gen, dtype=torch.float64)
    t1 = time.perf_counter()
    T = Ot @ Ot.T
    t_gram = time.perf_counter() - t1
    y = torch.randn(2 * Wg, generator=gen, dtype=torch.float64)
    t1 = time.perf_counter()
    Ot.T @ y
    t_push = time.perf_counter() - t1
    n_lu = 4096
    B = torch.randn(n_lu, n_lu, generator=gen, dtype=torch.float64)
    B = B @ B.T + n_lu * torch.eye(n_lu, dtype=torch.float64)
    t1 = time.perf_counter()
    torch.linalg.solve(B, torch.randn(n_lu, 1, generator=gen, dtype=torch.float64))
    t_lu = time.perf_counter() - t1
    del Ot, T, B
    per_walker = (tm["sample"] + tm["eloc"] + tm["grad"]) / n_walkers
    ext = {"walker_stages_s": per_walker * W_full, "gram_s": t_gram * (W_full / Wg) ** 2,
           "pushthrough_s": t_push * (W_full / Wg), "lu_s": t_lu * (2 * W_full / n_lu) ** 3}
    t_full = sum(ext.values())
    return {"value": W_full / t_full, "unit": "walker-steps/s", "cores": cores, "kind": "port",
            "sample": f"the float64 torch oracle's step (10 sweeps with log psi recomputed per proposal, "
                      f"dense-Hessian E_loc, reverse-mode O, LU MinSR): value = {W_full} walkers / the "
                      f"extrapolated step time; measured: one full step on {n_walkers} walkers "
                      f"({n_walkers / dt:.3f} walker-steps/s, stage seconds "
                      f"{json.dumps({k: round(v, 3) for k, v in tm.items()})}), the Gram Otil Otil^T on "
                      f"{2 * Wg} x {P} ({t_gram:.3f} s), the push-through ({t_push:.4f} s), a {n_lu}^2 LU solve "
                      f"({t_lu:.3f} s)",
            "measured_sample_walker_steps_per_s": n_walkers / dt,
            "extrapolated_s": {k: round(v, 1) for k, v in ext.items()}}


def _kernel_sha(path):
    import hashlib
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()[:16]


def gram_pmc(precision):
    """rocprofv3 PMC figures of the Gram on this workload (traffic_bytes: FETCH_SIZE / WRITE_SIZE
    scaled by a plain copy kernel; mfma_util_pct: MfmaUtil), read from the profile summary they were
    written to.  Returned only when the summary names the sha256 of the Gram's source file that is
    built now, with that provenance; otherwise None."""
    f = os.path.join(ROOT, "profiles", "gram_hbm_traffic_n16_w8192.json")
    try:
        with open(f) as fh:
            js = json.load(fh)
        rec = js[precision]
        src = os.path.join(ROOT, "continuum-nqs_amd", "csrc", "minsr.hip")
        if js.get("minsr_hip_sha256_16") != _kernel_sha(src):
            return None
        return {"traffic": float(rec["traffic_bytes"]), "mfma_util_pct": float(rec["mfma_util_pct"]),
                "source": "profiles/gram_hbm_traffic_n16_w8192.json (rocprofv3 --pmc on this source revision, "
                          "minsr.hip sha256 " + js["minsr_hip_sha256_16"] + ")"}
    except (OSError, KeyError, ValueError):
        return None


def f_gemm(N=N_E, d=ARCH["embed_dim"], depth=ARCH["depth"], F=ARCH["n_freq"], K=ARCH["n_dets"]):
    """Dense-layer + attention FLOP of one log psi forward (one channel): embed, per block QKV, output
    projection and the d -> 2d -> d MLP (16 N d^2) plus QK^T and PV (4 N^2 d), complex backflow (two
    real GEMMs, 4 N d K N)."""
    return 2 * N * (4 * F + 2) * d + depth * (16 * N * d * d + 4 * N * N * d) + 4 * N * d * K * N


def step_roofline(precision, W, P, ldo, ms, n_sweeps=10, n_flip=2, refine=1):
    """Whole-step speed-of-light: sum over the step's stages of max(flop / peak, bytes / HBM), each
    term from the algorithm (not from counters), over the measured step time.  Stages: walker GEMMs
    (n_sweeps (1 + n_flip) forwards, E_loc jets with 2N + 2 channels for the Laplacian and 3 per
    flipped configuration for the SOC term, about 4 forwards for the backward) at the bf16 MFMA peak
    x 3 (bf16x3 runs three bf16 products per float32 product) or the f64 peak; determinant LUs and
    inverses (f64); O written once; the Gram operand (bf16: O read once, bf16 copy written; float64:
    centred in place, read + write); the Gram (2W)(2W+1)P at bf16 / f64 MFMA peak; the Cholesky
    (2W)^3/3 at the f64 peak; the refinement residual and the push-through (three reads of O)."""
    HBM = 8.0e12
    N, K = N_E, ARCH["n_dets"]
    mixed = precision == "mixed"
    b = 4 if mixed else 8
    gemm_peak = PEAK["bf16"] / 3 if mixed else PEAK["fp64"]
    walker_flop = W * f_gemm() * (n_sweeps * (1 + n_flip) + (2 * N + 2) + 3 * N + 4)
    lu = 8.0 * N ** 3 / 3
    det_flop = W * K * (n_sweeps * (1 + n_flip) * lu + (lu + 8 * N ** 3 + 2 * N * 8 * N ** 3)
                        + N * (lu + 8 * N ** 3 + 2 * 8 * N ** 3))
    o_bytes = 2.0 * W * ldo * b
    n = 2 * W
    terms = {
        "walker_gemms": walker_flop / gemm_peak,
        "determinants": det_flop / PEAK["fp64"],
        "O_write": o_bytes / HBM,
        "gram_operand": (o_bytes + 2.0 * W * ldo * 2) / HBM if mixed else 2 * o_bytes / HBM,
        "gram": float(n) * (n + 1) * P / (PEAK["bf16"] if mixed else PEAK["fp64"]),
        "cholesky": float(n) ** 3 / 3 / PEAK["fp64"],
        "refine_pushthrough": (2 * refine + 1) * o_bytes / HBM,
    }
    floor = sum(terms.values())
    return {"floor_ms": 1e3 * floor, "measured_ms": ms, "frac": 1e3 * floor / ms,
            "terms_ms": {k: round(1e3 * v, 2) for k, v in terms.items()}}


def accuracy(vs, sysd, coul, gemm, n_oracle=2, n_ewald=64):
    """Mixed vs float64: the float64 GPU path on every walker (same walkers and parameters as the
    last timed step), and the float64 oracle on a subset (log psi of 8 walkers, E_loc of n_oracle
    walkers, Ewald of n_ewald walkers) for both GPU paths."""
    from nqs.psiformer import ContinuumSpinorPsiformer, LogPsi
    from nqs.local_energy import make_local_energy
    from nqs.minsr import grad_params, minsr_direction, clip_energies
    from oracle import psiformer as opf, local_energy as ole, ewald as oew
    m64 = ContinuumSpinorPsiformer(n_elec=N_E, Gs=sysd["Gs"], W=sysd["Wenv"], L=sysd["L"], precision="fp64", **ARCH)
    R, S, theta = vs.chain.R, vs.chain.S, vs.theta
    lps = {"mixed": LogPsi(vs.model, theta), "fp64": LogPsi(m64, theta)}
    res = {}
    for k, lp in lps.items():
        lv = lp(R, S)
        e = make_local_energy(lp, m_star=1.0, lambda_r=LAM, h_z=HZ)(R, S)
        res[k] = dict(lp=lv, e=e)
    ec = coul(R)
    dlp = res["mixed"]["lp"] - res["fp64"]["lp"]
    dim = torch.remainder(dlp.imag + math.pi, 2 * math.pi) - math.pi
    de = res["mixed"]["e"] - res["fp64"]["e"]
    E_m = float((res["mixed"]["e"].real + ec).mean())
    E_d = float((res["fp64"]["e"].real + ec).mean())
    out = {"reference": "float64 GPU path on all walkers (same walkers and parameters); the float64 oracle on "
                        "the subset in 'oracle'",
           "gemm": gemm, "walkers": int(R.shape[0]), "E_step_fp64": E_d, "dE_step": abs(E_m - E_d),
           "logpsi_re_max_abs": float(dlp.real.abs().max()), "logpsi_im_max_abs": float(dim.abs().max()),
           "eloc_max_abs": float(de.abs().max()), "eloc_rms": float(de.abs().pow(2).mean().sqrt())}
    # the oracle on a subset of the same walkers
    a = opf.Arch(N_E, sysd["Gs"], sysd["Wenv"], sysd["L"], **ARCH)
    th = theta.cpu()
    Rc, Sc = R[:max(8, n_oracle)].cpu(), S[:max(8, n_oracle)].cpu().long()
    lo = torch.stack([opf.logpsi(a, th, Rc[w], Sc[w]) for w in range(8)]).numpy()
    Eo = ole.make_local_energy(opf.make_logpsi(a, th), 1.0, LAM, HZ)
    eo = np.array([Eo(Rc[w], Sc[w]) for w in range(n_oracle)])
    _, _, ceo = oew.make_ewald(sysd["L"], 1.0, n_real=4, n_recip=8)
    ewo = np.array([ceo(r) for r in R[:n_ewald].cpu().numpy()])
    ewg = ec[:n_ewald].cpu().numpy()
    orc = {"logpsi_walkers": 8, "eloc_walkers": n_oracle, "ewald_walkers": n_ewald,
           "ewald_max_abs": float(np.max(np.abs(ewg - ewo))), "ewald_max_rel": float(np.max(np.abs(ewg - ewo) / np.abs(ewo)))}
    for k in ("mixed", "fp64"):
        lg = res[k]["lp"][:8].cpu().numpy()
        d = lg - lo
        orc[f"logpsi_{k}_max_abs"] = float(max(np.abs(d.real).max(),
                                               np.abs((d.imag + np.pi) % (2 * np.pi) - np.pi).max()))
        orc[f"eloc_{k}_max_abs"] = float(np.abs(res[k]["e"][:n_oracle].cpu().numpy() - eo).max())
    # bound on the mixed per-step energy against the oracle: |E_mixed - E_fp64| over all walkers plus
    # the float64 path's largest per-walker deviation from the oracle on the subset
    orc["dE_step_bound"] = out["dE_step"] + orc["eloc_fp64_max_abs"] + orc["ewald_max_abs"]
    out["oracle"] = orc
    out["ewald_max_abs"] = orc["ewald_max_abs"]
    # MinSR direction from each path's own O, clip and Gram
    dth = {}
    for k, lp in lps.items():
        e_clip, _ = clip_energies(res[k]["e"] + ec, 5.0)
        O, _ = grad_params(lp, R, S, None)
        dth[k] = minsr_direction(O, e_clip, vs.lam, "bf16" if k == "mixed" else "fp64", 1)
        del O
        torch.cuda.empty_cache()
    out["dtheta_rel"] = float((dth["mixed"] - dth["fp64"]).norm() / dth["fp64"].norm())
    return out


def time_path(model, theta, R, S, coul, args, gram, steps, warmup, graph=False, burn=0):
    """Warm-up and timed VMC steps of one precision; returns (VMCStep, ms per step, stage means,
    MinSR sub-stage means)."""
    from nqs import _lib
    from nqs.minsr import VMCStep
    vs = VMCStep(model, theta, R, S, m_star=1.0, lambda_r=LAM, h_z=HZ, coulomb=coul, lr=0.05, lam=1e-4,
                 n_sweeps=args.sweeps, step_size=0.3 * RS, clip_k=5.0, gram=gram, refine=1, graph=graph,
                 overlap=args.overlap == "on")
    if burn:
        vs.chain.run(burn, 0.3 * RS, 2, key=7)
    lib = _lib.load()
    for k in range(warmup):
        vs.step(key=100 + k)
    torch.cuda.synchronize()
    lib.nqs_minsr_profile(1, None)
    vs.timers = []
    gram_ms = []
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for k in range(steps):
        vs.step(key=1000 + k)
        ms5 = (ctypes.c_float * 5)()
        _lib.check(lib.nqs_minsr_profile(-1, ms5), "nqs_minsr_profile")
        gram_ms.append(list(ms5))
    t1.record()
    t1.synchronize()
    lib.nqs_minsr_profile(0, None)
    ms = t0.elapsed_time(t1) / steps
    stages = np.array([[a.elapsed_time(b) for a, b in zip(t[:-1], t[1:])] for t in vs.timers]).mean(0)
    vs.timers = None
    return vs, ms, stages, np.array(gram_ms).mean(0)


def _tile_choices():
    """The autotuner's choices so far: dense-layer GEMM tile per (N, K, log2 M) class, waves per
    workgroup of the attention jet and determinant heads per shape class."""
    from nqs import _lib
    names = {"1": "attention", "2": "det16", "3": "det32", "4": "gemm_tile"}
    out = []
    for line in _lib.autotune_report().splitlines():
        if not line.startswith("kernel "):
            continue
        k = line.split()[1]
        shape = line[line.index("(") + 1:line.index(")")]
        out.append(f"{names.get(k, k)} ({shape}) -> {line.rsplit('-> ', 1)[-1]}")
    return out


def _stage_dict(stages, mins):
    """The step's stages in order (HIP events on the main stream): sweeps, O rows, centred Gram, the local
    energies with Ewald and clip (the Cholesky factorisation runs beside them on a side stream), solve and
    push-through, update; and the MinSR sub-stages timed on the streams they ran on."""
    return {"sample": float(stages[0]), "grad": float(stages[1]), "center_gram": float(stages[2]),
            "eloc_ewald_clip": float(stages[3]), "solve": float(stages[4]), "update_repack": float(stages[5]),
            "minsr_center": float(mins[0]), "minsr_gram": float(mins[1]), "minsr_cholesky_beside_eloc": float(mins[2]),
            "minsr_solve_refine": float(mins[3]), "minsr_pushthrough": float(mins[4])}


C4 = dict(N=32, rs=5.0, lam=0.5, hz=0.0, arch=dict(embed_dim=128, depth=4, n_heads=8, n_freq=8, n_dets=16))


def bench_c4(args):
    """The large configuration: N = 32, rs = 5, lambda = 0.5, h_z = 0, d = 128, 4 layers, 8 heads,
    K = 16 (P = 666,784), 16384 walkers, Ewald 5 x 5 real-space images and |m| <= 8, two Gaussian
    z-field pins (0.05 Ha, width rs), float32 MFMA Gram; mixed-precision backbone."""
    from nqs.rashba_reference import noninteracting_gs, occupied_orbital_matrix, planewave_envelope
    from nqs.psiformer import ContinuumSpinorPsiformer, ravel_params
    from nqs.ewald import make_ewald
    from nqs.pinning import make_gaussian_pins
    from nqs.minsr import VMCStep
    from nqs import _lib
    N, rs = C4["N"], C4["rs"]
    W = args.walkers if args.walkers != 8192 else 16384
    L = rs * math.sqrt(math.pi * N)
    g = noninteracting_gs(N, L, 1.0, C4["lam"], C4["hz"], n_cut=10, twist=(0.1, 0.17))
    Gs, Wenv = planewave_envelope(*occupied_orbital_matrix(g))
    model = ContinuumSpinorPsiformer(n_elec=N, Gs=Gs, W=Wenv, L=L, precision="mixed", gemm=args.gemm, **C4["arch"])
    params = model.init(0)
    rng = np.random.default_rng(1)
    kr = params["params"]["backflow"]["kernel"]
    z = torch.tensor(rng.standard_normal(tuple(kr.shape)) * 0.05)
    params["params"]["backflow"]["kernel"] = (z + 1j * z).to(torch.complex128)
    theta = ravel_params(params)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(2)
    R = torch.rand(W, N, 2, generator=gen, dtype=torch.float64, device="cuda") * L
    S = torch.randint(0, 2, (W, N), generator=gen, device="cuda").to(torch.int32)
    pins = make_gaussian_pins([(0.25 * L, 0.3 * L), (0.7 * L, 0.65 * L)], 0.05, rs, kind="spin")
    _, _, coul = make_ewald(L, 1.0, n_real=2, n_recip=8)
    vs = VMCStep(model, theta, R, S, m_star=1.0, lambda_r=C4["lam"], h_z=C4["hz"], coulomb=coul, lr=0.05, lam=1e-4,
                 n_sweeps=args.sweeps, step_size=0.3 * rs, clip_k=5.0, gram="fp32", refine=1, pinning=pins,
                 overlap=args.overlap == "on")
    if args.burn:
        vs.chain.run(args.burn, 0.3 * rs, 2, key=7)
    lib = _lib.load()
    for k in range(args.warmup):
        vs.step(key=100 + k)
    torch.cuda.synchronize()
    lib.nqs_minsr_profile(1, None)
    vs.timers = []
    gram_ms = []
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for k in range(args.steps):
        vs.step(key=1000 + k)
        ms5 = (ctypes.c_float * 5)()
        _lib.check(lib.nqs_minsr_profile(-1, ms5), "nqs_minsr_profile")
        gram_ms.append(list(ms5))
    t1.record()
    t1.synchronize()
    lib.nqs_minsr_profile(0, None)
    ms = t0.elapsed_time(t1) / args.steps
    stages = np.array([[a.elapsed_time(b) for a, b in zip(t[:-1], t[1:])] for t in vs.timers]).mean(0)
    mins = np.array(gram_ms).mean(0)
    P = model.n_params()
    n = 2 * W
    achieved = float(n) * (n + 1) * P / (mins[1] / 1e3)
    stats = vs.last["stats"].cpu().numpy()
    return {"metric": "walker-steps/s incl. MinSR at N=32 (large configuration)", "value": W / (ms / 1000.0),
            "unit": "walker-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16+f64",
            "data": "synthetic",
            "config": {"workload": f"N=32 rs=5 lambda=0.5 h_z=0 d=128 layers=4 heads=8 K=16 F=8 walkers={W} "
                                   "Ewald 5x5 images |m|<=8, two Gaussian z pins 0.05 Ha width rs",
                       "walkers": W, "params": P, "sweeps": args.sweeps, "gram": "fp32", "gemm": args.gemm},
            "roofline": {"bound": "mfma", "kernel": "Gram (k_gemm_nt<float>, lower tiles)",
                         "achieved": achieved / 1e12, "peak": PEAK["fp32"] / 1e12, "unit": "TFLOP/s",
                         "frac": achieved / PEAK["fp32"], "traffic": None, "gram_ms": float(mins[1])},
            "stage_ms": _stage_dict(stages, mins),
            "E_mean": float(stats[0]), "E_err": float(stats[1]), "cpu_baseline": None}


def bench_c5(args):
    """The batched uneven ensemble: 24 systems, N from {4, 6, 8, 12, 16, 24, 32}, rs in [1, 6],
    lambda in [0, 1], h_z in [0, 0.2], d = 64, 3 layers, 4 heads, K = 4, 1024 walkers each."""
    from nqs.ensemble import Ensemble, make_c5_systems
    W = args.walkers if args.walkers != 8192 else 1024
    systems = make_c5_systems(24, seed=0)
    ens = Ensemble(systems, n_walkers=W, n_sweeps=args.sweeps, graph=True)
    if args.burn:
        ens.burn(args.burn)
    for k in range(args.warmup):
        ens.step(key=100 + k)
    torch.cuda.synchronize()
    timers = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ens.step(key=1000 + k, timers=timers)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    st = {k: float(np.mean([t[k] for t in timers])) for k in timers[0]}
    return {"metric": "walker-steps/s incl. MinSR, 24-system uneven ensemble", "value": 24 * W / (ms / 1000.0),
            "unit": "walker-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16+f64",
            "data": "synthetic",
            "config": {"workload": f"24 systems, N in {{4..32}}, d=64 layers=3 heads=4 K=4, {W} walkers each",
                       "systems": [[s["N"], round(s["rs"], 3)] for s in systems], "walkers": W,
                       "sweeps": args.sweeps},
            "stage_ms": st,
            "roofline": None, "cpu_baseline": None}


def _gpu_ms(fn, reps=3):
    """Mean milliseconds per call of fn() over reps calls after one warm-up call (HIP events)."""
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / reps


def _cpu_s(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def bench_next(args):
    """The rows next to the step, at the headline workload (N = 16, 8192 walkers, mixed): the
    first-order Adam step (surrogate gradient as one column GEMV over O, Adam in place, and a whole
    train step), the spin estimators, E_loc with a skyrmion texture field, and the overlap /
    rotation estimators.  Each row carries its GPU time, the roofline of the kernel that is its own
    (the rest is log psi / E_loc work measured on the main line) and the oracle timed on a few
    walkers of the same workload."""
    from nqs.psiformer import LogPsi
    from nqs.local_energy import make_local_energy
    from nqs.ewald import make_ewald
    from nqs.minsr import grad_params
    from nqs.observables import make_spin_estimators
    from nqs.manybody_chern import mc_overlap, ratio_stats
    from nqs.topology import rotation_eigenvalue_vmc
    from nqs import train as ntr, pinning as npn
    from oracle import psiformer as opf, train as otr, observables as oob, local_energy as ole, pinning as opn

    W = args.walkers
    sysd = build_system("mixed", W, gemm=args.gemm)
    model, L, R, S = sysd["model"], sysd["L"], sysd["R"], sysd["S"]
    theta = sysd["theta"].cuda().contiguous()
    _, _, coul = make_ewald(L, 1.0, n_real=4, n_recip=8)
    lp = LogPsi(model, theta)
    HBM = 8.0e12
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    arch = opf.Arch(N_E, sysd["Gs"], sysd["Wenv"], L, **ARCH)
    th_cpu = sysd["theta"].cpu()
    olp = opf.make_logpsi(arch, th_cpu)
    Rc, Sc = R[:4].cpu().numpy(), S[:4].cpu().numpy().astype(np.int64)
    rows = {}

    # f1: surrogate gradient + Adam + whole first-order step
    E = make_local_energy(lp, 1.0, LAM, HZ)
    e = E(R, S) + coul(R)
    O, _ = grad_params(lp, R, S)
    t_g = _gpu_ms(lambda: ntr.vmc_gradient(O, e))
    g_bytes = float(O.data.shape[0]) * O.data.shape[1] * O.data.element_size()
    opt = ntr.Adam(1e-3)
    th2 = theta.clone()
    st = opt.init(th2)
    part = opt.partner(model)
    g = ntr.vmc_gradient(O, e)
    P = theta.shape[0]
    t_a = _gpu_ms(lambda: opt.update_(th2, g, st, part), reps=10)
    a_bytes = 64.0 * P  # theta, g, mu, nu, partner read; theta, mu, nu written (8 bytes each)
    del O
    torch.cuda.empty_cache()
    optm, step = ntr.make_train_step(model, L, N_E, 1.0, LAM, HZ, coul, ntr.Adam(1e-3), n_sweeps=args.sweeps,
                                     step_size=0.3 * RS)
    th3 = theta.clone()
    ost = optm.init(th3)
    R2, S2 = R.clone(), S.clone()
    step(th3, ost, R2, S2, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        th3, ost, R2, S2 = step(th3, ost, R2, S2, 2 + k)[:4]
    torch.cuda.synchronize()
    t_step = 1e3 * (time.perf_counter() - t0) / args.steps
    nw = 2
    c_g = _cpu_s(lambda: otr.surrogate_gradient(arch, th_cpu, Rc[:nw], Sc[:nw], e[:nw].cpu().numpy()))
    c_a = _cpu_s(lambda: otr.adam_step(arch, th_cpu.numpy(), g.cpu().numpy(), None))
    rows["f1_adam_step"] = {
        "train_step_ms": t_step, "walker_steps_per_s": W / (t_step / 1e3),
        "vmc_gradient_ms": t_g, "adam_update_ms": t_a,
        "roofline": {"kernel": "nqs_vmc_gradient (column GEMV over O)", "bound": "hbm",
                     "achieved": g_bytes / (t_g / 1e3) / 1e9, "peak": HBM / 1e9, "unit": "GB/s",
                     "frac": g_bytes / (t_g / 1e3) / HBM, "algorithmic_bytes": g_bytes},
        "adam_roofline": {"bound": "hbm", "achieved": a_bytes / (t_a / 1e3) / 1e9, "peak": HBM / 1e9,
                          "unit": "GB/s", "frac": a_bytes / (t_a / 1e3) / HBM, "algorithmic_bytes": a_bytes},
        "cpu": {"surrogate_gradient_s_per_walker": c_g / nw, "adam_step_s": c_a, "cores": cores, "kind": "port"}}

    # f2: spin estimators (log psi of every walker and of its N single-spin flips)
    est = make_spin_estimators(lp)
    t_s = _gpu_ms(lambda: est(R, S))
    c_s = _cpu_s(lambda: [oob.spin_estimators(olp, Rc[w], Sc[w]) for w in range(nw)])
    rows["f2_spin_estimators"] = {
        "ms": t_s, "walkers_per_s": W / (t_s / 1e3), "configurations_per_s": W * (N_E + 1) / (t_s / 1e3),
        "roofline": None, "note": "N + 1 log psi forwards per walker; their kernels are the log psi kernels of the "
                                  "main line",
        "cpu": {"walkers_per_s": nw / c_s, "cores": cores, "kind": "port"}}

    # f3: E_loc with the skyrmion texture (ext_field) against plain E_loc
    q1 = np.array([2 * np.pi / L, 0.0])
    q2 = np.array([-np.pi / L, np.sqrt(3) * np.pi / L])
    Et = make_local_energy(lp, 1.0, LAM, HZ, ext_field=npn.make_pin("skx", q1, q2, 0.3))
    t_e0 = _gpu_ms(lambda: E(R, S), reps=2)
    t_e1 = _gpu_ms(lambda: Et(R, S), reps=2)
    Eo = ole.make_local_energy(olp, 1.0, LAM, HZ, ext_field=opn.make_pin("skx", q1, q2, 0.3))
    c_e = _cpu_s(lambda: [Eo(torch.tensor(Rc[w]), torch.tensor(Sc[w])) for w in range(1)])
    rows["f3_texture_eloc"] = {
        "eloc_ms": t_e1, "plain_eloc_ms": t_e0, "texture_overhead_ms": t_e1 - t_e0,
        "walkers_per_s": W / (t_e1 / 1e3), "roofline": None,
        "note": "the texture is a Fourier table evaluated inside the E_loc physics kernel",
        "cpu": {"walkers_per_s": 1.0 / c_e, "cores": cores, "kind": "port"}}

    # f4: overlap of two states and the C4 rotation eigenvalue
    rng = np.random.default_rng(5)
    lpb = LogPsi(model, (theta + 1e-3 * torch.tensor(rng.standard_normal(P), device=theta.device)).contiguous())
    t_o = _gpu_ms(lambda: mc_overlap(lp, lpb, R, S))
    t_r = _gpu_ms(lambda: rotation_eigenvalue_vmc(lp, R, S, 4, L))
    la, lb = lp(R, S), lpb(R, S)
    t_rs = _gpu_ms(lambda: ratio_stats(la, lb), reps=10)
    olpb = opf.make_logpsi(arch, lpb.theta.cpu())
    def cpu_overlap():
        a = [complex(olp(torch.tensor(Rc[w]), torch.tensor(Sc[w]))) for w in range(nw)]
        b = [complex(olpb(torch.tensor(Rc[w]), torch.tensor(Sc[w]))) for w in range(nw)]
        oob.ratio_stats(a, b)
    c_o = _cpu_s(cpu_overlap)
    rs_bytes = 32.0 * W
    rows["f4_overlap_rotation"] = {
        "mc_overlap_ms": t_o, "rotation_eigenvalue_ms": t_r, "ratio_stats_ms": t_rs,
        "walkers_per_s": W / (t_o / 1e3),
        "roofline": {"kernel": "nqs_ratio_stats", "bound": "hbm", "achieved": rs_bytes / (t_rs / 1e3) / 1e9,
                     "peak": HBM / 1e9, "unit": "GB/s", "frac": rs_bytes / (t_rs / 1e3) / HBM,
                     "algorithmic_bytes": rs_bytes, "note": "launch-bound at this size (256 KB)"},
        "cpu": {"walkers_per_s": nw / c_o, "cores": cores, "kind": "port"}}

    f1 = rows["f1_adam_step"]
    return {"metric": "first-order (Adam) VMC walker-steps/s at N=16, and the next rows' kernels",
            "value": f1["walker_steps_per_s"], "unit": "walker-steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": 1, "ms_per_step": f1["train_step_ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32+bf16+f64", "data": "synthetic",
            "config": {"workload": "N=16 rs=3 lambda=0.8 h_z=0.05 d=128 layers=4 heads=8 K=8 F=8 walkers=8192",
                       "walkers": W, "params": P, "sweeps": args.sweeps},
            "roofline": dict(f1["roofline"], traffic=None), "rows": rows,
            "cpu_baseline": {"value": 1.0 / (f1["cpu"]["surrogate_gradient_s_per_walker"]), "unit": "walkers/s",
                             "cores": cores, "kind": "port",
                             "sample": f"the oracle's surrogate gradient (reverse mode) on {nw} walkers; each row "
                                       "carries its own oracle timing"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--walkers", type=int, default=8192)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--sweeps", type=int, default=10)
    ap.add_argument("--burn", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-walkers", type=int, default=8)
    ap.add_argument("--gemm", default="bf16x3", choices=["bf16x3", "fp32", "bf16"],
                    help="dense-layer MFMA mode of the mixed backbone")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the mixed-vs-float64 accuracy leg")
    ap.add_argument("--fp64-steps", type=int, default=3,
                    help="timed steps of the plain float64 path after the mixed run (0 = skip; mixed only)")
    ap.add_argument("--graph", action="store_true", help="replay the sweeps from a captured HIP graph")
    ap.add_argument("--overlap", default="on", choices=["on", "off"],
                    help="factorise the Gram on a side stream beside the local energies (VMCStep(overlap=...))")
    ap.add_argument("--config", default="c3", choices=["c3", "c4", "c5", "next"],
                    help="c3: the headline N=16 workload (default); c4: the N=32 large configuration; c5: the "
                         "24-system uneven ensemble; next: the Adam step, spin estimators, texture E_loc and "
                         "overlap / rotation estimators at the c3 workload (each prints its own line)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy 5 times (the calibration kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: runs on several GPUs are not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 << 28, "copies": 5}))
        return
    if args.config == "c4":
        print(json.dumps(bench_c4(args)))
        return
    if args.config == "c5":
        print(json.dumps(bench_c5(args)))
        return
    if args.config == "next":
        print(json.dumps(bench_next(args)))
        return
    from nqs import _lib
    from nqs.ewald import make_ewald
    from nqs.psiformer import ContinuumSpinorPsiformer

    W = args.walkers
    sysd = build_system(args.precision, W, gemm=args.gemm)
    _, _, coul = make_ewald(sysd["L"], 1.0, n_real=4, n_recip=8)
    # burnt-in walkers shared by both precisions
    from nqs.psiformer import LogPsi
    from nqs.sampler import Chain
    ch = Chain(LogPsi(sysd["model"], sysd["theta"].cuda()), sysd["R"], sysd["S"])
    if args.burn:
        ch.run(args.burn, 0.3 * RS, 2, key=7)
    R0, S0 = ch.R.clone(), ch.S.clone()
    del ch
    gram = "bf16" if args.precision == "mixed" else "fp64"
    vs, ms, stages, mins = time_path(sysd["model"], sysd["theta"].clone(), R0, S0, coul, args, gram, args.steps,
                                     args.warmup, graph=args.graph)
    P = sysd["model"].n_params()
    ldo = vs.O.ldo
    n = 2 * W
    gram_flops = float(n) * (n + 1) * P          # n(n+1)/2 outputs of the symmetric rank-P update, 2P flops each
    achieved = gram_flops / (mins[1] / 1e3)
    peak = PEAK["bf16" if gram == "bf16" else "fp64"]
    stats = vs.last["stats"].cpu().numpy()
    pmc = gram_pmc(args.precision) if W == 8192 else None
    out = {"metric": METRIC, "value": W / (ms / 1000.0), "unit": "walker-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+bf16+f64" if args.precision == "mixed" else "f64",
           "data": "synthetic",
           "config": {"workload": f"N=16 rs=3 lambda=0.8 h_z=0.05 d=128 layers=4 heads=8 K=8 F=8 walkers={W}",
                      "walkers": W, "params": P, "sweeps": args.sweeps, "precision": args.precision, "gram": gram,
                      "gemm": args.gemm if args.precision == "mixed" else "fp64", "graph_sweeps": bool(args.graph)},
           "roofline": {"bound": "mfma", "kernel": "Gram (k_syrk_bf16)" if gram == "bf16" else "Gram (k_gemm_nt<double>)",
                        "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                        "frac": achieved / peak, "traffic": pmc["traffic"] if pmc else None,
                        "algorithmic_flop": gram_flops, "gram_ms": float(mins[1]),
                        "mfma_util_pct": pmc["mfma_util_pct"] if pmc else None,
                        "pmc_source": pmc["source"] if pmc else "not measured for this source revision"},
           "step_roofline": step_roofline(args.precision, W, P, ldo, ms, args.sweeps),
           "stage_ms": _stage_dict(stages, mins),
           "E_mean": float(stats[0]), "E_err": float(stats[1]), "failed_steps": vs.failed_steps,
           "autotune": _tile_choices()}
    if args.precision == "mixed" and not args.no_accuracy:
        out["accuracy"] = accuracy(vs, sysd, coul, args.gemm)
    if args.precision == "mixed" and args.fp64_steps > 0:
        vs.O = None
        _lib.release_workspaces()
        torch.cuda.empty_cache()
        m64 = ContinuumSpinorPsiformer(n_elec=N_E, Gs=sysd["Gs"], W=sysd["Wenv"], L=sysd["L"], precision="fp64",
                                       **ARCH)
        v64, ms64, st64, mi64 = time_path(m64, sysd["theta"].clone(), R0, S0, coul, args, "fp64", args.fp64_steps, 1)
        out["fp64"] = {"value": W / (ms64 / 1000.0), "ms_per_step": ms64, "steps": args.fp64_steps, "warmup": 1,
                       "stage_ms": _stage_dict(st64, mi64), "failed_steps": v64.failed_steps,
                       "gram_frac_of_f64_peak": gram_flops / (mi64[1] / 1e3) / PEAK["fp64"],
                       "step_roofline": step_roofline("fp64", W, P, v64.O.ldo, ms64, args.sweeps)}
        out["fp64_value"] = out["fp64"]["value"]
        out["ratio_mixed_over_fp64"] = out["value"] / out["fp64_value"]
        del v64
        _lib.release_workspaces()
        torch.cuda.empty_cache()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sysd, args.cpu_walkers, W_full=W)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
