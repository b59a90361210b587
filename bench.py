"""Benchmark: LLG RK4 steps/s of the 1024 x 1024 x 4 film (4,194,304 cells,
32x32-cell random in-plane domains, seed 2010) with the GPU FMM demag
(P = 6) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--m2l fp32|bf16|bf16x3]
                    [--leaf 4|8] [--probes N] [--no-cpu] [--no-error] [--no-forms]

One step = one fixed-dt (5e-14 s) RK4 step = 4 demag evaluations + 4 fused
field/LLG stage passes, state resident in HBM.  Prints ONE JSON line with the
steps/s, the roofline of the dominant kernel (algorithmic work per launch over
its mean HIP-event duration, measured in this run on the solver's stream),
the demag-field error against an FP64 direct sum on seeded probes, and the
reference CPU path timed on a bounded sample on this host.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mms-gpu_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# ----------------------------------------------------- algorithmic work --
def stage_bytes_per_cell(stage, all_magnetic=False):
    """Compulsory HBM bytes of the fused field+LLG stage kernel per cell:
    phi 4 + m_s 12 + material 1 + FP32 m_{s+1} 12, the FP32 slope
    accumulator (write 12 at stages 1-3, read 12 at stages 2-4), and M_n:
    none at stage 1 (its input is M_n), the FP32 copy (12) at stages 2-3,
    the FP64 base read and write (24 + 24) at stage 4: 41, 65, 65, 89 bytes.
    An all-magnetic film does not read the material byte."""
    return (41, 65, 65, 89)[stage] - (1 if all_magnetic else 0)


def m2l_flops_per_eval(side, zd, P, w, m2l):
    """M2L GEMMs of one evaluation (one launch covers every level): each
    target box and layer contracts 27 offsets x zd source layers of operator
    blocks.  Real embedding: only coefficients with n + m even are kept, K =
    (P+1)(P+2)/2 reals per box.  Returns (useful, issued): useful counts
    2 K^2 per block product; issued counts the MFMA work actually run, blocks
    padded to KP = 16 (K <= 16) or 32 ceil(K / 32), three BF16 products per
    block in the split (bf16x3) form, two in the FP16 hi/lo forms; the 16-bit
    forms of 16-slot blocks run the 27 offsets as 14 pairs (one zero block)."""
    K = (P + 1) * (P + 2) // 2
    KP = 16 if K <= 16 else 32 * ((K + 31) // 32)
    Lleaf = int(math.log2(side)) - int(math.log2(w))
    blocks = sum((4 ** L) * zd * 27 * zd for L in range(2, Lleaf + 1))
    products = {"bf16x3": 3, "bf16x3_f16near": 3, "fp16x2": 2, "fp16x2_bf16x3near": 2,
                "fp16x2_f16x2near": 2}.get(m2l, 1)
    issued_blocks = blocks * 28 / 27 if (KP == 16 and m2l != "fp32") else blocks
    return blocks * 2 * K * K, issued_blocks * 2 * KP * KP * products


def fmm_pass_bytes(side, zd, P, w):
    """Algorithmic HBM bytes of the multipole passes of one evaluation, with
    K = (P+1)(P+2)/2 useful coefficients per box and layer (the stored
    vectors are padded to 32; the per-source-layer M2L partial sums are this
    design's extra traffic, not algorithmic):
      upward   -- read every charge once (4 B/cell), write the multipoles of
                  the subtree levels R..leaf;
      m2m      -- the coarse levels 2..R-1 from level R (read level R, write
                  the coarser levels);
      downward -- read one copy of the locals of every level, read and write
                  the potential (8 B/cell) with L2P fused at the leaves.
    R = leaf level - D, D = 3 (4 on trees deeper than 8 levels below the
    root's subtrees), as the library splits the tree."""
    K = (P + 1) * (P + 2) // 2
    Lleaf = int(math.log2(side)) - int(math.log2(w))
    D = 3 if Lleaf - 3 <= 5 else 4
    if Lleaf - D < 2:
        D = Lleaf - 2
    R = Lleaf - D
    box = lambda L: (4 ** L) * zd * K * 4
    n = side * side * zd
    return {"upward": 4 * n + sum(box(L) for L in range(R, Lleaf + 1)),
            "m2m": sum(box(L) for L in range(2, R + 1)),
            "downward": sum(box(L) for L in range(2, Lleaf + 1)) + 8 * n}


def p2p_pairs_per_eval(side, zd, w):
    """Near-field pairs: each target against its 3x3 neighbour leaves in
    every layer (self excluded), edge leaves counted with their full window."""
    return side * side * zd * (9 * w * w * zd - 1)


def p2p_flops_per_eval(side, zd, w, m2l):
    """MFMA work of the near-field kernel (L2P is added by the downward
    pass).  FP32 form: one MAC per (target, window source) pair -- the 1/r
    table lookup replaces the distance arithmetic.  16-bit forms (leaf 8):
    the 24 x 24 source window of a target is packed into 18 k steps of 32
    with no padding, three BF16 products each in the BF16 split, one FP16
    product in the FP16 form.  Returns (flops, peak TFLOP/s)."""
    cells = side * side * zd
    if w == 8 and m2l != "fp32":
        prods = {"fp16": 1, "bf16x3_f16near": 1, "fp16x2": 1, "fp16x2_f16x2near": 2}.get(m2l, 3)
        return 2 * cells * zd * (3 * w) * (3 * w) * prods, BF16_PEAK_TFLOPS
    return 2 * cells * zd * 9 * w * w, FP32_PEAK_TFLOPS


def pmc_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by the reading of a
    plain device copy of known size); None when absent."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 4 and f[0] == "traffic" and f[1] == kernel and f[2] == workload:
            return float(f[3])
    return None


# ----------------------------------------------------------- CPU leg --
def host_cpu():
    """CPU model, physical cores and the threads this process may use."""
    model, phys = "unknown", set()
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if "core id" in cur:
                    phys.add((cur.get("physical id", "0"), cur["core id"]))
                cur = {}
                continue
            k, v = (t.strip() for t in line.split(":", 1))
            cur[k] = v
            if k == "model name":
                model = v
        if "core id" in cur:
            phys.add((cur.get("physical id", "0"), cur["core id"]))
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(avail, omp) if omp > 0 else avail
    return model, (len(phys) or None), threads


def cpu_baseline(w, ch, budget_s=20.0):
    """The reference tree-code (fmm_calc CPU branch, P = 4: the reference's
    maximum order) restated in oracle/mms_oracle.c -- the same body as the
    oracle's full evaluation -- timed on one whole level of one charge layer
    on this host's cores (the boxes of the level split over the threads), then
    extrapolated with the exact M2P term count of a whole evaluation (every
    level and layer) to steps/s (4 evaluations per step).  The level is the
    finest above the leaves (most boxes to share among the threads) whose
    estimated time fits the budget."""
    import oracle
    L = oracle.lib()
    side, zd, P = w.xdim, w.zdim, 4
    K = (P + 1) ** 2
    logN = int(math.log2(side))
    q = ch.reshape(zd, side, side)
    model, phys, threads = host_cpu()
    # exact M2P term counts per level: sum over non-empty source boxes of
    # (valid interaction-list cells) x zd target layers x K coefficients
    level_terms = {}
    for lev in range(2, logN + 1):
        wl_ = 2 ** (logN - lev)
        nb = 2 ** lev
        nz = (np.abs(q).reshape(zd, nb, wl_, nb, wl_).sum(axis=(2, 4)) != 0)  # non-empty boxes per layer
        bx = np.arange(nb)
        lo = np.maximum(2 * (bx // 2) - 2, 0)
        hi = np.minimum(2 * (bx // 2) + 3, nb - 1)
        span = hi - lo + 1
        nbr = np.minimum(bx + 1, nb - 1) - np.maximum(bx - 1, 0) + 1
        il = span[None, :] * span[:, None] - nbr[None, :] * nbr[:, None]
        per_box = il * wl_ * wl_ * zd * K
        level_terms[lev] = [float((per_box * nz[z]).sum()) for z in range(zd)]
    total_terms = sum(sum(v) for v in level_terms.values())
    est_rate = 1.4e7 * threads  # rough terms/s, only to pick the level
    cand = [lev for lev in range(2, logN) if level_terms[lev][0] / est_rate <= budget_s]
    lev = cand[-1] if cand else 2
    sample_terms = level_terms[lev][0]
    pot = np.zeros(ch.size, np.float32)
    t0 = time.perf_counter()
    st = L.or_fmm_bfs_sample(ch, pot, side, side, zd, float(w.dx), float(w.dy), float(w.dz), P, 0, lev, lev, 1,
                             threads)
    dt = time.perf_counter() - t0
    if st != 0 or sample_terms == 0:
        return None
    t_eval = dt / sample_terms * total_terms
    return {"value": 1.0 / (4.0 * t_eval), "unit": "steps/s", "cores": threads, "kind": "port",
            "cpu_model": model, "physical_cores": phys,
            "sample": (f"reference tree-code (fmm_calc CPU branch, P=4) restated in oracle/mms_oracle.c; one "
                       f"whole level ({lev}, {4 ** lev} boxes) of charge layer 0 on {threads} threads "
                       f"({sample_terms:.3g} M2P terms in {dt:.1f} s), extrapolated by exact term count "
                       f"({total_terms:.3g} terms per evaluation over {logN - 1} levels and {zd} layers, "
                       f"4 evaluations per step)"),
            "seconds_per_eval_extrapolated": t_eval}


# --------------------------------------------------------- field error --
def field_error(mg, f, w, nprobe=8192):
    """Relative L2 (and max) error of the demag field at seeded interior
    probes against the device FP64 direct sum (mms_gpu.workloads)."""
    from mms_gpu import workloads as wl
    return wl.demag_field_error(f, w, nprobe)


# ---------------------------------------------------------------- main --
PREC_NAMES = ("fp32", "bf16", "bf16x3", "fp16", "bf16x3_f16near", "fp16x2", "fp16x2_bf16x3near", "fp16x2_f16x2near")
FORM_DTYPE = {"fp32": "f32",
              "bf16": "f32 (BF16 M2L operands, BF16x3 near field)",
              "bf16x3": "f32 (BF16x3 M2L and near field)",
              "fp16": "f32 (FP16 M2L and near field)",
              "bf16x3_f16near": "f32 (BF16x3 M2L, FP16 near field)",
              "fp16x2": "f32 (FP16x2 M2L, FP16 near field)",
              "fp16x2_bf16x3near": "f32 (FP16x2 M2L, BF16x3 near field)",
              "fp16x2_f16x2near": "f32 (FP16x2 M2L, FP16x2 near field)"}
# Forms inside the 1024^2 x 4 workload's precision split: FP32 everywhere
# except 16-bit M2L inputs with FP32 accumulation, and a near field in FP32
# or a compensated hi/lo split (BF16x3: both operands; FP16x2: the 1/r
# operator against scaled FP16 charges).  The single-product FP16 near field
# and the plain BF16 M2L are outside it (reported as extras only).
CONFORMING = ("fp32", "bf16x3", "fp16x2_bf16x3near", "fp16x2_f16x2near")
RUN_CHECKPOINTS = (250, 500)  # steps of the C3 run at which the field-error gate is checked again


def prec_of(mg, m2l):
    return {"fp32": mg.M2L_FP32, "bf16": mg.M2L_BF16, "bf16x3": mg.M2L_BF16X3, "fp16": mg.M2L_FP16,
            "bf16x3_f16near": mg.M2L_BF16X3_FP16NEAR, "fp16x2": mg.M2L_FP16X2,
            "fp16x2_bf16x3near": mg.M2L_FP16X2_BF16X3NEAR, "fp16x2_f16x2near": mg.M2L_FP16X2_FP16X2NEAR}[m2l]


def make_film(mg, wl, w, P, m2l, leaf):
    prec = prec_of(mg, m2l)
    f = mg.Film(*w.dims, *w.d, solver=mg.SOLVER_FMM, P=P, m2l_precision=prec, leaf_width=leaf,
                mu_0=wl.MU_0, Ms=wl.MS, Aexch=wl.AEXCH, alfa=wl.ALFA, gamma=wl.GAMMA, material=w.material)
    f.set_m(w.M)
    return f


def timed_steps(f, w, warmup, steps, graph):
    """Steps/s over `steps` fixed-dt RK4 steps after `warmup` untimed ones,
    synchronised on both sides of the timed region."""
    f.steps(warmup, w.dt, use_graph=graph)
    f.sync()
    t0 = time.perf_counter()
    f.steps(steps, w.dt, use_graph=graph)
    f.sync()
    return steps / (time.perf_counter() - t0), time.perf_counter() - t0


def kernel_rooflines(prof, side, zd, P, leaf, m2l):
    """Achieved rate of each kernel class: algorithmic work per launch over
    its mean HIP-event duration on the solver's stream."""
    n = side * side * zd
    kernels = {}
    if "stage" in prof:
        # the measured film is all magnetic: no material byte is read
        avg_b = sum(stage_bytes_per_cell(st, all_magnetic=True) for st in range(4)) / 4 * n
        ms, _ = prof["stage"]
        kernels["stage"] = dict(bound="hbm", achieved=avg_b / (ms * 1e-3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                                per_launch=avg_b, ms=ms)
    if "div" in prof:
        ms, _ = prof["div"]
        kernels["div"] = dict(bound="hbm", achieved=16 * n / (ms * 1e-3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                              per_launch=16 * n, ms=ms)
    if "m2l" in prof:
        ms, _ = prof["m2l"]
        useful, issued = m2l_flops_per_eval(side, zd, P, leaf, m2l)
        peak = BF16_PEAK_TFLOPS if m2l != "fp32" else FP32_PEAK_TFLOPS
        kernels["m2l"] = dict(bound="mfma", achieved=issued / (ms * 1e-3) / 1e12, peak=peak, unit="TFLOP/s",
                              per_launch=issued, ms=ms, useful_tflops=useful / (ms * 1e-3) / 1e12)
    for name, by in fmm_pass_bytes(side, zd, P, leaf).items():
        if name in prof:
            ms, _ = prof[name]
            kernels[name] = dict(bound="hbm", achieved=by / (ms * 1e-3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                                 per_launch=by, ms=ms)
    if "p2p" in prof:
        ms, _ = prof["p2p"]
        fl, peak = p2p_flops_per_eval(side, zd, leaf, m2l)
        pairs = p2p_pairs_per_eval(side, zd, leaf)
        kernels["p2p"] = dict(bound="mfma", achieved=fl / (ms * 1e-3) / 1e12, peak=peak,
                              unit="TFLOP/s", per_launch=fl, ms=ms, pairs_per_s=pairs / (ms * 1e-3))
    return kernels


def pmc_calibration_copies():
    """Device-to-device copies of known size (256 MiB, then 1 GiB, beyond the
    256 MB MALL) whose FETCH_SIZE / WRITE_SIZE readings scale the counters
    of the film kernels to bytes.  Plain HIP runtime calls through ctypes:
    hipMemset fills the source, hipMemcpy (device to device) is the copy."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    for nbytes in (256 << 20, 1 << 30):
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(src), nbytes) or hip.hipMalloc(ctypes.byref(dst), nbytes):
            raise RuntimeError("hipMalloc failed for the calibration copy")
        ok = (hip.hipMemset(src, 0, nbytes) == 0 and hip.hipDeviceSynchronize() == 0
              and hip.hipMemcpy(dst, src, nbytes, 3) == 0 and hip.hipDeviceSynchronize() == 0)
        hip.hipFree(src)
        hip.hipFree(dst)
        if not ok:
            raise RuntimeError("calibration copy failed")


def surrogate_flops_per_eval(side, zd):
    """MFMA work of one surrogate inference: lift (27 -> 32 padded k per
    output channel), three 3x3 64 -> 64 convolutions, the projection, and
    the spectral branch's four basis transforms (32 modes per direction)
    and the per-mode 64 x 64 mixing."""
    n = side * side * zd
    conv = 2 * 64 * 64 * 9 * n
    lift = 2 * 64 * 32 * n
    spec = 2 * (2 * 32 * side * 64 * side * zd) + 2 * (2 * 32 * side * 32 * 64 * zd) + 2 * 32 * 32 * 64 * 64 * zd
    return 3 * conv + lift + 2 * 64 * n + spec


def surrogate_training(mg, wl, args):
    """Training of the surrogate on its own workload: args.sur_pairs seeded
    128 x 128 x 3 film states (disc / rectangle / ring masks, vortex or domain
    states) with their FMM (P = 6) potentials generated on the GPU, the last
    512 (at most a quarter) held out; args.sur_steps Adam steps of 8 films (rate 1e-3, 3e-4 over
    the last third) over a seeded shuffle.  Returns (weights, dict)."""
    dims, B = (128, 128, 3), 8
    if args.sur_pairs < 2 * B:
        raise SystemExit("--sur-pairs must be at least 16")
    hold = min(512, (args.sur_pairs // 4) // B * B)  # held-out pairs, whole batches
    t = mg.SurrogateTrainer(*dims, wl.CELL, wl.MS, B)
    t0 = time.perf_counter()
    t.generate(args.sur_pairs, 2010, P=6)
    gen_s = time.perf_counter() - t0
    ntrain = args.sur_pairs - hold
    rng = np.random.default_rng(2010)
    held = [np.arange(ntrain + B * k, ntrain + B * k + B) for k in range(hold // B)]
    before = float(np.mean([t.eval(b) for b in held]))
    order = np.zeros(0, np.int64)
    losses = []
    t0 = time.perf_counter()
    for s in range(args.sur_steps):
        if order.size < B:
            order = np.concatenate([order, rng.permutation(ntrain)])
        idx, order = order[:B], order[B:]
        losses.append(t.step(idx, 1e-3 if s < 2 * args.sur_steps // 3 else 3e-4))
    train_s = time.perf_counter() - t0
    after = [float(np.mean(t.eval(b))) for b in held]
    w = t.weights()
    t.close()
    out = {"workload": "128x128x3 seeded states, FMM P=6 potentials", "pairs": args.sur_pairs,
           "held_out": hold, "batch": B, "steps": args.sur_steps, "generate_s": round(gen_s, 2),
           "pairs_generated_per_s": round(args.sur_pairs / gen_s, 1),
           "ms_per_step": round(1e3 * train_s / max(1, args.sur_steps), 3),
           "films_per_s": round(B * args.sur_steps / train_s, 1),
           "train_loss_last_100": round(float(np.mean(losses[-100:])), 5) if losses else None,
           "held_out_residual_before": round(before, 4), "held_out_residual": round(float(np.mean(after)), 4)}
    log(f"surrogate training: {args.sur_pairs} pairs in {gen_s:.1f} s; {args.sur_steps} steps at "
        f"{out['ms_per_step']} ms; held-out residual {before:.3f} -> {out['held_out_residual']:.4f}")
    return w, out


def surrogate_timing(mg, wl, w, args, plain_sps, weights=None):
    """The learned surrogate on the same film: one inference timed with HIP
    events beside one solver evaluation; RK4 steps/s with the surrogate at
    stages 2-3 (and the stage-1 judge evaluation) in every step; the judged
    protocol (stages 2-3 take the surrogate only when its stage-1 residual
    is within 1e-3, else the solver); the residual of the network (trained
    when weights are given).  Returns a dict for the JSON line."""
    f = make_film(mg, wl, w, args.P, args.m2l, args.leaf)
    f.surrogate_init(weights)
    ch = wl.charge_of(w.M, w.dims, w.d)
    f.profile_begin()
    for _ in range(3):
        f.surrogate_potential(ch)
        f.potential(ch)
    prof = f.profile_end()
    sur_ms = prof["surrogate"][0]
    fmm_ms = sum(v[0] * v[1] for k, v in prof.items() if k not in ("surrogate", "div", "stage")) / 3.0
    f.surrogate_mode(1)
    sps23, _ = timed_steps(f, w, 2, max(5, args.steps // 4), bool(args.graph))
    res = f.surrogate_stats()["last_residual"]
    f.surrogate_mode(2, 1e-3)
    spsj, _ = timed_steps(f, w, 1, 5, False)
    st = f.surrogate_stats()
    f.close()
    fl = surrogate_flops_per_eval(w.xdim, w.zdim)
    out = {"trained": weights is not None, "inference_ms": round(sur_ms, 4), "solver_eval_ms": round(fmm_ms, 4),
           "inference_tflops": round(fl / (sur_ms * 1e-3) / 1e12, 2), "inference_flops": fl,
           "inference_frac_bf16_peak": round(fl / (sur_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
           "steps_per_s_stages23": round(sps23, 3), "ratio_stages23": round(sps23 / plain_sps, 4),
           "steps_per_s_judged": round(spsj, 3), "ratio_judged": round(spsj / plain_sps, 4),
           "stage1_residual": res, "judged_steps": st["judged"], "accepted_steps": st["accepted"]}
    log(f"surrogate: inference {sur_ms:.3f} ms vs solver {fmm_ms:.3f} ms; stages 2-3 {sps23:.1f} steps/s "
        f"({sps23 / plain_sps:.3f}x), judged {spsj:.1f} steps/s, residual {res:.3g}")
    return out


VALU_FP32_PEAK_TFLOPS = 78.6  # non-packed FP32 FMA (packed FP32 is not used by this build)


STREAM_CODE = r"""
import json, torch
n = 256 * 1024 * 1024
a = torch.ones(n, dtype=torch.float32, device="cuda")
b = torch.ones_like(a)
c = torch.empty_like(a)
out = {}
for name, fn, nbytes in (("copy", lambda: c.copy_(a), 8 * n), ("add", lambda: torch.add(a, b, out=c), 12 * n)):
    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(16):
        fn()
    e1.record()
    torch.cuda.synchronize()
    out[name + "_GBs"] = round(nbytes / (e0.elapsed_time(e1) / 16 * 1e-3) / 1e9, 1)
print(json.dumps(out))
"""


def hbm_stream_rates():
    """Practical HBM rates of plain streaming kernels on this GPU, for
    context beside the published 8 TB/s: PyTorch's device copy (c = a) and
    add (c = a + b) over 1 GiB FP32 arrays, bytes moved / time (HIP events,
    16 launches after 3 warm-up launches).  Run in a child process: PyTorch
    brings its own HIP runtime, which does not share a process with the
    library's."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, "-c", STREAM_CODE], capture_output=True, text=True, timeout=180)
        return json.loads(r.stdout.strip().splitlines()[-1])
    except (subprocess.SubprocessError, ValueError, IndexError) as e:
        log(f"HBM streaming rates not measured: {e}")
        return None


def bench_c5(args, mg, wl):
    """512 independent films (16^2..96^2 x 3, discs / squares / U shapes,
    field ramp -50..+50 mT): direct sum below 4096 cells, FMM P = 4 above,
    advanced together one fixed-dt RK4 step at a time.  value = batch RK4
    steps/s (every film advances one step per batch step)."""
    films = wl.c5_films()
    mat = np.concatenate([f.material for f in films])
    M0 = np.concatenate([f.M for f in films])
    prec = prec_of(mg, args.m2l)
    b = mg.FilmBatch(films, 3, *films[0].d, solver=mg.SOLVER_FMM, P=4, m2l_precision=prec, material=mat,
                     mu_0=wl.MU_0, Ms=wl.MS, Aexch=wl.AEXCH, alfa=wl.ALFA, gamma=wl.GAMMA)
    info = b.info()
    b.set_m(M0)
    tune = b.autotune(reps=2)
    log(f"c5: {info}; autotune: {tune.strip()}")
    dt = films[0].dt
    b.steps(args.warmup, dt, use_graph=bool(args.graph))
    b.sync()
    t0 = time.perf_counter()
    b.steps(args.steps, dt, use_graph=bool(args.graph))
    b.sync()
    elapsed = time.perf_counter() - t0
    sps = args.steps / elapsed
    nprof = max(2, min(args.steps, 5))
    b.profile_begin()
    b.steps(nprof, dt, use_graph=False)
    prof = b.profile_end()
    per_step = {k: v[0] * v[1] / nprof for k, v in prof.items()}
    log(f"c5: {sps:.2f} batch steps/s; per-step kernel ms:", {k: round(v, 4) for k, v in per_step.items()})
    M = b.get_m().reshape(-1, 3)
    b.close()
    mag_err = float(np.abs(np.linalg.norm(M[mat == 1].astype(np.float64), axis=1) / float(wl.MS) - 1).max())
    ncell = int(info["cells"])
    kernels = {}
    if "stage" in prof:
        # the films carry material masks: every stage reads the material byte
        ms, _ = prof["stage"]
        by = sum(stage_bytes_per_cell(st, False) for st in range(4)) / 4 * ncell
        kernels["stage"] = dict(bound="hbm", achieved=by / (ms * 1e-3) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s", ms=ms)
    if "direct" in prof:
        # the small films' direct sum reads 1/r from a per-film table: per
        # pair one FMA (2 flops), one table read from LDS, one integer
        # absolute difference for its address
        ms, _ = prof["direct"]
        pairs = sum(f.n * (f.n - 1) for f in films if f.solver == "direct")
        kernels["direct"] = dict(bound="valu", achieved=2 * pairs / (ms * 1e-3) / 1e12, peak=VALU_FP32_PEAK_TFLOPS,
                                 unit="TFLOP/s", ms=ms, pairs_per_s=pairs / (ms * 1e-3))
    dom = max(kernels, key=lambda k: per_step.get(k, 0.0))
    k = kernels[dom]
    # the FMM classes mix padded sizes and tuned leaf widths: their kernels are
    # reported by time only, so the roofline names the largest kernel with a
    # per-unit work figure and the largest kernel overall
    roof = {"kernel": dom, "bound": k["bound"], "achieved": round(k["achieved"], 3), "peak": k["peak"],
            "unit": k["unit"], "frac": round(k["achieved"] / k["peak"], 4), "traffic": None,
            "mean_launch_ms": round(k["ms"], 5),
            "dominant_by_time": max(per_step, key=per_step.get)}
    cpu = None
    if not args.no_cpu:
        import oracle
        sample = films[::16]
        t_eval = 0.0
        for f in sample:
            ch = oracle.divergence_3d(f.M, f.dims, f.d)
            t0 = time.perf_counter()
            oracle.calc_potential_exact(ch, f.dims, f.d)
            t_eval += time.perf_counter() - t0
        t_step = 4 * t_eval * len(films) / len(sample)
        cpu = {"value": 1.0 / t_step, "unit": "batch steps/s", "cores": 1, "kind": "port",
               "sample": (f"{len(sample)} of {len(films)} films (every 16th): one CPU direct-sum potential "
                          f"(calc_potential_exact restated in oracle/mms_oracle.c, single thread) per film in "
                          f"{t_eval:.1f} s, scaled by {len(films)}/{len(sample)} films and 4 evaluations per step; "
                          f"the reference's fmm_calc is not used because at P=4 on 5 nm cells it returns NaN")}
        log(f"c5 cpu baseline {cpu['value']:.3g} batch steps/s")
    out = {
        "metric": "batched many-film sweep: RK4 steps/s of 512 films advanced together",
        "value": round(sps, 3), "unit": "batch steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (FP64 RK4 combine)", "data": "synthetic (seeded SplitMix64, seed 2010)",
        "config": {"workload": "c5_films_512", "films": len(films), "cells": ncell,
                   "direct_films": info["direct_films"], "fmm_films": info["fmm_films"],
                   "fmm_size_classes": info["fmm_size_classes"], "P": 4, "m2l": args.m2l, "dt": float(dt),
                   "graph": bool(args.graph),
                   "film_steps_per_s": round(sps * len(films), 1)},
        "roofline": roof,
        "kernel_ms_per_step": {k: round(v, 4) for k, v in per_step.items()},
        "max_abs_m_norm_error": mag_err,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--m2l", default="fp16x2_f16x2near", choices=PREC_NAMES,
                    help="operand form; the default is the fastest form inside the workload's precision split")
    ap.add_argument("--leaf", type=int, default=8)
    ap.add_argument("--P", type=int, default=6)
    ap.add_argument("--side", type=int, default=1024)
    ap.add_argument("--probes", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the timing of the other M2L precisions")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-surrogate legs")
    ap.add_argument("--sur-pairs", type=int, default=20000, help="surrogate training pairs (128x128x3)")
    ap.add_argument("--sur-steps", type=int, default=2500, help="surrogate Adam steps of 8 films (0: untrained)")
    ap.add_argument("--workload", default="c3", choices=["c3", "c5"],
                    help="c3: the 1024x1024x4 film (default); c5: the 512-film sweep")
    ap.add_argument("--serial", action="store_true",
                    help="run every kernel in order on one stream (MMS_GPU_SERIAL=1), e.g. under a kernel trace")
    ap.add_argument("--no-stream", action="store_true", help="skip the plain streaming-kernel HBM rates")
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="end with plain device copies of 256 MiB and 1 GiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.serial:
        os.environ["MMS_GPU_SERIAL"] = "1"
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path "
              "(films, layers or subtrees would be sharded across processes; none is implemented)",
              file=sys.stderr)
        sys.exit(2)

    import mms_gpu as mg
    from mms_gpu import workloads as wl
    if args.workload == "c5":
        return bench_c5(args, mg, wl)
    w = wl.c3_domains(side=args.side, zd=4)
    log(f"workload {w.name}: {w.n} cells, FMM P={args.P} leaf={args.leaf} m2l={args.m2l}")
    f = make_film(mg, wl, w, args.P, args.m2l, args.leaf)
    steps_per_s, elapsed = timed_steps(f, w, args.warmup, args.steps, bool(args.graph))
    log(f"timed {args.steps} steps in {elapsed:.3f} s -> {steps_per_s:.2f} steps/s")

    # per-kernel HIP-event durations over a profiled window of the same steps
    nprof = max(2, min(args.steps, 5))
    f.profile_begin()
    f.steps(nprof, w.dt, use_graph=False)
    prof = f.profile_end()
    per_step = {k: v[0] * v[1] / nprof for k, v in prof.items()}
    log("per-step kernel ms:", {k: round(v, 4) for k, v in per_step.items()})
    m = f.get_m().reshape(-1, 3).astype(np.float64)
    mag_err = float(np.abs(np.linalg.norm(m, axis=1) / float(wl.MS) - 1).max())

    side, zd = args.side, 4
    kernels = kernel_rooflines(prof, side, zd, args.P, args.leaf, args.m2l)
    dom = max(kernels, key=lambda k: per_step.get(k, 0.0)) if kernels else None
    roof = None
    if dom:
        k = kernels[dom]
        roof = {"kernel": dom, "bound": k["bound"], "achieved": round(k["achieved"], 3), "peak": k["peak"],
                "unit": k["unit"], "frac": round(k["achieved"] / k["peak"], 4),
                "traffic": pmc_traffic(dom, f"c3_leaf{args.leaf}_{args.m2l}"),
                "mean_launch_ms": round(k["ms"], 5)}
    others = {}
    for name, k in kernels.items():
        e = {"achieved": round(k["achieved"], 3), "unit": k["unit"], "frac": round(k["achieved"] / k["peak"], 4),
             "mean_launch_ms": round(k["ms"], 5)}
        if "pairs_per_s" in k:
            e["pairs_per_s"] = float(f"{k['pairs_per_s']:.4g}")
        if "useful_tflops" in k:
            e["useful_tflops"] = round(k["useful_tflops"], 3)
        others[name] = e

    err = None
    done = args.warmup + args.steps + nprof
    if not args.no_error:
        e_l2, e_max, npr = field_error(mg, f, w, args.probes)
        log(f"demag field error vs FP64 after {done} steps: rel L2 {e_l2:.2e}, max {e_max:.2e} on {npr} probes")
        # the gate along the C3 run (500 steps of 5e-14 s): the error grows
        # as the film relaxes for some operand forms
        run = [{"step": done, "rel_l2": e_l2, "rel_max": e_max}]
        for target in RUN_CHECKPOINTS:
            if target > done:
                f.steps(target - done, w.dt, use_graph=bool(args.graph))
                done = target
                a, b, _ = field_error(mg, f, w, args.probes)
                run.append({"step": done, "rel_l2": a, "rel_max": b})
                log(f"demag field error vs FP64 after {done} steps: rel L2 {a:.2e}, max {b:.2e}")
        err = {"demag_field_rel_l2": e_l2, "demag_field_rel_max": e_max, "probes": npr, "oracle": "FP64 direct sum (device)",
               "along_run": run, "worst_rel_l2": max(r["rel_l2"] for r in run),
               "worst_rel_max": max(r["rel_max"] for r in run),
               "gate_1e-3": bool(max(r["rel_l2"] for r in run) <= 1e-3)}
    f.close()

    # the other M2L input precisions on the same workload (short timing + error)
    forms = {args.m2l: {"steps_per_s": round(steps_per_s, 3), "conforming": args.m2l in CONFORMING,
                        "field_rel_l2": err["demag_field_rel_l2"] if err else None,
                        "field_rel_l2_step_500": err["along_run"][-1]["rel_l2"] if err else None}}
    if not args.no_forms:
        for m2l in PREC_NAMES:
            if m2l == args.m2l:
                continue
            g = make_film(mg, wl, w, args.P, m2l, args.leaf)
            sps, _ = timed_steps(g, w, 2, 10, bool(args.graph))
            e = e500 = None
            if not args.no_error:
                e = field_error(mg, g, w, args.probes)[0]
                g.steps(RUN_CHECKPOINTS[-1] - 12, w.dt, use_graph=bool(args.graph))
                e500 = field_error(mg, g, w, args.probes)[0]
            g.close()
            forms[m2l] = {"steps_per_s": round(sps, 3), "conforming": m2l in CONFORMING, "field_rel_l2": e,
                          "field_rel_l2_step_500": e500}
            log(f"form m2l={m2l}: {sps:.2f} steps/s, field rel L2 {e} (after 12 steps), {e500} (after 500)")

    sur = None
    if not args.no_surrogate:
        weights, training = None, None
        if args.sur_steps > 0:
            weights, training = surrogate_training(mg, wl, args)
        sur = surrogate_timing(mg, wl, w, args, steps_per_s, weights)
        sur["training"] = training

    cpu = None
    if not args.no_cpu:
        import oracle
        ch = oracle.divergence_3d(w.M, w.dims, w.d)
        cpu = cpu_baseline(w, ch)
        if cpu:
            log(f"cpu baseline {cpu['value']:.3g} steps/s ({cpu['sample']})")

    if args.pmc_calibration:
        pmc_calibration_copies()

    stream = None
    if not args.no_stream:
        stream = hbm_stream_rates()
    if stream:
        stream["note"] = ("plain streaming kernels (PyTorch device copy and add over 1 GiB FP32 arrays): the "
                          "practical HBM rate beside the published 8000 GB/s peak")
        if roof and roof["unit"] == "GB/s":
            roof["frac_of_stream_add"] = round(roof["achieved"] / stream["add_GBs"], 4)
        log(f"HBM streaming rates: copy {stream['copy_GBs']} GB/s, add {stream['add_GBs']} GB/s")

    out = {
        "metric": "LLG RK4 steps/s at 1024²×4 cells, surrogate+FMM vs plain FMM at ≤1e-3 field err",
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": FORM_DTYPE[args.m2l],
        "data": "synthetic (seeded SplitMix64 domain pattern, seed 2010)",
        "config": {"workload": w.name, "form": "plain GPU FMM (the metric's denominator; no surrogate)",
                   "conforming_precision": args.m2l in CONFORMING, "cells": w.n, "P": args.P,
                   "near_field_stream": "serial" if args.serial else "concurrent",
                   "leaf_width": args.leaf, "m2l": args.m2l, "dt": float(w.dt), "graph": bool(args.graph),
                   "precision_split": "FP32 fields and stage M, FP64 RK4 combine and renormalisation"},
        "surrogate_ratio": sur["ratio_judged"] if sur else None,
        "surrogate_note": ("the metric's ratio: surrogate+FMM steps/s (stage-1 judge, fallback to the FMM when "
                           "the residual exceeds 1e-3) over plain-FMM steps/s, with the network trained in this "
                           "run (`surrogate.training`); the judge rejects it when its residual on this film is "
                           "above 1e-3 and the step falls back; `surrogate.ratio_stages23` is the ratio with the "
                           "surrogate used at stages 2-3 unconditionally (an upper bound on any network of this "
                           "size, whatever its accuracy); `value` is the plain GPU FMM form"),
        "surrogate": sur,
        "roofline": roof,
        "hbm_stream": stream,
        "kernels": others,
        "kernel_ms_per_step": {k: round(v, 4) for k, v in per_step.items()},
        "field_error": err,
        "m2l_forms": forms,
        "max_abs_m_norm_error": mag_err,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
