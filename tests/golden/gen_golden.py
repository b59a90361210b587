"""Write the golden fixtures under tests/golden/ (run in the build container,
where the reference sources are readable; the GPU box only reads the output).

    python tests/golden/gen_golden.py            # everything
    python tests/golden/gen_golden.py ref logs   # subsets: ref, logs, traj

ref   -> ref_vectors.npz: outputs of the REFERENCE's own functions compiled from
         its sources (oracle/_ref/libmmsref.so): divergence_3d, gradient_3d,
         add_exchange_field (vector_functions.cpp), associated_legendre and
         spherical_harmonic (numerics.cu), LLG_Mprime (ode_functions.cpp),
         neighbour/interaction lists and box centres (Box.cu) -- on seeded inputs.
logs  -> log_kats.json + nanowire4_mask.txt: the reference's own logged step
         records (log-files/*.log; data, not code) and the decoded
         test-structures/nanowire4.png mask (material = grey == 0, y = 0 at the
         bottom row as FreeImage reads it).
traj  -> c1_traj.npz: the 32x32x3 direct-sum film (seed 2010) after 1000 fixed
         1e-13 s RK4 steps through the oracle's restatement of the reference
         path (FP32, rk4_step + Hfield + calc_potential_exact CPU branch), and
         through the same restatement with FP64 stage combination.
"""
import json
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "mms-gpu_amd"))
REF = "/root/reference"


def gen_ref():
    import ctypes as C
    import oracle
    R = oracle.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libmmsref.so missing: run make -C oracle ref")
    rng = np.random.default_rng(2010)
    out = {}
    dims = (9, 7, 4)
    d = (np.float32(5e-9), np.float32(4e-9), np.float32(3e-9))
    n = int(np.prod(dims))
    M = (rng.normal(size=3 * n) * 8.6e5).astype(np.float32)
    S = np.zeros(n, np.float32)
    R.ref_divergence_3d(M, *dims, *d, S)
    V = np.zeros(3 * n, np.float32)
    cm = np.float32(1 / (4.0 * np.pi))
    R.ref_gradient_3d(S, *dims, *d, cm, V)
    H0 = (rng.normal(size=3 * n) * 1e4).astype(np.float32)
    H = H0.copy()
    R.ref_add_exchange_field(M, np.float32(8.6e5), np.float32(1.3e-11), np.float32(4e-7 * np.pi), *dims, *d, H)
    out.update(vf_dims=np.array(dims, np.int32), vf_d=np.array(d, np.float32), vf_M=M, vf_div=S,
               vf_cm=np.array([cm], np.float32), vf_grad=V, vf_H0=H0, vf_exch=H)
    xs = rng.uniform(-1, 1, 64).astype(np.float32)
    leg = np.zeros((5, 5, 64), np.float32)
    for l in range(5):
        for m in range(l + 1):
            for i, x in enumerate(xs):
                leg[l, m, i] = R.ref_associated_legendre(l, m, float(x))
    ang = rng.uniform(0, np.pi, (64, 2)).astype(np.float32)
    sph = np.zeros((5, 9, 64, 2), np.float32)
    a, b = C.c_float(), C.c_float()
    for l in range(5):
        for m in range(-l, l + 1):
            for i, (th, ph) in enumerate(ang):
                R.ref_spherical_harmonic(l, m, float(th), float(ph), C.byref(a), C.byref(b))
                sph[l, m + 4, i] = (a.value, b.value)
    out.update(leg_x=xs, leg=leg, sph_ang=ang, sph=sph)
    Ms_ = (rng.normal(size=(256, 3)) * 8.6e5).astype(np.float32)
    Hs = (rng.normal(size=(256, 3)) * 1e5).astype(np.float32)
    llg = np.zeros((256, 3), np.float32)
    o = np.zeros(3, np.float32)
    for i in range(256):
        R.ref_llg_mprime(np.ascontiguousarray(Ms_[i]), np.ascontiguousarray(Hs[i]), 0.5, 2.211e5, 8.6e5, o)
        llg[i] = o
    out.update(llg_M=Ms_, llg_H=Hs, llg_out=llg, llg_par=np.array([0.5, 2.211e5, 8.6e5], np.float32))
    limit = 5
    for level in range(0, limit + 1):
        nb = 4 ** level
        nbr = np.zeros(nb * 8, np.int32)
        il = np.zeros(nb * 27, np.int32)
        cen = np.zeros(nb * 2, np.float32)
        R.ref_tree_lists(limit, level, nbr, il, cen)
        out[f"tree_L{level}_nbr"] = nbr.reshape(nb, 8)
        out[f"tree_L{level}_il"] = il.reshape(nb, 27)
        out[f"tree_L{level}_centre"] = cen.reshape(nb, 2)
    out["tree_limit"] = np.array([limit], np.int32)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote ref_vectors.npz")


def _log_records(path):
    recs = []
    pat = re.compile(r"^\s*(\d+), ([-+0-9.e]+), ([-+0-9.e]+), ([-+0-9.e]+), ([-+0-9.e]+), ([-+0-9.e]+), "
                     r"([-+0-9.e]+), ([-+0-9.e]+), ([-+0-9.e]+)\s*$")
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            m = pat.match(line)
            if m:
                recs.append([lineno] + [float(v) for v in m.groups()])
    return recs


def gen_logs():
    from mms_gpu import geometry
    kats = {"_provenance": "step records 'tindex, t, dt, E2[eV], <Mx>/Ms, <My>/Ms, <Mz>/Ms, <|M|>/Ms, "
                           "max torque' copied from the reference's log-files (GPU direct-sum runs, "
                           "use_fmm=0, P=3, FP32); first column is the log line number"}
    cases = {
        "nanowire4_fixed_1e-13": ("sim_nanowire4_42x26x3_5nm_noHext_fixedstep_1e-13.log", 1000,
                                  dict(xdim=42, ydim=26, zdim=3, width=2.1e-7, height=1.3e-7, depth=1.5e-8,
                                       mask="nanowire4", timestep=1e-13)),
        "square32_variable": ("sim_square_variablestep_32x32x3_5nm.log", 1,
                              dict(xdim=32, ydim=32, zdim=3, width=1.6e-7, height=1.6e-7, depth=1.5e-8,
                                   mask="border", timestep=1e-13)),
        # built-in border mask, vortex (IC 1) and random (IC 2, srand(2010)) initial states
        "square32_vortex": ("sim_square_variablestep_32x32x3_5nm_ICvortex.log", 3,
                            dict(xdim=32, ydim=32, zdim=3, width=1.6e-7, height=1.6e-7, depth=1.5e-8,
                                 mask="border", timestep=1e-14, ic=1, seed=2010)),
        "square32_random": ("sim_square_variablestep_32x32x3_5nm_ICrandom.log", 3,
                            dict(xdim=32, ydim=32, zdim=3, width=1.6e-7, height=1.6e-7, depth=1.5e-8,
                                 mask="border", timestep=1e-14, ic=2, seed=2010)),
        "square16_random": ("sim_square_variablestep_16x16x3_5nm_ICrandom.log", 3,
                            dict(xdim=16, ydim=16, zdim=3, width=1.6e-7, height=1.6e-7, depth=1.5e-8,
                                 mask="border", timestep=1e-14, ic=2, seed=2010)),
    }
    for key, (fname, nkeep, meta) in cases.items():
        recs = _log_records(os.path.join(REF, "log-files", fname))
        kats[key] = dict(meta, log=fname, gamma=2.21e5, records=recs[:nkeep])
    with open(os.path.join(HERE, "log_kats.json"), "w") as f:
        json.dump(kats, f, indent=0)
    mask = geometry.material_from_png(os.path.join(REF, "test-structures", "nanowire4.png"))
    np.savetxt(os.path.join(HERE, "nanowire4_mask.txt"), mask, fmt="%d")
    print("wrote log_kats.json, nanowire4_mask.txt")


def gen_traj(nsteps=1000):
    import oracle
    from mms_gpu import workloads as wl
    w = wl.c1_square()
    res = {}
    for combine64 in (0, 1):
        t0 = time.time()
        M = w.M.copy()
        H, _, _ = oracle.hfield(M, w.dims, w.d, 0, wl.MU_0, wl.MS, wl.AEXCH)
        E = []
        for s in range(nsteps):
            M2, _ = oracle.rk4_step(M, H, w.material, w.dims, w.d, w.dt, 0, wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA,
                                    wl.GAMMA, combine64=combine64)
            H, _, _ = oracle.hfield(M2, w.dims, w.d, 0, wl.MU_0, wl.MS, wl.AEXCH)
            e, tq = oracle.energy_torque(w.material, M2, H, wl.MU_0, wl.MS, w.d)
            E.append((e, tq))
            M = M2
        tag = "f64c" if combine64 else "f32"
        res[f"M_{tag}"] = M
        res[f"Etorque_{tag}"] = np.array(E, np.float32)
        print(f"trajectory {tag}: {time.time() - t0:.1f}s", flush=True)
    res["M0"] = w.M
    res["nsteps"] = np.array([nsteps], np.int32)
    np.savez_compressed(os.path.join(HERE, "c1_traj.npz"), **res)
    print("wrote c1_traj.npz")


if __name__ == "__main__":
    what = sys.argv[1:] or ["ref", "logs", "traj"]
    if "ref" in what:
        gen_ref()
    if "logs" in what:
        gen_logs()
    if "traj" in what:
        gen_traj()
