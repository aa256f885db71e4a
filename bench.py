"""Throughput of the BROMOC BD step path on one MI355X.

Workload (default): the large translocation system -- a 1000-bp CG-DNA (6000 sites)
threaded through a 15 A pore in a 200 x 200 x 600 A box with 50,000 K+/Cl- ions (56,000
particles), 321 x 321 x 641 repulsion, static-field and reaction-field grids at 0.625 A,
D(z) from a PROFILE spline, GCMC bath buffers at both box ends; FP64, Philox noise.  A step
is ENERGY + DYNAMICS2 + DYNAMICS0NUC over every moving particle; every 10 steps GRAND
makes 20 insertion / deletion attempts, and each 10-step call returns its last step's
energies, as the drop-in's BD call does.

Prints one JSON line: particle-steps/s over the timed steps, the roofline of the dominant
kernel (its duration from HIP events on its own stream), and the CPU baseline: the
reference Fortran routines (oracle/_ref) at one thread and with the reference's OpenMP
ENERGY on all cores, timed on this host on the largest sample of the workload the
reference's fixed arrays hold, scaled to the full system by its pair count.  The other
workloads (the 2300-particle pore system, replica batches, the adjoint) are extra keys.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s + % HBM roofline, batched vs plain, FP64 rel err<=1e-10"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per launch of the dominant kernel, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch) on each workload and launch shape (the pore_dna step kernel
# with BROMOC_GRAPH=0, so that its launches are counted outside the graphs), scaled by what
# the same counters read for a 256 MiB device copy (bench.py --copy-probe 256): FETCH_SIZE
# x 2.0, WRITE_SIZE x 1.0.  The counters cannot be read from inside this process; a shape
# without a measurement reports null.
TRAFFIC_KIB = {("translocation", 18): (378440.4, 732727.8), ("pore_dna", 1): (3541.9, 215.0)}
# L2 hit rate of the same launches, TCC_HIT_sum / (TCC_HIT_sum + TCC_MISS_sum) from the
# WRITE_SIZE run's two L2 counters
L2_HIT = {("translocation", 18): 0.777, ("pore_dna", 1): 0.930}
TRAFFIC_SCALE = (2.0, 1.0)
# FP64 operations per pair evaluation: a table pair ~34 (r^2 8, sqrt 1, index 2, U 4, F/r 3,
# energy 1, force update on both particles ~15); the reaction-field cross term of an ion
# pair ~30 (reff and srfe products 2, 1/sqrt 1, rfcf, aux0-aux3 9, energy 2, the srfe and
# reff variation of both ions 4, the radial force on both 12)
FLOP_PAIR, FLOP_RF = 34.0, 30.0
TPB_NAMES = {1: "k_step, 1 target per block", 2: "k_step, 2 targets per block",
             4: "k_step, 4 targets per block", 8: "k_wave, 1 target per wavefront",
             17: "tile mode, tiles of 64", 18: "tile mode, tiles of 128",
             20: "tile mode, tiles of 256"}


def step_bytes(s, injected):
    """Algorithmic HBM bytes per particle-step: read x,y,z (24) and the type word (4), write
    x,y,z (24), 8 trilinear corners of each grid present (phix float32: 32, phiv int8: 8, the
    reaction-field pair gsrfen + greff: 64), injected Gaussians (24).  Partner positions, pair
    tables, DNA lists and type tables are re-read from L2 and are not HBM traffic."""
    from bromoc_amd import abi
    b = 24 + 4 + 24
    if s.flags & abi.QPHIX:
        b += 32
    if s.flags & abi.QPHIV:
        b += 8
    if s.flags & abi.QRFPAR:    # 8 corners of gsrfen and greff (float32 each), per ion
        b += 64
    if injected:
        b += 24
    return b


def rf_pairs(s, pos):
    """Ion pairs whose reaction-field cross term counts (rfpar.f90:177-223: the lower ion
    has srfe != 0 and the higher one moves inside the grid box), taking srfe != 0 as inside
    the grid box (true of grids without zero nodes): C(m, 2) + f m for m moving and f fixed
    ions inside the box at positions pos."""
    import numpy as np
    shp = s.rf.data.shape[-3:]
    cen = np.array([s.rf.xbcen, s.rf.ybcen, s.rf.zbcen])
    tran = 0.5 * (np.array(shp) - 1) * s.rf.dcel
    inside = np.all(np.abs(pos - cen) <= tran, axis=1)
    idx = np.arange(len(pos))
    m = int(np.sum(inside & (idx >= s.nsites + s.nfix)))
    f = int(np.sum(inside & (idx >= s.nsites) & (idx < s.nsites + s.nfix)))
    return m * (m - 1) // 2 + f * m


def pair_counts(s, ntot=None, pos=None):
    """Pair evaluations of one ENERGY as the reference makes them (each pair once): ion-ion
    pairs with a moving ion (energy.mono.f90:120-123), site-ion pairs (:586-588), the
    reaction-field cross terms that count (rf_pairs, at the positions pos, by default the
    system's own), and the CG-DNA list terms."""
    import numpy as np
    from bromoc_amd import abi
    n = s.ntot if ntot is None else ntot
    ni, nf = n - s.nsites, s.nfix
    ii = ni * (ni - 1) // 2 - nf * (nf - 1) // 2
    si = s.nsites * ni
    rf = 0
    if s.flags & abi.QRFPAR:
        if pos is None:
            pos = np.stack([s.x, s.y, s.z], 1)[:n]
        rf = rf_pairs(s, pos)
    lists = int(sum(s.count(k) for k in s.LIST_NAMES)) if s.nsites else 0
    return {"ion_ion": ii, "site_ion": si, "rf_cross": rf, "dna_list_terms": lists}


def fp64_rate(pc, ms):
    flops = FLOP_PAIR * (pc["ion_ion"] + pc["site_ion"]) + FLOP_RF * pc["rf_cross"]
    t = flops / (ms * 1e-3) / 1e12
    return {"achieved_tflops": t, "peak_tflops": FP64_PEAK_TFLOPS, "frac": t / FP64_PEAK_TFLOPS,
            "pairs_per_step": pc, "flops_per_step": flops}


def roofline(workload, tpb, pc, kernel_ms, bpp, nmov):
    """Roofline of the dominant kernel (all pair work of a step is in it): FP64-bound, so
    achieved = the algorithmic FP64 operations of one launch (each pair once, at FLOP_PAIR /
    FLOP_RF) over its event-timed duration, against MI355X's FP64 peak; the HBM figures
    beside it (bpp algorithmic bytes per moving particle) and, when counted, traffic = the
    HBM bytes of one launch from FETCH_SIZE / WRITE_SIZE, its rate over the same duration,
    and the launch's L2 hit rate."""
    f = fp64_rate(pc, kernel_ms)
    traffic = None
    if TRAFFIC_KIB.get((workload, tpb)):
        fk, wk = TRAFFIC_KIB[(workload, tpb)]
        traffic = 1024.0 * (fk * TRAFFIC_SCALE[0] + wk * TRAFFIC_SCALE[1])
    hbm = bpp * nmov / (kernel_ms * 1e-3) / 1e9
    traffic_gbs = traffic / (kernel_ms * 1e-3) / 1e9 if traffic is not None else None
    return {"bound": "mfma", "achieved": f["achieved_tflops"], "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": f["frac"], "traffic": traffic,
            "pipe": "FP64 VALU (pair math is not GEMM-shaped; MI355X's FP64 vector and "
                    "matrix peaks are both 78.6 TFLOP/s)",
            "kernel": "k_tiles" if tpb > 16 else "k_step", "kernel_ms": kernel_ms,
            "launch_shape": TPB_NAMES.get(tpb, str(tpb)), "targets_per_block": tpb,
            "flops_per_launch": f["flops_per_step"], "pairs_per_launch": pc,
            "hbm": {"achieved_gbs": hbm, "peak_gbs": HBM_PEAK_GBS, "frac": hbm / HBM_PEAK_GBS,
                    "bytes_per_particle_step": bpp, "algorithmic_bytes_per_launch": bpp * nmov,
                    "traffic_gbs": traffic_gbs, "l2_hit_rate": L2_HIT.get((workload, tpb))}}


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def time_reference(s, lib, target_s, min_steps=1):
    """Seconds per BD step of the reference routines (or the C oracle) on system s."""
    import numpy as np
    from bromoc_amd import abi
    from oracle.oracle import Oracle, RefFortran
    eng = Oracle(s) if lib == "port" else (RefFortran(s) if lib is None else RefFortran(s, lib=lib))
    rng = np.random.default_rng(99)
    eng.upload_noise(rng.standard_normal((1, 3 * s.ntot)))
    t0 = time.perf_counter()
    eng.step(1, abi.NOISE_INJECTED, 0)
    probe = time.perf_counter() - t0
    nsteps = max(min_steps, min(2000, int(target_s / max(probe, 1e-6))))
    eng.upload_noise(rng.standard_normal((nsteps, 3 * s.ntot)))
    t0 = time.perf_counter()
    eng.step(nsteps, abi.NOISE_INJECTED, 0)
    dt = time.perf_counter() - t0
    del eng
    return dt / nsteps, nsteps, dt


def cpu_baseline(s, sample=None, target_s=10.0):
    """The reference on this host's cores: one thread (the .mono routines) and all cores
    (the OpenMP ENERGY, energy.para.f90), on `sample` -- the same workload shrunk to what
    the reference's fixed arrays hold (datom = 23000 particles, dtype = 1000 types,
    grandmod.f90:24) -- scaled to s by the pair count of one step; on s itself when no
    sample is given.  Falls back to the C oracle where the reference build is absent."""
    from oracle.oracle import REF_OMP_SO, ref_available
    base = sample or s
    kind = "reference" if ref_available() else "port"
    pc_full, pc_base = pair_counts(s), pair_counts(base)
    work = lambda pc: FLOP_PAIR * (pc["ion_ion"] + pc["site_ion"]) + FLOP_RF * pc["rf_cross"]
    scale = work(pc_full) / work(pc_base)
    nmov = s.ntot - s.nfix
    info = cpu_info()
    out = {}
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info["affinity"] or info["nproc"]
    for key, lib, cores in (("one_thread", None if kind == "reference" else "port", 1),
                            ("all_cores", REF_OMP_SO if kind == "reference" else None, threads)):
        if key == "all_cores" and kind != "reference":
            continue
        per_step, nsteps, dt = time_reference(base, lib, target_s)
        full_step = per_step * scale
        out[key] = {"value": nmov / full_step, "unit": "particle-steps/s", "cores": cores,
                    "kind": kind, "s_per_step_sample": per_step,
                    "sample": (f"{nsteps} BD steps of a {base.ntot}-particle sample "
                               f"({base.nsites} sites) of the workload, {dt:.1f} s on {cores} "
                               f"thread(s), scaled x{scale:.2f} by the pair work of one step"
                               if sample else
                               f"{nsteps} BD steps of the same {base.ntot}-particle system, "
                               f"{dt:.1f} s on {cores} thread(s)")}
    res = dict(out["one_thread"])
    res.update(info)
    if "all_cores" in out:
        res["all_cores"] = out["all_cores"]
    return res


def host_roundtrip(eng, nmov, nsteps=200):
    """Rate of the per-routine call pattern a Fortran driver uses without the device-resident
    loop: upload positions, ENERGY, DYNAMICS, download positions and forces, every step."""
    from bromoc_amd import abi
    pos, typ = eng.positions()
    t0 = time.perf_counter()
    for k in range(nsteps):
        eng.set_positions(pos, typ)
        eng.energy(forces=True)
        eng.forces()
        eng.dynamics(abi.NOISE_PHILOX, 0, seed=7, counter=k)
        pos, typ = eng.positions()
    dt = time.perf_counter() - t0
    return {"value": nmov * nsteps / dt, "unit": "particle-steps/s", "steps": nsteps,
            "pattern": "set_positions + energy + get_forces + dynamics + get_positions per step"}


def gcmc_cycles(e, s, nsteps, g, nat, prob, rng, counter0, every=10, ngcmc=20, cycle0=0,
                ener=0.0):
    """nsteps BD steps in cycles of GRAND (ngcmc attempts) + `every` steps, each step call
    returning its last step's energies.  Returns (ener, counter, moving-particle steps)."""
    from bromoc_amd import abi
    done, ctr, cyc, psteps = 0, counter0, cycle0, 0
    while done < nsteps:
        _, ener = e.grand(ngcmc, rng.random(6 * ngcmc + 8), nat, prob, cyc + 1, ener)
        n = min(every, nsteps - done)
        e.step(n, abi.NOISE_PHILOX, seed=11, counter0=ctr, energies=True)
        psteps += n * (e.ntot - s.nfix)
        ctr += n
        done += n
        cyc += 1
    return ener, ctr, psteps


def translocation_leg(steps, warmup, cpu=True, parity=True):
    """The headline workload (see the module docstring)."""
    import numpy as np
    from bromoc_amd import abi, configs
    from bromoc_amd.engine import Engine
    t0 = time.perf_counter()
    s = configs.translocation()
    build_s = time.perf_counter() - t0
    g = s.meta["gcmc"]
    ib, nat = configs.assign_buffers(s, g)
    e = Engine(s)
    par = None
    if parity:
        from oracle.oracle import Oracle
        o = Oracle(s)
        eg, eo = e.energy(), o.energy()
        par = {"force_normwise_relerr": float(np.max(np.abs(e.forces() - o.forces())) /
                                              np.max(np.abs(o.forces()))),
               "energy_relerr": abs(eg["ener"] - eo["ener"]) / abs(eo["ener"])}
        del o
    e.set_gcmc(g, ib)
    prob = np.zeros((g.nbuffer, s.capacity + 1), np.int32)
    rng = np.random.default_rng(5)
    tpb = e.tuning()
    ener, ctr, _ = gcmc_cycles(e, s, max(warmup, 1), g, nat, prob, rng, 0)
    ncyc0 = (max(warmup, 1) + 9) // 10
    t0 = time.perf_counter()
    ener, ctr, psteps = gcmc_cycles(e, s, steps, g, nat, prob, rng, ctr, cycle0=ncyc0, ener=ener)
    elapsed = time.perf_counter() - t0
    nk = 10
    total_ms, kernel_ms = e.time_steps(nk, abi.NOISE_PHILOX, seed=11, counter0=ctr)
    # the step without GRAND, for the cost of the GCMC cycles
    t0 = time.perf_counter()
    e.step(steps, abi.NOISE_PHILOX, seed=11, counter0=ctr + nk, energies=True)
    plain_ms = 1e3 * (time.perf_counter() - t0) / steps
    n_now = e.ntot
    pos_now = e.positions()[0]
    e.close()
    pc = pair_counts(s, n_now, pos_now)
    roof = roofline("translocation", tpb, pc, kernel_ms, step_bytes(s, injected=False),
                    n_now - s.nfix)
    wl = {"workload": "translocation", "ions": s.ntot - s.nsites, "dna_sites": s.nsites,
          "dna_bp": s.meta["n_bp"], "box_A": [200, 200, 600], "grid": list(s.meta["grid"]),
          "grid_spacing_A": s.meta["dcel"], "rfpar": True, "diffusion": "profile",
          "pair_tables": s.npair, "dt_ps": s.meta["dt"], "noise": "philox",
          "gcmc": "20 GRAND attempts every 10 steps", "energies": "last step of each call"}
    out = {"value": psteps / elapsed, "elapsed_s": elapsed, "ms_per_step": 1e3 * elapsed / steps,
           "config": wl, "roofline": roof, "step_ms_without_gcmc": plain_ms,
           "step_kernels_ms": total_ms / nk, "particles": n_now, "setup_s": build_s,
           "parity_vs_oracle": par}
    if cpu:
        sample = configs.translocation(n_bp=300, n_ions=20000, gcmc=False, capacity=1)
        sample.capacity = sample.ntot
        out["cpu_baseline"] = cpu_baseline(s, sample=sample)
    return out


def pore_dna_leg(steps=2000, warmup=200, cpu=True):
    """The 2300-particle pore workload: 2000 K+/Cl- ions and a 50-bp CG-DNA (300 sites) in a
    100 x 100 x 200 A box, 201 x 201 x 401 repulsion and static-field grids at 0.5 A, 15 EFPOT
    tables, Philox noise, no GCMC."""
    import numpy as np
    from bromoc_amd import abi, configs
    from bromoc_amd.engine import Engine
    s = configs.pore_dna(seed=1)
    nmov = s.nsites + (s.ntot - s.nsites - s.nfix)
    e = Engine(s)
    e.step(warmup, abi.NOISE_PHILOX, seed=2024, counter0=0, energies=False)
    tpb = e.tuning()
    t0 = time.perf_counter()
    e.step(steps, abi.NOISE_PHILOX, seed=2024, counter0=warmup, energies=True)
    el = time.perf_counter() - t0
    _, kms = e.time_steps(200, abi.NOISE_PHILOX, seed=2024, counter0=warmup + steps)
    host = host_roundtrip(e, nmov)
    e.close()
    out = {"value": nmov * steps / el, "unit": "particle-steps/s", "ms_per_step": 1e3 * el / steps,
           "steps": steps, "ions": s.ntot - s.nsites, "dna_sites": s.nsites,
           "roofline": roofline("pore_dna", tpb, pair_counts(s), kms,
                                step_bytes(s, injected=False), nmov),
           "host_roundtrip": host}
    if cpu:
        out["cpu_baseline"] = cpu_baseline(s, target_s=8.0)
    return out


def batch_legs(steps=300, gcmc_every=10, ngcmc=20, small_n=200, small_steps=400):
    """Replica batching: (1) 256 independent ion + pore systems (log-uniform 100..4000 ions,
    own 129x129x257 static field at 0.75 A, shared repulsion map, GCMC bath buffers at both
    ends) advanced together with GRAND every `gcmc_every` steps; (2) at small N, 256
    replicas of `small_n` ions against the single-system path on one of them."""
    import numpy as np
    from bromoc_amd import abi, configs
    from bromoc_amd.engine import Batch, Engine
    from bromoc_amd.system import replica_system
    t0 = time.perf_counter()
    sh, reps = configs.replica_batch(seed=3)
    gen_s = time.perf_counter() - t0
    b = Batch(sh, reps)
    b.run(20, seed=7, counter0=0, gcmc_every=gcmc_every, ngcmc=ngcmc)
    nmov = sum(b.get(r)[2] - rep.nfix for r, rep in enumerate(reps))
    t0 = time.perf_counter()
    b.run(steps, seed=7, counter0=20, gcmc_every=gcmc_every, ngcmc=ngcmc, cycle0=2)
    el = time.perf_counter() - t0
    nmov_end = sum(b.get(r)[2] - rep.nfix for r, rep in enumerate(reps))
    _, kms = b.time_steps(10, seed=7, counter0=10000)
    full = {"value": 0.5 * (nmov + nmov_end) * steps / el, "unit": "particle-steps/s",
            "replicas": len(reps), "ions": [int(nmov), int(nmov_end)],
            "ion_range": [min(r.ntot for r in reps), max(r.ntot for r in reps)],
            "steps": steps, "gcmc_every": gcmc_every, "ngcmc": ngcmc,
            "ms_per_step": 1e3 * el / steps, "step_kernel_ms": kms,
            "launch_shape": dict(zip(("blocks", "lanes_per_target"), b.tuning())),
            "grid": list(sh.phix.data.shape),
            "setup_s": gen_s}
    b.close()
    del b
    sh2, reps2 = configs.replica_batch(seed=4, n_range=(small_n, small_n), gcmc=False)
    e = Engine(replica_system(sh2, reps2[0]))
    e.step(50, abi.NOISE_PHILOX, seed=1, counter0=0, energies=False)
    t0 = time.perf_counter()
    e.step(small_steps, abi.NOISE_PHILOX, seed=1, counter0=50, energies=False)
    plain = reps2[0].ntot * small_steps / (time.perf_counter() - t0)
    e.close()
    b2 = Batch(sh2, reps2)
    b2.run(20, seed=1)
    t0 = time.perf_counter()
    b2.run(small_steps, seed=1, counter0=20)
    batched = sum(r.ntot for r in reps2) * small_steps / (time.perf_counter() - t0)
    small = {"n": small_n, "replicas": len(reps2), "plain": plain, "batched": batched,
             "ratio": batched / plain, "unit": "particle-steps/s", "steps": small_steps}
    return full, small


def adjoint_leg(nsteps=400, checkpoint_every=20, reps=3):
    """The differentiation workload: 200 K+ + 200 Cl- + a fixed wall site in a 60 A box,
    4 x 1851 knot parameters, 400 BD steps of fixed injected noise checkpointed every 20,
    smoothed-RDF loss against a second seed's RDFs.  Times the forward pass alone and the
    full gradient (forward with checkpoints + reverse pass with recomputation + chain rule
    through the table build); a central-difference gradient would take 2 forward passes
    per knot (14,808 passes)."""
    import numpy as np
    from bromoc_amd import configs
    from bromoc_amd.adjoint import RdfSpec, rdf_loss
    from bromoc_amd.engine import Engine
    spec = RdfSpec(groups=[(1, 1), (1, 2), (2, 2), (1, 3)])
    s = configs.adjoint_system(seed=10, n_per_type=200, box=60.0)
    noise = np.random.default_rng(11).standard_normal((nsteps, 3 * s.ntot))
    t = configs.with_knots(s, {is_: u for is_, (u, _, _) in configs.adjoint_system(
        seed=10, n_per_type=200, box=60.0, curve_seed=12).meta["params"].items()})
    target = rdf_loss(t, noise, nsteps, spec, grad=False)[1]
    e = Engine(s)
    rdf_loss(s, noise, nsteps, spec, target=target, grad=True, engine=e)     # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        rdf_loss(s, noise, nsteps, spec, target=target, grad=False, engine=e)
    fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        _, _, grads = rdf_loss(s, noise, nsteps, spec, target=target, grad=True,
                               checkpoint_every=checkpoint_every, engine=e)
    full = (time.perf_counter() - t0) / reps
    nknots = sum(len(g) for g in grads.values())
    e.close()
    return {"particles": s.ntot, "steps": nsteps, "checkpoint_every": checkpoint_every,
            "knot_parameters": nknots, "forward_s": fwd, "gradient_s": full,
            "gradient_over_forward": full / fwd,
            "central_difference_equivalent_s": 2 * nknots * fwd}


def copy_probe(mib):
    """A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8, "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="translocation",
                    choices=["translocation", "pore_dna", "adjoint"],
                    help="the headline workload; adjoint prints the adjoint leg alone")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-parity", action="store_true", help="skip the start-state parity check")
    ap.add_argument("--no-batch", action="store_true", help="skip the replica-batch legs")
    ap.add_argument("--no-pore", action="store_true", help="skip the 2300-particle pore leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload == "adjoint":
        print(json.dumps({"adjoint_rdf_gradient": adjoint_leg()}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "translocation":
        head = translocation_leg(args.steps, args.warmup, cpu=not args.no_cpu,
                                 parity=not args.no_parity)
        pore = None if args.no_pore else pore_dna_leg(cpu=not args.no_cpu)
    else:
        pore = pore_dna_leg(steps=args.steps, warmup=args.warmup, cpu=not args.no_cpu)
        head = dict(pore)
        head["config"] = {"workload": "pore_dna", "ions": pore["ions"],
                          "dna_sites": pore["dna_sites"], "noise": "philox"}
        head["elapsed_s"] = head["ms_per_step"] * args.steps * 1e-3
        pore = None
    batched, small_n = (None, None) if args.no_batch else batch_legs()
    adjoint = None if args.no_adjoint else adjoint_leg()
    cpu = head.pop("cpu_baseline", None)
    out = {
        "metric": METRIC, "value": head["value"], "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded positions, grids and EFPOT tables)", "config": head["config"],
        "roofline": head["roofline"], "cpu_baseline": cpu,
    }
    for k in ("step_ms_without_gcmc", "step_kernels_ms", "particles", "setup_s",
              "parity_vs_oracle", "host_roundtrip"):
        if k in head:
            out[k] = head[k]
    out.update({"pore_dna": pore, "batched": batched, "batched_vs_plain_small_n": small_n,
                "adjoint_rdf_gradient": adjoint})
    print(json.dumps(out))


if __name__ == "__main__":
    main()
