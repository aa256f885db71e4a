"""Throughput of the BROMOC BD step path on one MI355X.

Workload (default): ions in a cylindrical pore through a slab plus a 50-bp CG-DNA duplex
-- 2000 K+/Cl- ions and 300 DNA sites in a 100x100x200 A box, 201x201x401 repulsion and
static-field grids at 0.5 A, 15 EFPOT tables (r = 1.5..20 A at 0.01 A), FP64, Philox noise.
One step = ENERGY + DYNAMICS0 + DYNAMICS0NUC over every moving particle (one kernel launch).

Prints one JSON line: particle-steps/s over the timed steps, the roofline of the step
kernel (bytes from HIP events on the kernel's own stream), and a CPU baseline: the
reference Fortran routines (oracle/_ref, one thread) -- or the C oracle where that build
is absent -- timed on this host over a bounded sample of the same workload.
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
# HBM bytes per launch of the step kernel, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (KiB per dispatch, one counter per run) on this workload and launch shape (targets per
# block the autotuner picks), scaled by what the same counters read for a 256 MiB device
# copy (bench.py --copy-probe 256): FETCH_SIZE x 2.0, WRITE_SIZE x 1.0.  The counters cannot
# be read from inside this process; a shape without a measurement reports null.
TRAFFIC_KIB = {("pore_dna", 1): (3073.5, 145.3), ("pore_dna", 2): (3198.4, 144.5)}
TRAFFIC_SCALE = (2.0, 1.0)


def step_bytes(s, injected):
    """Algorithmic HBM bytes per particle-step of the fused step kernel: read x,y,z (24)
    and the type word (4), write x,y,z (24), 8 trilinear corners of each grid present
    (phix float32: 32, phiv int8: 8, the reaction-field pair gsrfen + greff: 64), injected
    Gaussians (24).  Partner positions, pair
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


def cpu_baseline(s, target_s=12.0):
    """Time the reference routines (or the C oracle) on a bounded number of steps."""
    import numpy as np
    from bromoc_amd import abi
    try:
        from oracle.oracle import RefFortran, ref_available, Oracle
    except ImportError:
        return None
    kind = "reference" if ref_available() else "port"
    eng = RefFortran(s) if kind == "reference" else Oracle(s)
    nmov = s.nsites + (s.ntot - s.nsites - s.nfix)
    rng = np.random.default_rng(99)
    # one probe step sizes the sample to about target_s seconds
    eng.upload_noise(rng.standard_normal((1, 3 * s.ntot)))
    t0 = time.perf_counter()
    eng.step(1, abi.NOISE_INJECTED, 0)
    probe = time.perf_counter() - t0
    nsteps = max(1, min(2000, int(target_s / max(probe, 1e-6))))
    eng.upload_noise(rng.standard_normal((nsteps, 3 * s.ntot)))
    t0 = time.perf_counter()
    eng.step(nsteps, abi.NOISE_INJECTED, 0)
    dt = time.perf_counter() - t0
    return {"value": nmov * nsteps / dt, "unit": "particle-steps/s", "cores": 1, "kind": kind,
            "sample": f"{nsteps} BD steps (ENERGY+DYNAMICS0+DYNAMICS0NUC) of the same "
                      f"{s.ntot}-particle system, 1 thread, {dt:.1f} s"}


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


def large_leg(cycles=4, gcmc_every=10, ngcmc=20, cpu=True):
    """The large translocation workload (56,000 particles: 50,000 ions and a 1000-bp
    CG-DNA of 6000 sites through a 15 A pore; 321x321x641 repulsion, static-field and
    reaction-field grids; PROFILE diffusion; GCMC bath buffers): `cycles` cycles of GRAND
    (`ngcmc` attempts) + `gcmc_every` BD steps on the device, and, with cpu, one step of the
    C oracle on this host (the reference's fixed arrays hold 23000 particles, fewer than this
    system has)."""
    import numpy as np
    from bromoc_amd import abi, configs
    from bromoc_amd.engine import Engine
    t0 = time.perf_counter()
    s = configs.translocation()
    gen_s = time.perf_counter() - t0
    g = s.meta["gcmc"]
    ib, nat = configs.assign_buffers(s, g)
    e = Engine(s)
    e.set_gcmc(g, ib)
    prob = np.zeros((g.nbuffer, s.capacity + 1), np.int32)
    rng = np.random.default_rng(5)
    e.step(4, abi.NOISE_PHILOX, seed=11, counter0=0, energies=False)
    nmov0 = e.ntot - s.nfix
    ener, ctr = 0.0, 4
    t0 = time.perf_counter()
    for c in range(cycles):
        _, ener = e.grand(ngcmc, rng.random(6 * ngcmc + 8), nat, prob, c + 1, ener)
        e.step(gcmc_every, abi.NOISE_PHILOX, seed=11, counter0=ctr, energies=False)
        ctr += gcmc_every
    el = time.perf_counter() - t0
    nmov1 = e.ntot - s.nfix
    _, kms = e.time_steps(5, abi.NOISE_PHILOX, seed=11, counter0=ctr)
    nsteps = cycles * gcmc_every
    out = {"value": 0.5 * (nmov0 + nmov1) * nsteps / el, "unit": "particle-steps/s",
           "particles": [int(nmov0), int(nmov1)], "ions": s.ntot - s.nsites,
           "dna_sites": s.nsites, "box_A": [200, 200, 600], "grid": list(s.meta["grid"]),
           "grid_spacing_A": s.meta["dcel"], "rfpar": True, "diffusion": "profile",
           "steps": nsteps, "gcmc_every": gcmc_every, "ngcmc": ngcmc,
           "ms_per_step": 1e3 * el / nsteps, "step_kernel_ms": kms,
           "dna_list_terms": int(sum(s.count(k) for k in s.LIST_NAMES)), "setup_s": gen_s}
    e.close()
    if cpu:
        from oracle.oracle import Oracle
        o = Oracle(s)
        t0 = time.perf_counter()
        o.step(1, abi.NOISE_PHILOX, seed=11, counter0=0)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": (s.ntot - s.nfix) / dt, "unit": "particle-steps/s",
                               "cores": 1, "kind": "port",
                               "sample": f"1 BD step of the same system, C oracle, 1 thread, "
                                         f"{dt:.1f} s"}
    return out


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
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="pore_dna",
                    choices=["pore_dna", "plain_ions", "translocation"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-parity", action="store_true", help="skip the start-state parity check")
    ap.add_argument("--no-host", action="store_true", help="skip the host round-trip rate")
    ap.add_argument("--no-batch", action="store_true", help="skip the replica-batch legs")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the large translocation leg (56,000 particles)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from bromoc_amd import abi, configs
    from bromoc_amd.engine import Engine

    t0 = time.perf_counter()
    if args.workload == "pore_dna":
        s = configs.pore_dna(seed=1)
        wl = {"workload": "pore_dna", "ions": s.ntot - s.nsites, "dna_sites": s.nsites,
              "dna_bp": s.meta["n_bp"], "box_A": [100, 100, 200], "grid": list(s.meta["grid"]),
              "grid_spacing_A": s.meta["dcel"], "pair_tables": s.npair,
              "table_knots": int(len(s.ep) // 3), "dt_ps": s.meta["dt"], "noise": "philox"}
    elif args.workload == "translocation":
        s = configs.translocation()
        wl = {"workload": "translocation", "ions": s.ntot - s.nsites, "dna_sites": s.nsites,
              "box_A": [200, 200, 600], "grid": list(s.meta["grid"]),
              "grid_spacing_A": s.meta["dcel"], "rfpar": True, "diffusion": "profile",
              "pair_tables": s.npair, "dt_ps": s.meta["dt"], "noise": "philox"}
    else:
        s = configs.plain_ions(seed=0)
        wl = {"workload": "plain_ions", "ions": s.ntot, "box_A": 60, "pair_tables": s.npair,
              "dt_ps": s.meta["dt"], "noise": "philox"}
    build_s = time.perf_counter() - t0
    nmov = s.nsites + (s.ntot - s.nsites - s.nfix)
    eng = Engine(s)

    parity = None
    if not args.no_parity:
        from oracle.oracle import Oracle
        o = Oracle(s)
        eg, eo = eng.energy(), o.energy()
        fg, fo = eng.forces(), o.forces()
        parity = {"force_normwise_relerr": float(np.max(np.abs(fg - fo)) / np.max(np.abs(fo))),
                  "energy_relerr": abs(eg["ener"] - eo["ener"]) / abs(eo["ener"])}
        del o

    eng.step(args.warmup, abi.NOISE_PHILOX, seed=2024, counter0=0, energies=False)
    tpb = eng.tuning()
    # timed region: the steps as production runs them (captured graphs), host clock around
    # a synchronous call
    t0 = time.perf_counter()
    eng.step(args.steps, abi.NOISE_PHILOX, seed=2024, counter0=args.warmup, energies=False)
    elapsed = time.perf_counter() - t0
    value = nmov * args.steps / elapsed
    # per-launch duration of the step kernel, HIP events on its own stream
    nk = min(args.steps, 500)
    _, kernel_ms = eng.time_steps(nk, abi.NOISE_PHILOX, seed=2024,
                                  counter0=args.warmup + args.steps)
    bpp = step_bytes(s, injected=False)
    traffic = None
    if TRAFFIC_KIB.get((args.workload, tpb)):
        fk, wk = TRAFFIC_KIB[(args.workload, tpb)]
        traffic = 1024.0 * (fk * TRAFFIC_SCALE[0] + wk * TRAFFIC_SCALE[1])
    achieved = bpp * nmov / (kernel_ms * 1e-3) / 1e9
    # FP64 work of the kernel: ~34 flops per table pair (r^2 8, sqrt 1, index 2, U 4,
    # F/r 3, energy 1, force update ~15), every ordered pair of a target with its partners
    nions = s.ntot - s.nsites
    pairs = (s.ntot - s.nsites - s.nfix) * (s.ntot - 1) + s.nsites * nions
    gflops = 34.0 * pairs / (kernel_ms * 1e-3) / 1e9
    host = None if args.no_host else host_roundtrip(eng, nmov)
    eng.close()
    batched, small_n = (None, None) if args.no_batch else batch_legs()
    large = None if args.no_large else large_leg(cpu=not args.no_cpu)
    adjoint = None if args.no_adjoint else adjoint_leg()
    cpu = None if args.no_cpu else cpu_baseline(s)
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded positions, grids and EFPOT tables)", "config": wl,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_step", "kernel_ms": kernel_ms, "bytes_per_particle_step": bpp,
                     "fp64_valu": {"achieved_tflops": gflops / 1e3, "peak_tflops": 78.6,
                                   "frac": gflops / 1e3 / 78.6, "pair_evals_per_step": pairs},
                     "targets_per_block": tpb},
        "cpu_baseline": cpu, "host_roundtrip": host, "batched": batched,
        "batched_vs_plain_small_n": small_n, "large_translocation": large,
        "adjoint_rdf_gradient": adjoint,
        "parity_vs_oracle": parity,
        "setup_s": build_s,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
