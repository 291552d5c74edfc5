#!/usr/bin/env python3
"""Benchmark of the linearized-PB hot path on one MI355X.

Workload (default): the synthetic globular protein — 2,500 atoms seeded (20502) in a
20 A ball, radii U[1.2, 2.0] A, charges N(0, 0.35) e shifted to a net -3 e,
eps_in 2 / eps_out 80, 0.145 M, T 298.15 K, domain edge 3x the molecular extent,
adaptive 2:1 octree with 0.5 A leaves near the molecule coarsening to 2 A at the box
boundary (about 8.0M unknowns with hanging nodes), Coulomb Dirichlet data.

A step is one PB solve on resident inputs: in/out markers -> edge fractions ->
numeric assembly (Laplacian, kappa^2 mass, hanging-node constraints, charge
projection, Dirichlet elimination) -> Krylov solve to ||r|| <= 1e-10 ||b|| ->
polarization + ionic energy. The mesh topology (symbolic assembly) is built once
before timing. `value` is mixed-precision solves/s; the plain FP64 solve of the
same system is timed beside it, with the relative energy error between the two.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def pct(a, q):
    return float(np.percentile(np.asarray(a), q))


def measured_traffic(bytes_per_launch, name="compact_spmv_globular_8M_traffic.txt", key="compact_spmv_bytes_per_launch"):
    """HBM bytes per SpMV launch from a committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (copy-kernel calibrated), if it was measured on an operator of exactly this
    many algorithmic bytes; otherwise None."""
    path = os.path.join(REPO, "profiles", name)
    try:
        for line in open(path):
            if line.startswith(key + "="):
                kv = dict(t.split("=") for t in line.split())
                if int(kv[key]) == int(bytes_per_launch):
                    return float(kv["traffic_bytes_per_launch"])
    except OSError:
        pass
    return None


def run_steps(pb, precision, steps, tol, check_every, profile):
    from cudanextgenpb_amd import native

    ctx = pb.ctx
    ctx.set_profiling(profile)
    lat, iters, st, E = [], [], None, None
    t0 = time.perf_counter()
    for k in range(steps):
        s0 = time.perf_counter()
        pb.create_markers()
        ctx.assemble()
        _, st = ctx.solve(precision, tol=tol, maxiter=200000, drop=1e-3, check_every=check_every, want_phi=False)
        E = ctx.energy()
        lat.append((time.perf_counter() - s0) * 1e3)
        iters.append(st["iterations"])
        log(f"  {'mixed' if precision == native.PREC_MIXED else 'fp64 '} step {k + 1}/{steps}: {lat[-1]:.1f} ms, "
            f"{st['iterations']} it, solve {st['ms_solve']:.1f} ms, relres {st['relres']:.2e}, "
            f"E_polar {E['polar']:.10f}")
    total = time.perf_counter() - t0
    kst = ctx.kernel_stats(precision) if profile else None
    ctx.set_profiling(False)
    return dict(total_s=total, lat_ms=lat, iters=iters, stats=st, energy=E, kernel=kst)


def cpu_baseline(pb, fp64_iters, seconds):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    rp, col, val, rhs = pb.ctx.export_csr()
    S = O.System(len(rhs), rp, col, val, rhs)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    # each call allocates and first-touches its vectors: the per-iteration time is
    # taken from the difference of two pilot lengths, which cancels that fixed cost
    O.pcg_time_iters(S, 3, threads)
    t5 = O.pcg_time_iters(S, 5, threads)
    t25 = O.pcg_time_iters(S, 25, threads)
    per_it = max((t25 - t5) / 20, 1e-6)
    n_it = int(max(5, min(20000, seconds / per_it)))
    t = O.pcg_time_iters(S, n_it, threads)
    per_it = t / n_it
    solve_s = per_it * fp64_iters
    return {
        "value": 1.0 / solve_s,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{n_it} Jacobi-PCG iterations (restated LIS -i cg -p jacobi, OpenMP, {threads} threads) on the "
                   f"same {S.n}-row FP64 CSR in {t:.1f} s = {per_it * 1e3:.2f} ms/iteration, scaled to the "
                   f"{fp64_iters} iterations of the FP64 solve; markers, assembly and energy not included"),
        "dof_iter_per_s": S.n / per_it,
    }


def batch_leg(n_systems, tol, prep_threads=8, tiling=None):
    """The batched ligand systems (config 4's generator): one batch (mixed precision,
    hipGraph replays, autotuned tiles) against one-at-a-time FP64 solves of the same
    systems; the batched SpMV's roofline from the autotuner's timing. tiling = (tile
    slices, waves per block) skips the autotuner (no roofline then): the form the
    counter profiles use, with prep_threads = 1."""
    from cudanextgenpb_amd import native, workload as W
    from cudanextgenpb_amd.batch import BatchedSolver, prepare, prepare_many

    # every system from its atoms on the device (octree, rays, owners, boundary values),
    # first one system at a time, then with one grouped assembly over all of them
    specs = [W.ligand_spec(i) for i in range(n_systems)]
    t = time.perf_counter()
    preps, tot_n = [], 0
    for i, spec in enumerate(specs):
        preps.append(prepare(spec))
        tot_n += preps[-1].n
        if (i + 1) % 64 == 0:
            log(f"prepared {i + 1}/{n_systems} systems one at a time ({tot_n / 1e6:.1f}M unknowns, "
                f"{time.perf_counter() - t:.1f} s)")
    setup_s = time.perf_counter() - t
    asm_one_s = sum(p.meta["assemble_ms"] for p in preps) * 1e-3

    for p in preps:
        p.ctx.close()
    t = time.perf_counter()
    preps = prepare_many(specs, threads=prep_threads)
    grouped_s = time.perf_counter() - t
    asm_many_s = sum(p.meta["assemble_ms"] for p in preps) * 1e-3
    log(f"prepared {n_systems} systems with one grouped assembly in {grouped_s:.2f} s (assembly {asm_many_s:.3f} s; "
        f"one at a time {setup_s:.2f} s, assembly {asm_one_s:.3f} s)")

    # one-at-a-time FP64 (single-system path), solve + energy per system, on the same
    # contexts the batch then solves
    t = time.perf_counter()
    e64, it64 = [], 0
    for p in preps:
        _, st = p.ctx.solve(native.PREC_FP64, tol=tol, want_phi=False)
        e64.append(p.ctx.energy()["polar"])
        it64 += st["iterations"]
    fp64_s = time.perf_counter() - t
    log(f"one-at-a-time FP64: {n_systems} solves in {fp64_s:.2f} s")

    spmv_bytes = sum(p.ctx.operator_info()["compact_spmv_bytes"] for p in preps)
    bs = BatchedSolver(0, max_systems=max(n_systems, 1), tol=tol)
    for p in preps:
        bs.add(p)
    if tiling:
        bs.batch.set_tiling(*tiling)
        tune = {"tile_slices": tiling[0], "waves_per_block": tiling[1], "ms_per_spmv": None}
    else:
        tune = bs.autotune()
    log(f"autotune: {tune}")
    t = time.perf_counter()
    res = bs.solve(energies=True)
    batch_s = time.perf_counter() - t
    timing = bs.batch.timing()
    log(f"batched mixed: {n_systems} solves in {batch_s:.2f} s ({timing['replays']} replays)")
    errs, iters, dof_it = [], 0, 0.0
    for p, e in zip(preps, e64):
        r = res[p.name]
        errs.append(abs(r["energy"]["polar"] - e) / abs(e))
        iters += r["stats"]["iterations"]
        dof_it += p.n * r["stats"]["iterations"]
        assert r["stats"]["status"] == 0, (p.name, r["stats"])
    last = bs.last_timing
    bs.close()
    for p in preps:
        p.ctx.close()
    gbs = spmv_bytes / (tune["ms_per_spmv"] * 1e-3) / 1e9 if tune["ms_per_spmv"] else None
    return {
        "systems": n_systems,
        "unknowns_total": tot_n,
        "solves_per_s": n_systems / batch_s,
        "batch": {
            "solve_s": batch_s,
            "replays": timing["replays"],
            "device_ms": timing["ms_total"],
            "solve_all_s": last["solve_all_s"],
            "energies_s": last["energies_s"],
            "inner_iterations_total": iters,
            "dof_iter_per_s": dof_it / batch_s,
            "autotune": tune,
            "spmv_bytes_all_systems": spmv_bytes,
            "max_rel_energy_err_vs_fp64": max(errs),
        },
        "roofline": {
            "kernel": "batched compact FP32 SpMV (kb_spmv: tiles of the batch's systems, autotuned tile and waves)",
            "bound": "hbm",
            "achieved": gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS if gbs else None,
            "traffic": measured_traffic(spmv_bytes, "batch_spmv_128_traffic.txt", "batch_spmv_bytes_per_launch"),
            "traffic_source": "profiles/batch_spmv_128_traffic.txt (rocprofv3 FETCH_SIZE x2.000 + WRITE_SIZE x1.000)",
            "bytes_per_launch": spmv_bytes,
            "avg_launch_ms": tune["ms_per_spmv"],
            "timing": "autotuner: HIP events around 5 launches of the chosen tiling over every system",
        },
        "fp64_one_at_a_time": {
            "solves_per_s": n_systems / fp64_s,
            "solve_s": fp64_s,
            "iterations_total": it64,
        },
        "batch_over_fp64_one_at_a_time": fp64_s / batch_s,
        "setup_s": setup_s,
        # preparation from atoms on the device (not in the timed solve): one system at a
        # time, and with one grouped assembly (ngpb_assemble_many) over all systems
        "prepare": {
            "one_at_a_time_s": setup_s,
            "one_at_a_time_assembly_s": asm_one_s,
            "grouped_s": grouped_s,
            "grouped_assembly_s": asm_many_s,
        },
    }


def bench_batch(args):
    """Batched ligand systems: one batch (mixed precision, hipGraph replays, autotuned
    tiles) against one-at-a-time FP64 solves of the same systems."""
    tiling = tuple(int(v) for v in args.tiling.split("x")) if args.tiling else None
    L = batch_leg(args.systems, args.tol, args.prep_threads, tiling)
    out = {
        "metric": "solves/s & DOF-iter/s at 1e-6 rel energy err, mixed vs FP64 GPU; p99 latency",
        "value": L["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": L["batch"]["solve_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.systems} ligands (seed 512*1000+i): U{{20..80}} atoms in 6-10 A balls, radii "
                         f"U[1.2,1.9], q N(0,0.3); uniform 0.4 A leaves, 48^3-96^3 per system; Coulomb Dirichlet; "
                         f"tol {args.tol:g}; a step = solve + energy of every system"),
            "systems": args.systems,
            "unknowns_total": L["unknowns_total"],
            "parallelism": "single GPU",
        },
    }
    for k in ("batch", "roofline", "fp64_one_at_a_time", "batch_over_fp64_one_at_a_time", "setup_s", "prepare"):
        out[k] = L[k]
    return out


def serving_leg(n_requests, rate, tol, device_mesh, probe=0.0):
    """Serving stream: Poisson arrivals, continuous batching with HBM-footprint
    admission, against the same stream served one at a time in FP64. Returns the two
    summaries, the peak admitted bytes, the largest relative energy difference and
    the stream's size."""
    from cudanextgenpb_amd import serve

    t = time.perf_counter()
    reqs = serve.make_request_stream(n_requests=n_requests, rate_hz=rate, seed=5)
    gen_s = time.perf_counter() - t
    tot_n = sum(r.n for r in reqs)
    log(f"{len(reqs)} requests generated in {gen_s:.1f} s: N {min(r.n for r in reqs)}..{max(r.n for r in reqs)}, "
        f"total {tot_n / 1e6:.1f}M unknowns, span {reqs[-1].arrival:.1f} s")
    srv = serve.Server(budget_bytes=250e9, small_max_unknowns=1_500_000, max_batch=256, tol=tol,
                       device_mesh=device_mesh, probe=probe)
    # warm-up: one small request through both paths
    warm = [r for r in reqs if r.n <= 200_000][:1]
    if warm:
        srv.run(warm, realtime=False)
        serve.run_sequential_fp64(warm, realtime=False, tol=tol, device_mesh=device_mesh, probe=probe)
    res = srv.run(reqs, realtime=True)
    peak = srv.peak_bytes
    srv.close()
    summ = serve.summarize(res)
    log(f"continuous batching: {summ}")
    seq = serve.run_sequential_fp64(reqs, realtime=True, tol=tol, device_mesh=device_mesh, probe=probe)
    seq_summ = serve.summarize(seq)
    log(f"sequential FP64: {seq_summ}")
    ok = [k for k in res if res[k].energy is not None and seq[k].energy is not None]
    errs = [abs(res[k].energy["polar"] - seq[k].energy["polar"]) / abs(seq[k].energy["polar"]) for k in ok]
    return dict(summary=summ, fp64=seq_summ, peak=peak, max_err=max(errs) if errs else None, requests=len(reqs),
                unknowns=tot_n, warm=bool(warm))


def bench_serve(args):
    dm = args.device_mesh
    L = serving_leg(args.requests, args.rate, args.tol, dm, args.probe)
    summ, seq_summ, peak = L["summary"], L["fp64"], L["peak"]
    reqs_n, tot_n, warm = L["requests"], L["unknowns"], L["warm"]
    return {
        "metric": "solves/s & DOF-iter/s at 1e-6 rel energy err, mixed vs FP64 GPU; p99 latency",
        "value": summ["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1 if warm else 0,
        "ms_per_step": summ["makespan_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {
            "workload": (f"serving stream: {reqs_n} requests, Poisson arrivals at {args.rate:g}/s (seed 5), atoms "
                         f"log-uniform in [50, 20000], globular generator, 0.5 A leaves within 3 A of the atoms, 1 A "
                         f"within 6 A, 2 A beyond; mixed solve tol {args.tol:g} + energy; admission by predicted HBM "
                         f"footprint within 250 GB; systems <= 1.5M unknowns batched, larger streamed; "
                         + ("requests arrive as atoms and box, octree, rays and owners built on the device"
                            if dm else "requests arrive with host-built mesh, ray table and owners")),
            "requests": reqs_n,
            "unknowns_total": tot_n,
            "parallelism": "single GPU",
        },
        "p50_latency_ms": summ["p50_latency_ms"],
        "p99_latency_ms": summ["p99_latency_ms"],
        "serving": dict(summ, peak_admitted_bytes=peak, max_rel_energy_err_vs_fp64=L["max_err"]),
        "fp64_sequential": seq_summ,
    }


def bench_mesh(args):
    """Per-request mesh preparation on the device (octree, ray table, atom owners) for
    the globular (or large) system, against the host generator of the same mesh."""
    import numpy as np
    from cudanextgenpb_amd import native, workload as W

    big = args.mesh_size == "large"
    t = time.perf_counter()
    P = W.large() if big else W.globular()
    host_problem_s = time.perf_counter() - t
    m, a = P.mesh, P.atoms
    t = time.perf_counter()
    W.make_mesh(m.origin, m.hf, m.n, m.smax, m.bands, a).close()
    host_mesh_s = time.perf_counter() - t
    log(f"host: problem {host_problem_s:.2f} s, mesh alone {host_mesh_s:.3f} s; N={m.n_nodes} L={m.n_leaves}")
    ctx = native.Context(0)
    ctx.set_physics(P.physics)
    tb, tr, tl, ts = [], [], [], []
    for it in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        sizes = ctx.build_octree(m.origin, m.hf, m.n, m.smax, m.bands, a.xyz, a.r)
        t1 = time.perf_counter()
        ctx.cast_rays(a.xyz, a.r)
        t2 = time.perf_counter()
        owner = ctx.locate_atoms(a.xyz, P.atom_tol)
        t3 = time.perf_counter()
        if args.probe > 0.0:
            ctx.cast_rays_ses(a.xyz, a.r, args.probe)
        t4 = time.perf_counter()
        log(f"  build {it + 1}/{args.warmup + args.steps}: octree {(t1 - t0) * 1e3:.1f} ms, rays {(t2 - t1) * 1e3:.1f} ms, "
            f"owners {(t3 - t2) * 1e3:.1f} ms" + (f", SES rays {(t4 - t3) * 1e3:.1f} ms" if args.probe > 0 else ""))
        if it >= args.warmup:
            tb.append(t1 - t0)
            tr.append(t2 - t1)
            tl.append(t3 - t2)
            ts.append(t4 - t3)
    ses_rays = [int(ctx.get_rays(axis)[1].size - 1) for axis in range(3)] if args.probe > 0.0 else None
    ctx.cast_rays(a.xyz, a.r)  # the sphere-union table is what the host caster is compared with
    g = ctx.get_mesh()
    same = all(np.array_equal(g[k], getattr(m, k))
               for k in ("node_xyz", "hang_xyz", "hang_np", "hang_par", "leaf_corner", "bnd_nodes"))
    same = same and np.array_equal(owner, P.owner)
    for axis in range(3):
        key, ptr, inters, normals = ctx.get_rays(axis)
        same = same and np.array_equal(key, P.rays[axis].key) and np.array_equal(inters, P.rays[axis].inters)
    ctx.close()
    if not same:
        raise SystemExit("device mesh, rays or owners differ from the host generator's")
    med = lambda v: float(np.median(v)) * 1e3
    total = med([x + y + z for x, y, z in zip(tb, tr, tl)])
    return {
        "metric": "per-request mesh preparation time (octree + rays + atom owners)",
        "value": total,
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+int32",
        "data": "synthetic",
        "config": {
            "workload": ("large complex, 40000 atoms, 0.35 A finest leaves" if big else
                         "globular protein, 2500 atoms, 0.5 A finest leaves") + ", 2 refinement levels, 2:1 balanced",
            "nodes": sizes[0], "hanging": sizes[1], "leaves": sizes[2], "boundary_nodes": sizes[3],
            "parallelism": "single GPU",
        },
        "octree_ms": med(tb),
        "rays_ms": med(tr),
        "locate_ms": med(tl),
        "ses_rays_ms": med(ts) if args.probe > 0.0 else None,
        "ses_rays_per_axis": ses_rays,
        "identical_to_host": True,
        "cpu_baseline": {"value": host_mesh_s * 1e3, "unit": "ms", "kind": "port",
                         "cores": int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1),
                         "sample": "host generator of the same octree (csrc/workload.cpp, OpenMP), one build; "
                                   f"with rays and owners: {host_problem_s * 1e3:.0f} ms"},
    }


REQUEST_STAGES = ("octree", "rays", "markers", "locate", "charges", "dirichlet", "assemble", "solve", "energy")


def request_leg(P, warmup, steps, tol, check_every, probe=0.0):
    """Requests from atoms to energy on fresh contexts (octree, rays, markers, owners,
    Dirichlet data, assembly, mixed solve, energy); per-request stage times in ms and
    the energy (identical across requests)."""
    from cudanextgenpb_amd import native

    m, a = P.mesh, P.atoms
    stages = REQUEST_STAGES
    rows, energy = [], None
    for it in range(warmup + steps):
        c = native.Context(0)
        c.set_physics(P.physics)
        tt = [time.perf_counter()]
        c.build_octree(m.origin, m.hf, m.n, m.smax, m.bands, a.xyz, a.r)
        tt.append(time.perf_counter())
        if probe > 0.0:
            c.cast_rays_ses(a.xyz, a.r, probe)
        else:
            c.cast_rays(a.xyz, a.r)
        tt.append(time.perf_counter())
        c.create_markers()
        tt.append(time.perf_counter())
        owner = c.locate_atoms(a.xyz, P.atom_tol)
        tt.append(time.perf_counter())
        c.set_charges(a.xyz, a.q, owner)
        tt.append(time.perf_counter())
        c.dirichlet_values(P.bc_type)
        tt.append(time.perf_counter())
        c.assemble()
        tt.append(time.perf_counter())
        _, st = c.solve(native.PREC_MIXED, tol=tol, check_every=check_every, want_phi=False)
        tt.append(time.perf_counter())
        e = c.energy()
        tt.append(time.perf_counter())
        c.close()
        if energy is not None and e["polar"] != energy["polar"]:
            raise SystemExit("requests gave different energies")
        energy = e
        if it >= warmup:
            rows.append([(tt[k + 1] - tt[k]) * 1e3 for k in range(len(stages))])
            log(f"  request {it + 1 - warmup}/{steps}: {sum(rows[-1]):.1f} ms, {st['iterations']} it")
    return np.array(rows), energy


def bench_request(args):
    """One request from atoms to energy on a fresh context: octree, rays, markers,
    owners, Dirichlet data, assembly (symbolic + numeric), mixed solve, energy."""
    from cudanextgenpb_amd import workload as W

    big = args.mesh_size == "large"
    P = W.large() if big else W.globular()
    m = P.mesh
    stages = REQUEST_STAGES
    r, energy = request_leg(P, args.warmup, args.steps, args.tol, args.check_every, args.probe)
    tot = r.sum(axis=1)
    return {
        "metric": "request latency from atoms to energy (mesh, assembly, mixed solve, energy on the GPU)",
        "value": float(np.median(tot)),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.median(tot)),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": ("large complex, 40000 atoms" if big else "globular protein, 2500 atoms")
                   + f", {m.n_nodes} nodes, {m.n_leaves} leaves; mixed solve tol {args.tol:g} + energy",
                   "parallelism": "single GPU"},
        "stages_ms": {k: float(np.median(r[:, i])) for i, k in enumerate(stages)},
        "energy_polar": energy["polar"],
    }


def bench_spmv(args):
    """SpMV kernels alone on an operator assembled from atoms on the device: the
    compact FP32 SpMV under the chosen schedule and the FP64 CSR SpMV, each timed
    over back-to-back launches with HIP events on the context's stream, and a plain
    copy kernel of known size for reference (also the calibration kernel of the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs)."""
    from cudanextgenpb_amd import native, workload as W
    from cudanextgenpb_amd.batch import prepare

    spec = W.large_spec() if args.mesh_size == "large" else W.globular_spec()
    p = prepare(spec, probe=args.probe)  # the molecules' surface: SES with the default probe
    ctx = p.ctx
    info = ctx.operator_info()
    x = np.random.default_rng(1).standard_normal(p.n)
    ctx.set_row_order(args.row_band)
    ctx.set_spmv_schedule(4)
    y_ref = ctx.spmv(x, native.PREC_MIXED)
    ctx.set_spmv_schedule(args.spmv_schedule)
    y = ctx.spmv(x, native.PREC_MIXED)
    if not np.array_equal(y, y_ref):
        raise SystemExit("SpMV schedule changed the result")
    ms32, b32 = ctx.time_spmv(native.PREC_MIXED, reps=args.spmv_reps)
    ms64, b64 = ctx.time_spmv(native.PREC_FP64, reps=args.spmv_reps)
    copy_bytes = 1 << 30
    copy_ms = ctx.time_copy(copy_bytes, reps=args.spmv_reps)
    ctx.close()
    gbs = b32 / (ms32 * 1e-3) / 1e9
    log(f"compact SpMV {ms32 * 1e3:.1f} us ({b32 / 1e6:.1f} MB, {gbs:.0f} GB/s); FP64 CSR {ms64 * 1e3:.1f} us "
        f"({b64 / (ms64 * 1e-3) / 1e9:.0f} GB/s); copy {2 * copy_bytes / (copy_ms * 1e-3) / 1e9:.0f} GB/s")
    return {
        "metric": "compact FP32 SpMV achieved HBM bandwidth (algorithmic bytes per launch / launch time)",
        "value": gbs,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.spmv_reps,
        "warmup": 1,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"{spec.name}: operator of {p.n} rows, {p.nnz} nnz assembled from atoms on the device "
                               f"(SES, probe {args.probe:g} A); SpMV schedule {args.spmv_schedule}, row order {args.row_band}",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": None, "bytes_per_launch": b32, "avg_launch_ms": ms32},
        "fp64_spmv": {"ms": ms64, "bytes": b64, "gbs": b64 / (ms64 * 1e-3) / 1e9},
        "copy_kernel_gbs": 2 * copy_bytes / (copy_ms * 1e-3) / 1e9,
        "operator": info,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["globular", "born", "large", "batch", "serve", "mesh", "request", "spmv"],
                    default="globular")
    ap.add_argument("--spmv-schedule", type=int, default=6, help="spmv workload: compact SpMV schedule (0-6)")
    ap.add_argument("--spmv-reps", type=int, default=50, help="spmv workload: timed launches per kernel")
    ap.add_argument("--row-band", type=int, default=0,
                    help="spmv workload: row order of the compact operator (0 node order, >0 band height)")
    ap.add_argument("--mesh-size", choices=["globular", "large"], default="globular",
                    help="mesh and request workloads: which system")
    ap.add_argument("--requests", type=int, default=48, help="serve workload: number of requests")
    ap.add_argument("--rate", type=float, default=8.0, help="serve workload: Poisson arrival rate [1/s]")
    ap.add_argument("--systems", type=int, default=512, help="batch workload: number of ligand systems")
    ap.add_argument("--prep-threads", type=int, default=8, help="batch workload: threads preparing the systems")
    ap.add_argument("--tiling", default=None, help="batch workload: fixed tiling SLICESxWAVES (skips the autotuner)")
    ap.add_argument("--batch-systems", type=int, default=128,
                    help="default run: ligand systems of the batched leg (0: skip it)")
    ap.add_argument("--device-mesh", action="store_true",
                    help="serve workload: requests carry only atoms and box; mesh, rays and owners built on the GPU")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--probe", type=float, default=1.4,
                    help="probe radius of the solvent-excluded surface the rays are cast against (0: the union of "
                         "the atom spheres)")
    ap.add_argument("--check-every", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--no-serve", action="store_true",
                    help="globular workload: skip the serving-stream and request-from-atoms legs")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (each system is solved on one device; "
              "independent systems would be spread over devices, a single large one split into Morton ranges "
              "of leaves with a halo exchange per iteration). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from cudanextgenpb_amd import native, workload as W
    from cudanextgenpb_amd.pb import PoissonBoltzmann

    if native.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.workload == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    if args.workload == "mesh":
        print(json.dumps(bench_mesh(args)), flush=True)
        return
    if args.workload == "request":
        print(json.dumps(bench_request(args)), flush=True)
        return
    if args.workload == "spmv":
        print(json.dumps(bench_spmv(args)), flush=True)
        return

    t = time.perf_counter()
    if args.workload == "globular":
        P = W.globular()
        wname = ("globular protein: 2500 atoms (seed 20502) in a 20 A ball, radii U[1.2,2.0], q N(0,0.35) net -3 e; "
                 "eps 2/80, 0.145 M; box 3x extent; 2:1 octree 0.5 A near the molecule -> 2 A at the boundary; "
                 "Coulomb Dirichlet; tol 1e-10")
    elif args.workload == "large":
        P = W.large()
        wname = ("large complex: 40000 atoms (seed 40003) in a 60x45x45 A ellipsoid, radii U[1.2,2.0], q N(0,0.35) "
                 "net -20 e; eps 2/80, 0.145 M; box 3x extent; 2:1 octree 0.35 A near the molecule -> 1.4 A at the "
                 "boundary; Coulomb Dirichlet; tol 1e-10")
    else:
        P = W.born(h=0.25)
        wname = "Born sphere: r 2 A, q +1, 32 A box, uniform 0.25 A (128^3 leaves), analytic Dirichlet; tol 1e-10"
    m = P.mesh
    log(f"workload {P.name}: N={m.n_nodes} leaves={m.n_leaves} hanging={m.n_hang} atoms={P.atoms.n} "
        f"(host generation {time.perf_counter() - t:.1f} s)")

    # the Born configuration's surface is the sphere itself; the molecules' is the SES
    probe = 0.0 if args.workload == "born" else args.probe
    surface = (f"solvent-excluded surface, probe {probe:g} A, rays cast on the device" if probe > 0
               else "union of the atom spheres")
    t = time.perf_counter()
    pb = PoissonBoltzmann(P, probe_radius=probe)
    pb.create_markers()
    pb.create_density_map()
    pb.assemple_system_matrix()
    ctx = pb.ctx
    setup_ms = (time.perf_counter() - t) * 1e3
    log(f"resident on device: n={ctx.n} nnz={ctx.nnz} device bytes={ctx.device_bytes() / 2**30:.2f} GiB "
        f"(upload + symbolic + first assembly {setup_ms:.0f} ms)")

    for _ in range(args.warmup):
        run_steps(pb, native.PREC_MIXED, 1, args.tol, args.check_every, False)
        if not args.no_fp64:
            run_steps(pb, native.PREC_FP64, 1, args.tol, args.check_every, False)

    mixed = run_steps(pb, native.PREC_MIXED, args.steps, args.tol, args.check_every, True)
    fp64 = None if args.no_fp64 else run_steps(pb, native.PREC_FP64, args.steps, args.tol, args.check_every, True)

    n = ctx.n
    op_info = ctx.operator_info()
    copy_bytes = 1 << 30
    copy_ms = ctx.time_copy(copy_bytes, reps=20)
    value = args.steps / mixed["total_s"]
    mixed_solve_ms = mixed["stats"]["ms_solve"]
    mk = mixed["kernel"]
    avg_ms = mk["total_ms"] / max(mk["launches"], 1)
    achieved = mk["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    csr32 = float(ctx.nnz * 8 + (n + 1) * 4 + n * 8)
    out = {
        "metric": "solves/s & DOF-iter/s at 1e-6 rel energy err, mixed vs FP64 GPU; p99 latency",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["total_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {
            "workload": wname,
            "surface": surface,
            "unknowns": n,
            "nnz": ctx.nnz,
            "leaves": m.n_leaves,
            "hanging_nodes": m.n_hang,
            "atoms": P.atoms.n,
            "parallelism": "single GPU",
        },
        "mixed": {
            "step_p50_ms": pct(mixed["lat_ms"], 50),
            "step_p99_ms": pct(mixed["lat_ms"], 99),
            "iterations": mixed["iters"][-1],
            "replacements": mixed["stats"]["replacements"],
            "solve_ms": mixed_solve_ms,
            "dof_iter_per_s": n * mixed["iters"][-1] / (mixed_solve_ms * 1e-3),
            "relres_fp64": mixed["stats"]["relres"],
            "energy_polar_kT": mixed["energy"]["polar"],
        },
        "roofline": {
            "kernel": "compact FP32 SpMV of the mixed inner CG (k_sell_lean<true>: SELL-64 delta, constant and packed slices, XCD-segmented, guard-banded input)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(mk["bytes_per_launch"]),
            "traffic_source": "profiles/compact_spmv_globular_8M_traffic.txt (rocprofv3 FETCH_SIZE x2.000 + WRITE_SIZE x1.000)",
            "bytes_per_launch": mk["bytes_per_launch"],
            "avg_launch_ms": avg_ms,
            "launches": mk["launches"],
            "copy_kernel_gbs": 2.0 * copy_bytes / (copy_ms * 1e-3) / 1e9,
        },
        # bytes of a plain FP32 CSR of the same operator (nnz (4 + 4) + (N + 1) 4 + N (4 + 4)):
        # the compression the compact format achieves, not a roofline
        "fp32_csr_bytes": csr32,
        "operator": op_info,
    }
    if fp64 is not None:
        fk = fp64["kernel"]
        f_avg = fk["total_ms"] / max(fk["launches"], 1)
        e64 = fp64["energy"]["polar"]
        out["fp64"] = {
            "solves_per_s": args.steps / fp64["total_s"],
            "iterations": fp64["iters"][-1],
            "solve_ms": fp64["stats"]["ms_solve"],
            "dof_iter_per_s": n * fp64["iters"][-1] / (fp64["stats"]["ms_solve"] * 1e-3),
            "p50_latency_ms": pct(fp64["lat_ms"], 50),
            "p99_latency_ms": pct(fp64["lat_ms"], 99),
            "energy_polar_kT": e64,
            "spmv_avg_launch_ms": f_avg,
            "spmv_gbs": fk["bytes_per_launch"] / (f_avg * 1e-3) / 1e9 if f_avg > 0 else 0.0,
        }
        out["mixed_over_fp64_solves"] = value / out["fp64"]["solves_per_s"]
        out["mixed_over_fp64_solve_time"] = fp64["stats"]["ms_solve"] / mixed_solve_ms
        out["rel_energy_err_vs_fp64"] = abs(mixed["energy"]["polar"] - e64) / abs(e64)
    if not args.no_serve and args.workload == "globular":
        # the metric's latency: a request stream (config 5) and one request from atoms
        log("request from atoms to energy (globular, fresh context per request)")
        r, _ = request_leg(P, 1, 5, args.tol, args.check_every, probe)
        tot = r.sum(axis=1)
        out["request_from_atoms"] = {
            "median_ms": float(np.median(tot)),
            "max_ms": float(tot.max()),
            "requests": int(len(tot)),
            "stages_ms": {k: float(np.median(r[:, i])) for i, k in enumerate(REQUEST_STAGES)},
        }
        log(f"serving stream: {args.requests} requests at {args.rate:g}/s, atoms in [50, 20000], 0.5 A leaves")
        L = serving_leg(args.requests, args.rate, args.tol, True, probe)
        summ = L["summary"]
        out["p50_latency_ms"] = summ["p50_latency_ms"]
        out["p99_latency_ms"] = summ["p99_latency_ms"]
        out["latency_source"] = ("serving stream: arrival to potential and energy on the host, continuous batching "
                                 "(mixed precision) with HBM-footprint admission, requests arriving as atoms")
        out["serving"] = dict(summ, rate_hz=args.rate, unknowns_total=L["unknowns"], peak_admitted_bytes=L["peak"],
                              max_rel_energy_err_vs_fp64=L["max_err"], fp64_sequential=L["fp64"])
        if args.batch_systems > 0:
            # config 4's batched ligands (the first batch_systems of its 512)
            log(f"batched ligands: {args.batch_systems} systems, one batch against one-at-a-time FP64")
            out["batched_ligands"] = batch_leg(args.batch_systems, args.tol)
    else:
        out["p50_latency_ms"] = pct(mixed["lat_ms"], 50)
        out["p99_latency_ms"] = pct(mixed["lat_ms"], 99)
        out["latency_source"] = "solve steps of the timed loop"
    if not args.no_cpu:
        log("CPU baseline (restated LIS PCG-Jacobi, OpenMP)")
        fp64_iters = fp64["iters"][-1] if fp64 is not None else mixed["iters"][-1]
        out["cpu_baseline"] = cpu_baseline(pb, fp64_iters, args.cpu_seconds)
    pb.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
