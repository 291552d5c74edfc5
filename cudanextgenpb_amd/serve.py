"""Continuous-batching PB solve server on one GPU.

Requests (a molecule with its octree mesh and ray table) arrive over time. The
server admits them while their predicted HBM footprint fits the budget:

* small systems (up to `small_max_unknowns`) are uploaded and assembled by worker
  threads on their own context streams, join the running mixed-precision batch
  (``native.Batch``) at its next replay boundary and leave it as soon as they are done,
  so new requests neither wait for the batch to drain nor stall it;
* large systems are streamed: each is prepared and solved by the single-system mixed
  solver on its own context stream in a worker thread, overlapping the batch.

Each system keeps its mesh, operator and solver state resident on the device until
its energy has been read back. A result does not depend on what else was in flight:
batched systems use per-system reductions in a fixed order and join only at replay
boundaries, and streamed systems run the deterministic single-system solver.

Latency is measured from a request's arrival to its potential and energy being
available on the host. The comparison leg (`run_sequential_fp64`) serves the same
arrival stream one request at a time with the plain FP64 solver.

Reference counterpart: per request, the chain of src/poisson_boltzmann.cpp:197-299;
the reference has no server, it solves one system per process.
"""
from __future__ import annotations

import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np

from . import native
from . import workload as W
from .batch import prepare


@dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the stream
    problem: W.Problem
    n: int                    # unknowns
    leaves: int
    predicted_bytes: int


def predict_bytes(problem: W.Problem) -> int:
    """Upper estimate of the device bytes one system holds at its peak (the symbolic
    assembly, when the sorted emission lists and their repacked copy coexist) through
    markers, assembly, mixed solve and energy, from the mesh sizes and the atoms alone
    (before anything is uploaded).
    Per node: coordinates 24, markers 16, Dirichlet 9, row pointers and diagonal slots
    12, right-hand sides 24, scaling 8, FP64 and FP32 solver vectors 64, scratch 10,
    symbolic temporaries 210; per matrix entry (at most 8 per node outside hanging
    regions, 12 allowed): CSR 12, block-relative slot offset 4, compact operator <= 6,
    24 for the symbolic phase's per-slot temporaries (8 are used: the slot starts); per
    leaf: corners 32, marker and border
    map 5, conductances 104, and 56 emissions of 5 bytes kept plus 10 bytes while
    sorting; per hanging node 64; per border leaf (fractions, normals and the energy's
    sources and accumulators) 3.6 KB, with the border leaves bounded by the atoms'
    summed sphere areas over the finest leaf face (the surface of their union is
    smaller)."""
    mesh = problem.mesh
    N, L, H = mesh.n_nodes, mesh.n_leaves, mesh.n_hang
    per_node = 24 + 16 + 9 + 12 + 24 + 8 + 64 + 10 + 210 + 12 * (12 + 4 + 6 + 24)
    per_leaf = 32 + 5 + 104 + 56 * (5 + 10)
    area = float(np.sum(4.0 * np.pi * problem.atoms.r ** 2))
    border = min(L, int(area / (mesh.hf * mesh.hf)) + 1)
    return int(1.25 * (per_node * N + per_leaf * L + 64 * H + 3600 * border) + (8 << 20))


def make_request_stream(n_requests: int = 48, rate_hz: float = 4.0, seed: int = 5, atoms_lo: int = 50,
                        atoms_hi: int = 20000, h_fine: float = 0.5, fine_band: float = 3.0,
                        mid_band: float = 6.0) -> List[Request]:
    """Poisson arrivals at `rate_hz`; atom counts log-uniform in [atoms_lo, atoms_hi];
    molecules from the globular generator (ball radius scaled with the atom count at
    the globular configuration's density, same radius/charge distributions), leaves of
    h_fine within fine_band of the atoms, 2 h_fine within mid_band, 4 h_fine beyond,
    box 3x the molecule extent, Coulomb Dirichlet data."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for k in range(n_requests):
        t += float(rng.exponential(1.0 / rate_hz))
        na = int(round(math.exp(rng.uniform(math.log(atoms_lo), math.log(atoms_hi)))))
        radius = 20.0 * (na / 2500.0) ** (1.0 / 3.0)
        qnet = -3.0 * na / 2500.0
        P = W.globular(seed=seed * 100000 + k, n_atoms=na, ball_radius=radius, min_sep=2.0, h_fine=h_fine,
                       levels=2, fine_band=fine_band, mid_band=mid_band, qnet=round(qnet))
        out.append(Request(k, t, P, P.mesh.n_nodes, P.mesh.n_leaves, predict_bytes(P)))
    return out


@dataclass
class Result:
    rid: int
    mode: str                 # "batch" or "stream"
    arrival: float
    admitted: float
    done: float
    latency_ms: float
    iterations: int
    status: int
    energy: Optional[dict]
    device_bytes: int
    phi: Optional[np.ndarray] = None
    error: Optional[str] = None   # set when the request failed (status != 0 then)


class Server:
    def __init__(self, device: int = 0, budget_bytes: float = 250e9, small_max_unknowns: int = 1_500_000,
                 max_batch: int = 256, tol: float = 1e-10, inner_per_replay: int = 8, stream_workers: int = 2,
                 prep_workers: int = 2, keep_phi: bool = False, device_mesh: bool = False, probe: float = 0.0):
        self.device = device
        self.device_mesh = device_mesh
        self.probe = probe  # > 0: solvent-excluded surface rays with this probe radius
        self.budget = float(budget_bytes)
        self.small_max = small_max_unknowns
        self.tol = tol
        self.keep_phi = keep_phi
        self.batch = native.Batch(device, max_batch, tol=tol, inner_per_replay=inner_per_replay)
        self.max_batch = max_batch
        self.pool = ThreadPoolExecutor(max_workers=stream_workers)
        self.prep_pool = ThreadPoolExecutor(max_workers=prep_workers)
        # finished contexts are released on their own thread: freeing device memory
        # waits for the work queued on every stream of the device, which held the
        # serving loop (replays, admissions) for up to ~165 ms per context and delayed
        # streamed results by ~16 ms
        self.close_pool = ThreadPoolExecutor(max_workers=1)
        self.peak_bytes = 0.0

    def close(self):
        self.pool.shutdown(wait=True)
        self.prep_pool.shutdown(wait=True)
        self.close_pool.shutdown(wait=True)
        self.batch.close()

    def _stream_solve(self, req: Request):
        """Large request: prepared and solved on its own context (worker thread). A
        solve that stops at maxiter or stagnates is reported through its status, as a
        batched one is; the context is released whatever happens, after the result is
        in hand (on the releasing thread, so the result does not wait for it)."""
        p = prepare(req.problem, self.device, device_mesh=self.device_mesh, probe=self.probe)
        try:
            _, st = p.ctx.solve(native.PREC_MIXED, tol=self.tol, want_phi=False, raise_on_noconv=False)
            e = p.ctx.energy()
            phi = p.ctx.get_phi() if self.keep_phi else None
            nbytes = p.ctx.device_bytes()
        finally:
            self.close_pool.submit(p.ctx.close)
        return st, e, phi, nbytes

    @staticmethod
    def _failed(req: Request, mode: str, adm: float, t: float, exc: BaseException) -> Result:
        status = exc.code if isinstance(exc, native.NGPBError) else -1
        return Result(req.rid, mode, req.arrival, adm, t, (t - req.arrival) * 1e3, 0, status or -1, None, 0,
                      None, f"{type(exc).__name__}: {exc}")

    def run(self, requests: List[Request], realtime: bool = True) -> Dict[int, Result]:
        """Serves `requests` (arrival times relative to the start). Small requests are
        prepared by worker threads and join the batch at a replay boundary; the main
        loop keeps replaying the batch meanwhile. Large requests are prepared and
        solved by stream workers. realtime=False admits every request at once."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: 0.0)
        nxt = 0
        pending: List[Request] = []
        preparing: List[tuple] = []       # (req, admitted, future -> Prepared)
        in_batch: Dict[int, tuple] = {}   # slot -> (req, ctx, admitted)
        streaming: List[tuple] = []       # (req, admitted, future)
        used = 0.0
        results: Dict[int, Result] = {}
        while len(results) < len(reqs):
            now = clock() if realtime else float("inf")
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                pending.append(reqs[nxt])
                nxt += 1
            # admission: arrival order, any request whose footprint fits now
            for req in list(pending):
                small = req.n <= self.small_max
                if used + req.predicted_bytes > self.budget:
                    continue
                if small and len(in_batch) + len(preparing) >= self.max_batch:
                    continue
                pending.remove(req)
                used += req.predicted_bytes
                self.peak_bytes = max(self.peak_bytes, used)
                if small:
                    fut = self.prep_pool.submit(prepare, req.problem, self.device, None, self.device_mesh, self.probe)
                    preparing.append((req, clock(), fut))
                else:
                    streaming.append((req, clock(), self.pool.submit(self._stream_solve, req)))
            progressed = False
            # prepared small systems join at this replay boundary
            for item in list(preparing):
                req, adm, fut = item
                if not fut.done():
                    continue
                preparing.remove(item)
                progressed = True
                try:
                    p = fut.result()
                except Exception as exc:  # this request fails; the others go on
                    used -= req.predicted_bytes
                    results[req.rid] = self._failed(req, "batch", adm, clock(), exc)
                    continue
                try:
                    slot = self.batch.add(p.ctx)
                except Exception as exc:
                    p.ctx.close()
                    used -= req.predicted_bytes
                    results[req.rid] = self._failed(req, "batch", adm, clock(), exc)
                    continue
                p.ctx._keep_mesh = None
                in_batch[slot] = (req, p.ctx, adm)
            if in_batch:
                self.batch.run(1)
                progressed = True
                for slot, (req, ctx, adm) in list(in_batch.items()):
                    done, st = self.batch.status(slot)
                    if not done:
                        continue
                    del in_batch[slot]
                    used -= req.predicted_bytes
                    try:
                        self.batch.remove(slot)
                    except Exception as exc:  # the context stays with the batch
                        results[req.rid] = self._failed(req, "batch", adm, clock(), exc)
                        continue
                    try:
                        e = ctx.energy()
                        phi = ctx.get_phi() if self.keep_phi else None
                        t = clock()
                        nbytes = ctx.device_bytes()
                    except Exception as exc:
                        results[req.rid] = self._failed(req, "batch", adm, clock(), exc)
                        continue
                    finally:
                        self.close_pool.submit(ctx.close)
                    results[req.rid] = Result(req.rid, "batch", req.arrival, adm, t, (t - req.arrival) * 1e3,
                                              st["iterations"], st["status"], e, nbytes, phi)
            for item in list(streaming):
                req, adm, fut = item
                if not fut.done():
                    continue
                t = clock()
                used -= req.predicted_bytes
                streaming.remove(item)
                progressed = True
                try:
                    st, e, phi, nbytes = fut.result()
                except Exception as exc:
                    results[req.rid] = self._failed(req, "stream", adm, t, exc)
                    continue
                results[req.rid] = Result(req.rid, "stream", req.arrival, adm, t, (t - req.arrival) * 1e3,
                                          st["iterations"], st["status"], e, nbytes, phi)
            if not progressed:
                if pending and not in_batch and not streaming and not preparing:
                    raise RuntimeError("a request does not fit the HBM budget on its own")
                if realtime and nxt < len(reqs) and not pending and not streaming and not preparing:
                    time.sleep(max(0.0, min(0.005, reqs[nxt].arrival - clock())))
                else:
                    time.sleep(0.0005)
        return results


def run_sequential_fp64(requests: List[Request], device: int = 0, tol: float = 1e-10,
                        realtime: bool = True, device_mesh: bool = False, probe: float = 0.0) -> Dict[int, Result]:
    """The same arrival stream served one request at a time with the FP64 solver."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    out = {}
    for req in reqs:
        if realtime:
            wait = req.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        adm = time.perf_counter() - t0
        p = prepare(req.problem, device, device_mesh=device_mesh, probe=probe)
        _, st = p.ctx.solve(native.PREC_FP64, tol=tol, want_phi=False)
        e = p.ctx.energy()
        t = time.perf_counter() - t0
        nbytes = p.ctx.device_bytes()
        p.ctx.close()
        out[req.rid] = Result(req.rid, "fp64", req.arrival, adm, t, (t - req.arrival) * 1e3, st["iterations"],
                              st["status"], e, nbytes)
    return out


def summarize(results: Dict[int, Result]) -> dict:
    lat = np.array([r.latency_ms for r in results.values()])
    first = min(r.arrival for r in results.values())
    last = max(r.done for r in results.values())
    return {
        "requests": len(results),
        "solves_per_s": len(results) / max(last - first, 1e-9),
        "makespan_s": last - first,
        "p50_latency_ms": float(np.percentile(lat, 50)),
        "p99_latency_ms": float(np.percentile(lat, 99)),
        "max_latency_ms": float(lat.max()),
        "batched": sum(r.mode == "batch" for r in results.values()),
        "streamed": sum(r.mode == "stream" for r in results.values()),
        "failed": sum(r.status != 0 for r in results.values()),
        "errors": sum(r.error is not None for r in results.values()),
    }
