"""Continuous-batching server (cudanextgenpb_amd.serve) on the GPU.

* Each result is the same alone and in the mix: every request served on its own and
  within a stream of overlapping requests (batched and streamed) gives bit-identical
  potentials and energies, also when the requests are prepared from their atoms on the
  device.
* Mixed-precision results agree with the FP64 solve to 1e-6 relative in energy.
* The footprint prediction used for admission bounds the device bytes a request holds.
* A real-time stream completes with every solve converged and reports p50/p99.
"""
import functools

import numpy as np
import pytest

from cudanextgenpb_amd import serve

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def stream():
    reqs = serve.make_request_stream(n_requests=6, rate_hz=40.0, seed=5, atoms_lo=30, atoms_hi=400, h_fine=1.0,
                                     fine_band=2.0, mid_band=4.0)
    for r in reqs:
        print(f"[serve case] request {r.rid}: atoms={r.problem.atoms.n} N={r.n} predicted={r.predicted_bytes}",
              flush=True)
    return reqs


def small_max():
    # median size: about half of the requests are batched, the rest streamed
    return int(np.median([r.n for r in stream()]))


@functools.lru_cache(maxsize=None)
def mix():
    s = serve.Server(budget_bytes=250e9, small_max_unknowns=small_max(), max_batch=8, keep_phi=True)
    try:
        return s.run(list(stream()), realtime=False)
    finally:
        s.close()


def test_alone_equals_mix():
    m = mix()
    assert {r.mode for r in m.values()} == {"batch", "stream"}
    for req in stream():
        s = serve.Server(budget_bytes=250e9, small_max_unknowns=small_max(), max_batch=8, keep_phi=True)
        try:
            a = s.run([req], realtime=False)[req.rid]
        finally:
            s.close()
        b = m[req.rid]
        assert a.mode == b.mode
        assert a.status == 0 and b.status == 0
        assert np.array_equal(a.phi, b.phi), req.rid
        assert a.energy["polar"] == b.energy["polar"], req.rid
        assert a.iterations == b.iterations


def test_mix_from_atoms_equals_mix_from_host_meshes():
    """The same stream with every request prepared from its atoms on the device (octree,
    rays and owners built by the preparing threads while the batch replays) gives
    bit-identical potentials and energies."""
    m = mix()
    s = serve.Server(budget_bytes=250e9, small_max_unknowns=small_max(), max_batch=8, keep_phi=True,
                     device_mesh=True)
    try:
        d = s.run(list(stream()), realtime=False)
    finally:
        s.close()
    for req in stream():
        a, b = d[req.rid], m[req.rid]
        assert a.mode == b.mode and a.status == 0
        assert np.array_equal(a.phi, b.phi), req.rid
        assert a.energy["polar"] == b.energy["polar"], req.rid


def test_mixed_energy_vs_fp64_and_footprint():
    m = mix()
    seq = serve.run_sequential_fp64(list(stream()), realtime=False)
    for req in stream():
        e, e64 = m[req.rid].energy["polar"], seq[req.rid].energy["polar"]
        assert abs(e - e64) <= 1e-6 * abs(e64), (req.rid, e, e64)
        assert req.predicted_bytes >= m[req.rid].device_bytes, (req.rid, req.predicted_bytes, m[req.rid].device_bytes)
        assert req.predicted_bytes <= 4 * m[req.rid].device_bytes + (16 << 20)


def test_realtime_stream_reports_latency():
    s = serve.Server(budget_bytes=250e9, small_max_unknowns=small_max(), max_batch=8)
    try:
        res = s.run(list(stream()), realtime=True)
    finally:
        s.close()
    summ = serve.summarize(res)
    print(summ, flush=True)
    assert summ["requests"] == len(stream()) and summ["failed"] == 0
    assert 0 < summ["p50_latency_ms"] <= summ["p99_latency_ms"]
    assert all(r.done >= r.admitted >= r.arrival - 1e-9 for r in res.values())


def test_admission_respects_budget():
    # a budget that holds about one request at a time still serves everything
    big = max(r.predicted_bytes for r in stream())
    s = serve.Server(budget_bytes=big * 1.05, small_max_unknowns=small_max(), max_batch=8)
    try:
        res = s.run(list(stream()), realtime=False)
        assert s.peak_bytes <= big * 1.05
    finally:
        s.close()
    assert len(res) == len(stream()) and all(r.status == 0 for r in res.values())
