"""Serving loop: a stream of reconstruction requests, admitted against an HBM
budget and reconstructed in batches of one octree depth.

The reference's driver reconstructs one cloud per process run
(main.cpp:8-41 builds a PoissonReconstruction and calls
SolvePoissionReconstructionMesh once, mesh/PoissonReconstruction.cpp:195-245).
Here requests of any size arrive over time; the server keeps one context per
depth and, whenever the GPU is free, takes the depth of the oldest waiting
request and reconstructs every waiting request of that depth that fits, in
arrival order, with one batched call (psr_reconstruct_batch).

Admission.  A request's device-memory need is estimated as its point count
times the largest bytes-per-point seen so far at its depth (the run's
psr_stats.device_bytes over its points), starting from a conservative prior.
A batch takes requests while the sum of their estimates stays within the
budget (default: half the free device memory when the server starts); a
request larger than the whole budget is served alone.
"""
import collections
import time

from . import binding as B

# Device bytes per input point assumed before a depth has served a request:
# about 1.5 times what the runs' buffers need for closed surfaces of 10k-1M
# points (measured: 7.3k, 9.2k, 17.9k and 9.3k bytes per point at depths 6-9).
PRIOR_BYTES_PER_POINT = {6: 12000, 7: 14000, 8: 27000, 9: 14000}


class Server:
    def __init__(self, depths=(6, 7, 8, 9), precision=B.PRECISION_MIXED, budget_bytes=None, max_batch=64,
                 batch_cg_plan=0):
        import torch
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info()
            budget_bytes = free // 2
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
        self.ctx = {d: B.PoissonReconstruction(depth=d, precision=precision, batch_workers=max_batch,
                                               batch_cg_plan=batch_cg_plan) for d in depths}
        self.bytes_per_point = dict(PRIOR_BYTES_PER_POINT)
        self._seen = set()

    def close(self):
        for c in self.ctx.values():
            c.close()
        self.ctx = {}

    def estimate(self, n, depth):
        return int(n * self.bytes_per_point[depth])

    def _learn(self, depth, n, used):
        per = used / max(n, 1)
        if depth not in self._seen:
            self._seen.add(depth)
            self.bytes_per_point[depth] = per * 1.25
        else:
            self.bytes_per_point[depth] = max(self.bytes_per_point[depth], per * 1.25)

    def _admit(self, waiting):
        depth = waiting[0]["depth"]
        batch, used = [], 0
        for r in waiting:
            if r["depth"] != depth:
                continue
            need = self.estimate(r["points"], depth)
            if batch and (len(batch) == self.max_batch or used + need > self.budget):
                break
            r["estimate"] = need
            batch.append(r)
            used += need
        return depth, batch

    def serve(self, requests, realtime=True, on_mesh=None):
        """requests: list of (arrival_s, xyz, normals, depth) with xyz and
        normals float32 CUDA tensors of shape (N, 3), arrivals increasing.
        With realtime the loop waits for each arrival on the wall clock
        (latency = completion - arrival); without, every request is waiting
        from the start.  on_mesh(index, psr_mesh) is called for each request
        while its mesh buffers are valid.  Returns one record per request."""
        import torch
        recs = [{"index": i, "arrival": float(a), "xyz": x, "nrm": n, "depth": int(d), "points": int(x.shape[0])}
                for i, (a, x, n, d) in enumerate(requests)]
        for r in recs:
            if r["depth"] not in self.ctx:
                raise ValueError("request %d: depth %d has no context" % (r["index"], r["depth"]))
        torch.cuda.synchronize()
        waiting = collections.deque()
        nxt = 0
        t0 = time.perf_counter()
        batches = 0
        while nxt < len(recs) or waiting:
            now = time.perf_counter() - t0
            while nxt < len(recs) and (not realtime or recs[nxt]["arrival"] <= now):
                waiting.append(recs[nxt])
                nxt += 1
            if not waiting:
                time.sleep(max(0.0, recs[nxt]["arrival"] - now))
                continue
            depth, batch = self._admit(waiting)
            for r in batch:
                waiting.remove(r)
            start = time.perf_counter() - t0
            ctx = self.ctx[depth]
            meshes = ctx.ReconstructBatch([(r["xyz"], r["nrm"]) for r in batch])
            done = time.perf_counter() - t0
            batches += 1
            for k, (r, m) in enumerate(zip(batch, meshes)):
                st = ctx.batch_stats(k)
                self._learn(depth, r["points"], st["device_bytes"])
                r.update(start=start, done=done, latency=done - (r["arrival"] if realtime else 0.0),
                         batch=batches, batch_size=len(batch), n_vertices=m.n_vertices, n_triangles=m.n_triangles,
                         device_bytes=st["device_bytes"])
                if on_mesh is not None:
                    on_mesh(r["index"], m)
        for r in recs:
            del r["xyz"], r["nrm"]
        return recs


def summary(recs, rate_hz):
    """Throughput and latency percentiles of a served stream."""
    import numpy as np
    lat = np.array([r["latency"] for r in recs]) * 1000.0
    span = max(r["done"] for r in recs) - min(r["arrival"] for r in recs)
    return {"requests": len(recs), "arrival_rate_hz": rate_hz, "recon_per_s": len(recs) / span,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_batch": float(np.mean([r["batch_size"] for r in recs])),
            "batches": int(max(r["batch"] for r in recs)),
            "points": int(sum(r["points"] for r in recs)),
            "per_depth": {int(d): sum(1 for r in recs if r["depth"] == d) for d in sorted({r["depth"] for r in recs})}}
