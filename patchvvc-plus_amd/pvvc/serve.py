"""Open-loop load driver for the serving loop.

Streams send frames at Poisson arrival times (exponential gaps, mean
1/fps, a random phase per stream).  A submitter thread hands each frame to
the server at its arrival time, stamped with that scheduled time; a
consumer thread polls completions and hands the output slots back at once.
The server measures arrival -> completion per frame.

Each stream cycles through a GoP of pre-parsed frames (frame k of the
stream is gop[k % len(gop)]), so a stream is a valid decode order of
intra and predictive frames for any length.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass

import numpy as np

from . import Frame, Server


@dataclass
class StreamSpec:
    name: str
    gop: list[Frame]

    def dims(self) -> tuple[int, int, int]:
        metas = [f.meta() for f in self.gop]
        return (max(int(m["index"].max()) for m in metas) + 1, max(int(m["size"].max()) for m in metas),
                max(f.points for f in self.gop))


def poisson_schedule(n_streams: int, fps: float, duration: float, seed: int) -> list[tuple[float, int, int]]:
    """(time, stream, frame number) sorted by time."""
    rng = np.random.default_rng(seed)
    ev = []
    for s in range(n_streams):
        t, k = float(rng.uniform(0.0, 1.0 / fps)), 0
        while t < duration:
            ev.append((t, s, k))
            k += 1
            t += float(rng.exponential(1.0 / fps))
    ev.sort()
    return ev


def run(server: Server, specs: list[StreamSpec], fps: float = 30.0, duration: float = 3.0, seed: int = 0,
        check=None) -> dict:
    """Drives the server with Poisson arrivals; returns the server's stats
    plus offered load.  `check(stream_index, frame_number, completion)` is
    called for every completion before its slot is released, if given."""
    ids = [server.open_stream(*sp.dims()) for sp in specs]
    sched = poisson_schedule(len(specs), fps, duration, seed)
    total = len(sched)
    done = {"n": 0, "bad": 0}
    waits, services, per_kind = [], [], {}
    idx_of = {sid: i for i, sid in enumerate(ids)}
    stop = threading.Event()

    def consume():
        while done["n"] < total and not stop.is_set():
            for c in server.poll(512, 50):
                if c["status"] != 0:
                    done["bad"] += 1
                waits.append(c["launch_s"] - c["arrival_s"])
                services.append(c["done_s"] - c["launch_s"])
                per_kind.setdefault(specs[idx_of[c["stream_id"]]].name, []).append(c["done_s"] - c["arrival_s"])
                if check is not None:
                    check(idx_of[c["stream_id"]], c["seq"], c)
                server.release(c)
                done["n"] += 1

    th = threading.Thread(target=consume, daemon=True)
    th.start()
    server.reset_stats()
    t0 = server.clock() + 0.05
    offered_pts = 0
    late = []
    for t, s, k in sched:
        target = t0 + t
        while True:
            d = target - server.clock()
            if d <= 0:
                break
            time.sleep(d if d > 0.002 else 0)
        late.append(server.clock() - target)
        f = specs[s].gop[k % len(specs[s].gop)]
        server.submit(ids[s], f, arrival=target)
        offered_pts += f.points
    th.join(timeout=max(30.0, 4 * duration))
    stop.set()
    st = server.stats()
    if st["streams_waiting"]:
        # streams that never fit the budget / slot table / output pool stay waiting
        st["note"] = f"{st['streams_waiting']} streams were never admitted"
    for sid in ids:
        server.close_stream(sid)
    q = lambda a, p: float(np.percentile(np.asarray(a) * 1e3, p)) if len(a) else 0.0  # noqa: E731
    st.update(queue_wait_p50_ms=q(waits, 50), queue_wait_p99_ms=q(waits, 99),
              service_p50_ms=q(services, 50), service_p99_ms=q(services, 99),
              p99_ms_by_kind={k: round(q(v, 99), 3) for k, v in per_kind.items()},
              frames_per_batch=(done["n"] / st["batches"]) if st["batches"] else 0.0)
    st.update(frames_offered=total, frames_completed=done["n"], frames_failed=done["bad"], streams=len(specs),
              fps_per_stream=fps, duration_s=duration,
              offered_mpts_per_s=offered_pts / duration / 1e6,
              submit_late_p99_ms=float(np.percentile(np.asarray(late) * 1e3, 99)) if late else 0.0)
    return st
