"""The continuous-batching serving loop (sieve_server, include/sieve_gpu.h) and an open-loop
load driver: Poisson arrivals at a target rate, per-request latency, p50/p99 (the reference's
percentile, src/bin/search.rs:400-402) and achieved QPS."""
import ctypes
import threading
import time

import numpy as np

from . import abi
from .evaluate import pct


class Server:
    def __init__(self, index, params, max_batch=1024, max_inflight=0, hbm_budget=0, live=None):
        self._ix = index
        self._live = live
        self.k = params.k
        h = ctypes.c_void_p()
        o = abi.ServerOpts(max_batch=max_batch, max_inflight=max_inflight, hbm_budget=hbm_budget,
                           live=live._h.value if live is not None else None)
        abi.check(abi.lib().sieve_server_create(index._h, ctypes.byref(abi.make_params(params)), ctypes.byref(o),
                                                ctypes.byref(h)))
        self._h = h

    def submit(self, q768, q256):
        q7 = np.ascontiguousarray(q768, np.float32)
        q2 = np.ascontiguousarray(q256, np.float32)
        t = ctypes.c_uint64()
        abi.check(abi.lib().sieve_server_submit(self._h, abi.ptr(q7), abi.ptr(q2), ctypes.byref(t)))
        return t.value

    def submit_filtered(self, q768, q256, filt, plan="auto"):
        q7 = np.ascontiguousarray(q768, np.float32)
        q2 = np.ascontiguousarray(q256, np.float32)
        t = ctypes.c_uint64()
        abi.check(abi.lib().sieve_server_submit_filtered(self._h, abi.ptr(q7), abi.ptr(q2), filt._h, abi.PLANS[plan],
                                                         ctypes.byref(t)))
        return t.value

    def insert(self, ext, vecs):
        e = np.ascontiguousarray(ext, np.uint32)
        v = np.ascontiguousarray(vecs, np.float32)
        t = ctypes.c_uint64()
        abi.check(abi.lib().sieve_server_insert(self._h, abi.ptr(e), abi.ptr(v), e.shape[0], ctypes.byref(t)))
        return t.value

    def delete(self, ext):
        e = np.ascontiguousarray(ext, np.uint32)
        t = ctypes.c_uint64()
        abi.check(abi.lib().sieve_server_delete(self._h, abi.ptr(e), e.shape[0], ctypes.byref(t)))
        return t.value

    def wait(self, ticket):
        s = np.empty(self.k, np.float32)
        r = np.empty(self.k, np.uint32)
        n = ctypes.c_uint32()
        info = abi.RequestInfo()
        abi.check(abi.lib().sieve_server_wait(self._h, ticket, abi.ptr(s), abi.ptr(r), ctypes.byref(n),
                                              ctypes.byref(info)))
        return s[:n.value], r[:n.value], info

    def stats(self):
        st = abi.ServerStats()
        abi.check(abi.lib().sieve_server_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def close(self):
        if getattr(self, "_h", None):
            abi.lib().sieve_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def open_loop(server, q768, q256, rate, seconds, seed=0, truth=None):
    """Submits queries (cycling through the given ones) at Poisson arrivals of `rate`/s for
    `seconds`, waits for every result, and returns latency percentiles and throughput. With
    `truth` (exact top-10 row lists of the first len(truth) queries), also the mean R@10 of the
    results served for those queries (eval::recall_at, src/eval.rs:7-10)."""
    from .evaluate import recall_at
    rng = np.random.default_rng(seed)
    tickets = []
    lock = threading.Lock()
    done = threading.Event()
    lat, batch, rec = [], [], []

    def collect():
        i = 0
        while True:
            with lock:
                have = i < len(tickets)
                t = tickets[i] if have else None
            if not have:
                if done.is_set():
                    with lock:
                        if i >= len(tickets):
                            return
                    continue
                time.sleep(0.0002)
                continue
            _, rows, info = server.wait(t)
            lat.append(info.total_ms)
            batch.append(info.batch_size)
            j = i % len(q768)
            if truth is not None and j < len(truth):
                rec.append(recall_at(10, rows, truth[j]))
            i += 1

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    nxt = t0
    i = 0
    while True:
        now = time.perf_counter()
        if now - t0 >= seconds:
            break
        if now < nxt:
            time.sleep(min(nxt - now, 0.001))
            continue
        j = i % len(q768)
        tk = server.submit(q768[j], q256[j])
        with lock:
            tickets.append(tk)
        i += 1
        nxt += rng.exponential(1.0 / rate)
    done.set()
    th.join()
    wall = time.perf_counter() - t0
    lat.sort()
    out = dict(offered_qps=rate, achieved_qps=round(len(lat) / wall, 1), requests=len(lat),
               p50_ms=round(pct(lat, 0.5), 3), p99_ms=round(pct(lat, 0.99), 3),
               mean_batch=round(float(np.mean(batch)), 1))
    if rec:
        out["recall_at_10"] = round(float(np.mean(rec)), 4)
        out["recall_requests"] = len(rec)
    return out
