"""A serving loop for scene requests on one GPU, with continuous batching.

A request is one scene: an ordered list of fillPath / strokePath calls on a
fresh transparent canvas of its own size, optionally blurred afterwards.
Pixie has no server (it is a library); this module is the serving form of
the raster path.

Continuous batching works like this:

* Requests queue in arrival order.
* A request is admitted as soon as an execution lane is free and its canvas
  fits in the HBM arena. The arena is one device allocation sized by the
  HBM budget, carved first-fit.
* Each admitted request runs on its own lane, an independent device stream
  driven by its own host thread. Requests that arrive while others are on
  the GPU therefore join while the others are still running, and their
  kernels overlap on the device.
* Recording a request's scene on the host overlaps with the other lanes'
  device work.

Canvases and scene buffers stay resident:

* The arena is allocated once.
* Each lane keeps one scene object whose device buffers only grow.

The plain baseline serves the same arrivals one request at a time, with one
fillPath / strokePath launch per call.
"""
import ctypes
import threading
import time

import numpy as np

from . import Image, PixieError, Scene, lane_synchronize, lanes, lib


class Request:
    __slots__ = ("calls", "side", "blur", "arrival")

    def __init__(self, calls, side, blur, arrival):
        self.calls = calls        # list of workloads.FillCmd
        self.side = int(side)     # square canvas side in px
        self.blur = float(blur)   # blur radius after the paths (0 = none)
        self.arrival = float(arrival)


class Result:
    """Outcome of one request.  `error` is None, or the PixieError the
    request raised (for example a capacity error): a failing request is
    answered with its error and the others are served as usual."""
    __slots__ = ("index", "admitted", "done", "latency", "pixels", "error")

    def __init__(self, index):
        self.index = index
        self.admitted = self.done = self.latency = 0.0
        self.pixels = None
        self.error = None


class _Arena:
    """First-fit allocator over one device allocation (byte offsets)."""

    def __init__(self, nbytes):
        rows = (nbytes + 65535) // 65536
        self.image = Image(65536, rows, a8=True)
        self.base = self.image.device_ptr()
        self.size = 65536 * rows
        self.free = [(0, self.size)]  # sorted (offset, length)

    def alloc(self, n):
        n = (n + 255) & ~255
        for i, (off, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (off + n, ln - n)
                return off, n
        return None

    def release(self, off, n):
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((o, ln))
        self.free = merged


def _record(scene, calls):
    for c in calls:
        if c.stroke:
            scene.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap, lineJoin=c.join,
                             miterLimit=c.miter, blendMode=c.mode)
        else:
            scene.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)


class Server:
    """Serves requests on `n_lanes` execution lanes within `hbm_budget`
    bytes of canvas memory."""

    def __init__(self, n_lanes=None, hbm_budget=1 << 30, keep_pixels=False):
        self.n_lanes = n_lanes or (lanes() - 1)
        if not 1 <= self.n_lanes <= lanes() - 1:
            raise ValueError("n_lanes must be in 1..%d" % (lanes() - 1))
        self.arena = _Arena(hbm_budget)
        self.keep_pixels = keep_pixels
        self.scenes = [Scene() for _ in range(self.n_lanes)]

    def run(self, requests):
        """Serves `requests` (arrival times in seconds from the start) and
        returns one Result per request, in request order."""
        n = len(requests)
        results = [Result(i) for i in range(n)]
        cv = threading.Condition()
        state = {"next": 0, "error": None}
        biggest = max(r.side * r.side * 4 for r in requests) if requests else 0
        if biggest > self.arena.size:
            raise ValueError("a canvas is larger than the HBM budget")
        t0 = time.perf_counter() + 0.01

        def take():
            # FIFO admission: the oldest arrived request, once its canvas fits
            with cv:
                while True:
                    if state["error"] is not None or state["next"] >= n:
                        return None
                    i = state["next"]
                    r = requests[i]
                    wait = t0 + r.arrival - time.perf_counter()
                    if wait > 0:
                        cv.wait(timeout=wait)
                        continue
                    blk = self.arena.alloc(r.side * r.side * 4)
                    if blk is None:
                        cv.wait(timeout=0.001)
                        continue
                    state["next"] = i + 1
                    cv.notify_all()
                    return i, blk

        def worker(k):
            lane = k + 1
            scene = self.scenes[k]
            try:
                while True:
                    got = take()
                    if got is None:
                        return
                    i, (off, nbytes) = got
                    r = requests[i]
                    res = results[i]
                    res.admitted = time.perf_counter() - t0
                    img = Image.view(self.arena.base + off, r.side, r.side)
                    try:
                        img.fill((0, 0, 0, 0), lane=lane)
                        scene.clear()
                        _record(scene, r.calls)
                        scene.render(img, lane=lane)
                        if r.blur:
                            img.blur(r.blur, lane=lane)
                        lane_synchronize(lane)
                        res.done = time.perf_counter() - t0
                        if self.keep_pixels:
                            res.pixels = img.download()
                    except PixieError as e:
                        # this request fails alone: its lane is drained so
                        # that nothing of it is left on the device, and the
                        # lane goes on with the next request
                        res.error = e
                        try:
                            lane_synchronize(lane)
                        except PixieError:
                            pass
                        res.done = time.perf_counter() - t0
                    res.latency = res.done - r.arrival
                    del img
                    with cv:
                        self.arena.release(off, nbytes)
                        cv.notify_all()
            except Exception as e:  # surfaced to the caller below
                with cv:
                    state["error"] = e
                    cv.notify_all()

        threads = [threading.Thread(target=worker, args=(k,)) for k in range(self.n_lanes)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if state["error"] is not None:
            raise state["error"]
        return results


class NativeServer:
    """The same continuous batching, run by the library's own server
    (pxh_server_*, the C-ABI served form): worker threads per lane in C++,
    FIFO admission by lane and arena budget.  Requests are recorded into
    scenes as they arrive and then enqueued; results are read back and
    released as they finish."""

    def __init__(self, n_lanes=None, hbm_budget=1 << 30, keep_pixels=False):
        self.n_lanes = n_lanes or (lanes() - 1)
        h = lib().pxh_server_new(self.n_lanes, int(hbm_budget))
        if not h:
            raise PixieError(6, "pxh_server_new failed (lanes 1..%d, a budget that fits HBM)"
                             % (lanes() - 1))
        self._h = ctypes.c_void_p(h)
        self.budget = int(hbm_budget)
        self.keep_pixels = keep_pixels

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None:
            lib().pxh_server_free(h)
            self._h = None

    def _finish(self, ticket, res, offset, arrival):
        L = lib()
        st = ctypes.c_int32()
        L.pxh_server_wait(self._h, ticket, ctypes.byref(st))
        e, a, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        L.pxh_server_times(self._h, ticket, ctypes.byref(e), ctypes.byref(a), ctypes.byref(d))
        res.admitted = a.value + offset
        res.done = d.value + offset
        res.latency = res.done - arrival
        if st.value:
            res.error = PixieError(st.value, L.pxh_server_message(self._h, ticket).decode())
        elif self.keep_pixels:
            img = L.pxh_server_image(self._h, ticket)
            w = L.pxh_image_width(img)
            h = L.pxh_image_height(img)
            out = np.empty((h, w, 4), np.uint8)
            L.pxh_image_download(img, out.ctypes.data)
            res.pixels = out
        L.pxh_server_release(self._h, ticket)

    def run(self, requests):
        """Serves `requests` (arrival times in seconds from the start) and
        returns one Result per request, in request order.  Each request
        arrives as a recorded scene, as a C caller would hand it over: the
        scenes are recorded before the clock starts (recording through
        ctypes from one Python thread would otherwise bound the arrival
        rate), so latency runs from arrival to the finished canvas."""
        L = lib()
        n = len(requests)
        results = [Result(i) for i in range(n)]
        if any(r.side * r.side * 4 > self.budget for r in requests):
            raise ValueError("a canvas is larger than the HBM budget")
        recorded = []
        for r in requests:
            sc = Scene()
            _record(sc, r.calls)
            recorded.append(sc)
        t0 = time.perf_counter() + 0.01
        offset = time.perf_counter() - L.pxh_server_now(self._h) - t0
        pending = []   # (ticket, index, scene): a scene lives until its request is done
        for i, r in enumerate(requests):
            wait = t0 + r.arrival - time.perf_counter()
            if wait > 0:
                # finish what is done while waiting for the next arrival
                while pending and wait > 0.0005:
                    done = ctypes.c_int32()
                    L.pxh_server_poll(self._h, pending[0][0], ctypes.byref(done), None)
                    if not done.value:
                        break
                    tk, j, _sc = pending.pop(0)
                    self._finish(tk, results[j], offset, requests[j].arrival)
                    wait = t0 + r.arrival - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
            sc = recorded[i]
            tk = L.pxh_server_enqueue(self._h, sc._h, r.side, r.side, float(r.blur))
            if tk < 0:
                raise PixieError(-tk, "pxh_server_enqueue failed")
            results[i].admitted = 0.0
            pending.append((tk, i, sc))
        for tk, j, _sc in pending:
            self._finish(tk, results[j], offset, requests[j].arrival)
        return results


def serve_plain(requests, keep_pixels=False):
    """The same arrivals served one request at a time with one launch per
    call (the plain per-call form) on lane 0."""
    results = [Result(i) for i in range(len(requests))]
    t0 = time.perf_counter() + 0.01
    for i, r in enumerate(requests):
        wait = t0 + r.arrival - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        res = results[i]
        res.admitted = time.perf_counter() - t0
        img = Image(r.side, r.side)
        try:
            for c in r.calls:
                if c.stroke:
                    img.strokePath(c.cmds, c.color, strokeWidth=c.width, lineCap=c.cap,
                                   lineJoin=c.join, miterLimit=c.miter, blendMode=c.mode)
                else:
                    img.fillPath(c.cmds, c.color, windingRule=c.winding, blendMode=c.mode)
            if r.blur:
                img.blur(r.blur)
            lane_synchronize(0)
        except PixieError as e:
            res.error = e
        res.done = time.perf_counter() - t0
        res.latency = res.done - r.arrival
        if keep_pixels and res.error is None:
            res.pixels = img.download()
    return results


def latency_stats(results, requests):
    import numpy as np
    ok = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in ok]) * 1e3
    span = max(r.done for r in results) - min(q.arrival for q in requests)
    return {"scenes": len(ok), "failed": len(results) - len(ok),
            "scenes_per_s": round(len(ok) / span, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "max_ms": round(float(lat.max()), 3)}
