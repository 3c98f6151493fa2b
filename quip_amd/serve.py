"""Continuous-batching server over libgapx.so (gapx_server_* entry points).

    srv = Server(pot, n_slots=2, hbm_budget=32 << 30, max_batch_atoms=65536)
    t = srv.submit(at)                  # returns at once
    r = srv.wait(t)                     # {"energy", "force", "virial", "latency_s"}

Requests are evaluated with the physics of Potential.calc; a free slot batches
whatever is queued (first in, first out) within its share of the HBM budget.

A client that resubmits one structure as it moves opens a session, which keeps its
neighbour list (with a skin) and descriptor buffers resident on the device:

    with srv.session(skin=0.5) as s:
        t = s.submit(at)                # evaluated in order, list reused within the skin
        r = srv.wait(t)
        s.stats()                       # {"requests", "nl_reused", "results_reused"}
"""
import ctypes as C

import numpy as np

from . import _lib as L


class Server:
    def __init__(self, pot, n_slots=2, hbm_budget=32 << 30, max_batch_atoms=65536):
        self.pot = pot   # keeps the model alive while the server runs
        h = C.c_void_p()
        L.check(L.lib().gapx_server_create(pot.handle, n_slots, hbm_budget, max_batch_atoms, C.byref(h)),
                "gapx_server_create")
        self._h = h
        self._n = {}

    @staticmethod
    def _want(want):
        w = L.WANT_E
        for key, bit in (("force", L.WANT_F), ("virial", L.WANT_V), ("local_energy", L.WANT_LOCAL_E),
                         ("local_virial", L.WANT_LOCAL_V)):
            if key in want:
                w |= bit
        return w

    def submit(self, at, want=("energy", "force", "virial")):
        w = self._want(want)
        t = C.c_int64()
        pos, Z, lat, pbc = L.f64(at.positions), L.i32(at.numbers), L.f64(at.cell), L.i32(at.pbc)
        L.check(L.lib().gapx_server_submit(self._h, len(at), L.ptr(pos), L.ptr(Z), L.ptr(lat), L.ptr(pbc), w,
                                           C.byref(t)), "gapx_server_submit")
        self._n[t.value] = (len(at), w)
        return t.value

    def session(self, skin=0.5):
        """A resident client structure (gapx_server_session_open)."""
        return Session(self, skin)

    def done(self, ticket):
        d = C.c_int()
        L.check(L.lib().gapx_server_poll(self._h, ticket, C.byref(d)), "gapx_server_poll")
        return bool(d.value)

    def wait(self, ticket):
        n, w = self._n.pop(ticket)
        e, v = np.zeros(1), np.zeros(9)
        f = np.zeros((n, 3)) if w & L.WANT_F else None
        le = np.zeros(n) if w & L.WANT_LOCAL_E else None
        lv = np.zeros((n, 9)) if w & L.WANT_LOCAL_V else None
        lat = C.c_double()
        L.check(L.lib().gapx_server_wait(self._h, ticket, L.ptr(e), L.ptr(le), L.ptr(f), L.ptr(v), L.ptr(lv),
                                         C.byref(lat)), "gapx_server_wait")
        r = {"energy": e[0], "virial": v.reshape(3, 3).T, "latency_s": lat.value}
        if f is not None:
            r["force"] = f
        if le is not None:
            r["local_energy"] = le
        if lv is not None:
            r["local_virial"] = lv
        return r

    def stats(self):
        nd, nb, na, busy = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        L.check(L.lib().gapx_server_stats(self._h, C.byref(nd), C.byref(nb), C.byref(na), C.byref(busy)),
                "gapx_server_stats")
        return dict(requests=nd.value, batches=nb.value, atoms=na.value, busy_s=busy.value)

    def close(self):
        if self._h:
            L.check(L.lib().gapx_server_free(self._h), "gapx_server_free")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Session:
    """A client's resident structure on a Server: requests are evaluated in order with
    the neighbour list (model cutoff + skin) kept within the skin."""

    def __init__(self, server, skin=0.5):
        self.server = server
        h = C.c_int64()
        L.check(L.lib().gapx_server_session_open(server._h, float(skin), C.byref(h)), "gapx_server_session_open")
        self._id = h.value

    def submit(self, at, want=("energy", "force", "virial")):
        w = Server._want(want)
        t = C.c_int64()
        pos, Z, lat, pbc = L.f64(at.positions), L.i32(at.numbers), L.f64(at.cell), L.i32(at.pbc)
        L.check(L.lib().gapx_server_session_submit(self.server._h, self._id, len(at), L.ptr(pos), L.ptr(Z),
                                                   L.ptr(lat), L.ptr(pbc), w, C.byref(t)),
                "gapx_server_session_submit")
        self.server._n[t.value] = (len(at), w)
        return t.value

    def stats(self):
        n, nl, rr = C.c_int64(), C.c_int64(), C.c_int64()
        L.check(L.lib().gapx_server_session_stats(self.server._h, self._id, C.byref(n), C.byref(nl), C.byref(rr)),
                "gapx_server_session_stats")
        return dict(requests=n.value, nl_reused=nl.value, results_reused=rr.value)

    def close(self):
        if self._id is not None and self.server._h:
            L.check(L.lib().gapx_server_session_close(self.server._h, self._id), "gapx_server_session_close")
        self._id = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
