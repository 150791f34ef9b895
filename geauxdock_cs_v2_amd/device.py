"""Python front end of the device C ABI: one GPU context, resident targets and
jobs, the plain chunk (MonteCarlo_d), the batched form, pose energies,
exchange, and the reference-shaped Docker driver."""
import ctypes as C

import numpy as np

from . import ctypes_defs as D
from . import lib as L


class Device:
    def __init__(self, index=0):
        self.lib = L.load()
        self.h = C.c_void_p()
        L.check(self.lib.gd_create(index, C.byref(self.h)))

    def close(self):
        if self.h:
            self.lib.gd_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        name = C.create_string_buffer(64)
        cu = C.c_int()
        mem = C.c_size_t()
        L.check(self.lib.gd_device_info(self.h, name, 64, C.byref(cu), C.byref(mem)))
        return {"arch": name.value.decode(), "cu": cu.value, "hbm_bytes": mem.value}

    def upload_target(self, cx):
        t = C.c_int()
        L.check(self.lib.gd_upload_target(self.h, C.byref(cx.view), C.byref(t)))
        return t.value

    def upload_job(self, target, cx, seed):
        j = C.c_int()
        L.check(self.lib.gd_upload_job(self.h, target, C.byref(cx.view), int(seed), C.byref(j)))
        return j.value

    def set_stream(self, job, seed, k_words=None):
        """Re-key the job's stream; k_words (uint32 [n_rep, n_steps, 7], values
        <= RAND_MAX) replaces the Philox MC draws for steps 1..n_steps."""
        if k_words is None:
            L.check(self.lib.gd_set_stream(self.h, job, int(seed), None, 0))
            return
        w = np.ascontiguousarray(k_words, np.uint32)
        if w.ndim != 3 or w.shape[2] != 7:
            raise ValueError("k_words must be [n_rep, n_steps, 7]")
        L.check(self.lib.gd_set_stream(self.h, job, int(seed), w.ctypes.data_as(C.POINTER(C.c_uint32)),
                                       w.shape[1]))

    def job_shape(self, job):
        """(waves, slot waves, point slices, replicas per workgroup) of the
        job's last batched launch."""
        s = (C.c_int * 4)()
        L.check(self.lib.gd_job_shape(self.h, job, s))
        return tuple(s)

    def batch_plan(self):
        """(tasks, workgroups launched, step slices per task) of the last
        persistent launch."""
        s = (C.c_int * 3)()
        L.check(self.lib.gd_batch_plan(self.h, s))
        return tuple(s)

    def free_job(self, job):
        L.check(self.lib.gd_free_job(self.h, job))

    def free_target(self, target):
        L.check(self.lib.gd_free_target(self.h, target))

    def montecarlo(self, job, s1, s2max):
        L.check(self.lib.gd_montecarlo(self.h, job, s1, s2max))

    def exchange(self, job, rnd):
        acc = C.c_int()
        L.check(self.lib.gd_exchange(self.h, job, rnd, C.byref(acc)))
        return acc.value

    def energy(self, job, rep_idx, move):
        rep_idx = np.ascontiguousarray(rep_idx, np.int32)
        move = np.ascontiguousarray(move, np.float32).reshape(-1, 6)
        out = np.zeros((len(rep_idx), 10), np.float32)
        L.check(self.lib.gd_energy(self.h, job, len(rep_idx), rep_idx.ctypes.data_as(C.POINTER(C.c_int)),
                                   move.ctypes.data_as(C.POINTER(C.c_float)),
                                   out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def replicas(self, job, n):
        arr = (D.ReplicaMC * n)()
        L.check(self.lib.gd_read_replicas(self.h, job, arr))
        return arr

    def write_replicas(self, job, arr):
        L.check(self.lib.gd_write_replicas(self.h, job, arr))

    def records(self, job, n, cap=100):
        rec = (D.ReplicaMC * (n * cap))()
        ne = np.zeros(n, np.int32)
        L.check(self.lib.gd_read_records(self.h, job, rec, ne.ctypes.data_as(C.POINTER(C.c_int)), cap))
        return rec, ne

    def batch_run(self, jobs, steps, exchange=0, graph=True, compensated=True, records=False, waves=0,
                  slots_per_wave=0, point_slices=0):
        """Advance every job by `steps` MC steps in fused launches; returns the
        device time in ms (hipEvents around the chunk sequence)."""
        o = D.BatchOpts(int(steps), int(exchange), int(bool(graph)), int(slots_per_wave), int(waves),
                        int(point_slices), int(bool(compensated)), int(bool(records)))
        arr = (C.c_int * len(jobs))(*jobs)
        ms = C.c_double()
        L.check(self.lib.gd_batch_run(self.h, arr, len(jobs), C.byref(o), C.byref(ms)))
        return ms.value


class Server:
    """The serving loop on a Device's resident target (gd_server_*): submit
    staged complexes, poll or drain, read timings on the server clock.  The
    server owns the device context until close()."""

    def __init__(self, dev, target, hbm_budget, max_active=64, chunk=100, exchange=0):
        self.lib = dev.lib
        self.h = C.c_void_p()
        o = D.ServerOpts(int(hbm_budget), int(max_active), int(chunk), int(exchange))
        L.check(self.lib.gd_server_create(dev.h, target, C.byref(o), C.byref(self.h)))
        self.n_rep = {}

    def submit(self, cx, steps, seed):
        i = C.c_int()
        L.check(self.lib.gd_server_submit(self.h, C.byref(cx.view), int(steps), int(seed), C.byref(i)))
        self.n_rep[i.value] = cx.size["n_rep"]
        return i.value

    def clock_ms(self):
        return self.lib.gd_server_clock_ms(self.h)

    def drain(self):
        L.check(self.lib.gd_server_drain(self.h))

    def poll(self, req, replicas=False):
        """(RequestResult, final replicas or None); result.status is
        GD_ERR_BUSY while the request is pending, 0 once done."""
        res = D.RequestResult()
        reps = (D.ReplicaMC * self.n_rep[req])() if replicas else None
        L.check(self.lib.gd_server_poll(self.h, req, C.byref(res), reps))
        return res, reps

    def close(self):
        if self.h:
            self.lib.gd_server_destroy(self.h)
            self.h = None


class Docker:
    """The reference's backend interface (src/backend_gpu/backend_gpu.h:10-23):
    a staged `complex`, the `record` window of the last chunk, and run(),
    which loops MonteCarlo_d chunks of steps_per_dump steps
    (src/backend_gpu/backend_gpu.cu:91-107) with the replica-exchange hook
    filled in (exchange_rounds > 0 exchanges after every chunk)."""

    def __init__(self, device=None, seed=2024, exchange=True):
        self.dev = device or Device(0)
        self.complex = None
        self.record = None
        self.next_entry = None
        self.seed = seed
        self.exchange = exchange
        self.replica = None

    def run(self):
        cx = self.complex
        n = cx.size["n_rep"]
        mc = cx.view.mcpara
        if mc.steps_per_dump <= 0 or mc.steps_total < 0:
            raise ValueError("Docker.run: bad step counts")
        t = self.dev.upload_target(cx)
        try:
            j = self.dev.upload_job(t, cx, self.seed)
            try:
                rnd = 0
                for s1 in range(0, mc.steps_total, mc.steps_per_dump):
                    s2 = min(mc.steps_per_dump, mc.steps_total - s1)
                    self.dev.montecarlo(j, s1, s2)
                    self.record, self.next_entry = self.dev.records(j, n, max(s2, 100))
                    if self.exchange and cx.size["n_tmp"] > 1:
                        self.dev.exchange(j, rnd)
                        rnd += 1
                self.replica = self.dev.replicas(j, n)
            finally:
                self.dev.free_job(j)
        finally:
            self.dev.free_target(t)
