"""ctypes binding of libmesobasic_lbm.so (include/mesobasic_lbm.h).

Python is plumbing here: tests and the bench drive the C ABI through this
module; all lattice arithmetic runs in the HIP kernels.  Loading fails loudly
when the library has not been built -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmesobasic_lbm.so")

FP64, FP32, FP16S, BF16S, FP16N, BF16N = 0, 1, 2, 3, 4, 5
KERNEL_AUTO, KERNEL_TWO_PASS, KERNEL_FUSED = 0, 1, 2
FIELDS_REFERENCE, FIELDS_CURRENT = 0, 1
JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2
ENOMEM, EJOB = -2, -6
PRECISIONS = {"fp64": FP64, "fp32": FP32, "fp16": FP16S, "bf16": BF16S, "fp16n": FP16N,
              "bf16n": BF16N}

_D = ctypes.POINTER(ctypes.c_double)


class LbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lbm error {code}: {msg}")
        self.code = code


class CapsuleDesc(ctypes.Structure):
    _fields_ = [("nnodes", ctypes.c_int), ("kstiff", ctypes.c_double), ("dt", ctypes.c_double),
                ("body_fx", ctypes.c_double), ("body_fy", ctypes.c_double)]


class Desc(ctypes.Structure):
    _fields_ = [("q", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("nz", ctypes.c_int), ("tau", ctypes.c_double), ("G", ctypes.c_double),
                ("order", ctypes.c_int), ("walls", ctypes.c_int), ("precision", ctypes.c_int),
                ("rho_shift", ctypes.c_double), ("kernel", ctypes.c_int),
                ("components", ctypes.c_int), ("tau_b", ctypes.c_double),
                ("rho_sol_a", ctypes.c_double), ("rho_sol_b", ctypes.c_double)]


class BatchDesc(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("rho_shift", ctypes.c_double),
                ("hbm_budget", ctypes.c_int64), ("steps_per_tick", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("autotune", ctypes.c_int)]


class JobDesc(ctypes.Structure):
    _fields_ = [("q", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("tau", ctypes.c_double), ("G", ctypes.c_double), ("steps", ctypes.c_int64)]


class JobStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("steps_done", ctypes.c_int64),
                ("steps_finite", ctypes.c_int64), ("mass", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int), ("chunk", ctypes.c_int), ("order", ctypes.c_int),
                ("nt_store", ctypes.c_int)]


_lib = None


def load():
    """Loads the built library (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m mesobasic_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    sig = {
        "lbm_create": (ctypes.c_int, [ctypes.POINTER(Desc), ctypes.POINTER(P)]),
        "lbm_destroy": (None, [P]),
        "lbm_set_rho": (ctypes.c_int, [P, _D]),
        "lbm_init_equilibrium": (ctypes.c_int, [P]),
        "lbm_step": (ctypes.c_int, [P, ctypes.c_int]),
        "lbm_step_timed": (ctypes.c_int, [P, ctypes.c_int, _D]),
        "lbm_get_fields": (ctypes.c_int, [P, ctypes.c_int, _D, _D, _D, _D]),
        "lbm_get_f": (ctypes.c_int, [P, _D]),
        "lbm_get_force": (ctypes.c_int, [P, _D, _D, _D]),
        "lbm_set_f": (ctypes.c_int, [P, _D]),
        "lbm_conservation": (ctypes.c_int, [P, _D]),
        "lbm_steps_done": (ctypes.c_int64, [P]),
        "lbm_set_stream": (ctypes.c_int, [P, P]),
        "lbm_get_stream": (P, [P]),
        "lbm_synchronize": (ctypes.c_int, [P]),
        "lbm_device_bytes": (ctypes.c_int64, [P]),
        "lbm_kernel_in_use": (ctypes.c_int, [P]),
        "lbm_set_tuning": (ctypes.c_int, [P, ctypes.POINTER(Tuning)]),
        "lbm_get_tuning": (ctypes.c_int, [P, ctypes.POINTER(Tuning)]),
        "lbm_last_error": (ctypes.c_char_p, []),
        "lbm_fill_noise": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_int64, ctypes.c_double,
                                          ctypes.c_double, _D]),
        "lbm_fill_droplet": (ctypes.c_int, [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [_D]),
        "lbm_abi_version": (ctypes.c_int, []),
        "lbm_probe_exp": (ctypes.c_int, [_D, _D, ctypes.c_int64]),
        "lbm_batch_create": (ctypes.c_int, [ctypes.POINTER(BatchDesc), ctypes.POINTER(P)]),
        "lbm_batch_destroy": (None, [P]),
        "lbm_batch_job_bytes": (ctypes.c_int64, [P, ctypes.POINTER(JobDesc)]),
        "lbm_batch_submit": (ctypes.c_int, [P, ctypes.POINTER(JobDesc), _D, ctypes.POINTER(ctypes.c_int64)]),
        "lbm_batch_tick": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int)]),
        "lbm_batch_tick_begin": (ctypes.c_int, [P]),
        "lbm_batch_tick_end": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int)]),
        "lbm_batch_finished": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int)]),
        "lbm_batch_collect": (ctypes.c_int, [P, ctypes.c_int64, _D, _D, _D, _D]),
        "lbm_batch_active": (ctypes.c_int, [P]),
        "lbm_batch_bytes_in_use": (ctypes.c_int64, [P]),
        "lbm_batch_steps_run": (ctypes.c_int64, [P]),
        "lbm_batch_tuning": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "lbm_batch_set_work_limit": (ctypes.c_int, [P, ctypes.c_int64]),
        "lbm_batch_set_max_wait": (ctypes.c_int, [P, ctypes.c_int]),
        "lbm_batch_set_concurrency": (ctypes.c_int, [P, ctypes.c_int]),
        "lbm_batch_set_launch": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "lbm_batch_get_launch": (ctypes.c_int, [P] + [ctypes.POINTER(ctypes.c_int)] * 3),
        "lbm_batch_job_status": (ctypes.c_int, [P, ctypes.c_int64, ctypes.POINTER(JobStatus)]),
        "lbm_batch_cancel": (ctypes.c_int, [P, ctypes.c_int64]),
        "lbm_testing_alloc_count": (ctypes.c_int64, []),
        "lbm_testing_fail_alloc": (None, [ctypes.c_int64]),
        "lbm_probe_copy": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_int, _D]),
        "lbm_components": (ctypes.c_int, [P]),
        "lbm_set_rho_component": (ctypes.c_int, [P, ctypes.c_int, _D]),
        "lbm_get_fields_component": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D]),
        "lbm_get_force_component": (ctypes.c_int, [P, ctypes.c_int, _D, _D, _D]),
        "lbm_get_f_component": (ctypes.c_int, [P, ctypes.c_int, _D]),
        "lbm_set_f_component": (ctypes.c_int, [P, ctypes.c_int, _D]),
        "lbm_conservation_component": (ctypes.c_int, [P, ctypes.c_int, _D]),
        "lbm_probe_memcpy": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, _D]),
        "lbm_capsule_attach": (ctypes.c_int, [P, ctypes.POINTER(CapsuleDesc)]),
        "lbm_capsule_set_nodes": (ctypes.c_int, [P, _D, _D]),
        "lbm_capsule_get_nodes": (ctypes.c_int, [P] + [_D] * 6),
        "lbm_capsule_nodes": (ctypes.c_int, [P]),
        "lbm_batch_read": (ctypes.c_int, [P, ctypes.c_int64] + [_D] * 7),
        "lbm_batch_extend": (ctypes.c_int, [P, ctypes.c_int64, ctypes.c_int64]),
        "lbm_snapshot_begin": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
        "lbm_snapshot_wait": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int64)] + [_D] * 7),
        "lbm_probe_tile_map": (ctypes.c_int, [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int)] * 3),
        "lbm_probe_march": (ctypes.c_int, [ctypes.c_int] * 4 + [_D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise LbmError(rc, _lib.lbm_last_error().decode())


def _dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def noise_rho(shape, seed, mean=0.7, amp=0.02):
    """Seeded noisy density (std::mt19937), reference loop order."""
    lib = load()
    n = int(np.prod(shape))
    out = np.empty(n)
    _check(lib.lbm_fill_noise(seed & 0xFFFFFFFF, n, float(mean), float(amp), _dptr(out)))
    return out.reshape(tuple(shape))


def droplet_rho(shape, R, W=2.0, rl=2.0, rv=0.15):
    lib = load()
    sh = tuple(int(s) for s in shape)
    nz = sh[2] if len(sh) == 3 else 1
    out = np.empty(int(np.prod(sh)))
    _check(lib.lbm_fill_droplet(sh[0], sh[1], nz, float(R), float(W), float(rl), float(rv), _dptr(out)))
    return out.reshape(sh)


def probe_exp(x):
    """The FP64 kernels' device exp() applied to x (test hook)."""
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    _check(lib.lbm_probe_exp(_dptr(x), _dptr(y), x.size))
    return y


def probe_copy(nbytes, reps=5, mode=0):
    """Runs a streaming reference kernel (0-2 copy variants, 3 read-only);
    returns ms per run."""
    lib = load()
    ms = ctypes.c_double()
    _check(lib.lbm_probe_copy(int(nbytes), int(reps), int(mode), ctypes.byref(ms)))
    return ms.value


def alloc_count():
    """Device allocations the library has made (test hook)."""
    return load().lbm_testing_alloc_count()


def fail_alloc(after):
    """Injects one device-allocation failure after `after` successful ones;
    a negative value disarms it (test hook)."""
    load().lbm_testing_fail_alloc(int(after))


def probe_memcpy(nbytes, reps=5):
    """hipMemcpyAsync device-to-device of nbytes; returns ms per copy."""
    lib = load()
    ms = ctypes.c_double()
    _check(lib.lbm_probe_memcpy(int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value


def probe_march(n, elem_bytes=8, layout=0, reps=10):
    """Memory skeleton of the fused D3Q19 march on an n^3 lattice (its loads
    and stores, no arithmetic): returns ms per run.  layout 0: one plane per
    population (this library's layout); 1: the 19 populations of each 64-node
    z segment together."""
    lib = load()
    ms = ctypes.c_double()
    _check(lib.lbm_probe_march(int(n), int(elem_bytes), int(layout), int(reps), ctypes.byref(ms)))
    return ms.value


def probe_tile_map(nty, ntz, ncx, order=2):
    """Block -> (ty, tz, cx) map of the fused kernel (host evaluation)."""
    lib = load()
    nb = nty * ntz * ncx
    out = [np.empty(nb, dtype=np.int32) for _ in range(3)]
    P = ctypes.POINTER(ctypes.c_int)
    rc = lib.lbm_probe_tile_map(nty, ntz, ncx, order, *(a.ctypes.data_as(P) for a in out))
    if rc != 0:
        raise ValueError("bad tile grid")
    return out


class Lattice:
    """One Shan-Chen lattice resident on the GPU (a C-ABI handle).
    components=2 with tau_b: the two-component model (mcmp2D for q=9,
    mcmp3D for q=19); tau is component A's, G couples the components, and
    the comp argument of the accessors picks A (0) or B (1).  With walls as
    well: the film models, rho_sol = the wall densities of A and B."""

    def __init__(self, q, shape, tau, G, order=None, walls=False, precision=FP64,
                 rho_shift=0.7, kernel=KERNEL_AUTO, components=1, tau_b=None, rho_sol=(0.5, 0.5)):
        lib = load()
        shape = tuple(int(s) for s in shape)
        if q == 9:
            nx, ny, nz = shape[0], shape[1], 1
            order = 2 if order is None else order
        else:
            nx, ny, nz = shape
            order = 3 if order is None else order
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        self.q, self.shape, self.order, self.precision = q, shape, order, precision
        self.n = nx * ny * nz
        self.components = int(components)
        d = Desc(q, nx, ny, nz, float(tau), float(G), int(order), int(bool(walls)),
                 int(precision), float(rho_shift), int(kernel), int(components),
                 float(tau if tau_b is None else tau_b), float(rho_sol[0]), float(rho_sol[1]))
        h = ctypes.c_void_p()
        _check(lib.lbm_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def kernel(self):
        return self._lib.lbm_kernel_in_use(self._h)

    @property
    def device_bytes(self):
        return self._lib.lbm_device_bytes(self._h)

    @property
    def steps_done(self):
        return self._lib.lbm_steps_done(self._h)

    @property
    def stream(self):
        return self._lib.lbm_get_stream(self._h)

    def set_tuning(self, shape=-1, chunk=64, order=3, nt_store=1):
        t = Tuning(int(shape), int(chunk), int(order), int(nt_store))
        _check(self._lib.lbm_set_tuning(self._h, ctypes.byref(t)))

    @property
    def tuning(self):
        t = Tuning()
        _check(self._lib.lbm_get_tuning(self._h, ctypes.byref(t)))
        return t.shape, t.chunk, t.order, t.nt_store

    def set_rho(self, rho0, comp=0):
        r = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1)
        assert r.size == self.n
        _check(self._lib.lbm_set_rho_component(self._h, int(comp), _dptr(r)))

    def init(self):
        _check(self._lib.lbm_init_equilibrium(self._h))

    def step(self, nsteps=1):
        _check(self._lib.lbm_step(self._h, int(nsteps)))

    def step_timed(self, nsteps):
        ms = ctypes.c_double()
        _check(self._lib.lbm_step_timed(self._h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        _check(self._lib.lbm_synchronize(self._h))

    def fields(self, kind=FIELDS_REFERENCE, comp=0):
        rho, u, v, w = (np.empty(self.n) for _ in range(4))
        _check(self._lib.lbm_get_fields_component(self._h, int(comp), kind, _dptr(rho), _dptr(u), _dptr(v),
                                                  _dptr(w)))
        sh = self.shape
        return dict(rho=rho.reshape(sh), u=u.reshape(sh), v=v.reshape(sh), w=w.reshape(sh))

    def force(self, comp=0):
        fx, fy, fz = (np.empty(self.n) for _ in range(3))
        _check(self._lib.lbm_get_force_component(self._h, int(comp), _dptr(fx), _dptr(fy), _dptr(fz)))
        sh = self.shape
        return fx.reshape(sh), fy.reshape(sh), fz.reshape(sh)

    def attach_capsule(self, nnodes, kstiff=1.0, dt=1.0, body=(0.000008, 0.0)):
        """Turns the lattice into the capsule2D model: nnodes IB nodes with
        membrane stiffness kstiff and a body force (call before init)."""
        d = CapsuleDesc(int(nnodes), float(kstiff), float(dt), float(body[0]), float(body[1]))
        _check(self._lib.lbm_capsule_attach(self._h, ctypes.byref(d)))

    def set_capsule_nodes(self, x, y):
        """Node positions, also their rest positions."""
        nn = self._lib.lbm_capsule_nodes(self._h)
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
        assert x.size == nn and y.size == nn
        _check(self._lib.lbm_capsule_set_nodes(self._h, _dptr(x), _dptr(y)))

    def capsule_nodes(self):
        """dict of node positions x, y, velocities vx, vy and forces fx, fy."""
        nn = self._lib.lbm_capsule_nodes(self._h)
        arrs = [np.empty(nn) for _ in range(6)]
        _check(self._lib.lbm_capsule_get_nodes(self._h, *(_dptr(a) for a in arrs)))
        return dict(zip(("x", "y", "vx", "vy", "fx", "fy"), arrs))

    def snapshot_begin(self, comp=0, kind=FIELDS_REFERENCE):
        """Starts an asynchronous copy of fields(kind, comp) and force(comp);
        steps may be enqueued while it runs."""
        _check(self._lib.lbm_snapshot_begin(self._h, int(comp), kind))

    def snapshot_wait(self):
        """Waits for the snapshot begun last; returns (step, fields dict with
        rho, u, v, w, fx, fy, fz)."""
        arrs = [np.empty(self.n) for _ in range(7)]
        step = ctypes.c_int64()
        _check(self._lib.lbm_snapshot_wait(self._h, ctypes.byref(step), *(_dptr(a) for a in arrs)))
        sh = self.shape
        names = ("rho", "u", "v", "w", "fx", "fy", "fz")
        return step.value, {k: a.reshape(sh) for k, a in zip(names, arrs)}

    def get_f(self, comp=0):
        f = np.empty(self.n * self.q)
        _check(self._lib.lbm_get_f_component(self._h, int(comp), _dptr(f)))
        return f.reshape(self.shape + (self.q,))

    def set_f(self, f, comp=0):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        assert a.size == self.n * self.q
        _check(self._lib.lbm_set_f_component(self._h, int(comp), _dptr(a)))

    def conservation(self, comp=0):
        out = np.empty(4)
        _check(self._lib.lbm_conservation_component(self._h, int(comp), _dptr(out)))
        return out


class Batch:
    """Many independent lattices advanced together (the batched / serving
    form of the C ABI).  Jobs are admitted against an HBM budget with
    submit(), advanced by tick(), and read back and freed by collect()."""

    def __init__(self, precision=FP16S, rho_shift=0.7, hbm_budget=0, steps_per_tick=20,
                 use_graph=True, autotune=False, work_limit=0, max_wait=0, concurrent=True):
        lib = load()
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        d = BatchDesc(int(precision), float(rho_shift), int(hbm_budget), int(steps_per_tick),
                      int(bool(use_graph)), int(bool(autotune)))
        h = ctypes.c_void_p()
        _check(lib.lbm_batch_create(ctypes.byref(d), ctypes.byref(h)))
        self._h, self._lib = h, lib
        self.shapes = {}
        if work_limit:
            _check(lib.lbm_batch_set_work_limit(h, int(work_limit)))
        if max_wait:
            _check(lib.lbm_batch_set_max_wait(h, int(max_wait)))
        if not concurrent:
            _check(lib.lbm_batch_set_concurrency(h, 0))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lbm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @staticmethod
    def _desc(q, shape, tau, G, steps):
        shape = tuple(int(s) for s in shape)
        nz = shape[2] if len(shape) == 3 else 1
        return JobDesc(int(q), shape[0], shape[1], nz, float(tau), float(G), int(steps))

    def job_bytes(self, q, shape, tau=1.0, G=-5.0, steps=1):
        d = self._desc(q, shape, tau, G, steps)
        return self._lib.lbm_batch_job_bytes(self._h, ctypes.byref(d))

    def submit(self, q, shape, tau, G, steps, rho0):
        """Returns the job id, or None when the job does not fit the budget."""
        d = self._desc(q, shape, tau, G, steps)
        r = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1)
        jid = ctypes.c_int64()
        rc = self._lib.lbm_batch_submit(self._h, ctypes.byref(d), _dptr(r), ctypes.byref(jid))
        if rc == ENOMEM:
            return None
        _check(rc)
        self.shapes[jid.value] = tuple(int(s) for s in shape)
        return jid.value

    def status(self, job_id):
        """dict(state, steps_done, steps_finite, mass) of an uncollected job."""
        st = JobStatus()
        _check(self._lib.lbm_batch_job_status(self._h, int(job_id), ctypes.byref(st)))
        return dict(state=st.state, steps_done=st.steps_done, steps_finite=st.steps_finite,
                    mass=st.mass)

    def set_work_limit(self, max_nodes):
        """Nodes advanced per tick at most (0: every job); the jobs with the
        least remaining work go first (lbm_batch_set_work_limit)."""
        _check(self._lib.lbm_batch_set_work_limit(self._h, int(max_nodes)))

    def cancel(self, job_id):
        """Drops an admitted job (running or finished) and frees its memory."""
        _check(self._lib.lbm_batch_cancel(self._h, int(job_id)))
        self.shapes.pop(job_id, None)

    def read(self, job_id):
        """The reference's r,u,v,w and fx,fy,fz arrays of an admitted job at
        the steps it has run (the job stays admitted)."""
        shape = self.shapes[job_id]
        n = int(np.prod(shape))
        arrs = [np.empty(n) for _ in range(7)]
        _check(self._lib.lbm_batch_read(self._h, int(job_id), *(_dptr(a) for a in arrs)))
        names = ("rho", "u", "v", "w", "fx", "fy", "fz")
        return {k: a.reshape(shape) for k, a in zip(names, arrs)}

    def extend(self, job_id, nsteps):
        """Gives a finished job nsteps more steps."""
        _check(self._lib.lbm_batch_extend(self._h, int(job_id), int(nsteps)))

    def tick(self):
        n = ctypes.c_int()
        _check(self._lib.lbm_batch_tick(self._h, ctypes.byref(n)))
        return n.value

    def tick_begin(self):
        """Launches a tick and returns (lbm_batch_tick_begin); tick_end()
        completes it.  Ticks of several batches then run on the device
        together."""
        _check(self._lib.lbm_batch_tick_begin(self._h))

    def tick_end(self):
        n = ctypes.c_int()
        _check(self._lib.lbm_batch_tick_end(self._h, ctypes.byref(n)))
        return n.value

    def finished(self, max_ids=1 << 16):
        ids = (ctypes.c_int64 * max_ids)()
        n = ctypes.c_int()
        _check(self._lib.lbm_batch_finished(self._h, ids, max_ids, ctypes.byref(n)))
        return [ids[i] for i in range(n.value)]

    def collect(self, job_id):
        """Fields of a finished job (frees it).  A job the engine retired as
        failed comes back with failed=True, its step counts and its
        (non-finite) fields; every other error raises."""
        shape = self.shapes[job_id]
        st = self.status(job_id)
        n = int(np.prod(shape))
        rho, u, v, w = (np.empty(n) for _ in range(4))
        rc = self._lib.lbm_batch_collect(self._h, int(job_id), _dptr(rho), _dptr(u), _dptr(v), _dptr(w))
        if rc not in (0, EJOB):
            _check(rc)
        self.shapes.pop(job_id)
        out = dict(rho=rho.reshape(shape), u=u.reshape(shape), v=v.reshape(shape), w=w.reshape(shape),
                   failed=rc == EJOB, steps_done=st["steps_done"],
                   steps_finite=st["steps_finite"], mass=st["mass"])
        if rc == EJOB:
            out["error"] = self._lib.lbm_last_error().decode()
        return out

    @property
    def active(self):
        return self._lib.lbm_batch_active(self._h)

    @property
    def bytes_in_use(self):
        return self._lib.lbm_batch_bytes_in_use(self._h)

    @property
    def steps_run(self):
        return self._lib.lbm_batch_steps_run(self._h)

    @property
    def tuning(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(self._lib.lbm_batch_get_launch(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return dict(d2q9_chunk=a.value, d2q9_rows_kernel=b.value, d3q19_march=c.value)

    def set_concurrency(self, on):
        """Kernel classes on streams of their own within a tick (True, the
        default) or one after another on the batch's stream."""
        _check(self._lib.lbm_batch_set_concurrency(self._h, int(bool(on))))

    def set_launch(self, d2q9_chunk=32, d2q9_rows_kernel=1, d3q19_march=1):
        """Fixes the launch configuration (stops the tuner).  d3q19_march: 0
        two-pass, 1 the fused march for node-shift jobs that fill 90% of its
        tiles (default), 2 the
        march wherever it fits; D3Q19 results of the two kernels agree to
        FP32 rounding."""
        _check(self._lib.lbm_batch_set_launch(self._h, int(d2q9_chunk), int(d2q9_rows_kernel),
                                              int(d3q19_march)))
