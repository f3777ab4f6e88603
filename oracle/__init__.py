"""ctypes binding of the CPU oracle (oracle/scmp_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, the smoke check and the CPU
baseline leg of bench.py, never by the product package.  The oracle restates
the reference SCMP step (flowzario/mesoBasic, src/lattice_boltzmann) in the
reference's own AoS layout and floating-point order; see scmp_oracle.h for
the file:line map.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libscmp_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)


def build() -> None:
    """Compiles libscmp_oracle.so (gcc, no FMA contraction)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.scmp_oracle_create.restype = ctypes.c_void_p
    lib.scmp_oracle_create.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 2 + [ctypes.c_int] * 3
    lib.scmp_oracle_destroy.argtypes = [ctypes.c_void_p]
    lib.scmp_oracle_set_rho.argtypes = [ctypes.c_void_p, _D]
    lib.scmp_oracle_init.argtypes = [ctypes.c_void_p]
    lib.scmp_oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.scmp_oracle_get.argtypes = [ctypes.c_void_p] + [_D] * 5
    lib.scmp_oracle_get_force.argtypes = [ctypes.c_void_p] + [_D] * 3
    lib.scmp_oracle_moments.argtypes = [ctypes.c_void_p] + [_D] * 4
    lib.scmp_oracle_set_f.argtypes = [ctypes.c_void_p, _D]
    lib.scmp_oracle_set_exp.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.scmp_oracle_nodes.restype = ctypes.c_int64
    lib.scmp_oracle_nodes.argtypes = [ctypes.c_void_p]
    lib.scmp_oracle_noise_rho.argtypes = [ctypes.c_uint32, ctypes.c_int64, ctypes.c_double,
                                          ctypes.c_double, _D]
    lib.scmp_oracle_mt19937.argtypes = [ctypes.c_uint32, ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_uint32)]
    lib.scmp_oracle_droplet_rho.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [_D]
    lib.mcmp_oracle_create.restype = ctypes.c_void_p
    lib.mcmp_oracle_create.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 3 + [ctypes.c_int] * 2
    lib.mcmp_oracle_destroy.argtypes = [ctypes.c_void_p]
    lib.mcmp_oracle_set_film.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    lib.mcmp_oracle_set_rho.argtypes = [ctypes.c_void_p, _D, _D]
    lib.mcmp_oracle_init.argtypes = [ctypes.c_void_p]
    lib.mcmp_oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.mcmp_oracle_get.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_D] * 5
    lib.mcmp_oracle_get_force.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_D] * 3
    lib.mcmp_oracle_moments.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_D] * 4
    lib.capsule_oracle_create.restype = ctypes.c_void_p
    lib.capsule_oracle_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int] + \
        [ctypes.c_double] * 4
    lib.capsule_oracle_destroy.argtypes = [ctypes.c_void_p]
    lib.capsule_oracle_set_rho.argtypes = [ctypes.c_void_p, _D]
    lib.capsule_oracle_set_nodes.argtypes = [ctypes.c_void_p, _D, _D]
    lib.capsule_oracle_move_nodes.argtypes = [ctypes.c_void_p, _D, _D]
    lib.capsule_oracle_circle.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 3
    lib.capsule_oracle_init.argtypes = [ctypes.c_void_p]
    lib.capsule_oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.capsule_oracle_get.argtypes = [ctypes.c_void_p] + [_D] * 6
    lib.capsule_oracle_get_nodes.argtypes = [ctypes.c_void_p] + [_D] * 6
    lib.capsule_oracle_get_ghost_force.argtypes = [ctypes.c_void_p] + [_D] * 2
    _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


class Oracle:
    """One SCMP lattice on the CPU, reference semantics.

    q: 9 or 19; shape (nx, ny) for q=9 or (nx, ny, nz) for q=19.
    order: 3 = scmp3D step order, 2 = LBfluid2D step order.
    walls: bounce-back walls on the last axis.
    """

    def __init__(self, q, shape, tau, G, order=None, walls=False, nthreads=1):
        lib = _load()
        shape = tuple(int(s) for s in shape)
        if q == 9:
            assert len(shape) == 2
            nx, ny, nz = shape[0], shape[1], 1
            order = 2 if order is None else order
        else:
            assert len(shape) == 3
            nx, ny, nz = shape
            order = 3 if order is None else order
        self.q, self.shape, self.order = q, shape, order
        self.n = nx * ny * nz
        self._h = lib.scmp_oracle_create(q, nx, ny, nz, float(tau), float(G),
                                         int(order), int(bool(walls)), int(nthreads))
        if not self._h:
            raise ValueError("oracle: bad lattice description")
        self._lib = lib

    def close(self):
        if self._h:
            self._lib.scmp_oracle_destroy(self._h)
            self._h = None

    __del__ = close

    def set_rho(self, rho0: np.ndarray):
        r = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1)
        assert r.size == self.n
        self._lib.scmp_oracle_set_rho(self._h, _ptr(r))

    def init(self):
        self._lib.scmp_oracle_init(self._h)

    def step(self, nsteps: int = 1):
        self._lib.scmp_oracle_step(self._h, int(nsteps))

    def fields(self):
        """Reference r,u,v,w arrays and f ([node][q]), shaped to the lattice."""
        rho, u, v, w = (np.empty(self.n) for _ in range(4))
        f = np.empty(self.n * self.q)
        self._lib.scmp_oracle_get(self._h, _ptr(rho), _ptr(u), _ptr(v), _ptr(w), _ptr(f))
        sh = self.shape
        return dict(rho=rho.reshape(sh), u=u.reshape(sh), v=v.reshape(sh),
                    w=w.reshape(sh), f=f.reshape(sh + (self.q,)))

    def force(self):
        fx, fy, fz = (np.empty(self.n) for _ in range(3))
        self._lib.scmp_oracle_get_force(self._h, _ptr(fx), _ptr(fy), _ptr(fz))
        return fx.reshape(self.shape), fy.reshape(self.shape), fz.reshape(self.shape)

    def moments(self):
        rho, u, v, w = (np.empty(self.n) for _ in range(4))
        self._lib.scmp_oracle_moments(self._h, _ptr(rho), _ptr(u), _ptr(v), _ptr(w))
        sh = self.shape
        return dict(rho=rho.reshape(sh), u=u.reshape(sh), v=v.reshape(sh), w=w.reshape(sh))

    def set_f(self, f: np.ndarray):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        assert a.size == self.n * self.q
        self._lib.scmp_oracle_set_f(self._h, _ptr(a))

    def set_exp(self, fnptr):
        """fnptr: ctypes function pointer double(double), or None for libm."""
        self._lib.scmp_oracle_set_exp(self._h, fnptr)


class MCOracle:
    """Two-component Shan-Chen (mcmp2D / mcmp3D) on the CPU, reference
    semantics: components A (comp 0) and B (comp 1) with taus tauA, tauB and
    cross coupling G.  order: 2 = mcmp2D's step (default for q=9), 3 =
    mcmp3D's (default for q=19)."""

    def __init__(self, q, shape, tauA, tauB, G, order=None, nthreads=1, film=None):
        """film: (rhoA_sol, rhoB_sol) for the film models (walls on the last
        axis), None for mcmp2D / mcmp3D."""
        lib = _load()
        shape = tuple(int(s) for s in shape)
        if q == 9:
            nx, ny, nz = shape[0], shape[1], 1
            order = 2 if order is None else order
        else:
            nx, ny, nz = shape
            order = 3 if order is None else order
        self.q, self.shape, self.order = q, shape, order
        self.n = nx * ny * nz
        self._h = lib.mcmp_oracle_create(q, nx, ny, nz, float(tauA), float(tauB), float(G),
                                         int(order), int(nthreads))
        if not self._h:
            raise ValueError("oracle: bad lattice description")
        self._lib = lib
        if film is not None:
            lib.mcmp_oracle_set_film(self._h, float(film[0]), float(film[1]))

    def close(self):
        if self._h:
            self._lib.mcmp_oracle_destroy(self._h)
            self._h = None

    __del__ = close

    def set_rho(self, rhoA: np.ndarray, rhoB: np.ndarray):
        a = np.ascontiguousarray(rhoA, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(rhoB, dtype=np.float64).reshape(-1)
        assert a.size == self.n and b.size == self.n
        self._lib.mcmp_oracle_set_rho(self._h, _ptr(a), _ptr(b))

    def init(self):
        self._lib.mcmp_oracle_init(self._h)

    def step(self, nsteps: int = 1):
        self._lib.mcmp_oracle_step(self._h, int(nsteps))

    def fields(self, comp):
        """Reference r,u,v,w arrays, force arrays and f of one component."""
        rho, u, v, w, fx, fy, fz = (np.empty(self.n) for _ in range(7))
        f = np.empty(self.n * self.q)
        self._lib.mcmp_oracle_get(self._h, int(comp), _ptr(rho), _ptr(u), _ptr(v), _ptr(w), _ptr(f))
        self._lib.mcmp_oracle_get_force(self._h, int(comp), _ptr(fx), _ptr(fy), _ptr(fz))
        sh = self.shape
        return dict(rho=rho.reshape(sh), u=u.reshape(sh), v=v.reshape(sh), w=w.reshape(sh),
                    fx=fx.reshape(sh), fy=fy.reshape(sh), fz=fz.reshape(sh),
                    f=f.reshape(sh + (self.q,)))

    def moments(self, comp):
        rho, u, v, w = (np.empty(self.n) for _ in range(4))
        self._lib.mcmp_oracle_moments(self._h, int(comp), _ptr(rho), _ptr(u), _ptr(v), _ptr(w))
        sh = self.shape
        return dict(rho=rho.reshape(sh), u=u.reshape(sh), v=v.reshape(sh), w=w.reshape(sh))


class CapsuleOracle:
    """Immersed-boundary capsule (capsule2D) on the CPU, reference semantics:
    a D2Q9 fluid of relaxation time tau in the LBfluid2D step order with
    bounce-back walls in y and no Shan-Chen force, and nnodes IB nodes tied
    to their rest positions with stiffness kstiff.  body: the body force on
    the rows between the walls (the reference's is (0.000008, 0))."""

    def __init__(self, shape, tau, nnodes, kstiff=1.0, dt=1.0, body=(0.000008, 0.0)):
        lib = _load()
        self.shape = (int(shape[0]), int(shape[1]))
        self.n = self.shape[0] * self.shape[1]
        self.nnodes = int(nnodes)
        self._h = lib.capsule_oracle_create(self.shape[0], self.shape[1], float(tau), self.nnodes,
                                            float(kstiff), float(dt), float(body[0]), float(body[1]))
        if not self._h:
            raise ValueError("oracle: bad capsule description")
        self._lib = lib

    def close(self):
        if self._h:
            self._lib.capsule_oracle_destroy(self._h)
            self._h = None

    __del__ = close

    def set_rho(self, rho0: np.ndarray):
        a = np.ascontiguousarray(rho0, dtype=np.float64).reshape(-1)
        assert a.size == self.n
        self._lib.capsule_oracle_set_rho(self._h, _ptr(a))

    def set_nodes(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        assert x.size == self.nnodes and y.size == self.nnodes
        self._lib.capsule_oracle_set_nodes(self._h, _ptr(x), _ptr(y))

    def move_nodes(self, x, y):
        """Node positions only; the rest positions stay."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        assert x.size == self.nnodes and y.size == self.nnodes
        self._lib.capsule_oracle_move_nodes(self._h, _ptr(x), _ptr(y))

    def circle(self, xo, yo, rad):
        self._lib.capsule_oracle_circle(self._h, float(xo), float(yo), float(rad))

    def init(self):
        self._lib.capsule_oracle_init(self._h)

    def step(self, nsteps: int = 1):
        self._lib.capsule_oracle_step(self._h, int(nsteps))

    def fields(self):
        """The fluid's r, u, v, force arrays and f."""
        rho, u, v, fx, fy = (np.empty(self.n) for _ in range(5))
        f = np.empty(self.n * 9)
        self._lib.capsule_oracle_get(self._h, _ptr(rho), _ptr(u), _ptr(v), _ptr(f), _ptr(fx), _ptr(fy))
        sh = self.shape
        return dict(rho=rho.reshape(sh), u=u.reshape(sh), v=v.reshape(sh), fx=fx.reshape(sh),
                    fy=fy.reshape(sh), f=f.reshape(sh + (9,)))

    def nodes(self):
        """IB node positions x, y, velocities vx, vy and membrane forces fx, fy."""
        arrs = [np.empty(self.nnodes) for _ in range(6)]
        self._lib.capsule_oracle_get_nodes(self._h, *(_ptr(a) for a in arrs))
        return dict(zip(("x", "y", "vx", "vy", "fx", "fy"), arrs))

    def ghost_force(self):
        """Force accumulated on the ghost rows j = 0 and j = ny+1: (gfx, gfy),
        each of shape (2, nx)."""
        gfx, gfy = np.empty(2 * self.shape[0]), np.empty(2 * self.shape[0])
        self._lib.capsule_oracle_get_ghost_force(self._h, _ptr(gfx), _ptr(gfy))
        return gfx.reshape(2, -1), gfy.reshape(2, -1)


def mt19937(seed: int, n: int) -> np.ndarray:
    """First n outputs of std::mt19937(seed) (uint32)."""
    out = np.empty(int(n), dtype=np.uint32)
    _load().scmp_oracle_mt19937(seed & 0xFFFFFFFF, out.size,
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def noise_rho(shape, seed: int, mean: float = 0.7, amp: float = 0.02) -> np.ndarray:
    """Seeded noisy density, reference loop order (x slowest)."""
    n = int(np.prod(shape))
    out = np.empty(n)
    _load().scmp_oracle_noise_rho(seed & 0xFFFFFFFF, n, float(mean), float(amp), _ptr(out))
    return out.reshape(tuple(shape))


def droplet_rho(shape, R: float, W: float = 2.0, rl: float = 2.0, rv: float = 0.15) -> np.ndarray:
    """tanh droplet centred at (N-1)/2 per axis."""
    sh = tuple(int(s) for s in shape)
    nx, ny = sh[0], sh[1]
    nz = sh[2] if len(sh) == 3 else 1
    out = np.empty(nx * ny * nz)
    _load().scmp_oracle_droplet_rho(nx, ny, nz, float(R), float(W), float(rl), float(rv), _ptr(out))
    return out.reshape(sh)
