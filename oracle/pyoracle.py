"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline leg
of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FLUX = {"ROE": 0, "HLLC": 1, "HLL": 2, "LLF": 3, "AUSM": 4, "AUSMPLUS": 5, "VANLEER": 6}
GRAD = {"NONE": 0, "GREENGAUSS": 1, "LEASTSQUARES": 2}
RECON = {"NONE": 0, "VANALBADA": 1, "VENKATAKRISHNAN": 2, "BARTHJESPERSEN": 3, "WENO": 4}
BC = {"slipwall": 0, "farfield": 1, "inflowoutflow": 2, "subsonicinflow": 3,
      "extrapolation": 4, "periodic": 5, "isothermalwall": 6, "adiabaticwall": 7}
PC = {"none": 0, "jacobi": 1, "sgs": 2, "ilu0": 3}


class BCConf(C.Structure):
    _fields_ = [("tag", C.c_int), ("kind", C.c_int), ("vals", C.c_double * 2)]


class Config(C.Structure):
    _fields_ = [("gamma", C.c_double), ("Minf", C.c_double), ("Tinf", C.c_double),
                ("Reinf", C.c_double), ("Pr", C.c_double), ("aoa", C.c_double),
                ("viscous", C.c_int), ("constvisc", C.c_int),
                ("flux", C.c_int), ("jacflux", C.c_int), ("gradient", C.c_int),
                ("recon", C.c_int), ("limiter_param", C.c_double), ("order2", C.c_int),
                ("nbc", C.c_int), ("bcs", BCConf * 8)]


class SolverConfig(C.Structure):
    _fields_ = [("cflinit", C.c_double), ("cflfin", C.c_double), ("tol", C.c_double),
                ("maxiter", C.c_int), ("implicit", C.c_int), ("fullupdate", C.c_int),
                ("min_relax", C.c_double), ("pc", C.c_int), ("krylov_restart", C.c_int),
                ("krylov_rtol", C.c_double), ("matrix_free", C.c_int), ("mf_eps", C.c_double),
                ("fixed_norm", C.c_double), ("pc_sweeps", C.c_int)]


_PATH = None


def use_library(path):
    """Loads the oracle from `path` (the CPU-baseline build tuned for the timing machine)
    instead of oracle/liboracle.so; must be called before the first use."""
    global _PATH
    if _LIB is not None and path != _PATH:
        raise RuntimeError("oracle library already loaded")
    _PATH = path


def set_threads(n):
    """OpenMP thread count of the oracle's parallel loops started from this thread."""
    lib()
    C.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def lib():
    global _LIB
    if _LIB is None:
        path = _PATH or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.fvo_last_error.restype = C.c_char_p
        L.fvo_mesh_free.restype = None
        L.fvo_disc_free.restype = None
        _LIB = L
    return _LIB


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().fvo_last_error().decode())


def make_config(gamma=1.4, Minf=0.5, Tinf=298.0, Reinf=float("inf"), Pr=float("nan"), aoa_deg=0.0,
                viscous=False, constvisc=False, flux="ROE", jacflux=None, gradient="LEASTSQUARES",
                recon="NONE", limiter_param=5.0, order2=True, bcs=()):
    c = Config()
    c.gamma, c.Minf, c.Tinf, c.Reinf, c.Pr = gamma, Minf, Tinf, Reinf, Pr
    c.aoa = np.deg2rad(aoa_deg)
    c.viscous, c.constvisc = int(viscous), int(constvisc)
    c.flux = FLUX[flux]
    c.jacflux = FLUX[jacflux or flux]
    c.gradient = GRAD[gradient]
    c.recon = RECON[recon]
    c.limiter_param = limiter_param
    c.order2 = int(order2 and gradient != "NONE")
    c.nbc = len(bcs)
    for i, b in enumerate(bcs):
        tag, kind = b[0], b[1]
        vals = b[2] if len(b) > 2 else (0.0, 0.0)
        c.bcs[i].tag = tag
        c.bcs[i].kind = BC[kind]
        c.bcs[i].vals[0] = vals[0]
        c.bcs[i].vals[1] = vals[1] if len(vals) > 1 else 0.0
    return c


class Mesh:
    def __init__(self, raw):
        L = lib()
        self.raw = raw
        coords = np.ascontiguousarray(raw.coords, np.float64)
        inpoel = np.ascontiguousarray(raw.inpoel, np.int32)
        nnode = np.ascontiguousarray(raw.nnode, np.int32)
        bface = np.ascontiguousarray(raw.bface, np.int32)
        btag = np.ascontiguousarray(raw.btag, np.int32)
        h = C.c_void_p()
        _check(L.fvo_mesh_new(raw.npoin, raw.nelem, raw.nbface, _p(coords, C.c_double),
                              _p(inpoel, C.c_int), _p(nnode, C.c_int), _p(bface, C.c_int),
                              _p(btag, C.c_int), C.byref(h)))
        self.h = h
        s = np.zeros(5, np.int32)
        L.fvo_mesh_sizes(h, _p(s, C.c_int))
        self.npoin, self.nelem, self.nbface, self.ninface, self.naface = map(int, s)
        ne, nf, nb = self.nelem, self.naface, self.nbface
        self.intfac = np.zeros((nf, 4), np.int32)
        self.facemetric = np.zeros((nf, 3))
        self.area = np.zeros(ne)
        self.rc = np.zeros((ne, 2))
        self.rcbp = np.zeros((nb, 2))
        self.gr = np.zeros((nf, 2))
        self.esuel = np.zeros((ne, 4), np.int32)
        self.elemface = np.zeros((ne, 4), np.int32)
        self.bface = np.zeros((nb, 2), np.int32)
        L.fvo_mesh_arrays(h, _p(self.intfac, C.c_int), _p(self.facemetric, C.c_double),
                          _p(self.area, C.c_double), _p(self.rc, C.c_double),
                          _p(self.rcbp, C.c_double), _p(self.gr, C.c_double),
                          _p(self.esuel, C.c_int), _p(self.elemface, C.c_int),
                          _p(self.bface, C.c_int))

    def level_schedule(self):
        out = np.zeros(self.nelem + 2, np.int32)
        cnt = C.c_int()
        _check(lib().fvo_level_schedule(self.h, _p(out, C.c_int), len(out), C.byref(cnt)))
        return out[:cnt.value].copy()

    def __del__(self):
        if getattr(self, "h", None) is not None and _LIB is not None:
            _LIB.fvo_mesh_free(self.h)
            self.h = None


class Disc:
    def __init__(self, mesh: Mesh, cfg: Config):
        self.mesh = mesh
        self.cfg = cfg
        h = C.c_void_p()
        _check(lib().fvo_disc_new(mesh.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and _LIB is not None:
            _LIB.fvo_disc_free(self.h)
            self.h = None

    def freestream(self):
        out = np.zeros(4)
        lib().fvo_freestream(C.byref(self.cfg), _p(out, C.c_double))
        return out

    def residual(self, u, gettimesteps=True, parallel=0, r0=None):
        u = np.ascontiguousarray(u, np.float64)
        r = np.zeros_like(u) if r0 is None else np.array(r0, np.float64)
        dt = np.zeros(self.mesh.nelem)
        _check(lib().fvo_residual(self.h, _p(u, C.c_double), _p(r, C.c_double),
                                  int(gettimesteps), _p(dt, C.c_double), int(parallel)))
        return r, dt

    def flux_magnitude(self, u):
        """Per cell and variable, sum over faces of |F_f| len_f (parity-metric scale)."""
        u = np.ascontiguousarray(u, np.float64)
        s = np.zeros_like(u)
        _check(lib().fvo_flux_magnitude(self.h, _p(u, C.c_double), _p(s, C.c_double)))
        return s

    def scalar_face_values(self, u, ug):
        """Gradients of a scalar field and unlimited linear face values (left, right)."""
        u = np.ascontiguousarray(u, np.float64)
        ug = np.ascontiguousarray(ug, np.float64)
        g = np.zeros((self.mesh.nelem, 2))
        ufl = np.zeros(self.mesh.naface)
        ufr = np.zeros(self.mesh.naface)
        _check(lib().fvo_scalar_face_values(self.h, _p(u, C.c_double), _p(ug, C.c_double),
                                            _p(g, C.c_double), _p(ufl, C.c_double),
                                            _p(ufr, C.c_double)))
        return g, ufl, ufr

    def face_states(self, u):
        """Primitive face states (left, right) of the second-order reconstruction."""
        u = np.ascontiguousarray(u, np.float64)
        ufl = np.zeros((self.mesh.naface, 4))
        ufr = np.zeros((self.mesh.naface, 4))
        _check(lib().fvo_face_states(self.h, _p(u, C.c_double), _p(ufl, C.c_double),
                                     _p(ufr, C.c_double)))
        return ufl, ufr

    def prim_gradients(self, u, parallel=0):
        u = np.ascontiguousarray(u, np.float64)
        g = np.zeros((self.mesh.nelem, 4, 2))
        _check(lib().fvo_prim_gradients(self.h, _p(u, C.c_double), _p(g, C.c_double),
                                        int(parallel)))
        return g

    def cons_gradients(self, u):
        u = np.ascontiguousarray(u, np.float64)
        g = np.zeros((self.mesh.nelem, 4, 2))
        _check(lib().fvo_cons_gradients(self.h, _p(u, C.c_double), _p(g, C.c_double)))
        return g

    def ghost(self, ib, uin):
        uin = np.ascontiguousarray(uin, np.float64)
        ug = np.zeros(4)
        _check(lib().fvo_ghost(self.h, int(ib), _p(uin, C.c_double), _p(ug, C.c_double)))
        return ug

    def ghost_jacobian(self, ib, uin):
        uin = np.ascontiguousarray(uin, np.float64)
        ug = np.zeros(4)
        dug = np.zeros((4, 4))
        _check(lib().fvo_ghost_jacobian(self.h, int(ib), _p(uin, C.c_double),
                                        _p(ug, C.c_double), _p(dug, C.c_double)))
        return ug, dug

    def local_jacobians(self, iface, ul, ur):
        ul = np.ascontiguousarray(ul, np.float64)
        ur = np.ascontiguousarray(ur, np.float64)
        L = np.zeros((4, 4))
        U = np.zeros((4, 4))
        _check(lib().fvo_local_jacobians(self.h, int(iface), _p(ul, C.c_double),
                                         _p(ur, C.c_double), _p(L, C.c_double),
                                         _p(U, C.c_double)))
        return L, U

    def jacobian(self, u):
        u = np.ascontiguousarray(u, np.float64)
        nnzb = C.c_int()
        _check(lib().fvo_jacobian_nnzb(self.h, C.byref(nnzb)))
        rowptr = np.zeros(self.mesh.nelem + 1, np.int32)
        colind = np.zeros(nnzb.value, np.int32)
        vals = np.zeros((nnzb.value, 4, 4))
        _check(lib().fvo_assemble_jacobian(self.h, _p(u, C.c_double), _p(rowptr, C.c_int),
                                           _p(colind, C.c_int), _p(vals, C.c_double)))
        return rowptr, colind, vals

    def functionals(self, u, walltag=2):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros(4)
        _check(lib().fvo_functionals(self.h, _p(u, C.c_double), int(walltag),
                                     _p(out, C.c_double)))
        return dict(CL=out[0], CDp=out[1], CDsf=out[2], entropy=out[3])

    def steady_solve(self, u, **kw):
        sc = SolverConfig()
        defaults = dict(cflinit=50.0, cflfin=500.0, tol=1e-10, maxiter=500, implicit=1,
                        fullupdate=1, min_relax=0.2, pc="ilu0", krylov_restart=30,
                        krylov_rtol=1e-1, matrix_free=0, mf_eps=1e-7, fixed_norm=0.0,
                        pc_sweeps=1)
        defaults.update(kw)
        for k, v in defaults.items():
            setattr(sc, k, PC[v] if k == "pc" else v)
        u = np.array(u, np.float64)
        st = np.zeros(7)
        _check(lib().fvo_steady_solve(self.h, C.byref(sc), _p(u, C.c_double),
                                      _p(st, C.c_double)))
        return u, dict(steps=int(st[0]), lin_iters=int(st[1]), converged=bool(st[2]),
                       status=int(st[3]), rel=st[4], abs=st[5], walltime=st[6])


def flux(cfg, name, ul, ur, n):
    ul, ur, n = (np.ascontiguousarray(a, np.float64) for a in (ul, ur, n))
    f = np.zeros(4)
    _check(lib().fvo_flux(C.byref(cfg), FLUX[name], _p(ul, C.c_double), _p(ur, C.c_double),
                          _p(n, C.c_double), _p(f, C.c_double)))
    return f


def flux_jacobian(cfg, name, ul, ur, n):
    ul, ur, n = (np.ascontiguousarray(a, np.float64) for a in (ul, ur, n))
    dl = np.zeros((4, 4))
    dr = np.zeros((4, 4))
    _check(lib().fvo_flux_jacobian(C.byref(cfg), FLUX[name], _p(ul, C.c_double),
                                   _p(ur, C.c_double), _p(n, C.c_double), _p(dl, C.c_double),
                                   _p(dr, C.c_double)))
    return dl, dr


def pc_apply_bsr(rowptr, colind, vals, pc, r, sweeps=1):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colind = np.ascontiguousarray(colind, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    z = np.zeros_like(r)
    _check(lib().fvo_pc_apply_bsr(len(rowptr) - 1, _p(rowptr, C.c_int), _p(colind, C.c_int),
                                  _p(vals, C.c_double), PC[pc], int(sweeps), _p(r, C.c_double),
                                  _p(z, C.c_double)))
    return z
