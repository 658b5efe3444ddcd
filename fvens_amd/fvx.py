"""ctypes binding of libfvx (include/fvx.h) for Python callers, tests and the benchmark.

Device buffers are torch tensors on the current HIP device; only their data pointers and the
current stream cross the C boundary. The library is required: importing this module on a
machine without a built libfvx.so raises immediately instead of falling back to anything.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .meshgen import RawMesh

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfvx.so")

FLUX = {"ROE": 0, "HLLC": 1, "HLL": 2, "LLF": 3, "AUSM": 4, "AUSMPLUS": 5, "VANLEER": 6}
GRAD = {"NONE": 0, "GREENGAUSS": 1, "LEASTSQUARES": 2}
RECON = {"NONE": 0, "VANALBADA": 1, "VENKATAKRISHNAN": 2, "BARTHJESPERSEN": 3, "WENO": 4}
BC = {"slipwall": 0, "farfield": 1, "inflowoutflow": 2, "subsonicinflow": 3,
      "extrapolation": 4, "periodic": 5, "isothermalwall": 6, "adiabaticwall": 7}

STATUS = {0: "OK", 1: "invalid argument", 2: "HIP error", 3: "non-finite result",
          4: "not converged", 5: "unsupported", 6: "allocation failed"}


class FvxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fvx status {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


class NumericalError(FvxError):
    """Non-finite residual (the reference's Numerical_error)."""


class ToleranceError(FvxError):
    """Iteration limit reached (the reference's Tolerance_error)."""


class MeshDesc(C.Structure):
    _fields_ = [("npoin", C.c_int), ("nelem", C.c_int), ("nbface", C.c_int),
                ("coords", C.POINTER(C.c_double)), ("inpoel", C.POINTER(C.c_int)),
                ("nnode", C.POINTER(C.c_int)), ("bface", C.POINTER(C.c_int)),
                ("btag", C.POINTER(C.c_int))]


class BCCfg(C.Structure):
    _fields_ = [("tag", C.c_int), ("kind", C.c_int), ("vals", C.c_double * 2)]


class PhysicsCfg(C.Structure):
    _fields_ = [("gamma", C.c_double), ("Minf", C.c_double), ("Tinf", C.c_double),
                ("Reinf", C.c_double), ("Pr", C.c_double), ("aoa", C.c_double),
                ("viscous", C.c_int), ("const_visc", C.c_int), ("nbc", C.c_int),
                ("bcs", BCCfg * 8)]


class NumericsCfg(C.Structure):
    _fields_ = [("conv_numflux", C.c_int), ("conv_numflux_jac", C.c_int),
                ("gradientscheme", C.c_int), ("reconstruction", C.c_int),
                ("limiter_param", C.c_double), ("order2", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libfvx.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.fvx_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().fvx_last_error().decode()
        cls = NumericalError if rc == 3 else ToleranceError if rc == 4 else FvxError
        raise cls(rc, msg)


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _dptr(t, numel=None):
    """Device pointer of a contiguous float64 device tensor, checked against the element
    count the kernels will touch before anything is launched."""
    if not (t.is_cuda and t.is_contiguous()):
        raise ValueError("expected a contiguous device tensor")
    if numel is not None:
        import torch
        if t.dtype != torch.float64 or t.numel() != numel:
            raise ValueError(f"expected {numel} float64 elements, got {t.numel()} {t.dtype}")
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def physics_cfg(gamma=1.4, Minf=0.5, Tinf=298.0, Reinf=float("inf"), Pr=float("nan"),
                aoa_deg=0.0, viscous=False, const_visc=False, bcs=()):
    p = PhysicsCfg()
    p.gamma, p.Minf, p.Tinf, p.Reinf, p.Pr = gamma, Minf, Tinf, Reinf, Pr
    p.aoa = float(np.deg2rad(aoa_deg))
    p.viscous, p.const_visc = int(viscous), int(const_visc)
    p.nbc = len(bcs)
    for i, b in enumerate(bcs):
        p.bcs[i].tag = b[0]
        p.bcs[i].kind = BC[b[1]]
        vals = b[2] if len(b) > 2 else (0.0, 0.0)
        p.bcs[i].vals[0] = vals[0]
        p.bcs[i].vals[1] = vals[1] if len(vals) > 1 else 0.0
    return p


def numerics_cfg(flux="ROE", jacflux=None, gradient="LEASTSQUARES", recon="NONE",
                 limiter_param=5.0, order2=True):
    n = NumericsCfg()
    n.conv_numflux = FLUX[flux]
    n.conv_numflux_jac = FLUX[jacflux or flux]
    n.gradientscheme = GRAD[gradient]
    n.reconstruction = RECON[recon]
    n.limiter_param = limiter_param
    n.order2 = int(order2 and gradient != "NONE")
    return n


def _desc(raw: RawMesh):
    arrs = dict(coords=np.ascontiguousarray(raw.coords, np.float64),
                inpoel=np.ascontiguousarray(raw.inpoel, np.int32),
                nnode=np.ascontiguousarray(raw.nnode, np.int32),
                bface=np.ascontiguousarray(raw.bface, np.int32),
                btag=np.ascontiguousarray(raw.btag, np.int32))
    d = MeshDesc(raw.npoin, raw.nelem, raw.nbface, _ptr(arrs["coords"], C.c_double),
                 _ptr(arrs["inpoel"], C.c_int), _ptr(arrs["nnode"], C.c_int),
                 _ptr(arrs["bface"], C.c_int), _ptr(arrs["btag"], C.c_int))
    return d, arrs


def rcm_permutation(raw: RawMesh) -> np.ndarray:
    """Reverse Cuthill-McKee cell order (perm[new] = old) computed by libfvx."""
    d, keep = _desc(raw)
    perm = np.zeros(raw.nelem, np.int32)
    check(lib().fvx_mesh_rcm(C.byref(d), _ptr(perm, C.c_int)))
    return perm


def hilbert_permutation(raw: RawMesh) -> np.ndarray:
    """Cells along a Hilbert curve through their centres (perm[new] = old), computed by libfvx."""
    d, keep = _desc(raw)
    perm = np.zeros(raw.nelem, np.int32)
    check(lib().fvx_mesh_hilbert(C.byref(d), _ptr(perm, C.c_int)))
    return perm


def find_lines(raw: RawMesh, threshold: float):
    """Lines of strongly coupled cells (findLines): a list of cell-index arrays."""
    d, keep = _desc(raw)
    n = C.c_int()
    ptr = np.zeros(raw.nelem // 2 + 2, np.int32)
    cells = np.zeros(raw.nelem, np.int32)
    check(lib().fvx_mesh_lines(C.byref(d), C.c_double(threshold), C.byref(n),
                               _ptr(ptr, C.c_int), _ptr(cells, C.c_int)))
    return [cells[ptr[i]:ptr[i + 1]].copy() for i in range(n.value)]


LINE_ORDERINGS = {"line": 0, "line_natural": 1, "line_rcm": 2}


def line_permutation(raw: RawMesh, threshold: float, ordering: str = "line") -> np.ndarray:
    """Cell order of "-mesh_reorder line | line_natural | line_rcm" (perm[new] = old)."""
    d, keep = _desc(raw)
    perm = np.zeros(raw.nelem, np.int32)
    check(lib().fvx_mesh_line_order(C.byref(d), C.c_double(threshold),
                                    LINE_ORDERINGS[ordering], _ptr(perm, C.c_int)))
    return perm


class Mesh:
    """Device mesh (UMesh after preprocessing) built by libfvx."""

    def __init__(self, raw: RawMesh):
        self.raw = raw
        d, keep = _desc(raw)
        h = C.c_void_p()
        check(lib().fvx_mesh_create(C.byref(d), C.byref(h)))
        self.h = h
        s = np.zeros(5, np.int32)
        check(lib().fvx_mesh_sizes(h, _ptr(s, C.c_int)))
        self.npoin, self.nelem, self.nbface, self.ninface, self.naface = map(int, s)

    def faces(self):
        intfac = np.zeros((self.naface, 4), np.int32)
        fm = np.zeros((self.naface, 3))
        area = np.zeros(self.nelem)
        rc = np.zeros((self.nelem, 2))
        check(lib().fvx_mesh_faces(self.h, _ptr(intfac, C.c_int), _ptr(fm, C.c_double),
                                   _ptr(area, C.c_double), _ptr(rc, C.c_double)))
        return intfac, fm, area, rc

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.fvx_mesh_destroy(self.h)
            self.h = None


class FlowFV:
    """Device spatial discretisation with the reference's FlowFV_base operations."""

    def __init__(self, mesh: Mesh, pcfg: PhysicsCfg, ncfg: NumericsCfg):
        self.mesh = mesh
        self.pcfg, self.ncfg = pcfg, ncfg
        h = C.c_void_p()
        check(lib().fvx_disc_create(mesh.h, C.byref(pcfg), C.byref(ncfg), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.fvx_disc_destroy(self.h)
            self.h = None

    def freestream(self) -> np.ndarray:
        out = np.zeros(4)
        check(lib().fvx_freestream_state(self.h, _ptr(out, C.c_double)))
        return out

    def compute_residual(self, u, r, gettimesteps=False, dt=None):
        """Adds -r(u) into r; fills dt with local time steps if requested (device tensors)."""
        n = self.mesh.nelem
        if gettimesteps and dt is None:
            raise ValueError("time steps requested without an output tensor")
        check(lib().fvx_compute_residual(self.h, _dptr(u, 4 * n), _dptr(r, 4 * n),
                                         int(gettimesteps),
                                         _dptr(dt, n) if dt is not None else C.c_void_p(0),
                                         _stream()))

    def get_gradients(self, u, grads):
        n = self.mesh.nelem
        check(lib().fvx_get_gradients(self.h, _dptr(u, 4 * n), _dptr(grads, 8 * n), _stream()))

    def get_primitive_gradients(self, u, grads):
        n = self.mesh.nelem
        check(lib().fvx_get_primitive_gradients(self.h, _dptr(u, 4 * n), _dptr(grads, 8 * n),
                                                _stream()))

    KERNELS = {"gradients": 0, "flux": 1, "jacobian": 2, "pc": 3, "flux_wave": 4}

    def set_launch(self, kernel, block, xcd=True):
        check(lib().fvx_set_launch(self.h, self.KERNELS[kernel], int(block), int(xcd)))

    def get_launch(self, kernel):
        b, x = C.c_int(), C.c_int()
        check(lib().fvx_get_launch(self.h, self.KERNELS[kernel], C.byref(b), C.byref(x)))
        return b.value, bool(x.value)

    def autotune(self, u, reps=20):
        """Times every launch shape of the cell kernels at state u, keeps the fastest and
        returns the timing table [(kernel, block, xcd, ms)]."""
        rows = (TuneRow * 64)()
        nr = C.c_int()
        check(lib().fvx_autotune(self.h, _dptr(u, 4 * self.mesh.nelem), int(reps), rows, 64,
                                 C.byref(nr), _stream()))
        names = {v: k for k, v in self.KERNELS.items()}
        return [(names[r.kernel], r.block, bool(r.xcd), r.ms) for r in rows[:nr.value]]

    def autotune_pc(self, u, cfl=1000.0, restart=30, rtol=0.1, precision="f64"):
        """Times preconditioner types and sweep counts on the pseudo-time system at u; returns
        ((best type, best sweeps), [(type, sweeps, iterations, ms)]), ms < 0 where the solve
        did not reach rtol."""
        rows = (TuneRow * 64)()
        nr, bp, bs = C.c_int(), C.c_int(), C.c_int()
        check(lib().fvx_autotune_pc(self.h, _dptr(u, 4 * self.mesh.nelem), C.c_double(cfl),
                                    int(restart), C.c_double(rtol), PREC[precision], rows, 64,
                                    C.byref(nr), C.byref(bp), C.byref(bs), _stream()))
        names = {v: k for k, v in PC.items()}
        table = [(names[r.kernel], r.block, r.xcd, r.ms) for r in rows[:nr.value]]
        return (names[bp.value], bs.value), table

    def profile_enable(self, on=True):
        check(lib().fvx_profile_enable(self.h, int(on)))

    def profile_read(self, kernel):
        """(total device ms, launches) of one kernel since profiling was enabled."""
        t = C.c_double()
        n = C.c_long()
        check(lib().fvx_profile_read(self.h, self.KERNELS[kernel], C.byref(t), C.byref(n)))
        return t.value, n.value

    def functionals(self, u, wall_tag=2):
        out = np.zeros(4)
        check(lib().fvx_functionals(self.h, _dptr(u, 4 * self.mesh.nelem), int(wall_tag),
                                    _ptr(out, C.c_double),
                                    _stream()))
        return dict(CL=out[0], CDp=out[1], CDsf=out[2], entropy=out[3])

    # ---- implicit solve -------------------------------------------------------------------

    def assemble_jacobian(self, u, A: "Matrix"):
        """Overwrites A with the first-order approximate Jacobian at u."""
        check(lib().fvx_assemble_jacobian(self.h, _dptr(u, 4 * self.mesh.nelem), A.h, _stream()))

    def add_pseudotime(self, A: "Matrix", dt, cfl, dtm):
        n = self.mesh.nelem
        check(lib().fvx_add_pseudotime(self.h, A.h, _dptr(dt, n), C.c_double(cfl), _dptr(dtm, n),
                                       _stream()))

    def matfree_apply(self, u, r_u, dtm, x, y, eps=1e-7):
        n = self.mesh.nelem
        check(lib().fvx_matfree_apply(self.h, _dptr(u, 4 * n), _dptr(r_u, 4 * n), _dptr(dtm, n),
                                      _dptr(x, 4 * n), _dptr(y, 4 * n), C.c_double(eps),
                                      _stream()))

    def matfree_apply_noalloc(self, u, r_u, dtm, x, y, eps=1e-7):
        """matfree_apply through the entry point that allocates no device memory (meshes of at
        most FVX_NOALLOC_MAX_CELLS cells)."""
        n = self.mesh.nelem
        check(lib().fvx_matfree_apply_noalloc(self.h, _dptr(u, 4 * n), _dptr(r_u, 4 * n),
                                              _dptr(dtm, n), _dptr(x, 4 * n), _dptr(y, 4 * n),
                                              C.c_double(eps), _stream()))

    def steady_solve(self, u, raise_on_failure=True, **kw):
        """Runs the pseudo-time loop on the device tensor u (updated in place); returns the
        statistics as a dict. Keyword arguments override solver_cfg() defaults."""
        cfg = solver_cfg(**kw)
        st = SolveStats()
        rc = lib().fvx_steady_solve(self.h, C.byref(cfg), _dptr(u, 4 * self.mesh.nelem),
                                    C.byref(st), _stream())
        out = {f: getattr(st, f) for f, _ in SolveStats._fields_}
        out["rc"] = rc
        if raise_on_failure:
            check(rc)
        return out

    def surface_data(self, u, wall_tag=2):
        """(nfaces, 4) array of face centre x, y, Cp and Cf on the wall marker, in boundary-face
        order (computeSurfaceData's output)."""
        n = C.c_int()
        check(lib().fvx_surface_data(self.h, _dptr(u, 4 * self.mesh.nelem), int(wall_tag), None,
                                     0, C.byref(n), _stream()))
        out = np.zeros((max(1, n.value), 4))
        check(lib().fvx_surface_data(self.h, _dptr(u, 4 * self.mesh.nelem), int(wall_tag),
                                     _ptr(out, C.c_double), n.value, C.byref(n), _stream()))
        return out[:n.value]

    def solve_history(self):
        """Per-step records (step, rel, abs, walltime, lin_walltime, lin_iters, cfl) of the last
        one-instance steady solve on this discretisation."""
        n = C.c_int()
        check(lib().fvx_solve_history(self.h, None, 0, C.byref(n)))
        recs = (StepRecord * max(1, n.value))()
        check(lib().fvx_solve_history(self.h, recs, n.value, C.byref(n)))
        return [{f: getattr(r, f) for f, _ in StepRecord._fields_} for r in recs[:n.value]]


class StepRecord(C.Structure):
    _fields_ = [("step", C.c_int), ("rel", C.c_double), ("abs", C.c_double),
                ("walltime", C.c_double), ("lin_walltime", C.c_double), ("lin_iters", C.c_int),
                ("cfl", C.c_double)]


PREC = {"f64": 0, "f32": 1, "bf16": 2}


class TuneRow(C.Structure):
    _fields_ = [("kernel", C.c_int), ("block", C.c_int), ("xcd", C.c_int), ("ms", C.c_double)]
PC = {"none": 0, "jacobi": 1, "sgs": 2, "ilu0": 3, "line": 4, "sgs_level": 5, "ilu0_level": 6,
      "bjacobi_ilu0": 7}


class SolverCfg(C.Structure):
    _fields_ = [("cflinit", C.c_double), ("cflfin", C.c_double), ("tol", C.c_double),
                ("maxiter", C.c_int), ("implicit", C.c_int), ("fullupdate", C.c_int),
                ("min_relax", C.c_double), ("pc", C.c_int), ("pc_precision", C.c_int),
                ("pc_sweeps", C.c_int), ("jac_precision", C.c_int), ("krylov_restart", C.c_int),
                ("krylov_rtol", C.c_double), ("matrix_free", C.c_int), ("mf_eps", C.c_double),
                ("fixed_norm", C.c_double), ("line_threshold", C.c_double),
                ("pc_subdomain_cells", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("steps", C.c_int), ("lin_iters", C.c_int), ("converged", C.c_int),
                ("status", C.c_int), ("final_rel", C.c_double), ("final_abs", C.c_double),
                ("first_norm", C.c_double), ("walltime", C.c_double)]


def solver_cfg(cflinit=500.0, cflfin=5000.0, tol=1e-7, maxiter=170, implicit=True,
               fullupdate=False, min_relax=0.2, pc="sgs", pc_precision="f64", pc_sweeps=1,
               jac_precision="f64", krylov_restart=30, krylov_rtol=1e-1, matrix_free=False,
               mf_eps=1e-7, fixed_norm=0.0, line_threshold=10.0, pc_subdomain_cells=256):
    """Defaults: testcases/naca0012/transonic-sanity-test-muscl.ctrl main solve with the
    linear-solver options of testcases/defaults.solverc."""
    c = SolverCfg()
    c.cflinit, c.cflfin, c.tol, c.maxiter = cflinit, cflfin, tol, maxiter
    c.implicit, c.fullupdate, c.min_relax = int(implicit), int(fullupdate), min_relax
    c.pc, c.pc_precision, c.jac_precision = PC[pc], PREC[pc_precision], PREC[jac_precision]
    c.pc_sweeps = pc_sweeps
    c.krylov_restart, c.krylov_rtol = krylov_restart, krylov_rtol
    c.matrix_free, c.mf_eps, c.fixed_norm = int(matrix_free), mf_eps, fixed_norm
    c.line_threshold = line_threshold
    c.pc_subdomain_cells = pc_subdomain_cells
    return c


class Matrix:
    """Device block matrix (SeqBAIJ, bs = 4) on a mesh's cell graph."""

    def __init__(self, mesh: Mesh, precision="f64"):
        self.mesh = mesh
        h = C.c_void_p()
        check(lib().fvx_matrix_create(mesh.h, PREC[precision], C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.fvx_matrix_destroy(self.h)
            self.h = None

    def apply(self, x, y):
        n = self.mesh.nelem
        check(lib().fvx_matrix_apply(self.h, _dptr(x, 4 * n), _dptr(y, 4 * n), _stream()))

    def to_bsr(self):
        """(rowptr, colind, vals[nnzb,4,4]) on the host, columns sorted."""
        import torch
        torch.cuda.synchronize()
        nnzb = C.c_int()
        check(lib().fvx_matrix_bsr_nnzb(self.h, C.byref(nnzb)))
        rowptr = np.zeros(self.mesh.nelem + 1, np.int32)
        colind = np.zeros(nnzb.value, np.int32)
        vals = np.zeros((nnzb.value, 4, 4))
        check(lib().fvx_matrix_export_bsr(self.h, _ptr(rowptr, C.c_int), _ptr(colind, C.c_int),
                                          _ptr(vals, C.c_double)))
        return rowptr, colind, vals


class Preconditioner:
    """Device preconditioner (none / block Jacobi / multicolour SGS / multicolour BILU(0) /
    line Gauss-Seidel)."""

    def __init__(self, mesh: Mesh, kind="ilu0", precision="f64", sweeps=1, line_threshold=10.0,
                 subdomain_cells=None):
        self.mesh = mesh
        h = C.c_void_p()
        if kind == "line":
            check(lib().fvx_pc_create_lines(mesh.h, C.c_double(line_threshold), PREC[precision],
                                            C.byref(h)))
        else:
            check(lib().fvx_pc_create(mesh.h, PC[kind], PREC[precision], C.byref(h)))
        self.h = h
        check(lib().fvx_pc_set_sweeps(h, int(sweeps)))
        if subdomain_cells is not None:
            self.set_subdomain_cells(subdomain_cells)
        self._A = None

    def setup(self, A: "Matrix"):
        check(lib().fvx_pc_setup(self.h, A.h, _stream()))
        self._A = A          # the sweeps read A: keep it alive with the preconditioner

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.fvx_pc_destroy(self.h)
            self.h = None

    def apply(self, r, z):
        n = self.mesh.nelem
        check(lib().fvx_pc_apply(self.h, _dptr(r, 4 * n), _dptr(z, 4 * n), _stream()))

    def ordering(self):
        col = np.zeros(self.mesh.nelem, np.int32)
        nc = C.c_int()
        check(lib().fvx_pc_ordering(self.h, _ptr(col, C.c_int), C.byref(nc)))
        return col, nc.value

    def set_subdomain_cells(self, cells):
        """Block Jacobi + ILU(0): re-partitions into subdomains of about `cells` cells."""
        check(lib().fvx_pc_set_subdomain_cells(self.h, int(cells)))

    def subdomains(self):
        """Block Jacobi + ILU(0): the subdomain of every cell and the subdomain count."""
        part = np.zeros(self.mesh.nelem, np.int32)
        ns = C.c_int()
        check(lib().fvx_pc_subdomains(self.h, _ptr(part, C.c_int), C.byref(ns)))
        return part, ns.value


def linear_solve(A: Matrix, pc: Preconditioner, b, x, restart=30, rtol=1e-1):
    """FGMRES(restart) from zero; returns the iteration count."""
    it = C.c_int()
    n = A.mesh.nelem
    if pc.mesh is not A.mesh:
        raise ValueError("preconditioner and matrix belong to different meshes")
    check(lib().fvx_linear_solve(A.h, pc.h, _dptr(b, 4 * n), _dptr(x, 4 * n), int(restart),
                                 C.c_double(rtol),
                                 C.byref(it), _stream()))
    return it.value


class Batch:
    """Operating points (Mach, angle of attack) solved together on one discretisation; state
    tensors are [ninst, ncell, 4]."""

    def __init__(self, disc: FlowFV, mach, aoa_deg):
        mach = np.ascontiguousarray(mach, np.float64)
        aoa = np.ascontiguousarray(np.deg2rad(aoa_deg), np.float64)
        if mach.shape != aoa.shape or mach.ndim != 1:
            raise ValueError("mach and aoa_deg must be 1-D arrays of equal length")
        self.disc, self.ninst = disc, len(mach)
        h = C.c_void_p()
        check(lib().fvx_batch_create(disc.h, self.ninst, _ptr(mach, C.c_double),
                                     _ptr(aoa, C.c_double), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.fvx_batch_destroy(self.h)
            self.h = None

    def freestream(self) -> np.ndarray:
        out = np.zeros((self.ninst, 4))
        check(lib().fvx_batch_freestream(self.h, _ptr(out, C.c_double)))
        return out

    def compute_residual(self, u, r, gettimesteps=False, dt=None):
        n = self.ninst * self.disc.mesh.nelem
        if gettimesteps and dt is None:
            raise ValueError("time steps requested without an output tensor")
        check(lib().fvx_batch_compute_residual(self.h, _dptr(u, 4 * n), _dptr(r, 4 * n),
                                               int(gettimesteps),
                                               _dptr(dt, n) if dt is not None else C.c_void_p(0),
                                               _stream()))

    def steady_solve(self, u, **kw):
        """Solves every instance in place; returns one statistics dict per instance."""
        cfg = solver_cfg(**kw)
        st = (SolveStats * self.ninst)()
        check(lib().fvx_batch_steady_solve(self.h, C.byref(cfg),
                                           _dptr(u, 4 * self.ninst * self.disc.mesh.nelem), st,
                                           _stream()))
        return [{f: getattr(s, f) for f, _ in SolveStats._fields_} for s in st]

    def set_norms(self, norms=None):
        """Fixed per-instance normalisers of the convergence test (None: first-step norms)."""
        if norms is None:
            check(lib().fvx_batch_set_norms(self.h, C.c_void_p(0)))
            return
        norms = np.ascontiguousarray(norms, np.float64)
        if norms.shape != (self.ninst,):
            raise ValueError("one normaliser per instance")
        check(lib().fvx_batch_set_norms(self.h, _ptr(norms, C.c_double)))

    def residual_norms(self, u):
        """sqrt(sum r_E^2 area) of every instance at u."""
        out = np.zeros(self.ninst)
        check(lib().fvx_batch_residual_norms(self.h, _dptr(u, 4 * self.ninst * self.disc.mesh.nelem),
                                             _ptr(out, C.c_double), _stream()))
        return out

    def functionals(self, u, wall_tag=2):
        out = np.zeros((self.ninst, 4))
        check(lib().fvx_batch_functionals(self.h, _dptr(u, 4 * self.ninst * self.disc.mesh.nelem),
                                          int(wall_tag), _ptr(out, C.c_double), _stream()))
        return [dict(CL=o[0], CDp=o[1], CDsf=o[2], entropy=o[3]) for o in out]


def _fptr(t, numel, dtype="float32"):
    """Device pointer of a contiguous device tensor of `numel` elements of `dtype`."""
    import torch
    want = getattr(torch, dtype)
    if not (t.is_cuda and t.is_contiguous()) or t.dtype != want or t.numel() != numel:
        raise ValueError(f"expected a contiguous device tensor of {numel} {dtype} elements")
    return C.c_void_p(t.data_ptr())


class MPNN:
    """Learned warm start on one mesh: a 4-layer message-passing network (width 64, bf16
    weights, fp32 accumulation) trained and evaluated by the library's device kernels.

    Per-graph tensors are [ngraph, ncell, 4]; angles of attack are in degrees here."""

    def __init__(self, mesh: Mesh, wall_tag: int, max_graphs: int, gamma: float = 1.4):
        h = C.c_void_p()
        check(lib().fvx_mpnn_create(mesh.h, int(wall_tag), int(max_graphs), C.byref(h)))
        self.h, self.mesh, self.max_graphs, self.gamma = h, mesh, int(max_graphs), gamma
        self.N = mesh.nelem
        self.npoints = 0

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.fvx_mpnn_destroy(self.h)
            self.h = None

    @staticmethod
    def nparams() -> int:
        return int(lib().fvx_mpnn_nparams())

    def features(self):
        N = self.N
        out = dict(stat=np.zeros((N, 6), np.float32), edge=np.zeros((N, 4, 4), np.float32),
                   nbr=np.zeros((N, 4), np.int32), rev=np.zeros((N, 4), np.int32),
                   invdeg=np.zeros(N, np.float32))
        check(lib().fvx_mpnn_features(self.h, _ptr(out["stat"], C.c_float),
                                      _ptr(out["edge"], C.c_float), _ptr(out["nbr"], C.c_int),
                                      _ptr(out["rev"], C.c_int), _ptr(out["invdeg"], C.c_float)))
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.shape != (self.nparams(),):
            raise ValueError("parameter vector has the wrong length")
        check(lib().fvx_mpnn_set_params(self.h, _ptr(w, C.c_float)))

    def get_params(self) -> np.ndarray:
        w = np.zeros(self.nparams(), np.float32)
        check(lib().fvx_mpnn_get_params(self.h, _ptr(w, C.c_float)))
        return w

    def set_points(self, mach, aoa_deg):
        mach = np.ascontiguousarray(mach, np.float64)
        aoa = np.ascontiguousarray(np.deg2rad(aoa_deg), np.float64)
        if mach.shape != aoa.shape or mach.ndim != 1:
            raise ValueError("mach and aoa_deg must be 1-D arrays of equal length")
        check(lib().fvx_mpnn_set_points(self.h, len(mach), _ptr(mach, C.c_double),
                                        _ptr(aoa, C.c_double), C.c_double(self.gamma), _stream()))
        self.npoints = len(mach)

    def predict(self, y, ngraph=None):
        G = ngraph or self.npoints
        check(lib().fvx_mpnn_predict(self.h, G, _fptr(y, G * self.N * 4), _stream()))

    def warm_state(self, u, ngraph=None, floor=0.05):
        G = ngraph or self.npoints
        check(lib().fvx_mpnn_warm_state(self.h, G, C.c_double(floor),
                                        _fptr(u, G * self.N * 4, "float64"), _stream()))

    def targets(self, u, t, ngraph=None):
        G = ngraph or self.npoints
        check(lib().fvx_mpnn_targets(self.h, G, _fptr(u, G * self.N * 4, "float64"),
                                     _fptr(t, G * self.N * 4), _stream()))

    def _ids(self, ids, targets):
        ids = np.ascontiguousarray(ids, np.int32)
        nsamp = targets.numel() // (self.N * 4)
        if targets.numel() != nsamp * self.N * 4 or ids.ndim != 1 or len(ids) > self.npoints:
            raise ValueError("targets must be [nsamples, ncell, 4] and ids at most one per point")
        if len(ids) and (ids.min() < 0 or ids.max() >= nsamp):
            raise ValueError("sample index out of range")
        return ids, nsamp

    def gradient(self, targets, ids, grad=None):
        ids, nsamp = self._ids(ids, targets)
        loss = C.c_float()
        gp = _fptr(grad, self.nparams()) if grad is not None else C.c_void_p(0)
        check(lib().fvx_mpnn_gradient(self.h, len(ids), _fptr(targets, nsamp * self.N * 4),
                                      _ptr(ids, C.c_int), gp, C.byref(loss), _stream()))
        return float(loss.value)

    def train_step(self, targets, ids, lr, want_loss=True):
        ids, nsamp = self._ids(ids, targets)
        loss = C.c_float()
        check(lib().fvx_mpnn_train_step(self.h, len(ids), _fptr(targets, nsamp * self.N * 4),
                                        _ptr(ids, C.c_int), C.c_double(lr),
                                        C.byref(loss) if want_loss else C.c_void_p(0), _stream()))
        return float(loss.value) if want_loss else None
