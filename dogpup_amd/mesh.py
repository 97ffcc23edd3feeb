"""DMesh / DOptode: the reference's dMesh and dOptode API for the hot path.

``DMesh`` mirrors ``DOGPUP/dMesh.m:1-247`` (constructor, update_properties,
add_optode, flu_solve, meas_flu, adj_flu_solve, J_complex) and ``DOptode``
mirrors ``DOGPUP/dOptode.m:34-120``.  Differences of form, not of meaning:
indices (elements, links, optodes) are 0-based; arrays are torch tensors on
the GPU; solver outputs stay device-resident in solver layout (see
``FluenceField``) and are converted to the reference's NN x NF x NS layout on
request.  All numerical work of the path (assembly, FSAI factors, BiCGSTAB,
detector sampling, transforms, Jacobian) runs in libdogpup_hip.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, grid, topology
from .transforms import td2fc

DEV = "cuda"
METHODS = {"bicgstab": _lib.METHOD_BICGSTAB, "cocg": _lib.METHOD_COCG}
MODES = {"c128": _lib.MODE_C128, "c64": _lib.MODE_C64_REFINE, "mixed": _lib.MODE_C64_REFINE,
         "plain": _lib.MODE_C128_PLAIN, "fp64": _lib.MODE_C128}


def _i32(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=DEV)


def _f64(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=DEV)


def default_method(mode: str) -> str:
    """COCG for the fused and mixed modes, BiCGSTAB (the reference's, the
    only one it has) for the reference-structure mode."""
    return "bicgstab" if MODES[mode] == _lib.MODE_C128_PLAIN else "cocg"


def group_width(ncols: int, mode: str = "mixed") -> int:
    """Default columns per solver group, measured on the 269k-node,
    32-source workload:
    - up to 8 columns: 8;
    - the complex64 refinement form ('mixed') with more than 16 columns: 32,
      whose SpMVs walk each chunk in two 16-column halves (the x halo of one
      half fits the LDS three times per CU; the matrix slice serves both):
      579 ms per solve against 601 ms with 16-wide groups;
    - otherwise 16: the FP64 forms' gather SpMVs are faster on 16-wide groups
      (reference-structure form 10.3 vs 10.8 s).
    Pass nsp explicitly for other widths."""
    if ncols <= 8:
        return 8
    if mode == "mixed" and ncols > 16:
        return 32
    return 16


class DOptode:
    """Sources, detectors, links and the source spectrum (dOptode.m:34-74)."""

    def __init__(self, s_pos, d_pos, link, avg_pow, tpsf, t_axis, nf):
        self.link = np.asarray(link, np.int32).reshape(-1, 2)
        self.Nf = int(nf)
        t_axis = np.asarray(t_axis, float).reshape(-1)
        self.tAxis = t_axis
        self.dt = t_axis[1] - t_axis[0]
        T = t_axis[-1]                                   # dOptode.m:57
        self.fAxis = np.arange(nf) * 2.0 * np.pi / T     # :58-59
        if nf < 2:
            raise ValueError("Nf must be at least 2 (df = fAxis(2) - fAxis(1))")
        self.df = self.fAxis[1] - self.fAxis[0]          # :60
        self.d_positions = np.asarray(d_pos, float).reshape(-1, 3)
        self.s_positions = np.asarray(s_pos, float).reshape(-1, 3)
        self.update_tpsf(avg_pow, tpsf, t_axis, nf)
        self.s_bary = None
        self.d_bary = None

    def update_tpsf(self, avg_pow, tpsf, t_axis, nf):
        """dOptode.m:101-120: TPSF scaled to mean avg_pow, s_fpsf = td2fc."""
        tpsf = np.asarray(tpsf, float).reshape(-1)
        self.s_avgPow = float(avg_pow)
        self.s_tpsf = tpsf * (avg_pow / tpsf.mean())
        sf = td2fc(torch.as_tensor(self.s_tpsf[None, :]), self.fAxis, t_axis, 1)
        self.s_fpsf = sf[0].cpu().numpy()

    def snap2mesh(self, mesh: "DMesh"):
        """dOptode.m:77-97: detectors snapped to the surface, sources snapped
        then moved 1/min(musp) along the inward normal."""
        d_pos, _ = topology.snap_to_surface(mesh.node, mesh.face, mesh.bnd, self.d_positions)
        s_pos, s_dir = topology.snap_to_surface(mesh.node, mesh.face, mesh.bnd, self.s_positions)
        s_in = s_pos + (1.0 / mesh.musp.min()) * s_dir
        eid, w, _ = mesh.locator().find(np.concatenate([d_pos, s_in]))
        eid, w = eid.cpu().numpy().astype(np.int64), w.cpu().numpy()
        nd = d_pos.shape[0]
        d_id, d_w, s_id, s_w = eid[:nd], w[:nd], eid[nd:], w[nd:]
        if (d_id < 0).any() or (s_id < 0).any():
            raise ValueError("an optode could not be located inside the mesh")
        self.d_positions, self.d_dirs = d_pos, None
        self.s_positions, self.s_dirs = s_in, s_dir
        self.d_bary = (d_id, topology.clean_bary(d_w))
        self.s_bary = (s_id, topology.clean_bary(s_w))
        return self


@dataclass
class FluenceField:
    """Device-resident fluence in solver layout: complex128
    [NF*NCB][NN][NSP], column c of frequency k at group k*NCB + c//NSP."""

    data: torch.Tensor
    nn: int
    nf: int
    ncols: int
    nsp: int
    iters: np.ndarray = None
    relres: np.ndarray = None
    status: np.ndarray = None
    stats: dict = field(default_factory=dict)

    def to_reference_layout(self) -> torch.Tensor:
        """NN x NF x NS complex128 (the reference's phi layout)."""
        out = torch.empty((self.ncols, self.nf, self.nn), dtype=torch.complex128, device=DEV)
        _lib.call("dogpup_layout_from_solver", _lib.ptr(torch.view_as_real(self.data)),
                  _lib.ptr(torch.view_as_real(out)), self.nn, self.nf, self.ncols, self.nsp,
                  _lib.stream_ptr())
        return out.permute(2, 1, 0)


class DMesh:
    """Mesh, optical properties and forward matrices (dMesh.m:48-88)."""

    cVac = topology.C_VAC

    def __init__(self, node, elem, nr, props, reference_faces=True):
        self.node = np.asarray(node, float)
        self.elem = np.asarray(elem, np.int64)
        self.nr = float(nr)
        nn = self.node.shape[0]
        props = np.atleast_2d(np.asarray(props, float))
        self.mua = np.broadcast_to(props[:, 0], (nn,)).copy()      # dMesh.m:68-73
        self.musp = np.broadcast_to(props[:, 1], (nn,)).copy()
        self.kappa = 1.0 / (3.0 * (self.mua + self.musp))          # :74
        self.reference_faces = reference_faces
        self.optode = None
        self.basis = None            # voxel basis from mesh2grid (None: nodal basis)
        self._locator = None
        self._solvers = {}
        self.prepare_mesh()
        self.update_properties()

    # -- geometry (dMesh.m:945-1030) ------------------------------------------
    def prepare_mesh(self):
        nn = self.node.shape[0]
        self.c = self.cVac / self.nr
        self.face, self.area = topology.true_faces(self.node, self.elem)
        self.bnd = np.zeros(nn, np.int32)
        self.bnd[np.unique(self.face.ravel())] = 1
        self.R = topology.robin_factor(self.nr) * self.bnd
        self.bface = topology.robin_faces(self.elem, self.bnd, self.reference_faces, self.face)
        self.pattern = topology.system_pattern(nn, self.elem, self.bface)
        rG, cG = topology.fsai_pattern(self.pattern.rptr, self.pattern.cptr)
        self.g_pattern = (rG, cG)
        d = self._dev = {}
        d["node"] = _f64(self.node)
        d["elem"] = _i32(self.elem)
        d["bface"] = _i32(self.bface)
        d["R"] = _f64(self.R)
        d["rA"], d["cA"] = _i32(self.pattern.rptr), _i32(self.pattern.cptr)
        d["slot_ptr"], d["slot_src"] = _i32(self.pattern.slot_ptr), _i32(self.pattern.slot_src)
        d["rG"], d["cG"] = _i32(rG), _i32(cG)
        self.nnzA = int(self.pattern.cptr.size)
        self.nnzG = int(cG.size)
        self.solver_order = topology.morton_order(self.node, self.elem)

    @property
    def nn(self):
        return self.node.shape[0]

    # -- assembly (dMesh.m:80-88, 1033-1194) -----------------------------------
    def update_properties(self, props=None, kappa=None):
        """Re-assemble K and M (dMesh.m:80-88).  ``kappa`` (optional, NN)
        overrides kappa = 1/(3(mua+musp)); the Jacobian and its finite
        differences hold kappa fixed."""
        if props is not None:
            props = np.asarray(props, float)
            self.mua[:] = props[:, 0]
            self.musp[:] = props[:, 1]
            self.kappa[:] = 1.0 / (3.0 * (self.mua + self.musp))
        if kappa is not None:
            self.kappa[:] = np.asarray(kappa, float)
        d = self._dev
        ne = self.elem.shape[0]
        self.K = torch.empty(self.nnzA, dtype=torch.float64, device=DEV)   # Re A (full symmetric)
        self.M = torch.empty(self.nnzA, dtype=torch.float64, device=DEV)   # mass matrix values
        work = torch.empty(ne * 17, dtype=torch.float64, device=DEV)
        mua_d, kappa_d = _f64(self.mua), _f64(self.kappa)
        _lib.call("dogpup_assemble", self.nn, ne, int(self.bface.shape[0]), self.nnzA,
                  _lib.ptr(d["node"]), _lib.ptr(d["elem"]), _lib.ptr(d["bface"]),
                  _lib.ptr(mua_d), _lib.ptr(kappa_d), _lib.ptr(d["R"]),
                  _lib.ptr(d["slot_ptr"]), _lib.ptr(d["slot_src"]), 1.0, _lib.ptr(work),
                  _lib.ptr(self.K), _lib.ptr(self.M), _lib.stream_ptr())
        self._matrix_version = getattr(self, "_matrix_version", 0) + 1
        return self

    def system_csr(self):
        """(rptr, cptr, K, M) of the full symmetric system on the device;
        A_k = K + i*k*(df/c)*M."""
        d = self._dev
        return d["rA"], d["cA"], self.K, self.M

    # -- point location, voxel basis, interpolation (dMesh.m:286-411) ---------
    def locator(self) -> grid.Locator:
        if self._locator is None:
            self._locator = grid.Locator(self._dev["node"], self._dev["elem"])
        return self._locator

    def mesh2grid(self, x, y, z, tol=1e-9):
        """Attach the voxel basis of grid axes x, y, z (dMesh.m:286-380):
        grid, gridSize, dxyz, m2g, g2m, gridinmesh (and ``basis``, which the
        Jacobian and the LM driver use)."""
        b = grid.mesh2grid(self._dev["node"], self._dev["elem"], x, y, z, tol, locator=self.locator())
        if self.basis is None:
            self._nodal_dxyz = getattr(self, "dxyz", None)      # restored by clear_grid
        self.basis = b
        self.grid, self.gridSize, self.dxyz = b.grid, b.gridSize, b.dxyz
        self.m2g, self.g2m, self.gridinmesh = b.m2g, b.g2m, b.gridinmesh
        return self

    def clear_grid(self):
        """Back to the nodal basis (m2g = g2m = I), with the dxyz the mesh had
        before mesh2grid."""
        if self.basis is not None:
            self.dxyz = self._nodal_dxyz
        self.basis = None
        return self

    def mesh2mesh(self, val, old_node, old_elem):
        """``val`` on the mesh (old_node, old_elem) interpolated to this
        mesh's nodes (dMesh.m:383-411)."""
        return grid.mesh2mesh(self.node, val, old_node, old_elem)

    def mesh_slice(self, plane):
        return grid.mesh_slice(self._dev["node"], self._dev["elem"], plane, locator=self.locator())

    def mesh_slice3(self):
        return grid.mesh_slice3(self._dev["node"], self._dev["elem"], locator=self.locator())

    # -- optodes -------------------------------------------------------------
    def add_optode(self, optode: DOptode):
        optode.snap2mesh(self)
        self.optode = optode
        return self

    def clear_optode(self):
        self.optode = None
        return self

    def _bary_dev(self, which):
        eid, w = self.optode.s_bary if which == "s" else self.optode.d_bary
        nodes = self.elem[eid]                      # N x 4
        return _i32(nodes), _f64(w)

    # -- solver -----------------------------------------------------------------
    def solver(self, ncols, mode="mixed", method=None, freqs=None):
        fk = None if freqs is None else tuple(int(k) for k in np.asarray(freqs).reshape(-1))
        key = (ncols, mode) if fk is None else (ncols, mode, fk)
        s = self._solvers.get(key)
        if s is None:
            s = FluenceSolver(self, ncols, MODES[mode], nsp=group_width(ncols, mode), freqs=freqs)
            self._solvers[key] = s
        s.set_method(default_method(mode) if method is None else method)
        return s

    def solve_point_sources(self, which="s", mode="mixed", tol=None, maxit=1000, extra=None,
                            method=None, adj_tol=None, cont=None, freqs=None):
        """Solve for the sources ('s'), detectors ('d') or both ('sd', forward
        and adjoint batched in one launch: columns 0..NS-1 sources,
        NS..NS+ND-1 detectors).  method: 'cocg' (conjugate orthogonal CG, the
        default of the fused and mixed modes) or 'bicgstab' (the reference's,
        the only method of the 'plain' mode); both converge to the same
        solution.  adj_tol: relative residual of the detector (adjoint)
        columns when it differs from tol (fused and mixed modes), as separate
        flu_solve / adj_flu_solve calls with two tolerances would have.
        cont (mixed mode): frequency continuation, each frequency of a later
        level started from the converged solutions of neighbouring
        frequencies solved before it, using the source spectrum s_fpsf that
        every right-hand side carries: True / None / 'galerkin' (the
        default) combines the six nearest by the Galerkin solution in their
        span, 'lagrange' interpolates the four nearest; False starts every
        right-hand side from 0, as the reference does.
        freqs: indices into the optode's frequency axis to solve (a frequency
        shard, such as one GPU's share when the frequencies are split across
        GPUs); None solves every frequency.  The field then holds those
        frequencies in the given order."""
        parts = [self._bary_dev(w) for w in which]
        nodes = torch.cat([p[0] for p in parts])
        w = torch.cat([p[1] for p in parts])
        ncols = nodes.shape[0]
        S = self.solver(ncols, mode, method, freqs)
        s_fpsf = np.asarray(self.optode.s_fpsf)[S.freqs]
        if MODES[mode] == _lib.MODE_C64_REFINE:
            S.set_continuation(cont, s_fpsf)
        elif cont:
            raise ValueError("frequency continuation is for the mixed mode")
        spec = torch.as_tensor(s_fpsf, dtype=torch.complex128, device=DEV)
        b = S.empty_block(zero=False)               # dogpup_point_sources clears it
        _lib.call("dogpup_point_sources", _lib.ptr(torch.view_as_real(b)), self.nn, S.nf, ncols, S.nsp,
                  _lib.ptr(nodes), _lib.ptr(w), _lib.ptr(torch.view_as_real(spec)), _lib.stream_ptr())
        if tol is None:
            tol = 1e-12 if MODES[mode] != _lib.MODE_C64_REFINE else 1e-10
        col_tol = None
        if adj_tol is not None and "d" in which:
            col_tol = np.concatenate([np.full(len(p[0]), adj_tol if w_ == "d" else tol)
                                      for w_, p in zip(which, parts)])
        return S.solve(b, tol=tol, maxit=maxit, col_tol=col_tol)

    # -- forward / adjoint (dMesh.m:106-202) ------------------------------------
    def flu_solve(self, display=False, mode="mixed", tol=None):
        """phi (FluenceField, NS columns) and data (NM x NF complex)."""
        phi = self.solve_point_sources("s", mode, tol)
        return phi, self.meas_flu(phi)

    def adj_flu_solve(self, display=False, mode="mixed", tol=None):
        return self.solve_point_sources("d", mode, tol)

    def meas_flu(self, phi: FluenceField):
        """data(m,k) = (D phi_k)(d_m, s_m) (dMesh.m:144-171); NM x NF."""
        dn, dw = self._bary_dev("d")
        link = _i32(self.optode.link)
        nm = link.shape[0]
        out = torch.empty((phi.nf, nm), dtype=torch.complex128, device=DEV)
        _lib.call("dogpup_meas", _lib.ptr(torch.view_as_real(phi.data)), _lib.ptr(torch.view_as_real(out)),
                  _lib.ptr(dn), _lib.ptr(dw), _lib.ptr(link), self.nn, phi.nf, phi.ncols, phi.nsp, nm,
                  int(dn.shape[0]), _lib.stream_ptr())
        return out.t()


class FluenceSolver:
    """Device-resident batched FSAI-BiCGSTAB for one mesh, column count and
    mode (wraps dogpup_solver_*)."""

    def __init__(self, mesh: DMesh, ncols: int, mode: int, order=None, nsp=None, freqs=None):
        if mesh.optode is None:
            raise ValueError("add an optode first (the frequency axis comes from it)")
        self.mesh = mesh
        nfa = mesh.optode.Nf
        # frequency k of the handle is index freqs[k] of the optode's axis
        self.freqs = np.arange(nfa) if freqs is None else np.asarray(freqs, np.int64).reshape(-1)
        if self.freqs.size == 0 or self.freqs.min() < 0 or self.freqs.max() >= nfa:
            raise ValueError(f"freqs must be indices of the optode's {nfa} frequencies")
        self.nf = int(self.freqs.size)
        self.ncols = ncols
        self.nsp = int(nsp) if nsp else group_width(ncols, "mixed" if mode == _lib.MODE_C64_REFINE else "fp64")
        self.ncb = -(-ncols // self.nsp)
        self.mode = mode
        d = mesh._dev
        if order is None:
            order = mesh.solver_order
        h = C.c_void_p()
        _lib.call("dogpup_solver_create", C.byref(h), mesh.nn, self.nf, ncols, self.nsp, mode,
                  mesh.nnzA, _lib.ptr(d["rA"]), _lib.ptr(d["cA"]), mesh.nnzG, _lib.ptr(d["rG"]),
                  _lib.ptr(d["cG"]), _lib.ptr(np.ascontiguousarray(order, np.int32)))
        self.h = h
        self._version = None
        self._col_tol = False
        self.method = "cocg" if mode != _lib.MODE_C128_PLAIN else "bicgstab"
        # continuation state: the library's default (Galerkin guesses, flat
        # spectrum, in the refinement mode)
        self._cont_key = (_lib.CONT_GALERKIN if mode == _lib.MODE_C64_REFINE else 0, None)

    def set_method(self, method: str):
        """Krylov method of the fused modes: 'bicgstab' or 'cocg'."""
        if method not in METHODS:
            raise ValueError(f"unknown method {method!r}; expected one of {sorted(METHODS)}")
        _lib.call("dogpup_solver_set_method", self.h, METHODS[method])
        self.method = method

    def set_continuation(self, on=True, spectrum=None):
        """Frequency continuation of the mixed mode
        (dogpup_solver_set_freq_continuation): on = True / None / 'galerkin',
        'lagrange' or False; spectrum: the complex factor per frequency that
        every right-hand side carries (None: flat)."""
        on = _lib.cont_kind(on)
        spec = None if spectrum is None else np.ascontiguousarray(np.asarray(spectrum, np.complex128).reshape(-1))
        key = (on, None if spec is None else spec.tobytes())
        if key == self._cont_key:
            return
        if on and self.mode != _lib.MODE_C64_REFINE:
            raise ValueError("frequency continuation is for the mixed (refinement) mode")
        if spec is not None and spec.size != self.nf:
            raise ValueError(f"spectrum needs {self.nf} values, got {spec.size}")
        _lib.call("dogpup_solver_set_freq_continuation", self.h, on,
                  None if spec is None else _lib.ptr(spec.view(np.float64)))
        self._cont_key = key

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.lib().dogpup_solver_destroy(self.h)
        except Exception:
            pass

    def empty_block(self, zero=True):
        """A right-hand-side block in the solver layout, [NF*NCB][NN][NSP]
        complex128 (zeros, or uninitialised with zero=False)."""
        make = torch.zeros if zero else torch.empty
        return make((self.nf * self.ncb, self.mesh.nn, self.nsp), dtype=torch.complex128, device=DEV)

    def ensure_matrix(self):
        m = self.mesh
        if self._version == m._matrix_version:
            return
        fscale = _f64(self.freqs * m.optode.df / m.c)    # A_k = K + i k df/c M
        _lib.call("dogpup_solver_set_matrix", self.h, _lib.ptr(m.K), _lib.ptr(m.M), _lib.ptr(fscale),
                  _lib.stream_ptr())
        torch.cuda.current_stream().synchronize()
        self._version = m._matrix_version

    def set_col_tol(self, col_tol):
        """Per-column tolerances (NCOLS, the same at every frequency, or
        NF x NCOLS); None restores the solve's single tol."""
        if col_tol is None:
            if self._col_tol:
                _lib.call("dogpup_solver_set_rhs_tol", self.h, None, 0)
                self._col_tol = False
            return
        t = np.asarray(col_tol, np.float64)
        t = np.ascontiguousarray(np.broadcast_to(t, (self.nf, self.ncols)))
        _lib.call("dogpup_solver_set_rhs_tol", self.h, _lib.ptr(t), t.size)
        self._col_tol = True

    def solve(self, b: torch.Tensor, tol=1e-10, maxit=1000, x: torch.Tensor | None = None, col_tol=None):
        """col_tol: per-column tolerances (see set_col_tol); None = tol for
        every column."""
        self.ensure_matrix()
        self.set_col_tol(col_tol)
        if x is None:
            x = torch.empty_like(b)
        n = self.nf * self.ncols
        iters = np.zeros(n, np.int32)
        relres = np.zeros(n, np.float64)
        status = np.zeros(n, np.int32)
        _lib.call("dogpup_solver_solve", self.h, _lib.ptr(torch.view_as_real(b)),
                  _lib.ptr(torch.view_as_real(x)), float(tol), int(maxit), _lib.ptr(iters),
                  _lib.ptr(relres), _lib.ptr(status), _lib.stream_ptr())
        st = np.zeros(4)
        _lib.call("dogpup_solver_stats", self.h, _lib.ptr(st), 4)
        shape = (self.nf, self.ncols)
        return FluenceField(x, self.mesh.nn, self.nf, self.ncols, self.nsp, iters.reshape(shape),
                            relres.reshape(shape), status.reshape(shape),
                            {"max_iters": st[0], "group_iters": st[1], "sweeps": st[2],
                             "inner_iters": st[3]})

    def fsai_values(self) -> torch.Tensor:
        """FSAI factors of every frequency, complex128 [NF][nnzG]."""
        self.ensure_matrix()
        g = torch.empty((self.nf, self.mesh.nnzG), dtype=torch.complex128, device=DEV)
        _lib.call("dogpup_solver_get_fsai", self.h, _lib.ptr(torch.view_as_real(g)), _lib.stream_ptr())
        return g
