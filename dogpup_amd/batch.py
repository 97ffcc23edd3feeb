"""Many independent meshes solved in one handle (``dogpup_batch_create``).

The meshes' CSR patterns are concatenated into one block-diagonal pattern;
every pass of the batched Krylov solve then walks one work list over all
meshes' (frequency, column-block) groups, so meshes from a few thousand to
hundreds of thousands of nodes fill one launch together, and the captured
iteration (a HIP graph) covers the whole batch.  Each mesh keeps its own
frequencies (from its optode), its own columns (its sources, or sources and
detectors) and its own convergence per RHS.  The reference solves one mesh
per ``flu_solve`` call (``dMesh.m:106-141``); the per-mesh results here are
those calls' results.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .mesh import DEV, METHODS, MODES, FluenceField, group_width


def _i32(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=DEV)


class MeshBatch:
    """Batched forward (or adjoint, or both) solves of several DMesh objects.

    ``meshes``: DMesh instances with optodes.  ``which``: 's', 'd' or 'sd' as
    in DMesh.solve_point_sources.  ``nsp``: columns per group (8, 16, 32 or
    64; default: default_width, the width with the least padded work).
    ``tuning``: (tile_rows, unroll_a, unroll_g, graph_k), e.g. from
    dogpup_amd.autotune.  ``cont`` (mixed mode): frequency continuation
    within each mesh, normalised by each mesh's source spectrum (the
    right-hand sides of rhs() carry it); False starts every RHS from 0.
    """

    def __init__(self, meshes, which="s", mode="mixed", nsp=None, method="cocg", tuning=None, cont=True):
        self.meshes = list(meshes)
        if not self.meshes:
            raise ValueError("empty batch")
        for m in self.meshes:
            if m.optode is None:
                raise ValueError("every mesh needs an optode (the frequency axis comes from it)")
        self.which = which
        self.mode = mode
        self.ncols = [sum(len((m.optode.s_bary if w == "s" else m.optode.d_bary)[0]) for w in which)
                      for m in self.meshes]
        self.nf = [m.optode.Nf for m in self.meshes]
        self.nsp = int(nsp) if nsp else self.default_width(self.meshes, self.ncols, self.nf, mode)
        self.ncb = [-(-c // self.nsp) for c in self.ncols]
        nn = [m.nn for m in self.meshes]
        self.row0 = np.concatenate([[0], np.cumsum(nn)]).astype(np.int32)
        # block-diagonal patterns (global row and column numbering)
        rA, cA, rG, cG, order = [], [], [], [], []
        pa = pg = 0
        for i, m in enumerate(self.meshes):
            d = m._dev
            r0 = int(self.row0[i])
            rA.append(d["rA"][:-1] + pa)
            cA.append(d["cA"] + r0)
            rG.append(d["rG"][:-1] + pg)
            cG.append(d["cG"] + r0)
            order.append(np.asarray(m.solver_order, np.int64) + r0)
            pa += m.nnzA
            pg += m.nnzG
        self.rA = torch.cat(rA + [torch.tensor([pa], dtype=torch.int32, device=DEV)]).contiguous()
        self.cA = torch.cat(cA).contiguous()
        self.rG = torch.cat(rG + [torch.tensor([pg], dtype=torch.int32, device=DEV)]).contiguous()
        self.cG = torch.cat(cG).contiguous()
        self.nnzA, self.nnzG = pa, pg
        h = C.c_void_p()
        _lib.call("dogpup_batch_create", C.byref(h), len(self.meshes), _lib.ptr(self.row0),
                  _lib.ptr(np.asarray(self.nf, np.int32)), _lib.ptr(np.asarray(self.ncols, np.int32)), self.nsp,
                  MODES[mode], pa, _lib.ptr(self.rA), _lib.ptr(self.cA), pg, _lib.ptr(self.rG),
                  _lib.ptr(self.cG), _lib.ptr(np.concatenate(order).astype(np.int32)))
        self.h = h
        self.set_method(method)
        if tuning is not None:
            self.set_tuning(*tuning)
        if MODES[mode] == _lib.MODE_C64_REFINE:
            spec = np.ascontiguousarray(np.concatenate([np.asarray(m.optode.s_fpsf, np.complex128).reshape(-1)
                                                        for m in self.meshes]))
            on = _lib.cont_kind(cont)
            _lib.call("dogpup_solver_set_freq_continuation", self.h, on,
                      _lib.ptr(spec.view(np.float64)) if on else None)
        self._versions = None
        # block of each mesh in the batch vectors: nf*ncb groups of nn rows
        sizes = [f * b * n * self.nsp for f, b, n in zip(self.nf, self.ncb, nn)]
        self.voff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)

    @staticmethod
    def default_width(meshes, ncols, nf, mode):
        """The group width (8, 16 or 32; at most mesh.group_width of the
        widest mesh) with the least padded work sum_i nf_i nn_i
        ceil(c_i/w) w; ties go to the wider group.  On the 48-mesh uneven
        batch (8-32 columns per mesh) this is 8, the width the autotuner
        picks."""
        cap = group_width(max(ncols), mode)
        best = None
        for w in (8, 16, 32):
            if w > cap:
                break
            work = sum(f * m.nn * (-(-c // w)) * w for m, c, f in zip(meshes, ncols, nf))
            if best is None or work <= best[0]:
                best = (work, w)
        return best[1]

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.lib().dogpup_solver_destroy(self.h)
        except Exception:
            pass

    def set_method(self, method: str):
        if method not in METHODS:
            raise ValueError(f"unknown method {method!r}; expected one of {sorted(METHODS)}")
        _lib.call("dogpup_solver_set_method", self.h, METHODS[method])
        self.method = method

    def set_tuning(self, tile_rows=0, unroll_a=0, unroll_g=0, graph_k=0):
        _lib.call("dogpup_solver_set_tuning", self.h, int(tile_rows), int(unroll_a), int(unroll_g), int(graph_k))

    def set_cont_stride(self, stride=0):
        """Level-0 spacing of the frequency continuation (a power of two >= 2;
        0: the library's default), dogpup_solver_set_cont_stride."""
        _lib.call("dogpup_solver_set_cont_stride", self.h, int(stride))

    def reset_matrix(self):
        """The next solve redoes the FSAI setup (as after new properties)."""
        self._versions = None

    def ensure_matrix(self):
        """Concatenated K, M and per-mesh frequency scales k*df/c; FSAI setup."""
        versions = [m._matrix_version for m in self.meshes]
        if versions == self._versions:
            return
        K = torch.cat([m.K for m in self.meshes]).contiguous()
        M = torch.cat([m.M for m in self.meshes]).contiguous()
        fs = np.concatenate([np.arange(m.optode.Nf) * m.optode.df / m.c for m in self.meshes])
        fs = torch.as_tensor(fs, dtype=torch.float64, device=DEV)
        _lib.call("dogpup_solver_set_matrix", self.h, _lib.ptr(K), _lib.ptr(M), _lib.ptr(fs), _lib.stream_ptr())
        torch.cuda.current_stream().synchronize()
        self._versions = versions

    def rhs(self) -> torch.Tensor:
        """Point-source right-hand sides of every mesh (s_fpsf spectra) in the
        batch layout."""
        b = torch.zeros(int(self.voff[-1]), dtype=torch.complex128, device=DEV)
        for i, m in enumerate(self.meshes):
            parts = [m._bary_dev(w) for w in self.which]
            nodes = torch.cat([p[0] for p in parts])
            w = torch.cat([p[1] for p in parts])
            spec = torch.as_tensor(m.optode.s_fpsf, dtype=torch.complex128, device=DEV)
            bi = b[int(self.voff[i]):int(self.voff[i + 1])]
            _lib.call("dogpup_point_sources", _lib.ptr(torch.view_as_real(bi)), m.nn, self.nf[i], self.ncols[i],
                      self.nsp, _lib.ptr(nodes), _lib.ptr(w), _lib.ptr(torch.view_as_real(spec)),
                      _lib.stream_ptr())
        return b

    def solve(self, b: torch.Tensor | None = None, tol=1e-8, maxit=1000):
        """One batched solve; returns a FluenceField per mesh (views into one
        device block) with that mesh's iterations, residuals and status."""
        self.ensure_matrix()
        if b is None:
            b = self.rhs()
        x = torch.empty_like(b)
        nout = sum(f * c for f, c in zip(self.nf, self.ncols))
        iters = np.zeros(nout, np.int32)
        relres = np.zeros(nout, np.float64)
        status = np.zeros(nout, np.int32)
        _lib.call("dogpup_solver_solve", self.h, _lib.ptr(torch.view_as_real(b)), _lib.ptr(torch.view_as_real(x)),
                  float(tol), int(maxit), _lib.ptr(iters), _lib.ptr(relres), _lib.ptr(status), _lib.stream_ptr())
        st = np.zeros(4)
        _lib.call("dogpup_solver_stats", self.h, _lib.ptr(st), 4)
        self.stats = {"max_iters": st[0], "group_iters": st[1], "sweeps": st[2], "inner_iters": st[3]}
        out, o = [], 0
        for i, m in enumerate(self.meshes):
            n = self.nf[i] * self.ncols[i]
            shape = (self.nf[i], self.ncols[i])
            data = x[int(self.voff[i]):int(self.voff[i + 1])].view(self.nf[i] * self.ncb[i], m.nn, self.nsp)
            out.append(FluenceField(data, m.nn, self.nf[i], self.ncols[i], self.nsp, iters[o:o + n].reshape(shape),
                                    relres[o:o + n].reshape(shape), status[o:o + n].reshape(shape)))
            o += n
        return out


class SplitBatch:
    """A batch solved as several MeshBatch handles of their own group width,
    one after the other: meshes with many columns in wide groups (each
    matrix entry read once for more RHS), the rest in narrow ones (little
    padding).  On the 48-mesh uneven batch (8-32 sources per mesh), 32-wide
    groups for the 18 meshes with more than 24 sources and 8-wide ones for
    the other 30 measured 606 ms per solve against 655 ms in one 8-wide
    handle.  (The parts on one host thread each took 577 ms against 616 ms
    one after the other on an earlier build, but a capture
    of a part's graph then fails now and then: another thread's device-wide
    synchronisation, such as a handle destroyed by the garbage collector,
    can fall inside it.)

    ``parts``: list of (mesh indices, nsp); default width_parts.  The other
    arguments are MeshBatch's and apply to every part.  solve() returns one
    FluenceField per mesh in the order of ``meshes``.
    """

    def __init__(self, meshes, which="s", mode="mixed", parts=None, method="cocg", tuning=None, cont=True):
        self.meshes = list(meshes)
        if not self.meshes:
            raise ValueError("empty batch")
        self.which, self.mode = which, mode
        self.ncols = [sum(len((m.optode.s_bary if w == "s" else m.optode.d_bary)[0]) for w in which)
                      for m in self.meshes]
        if parts is None:
            parts = self.width_parts(self.ncols)
        seen = sorted(i for idx, _ in parts for i in idx)
        if seen != list(range(len(self.meshes))):
            raise ValueError("parts must cover every mesh exactly once")
        self.parts = [(list(idx), int(nsp)) for idx, nsp in parts if len(idx)]
        self.subs = [MeshBatch([self.meshes[i] for i in idx], which, mode, nsp=nsp, method=method, tuning=tuning,
                               cont=cont) for idx, nsp in self.parts]
        self.method = method

    @staticmethod
    def width_parts(ncols, wide=32, narrow=8, threshold=None):
        """Meshes with more than ``threshold`` columns (default 3/4 of
        ``wide``: at most a quarter of their wide groups is padding) in
        ``wide`` groups, the others in ``narrow`` ones."""
        if threshold is None:
            threshold = 3 * wide // 4
        big = [i for i, c in enumerate(ncols) if c > threshold]
        small = [i for i, c in enumerate(ncols) if c <= threshold]
        return [p for p in ((big, wide), (small, narrow)) if p[0]]

    def set_method(self, method: str):
        for B in self.subs:
            B.set_method(method)
        self.method = method

    def set_tuning(self, tile_rows=0, unroll_a=0, unroll_g=0, graph_k=0):
        for B in self.subs:
            B.set_tuning(tile_rows, unroll_a, unroll_g, graph_k)

    def set_cont_stride(self, stride=0):
        for B in self.subs:
            B.set_cont_stride(stride)

    def reset_matrix(self):
        for B in self.subs:
            B.reset_matrix()

    def ensure_matrix(self):
        for B in self.subs:
            B.ensure_matrix()

    def rhs(self):
        """Point-source right-hand sides, one batch-layout block per part."""
        return [B.rhs() for B in self.subs]

    def solve(self, b=None, tol=1e-8, maxit=1000):
        """The parts' solves one after the other (b: one block per part, as
        rhs() returns, or None); one FluenceField per mesh in the caller's
        order."""
        if b is None:
            b = [None] * len(self.subs)
        if len(b) != len(self.subs):
            raise ValueError("one right-hand-side block per part")
        results = [B.solve(bb, tol=tol, maxit=maxit) for B, bb in zip(self.subs, b)]
        out = [None] * len(self.meshes)
        stats = {"max_iters": 0.0, "group_iters": 0.0, "sweeps": 0.0, "inner_iters": 0.0}
        for (idx, _), B, fields in zip(self.parts, self.subs, results):
            for i, f in zip(idx, fields):
                out[i] = f
            stats["max_iters"] = max(stats["max_iters"], B.stats["max_iters"])
            for k in ("group_iters", "sweeps", "inner_iters"):
                stats[k] += B.stats[k]
        self.stats = stats
        return out
