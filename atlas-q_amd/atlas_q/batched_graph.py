"""Batched circuits with the bond dimensions on the device, replayable from a
HIP graph.

``BatchedMPS`` decides every launch of a sub-layer on the host from the
current bond dimensions, which it learns from the rank results (one
device-to-host copy per sub-layer).  ``PlannedCircuit`` removes those round
trips: the bond dimensions stay on the device, and every buffer, size bucket
and launch shape is fixed once, before the circuit runs, from upper bounds
of the bond dimensions -- a gate's new bond is at most min(cap, 2 chi_left,
2 chi_right), the limit of the rank rule itself, propagated through the
circuit's steps.  With ``schmidt_bound`` the plan also uses r chi_middle,
r the gate's operator Schmidt rank: exact in exact arithmetic and much
tighter for low-rank gates, but a float32 theta has round-off singular
values past that rank which the rank rule at a small eps keeps (as the
host-planned path and the reference do), so there the plan lowers such ranks
to the bound and departs from the host-planned result.  The steps run one
after another on one stream, so their large buffers (theta, its FP64 copy,
the Gram matrices, eigenvectors and projections) are carved from three
arenas sized to the largest step, not to the whole circuit.  The library's
``aq_layer_*`` kernels write the size-dependent descriptor fields from the
device dimensions (theta and Gram before the eigensolve; projection, rebuild
and the tridiagonal buckets' kept ranks after the rank rule), and the
grouped kernels skip the tiles a problem below its bound does not have.  A
circuit is then a fixed sequence of launches: ``run`` executes it,
``capture`` records it into one ``torch.cuda.CUDAGraph`` whose replays
advance the population through the whole circuit.

Per gate the arithmetic is that of ``BatchedMPS.apply_2q_arrays``: theta
with the gate in the epilogue, the FP64 Gram matrix of the side chosen by
the bounds (``BatchedMPS`` chooses by the actual sizes), the Jacobi kernel
for orders up to 64 at each matrix's own order and the tridiagonal path
above (padded buckets, kept vectors only, FP64 orthogonality re-check), the
rank rule (adaptive_mps.py:255-274 semantics) and the rebuild with S
absorbed left.  A bucket row whose eigensolve fails cannot be refilled
inside a graph: the failures are counted on the device and ``finish``
raises if there was any.  A kept rank above the planned bound (possible only
with ``schmidt_bound``) is lowered to the bound by ``aq_layer_post``, and
``finish`` reports the number of such gates as ``clamped``.
"""
from __future__ import annotations

import ctypes
from typing import List

import numpy as np
import torch

from . import _ffi
from ._ffi import COPY_DESC, GEMM_DESC, LAYER_GATE, SITE1Q_DESC, RankResult, check, lib
from .batched import BatchedMPS, _cdiv, _merge_per_site, _prefix, _upload, _h2d
from .eig import JACOBI_MAX, ORTH_TILE, ORTH_TOL, GramTridiag

C64, C128 = torch.complex64, torch.complex128
_RR = ctypes.sizeof(RankResult)


def _vp(x) -> ctypes.c_void_p:
    return ctypes.c_void_p(x if isinstance(x, int) else x.data_ptr())


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptrs_rows(t: torch.Tensor) -> np.ndarray:
    """Addresses of the rows t[b] of a batched tensor."""
    return t.data_ptr() + t.element_size() * t[0].numel() * np.arange(t.shape[0], dtype=np.int64)


class _Grouped:
    """One grouped launch with a fixed work list: descriptors (written on the
    device), the prefix of their bound work counts, the total and the grid."""

    def __init__(self, counts: np.ndarray, dev, grid: int):
        pre, self.total = _prefix(np.asarray(counts, np.int64))
        self.prefix = _h2d(pre, dev)
        self.count = len(counts)
        self.grid = max(1, min(int(grid), 65536))


class _Step1Q:
    def __init__(self, bm: BatchedMPS, dims_ptr: np.ndarray, jj, qq, U, bound):
        dev = bm.dev
        d = np.zeros(len(jj), SITE1Q_DESC)
        p = bm._ptrs(jj, qq)
        d["T_in"] = p
        d["T_out"] = p
        d["chiL"] = bound[jj, qq]
        d["chiR"] = bound[jj, qq + 1]
        d["gate"] = np.ascontiguousarray(U, np.complex128).reshape(len(jj), 4).view(np.float64)
        self.desc = _upload(d, dev)
        self.dims = _h2d(dims_ptr[jj].astype(np.int64), dev)
        self.site = _h2d(qq.astype(np.int64), dev)
        cnt = d["chiL"] * d["chiR"]
        total = int(cnt.sum())
        self.launch = _Grouped(cnt, dev, max(1, min(-(-total // 256), 256 * 16)))

    def run(self):
        L = lib()
        g = self.launch
        check(L.aq_layer_1q(_vp(self.desc), _vp(self.dims), _vp(self.site), g.count, _stream()),
              "aq_layer_1q")
        check(L.aq_grouped_1q(_vp(self.desc), _vp(g.prefix), g.count, g.total, g.grid, _stream()),
              "aq_grouped_1q")


def _schmidt_rank(U4: np.ndarray) -> np.ndarray:
    """Operator Schmidt rank (1, 2 or 4) of two-qubit gates [G, 4, 4] across
    the two sites: a gate is sum_s E_s (x) F_s with that many terms, so it
    multiplies the rank of theta across the bond by at most that factor."""
    U = np.asarray(U4, np.complex128).reshape(-1, 2, 2, 2, 2)      # [g, p, q, p1, p2]
    M = U.transpose(0, 1, 3, 2, 4).reshape(-1, 4, 4)               # [g, (p, p1), (q, p2)]
    s = np.linalg.svd(M, compute_uv=False)
    return (s > 1e-12 * s[:, :1]).sum(1).astype(np.int64)


def _align(n: int) -> int:
    """Element offsets into the shared arenas in 256-byte steps."""
    return -(-int(n) // 16) * 16


class _Step2Q:
    """One two-qubit sub-layer: launch shapes from bounds, large buffers carved
    (by ``bind``) from arenas shared by all steps."""

    def __init__(self, bm: BatchedMPS, dims_ptr: np.ndarray, jj, bb, U4, bound, fails, overflow,
                 schmidt_bound: bool):
        dev = bm.dev
        self.bm, self.fails, self.overflow = bm, fails, overflow
        dl, dm, dr = bound[jj, bb], bound[jj, bb + 1], bound[jj, bb + 2]
        mb, nb = 2 * dl, 2 * dr
        left = mb <= nb
        # left-Gram gates first: each side is one contiguous grouped launch
        perm = np.argsort(~left, kind="stable")
        self.perm = perm
        jj, bb, U4 = jj[perm], bb[perm], U4[perm]
        dl, dm, dr, mb, nb, left = dl[perm], dm[perm], dr[perm], mb[perm], nb[perm], left[perm]
        G = self.G = len(jj)
        self.nleft = int(left.sum())
        sb = np.minimum(mb, nb)
        cap = bm.cap[jj]
        kb = np.minimum(sb, cap)
        if schmidt_bound:
            kb = np.minimum(kb, _schmidt_rank(U4) * dm)
        bound[jj, bb + 1] = kb
        self.kb = kb
        # bound-sized slots: theta (complex64) and the projection output
        # (complex128) share one arena -- theta is dead once its complex128
        # copy is made, before the projection writes
        self.xoff, xtot = _prefix(mb * nb)
        self.poff, ptot = _prefix(np.where(left, kb * nb, mb * kb))
        self.xtot = int(xtot)
        gates = np.zeros(G, LAYER_GATE)
        gates["siteA"] = bm._ptrs(jj, bb)
        gates["siteB"] = bm._ptrs(jj, bb + 1)
        gates["dims"] = dims_ptr[jj]
        gates["bond"] = bb
        gates["left"] = left
        gates["kslot"] = -1
        gates["kbound"] = kb
        gates["gate"] = np.ascontiguousarray(U4, np.complex128).reshape(G, 16).view(np.float64)
        # size buckets of the Gram matrices by their bound order; T and the
        # eigenvectors go to the matrix arena at offsets moff
        bsize = np.maximum(2, 2 ** np.ceil(np.log2(np.maximum(sb, 1))).astype(np.int64))
        self.buckets = []
        ntri, moff = 0, 0
        for sz in np.unique(bsize):
            sz = int(sz)
            idx = np.nonzero(bsize == sz)[0]
            B = len(idx)
            S64 = torch.zeros((B, sz), dtype=torch.float64, device=dev)
            Sinv = torch.zeros_like(S64)
            S32 = torch.zeros((B, sz), dtype=torch.float32, device=dev)
            bk = {"sz": sz, "idx": _h2d(idx.astype(np.int64), dev), "idx_np": idx, "S64": S64,
                  "Sinv": Sinv, "S32": S32, "B": B, "t_off": moff}
            moff = _align(moff + B * sz * sz)
            gates["S64"][idx] = _ptrs_rows(S64)
            gates["Sinv"][idx] = _ptrs_rows(Sinv)
            gates["sz"][idx] = sz
            if sz <= self.bm.jacobi_max:
                bk["W"] = torch.zeros((B, sz), dtype=torch.float64, device=dev)
                bk["orders"] = torch.zeros(B, dtype=torch.int64, device=dev)
                bk["v_off"] = moff
                moff = _align(moff + B * sz * sz)
                gates["vfmt"][idx] = 0
            else:
                kpad = int(-(-int(kb[idx].max()) // ORTH_TILE) * ORTH_TILE)
                bk["kpad"] = kpad
                bk["v_off"] = moff
                bk["v_len"] = B * kpad * sz + ORTH_TILE * sz
                moff = _align(moff + bk["v_len"])
                bk["koff"] = ntri
                gates["vfmt"][idx] = 1
                gates["kslot"][idx] = ntri + np.arange(B)
                ntri += B
            self.buckets.append(bk)
        self._gates = gates
        # arena needs: bytes of the theta/projection arena, complex128
        # elements of the theta copy and of the matrix arena
        self.need = {"xp": max(8 * self.xtot, 16 * int(ptot)), "x64": self.xtot, "mat": moff}
        self.kvec = torch.zeros(max(ntri, 1), dtype=torch.int32, device=dev)
        self.order = torch.zeros(G, dtype=torch.int64, device=dev)
        zd = np.zeros(G, GEMM_DESC)
        self.theta_d = _upload(zd, dev)
        self.gram_d = _upload(zd, dev)
        self.proj_d = _upload(zd, dev)
        self.copy_d = _upload(np.zeros(2 * G, COPY_DESC), dev)
        # rank rule inputs: each gate's float32 singular values (its bucket row)
        sptr = np.zeros(G, np.int64)
        for bk in self.buckets:
            sptr[bk["idx_np"]] = _ptrs_rows(bk["S32"])
        self.sptr = _h2d(sptr, dev)
        self.caps = _h2d(cap.astype(np.int64), dev)
        self.rr = torch.zeros(G * _RR, dtype=torch.uint8, device=dev)
        self.log_k = torch.zeros(G, dtype=torch.int64, device=dev)
        self.log_eps = torch.zeros(G, dtype=torch.float64, device=dev)
        self.log_ent = torch.zeros(G, dtype=torch.float64, device=dev)
        # launch shapes from the bounds
        tun = bm.tuner
        nl = self.nleft
        dsz = GEMM_DESC.itemsize
        tiles = _cdiv(mb, 64) * _cdiv(dr, 64)
        self.theta_l = _Grouped(tiles, dev, tun.heuristic("theta", int(tiles.sum())))
        gt = _cdiv(sb, 64) * _cdiv(sb, 128)
        self.gram_l = [(0, _Grouped(gt[:nl], dev, tun.heuristic("gram", int(gt[:nl].sum()))), 1 | 4),
                       (nl, _Grouped(gt[nl:], dev, tun.heuristic("gram", int(gt[nl:].sum()))), 2 | 4)]
        pt = np.where(left, _cdiv(kb, 64) * _cdiv(nb, 128), _cdiv(mb, 64) * _cdiv(kb, 128))
        self.proj_l = [(0, _Grouped(pt[:nl], dev, tun.heuristic("proj", int(pt[:nl].sum()))), 2 | 4),
                       (nl, _Grouped(pt[nl:], dev, tun.heuristic("proj", int(pt[nl:].sum()))), 3 | 4)]
        cc = np.concatenate([mb * kb, kb * nb])
        self.copy_l = _Grouped(cc, dev, max(1, min(-(-int(cc.sum()) // 256), 256 * 16)))
        self._dsz = dsz

    def bind(self, xp: torch.Tensor, x64: torch.Tensor, mat: torch.Tensor):
        """Carve this step's buffers from the shared arenas and upload the
        gate records that point into them."""
        gates = self._gates
        self.X = xp.view(C64)[:max(self.xtot, 1)]
        self.X64 = x64[:max(self.xtot, 1)]
        gates["X"] = xp.data_ptr() + 8 * self.xoff
        gates["X64"] = x64.data_ptr() + 16 * self.xoff
        gates["P"] = xp.data_ptr() + 16 * self.poff
        for bk in self.buckets:
            sz, B, idx = bk["sz"], bk["B"], bk["idx_np"]
            bk["T"] = mat[bk["t_off"]: bk["t_off"] + B * sz * sz].view(B, sz, sz)
            gates["G"][idx] = _ptrs_rows(bk["T"])
            if sz <= self.bm.jacobi_max:
                bk["V"] = mat[bk["v_off"]: bk["v_off"] + B * sz * sz].view(B, sz, sz)
                gates["V"][idx] = _ptrs_rows(bk["V"])
            else:
                bk["Et"] = mat[bk["v_off"]: bk["v_off"] + bk["v_len"]]
                gates["V"][idx] = bk["Et"].data_ptr() + 16 * bk["kpad"] * sz * np.arange(B)
        self.gates = _upload(gates, self.bm.dev)
        self._gates = None

    def _grouped_gemm(self, dtype, descs: torch.Tensor, first: int, g: _Grouped, layout: int):
        if g.count == 0:
            return
        check(lib().aq_grouped_gemm(dtype, _vp(descs.data_ptr() + first * self._dsz),
                                    _vp(g.prefix), g.count, g.total, layout, g.grid, _stream()),
              "aq_grouped_gemm")

    def run(self):
        L = lib()
        G, st = self.G, _stream()
        check(L.aq_layer_pre(_vp(self.gates), G, _vp(self.theta_d), _vp(self.gram_d),
                             _vp(self.order), st), "aq_layer_pre")
        self._grouped_gemm(_ffi.AQ_C64, self.theta_d, 0, self.theta_l, 3)
        self.X64.copy_(self.X)
        for first, g, layout in self.gram_l:
            self._grouped_gemm(_ffi.AQ_C128, self.gram_d, first, g, layout)
        tris = []
        for bk in self.buckets:
            sz, B, T = bk["sz"], bk["B"], bk["T"]
            if sz <= self.bm.jacobi_max:
                torch.index_select(self.order, 0, bk["idx"], out=bk["orders"])
                check(L.aq_heev_small_orders(sz, _vp(bk["orders"]), _vp(T), sz, sz * sz,
                                             _vp(bk["W"]), _vp(bk["V"]), sz, sz * sz, B, 30,
                                             1e-15, None, st), "aq_heev_small_orders")
                W = bk["W"]
                bad = ~(torch.isfinite(W).all(1)
                        & torch.isfinite(torch.view_as_real(bk["V"])).reshape(B, -1).all(1))
                self.fails.add_(bad.sum())
            else:
                check(L.aq_layer_pad(_vp(T), sz, B, _vp(bk["idx"]), _vp(self.order), st),
                      "aq_layer_pad")
                tri = GramTridiag(T)
                W = tri.eigenvalues()
                tris.append((bk, tri))
            torch.sqrt(torch.clamp(W, min=0.0), out=bk["S64"])
            torch.where(bk["S64"] > 0, 1.0 / bk["S64"], torch.zeros_like(bk["S64"]),
                        out=bk["Sinv"])
            bk["S32"].copy_(bk["S64"])
        check(L.aq_rank_select_batched(_ffi.AQ_C64, _vp(self.sptr), _vp(self.order),
                                       _vp(self.caps), self.bm.eps, G, _vp(self.rr), st),
              "aq_rank_select_batched")
        check(L.aq_layer_post(_vp(self.gates), G, _vp(self.rr), _vp(self.proj_d),
                              _vp(self.copy_d), _vp(self.kvec), _vp(self.log_k),
                              _vp(self.log_eps), _vp(self.log_ent), _vp(self.overflow), st),
              "aq_layer_post")
        for bk, tri in tris:
            kv = self.kvec[bk["koff"]: bk["koff"] + bk["B"]]
            _, err = tri.vectors(kvec=kv, kpad=bk["kpad"], out=bk["Et"])
            self.fails.add_((~(err <= ORTH_TOL)).sum())
        for first, g, layout in self.proj_l:
            self._grouped_gemm(_ffi.AQ_C128, self.proj_d, first, g, layout)
        c = self.copy_l
        check(L.aq_grouped_copy(_vp(self.copy_d), _vp(c.prefix), c.count, c.total, c.grid, st),
              "aq_grouped_copy")


class PlannedCircuit:
    """``BatchedMPS.compile_circuits`` output run with the bond dimensions on
    the device (see the module docstring).  ``run()`` launches the whole
    circuit on the current stream, ``capture()`` records it into a HIP graph
    (replays start from the state the population is in), ``finish()``
    synchronises, copies the bond dimensions and the per-gate statistics back
    into the ``BatchedMPS`` and raises if an eigensolve failed."""

    def __init__(self, bm: BatchedMPS, compiled, schmidt_bound: bool = False):
        self.bm = bm
        dev = bm.dev
        self.dims = torch.from_numpy(np.ascontiguousarray(bm.dims, np.int64)).to(dev)
        row = bm.dims.shape[1]
        dims_ptr = self.dims.data_ptr() + 8 * row * np.arange(bm.M, dtype=np.int64)
        self.fails = torch.zeros((), dtype=torch.int64, device=dev)
        self.overflow = torch.zeros((), dtype=torch.int64, device=dev)
        bound = bm.dims.copy()
        self.steps: List = []
        self.gate_order = []           # per 2q step: (mps, bond) in the eager path's order
        nsteps = max((len(c) for c in compiled), default=0)
        for t in range(nsteps):
            live = [(j, c[t]) for j, c in enumerate(compiled) if t < len(c)]
            j1 = np.concatenate([np.full(len(st[0]), j, np.int64) for j, st in live])
            if len(j1):
                q1 = np.concatenate([st[0] for _, st in live])
                u1 = np.concatenate([st[1] for _, st in live])
                jm, qm, um = _merge_per_site(j1, q1, u1, bm.nmax)
                self.steps.append(_Step1Q(bm, dims_ptr, jm, qm, um, bound))
            j2 = np.concatenate([np.full(len(st[2]), j, np.int64) for j, st in live])
            if len(j2):
                b2 = np.concatenate([st[2] for _, st in live])
                u2 = np.concatenate([st[3] for _, st in live])
                step = _Step2Q(bm, dims_ptr, j2, b2, u2, bound, self.fails, self.overflow,
                               schmidt_bound)
                self.steps.append(step)
                self.gate_order.append((step, j2, b2))
        self.bound = bound
        # the steps run one after another on one stream: their large buffers
        # share three arenas sized to the largest step
        two = [s for s, _, _ in self.gate_order]
        need = {k: max([s.need[k] for s in two], default=1) for k in ("xp", "x64", "mat")}
        self.xp = torch.zeros(max(1, -(-need["xp"] // 16)), dtype=C128, device=dev)
        self.x64 = torch.zeros(max(1, need["x64"]), dtype=C128, device=dev)
        self.mat = torch.zeros(max(1, need["mat"]), dtype=C128, device=dev)
        for s in two:
            s.bind(self.xp, self.x64, self.mat)
        self.arena_bytes = 16 * (self.xp.numel() + self.x64.numel() + self.mat.numel())
        self.clamped = 0
        torch.cuda.current_stream(dev).synchronize()

    def run(self):
        old = lib().aq_hetrd_max_groups(1)     # the reduction on this stream only
        try:
            for s in self.steps:
                s.run()
        finally:
            lib().aq_hetrd_max_groups(old)

    def capture(self) -> torch.cuda.CUDAGraph:
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        return g

    def finish(self):
        torch.cuda.current_stream(self.bm.dev).synchronize()
        nf = int(self.fails.item())
        self.clamped = int(self.overflow.item())
        self.bm.dims[:] = self.dims.cpu().numpy()
        logs = self.bm.logs
        for step, j2, b2 in self.gate_order:
            inv = np.empty(step.G, np.int64)
            inv[step.perm] = np.arange(step.G)
            k = step.log_k.cpu().numpy()[inv]
            e = step.log_eps.cpu().numpy()[inv]
            s = step.log_ent.cpu().numpy()[inv]
            logs["mps"].extend(j2.tolist())
            logs["bond"].extend(b2.tolist())
            logs["k"].extend(k.tolist())
            logs["eps_local"].extend(e.tolist())
            logs["entropy"].extend(s.tolist())
        if nf:
            raise RuntimeError(f"PlannedCircuit: {nf} eigensolves failed (non-finite vectors or "
                               "the orthogonality re-check); rerun with BatchedMPS.run_compiled")
