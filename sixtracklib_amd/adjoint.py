"""Differentiable tracking behind torch.autograd (PyTorch-ROCm).

No counterpart in the reference (its tracker has no derivatives).  The
loss of n particles tracked n_turns turns through a lattice of Drift,
Multipole (no curvature), Cavity, XYShift, SRotation and aperture elements,

    loss = sum_i sum_t sum_c  w_c (X_ic(t) - target_c(t))^2,

X_i(t) = (x, px, y, py, zeta, delta) of particle i at the end of turn t, is
computed by a hand-written HIP forward kernel that stores the lane state
every K elements, and its gradient with respect to the initial coordinates
and every multipole's knl / ksl by hand-written reverse kernels that
recompute each K-element segment from its checkpoint
(include/sixtracklib_amd/adjoint.h).  PyTorch provides device memory, the
stream and the autograd plumbing; no PyTorch operator runs on the tracking
path.

    tr = AdjointTracker(lattice, n_turns=10, species=particles)
    coords = torch.tensor(..., dtype=torch.float64, device="cuda", requires_grad=True)
    knl, ksl = tr.knl0.clone().requires_grad_(), tr.ksl0.clone().requires_grad_()
    loss = tr.loss(coords, knl, ksl, target)
    loss.backward()          # coords.grad, knl.grad, ksl.grad
"""
from __future__ import annotations

import ctypes as ct
import math

import numpy as np
import torch

from . import stcommon as st
from .particles import particles_in

# opcodes of lattice.h the differentiable tracker accepts
OP_DRIFT, OP_MULTIPOLE = 0x01, 0x02
OP_CAVITY, OP_XYSHIFT, OP_SROTATION = 0x82, 0x83, 0x84
OP_LIMITS = (0x85, 0x86, 0x87)
ACCEPTED = {OP_DRIFT, OP_MULTIPOLE, OP_CAVITY, OP_XYSHIFT, OP_SROTATION} | set(OP_LIMITS)
MP_CURV_FLAG = 0x8000
MAX_ORDER = 10
STRIDE = 2 * MAX_ORDER + 2


class AdjProblem(ct.Structure):
    """st_adj_problem_t (include/sixtracklib_amd/adjoint.h)"""
    _fields_ = [("lat", ct.c_void_p), ("rec_off", ct.c_void_p), ("mp_index", ct.c_void_p),
                ("n_elem", ct.c_int64), ("K", ct.c_int64), ("n_turns", ct.c_int64),
                ("n", ct.c_int64),
                ("q0", ct.c_double), ("mass0", ct.c_double), ("beta0", ct.c_double),
                ("p0c", ct.c_double), ("chi", ct.c_double), ("charge_ratio", ct.c_double),
                ("weights", ct.c_double * 6), ("target", ct.c_void_p), ("ckpt", ct.c_void_p)]


_lib = st.sixtracklib
_lib.st_Adjoint_checkpoint_doubles.restype = ct.c_int64
_lib.st_Adjoint_checkpoint_doubles.argtypes = [ct.POINTER(AdjProblem)]
_lib.st_Adjoint_segment_doubles.restype = ct.c_int64
_lib.st_Adjoint_segment_doubles.argtypes = [ct.POINTER(AdjProblem)]
_lib.st_Adjoint_forward.restype = ct.c_int
_lib.st_Adjoint_forward.argtypes = [ct.POINTER(AdjProblem), ct.c_int, ct.c_void_p, ct.c_void_p,
                                    ct.c_void_p, ct.c_void_p]
_lib.st_Adjoint_backward_turn.restype = ct.c_int
_lib.st_Adjoint_backward_turn.argtypes = [ct.POINTER(AdjProblem), ct.c_int, ct.c_int64,
                                          ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                          ct.c_void_p]
_lib.st_Adjoint_reduce_coefficients.restype = ct.c_int
_lib.st_Adjoint_reduce_coefficients.argtypes = [ct.POINTER(AdjProblem), ct.c_int, ct.c_void_p,
                                                ct.c_int64, ct.c_void_p, ct.c_void_p,
                                                ct.c_int64, ct.c_void_p]


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, st.last_error()))


def _stream():
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


def _pack(ebuf):
    n_el = ct.c_int64(0)
    nw = st.st_Lattice_pack(ebuf.pointer, None, 0, None, ct.byref(n_el), None, None)
    if nw < 0:
        raise ValueError("lattice could not be packed")
    words = np.zeros(nw + 32, dtype=np.uint64)       # zero padding read ahead by kernels
    offs = np.zeros(n_el.value + 1, dtype=np.uint32)
    if st.st_Lattice_pack(ebuf.pointer, words.ctypes.data, nw, offs.ctypes.data,
                          ct.byref(n_el), None, None) != nw:
        raise ValueError("lattice could not be packed")
    return words, offs[:-1].copy()


class AdjointTracker:
    """Differentiable tracking of one lattice for ``n_turns`` turns.

    species: a Particles object or particle CBuffer whose first particle
    gives q0, mass0, beta0, p0c, chi, charge_ratio (the same for all
    tracked particles).  ``weights`` are the w_c of the loss (default: the
    four transverse coordinates).  K is the checkpoint interval.
    """

    def __init__(self, lattice, n_turns, species, K=64, fma=False,
                 weights=(1.0, 1.0, 1.0, 1.0, 0.0, 0.0), device="cuda"):
        self.device = torch.device(device)
        self.n_turns = int(n_turns)
        self.K = int(K)
        self.fma = 1 if fma else 0
        self.weights = tuple(float(w) for w in weights)
        p = species if hasattr(species, "p0c") else particles_in(species)
        self.species = {k: float(np.asarray(getattr(p, k))[0])
                        for k in ("q0", "mass0", "beta0", "p0c", "chi", "charge_ratio")}
        words, offs = _pack(lattice)
        ops = (words[offs] & np.uint64(0xff)).astype(np.int64)
        aux = ((words[offs] >> np.uint64(8)) & np.uint64(0xffff)).astype(np.int64)
        bad = sorted(set(ops.tolist()) - ACCEPTED)
        if bad:
            raise ValueError("differentiable tracking supports Drift, Multipole, Cavity, "
                             "XYShift, SRotation and apertures (opcodes found: %s)"
                             % ", ".join("0x%x" % b for b in bad))
        is_mp = ops == OP_MULTIPOLE
        if np.any(aux[is_mp] & MP_CURV_FLAG):
            raise ValueError("multipoles with hxl / hyl (curvature) are not differentiated")
        orders = aux[is_mp] & 0x7fff
        if orders.size and orders.max() > MAX_ORDER:
            raise ValueError("multipole order above %d" % MAX_ORDER)
        self.n_elem = len(offs)
        self.n_mp = int(is_mp.sum())
        self.max_order = int(orders.max()) if orders.size else 0
        mp_index = np.full(self.n_elem, -1, dtype=np.int32)
        mp_index[is_mp] = np.arange(self.n_mp, dtype=np.int32)
        no = self.max_order + 1
        # word positions of bal[2n] / bal[2n+1] for n <= order (padded slots masked)
        mask = np.zeros((self.n_mp, no), dtype=bool)
        pos_re = np.zeros((self.n_mp, no), dtype=np.int64)
        for m, (off, o) in enumerate(zip(offs[is_mp], orders)):
            mask[m, :o + 1] = True
            pos_re[m, :o + 1] = int(off) + 2 + 2 * np.arange(o + 1)
        wf = words.view(np.float64)
        inv_f = np.array([1.0 / math.factorial(k) for k in range(no)])
        fact = np.array([float(math.factorial(k)) for k in range(no)])
        knl0 = np.where(mask, wf[np.where(mask, pos_re, 0)] * fact, 0.0)
        ksl0 = np.where(mask, wf[np.where(mask, pos_re + 1, 0)] * fact, 0.0)
        dev = self.device
        self.orders = torch.from_numpy(orders.astype(np.int32)).to(dev)
        self.mask = torch.from_numpy(mask).to(dev)
        self.lat = torch.from_numpy(words.view(np.int64).copy()).to(dev)
        self.rec_off = torch.from_numpy(offs.astype(np.int32)).to(dev)
        self.mp_index = torch.from_numpy(mp_index).to(dev)
        self._pos_re = torch.from_numpy(pos_re[mask]).to(dev)
        self._inv_f = torch.from_numpy(inv_f).to(dev)
        # the lattice's own bal words: written back when the default
        # coefficients are used, so that knl0 / ksl0 track bit for bit like
        # the lattice (knl0 = bal * n! does not always round-trip)
        self._bal0_re = torch.from_numpy(wf[pos_re[mask]].copy()).to(dev)
        self._bal0_im = torch.from_numpy(wf[pos_re[mask] + 1].copy()).to(dev)
        self.knl0 = torch.from_numpy(knl0).to(dev)
        self.ksl0 = torch.from_numpy(ksl0).to(dev)

    # -- coefficients --------------------------------------------------------
    def write_coefficients(self, knl, ksl):
        """bal[2n] = knl[n] * (1/n!), bal[2n+1] = ksl[n] * (1/n!) into the
        packed lattice (be_multipole.h:437-469)."""
        lat = self.lat.view(torch.float64)
        if knl is self.knl0 and ksl is self.ksl0:
            lat[self._pos_re] = self._bal0_re
            lat[self._pos_re + 1] = self._bal0_im
            return
        lat[self._pos_re] = (knl * self._inv_f)[self.mask]
        lat[self._pos_re + 1] = (ksl * self._inv_f)[self.mask]

    def problem(self, n, target, ckpt):
        pr = AdjProblem()
        pr.lat = self.lat.data_ptr()
        pr.rec_off = self.rec_off.data_ptr()
        pr.mp_index = self.mp_index.data_ptr()
        pr.n_elem, pr.K, pr.n_turns, pr.n = self.n_elem, self.K, self.n_turns, int(n)
        for k, v in self.species.items():
            setattr(pr, k, v)
        for c in range(6):
            pr.weights[c] = self.weights[c]
        pr.target = target.data_ptr()
        pr.ckpt = 0 if ckpt is None else ckpt.data_ptr()
        return pr

    def checkpoint_size(self, n):
        pr = self.problem(n, self.knl0, None)
        return int(_lib.st_Adjoint_checkpoint_doubles(ct.byref(pr)))

    def target_tensor(self, target):
        t = torch.as_tensor(target, dtype=torch.float64, device=self.device)
        if t.shape == (self.n_turns, 4):
            t = torch.cat([t, torch.zeros(self.n_turns, 2, dtype=t.dtype, device=t.device)], 1)
        if t.shape != (self.n_turns, 6):
            raise ValueError("target must have shape (n_turns, 4) or (n_turns, 6)")
        return t.contiguous()

    def loss(self, coords, knl=None, ksl=None, target=None):
        """Scalar loss; differentiable in coords (n, 6), knl and ksl
        (n_multipoles, max_order + 1; entries above a multipole's order are
        ignored and get zero gradient)."""
        knl = self.knl0 if knl is None else knl
        ksl = self.ksl0 if ksl is None else ksl
        if target is None:
            target = torch.zeros(self.n_turns, 6, dtype=torch.float64, device=self.device)
        return _TrackLoss.apply(coords, knl, ksl, self.target_tensor(target), self)

    def per_particle_loss(self, coords, knl=None, ksl=None, target=None):
        """Forward pass only: the loss of every particle (no autograd)."""
        knl = self.knl0 if knl is None else knl
        ksl = self.ksl0 if ksl is None else ksl
        tgt = self.target_tensor(torch.zeros(self.n_turns, 6) if target is None else target)
        if not (knl is self.knl0 and ksl is self.ksl0):
            knl, ksl = knl.detach(), ksl.detach()
        loss, _ = _forward(self, coords.detach(), knl, ksl, tgt)
        return loss


def _forward(tr, coords, knl, ksl, tgt):
    n = coords.shape[0]
    tr.write_coefficients(knl, ksl)
    X = coords.to(torch.float64).t().contiguous()
    ckpt = torch.empty(max(tr.checkpoint_size(n), 1), dtype=torch.float64, device=tr.device)
    loss = torch.empty(n, dtype=torch.float64, device=tr.device)
    status = torch.zeros(1, dtype=torch.int32, device=tr.device)
    pr = tr.problem(n, tgt, ckpt)
    _check(_lib.st_Adjoint_forward(ct.byref(pr), tr.fma, X.data_ptr(), loss.data_ptr(),
                                   status.data_ptr(), _stream()), "adjoint forward")
    if int(status.item()) != 0:
        raise RuntimeError("a particle reached an aperture: the loss is not differentiable there")
    return loss, (X, ckpt)


class _TrackLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords, knl, ksl, tgt, tr):
        default = knl is tr.knl0 and ksl is tr.ksl0
        kn, ks = (tr.knl0, tr.ksl0) if default else (knl.detach(), ksl.detach())
        loss, (X, ckpt) = _forward(tr, coords.detach(), kn, ks, tgt)
        ctx.tr = tr
        ctx.n = coords.shape[0]
        ctx.default = default
        ctx.save_for_backward(kn, ks, tgt, ckpt)
        return loss.sum()

    @staticmethod
    def backward(ctx, grad_out):
        tr = ctx.tr
        knl, ksl, tgt, ckpt = ctx.saved_tensors
        n = ctx.n
        dev = tr.device
        # the lattice may have been rewritten since the forward pass
        tr.write_coefficients(*((tr.knl0, tr.ksl0) if ctx.default else (knl, ksl)))
        pr = tr.problem(n, tgt, ckpt)
        seg = torch.empty(max(int(_lib.st_Adjoint_segment_doubles(ct.byref(pr))), 1),
                          dtype=torch.float64, device=dev)
        coef = torch.empty(max(tr.n_mp * 4 * n, 1), dtype=torch.float64, device=dev)
        cot = torch.zeros(6, n, dtype=torch.float64, device=dev)
        gbal = torch.zeros(max(tr.n_mp, 1), STRIDE, dtype=torch.float64, device=dev)
        g = grad_out.to(torch.float64).reshape(1).contiguous()
        s = _stream()
        for turn in range(tr.n_turns - 1, -1, -1):
            _check(_lib.st_Adjoint_backward_turn(ct.byref(pr), tr.fma, turn, g.data_ptr(),
                                                 seg.data_ptr(), coef.data_ptr(),
                                                 cot.data_ptr(), s), "adjoint backward")
            if tr.n_mp:
                _check(_lib.st_Adjoint_reduce_coefficients(
                    ct.byref(pr), tr.fma, tr.orders.data_ptr(), tr.n_mp, coef.data_ptr(),
                    gbal.data_ptr(), STRIDE, s), "adjoint reduce")
        no = tr.max_order + 1
        gbal = gbal[:tr.n_mp]
        grad_knl = torch.where(tr.mask, gbal[:, 0:2 * no:2] * tr._inv_f, 0.0)
        grad_ksl = torch.where(tr.mask, gbal[:, 1:2 * no:2] * tr._inv_f, 0.0)
        return cot.t(), grad_knl, grad_ksl, None, None
