"""Differentiable tracking (AdjointTracker): forward loss against the plain
tracker, gradients against central finite differences, the one-turn
Jacobian from the adjoint against the finite-difference Jacobian of the
forward tracker, checkpoint-interval independence.

The reference has no derivatives; the forward pass is pinned to the
reference-equivalent tracker (TrackJob, exact build, itself bit-exact to
the oracle) and the gradients to the mathematics (finite differences,
Jacobian).  Tolerances are written next to each check.
"""
import math

import numpy as np
import pytest

from sixtracklib_amd import lattices
from sixtracklib_amd.cobjects import CBuffer
from sixtracklib_amd.particles import particles_in

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SIG = dict(sigma_x=3e-4, sigma_px=3e-6, sigma_y=3e-4, sigma_py=3e-6, sigma_zeta=0.05,
           sigma_delta=1e-4)
FIELDS = ("x", "px", "y", "py", "zeta", "delta")


def _beam(n, seed=7):
    pb, p = lattices.gaussian_beam(n, seed=seed, **SIG)
    coords = np.stack([np.array(getattr(p, f)) for f in FIELDS], 1)
    return pb, p, coords


def _target(turns, seed=11):
    return np.random.default_rng(seed).normal(0.0, [3e-5, 3e-7, 3e-5, 3e-7], (turns, 4))


def _tracker(eb, turns, p, **kw):
    from sixtracklib_amd.adjoint import AdjointTracker
    return AdjointTracker(eb, turns, species=p, **kw)


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


# Finite-difference steps.  The reference formulas resolve some
# dependencies only coarsely (zeta += L (rvv - (1 + (xp^2 + yp^2) / 2)) sees
# px through xp^2 ~ 1e-11 added to 1; the cavity forms delta from
# sqrt(...) - 1), so central differences with steps of 1e-7 of a
# coordinate are roundoff-dominated (1e-3 .. 1 relative error on the zeta
# and delta rows).  Steps of 1e-4 of the beam size sigma_c with Richardson
# extrapolation (4 D(h/2) - D(h)) / 3 make truncation O(h^4) and keep
# roundoff below 1e-9.
BEAM_SIGMA = np.array([3e-4, 3e-6, 3e-4, 3e-6, 0.05, 1e-4])
FD_REL = 1e-4


def _richardson(f, h):
    """f(step) -> value; returns (4 D(h/2) - D(h)) / 3, D central differences."""
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4 * d2 - d1) / 3


def test_forward_loss_matches_plain_tracker():
    """Per-particle loss from the adjoint forward pass vs the same loss from
    TrackJob states collected after every turn (exact build: the states are
    the same bits; the loss is summed in the same order -> 1e-15)."""
    from sixtracklib_amd.trackjob import TrackJob
    eb = lattices.adjoint_sublattice()
    turns = 3
    pb, p, coords = _beam(512)
    tgt = _target(turns)
    tr = _tracker(eb, turns, p)
    got = tr.per_particle_loss(_dev(coords), target=tgt).cpu().numpy()
    job = TrackJob(eb, pb, config_str="fma=0")
    want = np.zeros(len(coords))
    for t in range(turns):
        job.track_until(t + 1)
        job.collect_particles()
        q = particles_in(pb)
        for c, f in enumerate(FIELDS[:4]):
            d = np.array(getattr(q, f)) - tgt[t, c]
            want += d * d
    assert np.all(np.array(particles_in(pb).state) == 1)
    np.testing.assert_allclose(got, want, rtol=1e-15, atol=0)


def test_coordinate_gradients_match_central_differences():
    """d loss_i / d X0_i from the adjoint vs Richardson-extrapolated central
    differences of the forward loss (steps 1e-4 sigma_c); agreement within
    1e-8 relative (floor: 1e-8 of the particle's largest component)."""
    eb = lattices.adjoint_sublattice()
    turns = 4
    _, p, coords = _beam(256)
    tgt = _target(turns)
    tr = _tracker(eb, turns, p)
    X = _dev(coords).requires_grad_()
    loss = tr.loss(X, target=tgt)
    loss.backward()
    g_adj = X.grad.cpu().numpy()
    pick = [0, 17, 100, 255]
    steps = (1.0, -1.0, 0.5, -0.5)
    pert = []
    for i in pick:
        for c in range(6):
            for st in steps:
                v = coords[i].copy()
                v[c] += st * FD_REL * BEAM_SIGMA[c]
                pert.append(v)
    lp = tr.per_particle_loss(_dev(np.array(pert)), target=tgt).cpu().numpy().reshape(-1, 4)
    h = np.tile(FD_REL * BEAM_SIGMA, len(pick))
    d1 = (lp[:, 0] - lp[:, 1]) / (2 * h)
    d2 = (lp[:, 2] - lp[:, 3]) / h
    g_fd = ((4 * d2 - d1) / 3).reshape(len(pick), 6)
    ga = g_adj[pick]
    scale = np.max(np.abs(g_fd), axis=1, keepdims=True)
    err = np.abs(ga - g_fd) / (np.abs(g_fd) + scale)
    assert np.max(err) < 1e-8, (np.max(err), ga, g_fd)


def test_multipole_gradients_match_central_differences():
    """d loss / d knl, ksl for a sample of coefficients (every order) vs
    Richardson-extrapolated central differences (steps 1e-4 of the
    coefficient, or of its error scale when smaller); within 1e-8 relative
    plus the finite differences' own noise floor, 8e-16 |loss| / h)."""
    eb = lattices.adjoint_sublattice()
    turns = 3
    _, p, coords = _beam(512)
    tgt = _target(turns)
    tr = _tracker(eb, turns, p)
    knl = tr.knl0.clone().requires_grad_()
    ksl = tr.ksl0.clone().requires_grad_()
    loss = tr.loss(_dev(coords), knl, ksl, tgt)
    loss.backward()
    L = float(loss.detach())
    gk, gs = knl.grad.cpu().numpy(), ksl.grad.cpu().numpy()
    mask = tr.mask.cpu().numpy()
    assert np.all(gk[~mask] == 0) and np.all(gs[~mask] == 0)
    rng = np.random.default_rng(3)
    valid = np.argwhere(mask)
    sample = valid[rng.choice(len(valid), 12, replace=False)]
    k0, s0 = tr.knl0.cpu().numpy(), tr.ksl0.cpu().numpy()
    big = max(np.max(np.abs(gk)), np.max(np.abs(gs)))
    X = _dev(coords)
    for (m, n) in sample:
        err_scale = 1e-4 * 0.0207 * math.factorial(int(n)) * 0.017 ** (1.0 - n)
        for which, base, g in (("knl", k0, gk), ("ksl", s0, gs)):
            h = FD_REL * max(abs(base[m, n]), err_scale)

            def f(step):
                kk, ss = k0.copy(), s0.copy()
                (kk if which == "knl" else ss)[m, n] += step
                return float(tr.per_particle_loss(X, _dev(kk), _dev(ss), tgt).sum())
            fd = _richardson(f, h)
            # FD cannot resolve below its own noise: loss values carry
            # ~1e-15 relative rounding, i.e. ~8e-16 |L| / h on the quotient
            noise = 8 * 2.2e-16 * abs(L) / h
            assert abs(g[m, n] - fd) <= 1e-8 * abs(fd) + noise + 1e-12 * big, \
                (which, m, n, g[m, n], fd, noise)


def test_one_turn_jacobian_matches_forward_differences():
    """6x6 one-turn Jacobian J: from the adjoint, J^T e_c = (g(T) - g(T +
    d e_c)) / (2 d) with unit weights (exact for any d: the gradient is
    linear in the target); from the forward tracker (TrackJob, exact build)
    by Richardson-extrapolated central differences.  Steps: 1e-4 sigma_c for
    the x, px, y, py rows; 1e-2 sigma_c for the zeta and delta rows, whose
    reference formulas resolve small changes only to ~1e-14 absolute (the
    drift adds (xp^2 + yp^2) / 2 ~ 1e-11 to 1; the cavity's sqrt(...) - 1).
    Entries agree within 1e-8 of the largest entry of their row (zeta row:
    3e-6, see below)."""
    from sixtracklib_amd.trackjob import TrackJob
    eb = lattices.adjoint_sublattice()
    _, p, coords = _beam(1, seed=21)
    tr = _tracker(eb, 1, p, weights=(1.0,) * 6)
    x0 = coords[0]

    def grad_for(target):
        X = _dev(x0[None, :]).requires_grad_()
        tr.loss(X, target=target[None, :]).backward()
        return X.grad.cpu().numpy()[0]

    t0 = np.zeros(6)
    g0 = grad_for(t0)
    d = 1e-3
    J_adj = np.zeros((6, 6))
    for c in range(6):
        e = np.zeros(6)
        e[c] = d
        J_adj[c] = (g0 - grad_for(t0 + e)) / (2 * d)
    # forward differences of the tracked state
    def fd_jacobian(rel):
        steps = (1.0, -1.0, 0.5, -0.5)
        starts = []
        for c in range(6):
            for st in steps:
                v = x0.copy()
                v[c] += st * rel * BEAM_SIGMA[c]
                starts.append(v)
        starts = np.array(starts)
        pb, _ = lattices.make_particles(len(starts),
                                        **{f: starts[:, k] for k, f in enumerate(FIELDS)},
                                        p0c=float(p.p0c[0]), mass0=float(p.mass0[0]))
        job = TrackJob(eb, pb, config_str="fma=0")
        job.track_until(1)
        job.collect_particles()
        q = particles_in(pb)
        end = np.stack([np.array(getattr(q, f)) for f in FIELDS], 1).reshape(6, 4, 6)
        h = (rel * BEAM_SIGMA)[:, None]
        d1 = (end[:, 0] - end[:, 1]) / (2 * h)
        d2 = (end[:, 2] - end[:, 3]) / h
        return ((4 * d2 - d1) / 3).T                 # [output][input]
    J_fd = np.vstack([fd_jacobian(FD_REL)[:4], fd_jacobian(1e-2)[4:]])
    row_scale = np.max(np.abs(J_fd), axis=1, keepdims=True)
    err = np.abs(J_adj - J_fd) / row_scale
    # x, px, y, py and delta rows: 1e-8.  zeta row: the reference drift adds
    # (xp^2 + yp^2) / 2 ~ 1e-11 to 1 before forming dzeta, so zeta resolves
    # its dependence on the transverse coordinates only to ~1e-14 absolute
    # per drift; no difference quotient of the tracked zeta gets below
    # ~1e-6 of the row scale (observed 1.0e-6 at 1e-2 sigma steps).
    assert np.max(err[[0, 1, 2, 3, 5]]) < 1e-8, (err, J_adj, J_fd)
    assert np.max(err[4]) < 3e-6, (err, J_adj, J_fd)


def test_gradients_independent_of_checkpoint_interval():
    """The reverse pass recomputes the same states whatever K is: the
    gradients are bit-identical for K = 64, 7 and one segment per turn."""
    eb = lattices.adjoint_sublattice()
    turns = 2
    _, p, coords = _beam(128)
    tgt = _target(turns)
    grads = []
    for K in (64, 7, 1000):
        tr = _tracker(eb, turns, p, K=K)
        X = _dev(coords).requires_grad_()
        knl = tr.knl0.clone().requires_grad_()
        tr.loss(X, knl, tr.ksl0, tgt).backward()
        grads.append((X.grad.cpu().numpy(), knl.grad.cpu().numpy()))
    for g in grads[1:]:
        assert np.array_equal(g[0], grads[0][0]) and np.array_equal(g[1], grads[0][1])


def test_fma_build_gradients_close_to_exact():
    eb = lattices.adjoint_sublattice()
    turns = 3
    _, p, coords = _beam(256)
    tgt = _target(turns)
    out = []
    for fma in (False, True):
        tr = _tracker(eb, turns, p, fma=fma)
        X = _dev(coords).requires_grad_()
        tr.loss(X, target=tgt).backward()
        out.append(X.grad.cpu().numpy())
    scale = np.max(np.abs(out[0]), axis=0)
    assert np.max(np.abs(out[1] - out[0]) / scale) < 1e-10


S4 = np.array([[0.0, 1.0, 0.0, 0.0], [-1.0, 0.0, 0.0, 0.0],
               [0.0, 0.0, 0.0, 1.0], [0.0, 0.0, -1.0, 0.0]])


@pytest.mark.parametrize("turns", [1, 10])
def test_adjoint_transverse_jacobian_is_symplectic(turns):
    """With the cavity off (zero voltage: delta stays fixed) the sub-lattice
    is drifts and thin multipoles, a symplectic map of (x, px, y, py).  The
    adjoint's Jacobian (rows J^T e_c from gradient differences with the
    target moved in the last turn, exact for a quadratic loss) has a
    transverse block J4 with J4 S J4^T = S to rounding, after 1 and after
    10 turns."""
    eb = lattices.adjoint_sublattice(voltage=0.0)
    _, p, coords = _beam(1, seed=33)
    tr = _tracker(eb, turns, p, weights=(1.0,) * 6)
    x0 = coords[0]

    def grad_for(last):
        tg = np.zeros((turns, 6))
        tg[-1] = last
        X = _dev(x0[None, :]).requires_grad_()
        tr.loss(X, target=tg).backward()
        return X.grad.cpu().numpy()[0]

    g0 = grad_for(np.zeros(6))
    J = np.zeros((6, 6))
    for c in range(6):
        e = np.zeros(6)
        e[c] = 1e-3
        J[c] = (g0 - grad_for(e)) / 2e-3
    J4 = J[:4, :4]
    assert np.max(np.abs(J4 - np.eye(4))) > 1e-1
    big = float(np.max(np.abs(J4))) ** 2
    resid = J4 @ S4 @ J4.T - S4
    assert np.max(np.abs(resid)) < 1e-12 * big, (resid, J4)
    # delta is a constant of the motion and zeta does not act back
    np.testing.assert_array_equal(J[5, :5], np.zeros(5))
    assert abs(J[5, 5] - 1.0) < 1e-10
    assert np.max(np.abs(J[:4, 4])) == 0.0
