"""Chebyshev and Taylor propagation on the device around the SG4 H.psi
library, with piecewise-constant control fields, and the exact gradient of a
final-state overlap with respect to the field amplitudes (Chebyshev).

Chebyshev step (march_cheby, SRC/sub_propagation/sub_module_propa_march.f90
:4145-4436): with Hs = (H - E0)/Esc, E0 = (Hmin+Hmax)/2, Esc = (Hmax-Hmin)/2,
alpha = Esc dt and c_0 = J_0(alpha), c_k = 2 J_k(alpha) (cof, :5739-5760),
    w1 = psi, w2 = -i Hs w1, w_{k+1} = w_{k-1} - 2i Hs w_k,
    psi(t+dt) = exp(-i E0 dt) sum_k c_k w_k,
the series being cut when |c_k| ||w_k|| < poly_tol (default 1e-12 as in the
reference, sub_module_propagation.f90:1244).  A complex psi is applied as two
real vectors (sub_PrimOpPsi, SRC/sub_Operator/sub_OpPsi.f90:397-411): the
real and imaginary parts of all states form one block for the H.psi kernel.

Field: H(t) = H0 + eps_s mu on segment s (mu = sum_i q_i for the dipole
coupling); the reference has field-driven Taylor/RK propagators only
(march_noD_field, :1641), Chebyshev per constant-field segment is new here.

Taylor step (march_noD_field, :1641-1842, the field constant on a segment):
w_0 = psi, w_{k+1} = (-i dt/(k+1)) (H - E0) w_k, psi += w_{k+1} until the
squared norm of the new term is below poly_tol (default 1e-20), then
psi *= exp(-i E0 dt); see propagate_taylor_device.

Gradient of J = |<target|psi(T)>|^2 with respect to eps_s: exact for the
discrete map.  The adjoint state lambda_s = U_{>s}^dagger target is propagated
backwards with the conjugate Chebyshev series (the same polynomial, so U^dagger
is the exact adjoint of the forward map for a symmetric H); the derivative of
one segment is the forward tangent of the recurrence (dHs/deps = mu/Esc, E0
and Esc fixed), and dJ/deps_s = 2 Re(conj(a) <lambda_s | dpsi_s/deps_s>) with
a = <target|psi(T)>.  States are checkpointed every `checkpoint_every`
segments; the backward sweep recomputes the segments between checkpoints.
"""
import ctypes as C
import math

import numpy as np
import torch
from scipy.special import jv

from . import PREC_FP64, SG4Operator, SG4Batch, _capi
from ._capi import lib
from .psi_io import write_restart


def cheby_coefficients(alpha, n):
    """c_0 = J_0(alpha), c_k = 2 J_k(alpha) (cof + mmbsjn of the reference)."""
    c = jv(np.arange(n), alpha)
    c[1:] *= 2.0
    return c


class ChebyshevPropagator:
    """Propagates a block of complex states (ns, nb_ba) on the device."""

    def __init__(self, op, dt, Hmin, Hmax, poly_tol=1e-12, prec=PREC_FP64, mu=None,
                 device="cuda:0", npoly_max=20000):
        self.op, self.dt, self.prec = op, dt, prec
        self.Hmin, self.Hmax = Hmin, Hmax
        self.E0 = 0.5 * (Hmin + Hmax)
        self.Esc = 0.5 * (Hmax - Hmin)
        self.alpha = self.Esc * dt
        self.poly_tol = poly_tol
        self.dev = torch.device(device)
        self.coef = cheby_coefficients(self.alpha, npoly_max)
        # the series converges once k > alpha; keep the coefficients above tol
        big = np.nonzero(np.abs(self.coef) > poly_tol * 1e-3)[0]
        self.nmax = int(big.max()) + 2 if len(big) else 2
        self.phase = complex(math.cos(self.E0 * dt), -math.sin(self.E0 * dt))
        self.mu_grid = mu
        self.op_mu = None
        self.n_hpsi = 0

    # ---- operator pieces -------------------------------------------------
    def _apply_real(self, op, X):
        Y = torch.empty_like(X)
        for a in range(0, X.shape[0], op.max_nvec):
            b = min(a + op.max_nvec, X.shape[0])
            op.apply(X[a:b].contiguous(), Y[a:b], self.prec)
        self.n_hpsi += X.shape[0]
        return Y

    def _apply_c(self, op, x):
        ns = x.shape[0]
        X = torch.cat([x.real, x.imag], 0).contiguous()
        Y = self._apply_real(op, X)
        return torch.complex(Y[:ns], Y[ns:])

    def hs(self, x):
        return (self._apply_c(self.op, x) - self.E0 * x) / self.Esc

    def mu_s(self, x):
        """(dHs/deps) x = mu x / Esc with the SG4 representation of mu."""
        if self.op_mu is None:
            mu = self.mu_grid.cpu().numpy()
            if hasattr(self.op, "bases"):      # batch: one mu grid per instance
                q0 = self.op.q0
                self.op_mu = SG4Batch(self.op.bases, max_nvec=self.op.max_nvec,
                                      potential=[mu[q0[i]:q0[i + 1]] for i in range(len(q0) - 1)],
                                      ck=[np.zeros(b.D) for b in self.op.bases])
            else:
                b = self.op.basis
                self.op_mu = SG4Operator(b, max_nvec=self.op.max_nvec, potential=mu,
                                         ck=np.zeros(b.D))
        return self._apply_c(self.op_mu, x) / self.Esc

    def set_field(self, eps):
        self.op.set_field(self.mu_grid, eps)

    # ---- one step --------------------------------------------------------
    def step(self, psi, conj=False, npoly=None):
        """psi(t+dt) (conj=False) or the adjoint step U^dagger psi (conj=True).
        Returns (new psi, number of polynomials)."""
        r = 1j if conj else -1j
        w1 = psi
        acc = self.coef[0] * w1
        w2 = r * self.hs(w1)
        acc = acc + self.coef[1] * w2
        k = 2
        nlim = npoly if npoly is not None else self.nmax
        while k < nlim:
            w3 = w1 + 2 * r * self.hs(w2)
            acc = acc + self.coef[k] * w3
            w1, w2 = w2, w3
            k += 1
            if npoly is None:
                err = abs(self.coef[k - 1]) * torch.linalg.norm(w3).item()
                if err < self.poly_tol:
                    break
        ph = self.phase.conjugate() if conj else self.phase
        return acc * ph, k

    def step_tangent(self, psi, dpsi, npoly):
        """Forward tangent of one step: (U psi, U dpsi + (dU/deps) psi)."""
        r = -1j
        w1, t1 = psi, dpsi
        acc = self.coef[0] * w1
        tacc = self.coef[0] * t1
        w2 = r * self.hs(w1)
        t2 = r * (self.mu_s(w1) + self.hs(t1))
        acc = acc + self.coef[1] * w2
        tacc = tacc + self.coef[1] * t2
        for k in range(2, npoly):
            w3 = w1 + 2 * r * self.hs(w2)
            t3 = t1 + 2 * r * (self.mu_s(w2) + self.hs(t2))
            acc = acc + self.coef[k] * w3
            tacc = tacc + self.coef[k] * t3
            w1, w2, t1, t2 = w2, w3, t2, t3
        return acc * self.phase, tacc * self.phase


def spectral_bounds(op, nsteps=60, seed=1, device="cuda:0", margin=0.05):
    """Hmin, Hmax from a Lanczos run (the role of sub_Hmax,
    SRC/sub_propagation/sub_Hmax.f90:48) widened by `margin` of the range."""
    dev = torch.device(device)
    nb = op.nb_ba
    g = torch.Generator(device=dev).manual_seed(seed)
    q = torch.randn(1, nb, device=dev, dtype=torch.float64, generator=g)
    q /= torch.linalg.norm(q)
    Q = [q]
    al, be = [], []
    qprev = torch.zeros_like(q)
    b = 0.0
    for _ in range(min(nsteps, nb)):
        w = torch.empty_like(q)
        op.apply(q, w)
        a = torch.dot(w[0], q[0]).item()
        w = w - a * q - b * qprev
        for qq in Q:   # full reorthogonalisation (short runs)
            w = w - torch.dot(w[0], qq[0]) * qq
        b = torch.linalg.norm(w).item()
        al.append(a)
        be.append(b)
        if b < 1e-12:
            break
        qprev, q = q, w / b
        Q.append(q)
    T = np.diag(al) + np.diag(be[:-1], 1) + np.diag(be[:-1], -1)
    ev = np.linalg.eigvalsh(T)
    lo, hi = ev[0], ev[-1]
    d = hi - lo
    return lo - margin * d, hi + margin * d


def propagate_field(prop, psi0, eps, steps_per_segment, store_every=None, npolys=None,
                    restart_path=None, restart_every=1, t0=0.0):
    """Forward propagation through the segments; returns the final state, the
    checkpoints {segment index: state at its start} (if store_every is set)
    and the number of polynomials used per step (or the given ones).

    With ``restart_path``, the states and the time ``t0 + steps * dt`` are
    written as the reference's restart file (psi_io.write_restart) every
    ``restart_every`` steps and after the last one, as its march loop does
    after every step (SRC/sub_propagation/sub_module_propa_march.f90:226).
    A run is resumed from psi_io.read_restart's states with the segments
    not yet done and ``t0`` = the time read."""
    psi = psi0
    cps = {}
    used = []
    nstep = 0
    nall = len(eps) * steps_per_segment
    for s, e in enumerate(eps):
        if store_every and s % store_every == 0:
            cps[s] = psi
        prop.set_field(e)
        nps = []
        for j in range(steps_per_segment):
            psi, n = prop.step(psi, npoly=None if npolys is None else npolys[s][j])
            nps.append(n)
            nstep += 1
            if restart_path is not None and (nstep % restart_every == 0 or nstep == nall):
                write_restart(restart_path, t0 + nstep * prop.dt, psi)
        used.append(nps)
    npolys = used
    prop.set_field(0.0)
    return psi, cps, npolys


def overlap_gradient(prop, psi0, target, eps, steps_per_segment, checkpoint_every=10,
                     blocks=None):
    """J = sum_i |<target_i|psi_i(T)>|^2 over the row blocks of a batch
    (blocks = [(r0, r1), ...]; default: the whole vector) and dJ/deps_s for
    all segments, exact for the discrete Chebyshev map with the polynomial
    orders of the forward run.  The adjoint seed is sum_i a_i target_i with
    a_i = <target_i|psi_i(T)>, so that dJ/deps_s = 2 Re <lambda_s|dpsi_s>."""
    S = len(eps)
    psiT, cps, npolys = propagate_field(prop, psi0, eps, steps_per_segment, checkpoint_every)
    if blocks is None:
        blocks = [(0, psiT.shape[-1])]
    lam = torch.zeros_like(target)
    J = 0.0
    for r0, r1 in blocks:
        ai = torch.sum(target[..., r0:r1].conj() * psiT[..., r0:r1]).item()
        J += abs(ai) ** 2
        lam[..., r0:r1] = ai * target[..., r0:r1]
    grad = np.zeros(S)
    seg_states = {}
    for s in range(S - 1, -1, -1):
        if s not in seg_states:   # recompute the states of this checkpoint block
            c0 = (s // checkpoint_every) * checkpoint_every
            psi = cps[c0]
            for u in range(c0, s + 1):
                seg_states[u] = psi
                if u == s:
                    break
                prop.set_field(eps[u])
                for j in range(steps_per_segment):
                    psi, _ = prop.step(psi, npoly=npolys[u][j])
        prop.set_field(eps[s])
        psi = seg_states.pop(s)
        dpsi = torch.zeros_like(psi)
        for j in range(steps_per_segment):
            psi, dpsi = prop.step_tangent(psi, dpsi, npolys[s][j])
        grad[s] = 2.0 * torch.sum(lam.conj() * dpsi).item().real
        for j in range(steps_per_segment - 1, -1, -1):
            lam, _ = prop.step(lam, conj=True, npoly=npolys[s][j])
    prop.set_field(0.0)
    return J, grad


# ---- device propagator (sg4_cheby_propagate / sg4_cheby_field_gradient) ----
def to_rows(x):
    """complex (ns, nb_ba) -> the C ABI's rows layout: float64 (nb_ba, 2 ns),
    row r = real parts of the ns states, then their imaginary parts."""
    return torch.cat([x.real, x.imag], 0).T.contiguous()


def from_rows(r, ns):
    return torch.complex(r[:, :ns].T.contiguous(), r[:, ns:].T.contiguous())


def _params(prop, eps, steps_per_segment, checkpoint_every=10):
    coef = np.ascontiguousarray(prop.coef[:prop.nmax], dtype=np.float64)
    ev = np.ascontiguousarray(eps, dtype=np.float64)
    prm = _capi.ChebyParams()
    prm.dt, prm.E0, prm.Esc = prop.dt, prop.E0, prop.Esc
    prm.ncoef = len(coef)
    prm.coef = coef.ctypes.data_as(_capi.D_P)
    prm.nseg = len(ev)
    prm.steps_per_seg = steps_per_segment
    prm.eps = ev.ctypes.data_as(_capi.D_P)
    prm.ckpt_every = checkpoint_every
    prm.prec = prop.prec
    return prm, (coef, ev)


def propagate_device(prop, psi0, eps, steps_per_segment):
    """psi(T) through all segments with the device propagator: the same map as
    propagate_field with the fixed order count prop.nmax, the recurrence in
    device kernels with one H.psi launch per order and no host synchronisation
    per order."""
    ns = psi0.shape[0]
    x = to_rows(psi0)
    prm, keep = _params(prop, eps, steps_per_segment)
    mu = prop.mu_grid.contiguous()
    _capi.check(lib().sg4_cheby_propagate(prop.op._h, C.c_void_p(mu.data_ptr()), ns,
                                          C.c_void_p(x.data_ptr()), C.byref(prm)))
    return from_rows(x, ns)


def field_gradient_device(prop, psi0, target, eps, steps_per_segment, checkpoint_every=10,
                          blocks=None):
    """J = sum_i |<target_i|psi_i(T)>|^2 over the row blocks (default: the whole
    vector) and dJ/deps for every segment, computed on the device by
    sg4_cheby_field_gradient (forward with checkpoints, adjoint sweep with
    recomputation, forward tangents of the segments)."""
    ns = psi0.shape[0]
    if prop.op_mu is None:
        prop.mu_s(psi0[:1] * 0)          # builds the mu operator plan
    x0 = to_rows(psi0)
    tg = to_rows(target)
    if blocks is None:
        blocks = [(0, psi0.shape[-1])]
    br = (C.c_int64 * (2 * len(blocks)))(*[int(v) for b in blocks for v in b])
    prm, keep = _params(prop, eps, steps_per_segment, checkpoint_every)
    mu = prop.mu_grid.contiguous()
    J = C.c_double()
    grad = np.zeros(len(eps))
    _capi.check(lib().sg4_cheby_field_gradient(
        prop.op._h, prop.op_mu._h, C.c_void_p(mu.data_ptr()), ns, C.c_void_p(x0.data_ptr()),
        C.c_void_p(tg.data_ptr()), len(blocks), br, C.byref(prm), C.byref(J),
        grad.ctypes.data_as(_capi.D_P)))
    return J.value, grad


# ---- Taylor propagation (march_noD_field) --------------------------------------
def taylor_npoly(alpha, poly_tol=1e-20, max_poly=5000):
    """Maximal Taylor order for alpha = (Hmax - Hmin) dt / 2 (cof_nOD,
    SRC/sub_propagation/sub_module_propa_march.f90:5771-5827): first the order
    at which log(sum_k<=i alpha^k/k!) reaches alpha within poly_tol, then on
    until the term alpha^i/i! is below poly_tol, both at most max_poly + 1."""
    i, xi, reste = 1, float(alpha), 1.0 + alpha
    while abs(math.log(reste) - alpha) > poly_tol and i <= max_poly:
        i += 1
        xi *= alpha / i
        reste += xi
    while xi > poly_tol and i <= max_poly:
        i += 1
        xi *= alpha / i
    return i


def propagate_taylor_device(op, psi0, dt, Hmin, Hmax, eps=(0.0,), steps_per_segment=1, mu=None,
                            poly_tol=1e-20, npoly=None, max_norm2=0.0, prec=PREC_FP64,
                            max_poly=5000):
    """psi(T) by Taylor steps on the device (sg4_taylor_propagate): the series
    in H - E0 with E0 = (Hmin + Hmax)/2, cut per step when every state's new
    term has a squared norm below poly_tol (reference default 1e-20), at most
    npoly terms (default taylor_npoly of (Hmax - Hmin) dt / 2), H = H0 + eps_s mu
    on segment s (mu: device grid tensor, or None for no field).  Returns
    (psi(T), number of terms of every step)."""
    ns = psi0.shape[0]
    E0 = Hmin + 0.5 * (Hmax - Hmin)
    if npoly is None:
        npoly = taylor_npoly(0.5 * (Hmax - Hmin) * dt, poly_tol, max_poly)
    ev = np.ascontiguousarray(eps, dtype=np.float64)
    prm = _capi.TaylorParams()
    prm.dt, prm.E0, prm.npoly, prm.poly_tol = dt, E0, int(npoly), poly_tol
    prm.nseg, prm.steps_per_seg = len(ev), steps_per_segment
    prm.eps = ev.ctypes.data_as(_capi.D_P)
    prm.max_norm2, prm.prec = max_norm2, prec
    orders = np.zeros(len(ev) * steps_per_segment, dtype=np.int32)
    x = to_rows(psi0)
    mc = mu.contiguous() if mu is not None else None
    mup = C.c_void_p(mc.data_ptr()) if mc is not None else None
    _capi.check(lib().sg4_taylor_propagate(op._h, mup, ns, C.c_void_p(x.data_ptr()), C.byref(prm),
                                           orders.ctypes.data_as(C.POINTER(C.c_int32))))
    return from_rows(x, ns), orders
