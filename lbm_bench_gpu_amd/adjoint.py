"""Differentiable ZSC runs: a ``torch.autograd.Function`` over libzsc's
hand-written adjoint (include/zsc.h, zsc_adj_*).

The differentiated program is

    a_0 = init(phi0; theta)                  (init_f / init_g equilibria)
    a_{k+1} = step(a_k; theta),  k < N       (update, collide, stream)
    L = sum_x (phi(a_N)(x) - phi_tgt(x))^2

with theta = (kappa, beta, tau_f, tau_g, gravity): beta is the reference's
alpha (mu = 4 beta phi (phi^2 - phi*^2) - kappa lap phi), tau_f its tauPhi,
tau_g its tauRho, and gravity enters as 2 phi* gravity on phi > 0 cells
(src/paramread.c:141-170, src/collision_f.cu:96).

The forward pass keeps either a tape, the input state and phi of every step
(the backward pass then recomputes nothing: 108 B per cell and step at FP32,
so 128^3 cells x 400 steps take 90 GB of HBM), or a checkpoint every
``checkpoint_every`` steps, from which the backward pass recomputes each
segment before it runs the adjoint steps back through it.  By default the
tape is kept when it fits in half of the free device memory.  All arithmetic
runs in libzsc's HIP kernels; torch provides the device buffers.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import zsc
from ._lib import check, lib

THETA = ("kappa", "beta", "tau_f", "tau_g", "gravity")
SLOTS = ("alpha4", "kappa", "inv_tau_phi_one", "inv_tau_phi", "inv_eta2", "eta",
         "inv_tau_rho_one", "k0", "k1", "k2", "kc0", "kc1", "kc2", "grav", "loss")


def theta_of(params: zsc.Params):
    return [params.kappa, params.alpha, 1.0 / params.inv_tau_phi,
            1.0 / (1.0 - params.inv_tau_rho_one), params.gravity]


def coefficients(theta, base: zsc.Params) -> zsc.Params:
    """Step coefficients for free (kappa, beta, tau_f, tau_g, gravity); the
    other fields (phi*, Gamma, rho_mean, ...) come from ``base``.  Same
    expressions as zsc_derive_params."""
    kappa, beta, tau_f, tau_g, gravity = (float(t) for t in theta)
    itr = 1.0 / tau_g
    half = 1.0 - 0.5 * itr
    eta = 1.0 / (tau_f + 0.5)
    return base.replace(
        alpha=beta, alpha4=beta * 4.0, kappa=kappa, eta=eta, eta2=1.0 - eta,
        inv_eta2=0.5 / (1.0 - eta), inv_tau_phi=1.0 / tau_f, inv_tau_phi_one=1.0 - 1.0 / tau_f,
        inv_tau_rho_one=1.0 - itr, k0=itr * (1.0 / 3.0), k1=itr * (1.0 / 18.0),
        k2=itr * (1.0 / 36.0), kc0=half * 1.0, kc1=half * (1.0 / 6.0), kc2=half * (1.0 / 12.0),
        gravity=gravity, grav=2.0 * base.phi_star * gravity)


def chain_rule(g, theta, phi_star: float):
    """dL/dtheta from dL/d(step coefficients) (the 16 accumulated slots)."""
    _, _, tau_f, tau_g, _ = (float(t) for t in theta)
    G = dict(zip(SLOTS, g))
    eta = 1.0 / (tau_f + 0.5)
    d_kappa = G["kappa"]
    d_beta = 4.0 * G["alpha4"]
    d_tau_f = ((G["inv_tau_phi_one"] - G["inv_tau_phi"]) / tau_f ** 2
               - eta ** 2 * (G["eta"] + G["inv_eta2"] * 0.5 / (1.0 - eta) ** 2))
    d_tau_g = ((G["inv_tau_rho_one"] - G["k0"] / 3.0 - G["k1"] / 18.0 - G["k2"] / 36.0) / tau_g ** 2
               + (G["kc0"] + G["kc1"] / 6.0 + G["kc2"] / 12.0) * 0.5 / tau_g ** 2)
    d_grav = 2.0 * phi_star * G["grav"]
    return [d_kappa, d_beta, d_tau_f, d_tau_g, d_grav]


def _vp(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


def _own_queue_stream(device):
    """A stream on a hardware queue of its own, so that its kernels can run
    beside those of the runner's stream.  Streams share the process's few
    hardware queues (4 unless GPU_MAX_HW_QUEUES says otherwise), and two
    streams that share one run one after the other; a stream created with a
    compute-unit mask gets its own queue.  The mask names every compute unit,
    so the stream may use the whole GPU; its priority is the default.  Falls
    back to a pooled torch stream if the call is refused."""
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        ncu = torch.cuda.get_device_properties(device).multi_processor_count
        words = (ncu + 31) // 32
        mask = (ctypes.c_uint32 * words)(*([0xFFFFFFFF] * words))
        if ncu % 32:
            mask[words - 1] = (1 << (ncu % 32)) - 1
        s = ctypes.c_void_p()
        with torch.cuda.device(device):
            r = hip.hipExtStreamCreateWithCUMask(ctypes.byref(s), ctypes.c_uint32(words), mask)
        if r != 0 or not s.value:
            raise OSError(r)
    except (OSError, AttributeError):
        return torch.cuda.Stream(device=device)
    st = torch.cuda.ExternalStream(s.value, device=device)
    st._zsc_hip = hip  # the creator's handle, for the destroy below
    return st


def _destroy_stream(st):
    hip = getattr(st, "_zsc_hip", None)
    if hip is not None:
        hip.hipStreamSynchronize(ctypes.c_void_p(st.cuda_stream))
        hip.hipStreamDestroy(ctypes.c_void_p(st.cuda_stream))


class AdjointRunner:
    """Owns a context and the device buffers of one differentiable run."""

    def __init__(self, dims, base: zsc.Params, precision: str = "f32", checkpoint_every: int = 20,
                 gravity_on: bool = True, device: str = "cuda", timing: bool = False,
                 tape="auto", tape_fraction: float = 0.5, states: str | None = None,
                 overlap: bool = True):
        """precision: storage of the adjoint lattices; states: precision of
        the forward run and its stored states (default: the same).  FP32
        adjoints over FP64 states (precision="f32", states="f64") keep the
        FP64 forward and FP64 per-cell reverse arithmetic at FP32 adjoint
        storage.  FP32 adjoints over f32d states (states="f32d": FP32
        lattices whose g are deviations from w_i rho_mean) run at FP32 cost
        throughout and keep the density variations that FP32 states round
        away, with the coefficient sums formed without the rest state's
        part (it cancels in the tau_g derivative).  tape: True keeps every step's input state and phi (no
        recomputation), False a checkpoint every ``checkpoint_every`` steps
        and the states of the last segment (its backward steps come first),
        "auto" the tape when it takes at most ``tape_fraction`` of the device
        memory that is free (to the driver or in torch's cache) when the
        forward pass starts.  overlap: with checkpoints, recompute the next
        segment (on a second stream, into a second set of buffers) while
        the adjoint steps of this one run, when both sets of buffers take at
        most 80% of the free device memory (otherwise in line)."""
        states = states or precision
        if (precision, states) not in (("f32", "f32"), ("f64", "f64"), ("f32", "f64"),
                                       ("f32", "f32d")):
            raise ValueError("the adjoint runs at f32 or f64, or f32 over f64 or f32d states")
        if tape not in (True, False, "auto"):
            raise ValueError("tape is True, False or 'auto'")
        if int(checkpoint_every) < 1:
            raise ValueError("checkpoint_every must be >= 1")
        self.tape_mode = tape
        self.tape_fraction = float(tape_fraction)
        self.tape = None
        self.dims = tuple(dims)
        nx, ny, nz = self.dims
        self.n = nx * ny * nz
        self.base = base
        self.k = int(checkpoint_every)
        self.dtype = torch.float64 if states == "f64" else torch.float32    # states, phi
        self.ldtype = torch.float32 if precision == "f32" else torch.float64  # adjoint lattices
        self.device = device
        # A stream of our own (default priority): the library's kernels and
        # the torch ops on these buffers are ordered on it.  torch's default
        # stream has handle 0, which the C-ABI reads as "make a new stream".
        self.stream = torch.cuda.Stream(device=device)
        self.ctx = zsc.Context(nx, ny, nz, precision=states, stream=self.stream.cuda_stream)
        check(lib().zsc_adj_set_lambda_precision(self.ctx._h, zsc._lib.PRECISIONS[precision]))
        self.ctx.set_gravity(gravity_on)
        self.states_precision = states
        self.gravity_on = bool(gravity_on)
        self.overlap = bool(overlap)
        self.stream2 = self.ctx2 = None  # the recomputation's, made on first use
        nbytes = lib().zsc_adj_scratch_bytes(self.ctx._h)
        self.scratch = torch.empty(nbytes, dtype=torch.uint8, device=device)
        self.grads = torch.zeros(16, dtype=torch.float64, device=device)
        self.checkpoints = {}
        self.nsteps = 0
        self.phi0 = None
        # with timing, backward() leaves the device time of the segment
        # recomputations and of the adjoint steps in last_timing (ms)
        self.timing = timing
        self.last_timing = None

    def close(self):
        """Frees the recomputation's context and stream (the runner's own
        context frees itself)."""
        if self.ctx2 is not None:
            self.ctx2.close()
            self.ctx2 = None
        if self.stream2 is not None:
            _destroy_stream(self.stream2)
            self.stream2 = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _lattice(self):
        return torch.empty((26, self.n), dtype=self.dtype, device=self.device)

    def _lam(self):
        return torch.empty((26, self.n), dtype=self.ldtype, device=self.device)

    def _set_theta(self, theta):
        self.theta = [float(t) for t in theta]
        self.ctx.set_params(coefficients(self.theta, self.base))
        if self.ctx2 is not None:
            self.ctx2.set_params(coefficients(self.theta, self.base))

    def _recompute_ctx(self):
        """A second context on a second stream (default priority) for the
        recomputation, with this run's parameters and gravity."""
        if self.ctx2 is None:
            self.stream2 = _own_queue_stream(self.device)
            self.ctx2 = zsc.Context(*self.dims, precision=self.states_precision,
                                    stream=self.stream2.cuda_stream)
            self.ctx2.set_gravity(self.gravity_on)
            self.ctx2.set_params(coefficients(self.theta, self.base))
        return self.ctx2

    def forward(self, theta, phi0: torch.Tensor, nsteps: int, keep=True) -> torch.Tensor:
        """a_N; keeps checkpoints every k steps when ``keep``."""
        torch.cuda.current_stream().synchronize()
        with torch.cuda.stream(self.stream):
            a = self._forward(theta, phi0, nsteps, keep)
        self.stream.synchronize()
        return a

    def _phi(self):
        return torch.empty(self.n, dtype=self.dtype, device=self.device)

    def tape_bytes(self, nsteps: int) -> int:
        """Device bytes of the tape of an nsteps run: B_0 .. B_{N-2} and the
        phi of every step."""
        esize = 4 if self.dtype == torch.float32 else 8
        return (26 * max(nsteps - 1, 0) + nsteps) * self.n * esize

    def _use_tape(self, nsteps: int) -> bool:
        if self.tape_mode != "auto":
            return bool(self.tape_mode)
        free, _ = torch.cuda.mem_get_info(self.device)
        cached = torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
        return self.tape_bytes(nsteps) <= self.tape_fraction * (free + cached)

    def _forward(self, theta, phi0, nsteps, keep):
        # Steps run as fused launches on post-collision lattices:
        # B_0 = collide(A_0), B_n = fused(B_{n-1}), A_N = stream(B_{N-1}).
        # The checkpoint of step s is A_0 (s = 0) or B_{s-1}.
        self._set_theta(theta)
        self.phi0 = phi0.to(self.device, self.dtype).contiguous().view(-1)
        h = self.ctx._h
        self.checkpoints, self.tape, self.tail = {}, None, None
        a0 = self._lattice()
        check(lib().zsc_adj_init(h, _vp(self.phi0), _vp(a0)))
        self.checkpoints = {0: a0} if keep else {}
        self.nsteps = nsteps
        if nsteps == 0:
            self.final = a0
            return a0
        if keep and self._use_tape(nsteps):
            # tape: step n's input state (A_0, then B_{n-1}) and phi(A_n)
            states, phis = [a0], [self._phi()]
            b = self._lattice()
            check(lib().zsc_adj_collide(h, _vp(a0), _vp(b), _vp(phis[0])))
            for s in range(1, nsteps):
                nb, phi = self._lattice(), self._phi()
                check(lib().zsc_adj_fused(h, _vp(b), _vp(nb), _vp(phi)))
                states.append(b)
                phis.append(phi)
                b = nb
            a = self._lattice()
            check(lib().zsc_adj_stream(h, _vp(b), _vp(a)))
            self.tape = (states, phis)
            self.final = a
            return a
        # checkpoints; the last segment's states and phi are kept as they are
        # computed (its backward steps come first, and need no recomputation)
        last = (nsteps - 1) // self.k * self.k
        b, b2, phi = self._lattice(), self._lattice(), self._phi()
        check(lib().zsc_adj_collide(h, _vp(a0), _vp(b), _vp(phi)))
        tail = ([a0], [phi]) if keep and last == 0 else None
        for s in range(1, nsteps):
            if keep and s == last:
                self.checkpoints[s] = b  # kept in the tail, never written again
                tail = ([], [])
            elif keep and s % self.k == 0:
                self.checkpoints[s] = b.clone()
            if tail is not None:
                nb, ph = self._lattice(), self._phi()
                check(lib().zsc_adj_fused(h, _vp(b), _vp(nb), _vp(ph)))
                tail[0].append(b)
                tail[1].append(ph)
                b = nb
            else:
                check(lib().zsc_adj_fused(h, _vp(b), _vp(b2), None))
                b, b2 = b2, b
        a = b2
        check(lib().zsc_adj_stream(h, _vp(b), _vp(a)))
        self.tail = tail
        self.final = a
        return a

    def loss(self, a_n: torch.Tensor, phi_tgt: torch.Tensor, lam: torch.Tensor | None = None):
        torch.cuda.current_stream().synchronize()
        with torch.cuda.stream(self.stream):
            out = self._loss(a_n, phi_tgt, lam)
        self.stream.synchronize()
        return out

    def _loss(self, a_n, phi_tgt, lam):
        self.grads.zero_()
        tgt = phi_tgt.to(self.device, self.dtype).contiguous().view(-1)
        lam = self._lam() if lam is None else lam
        check(lib().zsc_adj_seed(self.ctx._h, _vp(a_n), _vp(tgt), _vp(lam), _vp(self.scratch),
                                 ctypes.cast(self.grads.data_ptr(), ctypes.POINTER(ctypes.c_double))))
        return float(self.grads[14].item()), lam

    def backward(self, lam: torch.Tensor, scale: float = 1.0):
        """dL/dtheta from the tape, or by recomputing each checkpoint segment;
        lam = dL/da_N."""
        torch.cuda.current_stream().synchronize()
        try:
            with torch.cuda.stream(self.stream):
                g = self._backward(lam, scale)
        finally:
            # also after an error: no recomputation may still write buffers
            # that are freed with this call's
            self.stream.synchronize()
            if self.stream2 is not None:
                self.stream2.synchronize()
        return g

    def _recompute(self, s0, s1, h=None, bank=None):
        """The input state and phi_n of every step of the segment s0 .. s1-1,
        recomputed from its checkpoint (through context handle h, default
        the runner's) into bank = (lattices, phis) or new buffers."""
        h = self.ctx._h if h is None else h
        out = iter(bank[0]) if bank is not None else None
        outp = iter(bank[1]) if bank is not None else None
        lat = (lambda: next(out)) if bank is not None else self._lattice
        ph = (lambda: next(outp)) if bank is not None else self._phi
        states, phis = [], []
        if s0 == 0:
            b, phi = lat(), ph()
            check(lib().zsc_adj_collide(h, _vp(self.checkpoints[0]), _vp(b), _vp(phi)))
            states.append(self.checkpoints[0])
            phis.append(phi)
            n0 = 1
        else:
            b = self.checkpoints[s0]
            n0 = s0
        for _ in range(n0, s1):
            nb, phi = lat(), ph()
            check(lib().zsc_adj_fused(h, _vp(b), _vp(nb), _vp(phi)))
            states.append(b)
            phis.append(phi)
            b = nb
        return states, phis

    def _adjoint_steps(self, s0, states, phis, lam, lam_out, lam_tmp):
        """lam_out = the adjoint at step s0 from lam at the segment's end:
        one zsc_adj_backward_steps call (coefficient sums kept in scratch)."""
        k = len(states)
        sp = (ctypes.c_void_p * k)(*[t.data_ptr() for t in states])
        pp = (ctypes.c_void_p * k)(*[t.data_ptr() if t is not None else None for t in phis])
        check(lib().zsc_adj_backward_steps(self.ctx._h, k, sp, pp, 0 if s0 == 0 else 1, _vp(lam),
                                           _vp(lam_out), _vp(lam_tmp), _vp(self.scratch), None))

    def _backward(self, lam, scale):
        if self.tape is None and 0 not in self.checkpoints:
            raise RuntimeError("backward needs the states of a forward pass run with keep=True "
                               "(the last forward pass kept none)")
        h = self.ctx._h
        gp = ctypes.cast(self.grads.data_ptr(), ctypes.POINTER(ctypes.c_double))
        loss_slot = float(self.grads[14].item())
        # drop coefficient sums a failed pass may have left in the scratch
        # buffer (a no-op otherwise), so that this pass starts from none
        check(lib().zsc_adj_flush(h, _vp(self.scratch), gp))
        self.grads.zero_()
        lam2, lam3 = self._lam(), self._lam()
        if self.tape is not None:
            # every step's input is on the tape: one segment, nothing to recompute
            segs = [(0, self.nsteps, self.tape)]
        else:
            starts = sorted(self.checkpoints)
            segs = [(s0, starts[i + 1] if i + 1 < len(starts) else self.nsteps, None)
                    for i, s0 in enumerate(starts)]
            if self.tail is not None:
                segs[-1] = segs[-1][:2] + (self.tail,)
        segs = segs[::-1]
        overlap = self.tape is None and self.overlap and len(segs) > 1
        if overlap:
            # two sets of segment buffers instead of one: only where they fit
            esize = 4 if self.dtype == torch.float32 else 8
            need = 2 * self.k * 27 * self.n * esize
            free, _ = torch.cuda.mem_get_info(self.device)
            cached = torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
            overlap = need <= 0.8 * (free + cached)
        ev = lambda: torch.cuda.Event(enable_timing=self.timing)
        marks = []
        if overlap:
            # segment i is recomputed on stream2 into bank i % 2 while segment
            # i - 1's adjoint steps run on the runner's stream; bank i % 2 is
            # rewritten only after segment i - 2's adjoint steps are done
            h2 = self._recompute_ctx()._h
            k = max(s1 - s0 for s0, s1, taped in segs if taped is None)
            banks = [([self._lattice() for _ in range(k)], [self._phi() for _ in range(k)])
                     for _ in range(2)]
            done_rec, done_adj, t_rec = {}, {}, []

            def issue(i):
                s0, s1, taped = segs[i]
                if taped is not None:  # the last segment, kept by the forward pass
                    done_rec[i] = None
                    return taped
                with torch.cuda.stream(self.stream2):
                    if i >= 2:
                        self.stream2.wait_event(done_adj[i - 2])
                    t0 = ev()
                    t0.record(self.stream2)
                    st = self._recompute(s0, s1, h2, banks[i % 2])
                    done_rec[i] = ev()
                    done_rec[i].record(self.stream2)
                    t_rec.append((t0, done_rec[i]))
                return st

            pending = issue(0)
            for i, (s0, s1, _) in enumerate(segs):
                states, phis = pending
                if i + 1 < len(segs):
                    pending = issue(i + 1)
                if done_rec[i] is not None:
                    self.stream.wait_event(done_rec[i])
                t0 = ev()
                t0.record(self.stream)
                self._adjoint_steps(s0, states, phis, lam, lam2, lam3)
                lam, lam2 = lam2, lam
                done_adj[i] = ev()
                done_adj[i].record(self.stream)
                marks.append((t0, done_adj[i]))
        else:
            t_rec = []
            for s0, s1, taped in segs:
                t0 = ev()
                t0.record(self.stream)
                if taped is not None:
                    states, phis = taped
                else:
                    states, phis = self._recompute(s0, s1)
                t1 = ev()
                t1.record(self.stream)
                self._adjoint_steps(s0, states, phis, lam, lam2, lam3)
                lam, lam2 = lam2, lam
                t2 = ev()
                t2.record(self.stream)
                t_rec.append((t0, t1))
                marks.append((t1, t2))
                del states, phis
        check(lib().zsc_adj_flush(h, _vp(self.scratch), gp))
        check(lib().zsc_adj_init_backward(h, _vp(self.phi0), _vp(lam), _vp(self.scratch), gp))
        if self.timing:
            self.stream.synchronize()
            # with overlap the two kinds of interval run at the same time
            self.last_timing = {
                "recompute_ms": sum(a.elapsed_time(b) for a, b in t_rec),
                "adjoint_steps_ms": sum(a.elapsed_time(b) for a, b in marks),
                "overlapped": overlap}
        g = self.grads.cpu().numpy()
        g[14] = loss_slot
        self.lam0 = lam
        return [scale * x for x in chain_rule(g, self.theta, self.base.phi_star)]

    def loss_and_grad(self, theta, phi0, phi_tgt, nsteps):
        a = self.forward(theta, phi0, nsteps)
        L, lam = self.loss(a, phi_tgt)
        return L, self.backward(lam)

    def loss_only(self, theta, phi0, phi_tgt, nsteps):
        a = self.forward(theta, phi0, nsteps, keep=False)
        return self.loss(a, phi_tgt)[0]


class ZSCLoss(torch.autograd.Function):
    """L(theta) = sum (phi_N - phi_tgt)^2 of a ZSC run; theta is a float64
    tensor (kappa, beta, tau_f, tau_g, gravity).  The runner holds the states
    of its last forward pass, so one loss per runner may await its backward
    pass at a time."""

    @staticmethod
    def forward(ctx, theta, phi0, phi_tgt, runner: AdjointRunner, nsteps: int):
        a = runner.forward(theta.detach().cpu().tolist(), phi0, nsteps)
        L, lam = runner.loss(a, phi_tgt)
        ctx.runner = runner
        ctx.lam = lam
        return torch.tensor(L, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.runner.backward(ctx.lam, scale=float(grad_out))
        return torch.tensor(g, dtype=torch.float64), None, None, None, None


def zsc_loss(theta: torch.Tensor, phi0, phi_tgt, runner: AdjointRunner, nsteps: int):
    return ZSCLoss.apply(theta, phi0, phi_tgt, runner, nsteps)


def theta_array(params: zsc.Params) -> np.ndarray:
    return np.array(theta_of(params))
