"""Plain energy solver on the GPU: gradient descent with legacy Adam.

Restates igm/processes/iceflow/solve/solve.py:11-119 MI355X-first:

* one fused kernel per iteration evaluates the energy, its adjoint and the
  Adam update (ping-pong velocity buffers, moments updated in place);
* each pass first reduces the previous iteration's float64 COST and runs
  the EarlyStopping rule (iceflow/utils.py:80-111) on the device, every
  block for itself, so an iteration is one kernel launch;
* the host never waits on a COST value inside the loop: it reads the
  device stop flag every ``check_every`` iterations.  Iterations enqueued
  after the stop are no-ops (the kernels test the flag), so the result is
  the one the reference's per-iteration ``COST.numpy()`` check gives.

Adam step constants follow TensorFlow's legacy Adam (see
``LegacyAdamState``); the step counter persists across solves like the
optimizer created once by ``initialize_iceflow_solver`` (solve.py:11-20).
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .energy import EnergyOperator


class LegacyAdamState:
    """Global step counter + hyper-parameters of tf.keras.optimizers.legacy.Adam."""

    def __init__(self, lr=1.0, beta1=0.9, beta2=0.999, epsilon=1e-7):
        self.lr, self.beta1, self.beta2, self.epsilon = lr, beta1, beta2, epsilon
        self.iterations = 0

    def alphas(self, t_first, lrs, dtype=torch.float32):
        """lr_i * sqrt(1 - beta2^t) / (1 - beta1^t) for the steps t = t_first,
        t_first + 1, ... (1-based) with step sizes lrs, in the working
        precision (float32 for the float32 solver, as TF computes it).
        Returns a numpy array of that precision."""
        if dtype == torch.float32:
            f = np.float32
            t = np.arange(t_first, t_first + len(lrs), dtype=np.float64).astype(f)
            b1p = np.power(f(self.beta1), t)
            b2p = np.power(f(self.beta2), t)
            return np.asarray(lrs, dtype=np.float64).astype(f) * np.sqrt(f(1) - b2p) / (f(1) - b1p)
        t = np.arange(t_first, t_first + len(lrs), dtype=np.float64)
        return np.asarray(lrs, dtype=np.float64) * np.sqrt(1 - self.beta2 ** t) / (1 - self.beta1 ** t)

    def alpha(self, t, dtype=torch.float32):
        """The step alpha at step t (1-based) with the current lr (see ``alphas``)."""
        return float(self.alphas(t, [self.lr], dtype)[0])


class SolveResult:
    """U, V (device), the COST of every iteration, the iteration count and
    whether EarlyStopping ended the solve.  ``costs_device`` is the float64
    COST history on the device; ``costs`` reads it to a numpy array on first
    use (so a caller that does not look at it never waits for the GPU)."""

    def __init__(self, U, V, costs, iterations, stopped):
        self.U, self.V, self.iterations, self.stopped = U, V, iterations, stopped
        self.costs_device = costs if isinstance(costs, torch.Tensor) else None
        self._costs = None if isinstance(costs, torch.Tensor) else costs

    @property
    def costs(self):
        if self._costs is None:
            self._costs = self.costs_device.cpu().numpy()
        return self._costs


# Pinned host buffers for the small uploads of a solve, in a ring: a buffer
# is refilled only once the copy out of it queued last time has run (its
# event), which with several buffers has long happened.
_RING = [[None, None] for _ in range(4)]
_ring_next = [0]


def upload(chunks, dev):
    """One host-to-device copy of the concatenated byte strings ``chunks``,
    queued on the current stream from pinned memory (the host does not wait
    for the GPU work before it).  Returns the device uint8 tensor."""
    n = sum(len(c) for c in chunks)
    slot = _RING[_ring_next[0]]
    _ring_next[0] = (_ring_next[0] + 1) % len(_RING)
    buf, ev = slot
    if ev is not None:
        ev.synchronize()
    if buf is None or buf.numel() < n:
        buf = torch.empty(max(n, 1 << 14), dtype=torch.uint8).pin_memory()
        slot[0] = buf
    a = buf.numpy()
    o = 0
    for c in chunks:
        a[o:o + len(c)] = np.frombuffer(c, dtype=np.uint8)
        o += len(c)
    d = torch.empty(max(n, 1), dtype=torch.uint8, device=dev)
    d[:n].copy_(buf[:n], non_blocking=True)
    ev = torch.cuda.Event()
    ev.record()
    slot[1] = ev
    return d


_TABLES = {}  # step tables of keyed schedules, the most recent _TABLES_MAX
_TABLES_MAX = 16


def step_table(optimizer, t0, n, lr_schedule, dt):
    """The Adam steps alpha of iterations t0 + 1 .. t0 + n (``optimizer``'s
    hyper-parameters, step ``lr_schedule(i)`` or the optimizer's lr) in the
    working precision, contiguous.  Tables of a schedule with a ``key`` (an
    ``adam_schedule``, or the optimizer's constant lr) are kept: a polish
    repeated every time step, with a fresh optimizer each time, builds its
    table once."""
    sk = getattr(lr_schedule, "key", None) if lr_schedule is not None else ("constant", optimizer.lr)
    ck = None if sk is None else (sk, t0, n, optimizer.beta1, optimizer.beta2, dt)
    tab = _TABLES.pop(ck, None) if ck is not None else None
    if tab is None:
        lrs = [lr_schedule(i) if lr_schedule is not None else optimizer.lr for i in range(n)]
        tab = np.ascontiguousarray(optimizer.alphas(t0 + 1, lrs, dt))
        tab.setflags(write=False)
        if ck is not None and len(_TABLES) >= _TABLES_MAX:
            del _TABLES[next(iter(_TABLES))]  # the least recently used
    if ck is not None:
        _TABLES[ck] = tab
    return tab


def solve(op: EnergyOperator, U0, V0, nbitmax, optimizer: LegacyAdamState, early_stopping=True,
          relative_min_delta=0.0002, patience=10, check_every=16, lr_schedule=None, pairs=True):
    """Minimise the energy of ``op`` from (U0, V0) (planar (Nz, Ny, Nx)).

    Returns SolveResult with unmasked U, V (device), the COST history (host
    float64 numpy array) and the number of iterations whose update was
    applied.  ``lr_schedule(i)`` (optional) overrides the learning rate of
    iteration i (the reference uses a constant step_size).

    The float32 solver runs device-driven passes (``run_solve_pass``): each
    runs two iterations whenever the first of them cannot trigger the
    EarlyStopping rule -- always, without early stopping -- and gives the
    same result as one iteration per pass; ``pairs=False`` keeps one
    iteration per launch.
    """
    dt = op.dtype
    assert U0.dtype == dt and U0.shape == V0.shape and U0.dim() == 3
    dev = U0.device
    U0, V0 = U0.detach().contiguous(), V0.detach().contiguous()
    # Pass 0 reads (U0, V0) as given and writes buffer 0; pass k >= 1 reads
    # the buffer pass k-1 wrote and writes the other one.  The input is never
    # written, so it needs no copy.
    Ub = [torch.empty_like(U0), torch.empty_like(U0)]
    Vb = [torch.empty_like(V0), torch.empty_like(V0)]

    def src(k):
        return (U0, V0) if k == 0 else (Ub[(k - 1) % 2], Vb[(k - 1) % 2])

    def dst(k):
        return Ub[k % 2], Vb[k % 2]

    def field_after(p):
        """The field once p passes have applied their update (a copy of the
        input for p = 0)."""
        return (U0.clone(), V0.clone()) if p == 0 else dst(p - 1)

    # Adam slots, one (m_U, v_U, m_V, v_V) record per node and level; a
    # fresh optimizer starts from zero moments
    mshape = tuple(U0.shape) + (4,)
    es_host = L.IgmEarlyStop(best_loss=0.0, relative_min_delta=relative_min_delta, has_best=0,
                             wait=0, patience=patience if early_stopping else 2**30, stopped=0,
                             stop_iter=-1, iters_done=0)
    # Two EarlyStopping states and two partial-sum buffers: pass i finalizes
    # iteration i-1 itself (every block sums its partials and applies the
    # rule), reading state E[i % 2] and writing E[(i - 1) % 2], so a state is
    # never written while blocks of the same pass read it; the pass writes its
    # own partials to P[i % 2].  The state of the last launched pass is
    # E[n % 2], where a separate finalize then accounts for iteration n - 1.
    es_size = C.sizeof(L.IgmEarlyStop)
    t0 = optimizer.iterations
    # the Adam step of every iteration, as the legacy optimizer computes it
    alphas = step_table(optimizer, t0, nbitmax, lr_schedule, dt)
    # the two states and the step table in one host-to-device copy (es_size
    # is a multiple of 8, so the table is aligned)
    up = upload([bytes(es_host) * 2, alphas.tobytes()], dev)
    es_dev = up[:2 * es_size]
    E = [es_dev.data_ptr(), es_dev.data_ptr() + es_size]
    P = [op.partials, op.partials_b]
    # every COST this returns is written by a pass or a finalize
    hist = torch.empty(max(1, nbitmax), dtype=torch.float64, device=dev)
    adam = L.IgmAdam(alpha=0.0, beta1=optimizer.beta1, beta2=optimizer.beta2,
                     epsilon=optimizer.epsilon)

    def result(last):
        raw = bytes(es_dev.cpu().numpy().tobytes())
        return L.IgmEarlyStop.from_buffer_copy(raw[last * es_size:(last + 1) * es_size])

    if pairs and op.pair_supported and nbitmax >= 1:
        # Device-driven passes: launch k reads buffer k % 2 of U, V and the
        # moments and writes buffer (k + 1) % 2; it accounts the previous
        # launch's iterations itself (reading state E[k % 2], writing
        # E[(k + 1) % 2]) and runs two iterations when the first cannot end
        # the solve, else one, else none.  Without early stopping that is
        # pairs throughout (the last one single when nbitmax is odd); with
        # it, the host reads the state every check_every launches and stops
        # launching once the solve has ended.  A last launch accounts the
        # final pass; the field is then in buffer passes % 2.
        # launch 0 reads no moments (the pass takes them as zero), so neither
        # buffer needs clearing
        M = [torch.empty(mshape, dtype=dt, device=dev), torch.empty(mshape, dtype=dt, device=dev)]
        assert dt == torch.float32
        atab = up[2 * es_size:].view(torch.float32)

        def launch(k, finalize_only=False, last=False, io=None):
            # launch 0 reads no moments (zero); the last launch of a solve
            # whose length is known writes none (nothing reads them)
            (Us, Vs), (Ud, Vd) = io if io is not None else (src(k), dst(k))
            op.run_solve_pass(Us, Vs, Ud, Vd, M[k % 2] if k > 0 else None,
                              None if last else M[(k + 1) % 2], adam, atab, nbitmax, P[k % 2],
                              P[(k - 1) % 2] if k > 0 else None, E[k % 2], E[(k + 1) % 2], hist,
                              finalize_only)

        k = 0
        if not early_stopping:
            # every launch runs one or two iterations and nothing stops the
            # solve: nbitmax iterations in k passes; the host does not wait
            # for the GPU.  Launch 0 reads the planar start field, the last
            # one writes the planar result to buffer 0, and the fields in
            # between stay interleaved, (Nz, Ny, Nx, 2): one 8-B access per
            # node and level instead of two 4-B ones.
            k = (nbitmax + 1) // 2
            UV = [torch.empty(tuple(U0.shape) + (2,), dtype=dt, device=dev) for _ in range(2)] \
                if k >= 2 else [None, None]

            def io_of(j):
                s_ = (U0, V0) if j == 0 else (UV[(j - 1) % 2], None)
                d_ = (Ub[0], Vb[0]) if j == k - 1 else (UV[j % 2], None)
                return s_, d_

            for j in range(k):
                launch(j, last=j == k - 1, io=io_of(j))
            # (the accounting launch runs no iteration: it reads and writes no
            # field, any two distinct buffers serve)
            launch(k, finalize_only=True, io=((Ub[0], Vb[0]), (Ub[1], Vb[1])))
            optimizer.iterations = t0 + nbitmax
            return SolveResult(Ub[0], Vb[0], hist[:nbitmax], nbitmax, False)
        else:
            while k < nbitmax:
                launch(k)
                k += 1
                if k >= 2 and k % check_every == 0:
                    st = result(k % 2)  # the state before launch k-1 ran
                    if st.stopped or st.iters_done >= nbitmax:
                        break
        launch(k, finalize_only=True)
        st = result((k + 1) % 2)
        n = int(st.iters_done)
        optimizer.iterations = t0 + n
        return SolveResult(*field_after(int(st.passes)), hist[:n], n, bool(st.stopped))

    mom = torch.zeros(mshape, dtype=dt, device=dev)  # updated in place by every pass
    n_launched = 0
    for i in range(nbitmax):
        adam.alpha = float(alphas[i])
        prev = None if i == 0 else (P[(i - 1) % 2], E[i % 2], E[(i - 1) % 2], hist, i - 1)
        (Us, Vs), (Ud, Vd) = src(i), dst(i)
        op.run_pass(2, Us, Vs, Uo=Ud, Vo=Vd, moments=mom, adam=adam, partials=P[i % 2], prev=prev)
        n_launched = i + 1
        if early_stopping and i > 0 and (i + 1) % check_every == 0:
            o = ((i - 1) % 2) * es_size + L.EARLYSTOP_STOPPED_OFFSET
            if int(es_dev[o:o + 4].view(torch.int32).item()) != 0:
                break
    if n_launched > 0:
        last = n_launched % 2
        op.finalize(es=E[last], cost_hist=hist, it=n_launched - 1, partials=P[(n_launched - 1) % 2])
    else:
        last = 0
    st = result(last)
    n = int(st.iters_done)
    optimizer.iterations = t0 + n
    costs = hist[:n].cpu().numpy()
    return SolveResult(*field_after(n), costs, n, bool(st.stopped))


class adam_schedule:
    """Step schedule: a geometric warm-up from lr_start to lr0 over n_warmup
    iterations (none by default), lr0 for n_flat iterations, then a geometric
    decay to lr_end over n_decay iterations (held at lr_end afterwards).
    ``schedule(i)`` is the step of iteration i (0-based); ``key`` names the
    schedule (equal keys, equal steps), so that ``solve`` can keep the step
    tables it has built."""

    def __init__(self, lr0, n_flat, n_decay, lr_end=1e-3, n_warmup=0, lr_start=None):
        self.lr0, self.n_flat, self.lr_end, self.n_warmup = lr0, n_flat, lr_end, n_warmup
        self.d = (lr_end / lr0) ** (1.0 / max(1, n_decay))
        self.w = 1.0 if n_warmup <= 0 else (lr0 / (lr0 if lr_start is None else lr_start)) ** (1.0 / n_warmup)
        self.ls = lr0 if lr_start is None else lr_start
        self.key = ("adam_schedule", lr0, n_flat, n_decay, lr_end, n_warmup, self.ls if n_warmup > 0 else None)

    def __call__(self, i):
        if i < self.n_warmup:
            return self.ls * self.w ** i
        i -= self.n_warmup
        return self.lr0 if i < self.n_flat else max(self.lr_end, self.lr0 * self.d ** (i - self.n_flat))


def rel_err(U, V, Ur, Vr, mask):
    """Relative L2 velocity error ||(U,V) - (Ur,Vr)|| / ||(Ur,Vr)|| over the
    cells where ``mask`` holds (all layers), in float64."""
    d = torch.stack([U, V]).double() - torch.stack([Ur, Vr]).double()
    r = torch.stack([Ur, Vr]).double()
    m = mask.to(torch.float64)
    return float(torch.linalg.vector_norm(d * m) / torch.linalg.vector_norm(r * m))


def converged_solution(op32, op64, U0, V0, n_flat=8000, n_decay=20000, n_refine=2000, lr_end=1e-3,
                       epsilon=1e-7):
    """Float64 minimiser of one field: float32 Adam from (U0, V0) at lr 10
    for n_flat iterations, a geometric decay to ``lr_end`` over n_decay, then
    float64 Adam from that point with the step decaying lr_end -> lr_end/100
    (Adam's ``epsilon`` in both stages).

    Returns (U, V, COST, change) with ``change`` the relative difference the
    float64 stage made.  Fields whose velocities are weakly determined (thin
    ice) leave runs of different schedules ~1e-3 apart at energies equal to
    ~1e-7: compare two schedules to know how far a minimiser can be trusted.
    """
    r = solve(op32, U0, V0, n_flat + n_decay, LegacyAdamState(lr=1.0, epsilon=epsilon), early_stopping=False,
              lr_schedule=adam_schedule(10.0, n_flat, n_decay, lr_end))
    r2 = solve(op64, r.U.double(), r.V.double(), n_refine, LegacyAdamState(lr=1.0, epsilon=epsilon),
               early_stopping=False, lr_schedule=adam_schedule(lr_end, 0, n_refine, lr_end=lr_end / 100))
    change = rel_err(r.U, r.V, r2.U, r2.V, torch.ones_like(r.U[0], dtype=torch.bool))
    return r2.U, r2.V, float(r2.costs[-1]), change
