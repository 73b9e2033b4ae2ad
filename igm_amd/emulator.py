"""CNN emulator of the ice flow on hand-written MFMA kernels.

Network (emulate/neural_network.py:9-66, default cfg): nb_layers x
[Conv2D 3x3 "same" + bias -> LeakyReLU(0.01)] -> Conv2D 1x1 + bias, mapping
X = [thk, usurf, arrhenius, slidingco, dX] (raw, iceflow/utils.py:147-158)
to Y = [U_0..U_{Nz-1}, V_0..V_{Nz-1}] (iceflow/utils.py:133-139).

Precision: layer 1 in fp32 (raw elevations need it), layers 2..nb_layers in
bf16 on MFMA with fp32 accumulation, output layer on MFMA with fp32 output;
fp32 master weights and legacy-Adam slots.  Training uses the energy itself
as the loss (emulate/emulate.py:155-281): Y -> fused energy/adjoint kernel ->
dJ/dY -> head backward -> per-layer MFMA wgrad + dgrad -> layer-1 wgrad ->
per-tensor clipnorm + Adam.

Process-module functions (same names as emulate/emulate.py):
initialize_iceflow_emulator, update_iceflow_emulated,
update_iceflow_emulator, split_into_patches, save_iceflow_model.
"""
import ctypes as C
import os

import numpy as np
import torch

from . import _lib as L
from .energy import EnergyOperator
from .fields import velocity_outputs

HERE = os.path.dirname(os.path.abspath(__file__))
EMULATOR_DIR = os.path.join(HERE, "emulators")
NF = 32   # filters: the kernels are specialised for 32
NIN = 5   # input fields


def glorot_uniform(shape, gen):
    kh, kw, ci, co = shape
    lim = np.sqrt(6.0 / (kh * kw * ci + kh * kw * co))
    return (torch.rand(shape, generator=gen, dtype=torch.float64) * 2 - 1) * lim


class CNN:
    """Parameters of the emulator (flat fp32 masters + Adam slots + bf16 operands)."""

    def __init__(self, nz, nb_layers=16, device=None, seed=0, weights=None):
        if nb_layers < 1:
            raise ValueError("nb_layers must be >= 1")
        if 2 * nz > NF:
            raise NotImplementedError("2*Nz must be <= 32 for the output layer kernel")
        self.nz, self.nl, self.cout = nz, nb_layers, 2 * nz
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        shapes = [(3, 3, NIN, NF)] + [(3, 3, NF, NF)] * (nb_layers - 1) + [(1, 1, NF, self.cout)]
        self.shapes = []
        for s in shapes:
            self.shapes += [s, (s[-1],)]
        sizes = [int(np.prod(s)) for s in self.shapes]
        self.offs_host = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.n = int(self.offs_host[-1])
        dev = self.device
        self.w = torch.zeros(self.n, dtype=torch.float32, device=dev)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.g = torch.zeros_like(self.w)
        self.offs = torch.as_tensor(self.offs_host, device=dev)
        # offsets of the 3x3 MFMA kernels (layers 1..nl-1); None when there are none
        self.offs3x3 = (torch.as_tensor(self.offs_host[2:2 * nb_layers:2], device=dev)
                        if nb_layers > 1 else None)
        self.sq = torch.zeros(len(self.shapes), dtype=torch.float32, device=dev)
        self.fwd = torch.empty((max(1, nb_layers - 1), 9, NF, NF), dtype=torch.bfloat16, device=dev)
        self.dgr = torch.empty_like(self.fwd)
        self.head = torch.empty((NF, NF), dtype=torch.bfloat16, device=dev)
        self.iterations = 0
        if weights is None:
            gen = torch.Generator().manual_seed(seed)
            weights = []
            for i in range(0, len(self.shapes), 2):
                weights.append((glorot_uniform(self.shapes[i], gen).numpy(), np.zeros(self.shapes[i + 1])))
        self.set_weights(weights)

    # views into the flat vectors
    def view(self, flat, i):
        return flat[self.offs_host[i]:self.offs_host[i + 1]].view(self.shapes[i])

    def kernel(self, layer, flat=None):
        return self.view(self.w if flat is None else flat, 2 * layer)

    def bias(self, layer, flat=None):
        return self.view(self.w if flat is None else flat, 2 * layer + 1)

    def set_weights(self, weights):
        assert len(weights) == self.nl + 1, "layer count mismatch"
        for li, (k, b) in enumerate(weights):
            k = np.asarray(k, np.float32)
            assert k.shape == self.shapes[2 * li], (k.shape, self.shapes[2 * li])
            self.kernel(li).copy_(torch.as_tensor(k))
            self.bias(li).copy_(torch.as_tensor(np.asarray(b, np.float32)))
        self.pack()

    def get_weights(self):
        return [(self.kernel(i).cpu().numpy(), self.bias(i).cpu().numpy()) for i in range(self.nl + 1)]

    def pack(self):
        L.call("igm_pack_conv_weights", L.ptr(self.w), L.ptr(self.offs3x3), self.nl - 1,
               L.ptr(self.kernel(self.nl)), self.cout, L.ptr(self.fwd), L.ptr(self.dgr),
               L.ptr(self.head), L.stream_ptr())

    # ------------------------------------------------------------ forward --
    def forward(self, fields, extents=None, keep=False, B=None):
        """fields: (f0..f4 device fp32 tensors, (fb, fr, fc) strides, B, H, W).
        Returns Y (B, H, W, 2Nz) fp32 (and the activations when keep)."""
        fptrs, (fb, fr, fc), B, H, W = fields
        dev = self.device
        ext = L.ptr(extents)
        st = L.stream_ptr()
        act = torch.empty((B, H, W, NF), dtype=torch.bfloat16, device=dev)
        L.call("igm_conv_in_fwd", *fptrs, fb, fr, fc, ext, B, H, W, L.ptr(self.kernel(0)),
               L.ptr(self.bias(0)), L.ptr(act), st)
        acts = [act]
        spare = None
        for li in range(1, self.nl):
            if keep or spare is None:
                out = torch.empty_like(act)
            else:
                out = spare
            L.call("igm_conv3x3_fwd_bf16", L.ptr(act), L.ptr(self.fwd[li - 1]), L.ptr(self.bias(li)),
                   ext, B, H, W, L.ptr(out), st)
            if keep:
                acts.append(out)
            else:
                spare = act
            act = out
        Y = torch.empty((B, H, W, self.cout), dtype=torch.float32, device=dev)
        L.call("igm_conv1x1_head_fwd", L.ptr(act), L.ptr(self.head), L.ptr(self.bias(self.nl)), ext,
               B, H, W, self.cout, L.ptr(Y), st)
        return (Y, acts) if keep else Y

    # ----------------------------------------------------------- backward --
    def backward(self, dY, acts, fields, extents=None, nblocks=512, accumulate=False, input_grad=None,
                 deferred=False):
        """Gradient of sum(dY * Y) w.r.t. all parameters, into self.g (added
        to it when ``accumulate``).  ``input_grad`` (optional, fp32 (B, H, W,
        5)) receives the gradient w.r.t. the input fields, the quantity data
        assimilation differentiates through the emulator.

        ``deferred``: for one batch split into sub-batches, each sub-batch's
        weight gradients stay as per-block partial sums in persistent
        per-layer buffers (``accumulate``: added to the previous
        sub-batches'), and ``reduce_deferred()`` sums them into self.g once:
        one reduction per layer per batch instead of one per sub-batch."""
        acc = 1 if accumulate else 0
        fptrs, (fb, fr, fc), B, H, W = fields
        ext = L.ptr(extents)
        st = L.stream_ptr()
        lib = L.load()
        if deferred:
            part = self._deferred_partials(nblocks)

            def wgrad(li, taps, a, d, cout):
                L.call("igm_conv_wgrad_partials_bf16", taps, L.ptr(a), L.ptr(d), ext, B, H, W, nblocks,
                       L.ptr(part[li]), acc, st)

            def wgrad_in(d):
                L.call("igm_conv_in_wgrad_partials", *fptrs, fb, fr, fc, ext, B, H, W, L.ptr(d), nblocks,
                       L.ptr(part[0]), acc, st)
        else:
            ws_n = max(lib.igm_wgrad_workspace_floats(9, nblocks), lib.igm_wgrad_workspace_floats(0, nblocks))
            ws = torch.empty(ws_n, dtype=torch.float32, device=self.device)

            def wgrad(li, taps, a, d, cout):
                L.call("igm_conv_wgrad_bf16", taps, L.ptr(a), L.ptr(d), ext, B, H, W, cout, nblocks,
                       L.ptr(ws), acc, L.ptr(self.kernel(li, self.g)), L.ptr(self.bias(li, self.g)), st)

            def wgrad_in(d):
                L.call("igm_conv_in_wgrad", *fptrs, fb, fr, fc, ext, B, H, W, L.ptr(d), nblocks, L.ptr(ws),
                       acc, L.ptr(self.kernel(0, self.g)), L.ptr(self.bias(0, self.g)), st)
        nl = self.nl
        dz = torch.empty((B, H, W, NF), dtype=torch.bfloat16, device=self.device)
        dYb = torch.empty_like(dz)
        L.call("igm_conv1x1_head_bwd", L.ptr(dY), L.ptr(acts[nl - 1]), L.ptr(self.kernel(nl)), ext, B, H,
               W, self.cout, L.ptr(dz), L.ptr(dYb), st)
        wgrad(nl, 1, acts[nl - 1], dYb, self.cout)
        below = dYb  # reuse as the ping-pong buffer
        for li in range(nl - 1, 0, -1):  # layer index li (0-based) is a 3x3 MFMA layer
            a_in = acts[li - 1]
            wgrad(li, 9, a_in, dz, NF)
            L.call("igm_conv3x3_dgrad_bf16", L.ptr(dz), L.ptr(self.dgr[li - 1]), L.ptr(a_in), ext, B, H,
                   W, L.ptr(below), st)
            dz, below = below, dz
        wgrad_in(dz)
        if input_grad is not None:
            assert input_grad.shape == (B, H, W, 5) and input_grad.dtype == torch.float32
            assert input_grad.is_contiguous()
            L.call("igm_conv_in_dgrad", L.ptr(dz), L.ptr(self.kernel(0)), ext, B, H, W, L.ptr(input_grad), st)
        return self.g

    def _deferred_partials(self, nblocks):
        """Persistent per-layer partial-sum buffers of the deferred backward
        (layer 0: layer 1's layout; 1..nl-1: 3x3; nl: the 1x1 head)."""
        if getattr(self, "_part_nblocks", None) != nblocks:
            lib = L.load()
            n = {0: lib.igm_wgrad_workspace_floats(0, nblocks), 9: lib.igm_wgrad_workspace_floats(9, nblocks),
                 1: lib.igm_wgrad_workspace_floats(1, nblocks)}
            self._part = [torch.empty(n[0 if li == 0 else (1 if li == self.nl else 9)], dtype=torch.float32,
                                      device=self.device) for li in range(self.nl + 1)]
            self._part_nblocks = nblocks
        return self._part

    def reduce_deferred(self, accumulate=False):
        """Sum the deferred partials of every layer into self.g (added to it
        when ``accumulate``), one fixed-order reduction per layer; the
        partials are consumed."""
        nb = self._part_nblocks
        st = L.stream_ptr()
        for li in range(self.nl + 1):
            taps = 0 if li == 0 else (1 if li == self.nl else 9)
            cout = self.cout if li == self.nl else NF
            L.call("igm_conv_wgrad_reduce_f32", taps, cout, nb, L.ptr(self._part[li]), 1 if accumulate else 0,
                   L.ptr(self.kernel(li, self.g)), L.ptr(self.bias(li, self.g)), st)
        return self.g

    # ---------------------------------------------------------- optimizer --
    def adam_step(self, lr, clipnorm=1.0, epsilon=1e-7, beta1=0.9, beta2=0.999):
        """Legacy Keras Adam (tf.keras.optimizers.legacy.Adam, emulate.py:24-35)
        with per-tensor clip_by_norm, on self.g; then repack the bf16 operands."""
        t = self.iterations + 1
        f = np.float32
        b1p = f(np.power(f(beta1), f(t)))
        b2p = f(np.power(f(beta2), f(t)))
        alpha = float(f(lr) * np.sqrt(f(1) - b2p) / (f(1) - b1p))
        adam = L.IgmAdam(alpha=alpha, beta1=beta1, beta2=beta2, epsilon=epsilon)
        L.call("igm_adam_clip_f32", L.ptr(self.w), L.ptr(self.g), L.ptr(self.m), L.ptr(self.v), self.n,
               L.ptr(self.offs), len(self.shapes), L.ptr(self.sq), float(clipnorm),
               C.addressof(adam), L.stream_ptr())
        self.iterations += 1
        self.pack()

    # -------------------------------------------------------------- files --
    def save(self, path, fieldin, nz, vert_spacing):
        from safetensors.numpy import save_file
        tensors = {}
        for i, (k, b) in enumerate(self.get_weights()):
            tensors[f"kernel_{i}"] = np.ascontiguousarray(k)
            tensors[f"bias_{i}"] = np.ascontiguousarray(b)
        save_file(tensors, path, metadata={"fieldin": ",".join(fieldin), "architecture": "cnn",
                                            "nb_layers": str(self.nl), "Nz": str(nz),
                                            "vert_spacing": str(vert_spacing)})


def load_weights(path):
    """[(kernel, bias), ...] and the metadata of a safetensors emulator file."""
    from safetensors import safe_open
    with safe_open(path, "np") as f:
        meta = f.metadata() or {}
        n = len([k for k in f.keys() if k.startswith("kernel_")])
        weights = [(f.get_tensor(f"kernel_{i}"), f.get_tensor(f"bias_{i}")) for i in range(n)]
    return weights, meta


class Fields(tuple):
    """(pointers of the 5 input fields, (fb, fr, fc) strides, B, H, W) that
    also holds references to the tensors, so they outlive every kernel that
    reads them through the raw pointers."""

    def __new__(cls, keep, ptrs, strides, B, H, W):
        obj = super().__new__(cls, (ptrs, strides, B, H, W))
        obj.keep = keep
        return obj


def planar_fields(tensors):
    """Descriptor for 5 planar (Ny, Nx) or (B, Ny, Nx) fp32 device tensors."""
    ts = [t if t.dim() == 3 else t[None] for t in tensors]
    ts = [t.to(torch.float32).contiguous() for t in ts]
    assert all(t.shape == ts[0].shape and t.stride() == ts[0].stride() for t in ts)
    B, H, W = ts[0].shape
    return Fields(ts, [t.data_ptr() for t in ts], tuple(ts[0].stride()), B, H, W)


def nhwc_fields(X):
    """Descriptor for X (B, H, W, 5) fp32 contiguous."""
    assert X.dim() == 4 and X.shape[-1] == NIN and X.is_contiguous() and X.dtype == torch.float32
    es = X.element_size()
    B, H, W, _ = X.shape
    return Fields([X], [X.data_ptr() + f * es for f in range(NIN)], (X.stride(0), X.stride(1), X.stride(2)),
                  B, H, W)


# ------------------------------------------------------------------------
# process-module functions (emulate/emulate.py)

def pretrained_name(cfg):
    """emulate.py:37-62."""
    ice = cfg.processes.iceflow
    L_ = "e" if ice.numerics.vert_basis == "Legendre" else "a"
    return ("pinnbp_" + str(ice.numerics.Nz) + "_" + str(int(ice.numerics.vert_spacing)) + "_"
            + ice.emulator.network.architecture + "_" + str(ice.emulator.network.nb_layers) + "_"
            + str(ice.emulator.network.nb_out_filter) + "_" + str(ice.physics.dim_arrhenius) + "_"
            + str(int(1)) + "_" + L_)


def initialize_iceflow_emulator(cfg, state):
    """emulate.py:22-104: optimizer settings, pretrained weights or a fresh CNN."""
    ice = cfg.processes.iceflow
    em = ice.emulator
    if int(em.network.nb_out_filter) != NF or list(em.fieldin) != ["thk", "usurf", "arrhenius", "slidingco", "dX"]:
        raise NotImplementedError("the emulator kernels need 32 filters and the 5 default fieldin")
    if int(em.network.conv_ker_size) != 3 or em.network.activation != "LeakyReLU" or em.network.dropout_rate > 0:
        raise NotImplementedError("the emulator kernels need 3x3 kernels, LeakyReLU and no dropout")
    state.opti_retrain = dict(lr=float(em.lr), epsilon=float(em.optimizer_epsilon),
                              clipnorm=float(em.optimizer_clipnorm))
    nz, nl = int(ice.numerics.Nz), int(em.network.nb_layers)
    weights = None
    if em.pretrained:
        path = None
        if em.name == "":
            cand = os.path.join(EMULATOR_DIR, pretrained_name(cfg) + ".safetensors")
            if os.path.exists(cand):
                path = cand
                print("Found pretrained emulator in the igm_amd package: " + pretrained_name(cfg))
            else:
                print("No pretrained emulator found in the igm_amd package")
        else:
            base = getattr(state, "original_cwd", os.getcwd())
            cand = em.name if os.path.isabs(em.name) else os.path.join(base, em.name)
            if os.path.isdir(cand) and os.path.exists(os.path.join(cand, "model.h5")):
                from .keras_h5 import load_keras_cnn
                weights, fieldin = load_keras_cnn(cand)
                assert list(em.fieldin) == fieldin  # emulate.py:89
            elif os.path.exists(cand):
                path = cand
            else:
                print("----------------------------------> No pretrained emulator found ")
        if path is not None:
            weights, meta = load_weights(path)
            if meta.get("fieldin"):
                assert list(em.fieldin) == meta["fieldin"].split(",")
    else:
        print("----------------------------------> No pretrained emulator, start from scratch.")
    state.iceflow_model = CNN(nz, nl, weights=weights)


def update_iceflow_emulated(cfg, state):
    """emulate.py:116-152: Y = model(X) -> mask thk>0 -> U, V and derived fields."""
    nz = int(cfg.processes.iceflow.numerics.Nz)
    fields = planar_fields([state.thk, state.usurf, state.arrhenius, state.slidingco, state.dX])
    Y = state.iceflow_model.forward(fields)
    out = velocity_outputs(Y, None, state.thk, state.vert_weight, nz)
    for k, v in out.items():
        setattr(state, k, v)
    # the unmasked network output (what the energy-gap check evaluates)
    state.iceflow_Y = Y
    state.iceflow_unmasked = None
    if cfg.processes.iceflow.force_max_velbar > 0:
        raise NotImplementedError("force_max_velbar is not supported")


def split_into_patches(Ny, Nx, nbmax):
    """emulate.py:283-308 ("sequential"): patch windows (j0, j1, i0, i1); the
    remainder rows/columns of the integer division are dropped, as there."""
    sy = Ny // nbmax + 1
    sx = Nx // nbmax + 1
    ly = int(Ny / sy)
    lx = int(Nx / sx)
    return [(j * ly, (j + 1) * ly, i * lx, (i + 1) * lx) for i in range(sx) for j in range(sy)]


def retrain_lr(lr, lr_decay, i):
    """emulate.py:209-210: learning rate of patch i of a retraining epoch,
    lr * lr_decay^(i / 1000) when lr_decay < 1."""
    return lr * (lr_decay ** (i / 1000)) if lr_decay < 1 else lr


def retrain_step(model, X, op, lr, clipnorm, epsilon, extents=None):
    """One energy-as-loss step on a batch X (B, H, W, 5): returns COST (sum
    over the batch entries), the global gradient norm is left in model.g."""
    fields = nhwc_fields(X)
    Y, acts = model.forward(fields, extents=extents, keep=True)
    e, gY = op.energy_grad(Y)
    model.backward(gY, acts, fields, extents=extents)
    model.adam_step(lr, clipnorm=clipnorm, epsilon=epsilon)
    return e


def update_iceflow_emulator(cfg, state, it, pertubate=False):
    """emulate.py:155-281: retrain every retrain_freq iterations (or during the
    warm-up) with the energy as the loss, patch by patch."""
    em = cfg.processes.iceflow.emulator
    run_it = False
    if em.retrain_freq > 0:
        run_it = (it % em.retrain_freq == 0)
    warm_up = int(it <= em.warm_up_it)
    if not (warm_up or run_it):
        return
    if pertubate:
        raise NotImplementedError("pertubate is not supported")
    if em.split_patch_method != "sequential":
        raise NotImplementedError("only split_patch_method 'sequential' is supported")
    state.COST_EMULATOR = []
    state.GRAD_EMULATOR = []
    X = torch.stack([state.thk, state.usurf, state.arrhenius, state.slidingco, state.dX], dim=-1)
    Ny, Nx = state.thk.shape
    wins = split_into_patches(Ny, Nx, int(em.framesizemax))
    if warm_up:
        nbit, lr = int(em.nbit_init), float(em.lr_init)
    else:
        nbit, lr = int(em.nbit), float(em.lr)
    model = state.iceflow_model
    opt = state.opti_retrain
    patches = [X[j0:j1, i0:i1].contiguous()[None] for (j0, j1, i0, i1) in wins]
    ops = [EnergyOperator(np.asarray(state.dzeta), cfg.processes.iceflow.physics, X=p) for p in patches]
    for epoch in range(nbit):
        cost_emulator = 0.0
        for i, (p, op) in enumerate(zip(patches, ops)):
            e = retrain_step(model, p, op, retrain_lr(lr, float(em.lr_decay), i), opt["clipnorm"],
                             opt["epsilon"])
            cost_emulator = cost_emulator + e[0, 3]
        state.COST_EMULATOR.append(float(cost_emulator))
        state.GRAD_EMULATOR.append(float(torch.linalg.vector_norm(model.g)))
    if len(em.save_cost) > 0:
        np.savetxt(em.save_cost + "-" + str(it) + ".dat",
                   np.array(list(zip(state.COST_EMULATOR, state.GRAD_EMULATOR))), fmt="%5.10f")


def save_iceflow_model(cfg, state, directory="iceflow-model"):
    """emulate.py:331-359: the weights as a Keras 2 HDF5 model file
    (``model.h5``, which the reference's ``tf.keras.models.load_model`` reads)
    and as safetensors (``model.safetensors``, this package's own format),
    ``fieldin.dat`` and ``vert_grid.dat``."""
    import shutil
    from .keras_h5 import write_keras_cnn
    ice = cfg.processes.iceflow
    net = ice.emulator.network
    if os.path.exists(directory):
        shutil.rmtree(directory)
    os.mkdir(directory)
    state.iceflow_model.save(os.path.join(directory, "model.safetensors"), list(ice.emulator.fieldin),
                             int(ice.numerics.Nz), float(ice.numerics.vert_spacing))
    write_keras_cnn(os.path.join(directory, "model.h5"), state.iceflow_model.get_weights(),
                    activation=net.activation, dropout_rate=float(net.dropout_rate),
                    weight_initialization=net.weight_initialization)
    with open(os.path.join(directory, "fieldin.dat"), "w") as fid:
        for key in ice.emulator.fieldin:
            fid.write("%s \n" % (key))
    with open(os.path.join(directory, "vert_grid.dat"), "w") as fid:
        fid.write("%4.0f  %s \n" % (ice.numerics.Nz, "# number of vertical grid point (Nz)"))
        fid.write("%2.2f  %s \n" % (ice.numerics.vert_spacing, "# param for vertical spacing (vert_spacing)"))
