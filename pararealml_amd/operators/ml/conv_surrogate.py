"""A learned coarse propagator: a small convolutional network trained on
trajectories of the device FDM solver, with BF16 matrix-core training and
inference kernels.

The reference's learned operators are Keras regressors stepped
autoregressively by their time step (pararealml/operators/ml/supervised/
supervised_ml_operator.py:91-128, trained on oracle trajectories at
381-526); this module keeps that protocol -- one network application advances
the state by the operator's ``d_t`` -- with a different model:

* six 3x3 "same" convolutions (zero padding, NHWC BF16 activations, FP32
  accumulation): C -> 64 -> 64 -> 64 -> 64 -> 64 -> C, tanh-form GELU after
  the first five;
* the input is scaled by its per-sample maximum magnitude s, x = u / s, and
  the output is the increment step ``u + s net(x)``: the identity part stays
  in the state's FP32/FP64 precision and only the learned increment passes
  through BF16 (training minimises ``(net(x) - (target - u) / s)^2``);
* optionally the network works on the 2x-coarsened vertex grid (injection in,
  bilinear prolongation out).

The 64 -> 64 layers (99% of the FLOPs) run on ``pml_conv64`` / ``pml_wgrad64``
(v_mfma_f32_32x32x16_bf16); the thin first and last layers on VALU kernels;
Adam keeps FP32 master weights.  Every kernel is checked against the NumPy
oracle ``oracle/surrogate_numpy.py``.
"""

import ctypes
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from pararealml_amd import _native as nat

N_LAYERS = 6
HIDDEN = 64


def _sp(stream):
    return ctypes.c_void_p(stream.cuda_stream)


class ConvSurrogate:
    """Device-resident network parameters, optimiser state and workspaces."""

    def __init__(self, n_fields: int, seed: int = 0, stream=None):
        import torch

        if not 1 <= n_fields <= 3:
            raise ValueError("the surrogate covers 1 to 3 fields")
        self.lib = nat.device_lib()
        self.torch = torch
        self.n_fields = n_fields
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        self.chans = [n_fields] + [HIDDEN] * (N_LAYERS - 1) + [n_fields]
        self.step = 0
        # The network learns the normalised increment divided by out_scale
        # (set by fit() from the training data), so a small correction is
        # still an O(1) output.
        self.out_scale = 1.0
        rng = np.random.RandomState(seed)
        params = []
        for layer in range(N_LAYERS):
            ci, co = self.chans[layer], self.chans[layer + 1]
            w = rng.normal(0.0, math.sqrt(2.0 / (9 * ci)), (3, 3, ci, co))
            if layer == N_LAYERS - 1:
                w[:] = 0.0
            params.append((w, np.zeros(co)))
        self._alloc()
        self._ws_cache = {}
        self.set_params(params)

    # ------------------------------------------------------------ parameters
    def _alloc(self):
        t = self.torch
        f32 = dict(dtype=t.float32, device="cuda")
        bf = dict(dtype=t.bfloat16, device="cuda")
        self.w32, self.b32, self.mw, self.vw, self.mb, self.vb = [], [], [], [], [], []
        self.gw, self.gb, self.plain, self.wf, self.wb = [], [], [], [], []
        # all gradients are views of one buffer, cleared by one fill per step
        sizes = [(9 * self.chans[k] * self.chans[k + 1], self.chans[k + 1]) for k in range(N_LAYERS)]
        self.gflat = t.zeros(sum(n + co for n, co in sizes), **f32)
        off = 0
        for layer in range(N_LAYERS):
            ci, co = self.chans[layer], self.chans[layer + 1]
            n = 9 * ci * co
            for lst, shape in ((self.w32, n), (self.mw, n), (self.vw, n)):
                lst.append(t.zeros(shape, **f32))
            for lst in (self.b32, self.mb, self.vb):
                lst.append(t.zeros(co, **f32))
            self.gw.append(self.gflat[off:off + n])
            self.gb.append(self.gflat[off + n:off + n + co])
            off += n + co
            mid = ci == HIDDEN and co == HIDDEN
            self.plain.append(None if mid else t.zeros(n, **bf))
            self.wf.append(t.zeros(n, **bf) if mid else None)
            self.wb.append(t.zeros(n, **bf) if mid else None)

    def set_params(self, params: Sequence[Tuple[np.ndarray, np.ndarray]]):
        """Sets weights ((3, 3, ci, co) arrays) and biases; the BF16 copies
        are the round-to-nearest-even images of the FP32 masters."""
        t = self.torch
        for layer, (w, b) in enumerate(params):
            self.w32[layer].copy_(t.from_numpy(np.asarray(w, dtype=np.float32).ravel()))
            self.b32[layer].copy_(t.from_numpy(np.asarray(b, dtype=np.float32)))
        self._refresh_bf16()

    def _refresh_bf16(self):
        for layer in range(N_LAYERS):
            ci, co = self.chans[layer], self.chans[layer + 1]
            w = self.w32[layer].view(9, ci, co)
            wb16 = w.to(self.torch.bfloat16)
            if self.plain[layer] is not None:
                self.plain[layer].copy_(wb16.reshape(-1))
            else:
                self.wf[layer].copy_(wb16.permute(0, 2, 1).reshape(-1))
                self.wb[layer].copy_(wb16.flip(0).reshape(-1))

    def params(self) -> List[Tuple[np.ndarray, np.ndarray]]:
        out = []
        for layer in range(N_LAYERS):
            ci, co = self.chans[layer], self.chans[layer + 1]
            out.append(
                (
                    self.w32[layer].cpu().numpy().reshape(3, 3, ci, co).astype(np.float64),
                    self.b32[layer].cpu().numpy().astype(np.float64),
                )
            )
        return out

    def state_dict(self) -> dict:
        sd = {"n_fields": self.torch.tensor(self.n_fields), "step": self.torch.tensor(self.step),
              "out_scale": self.torch.tensor(self.out_scale, dtype=self.torch.float64)}
        for layer in range(N_LAYERS):
            sd[f"w{layer}"] = self.w32[layer].cpu()
            sd[f"b{layer}"] = self.b32[layer].cpu()
        return sd

    def load_state_dict(self, sd: dict):
        for layer in range(N_LAYERS):
            self.w32[layer].copy_(sd[f"w{layer}"].to("cuda"))
            self.b32[layer].copy_(sd[f"b{layer}"].to("cuda"))
        self.step = int(sd.get("step", 0))
        self.out_scale = float(sd.get("out_scale", 1.0))
        self._refresh_bf16()

    def save(self, path: str):
        self.torch.save(self.state_dict(), path)

    @classmethod
    def load(cls, path: str, stream=None) -> "ConvSurrogate":
        import torch

        sd = torch.load(path, weights_only=True, map_location="cpu")
        net = cls(int(sd["n_fields"]), stream=stream)
        net.load_state_dict(sd)
        return net

    # ------------------------------------------------------------ workspaces
    def _workspace(self, batch: int, h: int, w: int, training: bool, stream=None):
        st = stream if stream is not None else self.stream
        key = (batch, h, w, training, st.cuda_stream)
        if key in self._ws_cache:
            return self._ws_cache[key]
        t = self.torch
        with t.cuda.stream(st):  # the buffers belong to the stream that uses them
            ws = self._alloc_workspace(batch, h, w, training)
        if len(self._ws_cache) >= 4:
            self._ws_cache.clear()
        self._ws_cache[key] = ws
        return ws

    def _alloc_workspace(self, batch: int, h: int, w: int, training: bool):
        t = self.torch
        px = batch * h * w
        bf = dict(dtype=t.bfloat16, device="cuda")
        f32 = dict(dtype=t.float32, device="cuda")
        c = self.n_fields
        ws = {
            "x": t.empty(px * c, **f32),  # normalised input, FP32 (first layer)
            "s": t.empty(batch, **f32),
            "so": t.empty(batch, **f32),  # s * out_scale
            "y": t.empty(px * c, **f32),
            "a": [t.empty(px * HIDDEN, **bf) for _ in range(N_LAYERS - 1)],
        }
        if training:
            ws["z"] = [t.empty(px * HIDDEN, **bf) for _ in range(N_LAYERS - 1)]
            ws["dy"] = t.empty(px * c, **bf)
            ws["dz"] = [t.empty(px * HIDDEN, **bf) for _ in range(2)]
            ws["loss"] = t.zeros(1, dtype=t.float64, device="cuda")
            # per-workgroup weight-gradient sums, added in a fixed order:
            # training is bitwise reproducible for a given seed
            need = 0
            for layer in range(N_LAYERS):
                n = ctypes.c_int64(0)
                nat.check(self.lib.pml_grad_partials_floats(
                    self.chans[layer], self.chans[layer + 1], batch, h, w, ctypes.byref(n)))
                need = max(need, n.value)
            ws["gpart"] = t.empty(need, **f32)
        return ws

    def _out_s(self, ws, stream):
        """Per-sample output scale s * out_scale (the input scale s alone
        when out_scale is 1)."""
        if self.out_scale == 1.0:
            return ws["s"]
        with self.torch.cuda.stream(stream):
            self.torch.mul(ws["s"], self.out_scale, out=ws["so"])
        return ws["so"]

    # --------------------------------------------------------------- forward
    def _forward(self, ws, batch, h, w, store_z: bool, stream=None):
        lib, sp = self.lib, _sp(stream if stream is not None else self.stream)
        c = self.n_fields
        z = ws.get("z") if store_z else None
        # layer 0: C -> 64
        nat.check(lib.pml_conv_small(
            0, ws["x"].data_ptr(), self.w32[0].data_ptr(), self.b32[0].data_ptr(), None,
            z[0].data_ptr() if z else None, ws["a"][0].data_ptr(), None, c, HIDDEN,
            batch, h, w, sp))
        for layer in range(1, N_LAYERS - 1):
            nat.check(lib.pml_conv64(
                0, ws["a"][layer - 1].data_ptr(), self.wf[layer].data_ptr(),
                self.b32[layer].data_ptr(), None, z[layer].data_ptr() if z else None,
                ws["a"][layer].data_ptr(), batch, h, w, sp))
        nat.check(lib.pml_conv_small(
            0, ws["a"][N_LAYERS - 2].data_ptr(), self.plain[N_LAYERS - 1].data_ptr(),
            self.b32[N_LAYERS - 1].data_ptr(), None, None, None, ws["y"].data_ptr(), HIDDEN,
            c, batch, h, w, sp))

    def predict(self, u, out=None, steps: int = 1, stream=None):
        """Applies the network ``steps`` times to field-major states
        ``u`` (B, C, H, W), FP32 or FP64; returns a tensor of u's dtype.

        Runs on ``stream`` (default: the surrogate's own stream), which must
        be the stream that produced ``u`` and consumes the result; it first
        waits for the surrogate's stream (where the parameters are updated),
        and the surrogate's stream waits for it afterwards."""
        t = self.torch
        batch, c, h, w = u.shape
        if c != self.n_fields:
            raise ValueError(f"expected {self.n_fields} fields, got {c}")
        prec = nat.PML_F64 if u.dtype == t.float64 else nat.PML_F32
        st = stream if stream is not None else self.stream
        other = st.cuda_stream != self.stream.cuda_stream
        if other:
            st.wait_stream(self.stream)
        ws = self._workspace(batch, h, w, training=False, stream=st)
        with t.cuda.stream(st):
            cur = u.contiguous()
            res = out if out is not None else t.empty_like(cur)
        sp = _sp(st)
        for k in range(steps):
            nat.check(self.lib.pml_surrogate_normalize(
                prec, cur.data_ptr(), ws["x"].data_ptr(), ws["s"].data_ptr(), batch, c, h * w, sp))
            self._forward(ws, batch, h, w, store_z=False, stream=st)
            so = self._out_s(ws, st)
            nat.check(self.lib.pml_surrogate_residual(
                prec, cur.data_ptr(), ws["y"].data_ptr(), so.data_ptr(),
                res.data_ptr(), batch, c, h * w, sp))
            cur = res
        if other:
            self.stream.wait_stream(st)
        return res

    # -------------------------------------------------------------- training
    def loss_and_grads(self, u, target) -> float:
        """Forward + backward on a batch (u, target: FP32 field-major
        (B, C, H, W)); gradients are left in ``gw`` / ``gb``."""
        t = self.torch
        lib, sp = self.lib, _sp(self.stream)
        batch, c, h, w = u.shape
        ws = self._workspace(batch, h, w, training=True)
        u = u.contiguous().to(t.float32)
        target = target.contiguous().to(t.float32)
        nat.check(lib.pml_surrogate_normalize(
            nat.PML_F32, u.data_ptr(), ws["x"].data_ptr(), ws["s"].data_ptr(), batch, c, h * w, sp))
        self._forward(ws, batch, h, w, store_z=True)
        self.gflat.zero_()
        ws["loss"].zero_()
        so = self._out_s(ws, self.stream)
        nat.check(lib.pml_surrogate_loss(
            u.data_ptr(), ws["y"].data_ptr(), so.data_ptr(), target.data_ptr(),
            ws["dy"].data_ptr(), ws["loss"].data_ptr(), batch, c, h * w, sp))
        last = N_LAYERS - 1
        # layer 5 (64 -> C): weight / bias gradients and the input gradient
        gpart, gfloats = ws["gpart"].data_ptr(), ws["gpart"].numel()
        nat.check(lib.pml_conv_small_wgrad(
            ws["a"][last - 1].data_ptr(), ws["dy"].data_ptr(), self.gw[last].data_ptr(),
            self.gb[last].data_ptr(), HIDDEN, c, batch, h, w, gpart, gfloats, sp))
        dz_cur, dz_next = ws["dz"]
        nat.check(lib.pml_conv_small(
            1, ws["dy"].data_ptr(), self.plain[last].data_ptr(), None,
            ws["z"][last - 1].data_ptr(), None, dz_cur.data_ptr(), None, HIDDEN, c,
            batch, h, w, sp))
        # layers 4 .. 1 (64 -> 64)
        for layer in range(last - 1, 0, -1):
            nat.check(lib.pml_wgrad64(
                ws["a"][layer - 1].data_ptr(), dz_cur.data_ptr(), self.gw[layer].data_ptr(),
                self.gb[layer].data_ptr(), batch, h, w, gpart, gfloats, sp))
            nat.check(lib.pml_conv64(
                1, dz_cur.data_ptr(), self.wb[layer].data_ptr(), None,
                ws["z"][layer - 1].data_ptr(), None, dz_next.data_ptr(), batch, h, w, sp))
            dz_cur, dz_next = dz_next, dz_cur
        # layer 0 (C -> 64)
        nat.check(lib.pml_conv_small_wgrad(
            ws["x"].data_ptr(), dz_cur.data_ptr(), self.gw[0].data_ptr(),
            self.gb[0].data_ptr(), c, HIDDEN, batch, h, w, gpart, gfloats, sp))
        return float(ws["loss"].item()) / (batch * c * h * w)

    def adam_step(self, lr: float, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        lib, sp = self.lib, _sp(self.stream)
        for layer in range(N_LAYERS):
            ci, co = self.chans[layer], self.chans[layer + 1]
            nat.check(lib.pml_adam(
                self.w32[layer].data_ptr(), self.mw[layer].data_ptr(), self.vw[layer].data_ptr(),
                self.gw[layer].data_ptr(), 9 * ci * co, lr, beta1, beta2, eps, self.step,
                None if self.plain[layer] is None else self.plain[layer].data_ptr(),
                None if self.wf[layer] is None else self.wf[layer].data_ptr(),
                None if self.wb[layer] is None else self.wb[layer].data_ptr(),
                ci, co, sp))
            nat.check(lib.pml_adam(
                self.b32[layer].data_ptr(), self.mb[layer].data_ptr(), self.vb[layer].data_ptr(),
                self.gb[layer].data_ptr(), co, lr, beta1, beta2, eps, self.step,
                None, None, None, 0, 0, sp))

    def fit(self, inputs, targets, steps: int, batch: int = 32, lr: float = 1e-3,
            seed: int = 0, log_every: int = 0, scale_output: bool = True,
            mix: float = 0.0) -> List[float]:
        """Adam on random minibatches of (inputs, targets) (device tensors
        (N, C, H, W)), cosine-decayed learning rate; returns the losses (in
        units of the output scale).  With ``scale_output``, an untrained
        network first sets ``out_scale`` to the RMS of the normalised
        increments (t - u) / max|u| over up to 256 pairs.

        ``mix`` (for maps that are linear in the state, such as the
        propagators of a linear equation): that fraction of every minibatch
        is replaced by linear combinations ``u_a + beta u_b`` of two pairs
        (targets combined alike, beta uniform in [-1.25, 1.25]), so the
        network also learns how the map responds to differences of states,
        which is what Parareal's correction G(U_{k+1}) - G(U_k) evaluates."""
        t = self.torch
        if not (bool(t.isfinite(inputs).all()) and bool(t.isfinite(targets).all())):
            raise ValueError("training pairs hold non-finite values (an unstable base or data "
                             "solve?)")
        gen = t.Generator(device="cpu")
        gen.manual_seed(seed)
        n = inputs.shape[0]
        if scale_output and self.step == 0:
            idx = t.randperm(n, generator=gen)[:256].to(inputs.device)
            u, tg = inputs[idx].double(), targets[idx].double()
            sc = u.abs().amax(dim=(1, 2, 3), keepdim=True).clamp_min(1e-30)
            rms = float(((tg - u) / sc).pow(2).mean().sqrt())
            self.out_scale = rms if rms > 0.0 and math.isfinite(rms) else 1.0
        losses = []
        n_mix = int(round(batch * mix))
        for k in range(steps):
            idx = t.randint(0, n, (batch,), generator=gen).to(inputs.device)
            x, y = inputs[idx], targets[idx]
            if n_mix:
                jdx = t.randint(0, n, (n_mix,), generator=gen).to(inputs.device)
                beta = (2.5 * t.rand((n_mix, 1, 1, 1), generator=gen) - 1.25).to(x)
                x = t.cat([x[n_mix:], x[:n_mix] + beta * inputs[jdx]])
                y = t.cat([y[n_mix:], y[:n_mix] + beta * targets[jdx]])
            loss = self.loss_and_grads(x.contiguous(), y.contiguous())
            cur_lr = lr * 0.5 * (1.0 + math.cos(math.pi * k / max(steps, 1)))
            self.adam_step(cur_lr)
            losses.append(loss)
            if log_every and (k + 1) % log_every == 0:
                print(f"surrogate step {k + 1}/{steps} loss {loss:.3e}", flush=True)
        return losses

    def fit_trajectories(self, snaps, steps: int, batch: int = 32, lr: float = 1e-3,
                         pushforward: float = 0.5, max_push: int = 1, seed: int = 0,
                         log_every: int = 0) -> List[float]:
        """Adam on trajectories ``snaps`` (device tensor (T, N, C, H, W),
        consecutive states one network step apart).  A ``pushforward``
        fraction of each minibatch starts from the network's own m-step
        prediction of a true state (m drawn from 1..max_push per minibatch)
        and is trained towards the true state one step further (the
        predictions are treated as data, no gradient flows through them),
        which teaches the network to damp its own errors under repeated
        application; the rest are plain one-step pairs.  Cosine-decayed
        learning rate; returns the losses."""
        t = self.torch
        gen = t.Generator(device="cpu")
        gen.manual_seed(seed)
        n_t, n_traj = snaps.shape[0], snaps.shape[1]
        max_push = max(1, min(int(max_push), n_t - 2))
        if n_t < 3 and pushforward > 0.0:
            raise ValueError("pushforward needs trajectories of at least 3 states")
        n_push = int(round(batch * pushforward))
        losses = []
        for k in range(steps):
            tr = t.randint(0, n_traj, (batch,), generator=gen)
            m = int(t.randint(1, max_push + 1, (1,), generator=gen))
            t1 = t.randint(0, n_t - 1, (batch - n_push,), generator=gen)
            t2 = t.randint(0, n_t - 1 - m, (n_push,), generator=gen)
            parts_in, parts_tg = [], []
            if batch - n_push:
                parts_in.append(snaps[t1, tr[n_push:]])
                parts_tg.append(snaps[t1 + 1, tr[n_push:]])
            if n_push:
                cur = snaps[t2, tr[:n_push]]
                for _ in range(m):
                    cur = self.predict(cur).to(t.float32)
                parts_in.append(cur)
                parts_tg.append(snaps[t2 + m + 1, tr[:n_push]])
            loss = self.loss_and_grads(t.cat(parts_in), t.cat(parts_tg))
            cur_lr = lr * 0.5 * (1.0 + math.cos(math.pi * k / max(steps, 1)))
            self.adam_step(cur_lr)
            losses.append(loss)
            if log_every and (k + 1) % log_every == 0:
                print(f"surrogate step {k + 1}/{steps} loss {loss:.3e}", flush=True)
        return losses
