"""ORACLE -- test infrastructure only (tests/ may import it; nothing in the package does).

NumPy restatement of the warm-start U-Net computed by platoanalyze_amd/csrc/pa_unet.hip:
same layer table, links, parameter layout and rounding points.  There is no reference
counterpart (the reference solve always starts from zero), so this is pinned by
  * stride-1 direct convolution == scipy.ndimage.correlate with zero padding,
  * stride-2 and transposed links against their defining index relations (brute force),
  * backward (data and weight gradients) == central differences of the forward in FP64,
and the GPU kernels are then checked against it with BF16 rounding applied at the
points where the kernels round (weights and activations to BF16, accumulation wide).
"""
from __future__ import annotations

import numpy as np

IN_CH, OUT_CH = 8, 16


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16, (u >> 16) << 16)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def levels(nx, ny, nz):
    g = [(nx, ny, nz)]
    for _ in range(3):
        g.append(tuple((n + 1) // 2 for n in g[-1]))
    return g


def layer_table(C):
    """(input buffer, col, cin, output buffer, col, cout, stride, transposed, relu); the
    buffer numbering and levels of pa_unet.hip's build_layers."""
    return [
        (0, 0, IN_CH, 1, 0, C, 1, 0, 1), (1, 0, C, 2, 0, C, 1, 0, 1), (2, 0, C, 3, 0, 2 * C, 2, 0, 1),
        (3, 0, 2 * C, 4, 0, 2 * C, 1, 0, 1), (4, 0, 2 * C, 5, 0, 4 * C, 2, 0, 1),
        (5, 0, 4 * C, 6, 0, 4 * C, 1, 0, 1), (6, 0, 4 * C, 7, 0, 8 * C, 2, 0, 1),
        (7, 0, 8 * C, 8, 0, 8 * C, 1, 0, 1), (8, 0, 8 * C, 6, 4 * C, 4 * C, 2, 1, 1),
        (6, 0, 8 * C, 9, 0, 4 * C, 1, 0, 1), (9, 0, 4 * C, 4, 2 * C, 2 * C, 2, 1, 1),
        (4, 0, 4 * C, 10, 0, 2 * C, 1, 0, 1), (10, 0, 2 * C, 2, C, C, 2, 1, 1),
        (2, 0, 2 * C, 11, 0, C, 1, 0, 1), (11, 0, C, -1, 0, OUT_CH, 1, 0, 0),
    ]


BUF_LEVEL = [0, 0, 0, 1, 1, 2, 2, 3, 3, 2, 1, 0]


def buf_channels(C):
    return [IN_CH, C, 2 * C, 2 * C, 4 * C, 4 * C, 8 * C, 8 * C, 8 * C, 4 * C, 2 * C, C]


def num_params(C):
    return sum(27 * l[2] * l[5] + l[5] for l in layer_table(C))


def links(gi, go, stride, transposed):
    """For each tap t: (output voxel indices, input voxel indices) of the valid links."""
    o = np.indices(go).reshape(3, -1)
    out = []
    for t in range(27):
        d = np.array([t // 9, (t // 3) % 3, t % 3]).reshape(3, 1)
        if not transposed:
            i = stride * o + d - 1
            ok = np.ones(o.shape[1], bool)
        else:
            i = o + 1 - d
            ok = np.all(i >= 0, axis=0) & np.all(i % stride == 0, axis=0)
            i = i // stride
        ok &= np.all(i >= 0, axis=0) & np.all(i < np.array(gi).reshape(3, 1), axis=0)
        ov = np.ravel_multi_index(o[:, ok], go)
        iv = np.ravel_multi_index(i[:, ok], gi)
        out.append((ov, iv))
    return out


class UNetRef:
    """exact=False: the GPU's rounding points (BF16 operands, float32 storage of fp32
    tensors); exact=True: float64 throughout (for finite-difference checks)."""

    def __init__(self, grid, C, exact=False):
        self.g = levels(*grid)
        self.C = C
        self.exact = exact
        self.L = layer_table(C)
        self.ch = buf_channels(C)
        self.links = [links(self.g[BUF_LEVEL[l[0]]], self.g[BUF_LEVEL[l[3]]] if l[3] >= 0 else self.g[0], l[6], l[7])
                      for l in self.L]

    def _rb(self, x):  # operand rounding
        return x.astype(np.float64) if self.exact else bf16(x).astype(np.float64)

    def _rf(self, x):  # fp32 storage
        return x if self.exact else x.astype(np.float32).astype(np.float64)

    def split(self, p):
        out, off = [], 0
        for l in self.L:
            cin, cout = l[2], l[5]
            W = p[off:off + 27 * cin * cout].reshape(27, cin, cout)
            off += 27 * cin * cout
            b = p[off:off + cout]
            off += cout
            out.append((W, b))
        return out

    def forward(self, x, p):
        """x: [B, V0, 8] -> (output [B, V0, 8], activations)."""
        B = x.shape[0]
        acts = [None] * len(self.ch)
        acts[0] = self._rb(x)
        for b in range(1, len(self.ch)):
            acts[b] = np.zeros((B, int(np.prod(self.g[BUF_LEVEL[b]])), self.ch[b]))
        out = None
        for k, (l, (W, bias)) in enumerate(zip(self.L, self.split(p))):
            bi, ci, cin, bo, co, cout, s, tr, relu = l
            X = acts[bi][:, :, ci:ci + cin]
            Wr = self._rb(W)
            nvo = acts[bo].shape[1] if bo >= 0 else X.shape[1]
            acc = np.zeros((B, nvo, cout))
            for t, (ov, iv) in enumerate(self.links[k]):
                np.add.at(acc, (slice(None), ov), X[:, iv] @ Wr[t])
            acc = acc + bias
            if relu:
                acc = np.maximum(acc, 0.0)
            if bo >= 0:
                acts[bo][:, :, co:co + cout] = self._rb(acc)
            else:
                out = self._rf(acc)
        return out, acts

    def loss_grad(self, x, target, p):
        """MSE over all output values and its gradient w.r.t. p (the kernels' order)."""
        out, acts = self.forward(x, p)
        n = out.size
        loss = float(np.sum((out - target) ** 2) / n)
        dout = self._rf(2.0 / n * (out - target))
        grads = [None] * len(self.ch)
        for b in range(1, len(self.ch)):
            grads[b] = np.zeros_like(acts[b])
        gp = []
        params = self.split(p)
        for k in range(len(self.L) - 1, -1, -1):
            bi, ci, cin, bo, co, cout, s, tr, relu = self.L[k]
            W, _ = params[k]
            dY = dout if bo < 0 else grads[bo][:, :, co:co + cout]
            X = acts[bi][:, :, ci:ci + cin]
            dYr = self._rb(dY)
            gW = np.zeros((27, cin, cout))
            for t, (ov, iv) in enumerate(self.links[k]):
                gW[t] = np.einsum("bvi,bvo->io", X[:, iv], dYr[:, ov])
            gb = dY.sum(axis=(0, 1))
            gp.append(np.concatenate([gW.ravel(), gb]))
            if k == 0:
                continue
            Wr = self._rb(W)
            dX = np.zeros_like(X)
            for t, (ov, iv) in enumerate(self.links[k]):
                np.add.at(dX, (slice(None), iv), dYr[:, ov] @ Wr[t].T)
            dX = self._rf(dX * (X > 0))
            if s == 2 and not tr:
                grads[bi][:, :, ci:ci + cin] += dX
            else:
                grads[bi][:, :, ci:ci + cin] = dX
        return loss, np.concatenate(gp[::-1]), out
