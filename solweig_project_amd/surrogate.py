"""Learned SVF surrogate: a 4-level convolutional encoder-decoder run and
trained on the GPU by the library's own BF16 MFMA kernels (``sw_cnn_fwd``,
``sw_cnn_bwd``, ``sw_adam``), gated by and falling back to the exact SVF
kernel.

The reference has no surrogate; this is a new path beside the exact sweep.
Input: 384 x 384 windows (256 x 256 core + 64-pixel halo) of the composed
march heights {DSM, vegetation DSM, trunk DSM} / 50 m. Output: the 10 SVF
fields total, E, S, W, N and their vegetation variants (the first 10
field planes of the exact sweep) over the core. Training data: windows of
seeded synthetic cities labelled by the exact SVF sweep on the whole city
(so labels see occluders beyond the halo). At inference the gate compares
the surrogate with the exact kernel on sampled 16 x 16 tiles of every
window core and re-solves the windows above the tolerance exactly.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L

LAYERS = ("e1a", "e1", "e2a", "e2", "e3a", "e3", "e4a", "e4", "d3a", "d3", "d2a", "d2", "d1a", "d1", "head")
CIN0 = 16           # input channels (3 heights, zero padded to one MFMA K chunk)
COUT_HEAD = 16      # head channels (10 fields, zero padded)
N_FIELDS = 10
FIELD_MAP = list(range(N_FIELDS))   # svf, svfE, svfS, svfW, svfN, svfveg, svfEveg, svfSveg, svfWveg, svfNveg
HALO = 64
CORE = 256
HW = CORE + 2 * HALO


def layer_shapes(ch):
    c1, c2, c3, c4 = ch
    return {"e1a": (c1, 3, CIN0), "e1": (c1, 3, c1), "e2a": (c2, 3, c1), "e2": (c2, 3, c2),
            "e3a": (c3, 3, c2), "e3": (c3, 3, c3), "e4a": (c4, 3, c3), "e4": (c4, 3, c4),
            "d3a": (c3, 3, c4 + c3), "d3": (c3, 3, c3), "d2a": (c2, 3, c3 + c2), "d2": (c2, 3, c2),
            "d1a": (c1, 3, c2 + c1), "d1": (c1, 3, c1), "head": (COUT_HEAD, 1, c1)}


#: resolution level of each layer (1 = full window, 4 = 1/8)
LAYER_LEVEL = {"e1a": 1, "e1": 1, "e2a": 2, "e2": 2, "e3a": 3, "e3": 3, "e4a": 4, "e4": 4,
               "d3a": 3, "d3": 3, "d2a": 2, "d2": 2, "d1a": 1, "d1": 1, "head": 1}


def macs_per_window(ch, hw: int = HW, padded: bool = False, core: int | None = None) -> int:
    """Multiply-accumulates of one forward pass over an hw x hw window. ``padded`` counts the
    zero-padded input and head channels the kernels compute; otherwise the 3 input and 10 output
    channels of the model. ``core``: inference, where d1 and the head run over the core x core
    square only and d1a over the core plus one pixel."""
    total = 0
    for name, (co, ks, ci) in layer_shapes(ch).items():
        if not padded and name == "e1a":
            ci = 3
        if not padded and name == "head":
            co = N_FIELDS
        side = hw >> (LAYER_LEVEL[name] - 1)
        if core is not None and name in ("d1", "head"):
            side = core
        elif core is not None and name == "d1a":
            side = core + 2
        total += side * side * co * ks * ks * ci
    return total


class SurrogateNet:
    """Parameters (f32 master + bf16 copy), gradients and Adam state in flat device buffers."""

    def __init__(self, channels=(32, 64, 128, 256), device="cuda:0", seed: int = 0):
        if any(c % 16 for c in channels) or len(channels) != 4:
            raise ValueError("four level channel counts, each a multiple of 16")
        self.channels = tuple(int(c) for c in channels)
        self.device = torch.device(device)
        self.shapes = layer_shapes(self.channels)
        self.offsets = {}
        off = 0
        for name in LAYERS:
            co, ks, ci = self.shapes[name]
            nw = co * ks * ks * ci
            self.offsets[name] = (off, nw, off + nw, co)
            off += nw + co
        self.n_params = off
        rng = np.random.Generator(np.random.PCG64(seed))
        host = np.zeros(off, np.float32)
        for name in LAYERS:
            co, ks, ci = self.shapes[name]
            w0, nw, b0, _ = self.offsets[name]
            fan = ks * ks * (3 if name == "e1a" else ci)
            w = rng.standard_normal((co, ks * ks, ci)).astype(np.float32) * np.float32(np.sqrt(2.0 / fan))
            if name == "e1a":
                w[:, :, 3:] = 0
            if name == "head":
                w *= np.float32(0.1)
                w[N_FIELDS:] = 0
                host[b0:b0 + N_FIELDS] = 0.4          # sigmoid(0.4) ~ a typical SVF
            host[w0:w0 + nw] = w.ravel()
        dev = self.device
        self.p = torch.from_numpy(host).to(dev)
        self.pb = self.p.to(torch.bfloat16)
        self.g = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.step = 0

    def w(self, name):
        w0, nw, _, _ = self.offsets[name]
        return self.pb[w0:w0 + nw]

    def b(self, name):
        _, _, b0, co = self.offsets[name]
        return self.p[b0:b0 + co]

    def gw(self, name):
        w0, nw, _, _ = self.offsets[name]
        return self.g[w0:w0 + nw]

    def gb(self, name):
        _, _, b0, co = self.offsets[name]
        return self.g[b0:b0 + co]

    def numpy_params(self):
        """{name: (w, b)} as float32 numpy arrays (weights as the kernels read them: bf16-rounded)."""
        pb = self.pb.float().cpu().numpy()
        p = self.p.cpu().numpy()
        out = {}
        for name in LAYERS:
            co, ks, ci = self.shapes[name]
            w0, nw, b0, _ = self.offsets[name]
            out[name] = (pb[w0:w0 + nw].reshape(co, ks * ks, ci).copy(), p[b0:b0 + co].copy())
        return out

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, stream=None):
        self.step += 1
        L.call("sw_adam", L.ptr(self.p), L.ptr(self.m), L.ptr(self.v), L.ptr(self.g), L.ptr(self.pb),
               self.n_params, lr, b1, b2, eps, self.step, _sp(stream))

    def state_dict(self):
        return {"channels": np.array(self.channels, np.int32), "p": self.p.cpu().numpy()}

    def load_state_dict(self, d):
        if tuple(int(c) for c in d["channels"]) != self.channels:
            raise ValueError("channel counts differ")
        p = torch.from_numpy(np.asarray(d["p"], np.float32)).to(self.device)
        if p.numel() != self.n_params:
            raise ValueError("parameter count differs")
        self.p.copy_(p)
        self.pb.copy_(p.to(torch.bfloat16))


def _sp(stream):
    return stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream


class Passes:
    """Activation/gradient buffers and the forward and backward op lists for one batch shape."""

    def __init__(self, net: SurrogateNet, batch: int, hw: int = HW, core: int = CORE, c0: int = HALO,
                 nchunk: int = 64, train: bool = True, fuse: bool | None = None):
        if hw % 8:
            raise ValueError("window size must be a multiple of 8")
        self.net, self.n, self.hw, self.core, self.c0, self.train = net, batch, hw, core, c0, train
        # with 16 first-level channels the pools and the head run in the convolutions' epilogues
        self.fuse = (net.channels[0] == 16) if fuse is None else bool(fuse)
        if self.fuse and net.channels[0] != 16:
            raise ValueError("the fused pool / head epilogues need 16 first-level channels")
        c1, c2, c3, c4 = net.channels
        dev = net.device
        bf = torch.bfloat16
        dims = {1: hw, 2: hw // 2, 3: hw // 4, 4: hw // 8}
        lvl = {"e1a": 1, "e1": 1, "e2a": 2, "e2": 2, "e3a": 3, "e3": 3, "e4a": 4, "e4": 4,
               "d3a": 3, "d3": 3, "d2a": 2, "d2": 2, "d1a": 1, "d1": 1}
        self.a, self.d = {}, {}

        def t(level, c):
            s = dims[level]
            return torch.zeros((batch, s, s, c), dtype=bf, device=dev)
        self.x = t(1, CIN0)
        for name, lv in lvl.items():
            co = net.shapes[name][0]
            if train or not (self.fuse and name == "d1"):
                self.a[name] = t(lv, co)
            if train:
                self.d[name] = t(lv, co)
        for k, lv, c in (("p1", 2, c1), ("p2", 3, c2), ("p3", 4, c3)):
            self.a[k] = t(lv, c)
            if train:
                self.d[k] = t(lv, c)
        if train:
            # data gradients of the concat layers: upsampled part (full res) and skip part
            self.d["up4"], self.d["sk3"] = t(3, c4), t(3, c3)
            self.d["up3"], self.d["sk2"] = t(2, c3), t(2, c2)
            self.d["up2"], self.d["sk1"] = t(1, c2), t(1, c1)
        self.gz = {}                 # ReLU-gradient scratch per (level, channels), reused layer to layer
        self.out = torch.zeros((batch, hw, hw, COUT_HEAD), dtype=torch.float32, device=dev)
        self.labels = torch.zeros((batch, core, core, N_FIELDS), dtype=torch.float32, device=dev)
        self.dims, self.lvl, self.nchunk = dims, lvl, nchunk
        self.fwd_ops = self._forward_ops()
        self.bwd_ops, ws = [], 0
        if train:
            self.dhead = t(1, COUT_HEAD)
            nparts = -(-(batch * hw * hw * COUT_HEAD) // 256)
            self.parts = torch.zeros(nparts, dtype=torch.float32, device=dev)
            self.bwd_ops, ws = self._backward_ops()
        self.ws = torch.zeros(max(ws // 4, 1), dtype=torch.float32, device=dev)
        for op in self.bwd_ops:
            if op.kind in (L.CNN_CONV_BWD_DATA, L.CNN_CONV_BWD_WEIGHT):
                op.ws = L.ptr(self.ws)
        self.fwd_arr = (L.CnnOp * len(self.fwd_ops))(*self.fwd_ops)
        self.bwd_arr = (L.CnnOp * max(len(self.bwd_ops), 1))(*self.bwd_ops)

    # -- op builders -----------------------------------------------------------
    def _conv(self, name, in0, in1=None, up=False, out=None, act=1, outf=None):
        net = self.net
        co, ks, ci = net.shapes[name]
        c1 = in0.shape[3]
        c2 = 0 if in1 is None else in1.shape[3]
        assert c1 + c2 == ci, (name, c1, c2, ci)
        s = self.dims[self.lvl.get(name, 1)]
        o = L.CnnOp(kind=L.CNN_CONV, n=self.n, h=s, w=s, c1=c1, c2=c2, up=1 if up else 0, o1=co, cout=co, ks=ks,
                    act=act)
        o.in0, o.in1 = L.ptr(in0), (L.ptr(in1) if in1 is not None else None)
        o.out0 = L.ptr(out) if out is not None else None
        o.outf = L.ptr(outf) if outf is not None else None
        o.weight, o.bias = L.ptr(net.w(name)), L.ptr(net.b(name))
        return o

    def _forward_ops(self):
        a, ops = self.a, []
        ops.append(self._conv("e1a", self.x, out=a["e1a"]))
        ops.append(self._conv("e1", a["e1a"], out=a["e1"]))
        for lv, (e, p) in enumerate((("e1", "p1"), ("e2", "p2"), ("e3", "p3")), start=1):
            s = self.dims[lv]
            if self.fuse:
                ops[-1].pool_out = L.ptr(a[p])          # the pool runs in the epilogue of the conv producing e
            else:
                o = L.CnnOp(kind=L.CNN_POOL2, n=self.n, h=s, w=s, c1=a[e].shape[3])
                o.in0, o.out0 = L.ptr(a[e]), L.ptr(a[p])
                ops.append(o)
            nxt = ("e2a", "e2") if lv == 1 else (("e3a", "e3") if lv == 2 else ("e4a", "e4"))
            ops.append(self._conv(nxt[0], a[p], out=a[nxt[0]]))
            ops.append(self._conv(nxt[1], a[nxt[0]], out=a[nxt[1]]))
        ops.append(self._conv("d3a", a["e4"], a["e3"], up=True, out=a["d3a"]))
        ops.append(self._conv("d3", a["d3a"], out=a["d3"]))
        ops.append(self._conv("d2a", a["d3"], a["e2"], up=True, out=a["d2a"]))
        ops.append(self._conv("d2", a["d2a"], out=a["d2"]))
        ops.append(self._conv("d1a", a["d2"], a["e1"], up=True, out=a["d1a"]))
        if self.fuse:
            o = self._conv("d1", a["d1a"], out=a.get("d1"), outf=self.out)
            o.head_w, o.head_b = L.ptr(self.net.w("head")), L.ptr(self.net.b("head"))
            if not self.train and self.c0 >= 1 and self.c0 + self.core < self.hw:
                # inference reads the core only: d1 (and the head) over the core, d1a one pixel beyond it
                o.c0, o.core = self.c0, self.core
                ops[-1].c0, ops[-1].core = self.c0 - 1, self.core + 2
            ops.append(o)
        else:
            ops.append(self._conv("d1", a["d1a"], out=a["d1"]))
            ops.append(self._conv("head", a["d1"], act=2, outf=self.out))
        return ops

    def _backward_ops(self):
        net, a, d = self.net, self.a, self.d
        ops = []
        ws_need = 0
        s1 = self.dims[1]
        o = L.CnnOp(kind=L.CNN_HEAD_LOSS, n=self.n, h=s1, w=s1, c1=COUT_HEAD, core=self.core, c0=self.c0,
                    nl=N_FIELDS)
        o.outf, o.in0, o.out0, o.ws = L.ptr(self.out), L.ptr(self.labels), L.ptr(self.dhead), L.ptr(self.parts)
        ops.append(o)

        def scratch(op):
            b = ctypes.c_int64()
            L.call("sw_cnn_bwd_scratch_bytes", ctypes.byref(op), ctypes.byref(b))
            return b.value

        def back(name, dz, in0, in1=None, up=False, dx0=None, dx1=None):
            nonlocal ws_need
            co, ks, ci = net.shapes[name]
            c1 = in0.shape[3]
            c2 = 0 if in1 is None else in1.shape[3]
            s = self.dims[self.lvl.get(name, 1)]
            w = L.CnnOp(kind=L.CNN_CONV_BWD_WEIGHT, n=self.n, h=s, w=s, c1=c1, c2=c2, up=1 if up else 0, o1=co,
                        cout=co, ks=ks, nchunk=self.nchunk)
            w.in0, w.in1 = L.ptr(in0), (L.ptr(in1) if in1 is not None else None)
            w.dz, w.gw, w.gb, w.weight = L.ptr(dz), L.ptr(net.gw(name)), L.ptr(net.gb(name)), L.ptr(net.w(name))
            ws_need = max(ws_need, scratch(w))
            ops.append(w)
            if dx0 is None:
                return
            dd = L.CnnOp(kind=L.CNN_CONV_BWD_DATA, n=self.n, h=s, w=s, c1=co, c2=0, up=0, o1=c1, cout=ci, ks=ks,
                         act=0)
            dd.in0, dd.weight = L.ptr(dz), L.ptr(net.w(name))
            dd.out0, dd.out1 = L.ptr(dx0), (L.ptr(dx1) if dx1 is not None else None)
            ws_need = max(ws_need, scratch(dd))
            ops.append(dd)

        def relu_b(dy, y, level):
            s = self.dims[level]
            c = y.shape[3]
            z = self.gz.get((level, c))
            if z is None:
                z = self.gz[(level, c)] = torch.zeros_like(y)
            o = L.CnnOp(kind=L.CNN_RELU_BWD, n=self.n, h=s, w=s, c1=c)
            o.in0, o.in1, o.out0 = L.ptr(dy), L.ptr(y), L.ptr(z)
            ops.append(o)
            return z

        def pool_b(dout, add, out, level):
            s = self.dims[level]
            o = L.CnnOp(kind=L.CNN_POOL2_BWD, n=self.n, h=s, w=s, c1=out.shape[3])
            o.in0, o.in1, o.out0 = L.ptr(dout), L.ptr(add), L.ptr(out)
            ops.append(o)

        def up_b(dfull, dhalf, level):
            s = self.dims[level]
            o = L.CnnOp(kind=L.CNN_UP2_BWD, n=self.n, h=s, w=s, c1=dfull.shape[3])
            o.in0, o.out0 = L.ptr(dfull), L.ptr(dhalf)
            ops.append(o)

        back("head", self.dhead, a["d1"], dx0=d["d1"])
        back("d1", relu_b(d["d1"], a["d1"], 1), a["d1a"], dx0=d["d1a"])
        back("d1a", relu_b(d["d1a"], a["d1a"], 1), a["d2"], a["e1"], up=True, dx0=d["up2"], dx1=d["sk1"])
        up_b(d["up2"], d["d2"], 1)
        back("d2", relu_b(d["d2"], a["d2"], 2), a["d2a"], dx0=d["d2a"])
        back("d2a", relu_b(d["d2a"], a["d2a"], 2), a["d3"], a["e2"], up=True, dx0=d["up3"], dx1=d["sk2"])
        up_b(d["up3"], d["d3"], 2)
        back("d3", relu_b(d["d3"], a["d3"], 3), a["d3a"], dx0=d["d3a"])
        back("d3a", relu_b(d["d3a"], a["d3a"], 3), a["e4"], a["e3"], up=True, dx0=d["up4"], dx1=d["sk3"])
        up_b(d["up4"], d["e4"], 3)
        back("e4", relu_b(d["e4"], a["e4"], 4), a["e4a"], dx0=d["e4a"])
        back("e4a", relu_b(d["e4a"], a["e4a"], 4), a["p3"], dx0=d["p3"])
        pool_b(d["p3"], d["sk3"], d["e3"], 3)
        back("e3", relu_b(d["e3"], a["e3"], 3), a["e3a"], dx0=d["e3a"])
        back("e3a", relu_b(d["e3a"], a["e3a"], 3), a["p2"], dx0=d["p2"])
        pool_b(d["p2"], d["sk2"], d["e2"], 2)
        back("e2", relu_b(d["e2"], a["e2"], 2), a["e2a"], dx0=d["e2a"])
        back("e2a", relu_b(d["e2a"], a["e2a"], 2), a["p1"], dx0=d["p1"])
        pool_b(d["p1"], d["sk1"], d["e1"], 1)
        back("e1", relu_b(d["e1"], a["e1"], 1), a["e1a"], dx0=d["e1a"])
        back("e1a", relu_b(d["e1a"], a["e1a"], 1), self.x)
        return ops, ws_need

    # -- passes ----------------------------------------------------------------
    def use_site(self, hpack, d_org, rows: int, cols: int):
        """Inference: the first convolution reads the windows straight from a site's composed heights
        (no packed input); ``d_org`` holds the window origins (int32 pairs) on the device."""
        op = self.fwd_arr[0]
        op.site, op.site_origins, op.site_rows, op.site_cols = L.ptr(hpack), L.ptr(d_org), int(rows), int(cols)

    def forward(self, stream=None):
        L.call("sw_cnn_fwd", self.fwd_arr, len(self.fwd_ops), _sp(stream))
        return self.out

    def backward(self, stream=None):
        if not self.train:
            raise RuntimeError("an inference Passes has no backward pass")
        L.call("sw_cnn_bwd", self.bwd_arr, len(self.bwd_ops), _sp(stream))

    def loss(self) -> float:
        return float(self.parts.sum().item())

    def train_step(self, lr, stream=None):
        self.forward(stream)
        self.backward(stream)
        self.net.adam(lr, stream=stream)


# ---------------------------------------------------------------------------
# training data, training loop, gated inference
# ---------------------------------------------------------------------------
class SiteField:
    """A device-resident city: composed heights (hpack) and exact SVF field planes."""

    def __init__(self, hpack, fields, rows, cols):
        self.hpack, self.fields, self.rows, self.cols = hpack, fields, rows, cols


def exact_sites(solver, n_sites: int, size: int, seed: int):
    """``n_sites`` seeded synthetic cities of size x size, labelled by the exact SVF sweep."""
    from . import synth
    from .solver import TileInputs
    out = []
    for k in range(n_sites):
        s = synth.make_site(size, size, seed=seed + k, with_aspect=False)
        ds = solver.upload(TileInputs.from_site(s))
        solver.svf(ds)
        out.append(SiteField(ds.hpack, ds.fields, size, size))
        del ds
    torch.cuda.synchronize()
    return out


def window_origins(n: int, rows: int, cols: int, seed: int, hw: int = HW, halo: int = HALO) -> np.ndarray:
    """``n`` window origins whose cores lie inside a rows x cols site (the halo may leave it)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    core = hw - 2 * halo
    r = rng.integers(-halo, rows - core - halo + 1, n)
    c = rng.integers(-halo, cols - core - halo + 1, n)
    return np.stack([r, c], axis=1).astype(np.int32)


def load_batch(ps: Passes, site: SiteField, origins: np.ndarray, with_labels: bool = True, stream=None):
    """Input windows (and labels) of ``origins`` into the pass buffers."""
    dev = ps.net.device
    org = torch.from_numpy(np.ascontiguousarray(origins, np.int32)).to(dev)
    n = len(origins)
    sp = _sp(stream)
    L.call("sw_cnn_pack_input", L.ptr(site.hpack), L.ptr(org), L.ptr(ps.x), n, ps.hw, site.rows, site.cols, sp)
    if with_labels:
        fm = _field_map(dev)
        L.call("sw_cnn_gather_labels", L.ptr(site.fields), L.ptr(org), L.ptr(fm), L.ptr(ps.labels), n, ps.core,
               ps.c0, N_FIELDS, site.rows, site.cols, sp)
    return org


_FM = {}


def _field_map(dev):
    t = _FM.get(str(dev))
    if t is None:
        t = _FM[str(dev)] = torch.tensor(FIELD_MAP, dtype=torch.int32, device=dev)
    return t


def train(net: SurrogateNet, sites, *, steps: int, batch: int = 8, lr: float = 2e-3, windows: int = 20000,
          seed: int = 0, log=None):
    """Adam with cosine decay over ``steps``; every batch is drawn from one site's fixed window list."""
    ps = Passes(net, batch)
    per = max(1, windows // len(sites))
    lists = [window_origins(per, s.rows, s.cols, seed + 101 * k) for k, s in enumerate(sites)]
    rng = np.random.Generator(np.random.PCG64(seed))
    losses = []
    t0 = time.perf_counter()
    for it in range(steps):
        k = int(rng.integers(0, len(sites)))
        pick = rng.choice(len(lists[k]), batch, replace=False)
        load_batch(ps, sites[k], lists[k][pick])
        cur = lr * 0.5 * (1.0 + np.cos(np.pi * it / max(steps, 1)))
        ps.train_step(float(cur))
        if it % 50 == 0 or it == steps - 1:
            losses.append((it, ps.loss()))
            if log:
                log(f"step {it}: loss {losses[-1][1]:.5f} ({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    return {"losses": losses, "seconds": time.perf_counter() - t0, "steps": steps, "batch": batch,
            "windows": per * len(sites)}


def core_origins(rows: int, cols: int, core: int = CORE, halo: int = HALO) -> np.ndarray:
    """Window origins whose cores tile a rows x cols site."""
    rr = np.arange(0, rows, core)
    cc = np.arange(0, cols, core)
    g = np.stack(np.meshgrid(rr, cc, indexing="ij"), axis=-1).reshape(-1, 2) - halo
    return g.astype(np.int32)


@dataclass
class GateReport:
    windows: int
    accepted: int
    sampled_tiles: int
    exact_tiles: int
    ms_total: float
    ms_network: float
    window_errors: np.ndarray = None


class SvfSurrogate:
    """Gated surrogate SVF of a device-resident scene.

    The network predicts fields 0..9 of every 256 x 256 core; SVFtotal
    (field 15) follows from them as in the fold (svf - (1 - svfveg) * 0.97).
    The gate runs the exact kernel on ``samples`` random 16 x 16 tiles of
    every core and accepts a core when its error over the sampled tiles and
    the 10 fields is at most ``tol``: with ``criterion="mean"`` the mean
    absolute difference, with ``criterion="max"`` the largest absolute
    difference of any field at any sampled pixel. Rejected cores are
    re-solved exactly. Sampled tiles and rejected cores carry exact values;
    the aveg fields (10..14) are exact only there. Unsampled pixels of an
    accepted core are bounded by neither criterion.
    """

    def __init__(self, net: SurrogateNet, tol: float = 0.02, samples: int = 4, seed: int = 0,
                 criterion: str = "mean"):
        if criterion not in ("mean", "max"):
            raise ValueError(f"gate criterion {criterion!r}: 'mean' or 'max'")
        self.net, self.tol, self.samples, self.seed = net, tol, samples, seed
        self.criterion = criterion
        self._geo = {}
        self.ps = None
        self._side = None

    def _geometry(self, rows, cols):
        key = (rows, cols)
        g = self._geo.get(key)
        if g is None:
            from .incremental import TileGrid
            dev = self.net.device
            org = core_origins(rows, cols)
            grid = TileGrid(rows, cols)
            rng = np.random.Generator(np.random.PCG64(self.seed))
            per = CORE // 16
            tiles, owner = [], []
            for w, (r0, c0) in enumerate(org + HALO):
                tr0, tc0 = r0 // 16, c0 // 16
                ty, tx = min(per, grid.ty - tr0), min(per, grid.tx - tc0)
                pick = rng.choice(ty * tx, min(self.samples, ty * tx), replace=False)
                for k in pick:
                    tiles.append((tr0 + int(k) // tx) * grid.tx + tc0 + int(k) % tx)
                    owner.append(w)
            tiles = np.array(tiles, np.int32)
            # pixel indices of the sampled tiles (inside the site) and their window
            pix, pw = [], []
            for t, w in zip(tiles, owner):
                ty, tx = divmod(int(t), grid.tx)
                rr = np.arange(ty * 16, min(rows, ty * 16 + 16))
                cc = np.arange(tx * 16, min(cols, tx * 16 + 16))
                idx = (rr[:, None] * cols + cc[None, :]).ravel()
                pix.append(idx)
                pw.append(np.full(idx.size, w, np.int64))
            pix = np.concatenate(pix)
            pw = np.concatenate(pw)
            seg = np.zeros(len(org) + 1, np.int32)
            np.cumsum(np.bincount(pw, minlength=len(org)), out=seg[1:])
            win = np.zeros((rows, cols), np.int64)
            for w, (r0, c0) in enumerate(org + HALO):
                win[r0:r0 + CORE, c0:c0 + CORE] = w
            sampled = np.zeros(rows * cols, bool)
            sampled[pix] = True
            g = dict(org=org, grid=grid, tiles=tiles, d_tiles=torch.from_numpy(tiles).to(dev),
                     pix=torch.from_numpy(pix).to(dev), pw=torch.from_numpy(pw).to(dev),
                     npix=torch.bincount(torch.from_numpy(pw), minlength=len(org)).to(dev).float(),
                     win=torch.from_numpy(win.ravel()).to(dev), sampled=torch.from_numpy(sampled).to(dev),
                     pix32=torch.from_numpy(pix.astype(np.int32)).to(dev), seg=torch.from_numpy(seg).to(dev),
                     d_org=torch.from_numpy(np.ascontiguousarray(org, np.int32)).to(dev))
            self._geo[key] = g
        return g

    def predict(self, ds, fields, stream=None):
        """Network outputs of every core of scene ``ds`` into fields 0..9 of ``fields`` (one batch)."""
        g = self._geometry(ds.rows, ds.cols)
        org = g["org"]
        if self.ps is None or self.ps.n != len(org):
            self.ps = Passes(self.net, len(org), train=False)
        site = SiteField(ds.hpack, fields, ds.rows, ds.cols)
        self.ps.fwd_arr[0].site = None
        d_org = load_batch(self.ps, site, org, with_labels=False, stream=stream)
        self.ps.forward(stream)
        L.call("sw_cnn_scatter_cores", L.ptr(self.ps.out), L.ptr(d_org), L.ptr(_field_map(self.net.device)),
               L.ptr(fields), len(org), self.ps.hw, COUT_HEAD, self.ps.core, self.ps.c0, N_FIELDS, ds.rows, ds.cols,
               _sp(stream))
        return org

    def svf(self, solver, ds, stream=None) -> GateReport:
        """Gated surrogate SVF into ``ds.fields``; ``ds`` must be composed (sw_compose) beforehand."""
        from .incremental import TileGrid
        t0 = time.perf_counter()
        rows, cols = ds.rows, ds.cols
        g = self._geometry(rows, cols)
        org = g["org"]
        dev = self.net.device
        main = stream if stream is not None else torch.cuda.current_stream(dev)
        if self._side is None:
            self._side = torch.cuda.Stream(dev)
        if self.ps is None or self.ps.n != len(org):
            self.ps = Passes(self.net, len(org), train=False)
        ps = self.ps
        solver.svf_tables(rows, cols, ds.tile.scale)          # built and uploaded on main before the fork
        # the exact sweep of the sampled tiles runs beside the network (different pixels and buffers)
        ready = torch.cuda.Event()
        ready.record(main)
        self._side.wait_event(ready)
        self._exact_tiles(solver, ds, g["d_tiles"], len(g["tiles"]), self._side)
        sampled_done = torch.cuda.Event()
        sampled_done.record(self._side)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(main)
        sp = _sp(main)
        ps.use_site(ds.hpack, g["d_org"], rows, cols)
        ps.forward(main)
        ev[1].record(main)
        main.wait_event(sampled_done)
        fm = _field_map(dev)
        werr = torch.empty(len(org), dtype=torch.float32, device=dev)
        L.call("sw_gate_max_error" if self.criterion == "max" else "sw_gate_error", L.ptr(ps.out), L.ptr(ds.fields),
               L.ptr(g["d_org"]), L.ptr(g["pix32"]),
               L.ptr(g["seg"]), L.ptr(fm), L.ptr(werr), len(org), ps.hw, COUT_HEAD, N_FIELDS, rows, cols, sp)
        L.call("sw_gate_write", L.ptr(ps.out), L.ptr(g["d_org"]), L.ptr(fm), L.ptr(g["sampled"]), L.ptr(ds.fields),
               len(org), ps.hw, COUT_HEAD, ps.core, ps.c0, N_FIELDS, rows, cols, sp)
        werr_h = werr.cpu().numpy() if stream is None else _to_host(werr, main)
        bad = np.nonzero(werr_h > self.tol)[0]
        ex = 0
        if len(bad):
            m = g["grid"].empty()
            for w in bad:
                r0, c0 = (org[w] + HALO).tolist()
                m[r0 // 16:min(g["grid"].ty, (r0 + CORE) // 16), c0 // 16:min(g["grid"].tx, (c0 + CORE) // 16)] = True
            ids = TileGrid.ids(m)
            ex = len(ids)
            self._exact_tiles(solver, ds, torch.from_numpy(ids).to(dev), ex, main)
        main.synchronize()
        return GateReport(windows=len(org), accepted=len(org) - len(bad), sampled_tiles=len(g["tiles"]),
                          exact_tiles=ex, ms_total=(time.perf_counter() - t0) * 1e3,
                          ms_network=ev[0].elapsed_time(ev[1]), window_errors=werr_h)

    @staticmethod
    def _exact_tiles(solver, ds, d_ids, count, stream=None):
        if count == 0:
            return
        P = L.ptr
        offs, tab = solver.svf_tables(ds.rows, ds.cols, ds.tile.scale)
        c = solver._fold_consts
        L.call("sw_svf_sweep_tiles", P(ds.hpack), P(ds.stats), P(offs), P(tab), P(ds.vegdsm2), P(c["w_iso"]),
               P(c["w_aniso"]), P(c["ring"]), P(c["nann"]), P(c["dirm"]), P(ds.counts), P(ds.hpad), P(ds.masks),
               P(ds.fields), P(d_ids), int(count), ds.rows, ds.cols, _sp(stream))


def _to_host(t, stream):
    """A device tensor's values on the host after ``stream`` reaches this point."""
    stream.synchronize()
    return t.cpu().numpy()


_DATA = __import__("pathlib").Path(__file__).resolve().parent / "data"
#: the fast network (0.49 M parameters, channels 16, 32, 64, 128)
DEFAULT_WEIGHTS = _DATA / "svf_surrogate_16_32_64_128.npz"
#: the network of the surrogate configuration (1.95 M parameters, channels 32, 64, 128, 256, trained on
#: 20 000 windows of 40 cities)
CONFIG_WEIGHTS = _DATA / "svf_surrogate_32_64_128_256.npz"


def load_weights(path, device="cuda:0") -> SurrogateNet:
    """A trained network from a weights file written by ``state_dict`` (channels and parameters)."""
    with np.load(path, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    net = SurrogateNet(tuple(int(c) for c in d["channels"]), device)
    net.load_state_dict(d)
    return net


def load_default(device="cuda:0") -> SurrogateNet:
    """The fast trained network shipped with the package (channels 16, 32, 64, 128)."""
    return load_weights(DEFAULT_WEIGHTS, device)


def load_config(device="cuda:0") -> SurrogateNet:
    """The trained network of the surrogate configuration (channels 32, 64, 128, 256)."""
    return load_weights(CONFIG_WEIGHTS, device)
