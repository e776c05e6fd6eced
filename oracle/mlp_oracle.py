"""CPU oracle of the learned closure's network (numpy).

TEST INFRASTRUCTURE ONLY (tests/, never the product path).  The closure has
no reference counterpart; this restates the GPU network exactly:

  12 -> 64 -> 64 -> 64 -> 9, ReLU or GELU (tanh form), inputs padded to 16 and
  outputs to 16; every GEMM operand (weights, activations, output and hidden
  gradients) rounded to BF16 (round to nearest even), FP32 accumulation, FP32
  bias, activation and Adam.

  Scale-free form: the scaled features x = f * fs of a point enter as the unit
  vector x / |x|, and the network's output is multiplied by |x|^2 (so the
  closure is homogeneous of degree 2 in the resolved gradients, as the
  gradient and eddy-viscosity stresses are); training targets are the scaled
  labels divided by |x|^2.  A point with x = 0 gets a zero input and output.

`bf16=False` gives the unrounded float64 network, used for finite-difference
gradient checks.  Parameter layout (FP32 block of 10448 + 32 scales):
W1[64][16] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[16][64] b4[16],
then input scales (16) and output scales (16).
"""
import numpy as np

NPARAM = 10448
OFF = dict(W1=0, b1=1024, W2=1088, b2=5184, W3=5248, b3=9344, W4=9408, b4=10432)
SHAPES = dict(W1=(64, 16), b1=(64,), W2=(64, 64), b2=(64,), W3=(64, 64), b3=(64,), W4=(16, 64), b4=(16,))
NIN, NOUT = 12, 9


def bf16(x):
    """round float32 values to bfloat16 (nearest even), returned as float32"""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    w = {k: p[OFF[k]:OFF[k] + int(np.prod(SHAPES[k]))].reshape(SHAPES[k]) for k in OFF}
    fs = p[NPARAM:NPARAM + 16]
    os_ = p[NPARAM + 16:NPARAM + 32]
    return w, fs, os_


def pack_grads(g):
    out = np.zeros(NPARAM)
    for k in OFF:
        out[OFF[k]:OFF[k] + g[k].size] = g[k].ravel()
    return out


RELU, GELU = 0, 1


def activation(x, kind=RELU):
    """hidden activation and its derivative.
    RELU: max(x, 0), derivative 1 for x > 0, else 0.
    GELU, tanh form: 0.5 x (1 + tanh(c (x + a x^3))), c = sqrt(2/pi), a = 0.044715."""
    if kind == GELU:
        c, a = 0.7978845608028654, 0.044715
        t = np.tanh(c * (x + a * x ** 3))
        dg = 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * c * (1 + 3 * a * x * x)
        return 0.5 * x * (1 + t), dg
    return np.maximum(x, 0.0), (x > 0).astype(np.float64)


class Net:
    def __init__(self, params, bf=True, act=RELU):
        self.w, self.fs, self.os = unpack(params)
        self.bf = bf
        self.act = act
        self.dt = np.float32 if bf else np.float64

    def r(self, x):
        return bf16(x) if self.bf else np.asarray(x, dtype=np.float64)

    def radius2(self, feats):
        """|x|^2 of the scaled features, computed in float32 like the GPU"""
        x = (np.asarray(feats, np.float64) * self.fs[:NIN]).astype(self.dt)
        return np.sum(x.astype(self.dt) * x.astype(self.dt), axis=1, dtype=self.dt)

    def forward(self, feats):
        """feats [n][12] raw -> dict of activations; out [n][16] normalised
        (scale-free: before the |x|^2 factor)"""
        dt = self.dt
        n = feats.shape[0]
        x = np.zeros((n, 16), dt)
        x[:, :NIN] = feats * self.fs[:NIN]
        r2 = self.radius2(feats)
        inv = np.where(r2 > 0, 1.0 / np.sqrt(np.where(r2 > 0, r2, 1.0)), 0.0).astype(dt)
        x = (x * inv[:, None]).astype(dt)
        act = {"a0": self.r(x), "r2": r2}
        a = act["a0"]
        for L, (W, b) in enumerate((("W1", "b1"), ("W2", "b2"), ("W3", "b3")), start=1):
            z = (a.astype(dt) @ self.r(self.w[W].astype(dt)).T + self.w[b].astype(dt)).astype(dt)
            g, dg = activation(z.astype(np.float64), self.act)
            act[f"z{L}"] = z
            act[f"dg{L}"] = dg.astype(dt)
            a = self.r(g.astype(dt))
            act[f"a{L}"] = a
        act["out"] = (a.astype(dt) @ self.r(self.w["W4"].astype(dt)).T + self.w["b4"].astype(dt)).astype(dt)
        return act

    def infer(self, feats):
        """physical outputs [n][9], tau made deviatoric"""
        act = self.forward(feats)
        o = act["out"][:, :NOUT].astype(np.float64) * act["r2"].astype(np.float64)[:, None] * self.os[:NOUT]
        tr = (o[:, 0] + o[:, 1] + o[:, 2]) / 3.0
        o[:, :3] -= tr[:, None]
        return o

    def loss_grad(self, feats, labels):
        """mean-square loss over normalised outputs and its parameter gradient"""
        dt = self.dt
        n = feats.shape[0]
        act = self.forward(feats)
        r2 = act["r2"]
        inv2 = np.where(r2 > 0, 1.0 / np.where(r2 > 0, r2, 1.0), 0.0).astype(dt)
        y = np.zeros((n, 16), dt)
        os_ = np.broadcast_to(self.os[:NOUT], (n, NOUT))
        scaled = np.divide(labels, os_, out=np.zeros((n, NOUT)), where=os_ > 0)   # zero column: target 0
        y[:, :NOUT] = scaled * inv2[:, None]
        d = np.zeros((n, 16), dt)
        d[:, :NOUT] = act["out"][:, :NOUT] - y[:, :NOUT]
        loss = float(np.sum(d.astype(np.float64) ** 2) / (n * NOUT))
        dout = self.r((2.0 * d / (n * NOUT)).astype(dt))
        g = {"W4": dout.T @ act["a3"], "b4": dout.sum(0)}
        da = dout @ self.r(self.w["W4"].astype(dt))
        for L, (W, b) in ((3, ("W3", "b3")), (2, ("W2", "b2")), (1, ("W1", "b1"))):
            dz = self.r((da * act[f"dg{L}"]).astype(dt))
            g[W] = dz.T @ act[f"a{L - 1}"]
            g[b] = dz.sum(0)
            if L > 1:
                da = dz @ self.r(self.w[W].astype(dt))
        # padded inputs / outputs carry no gradient
        g["W1"][:, NIN:] = 0
        g["W4"][NOUT:, :] = 0
        g["b4"][NOUT:] = 0
        return loss, pack_grads(g)
