"""FP64 numpy reference of the denoiser U-Net (TEST INFRASTRUCTURE).

Mirrors the architecture and parameter layout of kuafu_amd/csrc/denoise.hip:
19 3x3 convolutions (encoder 32/64/96/128/160 channels, decoder with skip
concatenations, output conv with 3 used channels), 2x2 average pooling,
nearest 2x upsampling, residual on log1p(noisy RGB), loss
mean|yhat - log1p(t)| + mean (expm1(yhat) - t)^2 / (t^2 + 0.01).

With bf16=True every tensor the GPU stores in BF16 is rounded to BF16 here
too (activations, gradients, weights), so the GPU's FP32-accumulated results
can be compared with an FP64 computation of the same rounded operands.  With
bf16=False it is a plain FP64 network whose backward is checked against
finite differences.
"""
import numpy as np

CH = [32, 64, 96, 128, 160]
IN_CH = 32
LOG_CLAMP = 10.0


def convs():
    c = []
    prev = IN_CH
    for ch in CH:
        c += [(prev, ch), (ch, ch)]
        prev = ch
    for l in range(len(CH) - 2, -1, -1):
        c += [(CH[l] + CH[l + 1], CH[l]), (CH[l], CH[l])]
    c.append((CH[0], 32))
    return c


def param_slices():
    out, off = [], 0
    for ci, co in convs():
        w = (off, off + 9 * ci * co)
        off = w[1]
        b = (off, off + co)
        off = b[1]
        out.append((w, b, ci, co))
    return out, off


def round_bf16(x):
    """Round to nearest-even BF16 (as float64)."""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def conv_fwd(x, w, b):
    """x [N,H,W,Ci], w [9,Co,Ci], b [Co] -> [N,H,W,Co] (zero padding)."""
    N, H, W, Ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((N, H, W, w.shape[1]))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        y += xp[:, ky:ky + H, kx:kx + W, :] @ w[tap].T
    return y + b


def conv_bwd(x, w, dy):
    N, H, W, Ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        dxp[:, ky:ky + H, kx:kx + W, :] += dy @ w[tap]
        dw[tap] = np.einsum("nhwo,nhwi->oi", dy, xp[:, ky:ky + H, kx:kx + W, :])
    return dxp[:, 1:H + 1, 1:W + 1, :], dw, dy.sum((0, 1, 2))


def pool(x):
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])


def pool_bwd(dp):
    return 0.25 * np.repeat(np.repeat(dp, 2, 1), 2, 2)


def up(x):
    return np.repeat(np.repeat(x, 2, 1), 2, 2)


def up_bwd(d):
    return d[:, 0::2, 0::2] + d[:, 0::2, 1::2] + d[:, 1::2, 0::2] + d[:, 1::2, 1::2]


def pack(inputs):
    """(N,H,W,9) records -> network input (N,H,W,32) and log1p(noisy rgb)."""
    r = np.nan_to_num(np.asarray(inputs, np.float64))
    r = np.clip(r, -1e4, 1e4)
    lin = np.log1p(np.maximum(r[..., 0:3], 0))
    x = np.zeros(r.shape[:3] + (IN_CH,))
    x[..., 0:3] = lin
    x[..., 3:6] = r[..., 3:6] * 3.141592
    x[..., 6:9] = r[..., 6:9]
    return x, lin


class Net:
    def __init__(self, params, bf16=True):
        self.sl, n = param_slices()
        p = np.asarray(params, np.float64)
        assert p.size == n
        self.emulate = bf16
        self.r = round_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
        f32 = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if bf16 else (lambda a: np.asarray(a, np.float64))
        self.f32 = f32
        self.W, self.B = [], []
        for (w, b, ci, co) in self.sl:
            self.W.append(self.r(p[w[0]:w[1]]).reshape(9, co, ci))
            self.B.append(f32(p[b[0]:b[1]]))

    def forward(self, inputs):
        r = self.r
        x0, lin = pack(inputs)
        x0 = r(x0)
        lin = self.f32(lin)
        s = dict(x0=x0, lin=lin, t=[], cat_skip=[], pl=[], da=[], u=[None] * 4)
        k = 0
        inp = x0
        L = len(CH)
        for l in range(L):
            t = r(np.maximum(conv_fwd(inp, self.W[k], self.B[k]), 0))
            k += 1
            s["t"].append(t)
            o = r(np.maximum(conv_fwd(t, self.W[k], self.B[k]), 0))
            k += 1
            if l < L - 1:
                s["cat_skip"].append(o)
                inp = r(pool(o))
                s["pl"].append(inp)
            else:
                s["bt"] = o
        low = s["bt"]
        s["cat"] = [None] * (L - 1)
        s["da"] = [None] * (L - 1)
        for l in range(L - 2, -1, -1):
            cat = np.concatenate([s["cat_skip"][l], up(low)], -1)
            s["cat"][l] = cat
            da = r(np.maximum(conv_fwd(cat, self.W[k], self.B[k]), 0))
            k += 1
            s["da"][l] = da
            u = r(np.maximum(conv_fwd(da, self.W[k], self.B[k]), 0))
            k += 1
            s["u"][l] = u
            low = u
        y = conv_fwd(s["u"][0], self.W[k], self.B[k])[..., 0:3]
        s["y"] = self.f32(y)
        return s

    def output(self, s):
        return np.expm1(np.minimum(s["y"] + s["lin"], LOG_CLAMP))

    def loss(self, s, target):
        yh = s["y"] + s["lin"]
        t = np.maximum(np.nan_to_num(np.asarray(target, np.float64)), 0)
        lt = np.log1p(t)
        o = np.expm1(np.minimum(yh, LOG_CLAMP))
        den = t * t + 0.01
        n = yh.size
        L = (np.abs(yh - lt) + (o - t) ** 2 / den).sum() / n
        d = np.sign(yh - lt) + np.where(yh > LOG_CLAMP, 0.0, 2 * (o - t) / den * (o + 1))
        return L, d / n

    def backward(self, s, dy3):
        """Gradients of every parameter (flat, same layout as the GPU)."""
        r = self.r
        L = len(CH)
        grads = [None] * len(self.W)
        k = len(self.W) - 1
        dy = np.zeros(dy3.shape[:3] + (32,))
        dy[..., 0:3] = dy3
        dy = r(dy)
        dx, grads_w, grads_b = conv_bwd(s["u"][0], self.W[k], dy)
        grads[k] = (grads_w, grads_b)
        du = [None] * (L - 1)
        du[0] = r(dx * (s["u"][0] > 0))
        k -= 1
        dcat = [None] * (L - 1)
        dbt = None
        for l in range(L - 1):
            dx, gw, gb = conv_bwd(s["da"][l], self.W[k], du[l])
            grads[k] = (gw, gb)
            dda = r(dx * (s["da"][l] > 0))
            k -= 1
            dx, gw, gb = conv_bwd(s["cat"][l], self.W[k], dda)
            grads[k] = (gw, gb)
            dcat[l] = r(dx * (s["cat"][l] > 0))
            k -= 1
            C = CH[l]
            low = s["u"][l + 1] if l + 1 < L - 1 else s["bt"]
            dlow = r(up_bwd(dcat[l][..., C:]) * (low > 0))
            if l + 1 < L - 1:
                du[l + 1] = dlow
            else:
                dbt = dlow
        dx, gw, gb = conv_bwd(s["t"][L - 1], self.W[k], dbt)
        grads[k] = (gw, gb)
        dt = r(dx * (s["t"][L - 1] > 0))
        k -= 1
        dx, gw, gb = conv_bwd(s["pl"][L - 2], self.W[k], dt)
        grads[k] = (gw, gb)
        dpl = r(dx)
        k -= 1
        for l in range(L - 2, -1, -1):
            C = CH[l]
            skip = s["cat"][l][..., :C]
            dskip = r(dcat[l][..., :C] + np.where(skip > 0, pool_bwd(dpl), 0.0))
            dx, gw, gb = conv_bwd(s["t"][l], self.W[k], dskip)
            grads[k] = (gw, gb)
            dt = r(dx * (s["t"][l] > 0))
            k -= 1
            inp = s["pl"][l - 1] if l > 0 else s["x0"]
            dx, gw, gb = conv_bwd(inp, self.W[k], dt)
            grads[k] = (gw, gb)
            dpl = r(dx) if l > 0 else None
            k -= 1
        flat = np.zeros(param_slices()[1])
        for (w, b, ci, co), (gw, gb) in zip(self.sl, grads):
            flat[w[0]:w[1]] = gw.ravel()
            flat[b[0]:b[1]] = gb
        return flat
