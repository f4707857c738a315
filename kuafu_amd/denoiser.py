"""Python binding of the denoise stage (U-Net on MFMA) of libkuafu_amd."""
import ctypes as C
import os

import numpy as np

from .abi import check, lib, ptr

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet.kfw")


def to_bf16_bits(x):
    """float -> BF16 bit patterns (round to nearest even)."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


class Denoiser:
    def __init__(self, weights=None, random_seed=None):
        h = C.c_void_p()
        path = None if weights is None else str(weights).encode()
        check(lib().kf_denoiser_create(path, C.byref(h)))
        self.h = h
        if random_seed is not None:
            check(lib().kf_denoiser_init_random(self.h, random_seed))
        n = C.c_uint64()
        check(lib().kf_denoiser_num_params(self.h, C.byref(n)))
        self.n = n.value

    def close(self):
        if self.h:
            lib().kf_denoiser_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.n, np.float32)
        check(lib().kf_denoiser_get_params(self.h, ptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n
        check(lib().kf_denoiser_set_params(self.h, ptr(p)))

    def save(self, path):
        check(lib().kf_denoiser_save(self.h, str(path).encode()))

    def load(self, path):
        check(lib().kf_denoiser_load(self.h, str(path).encode()))

    def gradients(self, inputs, targets):
        inputs = np.ascontiguousarray(inputs, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        n, h, w, _ = inputs.shape
        g = np.zeros(self.n, np.float32)
        loss = C.c_float()
        check(lib().kf_denoiser_gradients(self.h, ptr(inputs), ptr(targets), n, h, w, C.byref(loss), ptr(g)))
        return loss.value, g

    def train_step(self, inputs, targets, lr):
        inputs = np.ascontiguousarray(inputs, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        n, h, w, _ = inputs.shape
        loss = C.c_float()
        check(lib().kf_denoiser_train_step(self.h, ptr(inputs), ptr(targets), n, h, w, lr, C.byref(loss)))
        return loss.value

    def forward(self, inputs):
        inputs = np.ascontiguousarray(inputs, np.float32)
        n, h, w, _ = inputs.shape
        out = np.zeros((n, h, w, 3), np.float32)
        check(lib().kf_denoiser_forward(self.h, ptr(inputs), n, h, w, ptr(out)))
        return out

    def apply(self, renderer):
        """Denoise a renderer's rgba target in place (uses its albedo / normal)."""
        check(lib().kf_denoiser_apply(self.h, renderer.h))

    def infer_device(self, d_rgba, d_albedo, d_normal, width, height, d_out, n_images=1):
        """Denoise n_images stacked width x height images (device pointers); returns the kernel ms."""
        ms = C.c_float()
        check(lib().kf_denoiser_infer_device_batch(self.h, d_rgba, d_albedo, d_normal, width, height, n_images, d_out,
                                                   C.byref(ms)))
        return ms.value


def conv_forward(x, w, b, relu=True):
    """Single layer: x (n,h,w,cin) floats, w (9,cout,cin), b (cout,) -> BF16-stored output as float."""
    n, h, wd, cin = x.shape
    cout = w.shape[1]
    xb, wb = to_bf16_bits(x), to_bf16_bits(w)
    bb = np.ascontiguousarray(b, np.float32)
    y = np.zeros((n, h, wd, cout), np.uint16)
    check(lib().kf_conv3x3_forward(ptr(xb), cin, ptr(wb), ptr(bb), cout, n, h, wd, int(relu), ptr(y)))
    return from_bf16_bits(y)


def conv_backward(x, w, dy, mask_by_x=True):
    n, h, wd, cin = x.shape
    cout = w.shape[1]
    xb, wb, dyb = to_bf16_bits(x), to_bf16_bits(w), to_bf16_bits(dy)
    dx = np.zeros((n, h, wd, cin), np.uint16)
    dw = np.zeros(w.shape, np.float32)
    db = np.zeros(cout, np.float32)
    check(lib().kf_conv3x3_backward(ptr(xb), cin, ptr(wb), cout, ptr(dyb), n, h, wd, int(mask_by_x), ptr(dx),
                                    ptr(dw), ptr(db)))
    return from_bf16_bits(dx), dw, db
