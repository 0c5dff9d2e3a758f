"""Learned collision surrogate on the HIP path (host side).

The network, its training kernels and the encode / decode / fallback
pipeline are HIP code in libdgfs_hip (csrc/surrogate.hip); this module owns
the parameter vector and the optimiser state as device buffers, drives
training on (f, Q) pairs that the solver's own spectral operator produces,
and exposes the surrogate-with-fallback evaluation next to the exact
operator.  The reference has no learned operator; the drop-in boundary is the
same eval_contig(f, Q, npts) call the exact operator offers.
"""
import ctypes as C
import math

import numpy as np

from dgfs_hip import _lib

_P, _I, _L, _F, _D = C.c_void_p, C.c_int, C.c_long, C.c_float, C.c_double


class dgfs_sur_check(C.Structure):
    _fields_ = [("tol", C.c_double), ("safety", C.c_double), ("sample_frac", C.c_double),
                ("nbins", C.c_int), ("min_per_bin", C.c_int), ("seed", C.c_ulonglong)]


class dgfs_sur_check_stats(C.Structure):
    _fields_ = [("nfallback", C.c_int), ("nsampled", C.c_int), ("nsurrogate", C.c_int),
                ("nrejected_bins", C.c_int), ("est_rel_err", C.c_double)]


_lib.register({
    "dgfs_sur_create": [C.POINTER(_lib.dgfs_sur_desc), C.POINTER(C.c_void_p)],
    "dgfs_sur_destroy": [_P],
    "dgfs_sur_forward": [_P, _P, _P, _P, _I, _P, _I, _P],
    "dgfs_sur_backward": [_P, _P, _P, _P, _P],
    "dgfs_sur_adam": [_L, _P, _P, _P, _P, _F, _F, _F, _F, _I, _P],
    "dgfs_sur_encode": [_P, _P, _I, _P, _P, _P, _P],
    "dgfs_sur_targets": [_P, _P, _P, _I, _P, _P],
    "dgfs_sur_decode": [_P, _P, _P, _P, _I, _P, _I, _P],
    "dgfs_sur_eval": [_P, _P, _P, _P, _P, _I, _D, _D, _I, C.POINTER(_I), _P],
    "dgfs_sur_eval_checked": [_P, _P, _P, _P, _P, _I, _D, _D, C.POINTER(dgfs_sur_check), _I,
                              C.POINTER(dgfs_sur_check_stats), _P],
    "dgfs_nn_gemm": [_I, _I, _I, _I, _I, _F, _P, _L, _P, _L, _F, _P, _L, _I, _P],
    "dgfs_sur_param_count": [],
    "dgfs_sur_round_bf16": [_L, _P, _P, _P],
})

MODES = 512
WIDTH = 32


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def param_count():
    lib = _lib.load()
    fn = lib.dgfs_sur_param_count
    fn.restype = C.c_long
    fn.argtypes = []
    return int(fn())


def gemm(transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, Cm, ldc, bf16=True, stream=None):
    _lib.call("dgfs_nn_gemm", int(transA), int(transB), int(M), int(N), int(K), float(alpha), _ptr(A), int(lda),
              _ptr(B), int(ldb), float(beta), _ptr(Cm), int(ldc), int(bool(bf16)), stream)


def init_params(seed=0, scale=1.0):
    """Seeded initial parameter vector (numpy FP64; same recipe as the tests' checker)."""
    rng = np.random.default_rng(seed)
    D, L = MODES, 4
    layer = D * D + 2 * WIDTH * WIDTH + WIDTH
    th = np.zeros(2 * WIDTH + D * WIDTH + L * layer + WIDTH + D)
    th[0:WIDTH] = rng.normal(size=WIDTH) * scale
    th[WIDTH:2 * WIDTH] = rng.normal(size=WIDTH) * 0.1 * scale
    o = 2 * WIDTH
    th[o:o + D * WIDTH] = rng.normal(size=D * WIDTH) * 0.1 * scale
    o += D * WIDTH
    for _ in range(L):
        th[o:o + D * D] = rng.normal(size=D * D) / math.sqrt(D) * scale
        o += D * D
        th[o:o + WIDTH * WIDTH] = rng.normal(size=WIDTH * WIDTH) / math.sqrt(WIDTH) * scale
        o += WIDTH * WIDTH
        th[o:o + WIDTH * WIDTH] = rng.normal(size=WIDTH * WIDTH) / math.sqrt(WIDTH) * scale
        o += WIDTH * WIDTH
        o += WIDTH
    th[o:o + WIDTH] = rng.normal(size=WIDTH) / math.sqrt(WIDTH) * scale
    return th


class SpectralSurrogate:
    """Handle + FP32 parameter vector (+ Adam moments) on the device."""

    def __init__(self, Nv, L, cw, gamma, prefac, max_batch=512, device=None, seed=0, init_scale=1.0):
        import torch
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        desc = _lib.dgfs_sur_desc(Nv=int(Nv), max_batch=int(max_batch), L=float(L), cw=float(cw),
                                  gamma=float(gamma), prefac=float(prefac))
        h = C.c_void_p()
        _lib.call("dgfs_sur_create", C.byref(desc), C.byref(h))
        self._h = h
        self.Nv, self.n, self.L, self.cw = Nv, Nv ** 3, L, cw
        self.gamma, self.prefac, self.max_batch = gamma, prefac, max_batch
        self.nparams = param_count()
        th = init_params(seed, init_scale)
        assert th.size == self.nparams
        self.params = torch.as_tensor(th, dtype=torch.float32, device=self.device).contiguous()
        # BF16 copy of the weights (FP32 storage) that forward passes use; the
        # optimiser updates the FP32 master copy
        self.params_bf16 = torch.empty_like(self.params)
        self._round()
        self.grad = torch.zeros_like(self.params)
        self.m1 = torch.zeros_like(self.params)
        self.m2 = torch.zeros_like(self.params)
        self.step = 0

    @property
    def handle(self):
        return self._h

    def _round(self):
        _lib.call("dgfs_sur_round_bf16", int(self.nparams), _ptr(self.params), _ptr(self.params_bf16),
                  self._stream())

    def _stream(self):
        import torch
        return torch.cuda.current_stream(self.device).cuda_stream

    # -- pieces --------------------------------------------------------------
    def encode(self, f, npts, mom, x=None, lt=None):
        _lib.call("dgfs_sur_encode", self._h, _ptr(f), int(npts), _ptr(mom), _ptr(x), _ptr(lt), self._stream())

    def targets(self, Q, mom, npts, y):
        _lib.call("dgfs_sur_targets", self._h, _ptr(Q), _ptr(mom), int(npts), _ptr(y), self._stream())

    def decode(self, y, mom, nidx, Q, idx=None, project=True):
        _lib.call("dgfs_sur_decode", self._h, _ptr(y), _ptr(mom), _ptr(idx), int(nidx), _ptr(Q),
                  int(bool(project)), self._stream())

    def forward(self, x, lt, P, out, bf16=False, params=None):
        """Network outputs.  Default: the BF16 weights with FP32 activations on
        the FP32 MFMA (bf16=True rounds the activations to BF16 as well and
        runs the BF16 MFMA)."""
        _lib.call("dgfs_sur_forward", self._h, _ptr(self.params_bf16 if params is None else params), _ptr(x),
                  _ptr(lt), int(P), _ptr(out), int(bool(bf16)), self._stream())

    def backward(self, dout, grad=None, params=None):
        _lib.call("dgfs_sur_backward", self._h, _ptr(self.params_bf16 if params is None else params), _ptr(dout),
                  _ptr(self.grad if grad is None else grad), self._stream())

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        _lib.call("dgfs_sur_adam", int(self.nparams), _ptr(self.params), _ptr(self.grad), _ptr(self.m1),
                  _ptr(self.m2), float(lr), float(b1), float(b2), float(eps), int(self.step), self._stream())
        self._round()

    # -- training ------------------------------------------------------------
    def make_dataset(self, f, Q):
        """(f, Q) device tensors [npts][n] -> (x, lt, y, mom) device tensors."""
        import torch
        npts = f.shape[0]
        mom = torch.empty((npts, 8), dtype=torch.float64, device=self.device)
        x = torch.empty((npts, MODES), dtype=torch.float32, device=self.device)
        lt = torch.empty((npts,), dtype=torch.float32, device=self.device)
        y = torch.empty((npts, MODES), dtype=torch.float32, device=self.device)
        self.encode(f, npts, mom, x, lt)
        self.targets(Q, mom, npts, y)
        return x, lt, y, mom

    @staticmethod
    def mode_weights(device):
        """Loss weights of the modes: uniform, because the modes are
        coefficients in a grid-orthonormal basis, so their mean-square error
        is the L2 error of the decoded field (per unit prefac rho T^{gamma/2})."""
        import torch
        return torch.ones(MODES, dtype=torch.float32, device=device)

    def train(self, x, lt, y, epochs=10, batch=256, lr=1e-3, bf16=False, seed=0, log=None, lr_min=1e-5,
              relative=False):
        """Adam on the mean-square error of the modes (relative=True: each
        sample's squared error divided by its target's mean square, so small
        and large Q weigh alike); returns the per-epoch mean losses.  The
        gradient of the loss w.r.t. the outputs is formed elementwise;
        everything else runs in the HIP kernels."""
        import torch
        npts = x.shape[0]
        batch = min(batch, self.max_batch)
        w = self.mode_weights(self.device)
        if relative:
            ms = y.pow(2).mean(1)
            sw = 1.0 / (ms + 1e-3 * float(ms.mean()))
            sw = (sw / sw.mean()).view(-1, 1)
        else:
            sw = None
        out = torch.empty((batch, MODES), dtype=torch.float32, device=self.device)
        g = torch.Generator(device="cpu")
        g.manual_seed(seed)
        losses = []
        steps_total = epochs * max(1, npts // batch)
        k = 0
        for ep in range(epochs):
            perm = torch.randperm(npts, generator=g).to(self.device)
            # the epoch's loss is accumulated on the device: no host
            # synchronisation inside the step loop
            tot = torch.zeros((), dtype=torch.float64, device=self.device)
            nb = 0
            for b0 in range(0, npts - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                xb, ltb, yb = x[idx].contiguous(), lt[idx].contiguous(), y[idx].contiguous()
                self.forward(xb, ltb, batch, out, bf16=bf16)
                diff = out - yb
                if sw is not None:
                    diff = diff * sw[idx].sqrt()
                dout = (2.0 / (batch * MODES)) * w * (diff * sw[idx].sqrt() if sw is not None else diff)
                self.backward(dout.contiguous())
                lr_k = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * k / max(1, steps_total)))
                self.adam(lr_k)
                k += 1
                tot += (w * diff * diff).mean()
                nb += 1
            losses.append(float(tot) / max(nb, 1))
            if log:
                log(f"epoch {ep} loss {losses[-1]:.4e}")
        return losses

    def predict(self, x, lt, bf16=False):
        import torch
        npts = x.shape[0]
        y = torch.empty((npts, MODES), dtype=torch.float32, device=self.device)
        for b0 in range(0, npts, self.max_batch):
            nb = min(self.max_batch, npts - b0)
            self.forward(x[b0:b0 + nb].contiguous(), lt[b0:b0 + nb].contiguous(), nb, y[b0:b0 + nb], bf16=bf16)
        return y

    # -- evaluation with fallback ----------------------------------------------
    def eval_with_fallback(self, vhs_op, f, Q, npts, threshold=0.05, threshold_hi=1e-3, bf16=False, stream=None):
        """Q for npts contiguous points: surrogate where the indicator is
        <= threshold (and the high-mode indicator <= threshold_hi when that is
        > 0), the exact operator `vhs_op` (a VHSOperator of the same grid)
        elsewhere.  Returns the number of fallback points."""
        Nv, _, _, _, L = vhs_op.get_dims()
        if Nv != self.Nv or abs(L - self.L) > 1e-12 * self.L:
            raise ValueError(f"surrogate grid (N={self.Nv}, L={self.L}) differs from the exact operator's "
                             f"(N={Nv}, L={L})")
        nfb = C.c_int()
        _lib.call("dgfs_sur_eval", self._h, vhs_op._h, _ptr(self.params_bf16), _ptr(f), _ptr(Q), int(npts),
                  float(threshold), float(threshold_hi), int(bool(bf16)), C.byref(nfb),
                  stream if stream is not None else self._stream())
        return nfb.value

    def eval_checked(self, vhs_op, f, Q, npts, tol=1e-3, threshold=0.05, threshold_hi=1e-3, sample_frac=0.03,
                     nbins=16, min_per_bin=4, safety=0.7, seed=0, bf16=False, stream=None):
        """eval_with_fallback plus the sampled spectral check: a stratified
        sample of the surrogate candidates (ranked by indicator, cut into
        nbins bins) is evaluated exactly; bins whose sampled error would
        carry the batch's relative L2 error past safety * tol fall back to
        the exact operator.  Returns the statistics as a dict."""
        Nv, _, _, _, L = vhs_op.get_dims()
        if Nv != self.Nv or abs(L - self.L) > 1e-12 * self.L:
            raise ValueError(f"surrogate grid (N={self.Nv}, L={self.L}) differs from the exact operator's "
                             f"(N={Nv}, L={L})")
        chk = dgfs_sur_check(tol=float(tol), safety=float(safety), sample_frac=float(sample_frac),
                             nbins=int(nbins), min_per_bin=int(min_per_bin), seed=int(seed))
        st = dgfs_sur_check_stats()
        _lib.call("dgfs_sur_eval_checked", self._h, vhs_op._h, _ptr(self.params_bf16), _ptr(f), _ptr(Q), int(npts),
                  float(threshold), float(threshold_hi), C.byref(chk), int(bool(bf16)), C.byref(st),
                  stream if stream is not None else self._stream())
        return {"fallback_points": st.nfallback, "sampled_points": st.nsampled, "surrogate_points": st.nsurrogate,
                "rejected_bins": st.nrejected_bins, "estimated_rel_l2": st.est_rel_err}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.call("dgfs_sur_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
