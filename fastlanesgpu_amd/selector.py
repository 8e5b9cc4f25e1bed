"""Learned per-vector encoding selection and mixed-encoding columns.

No reference counterpart: the reference picks one bit width per column
(find_bw, fastlanes/src/include/ssb_utils.h:222-239) and never chooses an
encoding.  Here every 1024-value vector is stored with the smallest of four
encodings -- bit-pack, frame of reference, FastLanes delta, FastLanes RLE --
in FastLanes transposed row order, so the column decodes exactly whatever was
chosen (include/fls/fls.h, FLS_ENC_MIXED).

Three ways to choose, all on the GPU:
  * ``mixed_encode_trial``: try all encodings (four full encodes) and keep
    each vector's smallest -- the exhaustive baseline;
  * ``exhaustive_select``: the same choice from statistics only;
  * ``LearnedSelector``: 16 features of a 1/8 sample -> MLP 16-64-64-40 (BF16
    MFMA kernels) -> encoding; vectors the network is unsure of fall back to
    exhaustive search inside the planning pass.
The MLP is trained on the GPU by the library's own gradient and Adam kernels,
on labels from the exhaustive search.
"""
import ctypes

import numpy as np

from . import ENC_MIXED, VECTOR_SIZE, Column, _dev_u32, _dev_u64, _stream, call, dptr, load

MIX_BITPACK, MIX_FOR, MIX_DELTA, MIX_RLE = 0, 1, 2, 3
ENCODING_NAMES = ("bitpack", "for", "delta", "rle")
N_FEATURES, N_HIDDEN, N_CLASSES = 16, 64, 40
BUCKET_TOP = (0, 2, 4, 6, 8, 12, 16, 20, 24, 32)
VEC_META_BYTES = 16  # vmeta + aux per vector


def width_bucket(w):
    for b, top in enumerate(BUCKET_TOP):
        if w <= top:
            return b
    return len(BUCKET_TOP) - 1


def class_of(enc, width):
    return 10 * enc + width_bucket(width)


def class_encoding(c):
    return np.asarray(c) // 10


def n_vectors(n):
    return (n + VECTOR_SIZE - 1) // VECTOR_SIZE


# ------------------------------------------------------------- training data
def zipf_cdf(s=1.1, n_ranks=4096):
    """Inverse-CDF table of Zipf(s) over ranks 1..n_ranks, scaled to 2^32:
    entry r is the first u32 at which rank r+2 begins (last entry 2^32-1)."""
    p = np.arange(1, n_ranks + 1, dtype=np.float64) ** -s
    c = np.cumsum(p) / p.sum()
    t = np.minimum(np.floor(c * 2.0 ** 32), 2.0 ** 32 - 1).astype(np.uint32)
    t[-1] = 0xFFFFFFFF
    return t


def gen_train_vectors(seed, first_vector, n_vec, cdf=None, out=None, stream=None):
    """n_vec training vectors (uniform, Zipf s=1.1, sorted + Gaussian noise
    sigma 16, geometric runs of mean 32; family chosen per vector)."""
    import torch

    if cdf is None:
        cdf = torch.from_numpy(zipf_cdf().view(np.int32)).cuda()
    if out is None:
        out = torch.empty(n_vec * VECTOR_SIZE, dtype=torch.int32, device="cuda")
    call("fls_gen_train_vectors", seed, first_vector, n_vec, dptr(cdf), dptr(out), _stream(stream))
    return out


def vector_features(d_in, n, out=None, stream=None):
    """[n_vectors, 16] float32 features of a 1/8 sample of every vector."""
    import torch

    nv = n_vectors(n)
    if out is None:
        out = torch.empty((max(1, nv), N_FEATURES), dtype=torch.float32, device="cuda")
    call("fls_vector_features", dptr(d_in), n, dptr(out), _stream(stream))
    return out


def exhaustive_select(d_in, n, sizes=True, cls=None, stream=None):
    """Best class per vector and (optionally) the exact size in bytes of each
    of the four encodings, [n_vectors, 4]."""
    import torch

    nv = n_vectors(n)
    if cls is None:
        cls = torch.empty(max(1, nv), dtype=torch.uint8, device="cuda")
    s4 = torch.empty((max(1, nv), 4), dtype=torch.int32, device="cuda") if sizes else None
    call("fls_exhaustive_select", dptr(d_in), n, dptr(cls), dptr(s4), _stream(stream))
    return cls, s4


def build_training_set(seed, n_vec, chunk=1 << 16, stream=None):
    """Features, exhaustive labels and sizes of n_vec training vectors, built
    chunk by chunk on the GPU (only the chunk's values are ever resident)."""
    import torch

    feat = torch.empty((n_vec, N_FEATURES), dtype=torch.float32, device="cuda")
    label = torch.empty(n_vec, dtype=torch.uint8, device="cuda")
    sizes = torch.empty((n_vec, 4), dtype=torch.int32, device="cuda")
    cdf = torch.from_numpy(zipf_cdf().view(np.int32)).cuda()
    buf = torch.empty(min(chunk, n_vec) * VECTOR_SIZE, dtype=torch.int32, device="cuda")
    for v0 in range(0, n_vec, chunk):
        m = min(chunk, n_vec - v0)
        gen_train_vectors(seed, v0, m, cdf, buf, stream)
        n = m * VECTOR_SIZE
        call("fls_vector_features", dptr(buf), n, dptr(feat[v0:]), _stream(stream))
        call("fls_exhaustive_select", dptr(buf), n, dptr(label[v0:]), dptr(sizes[v0:]), _stream(stream))
    return feat, label, sizes


# ----------------------------------------------------------------------- MLP
class Mlp:
    """MLP 16 -> 64 -> 64 -> 40 with FP32 master weights, Adam state and the
    BF16 copy the kernels read, all on the device."""

    def __init__(self, seed=0, weights=None):
        import torch

        self.n_params = int(load().fls_mlp_param_count())
        if weights is None:
            weights = self.init_weights(seed)
        self.w = torch.from_numpy(np.asarray(weights, np.float32).copy()).cuda()
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.grad = torch.zeros_like(self.w)
        self.wb = torch.empty(self.n_params, dtype=torch.int16, device="cuda")
        self.step = 0
        self._work = None
        self._loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        call("fls_mlp_to_bf16", dptr(self.w), dptr(self.wb), _stream(None))

    @staticmethod
    def init_weights(seed):
        """He-normal weights (std sqrt(2/fan_in)), zero biases."""
        rng = np.random.default_rng(seed)
        parts = []
        for fi, fo in ((N_FEATURES, N_HIDDEN), (N_HIDDEN, N_HIDDEN), (N_HIDDEN, N_CLASSES)):
            parts.append(rng.normal(0.0, np.sqrt(2.0 / fi), fi * fo))
            parts.append(np.zeros(fo))
        return np.concatenate(parts).astype(np.float32)

    def weights(self):
        return self.w.cpu().numpy().copy()

    def forward(self, feat, n=None, logits=False, stream=None):
        """(class uint8 [n], confidence float32 [n], logits [n,40] or None)."""
        import torch

        n = feat.shape[0] if n is None else n
        cls = torch.empty(max(1, n), dtype=torch.uint8, device="cuda")
        conf = torch.empty(max(1, n), dtype=torch.float32, device="cuda")
        lg = torch.empty((max(1, n), N_CLASSES), dtype=torch.float32, device="cuda") if logits else None
        call("fls_mlp_forward", dptr(self.wb), dptr(feat), n, dptr(cls), dptr(conf), dptr(lg), _stream(stream))
        return cls[:n], conf[:n], (lg[:n] if logits else None)

    def _workspace(self, n):
        import torch

        need = int(load().fls_mlp_grad_workspace_bytes(n))
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.uint8, device="cuda")
        return self._work

    def accumulate_grad(self, feat, label, n=None, inv_batch=None, loss=None, stream=None):
        """grad += d(loss * inv_batch)/dw over n samples; *loss += summed loss."""
        n = feat.shape[0] if n is None else n
        inv = 1.0 / n if inv_batch is None else inv_batch
        call("fls_mlp_grad", dptr(self.wb), dptr(feat), dptr(label), n, ctypes.c_float(inv),
             dptr(self._workspace(n)), dptr(self.grad), dptr(loss), _stream(stream))

    def adam(self, lr=2e-3, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        self.step += 1
        call("fls_mlp_adam", dptr(self.w), dptr(self.m), dptr(self.v), dptr(self.grad), dptr(self.wb), self.step,
             ctypes.c_float(lr), ctypes.c_float(beta1), ctypes.c_float(beta2), ctypes.c_float(eps),
             _stream(stream))

    def train(self, feat, label, epochs=4, batch=2048, lr=2e-3, seed=0, lr_decay=0.5):
        """Mini-batch Adam over (feat, label); returns the mean loss per epoch.
        The sample order of each epoch is a seeded permutation; the gradient
        of a step is bit-reproducible, so the whole run is."""
        import torch

        n = feat.shape[0]
        gen = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).cuda()
            f, y = feat[perm].contiguous(), label[perm].contiguous()
            self._loss.zero_()
            for s0 in range(0, n, batch):
                m = min(batch, n - s0)
                self.accumulate_grad(f[s0:s0 + m], y[s0:s0 + m], m, 1.0 / m, self._loss)
                self.adam(lr * (lr_decay ** ep))
            losses.append(float(self._loss.item()) / n)
        return losses


# -------------------------------------------------------------- mixed column
class MixedWorkspace:
    """Device buffers of one mixed-column encode, sized for n values and
    reusable across encodes of at most that many (no allocation per call)."""

    def __init__(self, n, trial=False):
        import torch

        nv = max(1, n_vectors(n))
        self.n = n
        self.cls = torch.empty(nv, dtype=torch.uint8, device="cuda")
        self.vmeta, self.aux = _dev_u64(nv), _dev_u64(nv)
        self.words = _dev_u32(int(load().fls_mixed_max_words(n)))
        self.total = torch.zeros(1, dtype=torch.int64, device="cuda")
        self.fallbacks = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.scratch = torch.empty(max(16, int(load().fls_encode_scratch_bytes(n))), dtype=torch.uint8,
                                   device="cuda")
        self.feat = torch.empty((nv, N_FEATURES), dtype=torch.float32, device="cuda")
        self.conf = torch.empty(nv, dtype=torch.float32, device="cuda")
        self.trial = trial_buffer(n) if trial else None

    def column(self, n):
        """The encoded MIXED column (views of this workspace)."""
        total = int(self.total.item())
        nv = n_vectors(n)
        return Column(ENC_MIXED, n, self.words[:max(1, total)], vmeta=self.vmeta[:max(1, nv)],
                      aux=self.aux[:max(1, nv)], n_words=total)


def mixed_plan(d_in, n, ws, conf=None, conf_min=0.0, stream=None):
    call("fls_mixed_plan_device", dptr(d_in), n, dptr(ws.cls), dptr(conf), ctypes.c_float(conf_min),
         dptr(ws.vmeta), dptr(ws.aux), dptr(ws.total), dptr(ws.fallbacks), dptr(ws.scratch), _stream(stream))


def mixed_pack(d_in, n, ws, stream=None):
    call("fls_mixed_pack_device", dptr(d_in), n, dptr(ws.vmeta), dptr(ws.aux), dptr(ws.words), _stream(stream))


def mixed_encode(d_in, n, ws, conf=None, conf_min=0.0, stream=None):
    """Single-pass encode with the classes in ws.cls (and the confidence
    fallback when conf is given): plan, look-back offsets and pack in one
    launch after a small fill.  The same column as plan + pack; slower at
    120M rows (see mixed_encode_classes), kept for callers that want one
    launch."""
    call("fls_mixed_encode_device", dptr(d_in), n, dptr(ws.cls), dptr(conf), ctypes.c_float(conf_min),
         dptr(ws.vmeta), dptr(ws.aux), dptr(ws.words), dptr(ws.total), dptr(ws.fallbacks), dptr(ws.scratch),
         _stream(stream))


def mixed_encode_classes(d_in, n, ws, stream=None):
    """Encode with the classes already in ws.cls: plan + pack (no host sync),
    which at 120M rows measured faster than the single-pass mixed_encode."""
    mixed_plan(d_in, n, ws, stream=stream)
    mixed_pack(d_in, n, ws, stream)


def mixed_encode_exhaustive(d_in, n, ws, stream=None):
    """Statistics-only exhaustive choice, then plan + pack."""
    call("fls_exhaustive_select", dptr(d_in), n, dptr(ws.cls), None, _stream(stream))
    mixed_encode_classes(d_in, n, ws, stream)


def trial_buffer(n):
    """Candidate-column workspace of the try-all-encodings search for up to n
    values (shareable by encodes that run one after another)."""
    import torch

    return torch.empty(int(load().fls_trial_workspace_bytes(n)), dtype=torch.uint8, device="cuda")


def mixed_encode_trial(d_in, n, ws, stream=None, buf=None):
    """Try-all-encodings search (four full encodes, keep the smallest)."""
    buf = ws.trial if buf is None else buf
    if buf is None:
        raise ValueError("no trial workspace: build with trial=True or pass buf=trial_buffer(n)")
    call("fls_mixed_encode_trial", dptr(d_in), n, dptr(ws.cls), dptr(ws.vmeta), dptr(ws.aux), dptr(ws.words),
         dptr(ws.total), dptr(buf), _stream(stream))


def mixed_decode(col, out=None, stream=None):
    """Decode a MIXED column to [n_vectors * 1024] values in transposed order."""
    import torch

    nv = n_vectors(col.n_rows)
    if out is None:
        out = torch.empty(max(1, nv) * VECTOR_SIZE, dtype=torch.int32, device="cuda")
    call("fls_mixed_decode_u32", dptr(col.words), dptr(col.vmeta), dptr(col.aux), nv, dptr(out), _stream(stream))
    return out


class LearnedSelector:
    """Per-vector encoding choice by the MLP, with exhaustive search as the
    fallback for vectors whose confidence is below ``conf_min``."""

    def __init__(self, mlp=None, conf_min=0.9):
        self.mlp = mlp if mlp is not None else Mlp()
        self.conf_min = conf_min

    @classmethod
    def trained(cls, seed=1, n_train=1 << 18, epochs=4, batch=2048, lr=2e-3, conf_min=0.9):
        feat, label, _ = build_training_set(seed, n_train)
        mlp = Mlp(seed)
        mlp.train(feat, label, epochs=epochs, batch=batch, lr=lr, seed=seed)
        return cls(mlp, conf_min)

    def choose(self, d_in, n, ws, stream=None):
        """features -> MLP -> ws.cls / ws.conf."""
        nv = n_vectors(n)
        call("fls_vector_features", dptr(d_in), n, dptr(ws.feat), _stream(stream))
        call("fls_mlp_forward", dptr(self.mlp.wb), dptr(ws.feat), nv, dptr(ws.cls), dptr(ws.conf), None,
             _stream(stream))

    def encode(self, d_in, n, ws, stream=None):
        """The learned encode: choose, then plan (with the confidence fallback)
        and pack.  Asynchronous; ws.column(n) reads the result."""
        self.choose(d_in, n, ws, stream)
        mixed_plan(d_in, n, ws, conf=ws.conf, conf_min=self.conf_min, stream=stream)
        mixed_pack(d_in, n, ws, stream)
