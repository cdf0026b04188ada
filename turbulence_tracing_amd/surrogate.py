"""Learned correction of a cheap path-integrated ray estimate, with the full
RK4 tracer as the fallback and the measure of its error.

Not part of the reference (which has no surrogate).  The pieces:
  * ``tt_field_blocks`` / ``tt_path_features`` integrate the field along
    straight lines over z-blocks of k cells (with first-moment corrections)
    and advance each ray one drift-kick-drift step per block (the full
    tracer takes eight field evaluations per cell); the result is the
    estimate of the exit state [x, theta, y, phi, phase] plus 48 features;
  * a 4-layer MLP (48 -> 256 -> 256 -> 256 -> 8, GELU, bf16 weights and
    activations, fp32 accumulation; hand-written MFMA kernels in libtt for
    training and one fused kernel for inference) predicts the residual
    (full trace - estimate) for x, y, theta, phi, phase and two error
    bounds (position, angle);
  * the error heads are trained with a pinball loss, so each learns a
    quantile of the corrected error given the features (an upper bound that
    most rays stay under); ``Surrogate.calibrate`` then scales them on
    held-out traced rays so that every held-out ray whose corrected error
    exceeds a margin of the bar is flagged;
  * rays whose scaled bound exceeds the threshold (1 um / 1 urad by
    default), or whose estimate is not finite, are re-traced with the full
    tracer (``tt_trace_rk4_counted``, which reads the number of flagged rays
    from device memory, so the surrogate trace never waits on the host).
Training data are rays the tracer itself generates through seeded fields.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .pipeline import Config, Pipeline

NFEAT = 48
NOUT = 8  # 5 residuals, 2 error bounds, 1 pad
N_REG, N_ERR = 5, 2


def _p(t):
    return None if t is None else t.data_ptr()


_FRAG_CACHE = {}


def _frag_index(layer, fout):
    """Row / column indices of the packed fragments of one layer: fragment
    (nt, c) = 64 lanes x 8, lane l element j = W[16 nt + l % 16][k] with
    k = 32c + 8(l/16) + j for the input layer and k = 32c + 16(j/4) +
    4(l/16) + j%4 after a hidden layer (the order in which the previous
    layer's MFMA output sits in the lanes)."""
    key = (layer, fout)
    if key not in _FRAG_CACHE:
        nt_n = (fout + 15) // 16
        kc = 2 if layer == 0 else 8
        nt, c, lane, j = np.meshgrid(np.arange(nt_n), np.arange(kc), np.arange(64), np.arange(8), indexing="ij")
        n = 16 * nt + lane % 16
        g = lane // 16
        k = 32 * c + 8 * g + j if layer == 0 else 32 * c + 16 * (j // 4) + 4 * g + j % 4
        _FRAG_CACHE[key] = (torch.as_tensor(n.reshape(-1)), torch.as_tensor(k.reshape(-1)), nt_n, kc)
    return _FRAG_CACHE[key]


class MLP:
    """Weights [out][in] fp32 (master) + bf16 copies, Adam state, and the
    forward / backward passes as libtt GEMM calls."""

    def __init__(self, dims=(NFEAT, 256, 256, 256, NOUT), device="cuda", seed=0):
        self.dims = tuple(dims)
        self.device = device
        self.lib = _lib.load()
        g = torch.Generator().manual_seed(seed)
        self.W, self.b, self.Wb = [], [], []
        for i in range(len(dims) - 1):
            fan_in, fan_out = dims[i], dims[i + 1]
            bound = 1.0 / np.sqrt(fan_in)
            W = (torch.rand(fan_out, fan_in, generator=g, dtype=torch.float64) * 2 - 1) * bound
            b = (torch.rand(fan_out, generator=g, dtype=torch.float64) * 2 - 1) * bound
            self.W.append(W.float().to(device).contiguous())
            self.b.append(b.float().to(device).contiguous())
            self.Wb.append(self.W[-1].to(torch.bfloat16).contiguous())
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.dW = [torch.zeros_like(w) for w in self.W]
        self.db = [torch.zeros_like(b) for b in self.b]
        self.step = 0

    @property
    def nlayers(self):
        return len(self.W)

    def _gemm(self, A, sam, sak, B, sbk, sbn, M, N, K, bias=None, gelu=False, Cf=None, Cb=None, ldc=None, Zpre=None,
              ksplit=0):
        a = _lib.GemmArgs(_p(A), sam, sak, _p(B), sbk, sbn, M, N, K, _p(bias), 1 if gelu else 0, _p(Cf), _p(Cb),
                          ldc if ldc is not None else N, _p(Zpre), ksplit)
        self.lib.tt_gemm_bf16(C.byref(a), _lib.stream_ptr())

    def forward(self, X, out=None, keep=False):
        """X: bf16 [N, dims[0]] contiguous.  Returns fp32 [N, dims[-1]];
        keep=True stores activations and pre-activations for backward."""
        N = X.shape[0]
        H = X
        self.cache = [] if keep else None
        for l in range(self.nlayers):
            fin, fout = self.dims[l], self.dims[l + 1]
            last = l == self.nlayers - 1
            if last:
                Y = out if out is not None else torch.empty((N, fout), dtype=torch.float32, device=self.device)
                self._gemm(H, fin, 1, self.Wb[l], 1, fin, N, fout, fin, bias=self.b[l], Cf=Y, ldc=fout)
                if keep:
                    self.cache.append((H, None))
                return Y
            Hn = torch.empty((N, fout), dtype=torch.bfloat16, device=self.device)
            Z = torch.empty((N, fout), dtype=torch.float32, device=self.device) if keep else None
            self._gemm(H, fin, 1, self.Wb[l], 1, fin, N, fout, fin, bias=self.b[l], gelu=True, Cb=Hn, ldc=fout,
                       Zpre=Z)
            if keep:
                self.cache.append((H, Z))
            H = Hn

    def backward(self, dY):
        """dY: bf16 [N, dims[-1]] (gradient of the loss w.r.t. the output).
        Fills self.dW / self.db."""
        N = dY.shape[0]
        dZ = dY
        ks = 1024 if N >= 2048 else ((N + 31) // 32) * 32
        for l in range(self.nlayers - 1, -1, -1):
            fin, fout = self.dims[l], self.dims[l + 1]
            H, _ = self.cache[l]
            self.dW[l].zero_()
            self.db[l].zero_()
            # dW[out][in] = sum_n dZ[n][out] H[n][in]
            self._gemm(dZ, 1, fout, H, fin, 1, fout, fin, N, Cf=self.dW[l], ldc=fin, ksplit=ks)
            self.lib.tt_bias_grad(_p(dZ), N, fout, fout, _p(self.db[l]), _lib.stream_ptr())
            if l > 0:
                dH = torch.empty((N, fin), dtype=torch.float32, device=self.device)
                # dH[n][in] = sum_out dZ[n][out] W[out][in]
                self._gemm(dZ, fout, 1, self.Wb[l], fin, 1, N, fin, fout, Cf=dH, ldc=fin)
                Zp = self.cache[l - 1][1]
                dZn = torch.empty((N, fin), dtype=torch.bfloat16, device=self.device)
                self.lib.tt_gelu_backward(_p(dH), None, _p(Zp), _p(dZn), N * fin, _lib.stream_ptr())
                dZ = dZn

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step += 1
        st = _lib.stream_ptr()
        for l in range(self.nlayers):
            self.lib.tt_adam(_p(self.W[l]), _p(self.dW[l]), _p(self.mW[l]), _p(self.vW[l]), _p(self.Wb[l]),
                             self.W[l].numel(), lr, b1, b2, eps, self.step, wd, st)
            self.lib.tt_adam(_p(self.b[l]), _p(self.db[l]), _p(self.mb[l]), _p(self.vb[l]), None, self.b[l].numel(),
                             lr, b1, b2, eps, self.step, 0.0, st)

    def packed(self):
        """Weights and biases in the fragment layout of tt_mlp_infer (bf16
        [296 KiB], fp32 [784]); rebuilt after every Adam step."""
        if getattr(self, "_packed_step", None) == self.step and getattr(self, "_packed", None) is not None:
            return self._packed
        if len(self.dims) != 5 or self.dims[1:4] != (256, 256, 256) or self.dims[0] > 64 or self.dims[4] > 16:
            raise ValueError(f"tt_mlp_infer serves (<=64, 256, 256, 256, <=16) networks, not {self.dims}")
        ws, bs = [], []
        for l in range(self.nlayers):
            n_idx, k_idx, nt, kc = _frag_index(l, self.dims[l + 1])
            Wp = torch.zeros((nt * 16, kc * 32), dtype=torch.bfloat16, device=self.device)
            Wp[:self.dims[l + 1], :self.dims[l]] = self.Wb[l]
            ws.append(Wp[n_idx.to(self.device), k_idx.to(self.device)].reshape(-1))
            bp = torch.zeros(nt * 16, dtype=torch.float32, device=self.device)
            bp[:self.dims[l + 1]] = self.b[l]
            bs.append(bp)
        self._packed = (torch.cat(ws).contiguous(), torch.cat(bs).contiguous())
        self._packed_step = self.step
        return self._packed

    def infer(self, feat, mu, isd, out=None):
        """Fused forward pass on fp32 features [n, dims[0]] (standardised with
        mu / isd inside the kernel): fp32 [n, dims[-1]]."""
        n = feat.shape[0]
        w, b = self.packed()
        Y = out if out is not None else torch.empty((n, self.dims[-1]), dtype=torch.float32, device=self.device)
        self.lib.tt_mlp_infer(_p(feat), feat.stride(0), feat.stride(1), n, self.dims[0], _p(mu), _p(isd), _p(w),
                              _p(b), _p(Y), Y.stride(0), self.dims[-1], _lib.stream_ptr())
        return Y

    def state(self):
        return {f"W{l}": self.W[l].cpu().numpy() for l in range(self.nlayers)} | \
               {f"b{l}": self.b[l].cpu().numpy() for l in range(self.nlayers)}

    def load_state(self, st):
        for l in range(self.nlayers):
            self.W[l].copy_(torch.as_tensor(st[f"W{l}"]))
            self.b[l].copy_(torch.as_tensor(st[f"b{l}"]))
            self.Wb[l].copy_(self.W[l].to(torch.bfloat16))


@dataclass
class Norm:
    """Feature standardisation and target scales."""
    mu: np.ndarray
    isd: np.ndarray
    s_pos: float  # metres per unit of the x/y residual outputs
    s_ang: float  # radians per unit of the angle outputs
    s_ph: float   # radians per unit of the phase output
    e_pos: float = None  # metres per unit of the position error head (None: s_pos)
    e_ang: float = None  # radians per unit of the angle error head (None: s_ang); 0 flags nothing

    def __post_init__(self):
        if self.e_pos is None:
            self.e_pos = self.s_pos
        if self.e_ang is None:
            self.e_ang = self.s_ang


@dataclass
class Dataset:
    feat: torch.Tensor     # [N, 48] fp32
    target: torch.Tensor   # [N, 5] fp32, normalised residual (x, y, theta, phi, phase)
    resid: torch.Tensor    # [5, N] fp64 raw residual (m, m, rad, rad, rad)
    info: dict = field(default_factory=dict)


def block_fields(field4, grid, k=16):
    """Straight-line z-block integrals G and first moments M of the node field
    over blocks of k cells, column by column: [ny, nx, nb, 2, 4] float32
    ({G, M} per column and block), nb = ceil((nz - 1) / k)."""
    lib = _lib.load()
    nb = -(-(grid.nz - 1) // k)
    GM = torch.empty((grid.ny, grid.nx, nb, 2, 4), dtype=torch.float32, device=field4.device)
    lib.tt_field_blocks(_p(field4), C.byref(grid), k, _p(GM), _lib.stream_ptr())
    return GM


def path_estimate(field4, grid, s0, n=None, k=16, blocks=None, order=None, base=None, feat=None):
    """Cheap path estimate base [5, n] (fp64) and features [n, 48] (fp32,
    a feature-major view: strides (1, n)) over z-blocks of k cells.  ``blocks`` = block_fields(field4, grid, k)
    (built here when not given for k > 1; for k = 1 the steps read the node
    field itself unless blocks are given, with the same result); ``order`` (int32, optional) is the
    processing order of the rays; ``base`` / ``feat`` (optional) receive the
    results (contiguous, of those shapes)."""
    lib = _lib.load()
    n = s0.shape[1] if n is None else n
    ld = s0.stride(0)
    dev = s0.device
    GM = blocks if blocks is not None or k == 1 else block_fields(field4, grid, k)
    if GM is not None and GM.shape[2] != -(-(grid.nz - 1) // k):
        raise ValueError(f"path_estimate: blocks were built for another block size than k={k}")
    if base is None:
        base = torch.empty((5, n), dtype=torch.float64, device=dev)
    if feat is None:
        feat = torch.empty((NFEAT, n), dtype=torch.float32, device=dev).t()
    if n == 0:  # an empty ray set
        return base, feat
    if tuple(base.shape) != (5, n) or not base.is_contiguous():
        raise ValueError("path_estimate: base must be a contiguous [5, n] tensor")
    if tuple(feat.shape) != (n, NFEAT) or feat.stride(0) != 1 or feat.stride(1) < max(n, 1):
        raise ValueError("path_estimate: feat must be a feature-major [n, 48] view (strides (1, >= n))")
    lib.tt_path_features(_p(field4), _p(GM), k, C.byref(grid), _p(s0), ld, n, _p(order), _p(feat),
                         feat.stride(1), _p(base), n, _lib.stream_ptr())
    return base, feat


def choose_block(field4, grid, s0, rf_ref, phase_ref, candidates=(32, 16, 8, 4, 2, 1), p999_max=0.5e-6,
                 max_ang=1e-6, max_pos=1e-6):
    """Largest z-block size whose path estimate, on probe rays s0 already
    traced by the full tracer (rf_ref, phase_ref), keeps the angle error's
    99.9th percentile within p999_max and its maximum within max_ang
    (radians), and the position error within max_pos (metres).  Falls back
    to k = 1.  Returns (k, {k: (p999, max angle, max position)})."""
    stats = {}
    for k in sorted(candidates, reverse=True):
        base, _ = path_estimate(field4, grid, s0, k=k)
        ep, ea, _ = errors(torch.stack([base[0], base[1], base[2], base[3]]), base[4], rf_ref, phase_ref)
        ea = ea[torch.isfinite(ea)].double()
        ep = ep[torch.isfinite(ep)].double()
        sub = ea[:1 << 22]
        p999 = float(torch.quantile(sub, 0.999)) if sub.numel() else 0.0
        stats[k] = (p999, float(ea.max()) if ea.numel() else 0.0, float(ep.max()) if ep.numel() else 0.0)
        if stats[k][0] <= p999_max and stats[k][1] <= max_ang and stats[k][2] <= max_pos:
            return k, stats
    return 1, stats


def residual(rf, phase, base):
    """[5, n] fp64 residual full - estimate, in (x, y, theta, phi, phase) order."""
    return torch.stack([rf[0] - base[0], rf[2] - base[2], rf[1] - base[1], rf[3] - base[3], phase - base[4]])


def build_dataset(n_fields, rays_per_field, cfg=None, seed=1000, device="cuda", log=None, k=16):
    """Rays traced by the tracer through ``n_fields`` seeded fields: features
    of their cheap path estimate and the residual of the full trace."""
    cfg = cfg or Config()
    # few rays per field: the plane forms would not pay for their build (the
    # trace's results are the same with or without them)
    cfg = Config(**{**cfg.__dict__, "n_rays": rays_per_field, "plane_forms": False})
    p = Pipeline(cfg, device)
    feats, resids = [], []
    for f in range(n_fields):
        s = seed + 2 * f
        p.generate_field(s)
        p.gradient()
        p.beam(s + 1)
        p.sort()
        p.trace()
        base, feat = path_estimate(p.field4, p.grid, p.s0, k=k)
        feats.append(feat)
        resids.append(residual(p.rf, p.phase, base))
        if log and (f + 1) % max(1, n_fields // 8) == 0:
            log(f"dataset: field {f + 1}/{n_fields}")
    feat = torch.cat(feats)
    resid = torch.cat(resids, dim=1)
    return feat, resid, p


def make_norm(feat, resid):
    f = feat.double()
    mu = f.mean(0)
    sd = f.std(0)
    isd = torch.where(sd > 0, 1.0 / sd, torch.zeros_like(sd))
    r = resid.double()
    ok = torch.isfinite(r).all(0)
    r = r[:, ok]
    s_pos = float(r[0:2].std()) or 1e-9
    s_ang = float(r[2:4].std()) or 1e-12
    s_ph = float(r[4].std()) or 1e-9
    return Norm(mu.float().cpu().numpy(), isd.float().cpu().numpy(), s_pos, s_ang, s_ph)


def normalise_targets(resid, norm):
    s = torch.tensor([norm.s_pos, norm.s_pos, norm.s_ang, norm.s_ang, norm.s_ph], dtype=torch.float64,
                     device=resid.device)
    t = (resid / s[:, None]).T.float().contiguous()
    return torch.nan_to_num(t, nan=0.0, posinf=0.0, neginf=0.0)


class Surrogate:
    def __init__(self, norm, mlp=None, device="cuda", seed=0, k=16, tau=0.9):
        self.norm = norm
        self.k = k  # z-block size of the path estimate the network corrects
        self.tau = tau  # quantile the error heads learn
        self.device = device
        self.mlp = mlp or MLP(device=device, seed=seed)
        self.lib = _lib.load()
        self.mu = torch.as_tensor(norm.mu, device=device)
        self.isd = torch.as_tensor(norm.isd, device=device)
        self.w_reg = torch.tensor([1.0, 1.0, 1.0, 1.0, 0.5], dtype=torch.float32, device=device)
        self.err_lo = torch.tensor([0, 2], dtype=torch.int32, device=device)
        self.err_hi = torch.tensor([2, 4], dtype=torch.int32, device=device)

    def standardize(self, feat):
        n = feat.shape[0]
        X = torch.empty((n, NFEAT), dtype=torch.bfloat16, device=self.device)
        self.lib.tt_standardize_bf16(_p(feat), feat.stride(0), feat.stride(1), n, NFEAT, _p(self.mu), _p(self.isd),
                                     _p(X), NFEAT, _lib.stream_ptr())
        return X

    def train(self, feat, target, epochs=4, batch=8192, lr=2e-3, lr_min=5e-5, wd=0.0, log=None, seed=0):
        """Adam with cosine decay on (features, normalised targets)."""
        N = feat.shape[0]
        X = self.standardize(feat)
        gen = torch.Generator(device=self.device).manual_seed(seed)
        steps = epochs * (N // batch)
        loss = torch.zeros(2, dtype=torch.float64, device=self.device)
        dY = torch.empty((batch, NOUT), dtype=torch.bfloat16, device=self.device)
        Y = torch.empty((batch, NOUT), dtype=torch.float32, device=self.device)
        hist = []
        st = 0
        for ep in range(epochs):
            perm = torch.randperm(N, generator=gen, device=self.device)
            loss.zero_()
            for b in range(N // batch):
                idx = perm[b * batch:(b + 1) * batch]
                xb = X.index_select(0, idx)
                tb = target.index_select(0, idx)
                self.mlp.forward(xb, out=Y, keep=True)
                self.lib.tt_surrogate_loss(_p(Y), NOUT, _p(tb), N_REG, _p(dY), NOUT, batch, N_REG, _p(self.w_reg),
                                           N_ERR, _p(self.err_lo), _p(self.err_hi), float(self.tau), NOUT,
                                           _p(loss), _lib.stream_ptr())
                self.mlp.backward(dY)
                lr_t = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * st / max(1, steps)))
                self.mlp.adam(lr_t, wd=wd)
                st += 1
            l = (loss / (N // batch * batch)).cpu().numpy()
            hist.append(l)
            if log:
                log(f"epoch {ep + 1}/{epochs}: mse {l[0]:.4g}  err-head pinball {l[1]:.4g}")
        return hist

    def calibrate(self, feat, resid, bar_pos=1e-6, bar_ang=1e-6, margin=0.8, max_scale=1e6):
        """Scale the error heads on held-out rays (features ``feat`` [n, 48]
        and true residuals ``resid`` [5, n] of their path estimates, from
        the full tracer) so that each held-out ray whose corrected position
        or angle error exceeds ``margin`` times the bar gets a scaled bound
        above the bar (and so is flagged by apply / trace at thresholds
        bar_pos, bar_ang).  With no such ray the heads flag nothing.  A head
        output of zero on such a ray caps the scale at max_scale times the
        bar over the smallest nonzero head output, which flags nearly every
        ray.  Returns a dict with the chosen scales and, on the held-out rays,
        the flagged fraction and the largest unflagged errors."""
        Y = self.predict(feat).double()
        n = Y.shape[0]
        r = resid.double()
        ok = torch.isfinite(r).all(0)
        ep = torch.maximum((r[0] - Y[:, 0] * self.norm.s_pos).abs(), (r[1] - Y[:, 1] * self.norm.s_pos).abs())
        ea = torch.maximum((r[2] - Y[:, 2] * self.norm.s_ang).abs(), (r[3] - Y[:, 3] * self.norm.s_ang).abs())
        out = {}
        for name, err, head, bar in (("pos", ep, Y[:, 5].abs(), bar_pos), ("ang", ea, Y[:, 6].abs(), bar_ang)):
            bad = ok & (err > margin * bar)
            if bool(bad.any()):
                hmin = head[bad].min()
                floor = head[head > 0].min() if bool((head > 0).any()) else torch.tensor(1.0, device=head.device)
                scale = float(bar / torch.clamp(hmin, min=float(floor) / max_scale)) * (1 + 1e-6)
            else:
                scale = 0.0
            setattr(self.norm, "e_" + name, scale)
            flagged = head * scale > bar
            out[name] = {"scale": scale, "n_over_margin": int(bad.sum()), "flagged_frac": float(flagged.double().mean()),
                         "max_unflagged": float(err[ok & ~flagged].max()) if bool((ok & ~flagged).any()) else 0.0}
        both = (Y[:, 5].abs() * self.norm.e_pos > bar_pos) | (Y[:, 6].abs() * self.norm.e_ang > bar_ang) | ~ok
        out["flagged_frac"] = float(both.double().mean())
        out["rays"] = n
        return out

    def predict(self, feat, out=None):
        """Network output [n, 8] for fp32 features [n, 48] (fused kernel)."""
        return self.mlp.infer(feat, self.mu, self.isd, out=out)

    def apply(self, base, Y, t_pos=1e-6, t_ang=1e-6, want_list=True, out=None, first=0):
        """Corrected exit state and fallback flags.  ``out`` = (rf, phase,
        flag, list, count) views to write into (rf may be a column slice of a
        larger 4 x N tensor); ``first`` is added to the listed ray indices."""
        n = base.shape[1]
        if out is None:
            rf = torch.empty((4, n), dtype=torch.float64, device=self.device)
            phase = torch.empty(n, dtype=torch.float64, device=self.device)
            flag = torch.empty(n, dtype=torch.int32, device=self.device)
            lst = torch.empty(n, dtype=torch.int32, device=self.device) if want_list else None
            cnt = torch.zeros(1, dtype=torch.int32, device=self.device)
        else:
            rf, phase, flag, lst, cnt = out
        sc = (C.c_float * 5)(self.norm.s_pos, self.norm.s_ang, self.norm.s_ph, self.norm.e_pos, self.norm.e_ang)
        th = (C.c_double * 2)(t_pos, t_ang)
        self.lib.tt_surrogate_apply(_p(base), base.stride(0), _p(Y), Y.stride(0), n, sc, th, _p(rf), rf.stride(0),
                                    _p(phase), _p(flag), _p(lst), _p(cnt), int(first), _lib.stream_ptr())
        return rf, phase, flag, lst, cnt

    def _side_streams(self):
        if getattr(self, "_streams", None) is None:
            self._streams = (torch.cuda.Stream(device=self.device), torch.cuda.Stream(device=self.device))
        return self._streams

    def trace(self, field4, grid, s0, h=0.5, t_pos=1e-6, t_ang=1e-6, blocks=None, timings=None, chunks=1,
              cT=None):
        """Surrogate + fallback trace of rays s0 (6 x n, device): the block
        path estimate, the network's correction, and the full tracer on the
        rays the network flags (predicted position error above t_pos metres
        or angle error above t_ang radians, or a non-finite estimate, which
        includes every ray the full trace would hand over to the time
        integrator).  The fallback is the z-march, then the time integrator
        on the flagged rays it hands over, over the light-travel budget cT
        (metres; default sqrt(8) times the grid's largest z, as solve's time
        span, PT:312-331).  Returns rf (4 x n), phase (n) and the flag per ray
        (1 = re-traced); ``self.status`` holds the re-traced rays' status as
        tt_trace_rk4 / tt_trace_time leave it (0 for the others).

        With ``chunks`` > 1 the rays are cut into that many batches: the path
        estimate of batch i + 1 runs on one stream while the network and the
        correction of batch i run on another.  The two kernels then compete
        for the same CUs: on MI355X this measured the same as one batch at
        512^3 (k = 16) and ~1 ms slower per 1e7 rays at 1024^3 (k = 1), so
        the default is one batch.  ``timings`` (a dict) receives the total
        and fallback ms (and per-stage ms when chunks == 1)."""
        n = s0.shape[1]
        dev = self.device
        if n == 0:  # an empty ray set
            return (torch.empty((4, 0), dtype=torch.float64, device=dev), torch.empty(0, dtype=torch.float64, device=dev),
                    torch.empty(0, dtype=torch.int32, device=dev))
        main = torch.cuda.current_stream()
        t0 = None
        if timings is not None:
            t0 = torch.cuda.Event(enable_timing=True)
            t0.record(main)
        if blocks is None and self.k > 1:
            blocks = block_fields(field4, grid, self.k)
        rf = torch.empty((4, n), dtype=torch.float64, device=dev)
        phase = torch.empty(n, dtype=torch.float64, device=dev)
        flag = torch.empty(n, dtype=torch.int32, device=dev)
        lst = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        cnt = torch.zeros(1, dtype=torch.int32, device=dev)
        chunks = max(1, min(int(chunks), -(-n // 4096)))
        step = -(-(-(-n // chunks)) // 256) * 256
        bounds = [(a, min(n, a + step)) for a in range(0, n, step)]
        bases = [torch.empty((5, b - a), dtype=torch.float64, device=dev) for a, b in bounds]
        feats = [torch.empty((NFEAT, b - a), dtype=torch.float32, device=dev).t() for a, b in bounds]
        Ys = [torch.empty((b - a, NOUT), dtype=torch.float32, device=dev) for a, b in bounds]
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)] if (timings is not None and
                                                                          len(bounds) == 1) else None
        if len(bounds) == 1:
            if ev:
                ev[0].record(main)
            path_estimate(field4, grid, s0, k=self.k, blocks=blocks, base=bases[0], feat=feats[0])
            if ev:
                ev[1].record(main)
            self.predict(feats[0], out=Ys[0])
            if ev:
                ev[2].record(main)
            self.apply(bases[0], Ys[0], t_pos, t_ang, out=(rf, phase, flag, lst, cnt))
            if ev:
                ev[3].record(main)
        else:
            sp, sm = self._side_streams()
            ready = torch.cuda.Event()
            ready.record(main)
            sp.wait_event(ready)
            sm.wait_event(ready)
            for (a, b), base, feat, Y in zip(bounds, bases, feats, Ys):
                with torch.cuda.stream(sp):
                    path_estimate(field4, grid, s0[:, a:b], k=self.k, blocks=blocks, base=base, feat=feat)
                    done = torch.cuda.Event()
                    done.record(sp)
                with torch.cuda.stream(sm):
                    sm.wait_event(done)
                    self.predict(feat, out=Y)
                    self.apply(base, Y, t_pos, t_ang, out=(rf[:, a:b], phase[a:b], flag[a:b], lst, cnt), first=a)
            main.wait_stream(sp)
            main.wait_stream(sm)
        if timings is not None:
            t1 = torch.cuda.Event(enable_timing=True)
            t1.record(main)
        # the flagged rays, their number read by the kernel from cnt: no host
        # synchronisation, so waves of rays and graph capture stay unblocked;
        # then those of them the z-march handed over (status 1), in time
        if cT is None:
            cT = float(np.sqrt(8.0) * max(grid.z0 + (grid.nz - 1) * grid.dz, 0.5 * (grid.nz - 1) * grid.dz))
        opts = _lib.TraceOpts(h, cT, 0)
        status = torch.zeros(n, dtype=torch.int32, device=dev)
        self.lib.tt_trace_rk4_counted(_p(field4), C.byref(grid), C.byref(opts), _p(s0), s0.stride(0), n, _p(lst),
                                      _p(cnt), _p(rf), n, _p(phase), _p(status), _lib.stream_ptr())
        self.lib.tt_trace_time(_p(field4), C.byref(grid), C.byref(opts), _p(s0), s0.stride(0), n, None, _p(rf), n,
                               _p(phase), None, n, _p(status), 1, _lib.stream_ptr())
        self.n_fallback = cnt  # device int32: rays re-traced by the last call
        self.status = status
        if timings is not None:
            t2 = torch.cuda.Event(enable_timing=True)
            t2.record(main)
            torch.cuda.synchronize()
            timings.update(estimate_and_network=t0.elapsed_time(t1), fallback=t1.elapsed_time(t2),
                           total=t0.elapsed_time(t2), n_fallback=int(cnt.item()), chunks=len(bounds))
            if ev:
                timings.update(path=ev[0].elapsed_time(ev[1]), mlp=ev[1].elapsed_time(ev[2]),
                               apply=ev[2].elapsed_time(ev[3]))
        return rf, phase, flag

    def save(self, path):
        np.savez(path, **self.mlp.state(), mu=self.norm.mu, isd=self.norm.isd,
                 scales=np.array([self.norm.s_pos, self.norm.s_ang, self.norm.s_ph, self.norm.e_pos,
                                  self.norm.e_ang]), k=np.array(self.k), tau=np.array(self.tau))

    @classmethod
    def load(cls, path, device="cuda"):
        st = np.load(path)
        sc = st["scales"]
        norm = Norm(st["mu"], st["isd"], float(sc[0]), float(sc[1]), float(sc[2]), float(sc[3]), float(sc[4]))
        s = cls(norm, device=device, k=int(st["k"]), tau=float(st["tau"]) if "tau" in st.files else 0.9)
        s.mlp.load_state(st)
        return s


def errors(rf, phase, ref_rf, ref_phase):
    """Per-ray |error|: position (m, max of x, y), angle (rad, max of theta,
    phi), phase (rad)."""
    ep = torch.maximum((rf[0] - ref_rf[0]).abs(), (rf[2] - ref_rf[2]).abs())
    ea = torch.maximum((rf[1] - ref_rf[1]).abs(), (rf[3] - ref_rf[3]).abs())
    eph = (phase - ref_phase).abs()
    return ep, ea, eph


def summarize(e):
    e = e[torch.isfinite(e)].double()
    if e.numel() == 0:
        return {}
    q = torch.quantile(e[torch.randperm(e.numel(), device=e.device)[:1 << 20]], torch.tensor(
        [0.5, 0.99, 0.999], dtype=torch.float64, device=e.device)).cpu().numpy()
    return {"rms": float(e.pow(2).mean().sqrt()), "p50": float(q[0]), "p99": float(q[1]), "p999": float(q[2]),
            "max": float(e.max())}
