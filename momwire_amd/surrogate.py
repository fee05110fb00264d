"""Learned Sommerfeld surrogate: an MLP that replaces the six lambda-integrals
when a Sommerfeld grid is filled.

The network (4 hidden SiLU layers of 256, BF16 weights, FP32 accumulation on
MFMA; kernels in csrc/mw_mlp.hip) maps one node's (rho/lambda, (z+z')/lambda,
eps_r, sigma/(omega eps0)) to its four remainder surfaces I_rho^V, I_z^V,
I_rho^H, I_phi^H (normalised by -j omega mu/(4 pi) at lambda = 1).  It is
trained on the GPU by the project's own forward/backward/Adam kernels from
samples of the exact HIP engine, and it fills grid nodes inside its trained
domain; the engine fills every other node (R1 = 0 limits, nodes beyond the
trained extent, grounds outside the trained eps range), so the exact engine
is both the fallback and the yardstick of its error.

There is no counterpart in the reference.  Use:

    net = SommerfeldSurrogate.load()        # the shipped trained weights
    BSplineSolver(..., ground_model="sommerfeld", somm_surrogate=net)
"""

import ctypes
import math
import os
import time

import numpy as np

from . import _ground, _lib
from ._lib import MwMlp

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "somm_mlp.npz")

# trained domain (lambda = 1)
RHO_MAX = 4.0
H_MAX = 4.0
EPS_R = (5.0, 30.0)
SIGMA_RATIO = (0.6, 300.0)
# -j omega mu / (4 pi) of the normalised master: the output unit
OUT_UNIT = complex(0.0, -_ground.K2_REF * _ground.C_LIGHT * _ground.MU0_EXACT / (4 * np.pi))

FIN = [64, 256, 256, 256, 256]
FOUT = [256, 256, 256, 256, 64]
N_USED = 8      # outputs in use
NF_USED = 12    # input features in use


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise _lib.AcceleratorUnavailable("the surrogate runs on the GPU only")
    return torch


def init_params(seed=0):
    """Uniform(+-sqrt(6/fan_in)) weights, zero biases; pad features/outputs zero."""
    rng = np.random.default_rng(seed)
    flat = []
    for l, (i, f) in enumerate(zip(FIN, FOUT)):
        lim = math.sqrt(6.0 / (NF_USED if l == 0 else i))
        W = rng.uniform(-lim, lim, (f, i))
        if l == 0:
            W[:, NF_USED:] = 0.0
        if l == len(FIN) - 1:
            W *= 0.1
            W[N_USED:, :] = 0.0
        flat += [W.ravel(), np.zeros(f)]
    return np.concatenate(flat).astype(np.float32)


def sample_inputs(n, seed):
    """Training/validation nodes: 3/4 uniform over the (rho, h) box, 1/4 with
    R1 log-uniform in [1e-3, 1] (the steep near-field structure); eps_r
    uniform, sigma/(omega eps0) log-uniform over the trained ranges."""
    rng = np.random.default_rng(seed)
    n_box = (3 * n) // 4
    rho = np.empty(n)
    h = np.empty(n)
    rho[:n_box] = rng.uniform(0.0, RHO_MAX, n_box)
    h[:n_box] = rng.uniform(0.0, H_MAX, n_box)
    r1 = np.exp(rng.uniform(np.log(1e-3), 0.0, n - n_box))
    th = rng.uniform(0.0, 0.5 * np.pi, n - n_box)
    rho[n_box:] = r1 * np.cos(th)
    h[n_box:] = r1 * np.sin(th)
    er = rng.uniform(*EPS_R, n)
    s = np.exp(rng.uniform(np.log(SIGMA_RATIO[0]), np.log(SIGMA_RATIO[1]), n))
    x = np.stack([rho, h, er, s], axis=1)
    perm = rng.permutation(n)
    return x[perm]


def sample_grid_nodes(n, seed):
    """Nodes the grid fill actually queries: for random grounds over the
    trained eps range and extents log-uniform in [0.35, 6] wavelengths, the
    lattice nodes of each grid layout (_ground.grid_layout, the reference's
    region table) that lie inside the trained (rho, h) box, drawn until n.
    These weight the grazing (small theta) rows and the fine near-zone
    lattices the way the fill weights them, which a uniform box does not."""
    rng = np.random.default_rng(seed)
    parts, have = [], 0
    while have < n:
        er = rng.uniform(*EPS_R)
        s = np.exp(rng.uniform(np.log(SIGMA_RATIO[0]), np.log(SIGMA_RATIO[1])))
        r1 = np.exp(rng.uniform(np.log(0.35), np.log(6.0)))
        regions, scal = _ground.grid_layout(complex(er, -s), _ground.K2_REF, r1)
        _shapes, R, T = _ground.grid_nodes(regions, scal[3])
        rho, h = np.maximum(R * np.cos(T), 0.0), np.maximum(R * np.sin(T), 0.0)
        ok = (R > 0.0) & (rho <= RHO_MAX) & (h <= H_MAX)
        idx = np.nonzero(ok)[0]
        take = rng.choice(idx, size=min(idx.size, 2048, n - have), replace=False)
        parts.append(np.stack([rho[take], h[take], np.full(take.size, er),
                               np.full(take.size, s)], axis=1))
        have += take.size
    x = np.concatenate(parts)
    return x[rng.permutation(x.shape[0])]


def sample_training_inputs(n, seed, grid_frac=0.5):
    """The training set: grid_frac of the nodes from grid layouts
    (sample_grid_nodes), the rest from sample_inputs' box and near-field mix."""
    n_grid = int(round(grid_frac * n))
    x = np.concatenate([sample_inputs(n - n_grid, seed), sample_grid_nodes(n_grid, seed + 1000)])
    return x[np.random.default_rng(seed + 2000).permutation(n)]


def engine_targets(x, chunk=1 << 16, log=None):
    """Exact surfaces of the nodes x (n, 4) from the HIP engine, as the
    network's 8 normalised outputs (device float32 tensor (n, 8))."""
    torch = _torch()
    lib = _lib.load()
    n = x.shape[0]
    r1 = np.hypot(x[:, 0], x[:, 1])
    th = np.arctan2(x[:, 1], x[:, 0])
    eps = np.stack([x[:, 2], -x[:, 3]], axis=1)
    out = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    t0 = time.perf_counter()
    for i0 in range(0, n, chunk):
        i1 = min(n, i0 + chunk)
        dR = torch.from_numpy(r1[i0:i1]).cuda()
        dT = torch.from_numpy(th[i0:i1]).cuda()
        dE = torch.from_numpy(np.ascontiguousarray(eps[i0:i1])).cuda()
        S = torch.empty((i1 - i0, 4), dtype=torch.complex128, device="cuda")
        _lib.check(lib.mw_somm_surfaces_varied(dE.data_ptr(), _ground.K2_REF,
                                               _ground.K2_REF * _ground.C_LIGHT, _ground.MU0_EXACT,
                                               dR.data_ptr(), dT.data_ptr(), i1 - i0,
                                               _ground.GRID_RTOL, S.data_ptr(), None, st),
                   "engine samples")
        T = S / OUT_UNIT
        out[i0:i1] = torch.view_as_real(T).reshape(i1 - i0, 8).to(torch.float32)
        if log:
            torch.cuda.synchronize()
            log(f"engine samples {i1}/{n} ({time.perf_counter() - t0:.1f} s)")
    return out


class SommerfeldSurrogate:
    """The network, its device buffers and its training loop."""

    def __init__(self, params=None, seed=0):
        torch = _torch()
        self.lib = _lib.load()
        n_p = int(self.lib.mw_mlp_param_count())
        n_w = int(self.lib.mw_mlp_w16_count())
        if params is None:
            params = init_params(seed)
        params = np.asarray(params, dtype=np.float32)
        if params.size != n_p:
            raise ValueError(f"expected {n_p} parameters, got {params.size}")
        dev = torch.device("cuda")
        self.params = torch.from_numpy(params.copy()).to(dev)
        self.grads = torch.zeros(n_p, dtype=torch.float32, device=dev)
        self.m = torch.zeros(n_p, dtype=torch.float32, device=dev)
        self.v = torch.zeros(n_p, dtype=torch.float32, device=dev)
        self.w16 = torch.empty(n_w, dtype=torch.int16, device=dev)
        self.wt16 = torch.empty(n_w, dtype=torch.int16, device=dev)
        self.net = MwMlp(self.params.data_ptr(), self.grads.data_ptr(), self.m.data_ptr(),
                         self.v.data_ptr(), self.w16.data_ptr(), self.wt16.data_ptr())
        self.step = 0
        self._ws = None
        self._ws_n = 0
        _lib.check(self.lib.mw_mlp_sync_weights(ctypes.byref(self.net), self._st()), "sync weights")

    @staticmethod
    def _st():
        import torch

        return torch.cuda.current_stream().cuda_stream

    def workspace(self, n):
        torch = _torch()
        if self._ws is None or self._ws_n < n:
            nb = int(self.lib.mw_mlp_workspace_bytes(n))
            self._ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
            self._ws_n = n
        return self._ws

    # -- inference ------------------------------------------------------
    def predict(self, x, chunk=1 << 16):
        """Normalised surfaces (n, 4) complex for raw inputs x (n, 4)."""
        torch = _torch()
        x = np.asarray(x, dtype=np.float32)
        out = np.empty((x.shape[0], 8), dtype=np.float32)
        for i0 in range(0, x.shape[0], chunk):
            i1 = min(x.shape[0], i0 + chunk)
            dx = torch.from_numpy(np.ascontiguousarray(x[i0:i1])).cuda()
            dy = torch.empty((i1 - i0, 8), dtype=torch.float32, device="cuda")
            _lib.check(self.lib.mw_mlp_forward(ctypes.byref(self.net), dx.data_ptr(), i1 - i0,
                                               dy.data_ptr(), self.workspace(i1 - i0).data_ptr(),
                                               self._st()), "forward")
            out[i0:i1] = dy.cpu().numpy()
        return out.astype(np.float64).view(np.complex128).reshape(-1, 4)

    # -- training -------------------------------------------------------
    def loss_and_grads(self, dx, dt, n):
        """One forward/backward on device tensors; returns the loss (syncs)."""
        torch = _torch()
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        _lib.check(self.lib.mw_mlp_backward(ctypes.byref(self.net), dx.data_ptr(), dt.data_ptr(), n,
                                            loss.data_ptr(), self.workspace(n).data_ptr(),
                                            self._st()), "backward")
        return float(loss.cpu()[0])

    def train(self, x, t, steps, batch=4096, lr=2e-3, lr_min=2e-5, log=None, log_every=1000):
        """Adam with a cosine learning-rate decay over ``steps`` minibatches of
        the device data (x (n, 4) float32, t (n, 8) float32)."""
        torch = _torch()
        n = x.shape[0]
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        ws = self.workspace(batch)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(1234 + self.step)
        perm = torch.randperm(n, device="cuda", generator=gen)
        pos = 0
        hist = []
        t0 = time.perf_counter()
        st = self._st()
        for i in range(steps):
            if pos + batch > n:
                perm = torch.randperm(n, device="cuda", generator=gen)
                pos = 0
            sel = perm[pos:pos + batch]
            pos += batch
            xb = x.index_select(0, sel)
            tb = t.index_select(0, sel)
            _lib.check(self.lib.mw_mlp_backward(ctypes.byref(self.net), xb.data_ptr(), tb.data_ptr(),
                                                batch, loss.data_ptr(), ws.data_ptr(), st),
                       "backward")
            self.step += 1
            rate = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * i / max(steps - 1, 1)))
            _lib.check(self.lib.mw_mlp_adam_step(ctypes.byref(self.net), rate, 0.9, 0.999, 1e-8,
                                                 self.step, st), "adam")
            if (i + 1) % log_every == 0 or i + 1 == steps:
                lv = float(loss.cpu()[0])
                hist.append((i + 1, lv))
                if log:
                    log(f"step {i + 1}/{steps} loss {lv:.3e} lr {rate:.2e} "
                        f"({time.perf_counter() - t0:.1f} s)")
        return hist

    # -- weights --------------------------------------------------------
    def export(self):
        """BF16 weights (as uint16) and FP32 biases: what inference uses."""
        p = self.params.cpu().numpy()
        bits = p.view(np.uint32).astype(np.uint64)
        w16 = (((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)
        return dict(w16=w16, f32=p)

    def save(self, path=WEIGHTS, meta=None):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        e = self.export()
        is_w = np.zeros(e["f32"].size, dtype=bool)
        o = 0
        for i, f in zip(FIN, FOUT):
            is_w[o:o + f * i] = True
            o += f * i + f
        np.savez_compressed(path, w16=e["w16"][is_w], bias=e["f32"][~is_w],
                            meta=np.array(repr(meta or {})))

    @classmethod
    def load(cls, path=WEIGHTS):
        if not os.path.exists(path):
            raise FileNotFoundError(f"no trained surrogate at {path}")
        d = np.load(path, allow_pickle=False)
        n_p = sum(o * i + o for i, o in zip(FIN, FOUT))
        params = np.empty(n_p, dtype=np.float32)
        is_w = np.zeros(n_p, dtype=bool)
        o = 0
        for i, f in zip(FIN, FOUT):
            is_w[o:o + f * i] = True
            o += f * i + f
        params[is_w] = (d["w16"].astype(np.uint32) << 16).view(np.float32)
        params[~is_w] = d["bias"]
        net = cls(params)
        net.meta = str(d["meta"])
        return net

    # -- grid filling ---------------------------------------------------
    @staticmethod
    def in_domain(rho, h, eps_t):
        ok = (rho <= RHO_MAX) & (h <= H_MAX) & (np.hypot(rho, h) > 0.0)
        er, s = eps_t.real, -eps_t.imag
        ok_eps = (EPS_R[0] * (1 - 1e-9) <= er <= EPS_R[1] * (1 + 1e-9)
                  and SIGMA_RATIO[0] * (1 - 1e-9) <= s <= SIGMA_RATIO[1] * (1 + 1e-9))
        return ok & ok_eps

    def grid_filler(self):
        """Filler for ``_ground.GridMaster``: network on the nodes of the
        trained domain, exact engine on the rest; returns the engine's node
        count."""

        def fill(master, lib, st, cancel):
            torch = _torch()
            R = master.R1.cpu().numpy()
            T = master.theta.cpu().numpy()
            rho = np.maximum(R * np.cos(T), 0.0)
            h = np.maximum(R * np.sin(T), 0.0)
            learned = self.in_domain(rho, h, master.eps_t)
            idx_l = np.nonzero(learned)[0].astype(np.int32)
            idx_e = np.nonzero(~learned)[0]
            eps = master.eps_t
            if idx_e.size:
                dR = torch.from_numpy(R[idx_e]).cuda()
                dT = torch.from_numpy(T[idx_e]).cuda()
                S = torch.empty((idx_e.size, 4), dtype=torch.complex128, device="cuda")
                _lib.check(lib.mw_somm_surfaces(eps.real, eps.imag, _ground.K2_REF,
                                                _ground.K2_REF * _ground.C_LIGHT,
                                                _ground.MU0_EXACT, dR.data_ptr(), dT.data_ptr(),
                                                idx_e.size, _ground.GRID_RTOL, S.data_ptr(), cancel,
                                                st), "engine fallback")
                master.vals[torch.from_numpy(idx_e).cuda()] = S
            if idx_l.size:
                di = torch.from_numpy(idx_l).cuda()
                _lib.check(lib.mw_mlp_grid_values(ctypes.byref(self.net), eps.real, eps.imag,
                                                  master.R1.data_ptr(), master.theta.data_ptr(),
                                                  di.data_ptr(), idx_l.size, master.vals.data_ptr(),
                                                  OUT_UNIT.real, OUT_UNIT.imag,
                                                  self.workspace(idx_l.size).data_ptr(), st),
                           "surrogate grid")
            master.learned_nodes = int(idx_l.size)
            return int(idx_e.size)

        return fill


def main(argv=None):
    """Generate engine samples, train, and save the weights."""
    import argparse

    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--samples", type=int, default=1_000_000)
    ap.add_argument("--val", type=int, default=20_000)
    ap.add_argument("--steps", type=int, default=20_000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--grid-frac", type=float, default=0.0,
                    help="share of training nodes drawn from grid layouts (the shipped "
                         "weights: 0)")
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args(argv)

    def log(msg):
        print(f"[surrogate] {msg}", flush=True)

    torch = _torch()
    t0 = time.perf_counter()
    x = sample_training_inputs(a.samples, a.seed, a.grid_frac)
    t = engine_targets(x, log=log)
    xv = sample_training_inputs(a.val, a.seed + 1, a.grid_frac)
    tv = engine_targets(xv)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    log(f"{a.samples} engine samples in {t_gen:.1f} s")
    net = SommerfeldSurrogate(seed=a.seed)
    dx = torch.from_numpy(x.astype(np.float32)).cuda()
    t1 = time.perf_counter()
    hist = net.train(dx, t, a.steps, batch=a.batch, lr=a.lr, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    pv = net.predict(xv)
    tvh = tv.cpu().numpy().astype(np.float64).view(np.complex128).reshape(-1, 4)
    err = np.abs(pv - tvh).max(axis=1) / np.abs(tvh).max(axis=1)
    rms = float(np.sqrt(np.mean(np.abs(pv - tvh) ** 2) / np.mean(np.abs(tvh) ** 2)))
    meta = dict(samples=a.samples, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed,
                grid_frac=a.grid_frac,
                gen_s=round(t_gen, 2), train_s=round(t_train, 2), final_loss=hist[-1][1],
                val_rms_rel=rms, val_median_rel=float(np.median(err)),
                val_p99_rel=float(np.percentile(err, 99)))
    log(f"validation: {meta}")
    net.save(a.out, meta)
    log(f"saved {a.out}")
    return meta


if __name__ == "__main__":
    main()
