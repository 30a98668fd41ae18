"""Learned importance-sampling proposal: training driver.

The network (6 -> 64 -> 64 -> 64 direction bins + nB*nQ (band, g) bins) is
trained and evaluated by the library's own HIP kernels (BF16 MFMA forward and
backward, FP32 accumulation, Adam on FP32 master weights); device buffers come
from the library's allocator (dnsmc_amd.devarray).  Training data are the per-bundle tallies of plain runs:
every ray of a run in a deterministic tally mode gives (features of its
emitting cell, its direction bin, its (band, g) bin, |e - a|), and the loss is
the |e - a|-weighted cross-entropy of both heads -- the maximum-likelihood fit
of q to the zero-variance density p |x| (no reference counterpart; the
reference samples from the fixed Tmax proposal only).
"""
import ctypes as C

import numpy as np

from ._lib import check, lib
from .devarray import DeviceArray, concat

N_IN, N_HID, N_DIR = 6, 64, 64


def bf16_bits(x):
    """FP32 -> BF16 bit patterns, round to nearest even (as the kernels do)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def bf16_to_f32(h):
    return (np.asarray(h, np.uint32) << 16).view(np.float32)


def layout(nbin):
    noutp = ((N_DIR + nbin + 63) // 64) * 64
    o = {"W1": 0}
    o["b1"] = o["W1"] + N_IN * N_HID
    o["W2"] = o["b1"] + N_HID
    o["b2"] = o["W2"] + N_HID * N_HID
    o["W3"] = o["b2"] + N_HID
    o["b3"] = o["W3"] + N_HID * noutp
    o["P"] = o["b3"] + noutp
    o["noutp"] = noutp
    return o


def _ptr(t):
    return t.c


class ProposalNet:
    """FP32 master parameters + Adam state + BF16 copies, on the GPU."""

    def __init__(self, nbin, seed=0, theta=None):
        self.nbin = nbin
        self.L = layout(nbin)
        P = self.L["P"]
        assert lib().dnsmc_mlp_nparams(nbin) == P
        if theta is None:
            rng = np.random.default_rng(seed)
            th = np.zeros(P, np.float32)
            L, no = self.L, self.L["noutp"]
            th[L["W1"]:L["b1"]] = rng.uniform(-1, 1, N_IN * N_HID) * np.sqrt(6.0 / (N_IN + N_HID))
            th[L["W2"]:L["b2"]] = rng.uniform(-1, 1, N_HID * N_HID) * np.sqrt(6.0 / (2 * N_HID))
            # small last layer: the initial proposal is close to uniform in both heads
            th[L["W3"]:L["b3"]] = rng.uniform(-1, 1, N_HID * no) * 0.01
            theta = th
        th = np.ascontiguousarray(theta, np.float32)
        self.theta = DeviceArray.from_numpy(th)
        self.m = DeviceArray.from_numpy(np.zeros(P, np.float32))
        self.v = DeviceArray.from_numpy(np.zeros(P, np.float32))
        self.grad = DeviceArray.from_numpy(np.zeros(P, np.float32))
        self.step_count = 0
        L = self.L
        self.W1b = DeviceArray.from_numpy(bf16_bits(th[L["W1"]:L["b1"]]))
        self.W2b = DeviceArray.from_numpy(bf16_bits(th[L["W2"]:L["b2"]]))
        self.W3b = DeviceArray.from_numpy(bf16_bits(th[L["W3"]:L["b3"]]))
        self._work = None

    def host_weights(self):
        """(W1, b1, W2, b2, W3, b3) for Solver.set_mlp: BF16 bits + FP32 biases."""
        th = self.theta.numpy()
        L = self.L
        return (self.W1b.numpy(), th[L["b1"]:L["W2"]].copy(), self.W2b.numpy(), th[L["b2"]:L["W3"]].copy(),
                self.W3b.numpy(), th[L["b3"]:L["P"]].copy())

    def forward(self, X):
        """Logits [n, noutp] for features X (DeviceArray float32 [n, 6])."""
        n = X.shape[0]
        Z = DeviceArray((n, self.L["noutp"]), np.float32)
        check(lib().dnsmc_mlp_forward(self.nbin, _ptr(self.theta), _ptr(self.W1b), _ptr(self.W2b), _ptr(self.W3b),
                                      _ptr(X), n, _ptr(Z)))
        return Z

    def loss_grad(self, X, td, tb, wt):
        """Weighted cross-entropy and its gradient (into self.grad)."""
        n = X.shape[0]
        nblk = (n + 63) // 64
        need = nblk * (self.L["P"] + 1) + self.L["P"] + 2
        if self._work is None or self._work.shape[0] < need:
            self._work = DeviceArray(need, np.float32)
        loss = C.c_float(0.0)
        check(lib().dnsmc_mlp_grad(self.nbin, _ptr(self.theta), _ptr(self.W1b), _ptr(self.W2b), _ptr(self.W3b),
                                   _ptr(X), _ptr(td), _ptr(tb), _ptr(wt), n, _ptr(self._work),
                                   self._work.shape[0], _ptr(self.grad), C.byref(loss)))
        return loss.value

    def adam(self, lr):
        self.step_count += 1
        check(lib().dnsmc_mlp_adam(self.nbin, _ptr(self.theta), _ptr(self.m), _ptr(self.v), _ptr(self.grad),
                                   C.c_float(lr), self.step_count, _ptr(self.W1b), _ptr(self.W2b), _ptr(self.W3b)))


class SampleSet:
    """Training samples on the GPU: X [n, 6], direction bin, (band, g) bin, weight."""

    def __init__(self):
        self.parts = []

    def add_run(self, S, cost0=4.0):
        """Samples of the solver's last run (one snapshot, one batch of <= 2^22
        rays, records on).  Weights |x| / sqrt(steps + cost0) (cost0 < 0: |x|)."""
        n = S.nslots * S.rays_per_slot
        X = DeviceArray((n, N_IN), np.float32)
        td = DeviceArray(n, np.int32)
        tb = DeviceArray(n, np.int32)
        wt = DeviceArray(n, np.float32)
        check(lib().dnsmc_mlp_samples(S.h, _ptr(X), _ptr(td), _ptr(tb), _ptr(wt), n, C.c_float(cost0)))
        self.parts.append((X, td, tb, wt))
        return n

    def cat(self):
        return tuple(concat([p[i] for p in self.parts]) for i in range(4))


def train(net, samples, epochs=3, batch=8192, lr=2e-3, seed=0, log=None):
    """Adam over contiguous minibatches in a shuffled order; returns the loss history."""
    X, td, tb, wt = samples
    n = X.shape[0]
    nb = n // batch
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(nb)
        tot = 0.0
        for i, b in enumerate(order):
            sl = slice(int(b) * batch, (int(b) + 1) * batch)
            loss = net.loss_grad(X[sl], td[sl], tb[sl], wt[sl])
            net.adam(lr)
            tot += loss
        hist.append(tot / max(nb, 1))
        if log:
            log("epoch %d loss %.5f" % (ep, hist[-1]))
    return hist


def train_channel_proposal(n_snap=20, rays=8, cost0=16.0, epochs=2, batch=16384, lr=2e-3, log=None, **cfg_kw):
    """The channel's proposal: plain-run tallies of `n_snap` seeded 64x96x64
    snapshots (parabolic profile 573-955 K + N(0, 10 K), seeds 31..) with the
    synthetic 100x16x69 table, `rays` rays per cell, then `epochs` of Adam.
    cfg_kw: further configuration fields of the training runs (the proposal
    should be trained with the traversal it will serve, e.g. grid_num, maxs).
    Returns (weights for Solver.set_mlp, info dict)."""
    import time
    from . import workloads
    from .solver import Solver
    table = workloads.synth_ck_table(11)
    n1 = (64, 96, 64)
    t0 = time.time()
    samples = SampleSet()
    nbin = None
    for s in range(n_snap):
        S = Solver(workloads.channel_config(n1, rays_per_cell=rays, spectral=0, Tw=573.0, **cfg_kw))
        S.set_tables(workloads.table_arrays(table))
        S.set_record(True)
        S.set_field(0, workloads.ghosted(workloads.field_profile_noise(31 + s, n1)))
        S.run()
        S.wait()
        samples.add_run(S, cost0=cost0)
        sc = S.array("scalars")
        nbin = int(sc[3]) * int(sc[4])
        S.close()
    t1 = time.time()
    data = samples.cat()
    net = ProposalNet(nbin, seed=0)
    hist = train(net, data, epochs=epochs, batch=batch, lr=lr, seed=0, log=log)
    t2 = time.time()
    nsamp = int(data[0].shape[0])
    info = dict(train_snapshots=n_snap, rays_per_cell=rays, samples=nsamp, epochs=epochs, cost0=cost0,
                collect_s=round(t1 - t0, 2), train_s=round(t2 - t1, 2),
                train_samples_per_s=nsamp * epochs / max(t2 - t1, 1e-9), loss=[round(h, 5) for h in hist])
    return net.host_weights(), info
