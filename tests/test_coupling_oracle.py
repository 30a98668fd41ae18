"""The coupling oracle (fix_temp / fix_rad, code/fix_var.f:49-121) against
what fixes it: a real MPI_SUM-style all-reduce of the ranks' buffers over two
gloo processes, and the per-cell closed forms."""
import os
import socket

import numpy as np
import pytest

from oracle import coupling

N = (6, 8, 4)
TC, TPLUS = 573.0, 0.37


def _slabs(n, pj, pk, seed=3):
    """cnew arrays of a pj x pk x-pencil decomposition (all i on every rank)."""
    rng = np.random.default_rng(seed)
    glob = rng.uniform(0.0, 0.8, size=(n[0] + 2, n[1], n[2]))   # cnew(0:n0+1, global j, global k)
    out = []
    js = np.array_split(np.arange(1, n[1] + 1), pj)
    ks = np.array_split(np.arange(1, n[2] + 1), pk)
    for jr in js:
        for kr in ks:
            s = (1, int(jr[0]), int(kr[0]))
            e = (n[0], int(jr[-1]), int(kr[-1]))
            c = np.asfortranarray(glob[:, s[1] - 1:e[1], s[2] - 1:e[2]])
            out.append((c, s, e))
    return glob, out


def test_fix_temp_closed_form():
    glob, slabs = _slabs(N, 2, 2)
    T = coupling.fix_temp(slabs, N, TC, TPLUS)
    assert T.dtype == np.float32 and T.shape == (N[0] + 2, N[1] + 2, N[2] + 2)
    want = (glob[1:N[0] + 1] * TC / TPLUS + TC).astype(np.float32)
    np.testing.assert_array_equal(T[1:-1, 1:-1, 1:-1], want)
    # ghost layers stay zero (no rank writes them)
    assert not T[0].any() and not T[-1].any() and not T[:, 0].any() and not T[:, :, -1].any()


def test_fix_rad_closed_form():
    n0, nj, n2 = 5, 3, 6
    rng = np.random.default_rng(1)
    qr = np.asfortranarray(rng.normal(size=(n0 + 2, nj + 2, n2 + 2)).astype(np.float32))
    d = coupling.fix_rad(qr, (1, 1, 3), (n0, nj, 5), TC, n0, nj)
    assert d.shape == (n0 + 2, nj, 3)
    want = qr[1:n0 + 1, 1:nj + 1, 3:6].astype(np.float64) / 5.670373e-08 / TC ** 4
    np.testing.assert_allclose(d[1:n0 + 1], want, rtol=1e-15)
    assert not d[0].any() and not d[-1].any()


def _free_port():
    s = socket.socket()
    s.bind(("127.0.0.1", 0))
    p = s.getsockname()[1]
    s.close()
    return p


def _rank(rank, world, port, path):
    import torch
    import torch.distributed as dist
    os.environ["MASTER_ADDR"] = "127.0.0.1"
    os.environ["MASTER_PORT"] = str(port)
    dist.init_process_group("gloo", rank=rank, world_size=world)
    _, slabs = _slabs(N, 2, 1)
    c, s, e = slabs[rank]
    buf = torch.from_numpy(np.ascontiguousarray(coupling.rank_buffer(c, s, e, N, TC, TPLUS).ravel(order="F")))
    dist.all_reduce(buf, op=dist.ReduceOp.SUM)
    if rank == 0:
        np.save(path, buf.numpy())
    dist.destroy_process_group()


def test_fix_temp_matches_a_two_rank_allreduce(tmp_path):
    """Two gloo ranks, each with its slab, all-reduce their Tbuf as the
    reference does (code/fix_var.f:71): the oracle's field is the result."""
    mp = pytest.importorskip("torch.multiprocessing")
    path = str(tmp_path / "tmc.npy")
    mp.spawn(_rank, args=(2, _free_port(), path), nprocs=2, join=True)
    got = np.load(path).reshape((N[0] + 2, N[1] + 2, N[2] + 2), order="F").astype(np.float32)
    _, slabs = _slabs(N, 2, 1)
    np.testing.assert_array_equal(got, coupling.fix_temp(slabs, N, TC, TPLUS))
