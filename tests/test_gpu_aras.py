"""GPU parity of the ARAS kernels (fmlh_aras_kernels) against the oracle
(oracle/fml_oracle.c, bitwise equal to the reference Fortran build --
tests/test_oracle.py) and the committed reference outputs
(tests/golden/aras.npz).

Tolerance: every expression follows the reference's order; entries differ
by the device exp/sin/cos rounding (feeding l2 = self1 + self2 - 2 pd D,
where it is amplified by the cancellation and by 1/sigma^2) and by the
order of the final sum over atom pairs: 1e-12 relative (measured: <= 7e-16
on the reference molecules, <= 6e-14 on synthetic sets at sigma 0.5-20),
entries below 1e-290 compared as underflow.
"""
import os

import numpy as np
import pytest

import oracle
from fml_amd import synth
from fml_amd.aras import ARAS

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _close(got, ref, rel=1e-12, floor=1e-290):
    got, ref = np.asarray(got), np.asarray(ref)
    assert got.shape == ref.shape
    big = ref > floor
    assert np.all(np.abs(got[~big]) <= 1e-280)
    if big.any():
        err = np.max(np.abs(got[big] - ref[big]) / ref[big])
        assert err <= rel, err


def _set(nm, seed, nmin, nmax, amax, cut=5.0):
    ms = synth.generate(nm, seed=seed, nmin=nmin, nmax=nmax)
    ar = ARAS(maxMolSize=amax, maxAts=amax, cut=cut)
    X = np.array([ar.describe(ms.molecule(i)[1], ms.molecule(i)[0]) for i in range(ms.nmol)])
    Z = [ms.molecule(i)[0] for i in range(ms.nmol)]
    return X, Z, ms.natoms


def test_aras_golden(gpu):
    from fml_amd.kernels import (fget_kernels_aras, get_atomic_kernels_aras,
                                 get_atomic_symmetric_kernels_aras)
    d = np.load(os.path.join(GOLD, "aras.npz"))
    X, n, sig = d["X"], d["n"], d["sigmas"]
    Z = [[0] * int(k) for k in n]
    _close(get_atomic_kernels_aras(X[:5], X[5:], Z[:5], Z[5:], sig), d["kernels"])
    _close(get_atomic_symmetric_kernels_aras(X, Z, sig), d["symmetric"])
    nn = oracle.aras_neighbors(X, n, 4.0)
    _close(fget_kernels_aras(X[:5], X[5:], n[:5], n[5:], nn[:5], nn[5:], sig, t_width=2.0, d_width=0.3,
                             cut_distance=4.0, order=2), d["kernels_o2"])


@pytest.mark.parametrize("kw", [{}, dict(t_width=1.5, d_width=0.25, cut_distance=4.0, order=3, r_width=0.8,
                                         c_width=0.6, scale_angular=0.3)])
def test_aras_synthetic_vs_oracle(gpu, kw):
    from fml_amd.kernels import get_atomic_kernels_aras, get_atomic_symmetric_kernels_aras
    cut = kw.get("cut_distance", 5.0)
    X1, Z1, n1 = _set(30, 50, 1, 15, 15, cut)
    X2, Z2, n2 = _set(20, 51, 3, 15, 15, cut)
    sg = [0.5, 1.0, 4.0, 20.0]
    okw = dict(t_width=kw.get("t_width", np.pi), d_width=kw.get("d_width", 0.2), cut=cut,
               r_width=kw.get("r_width", 1.0), order=kw.get("order", 1), c_width=kw.get("c_width", 0.5),
               scale=kw.get("scale_angular", 0.1))
    _close(get_atomic_kernels_aras(X1, X2, Z1, Z2, sg, **kw), oracle.kernels_aras(X1, X2, n1, n2, sg, **okw))
    _close(get_atomic_symmetric_kernels_aras(X1, Z1, sg, **kw), oracle.kernels_aras(X1, None, n1, None, sg, **okw))


def test_aras_symmetric_equals_global_and_batch_invariance(gpu):
    """The symmetric form equals the global form on the same set for i >= j,
    bitwise; a molecule pair's value is the same in any batch."""
    from fml_amd.kernels import get_atomic_kernels_aras, get_atomic_symmetric_kernels_aras
    X, Z, n = _set(24, 52, 2, 12, 12)
    sg = [0.7, 3.0]
    full = get_atomic_kernels_aras(X, X, Z, Z, sg)
    sym = get_atomic_symmetric_kernels_aras(X, Z, sg)
    low = np.tril(np.ones((len(Z), len(Z)), dtype=bool))
    assert np.array_equal(sym[:, low], full[:, low])
    sub = get_atomic_kernels_aras(X[5:17], X[3:9], Z[5:17], Z[3:9], sg)
    assert np.array_equal(sub, full[:, 5:17, 3:9])


def test_aras_edges(gpu):
    """One-atom molecules (closed form), nine sigmas over two launches,
    device tensors in and out, and a charge outside the element table."""
    import torch
    from fml_amd.kernels import get_atomic_kernels_aras
    ar = ARAS(maxMolSize=2, maxAts=2)
    Zs = [1, 6, 8]
    X = np.array([ar.describe(np.zeros((1, 3)), [z]) for z in Zs])
    Z = [[z] for z in Zs]
    pd = oracle.aras_pd()
    K = get_atomic_kernels_aras(X, X, Z, Z, [1.5])[0]
    for i, zi in enumerate(Zs):
        for j, zj in enumerate(Zs):
            p = pd[zi - 1, zj - 1]
            assert K[i, j] == pytest.approx(np.exp(-(2.0 - 2.0 * p * p) / 2.25), rel=1e-14)
    X1, Z1, n1 = _set(8, 53, 2, 10, 10)
    sg = list(np.linspace(0.5, 8.0, 9))
    _close(get_atomic_kernels_aras(X1, X1, Z1, Z1, sg), oracle.kernels_aras(X1, X1, n1, n1, sg))
    Xd = torch.from_numpy(X1).cuda()
    kd = get_atomic_kernels_aras(Xd, Xd, Z1, Z1, [2.0])
    assert kd.is_cuda and np.array_equal(kd.cpu().numpy(), get_atomic_kernels_aras(X1, X1, Z1, Z1, [2.0]))
    bad = X1.copy()
    bad[0, 0, 1, 0] = 200.0            # no such element
    with pytest.raises(ValueError):
        get_atomic_kernels_aras(bad, X1, Z1, Z1, [1.0])


def test_aras_unsorted_neighbours(gpu):
    """Descriptors whose neighbour distances are not ascending (the
    reference's loops take them as they come): the device switches from the
    window walk to testing every neighbour pair and matches the oracle."""
    from fml_amd.kernels import get_atomic_kernels_aras, get_atomic_symmetric_kernels_aras
    X, Z, n = _set(10, 54, 4, 10, 10)
    rng = np.random.default_rng(7)
    for a in (0, 3):
        for i in range(int(n[a])):
            k = int(np.sum(X[a, i, 0, :] < 5.0))
            perm = np.concatenate([[0], 1 + rng.permutation(k - 1), np.arange(k, X.shape[3])])
            X[a, i, :2, :] = X[a, i, :2, perm].T
            X[a, i, 3:, :] = X[a, i, 3 + perm][:, perm]
    sg = [0.8, 5.0]
    _close(get_atomic_kernels_aras(X, X[2:7], Z, Z[2:7], sg), oracle.kernels_aras(X, X[2:7], n, n[2:7], sg))
    _close(get_atomic_symmetric_kernels_aras(X, Z, sg), oracle.kernels_aras(X, None, n, None, sg))


def test_aras_block_lists(gpu):
    """The C-ABI with the library's own molecule blocks (blk1 NULL), with
    greedy 64-atom blocks in input order and with molecules ordered by
    element set and size gives the same matrix, bitwise: a pair's value
    does not depend on how set-1 molecules are grouped."""
    import ctypes
    from fml_amd import _native
    from fml_amd.device import empty, to_device, workspace
    from fml_amd.kernels.aras_kernels import _pd_cached, aras_blocks
    X, Z, n = _set(21, 55, 2, 14, 14)
    nn = oracle.aras_neighbors(X, n, 5.0)
    dX, dn, dnn = to_device(X), to_device(n, np.int32), to_device(nn, np.int32)
    pd = _pd_cached(1.0, 0.5)
    sg = (ctypes.c_double * 2)(0.9, 3.0)
    lib = _native.load()
    p = _native.ptr
    res = []
    for blocks in (None, aras_blocks(n, 14), aras_blocks(n, 14, X[:, :, 1, 0])):
        out = empty((21, 21, 2), np.float64)
        st = empty((1,), np.int32)
        wsb = lib.fmlh_aras_ws_bytes(21, 21, 14, 1, 0)
        ws = workspace(wsb)
        blk, ordv, nblk = (None, None, 0) if blocks is None else blocks
        _native.call("fmlh_aras_kernels", 0, p(dX), p(dn), p(dnn), 21, p(dX), p(dn), p(dnn), 21, 14, sg, 2,
                     np.pi, 0.2, 5.0, 1, p(pd), int(pd.shape[0]), 0.1, p(blk), p(ordv), nblk, p(out), p(st),
                     p(ws), wsb, _native.stream_ptr())
        assert int(st.item()) == 0
        res.append(out.cpu().numpy())
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])
    _close(np.transpose(res[0], (2, 1, 0)), oracle.kernels_aras(X, X, n, n, [0.9, 3.0]))


@pytest.mark.parametrize("elements", [[1, 6, 7, 8, 9, 16], [1, 6, 7, 8, 9, 15, 16, 17],
                                      [1, 5, 6, 7, 8, 9, 14, 15, 16, 17, 35]])
def test_aras_many_elements(gpu, elements):
    """Data sets of 6, 8 and 11 elements (the 6-, 8- and 16-wide element
    lists, atoms with more than five neighbour elements) against the oracle,
    global and symmetric."""
    from fml_amd.kernels import get_atomic_kernels_aras, get_atomic_symmetric_kernels_aras
    rng = np.random.default_rng(len(elements))
    ms = synth.generate(14, seed=56, nmin=6, nmax=16)
    ar = ARAS(maxMolSize=16, maxAts=16)
    Z = [rng.choice(elements, size=int(k)).tolist() for k in ms.natoms]
    X = np.array([ar.describe(ms.molecule(i)[1], Z[i]) for i in range(ms.nmol)])
    n = ms.natoms
    sg = [0.7, 4.0]
    _close(get_atomic_kernels_aras(X[:9], X[4:], Z[:9], Z[4:], sg),
           oracle.kernels_aras(X[:9], X[4:], n[:9], n[4:], sg))
    _close(get_atomic_symmetric_kernels_aras(X, Z, sg), oracle.kernels_aras(X, None, n, None, sg))
