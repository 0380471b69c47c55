"""CPU pins of the oracle's FP64 contraction (src/mcc.jl:15-17):

    K3[n,m,l] = sum_ijk Psi[i,j,k] phi[i,n] phi[j,m] phi[k,l]

mcc3_fp64 (three mode products) and mcc3_fp64_slab (one output slab from a COO
Psi) are checked against a brute-force einsum of the definition, against the
closed form of a rank-1 tensor, and against each other, with inputs chosen so that
a transposed phi, a swapped mode or a dropped term fails: phi is not symmetric and
not orthogonal, Psi is not symmetric.
"""
import numpy as np
import pytest

from oracle import fc3_oracle as O


def _case(n, seed, symmetric=False):
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal((n, n, n))
    if symmetric:
        psi = sum(psi.transpose(p) for p in ((0, 1, 2), (1, 0, 2), (0, 2, 1), (2, 1, 0),
                                             (1, 2, 0), (2, 0, 1))) / 6
    phi = rng.standard_normal((n, n))
    return np.asfortranarray(psi), np.asfortranarray(phi)


def _coo(psi):
    i, j, k = np.nonzero(psi)
    return i, j, k, psi[i, j, k]


@pytest.mark.parametrize("n", [1, 2, 5, 9, 12])
def test_mcc3_fp64_matches_bruteforce_einsum(n):
    psi, phi = _case(n, 40 + n)
    ref = np.einsum("ijk,in,jm,kl->nml", psi, phi, phi, phi)
    got = O.mcc3_fp64(psi, phi)
    assert got.shape == (n, n, n)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_mcc3_fp64_rank1_closed_form():
    """Psi = a (x) b (x) c gives K3 = (phi^T a) (x) (phi^T b) (x) (phi^T c)."""
    n = 11
    rng = np.random.default_rng(5)
    a, b, c = rng.standard_normal((3, n))
    psi = np.asfortranarray(np.einsum("i,j,k->ijk", a, b, c))
    phi = np.asfortranarray(rng.standard_normal((n, n)))
    ref = np.einsum("n,m,l->nml", phi.T @ a, phi.T @ b, phi.T @ c)
    got = O.mcc3_fp64(psi, phi)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_mcc3_fp64_identity_and_permutation():
    n = 7
    psi, _ = _case(n, 3)
    assert np.array_equal(O.mcc3_fp64(psi, np.eye(n)), psi)
    perm = np.random.default_rng(8).permutation(n)
    P = np.zeros((n, n))
    P[perm, np.arange(n)] = 1.0
    assert np.array_equal(O.mcc3_fp64(psi, P), psi[np.ix_(perm, perm, perm)])


@pytest.mark.parametrize("n", [6, 10])
def test_mcc3_fp64_slab_matches_dense(n):
    psi, phi = _case(n, 70 + n)
    psi[np.abs(psi) < 0.8] = 0.0          # sparse, as a force-constant tensor is
    ref = np.einsum("ijk,in,jm,kl->nml", psi, phi, phi, phi)
    coo = _coo(psi)
    for l in range(n):
        got = O.mcc3_fp64_slab(coo, phi, l)
        assert np.allclose(got, ref[:, :, l], rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_mcc3_fp64_slab_duplicate_entries_add():
    """A COO entry listed twice counts twice (scipy's coo -> csr sums duplicates)."""
    n = 5
    psi, phi = _case(n, 9)
    i, j, k, v = _coo(psi)
    doubled = (np.concatenate([i, i]), np.concatenate([j, j]), np.concatenate([k, k]),
               np.concatenate([v, v]))
    for l in (0, 3):
        assert np.allclose(O.mcc3_fp64_slab(doubled, phi, l), 2 * O.mcc3_fp64_slab(_coo(psi), phi, l),
                           rtol=1e-13, atol=1e-13)


def test_mcc3_fp64_symmetric_input_gives_symmetric_output():
    n = 8
    psi, phi = _case(n, 12, symmetric=True)
    k3 = O.mcc3_fp64(psi, phi)
    for p in ((1, 0, 2), (0, 2, 1), (2, 1, 0)):
        assert np.allclose(k3, k3.transpose(p), rtol=1e-12, atol=1e-12 * np.abs(k3).max())
