"""The device-side generators against the host ones (CPU torch; no GPU needed)."""
import numpy as np
import torch

from zquatev_amd import gen


def test_uniform_t_is_bitwise_the_host_stream():
    idx = np.array([0, 1, 2, 1000, 2 ** 40 + 3], dtype=np.int64)
    a = gen.uniform_t(1001, torch.from_numpy(idx)).numpy()
    assert np.array_equal(a, gen.uniform(1001, idx))


def test_fock_case_device_matches_host_lower_triangles():
    n = 45
    D = gen.fock_case(n, 1007)
    out = torch.zeros(2 * n, 2 * n, dtype=torch.complex128)
    gen.fock_case_device(n, 1007, out, chunk=16)
    Dn = out.numpy().T
    low = np.tril(np.ones((n, n), dtype=bool))
    assert np.abs(Dn[:n, :n][low] - D[:n, :n][low]).max() <= 1e-14 * np.abs(D).max()
    assert np.abs(Dn[n:, :n][low] - D[n:, :n][low]).max() <= 1e-14 * np.abs(D).max()
    H = gen.full_hermitian_t(out, n).numpy()
    assert np.abs(H - gen.full_hermitian(D, n)).max() <= 1e-14 * np.abs(D).max()
    assert np.array_equal(H, H.conj().T)


def test_workload_size_draws():
    b = gen.batched_sizes(4096, 1002)
    assert b.min() >= 64 and b.max() <= 1024 and np.all(b % 2 == 0)
    assert abs(b.mean() - 544) < 15                 # uniform over even [64, 1024]
    s = gen.serving_sizes(4096, 1003)
    assert s.min() >= 128 and s.max() <= 4096 and np.all(s % 2 == 0)
    # log-uniform: the median sits near the geometric mean of the range
    assert abs(np.log(np.median(s)) - 0.5 * (np.log(128) + np.log(4096))) < 0.1


def test_uniform_case_device_is_bitwise_the_host_matrix():
    n = 37
    D = gen.uniform_case(n, 1004, diag_shift=10.0)
    out = torch.zeros(2 * n, 2 * n, dtype=torch.complex128)
    gen.uniform_case_device(n, 1004, out, diag_shift=10.0, chunk=8)
    Dn = out.numpy().T
    low = np.tril(np.ones((n, n), dtype=bool))
    assert np.array_equal(Dn[:n, :n][low], D[:n, :n][low])
    assert np.array_equal(Dn[n:, :n][low], D[n:, :n][low])
