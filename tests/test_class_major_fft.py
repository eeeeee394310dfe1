"""CPU: the residue-class split of the 256-point x transform used by the fused
Toeplitz passes (hastycucompute_amd/csrc/toeplitz_fused.hip), checked against
numpy.fft.  The kernels store frequency k = 4m + r at position r L/4 + m and
transform each class with one 64-point DFT:

* forward (first pass): X[4m + r] = sum_{n'<64} e^{-2 pi i r n'/256}
  (x[n'] + (-i)^r x[n' + 64]) e^{-2 pi i m n'/64}, for x zero-padded from 128;
* inverse (last pass): x[n] = sum_r e^{+2 pi i r n/256} inner_r(n mod 64) with
  inner_r the 64-point inverse DFT of class r, kept for n in [127, 255): from
  each n' the outputs n' + 128 (pixel n' + 1) and n' + 192 (pixel n' + 65;
  for n' = 63 instead n = 127, pixel 0);
* the z pass reads kernel line 4 (i mod 64) + i / 64 for grid line i.
"""
import numpy as np

L, N, M = 256, 128, 64


def class_major(k):
    return (k % 4) * M + k // 4


def test_forward_fold_matches_fft():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    ref = np.fft.fft(np.concatenate([x, np.zeros(L - N)]))
    got = np.empty(L, complex)
    n1 = np.arange(M)
    for r in range(4):
        u = np.exp(-2j * np.pi * r * n1 / L) * (x[:M] + (-1j) ** r * x[M:])
        got[r * M:(r + 1) * M] = np.fft.fft(u)               # class r, position r*64 + m
    assert np.allclose(got, ref[np.argsort([class_major(k) for k in range(L)])], atol=1e-9)
    # the same statement read the other way: position class_major(k) holds X[k]
    assert np.allclose(got[[class_major(k) for k in range(L)]], ref, atol=1e-9)


def test_inverse_classes_and_crop_match_ifft():
    rng = np.random.default_rng(1)
    X = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    full = np.fft.ifft(X) * L                                 # unnormalised inverse
    ref = full[N - 1:2 * N - 1]                               # the Toeplitz crop
    got = np.zeros(N, complex)
    n1 = np.arange(M)
    for r in range(4):
        inner = np.fft.ifft(X[4 * n1 + r]) * M                # inner_r(n'), unnormalised
        na = n1 + 128
        nb = np.where(n1 == 63, 127, n1 + 192)
        xa = n1 + 1
        xb = np.where(n1 == 63, 0, n1 + 65)
        got[xa] += np.exp(2j * np.pi * r * na / L) * inner
        got[xb] += np.exp(2j * np.pi * r * nb / L) * inner
    assert np.allclose(got, ref, atol=1e-9)
    # every pixel is written exactly once per class
    xs = np.concatenate([n1 + 1, np.where(n1 == 63, 0, n1 + 65)])
    assert sorted(xs.tolist()) == list(range(N))


def test_kernel_line_permutation_inverts_the_store():
    lines = np.arange(L)
    kcol = 4 * (lines % M) + lines // M
    assert all(class_major(int(k)) == i for i, k in zip(lines, kcol))


def test_a_wrong_twist_is_caught():
    """(+i)^r in place of (-i)^r in the fold (a plausible slip) breaks the forward check."""
    rng = np.random.default_rng(2)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    ref = np.fft.fft(np.concatenate([x, np.zeros(L - N)]))
    n1 = np.arange(M)
    r = 1
    u = np.exp(-2j * np.pi * r * n1 / L) * (x[:M] + (1j) ** r * x[M:])
    assert not np.allclose(np.fft.fft(u), ref[4 * n1 + r], atol=1e-6)
