"""The INT8 slice arithmetic of the mixed-precision Z product (Ozaki scheme),
restated in numpy: the five-slice split of split8_kernel and the exact reconstruction of
z8_value.  CPU only; pins the scheme's error bound, not the kernel."""
import numpy as np


def exponents(m):
    """e with m 2^-e <= 127 (0 for m = 0), per row maximum m."""
    return np.where(m > 0, np.frexp(m / 127.0)[1], 0)


def split_rows(x, nslices=5, e=None):
    """x (rows, K) -> (exponents (rows,), slices (nslices, rows, K) int8):
    x = 2^e sum_s x_s 128^-s + r, |x_0| <= 127, |x_s| <= 64 (s > 0); e from
    the row maxima of x unless given (it must satisfy max|x| 2^-e <= 127)."""
    if e is None:
        e = exponents(np.abs(x).max(axis=1))
    t = np.ldexp(x, -e[:, None])
    out = []
    for _ in range(nslices):
        w = np.rint(t)
        out.append(w.astype(np.int8))
        t = (t - w) * 128.0
    return e, np.array(out)


def product(A, B, nslices=5):
    """sum_k A[i,k] B[j,k] from the slices of A's rows and B's rows, keeping
    the slice pairs with a + b <= nslices - 1 (exact int products, reconstructed
    as in z8_value: (128 R0 + R1) + 128^-2 (128 R2 + R3 + R4 / 128), times 2^-7)."""
    ea, sa = split_rows(A, nslices)
    eb, sb = split_rows(B, nslices)
    R = [sum(sa[a].astype(np.int64) @ sb[d - a].astype(np.int64).T for a in range(d + 1)) for d in range(nslices)]
    assert all(np.abs(r).max() < 2 ** 31 for r in R)
    r01, r23 = 128 * R[0] + R[1], 128 * R[2] + R[3]
    assert np.abs(r01).max() < 2 ** 31 and np.abs(r23).max() < 2 ** 31
    t = r23.astype(np.float64) + R[4].astype(np.float64) / 128.0
    v = (r01.astype(np.float64) + t / 16384.0) / 128.0
    return np.ldexp(v, ea[:, None] + eb[None, :])


def test_slices_reconstruct_the_operand():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((64, 96)) * np.exp(-np.abs(rng.standard_normal((64, 96))) * 8)
    e, s = split_rows(x)
    assert np.abs(s[0]).max() <= 127 and np.abs(s[1:]).max() <= 64
    back = sum(s[i].astype(np.float64) * 128.0 ** -i for i in range(5))
    err = np.abs(np.ldexp(back, e[:, None]) - x) / np.abs(x).max(axis=1, keepdims=True)
    assert err.max() <= 0.5 * 128.0 ** -4 / 64  # half a unit of the last slice, relative to the row max


def test_product_error_bound_and_worst_case_k():
    rng = np.random.default_rng(2)
    for K in (17, 95, 256):
        A = rng.standard_normal((40, K)) * 10.0 ** rng.uniform(-3, 3, (40, 1))
        B = rng.standard_normal((24, K))
        exact = A @ B.T
        got = product(A, B)
        bound = np.abs(A).max(axis=1)[:, None] * np.abs(B).max(axis=1)[None, :] * K
        assert (np.abs(got - exact) / bound).max() < 2.0 ** -34, K
    # the largest K the int32 accumulators take: every slice at its maximum
    A = np.full((2, 256), 127.0)
    got = product(A, A)
    assert np.allclose(got, A @ A.T, rtol=1e-12)


def test_three_real_products_give_a_conj_b():
    """The 3M combination of z8_jobs: A conj(B) = (P1 + P2) + i (P3 - P1 + P2)
    with P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br - Bi)."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((16, 50)) + 1j * rng.standard_normal((16, 50))
    B = rng.standard_normal((16, 50)) + 1j * rng.standard_normal((16, 50))
    p1 = product(A.real, B.real)
    p2 = product(A.imag, B.imag)
    p3 = product(A.real + A.imag, B.real - B.imag)
    got = (p1 + p2) + 1j * (p3 - p1 + p2)
    bound = 4 * np.abs(A).max() * np.abs(B).max() * A.shape[1]   # (|re| + |im|) sums, normwise
    assert np.abs(got - A @ B.conj().T).max() < 2.0 ** -34 * bound


def diagonal_sums(A, B, ea, eb, nslices=5):
    """Integer sums R_d = sum_{a+b=d} sum_k A_a B_b of two real operands
    split with the given row exponents (the split8_kernel / z8_jobs layout)."""
    _, sa = split_rows(A, nslices, ea)
    _, sb = split_rows(B, nslices, eb)
    return [sum(sa[a].astype(np.int64) @ sb[d - a].astype(np.int64).T for a in range(d + 1)) for d in range(nslices)]


def value(R, ex):
    """z8_value: (128 R0 + R1 + (R2 + R3/128 + R4/128^2) / 128) / 128 * 2^ex, the
    head formed in int32 (asserted) and the tail in float32."""
    head = 128 * R[0] + R[1]
    assert np.abs(head).max() < 2 ** 31
    tail = (R[4].astype(np.float32) / np.float32(128.0) + R[3].astype(np.float32)) / np.float32(128.0) \
        + R[2].astype(np.float32)
    return np.ldexp((head.astype(np.float64) + tail.astype(np.float64) / 128.0) / 128.0, ex)


def test_shared_exponent_combines_the_three_products_in_int32():
    """The Z product's exact integer combination: one exponent per row of A
    (column of B) for the components re, im and re + im (re - im); X = P1 + P2
    and Y = P3 - P1 + P2 summed in int32 before one reconstruction each give
    Z = conj(A conj(B)) = X - i Y to the slices' normwise bound, also at
    the largest K (256) with every slice at its maximum."""
    rng = np.random.default_rng(4)
    for K, big in ((50, False), (256, False), (256, True)):
        if big:
            A = np.full((16, K), 127.0) * (1 + 1j)
            B = np.full((16, K), 127.0) * (1 - 1j)
        else:
            A = (rng.standard_normal((16, K)) + 1j * rng.standard_normal((16, K))) * 10.0 ** rng.uniform(-2, 2, (16, 1))
            B = rng.standard_normal((24, K)) + 1j * rng.standard_normal((24, K))
        ca = [A.real, A.imag, A.real + A.imag]
        cb = [B.real, B.imag, B.real - B.imag]
        ea = exponents(np.max([np.abs(c).max(axis=1) for c in ca], axis=0))
        eb = exponents(np.max([np.abs(c).max(axis=1) for c in cb], axis=0))
        R = [diagonal_sums(ca[c], cb[c], ea, eb) for c in range(3)]
        X = [R[0][d] + R[1][d] for d in range(5)]
        Y = [R[2][d] - R[0][d] + R[1][d] for d in range(5)]
        ex = ea[:, None] + eb[None, :]
        Z = value(X, ex) - 1j * value(Y, ex)
        exact = np.conj(A @ B.conj().T)
        bound = 4 * np.abs(A).max() * np.abs(B).max() * K
        assert np.abs(Z - exact).max() < 2.0 ** -33 * bound, K
