"""GF(2) linear algebra for building adversarial test inputs (test infrastructure only).

late_syndrome_cts() makes ciphertexts whose syndrome s_0 .. s_{K-1} is zero and s_K is not:
the syndrome is linear in the ciphertext bits (s_j = XOR of table rows of the set bits), so
such ciphertexts are the null space of the K*m bit rows, less those with s_K = 0.  In the
Berlekamp-Massey loop their first nonzero discrepancy comes at iteration K, when
B = x^(K+1): with K = 63 or 127 that coefficient sits on a 64-lane block boundary.
"""
import numpy as np

import _oracle as orc


def nullspace(A):
    """Basis (rows, uint8 0/1) of {x : A x = 0} over GF(2); A is (r, n) uint8 0/1."""
    r, n = A.shape
    P = np.packbits(A.astype(np.uint8), axis=1)
    piv = []
    row = 0
    for col in range(n):
        if row == r:
            break
        byte, bit = col // 8, 7 - col % 8
        below = (P[row:, byte] >> bit) & 1
        nz = np.flatnonzero(below)
        if nz.size == 0:
            continue
        p = row + int(nz[0])
        if p != row:
            P[[row, p]] = P[[p, row]]
        hit = ((P[:, byte] >> bit) & 1).astype(bool)
        hit[row] = False
        P[hit] ^= P[row]
        piv.append(col)
        row += 1
    R = np.unpackbits(P[:row], axis=1)[:, :n]
    free = [c for c in range(n) if c not in set(piv)]
    basis = np.zeros((len(free), n), np.uint8)
    for i, f in enumerate(free):
        basis[i, f] = 1
        basis[i, piv] = R[:, f]
    return basis


def late_syndrome_cts(pid, sk, K, count, seed=0):
    """`count` ciphertexts (rows of ct_bytes) of key sk with s_j = 0 for j < K, s_K != 0."""
    m, n, t, ct_bytes, sb, _ = orc.PARAMS[pid]
    _, _, table = orc.key_expand(pid, sk)          # sb x 2t uint16
    T = table.reshape(sb, 2 * t).astype(np.uint32)
    # row (j, k): bit k of s_j as a function of the sb ciphertext bits
    A = np.stack([(T[:, j] >> k) & 1 for j in range(K) for k in range(m)]).astype(np.uint8)
    basis = nullspace(A)
    assert basis.shape[0] > 0
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        pick = rng.integers(0, 2, basis.shape[0]).astype(bool)
        if not pick.any():
            continue
        x = np.bitwise_xor.reduce(basis[pick], axis=0)
        ct = np.packbits(x, bitorder="little")[:ct_bytes]
        s = orc.synd(pid, table, ct)
        if np.any(s[:K]) or s[K] == 0:
            continue
        out.append(ct)
    return np.stack(out)
