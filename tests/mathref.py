"""Independent textbook mathematics used to pin the oracle.

Nothing here follows the reference's code: these are the closed forms of the
negacyclic transform (evaluation of a(X) at the odd powers of a primitive
2N-th root psi, listed in bit-reversed order), its inverse (interpolation),
and schoolbook multiplication modulo X^N + 1, written with Python integers.
"""


def brv(i: int, width: int) -> int:
    return int(format(i, f"0{width}b")[::-1], 2) if width else 0


def negacyclic_eval_bitrev(a, q, psi):
    """out[k] = sum_j a[j] * psi^((2*brv(k)+1)*j) mod q."""
    N = len(a)
    logN = N.bit_length() - 1
    out = []
    for k in range(N):
        r = pow(psi, 2 * brv(k, logN) + 1, q)
        acc, p = 0, 1
        for j in range(N):
            acc += a[j] * p
            p = p * r % q
        out.append(acc % q)
    return out


def negacyclic_interp_bitrev(A, q, psi):
    """a[j] = N^-1 * sum_k A[k] * psi^(-(2*brv(k)+1)*j) mod q."""
    N = len(A)
    logN = N.bit_length() - 1
    ninv = pow(N, -1, q)
    psi_inv = pow(psi, -1, q)
    out = []
    for j in range(N):
        acc = 0
        for k in range(N):
            acc += A[k] * pow(psi_inv, (2 * brv(k, logN) + 1) * j, q)
        out.append(acc * ninv % q)
    return out


def negacyclic_mul(a, b, q):
    N = len(a)
    c = [0] * N
    for i in range(N):
        for j in range(N):
            k = i + j
            if k < N:
                c[k] += a[i] * b[j]
            else:
                c[k - N] -= a[i] * b[j]
    return [x % q for x in c]


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    small = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)
    for p in small:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def prime_factors(n: int):
    f, p = [], 2
    while p * p <= n:
        if n % p == 0:
            f.append(p)
            while n % p == 0:
                n //= p
        p += 1 if p == 2 else 2
    if n > 1:
        f.append(n)
    return f


def smallest_primitive_root(q: int) -> int:
    fs = prime_factors(q - 1)
    g = 2
    while any(pow(g, (q - 1) // p, q) == 1 for p in fs):
        g += 1
    return g


def largest_ntt_primes(count: int, two_n: int, bound: int = 1 << 30):
    """The `count` largest primes q < bound with q = 1 mod two_n, with roots."""
    out = []
    q = ((bound - 2) // two_n) * two_n + 1
    while len(out) < count and q > 2:
        if q < bound and is_prime(q):
            out.append((q, smallest_primitive_root(q)))
        q -= two_n
    return out
