"""Pure-Python restatement of the two library algorithms the reference's
k-means seed depends on (src/kmeans.cpp:165-167), used to pin the oracle and
the product's seeding on small sizes:

* std::mt19937: the Mersenne Twister MT19937 (Matsumoto & Nishimura 1998)
  with the C++ standard's seeding x[i] = 1812433253 (x[i-1] ^ (x[i-1] >> 30)) + i.
* libstdc++ 11 std::shuffle (bits/stl_algo.h): when (2^32 - 1) / N >= N it
  draws two swap positions per call from one uniform integer in
  [0, r (r + 1) - 1]; otherwise one position per element.  Each uniform draw
  uses Lemire's nearly-divisionless method on the 32-bit engine
  (bits/uniform_int_dist.h, _S_nd with a 64-bit product).
"""


class MT19937:
    N, M = 624, 397

    def __init__(self, seed: int = 5489):
        self.mt = [0] * self.N
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, self.N):
            prev = self.mt[i - 1]
            self.mt[i] = (1812433253 * (prev ^ (prev >> 30)) + i) & 0xFFFFFFFF
        self.idx = self.N

    def _twist(self):
        mt = self.mt
        for i in range(self.N):
            y = (mt[i] & 0x80000000) | (mt[(i + 1) % self.N] & 0x7FFFFFFF)
            v = mt[(i + self.M) % self.N] ^ (y >> 1)
            if y & 1:
                v ^= 0x9908B0DF
            mt[i] = v
        self.idx = 0

    def __call__(self) -> int:
        if self.idx >= self.N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def uniform_int(g: MT19937, a: int, b: int) -> int:
    """uniform_int_distribution<unsigned long>{a, b}(g) for the 32-bit engine."""
    urange = b - a
    rng = urange + 1  # < 2^32 here
    product = g() * rng
    low = product & 0xFFFFFFFF
    if low < rng:
        threshold = ((1 << 32) - rng) % rng
        while low < threshold:
            product = g() * rng
            low = product & 0xFFFFFFFF
    return (product >> 32) + a


def shuffle(seq: list, g: MT19937) -> None:
    n = len(seq)
    if n == 0:
        return
    urngrange = 0xFFFFFFFF
    if urngrange // n >= n:
        i = 1
        if n % 2 == 0:
            j = uniform_int(g, 0, 1)
            seq[i], seq[j] = seq[j], seq[i]
            i += 1
        while i != n:
            r = i + 1  # swap range
            x = uniform_int(g, 0, r * (r + 1) - 1)
            p1, p2 = x // (r + 1), x % (r + 1)
            seq[i], seq[p1] = seq[p1], seq[i]
            i += 1
            seq[i], seq[p2] = seq[p2], seq[i]
            i += 1
        return
    for i in range(1, n):
        j = uniform_int(g, 0, i)
        seq[i], seq[j] = seq[j], seq[i]


def initial_indices(n: int, k: int, seed: int) -> list:
    """KMeansSolver::guess_initial_centroids' index choice (kmeans.cpp:161-169)."""
    perm = list(range(n))
    shuffle(perm, MT19937(seed))
    return sorted(perm[:k])
