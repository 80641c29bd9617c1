"""The oracle's random numbers (rng.hpp:12-17, 70-93): Mufflon seeds PCG64
through SplitMix64's finalizer and steps a 64-bit LCG with the RXS-M-XS
output permutation.

Pins:
* SplitMix64 (Steele, Lea, Flood 2014, the paper rng.hpp cites): the first
  output of the generator started at 0 is mix(0 + golden gamma) =
  0xE220A8397B1DCDAF, the value Java's SplittableRandom(0).nextLong()
  returns (-2152535657050944081); scramble_seed(0) is exactly that output.
* PCG RXS-M-XS 64/64 (O'Neill 2014, pcg_output_rxs_m_xs_64_64 of the PCG
  reference library) on Knuth's MMIX LCG (multiplier 6364136223846793005,
  increment 1442695040888963407), evaluated here with Python integers and
  the LCG state obtained in closed form (the affine map's n-th power by
  repeated squaring) rather than by stepping.
"""
MASK = (1 << 64) - 1
A, C_INC = 6364136223846793005, 1442695040888963407


def _mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def _lcg_jump(state, n):
    """State after n LCG steps: compose the affine map x -> A x + C by squaring."""
    a, c = 1, 0          # accumulated map
    pa, pc = A, C_INC    # map of 2^k steps
    while n:
        if n & 1:
            a, c = (pa * a) & MASK, (pa * c + pc) & MASK
        pa, pc = (pa * pa) & MASK, (pa * pc + pc) & MASK
        n >>= 1
    return (a * state + c) & MASK


def _rxs_m_xs(state):
    word = ((state >> ((state >> 59) + 5)) ^ state) & MASK
    word = (word * 12605985483714917081) & MASK
    return (word >> 43) ^ word


def test_scramble_seed_is_splitmix64(oracle):
    assert oracle.scramble_seed(0) == 0xE220A8397B1DCDAF
    assert 0xE220A8397B1DCDAF - (1 << 64) == -2152535657050944081
    for x in (1, 7, 12345, 0xFFFFFFFF):
        assert oracle.scramble_seed(x) == _mix((x + 0x9E3779B97F4A7C15) & MASK)


def test_pcg64_sequence(oracle):
    for seed in (0, 1, 2, 65537, 0xFFFFFFFF):
        # PCG64(seed): s = scramble(seed); next(); s += seed; next()
        s = _lcg_jump(_mix((seed + 0x9E3779B97F4A7C15) & MASK), 1)
        s = _lcg_jump((s + seed) & MASK, 1)
        got = oracle.pcg_sequence(seed, 64)
        expect = [_rxs_m_xs(_lcg_jump(s, i)) for i in range(64)]
        assert got == expect, seed
