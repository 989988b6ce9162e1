"""The numpy restatement of the CV2_NL_VERIFY pair checksum (helpers), pinned
on the CPU: splitmix64 against its published first outputs from state 0, and
the expansion of a half-list pair set into ordered pairs."""
import numpy as np

from helpers import pair_checksum, splitmix64


def test_splitmix64_known_answers():
    # the generator seeded with 0 (Steele, Lea and Flood 2014; the seeding
    # routine of xoshiro/xoroshiro): its state before each output is k * gamma
    gamma = 0x9E3779B97F4A7C15
    z = np.array([0, gamma, (2 * gamma) % (1 << 64)], np.uint64)
    assert [int(v) for v in splitmix64(z)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                               0x06C45D188009454F]


def test_pair_checksum_expansion():
    one = lambda i, j, img: int(splitmix64(np.array([(i << 38) ^ (j << 6) ^ img], np.uint64))[0])
    # an owned pair counts both ways, with the no-shift image 0b010101
    s, c = pair_checksum([[3, 7, 13]])
    assert c == 2 and s == (one(3, 7, 0x15) + one(7, 3, 0x15)) % (1 << 64)
    # an image pair counts as listed: code 14 is vx = +1 (x_j - L), vy = vz = 0
    s, c = pair_checksum([[3, 7, 14]])
    assert c == 1 and s == one(3, 7, 0x16)
    # code 0 is the shift (-1, -1, -1), code 26 is (+1, +1, +1)
    assert pair_checksum([[1, 2, 0]])[0] == one(1, 2, 0x00)
    assert pair_checksum([[1, 2, 26]])[0] == one(1, 2, 0x2A)
    # ids replace indices; the order of rows does not matter
    ids = np.array([10, 11, 12, 13])
    a = pair_checksum([[0, 1, 13], [2, 3, 5]], ids)
    b = pair_checksum([[2, 3, 5], [0, 1, 13]], ids)
    assert a == b and a[1] == 3
    assert a[0] == (one(10, 11, 0x15) + one(11, 10, 0x15) + one(12, 13, 2 | (1 << 2) | (0 << 4))) % (1 << 64)
    assert pair_checksum(np.zeros((0, 3), np.int64)) == (0, 0)
