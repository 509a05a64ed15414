"""Pins the oracle's hashes (and the library's host build of the same code) against
published vectors and library routines -- CPU only."""
import hashlib

import numpy as np
import pytest

from oracle import oracle as O

M64 = (1 << 64) - 1

# RFC 1321, appendix A.5.
RFC1321 = [
    (b"", "d41d8cd98f00b204e9800998ecf8427e"),
    (b"a", "0cc175b9c0f1b6a831c399e269772661"),
    (b"abc", "900150983cd24fb0d6963f7d28e17f72"),
    (b"message digest", "f96b697d7cb7938d525a2f31aaf161d0"),
    (b"abcdefghijklmnopqrstuvwxyz", "c3fcd3d76192e4007dfb496cca67e13b"),
    (b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789", "d174ab98d277d9f5a5611c2c9f419d9f"),
    (b"1234567890" * 8, "57edf4a22be3c955ac49da2e2107b67a"),
]


@pytest.mark.parametrize("msg,hexd", RFC1321)
def test_md5_rfc1321(msg, hexd):
    assert O.md5(msg).hex() == hexd


def test_md5_matches_hashlib_all_block_boundaries():
    rng = np.random.default_rng(1)
    for n in list(range(0, 200)) + [1000, 4096, 4097]:
        a = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert O.md5(a) == hashlib.md5(a).digest(), n


def test_murmur3_body_pinned_by_smhasher_verification():
    """With a full 64-bit rotate the oracle's body is canonical MurmurHash3_x64_128:
    smhasher's VerificationTest (keys {0},{0,1},... hashed with seed 256-i, then the
    concatenation hashed with seed 0) gives 0x6384BA69 for that function."""
    key = np.zeros(256, np.uint8)
    hashes = bytearray()
    for i in range(256):
        key[i] = i
        hashes += O.murmur3_canonical(key[:i], 256 - i)
    final = O.murmur3_canonical(bytes(hashes), 0)
    assert int.from_bytes(final[:4], "little") == 0x6384BA69


def _rotl_lo32(x, r):
    # include/kokkos_murmur3.hpp:31: a uint64 rotate returned through uint32_t
    return ((x << r) | (x >> (64 - r))) & 0xFFFFFFFF


def _fmix(k):
    k ^= k >> 33
    k = (k * 0xFF51AFD7ED558CCD) & M64
    k ^= k >> 33
    k = (k * 0xC4CEB9FE1A85EC53) & M64
    return k ^ (k >> 33)


def _murmur_mtcheck_py(data: bytes) -> bytes:
    """Independent big-int restatement of MurmurHash3_x64_128 as MTCheck computes it
    (kokkos_murmur3.hpp:214-292): every rotl64 result truncated to 32 bits, seed 0."""
    c1, c2 = 0x87C37B91114253D5, 0x4CF5AD432745937F
    h1 = h2 = 0
    n = len(data)
    for i in range(n // 16):
        k1 = int.from_bytes(data[16 * i:16 * i + 8], "little")
        k2 = int.from_bytes(data[16 * i + 8:16 * i + 16], "little")
        k1 = _rotl_lo32((k1 * c1) & M64, 31) * c2 & M64
        h1 ^= k1
        h1 = (((_rotl_lo32(h1, 27) + h2) & M64) * 5 + 0x52DCE729) & M64
        k2 = _rotl_lo32((k2 * c2) & M64, 33) * c1 & M64
        h2 ^= k2
        h2 = (((_rotl_lo32(h2, 31) + h1) & M64) * 5 + 0x38495AB5) & M64
    tail = data[(n // 16) * 16:]
    if len(tail) > 8:
        k2 = int.from_bytes(tail[8:], "little")
        h2 ^= _rotl_lo32((k2 * c2) & M64, 33) * c1 & M64
    if tail:
        k1 = int.from_bytes(tail[:8], "little")
        h1 ^= _rotl_lo32((k1 * c1) & M64, 31) * c2 & M64
    h1 ^= n
    h2 ^= n
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    h1, h2 = _fmix(h1), _fmix(h2)
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    return h1.to_bytes(8, "little") + h2.to_bytes(8, "little")


def test_murmur3_mtcheck_truncation():
    rng = np.random.default_rng(2)
    for n in list(range(0, 70)) + [127, 128, 129, 512, 4096]:
        a = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert O.murmur3_mtcheck(a) == _murmur_mtcheck_py(a), n


def test_murmur3_mtcheck_closed_forms():
    # len 0: no blocks, no tail, fmix64(0) = 0 -> 16 zero bytes
    assert O.murmur3_mtcheck(b"") == bytes(16)
    # the truncation changes the digest of any non-empty input
    for n in (1, 16, 128):
        a = bytes((37 * i + 1) & 255 for i in range(n))  # a zero block would hide the rotate
        assert O.murmur3_mtcheck(a) != O.murmur3_canonical(a, 0)


def test_library_host_digest_matches_oracle():
    """mtc_digest_host (calculate_digest_host) is the library's own host build of
    its hash code; it must agree with the oracle."""
    import mtcheck_amd as M
    rng = np.random.default_rng(3)
    for n in list(range(0, 140)) + [1000, 4096, 10001]:
        a = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert M.digest_host(a, M.MURMUR3) == O.murmur3_mtcheck(a), n
        assert M.digest_host(a, M.MD5) == O.md5(a), n


def test_digest_string_format():
    import mtcheck_amd as M
    s = M.digest_str(bytes(range(16)))
    assert s == "00010203 04050607 08090A0B 0C0D0E0F "  # include/hash_functions.hpp:52-61
