"""Device-side seeded input generation matches the oracle's ChaCha20 stream."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,start,n", [(0, 0, 1000), (42, 12345, 777), (2**64 - 1, 0, 4096)])
def test_random_fr_matches_oracle(lib, co, seed, start, n):
    from halo2_amd import _lib, device as dv

    d = dv.empty_u64(4 * n)
    _lib.check(lib.h2amd_random_fr(dv.ptr(d), n, seed, start, dv.stream()))
    assert (dv.to_host(d) == co.chacha_fr(seed, start, n)).all()
