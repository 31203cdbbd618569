"""SRS files on the device (src/poly/kzg/commitment.rs:317-414): write/read round trips
in the three SerdeFormats, the raw section equal to the device base layout, compressed
bytes equal to a host encoding, RawBytes rejecting corrupted points (and the unchecked
form reading them), Processed rejecting bad x, and a read SRS committing as the original."""
import io

import numpy as np
import pytest

from conftest import rand_fr

pytestmark = pytest.mark.gpu


def _params(k=10):
    from halo2_amd import fields as F, msm

    return msm.ParamsKZG(k, 0x5EED5 % F.R_MOD, fixed_base=False)


def _fq_int(words) -> int:
    from halo2_amd.fields import P_MOD

    v = int.from_bytes(np.ascontiguousarray(words, dtype=np.uint64).tobytes(), "little")
    return v * pow(1 << 256, -1, P_MOD) % P_MOD


def test_srs_round_trip_all_formats(lib, co):
    from halo2_amd import device as dv, srs

    params = _params()
    n = params.n
    g = dv.to_host(params.g_device, cols=8)
    gl = dv.to_host(params.g_lagrange_device, cols=8)
    for fmt in srs.SerdeFormat:
        buf = io.BytesIO()
        srs.write_custom(params, buf, fmt)
        data = buf.getvalue()
        g1, g2 = (32, 64) if fmt == srs.SerdeFormat.Processed else (64, 128)
        assert len(data) == 4 + 2 * n * g1 + 2 * g2, fmt
        assert data[:4] == (10).to_bytes(4, "little")
        if fmt != srs.SerdeFormat.Processed:  # the raw section is the device layout itself
            assert data[4: 4 + 64 * n] == g.tobytes()
        back = srs.read_custom(io.BytesIO(data), fmt, fixed_base=False)
        assert back.k == 10 and (dv.to_host(back.g_device, cols=8) == g).all()
        assert (dv.to_host(back.g_lagrange_device, cols=8) == gl).all()
        assert back.g2 == params.g2 and back.s_g2 == params.s_g2
        again = io.BytesIO()
        srs.write_custom(back, again, fmt)
        assert again.getvalue() == data
    # write/read use RawBytesUnchecked
    buf = io.BytesIO()
    srs.write(params, buf)
    assert len(buf.getvalue()) == 4 + 128 * n + 256
    assert (dv.to_host(srs.read(io.BytesIO(buf.getvalue())).g_device, cols=8) == g).all()


def test_srs_compressed_bytes_and_commit(lib, co):
    from halo2_amd import device as dv, fields as F, srs

    params = _params()
    g = dv.to_host(params.g_device, cols=8)
    comp = srs.g1_compress_device(params.g_device).cpu().numpy().reshape(-1, 32)
    for i in (0, 1, 2, 500, params.n - 1):
        x, y = _fq_int(g[i, :4]), _fq_int(g[i, 4:])
        exp = bytearray(x.to_bytes(32, "little"))
        exp[31] |= (y & 1) << 7
        assert bytes(comp[i]) == bytes(exp), i
    # the identity compresses to zeros and back
    z = dv.to_device(np.zeros((3, 8), dtype=np.uint64))
    zc = srs.g1_compress_device(z)
    assert not zc.cpu().numpy().any()
    pts, bad, _ = srs.g1_decompress_device(zc)
    assert bad == 0 and not dv.to_host(pts).any()
    # a read SRS commits like the original: f(tau) G
    buf = io.BytesIO()
    srs.write_custom(params, buf, srs.SerdeFormat.Processed)
    back = srs.read_custom(io.BytesIO(buf.getvalue()), srs.SerdeFormat.Processed, fixed_base=False)
    poly = rand_fr(co, params.n, 5)
    assert (back.commit_device(dv.to_device(poly)) == params.commit_device(dv.to_device(poly))).all()
    assert (back.commit_lagrange(poly) == params.commit_lagrange(poly)).all()
    del F


def test_srs_rejects_bad_points(lib, co):
    from halo2_amd import device as dv, srs
    from halo2_amd.fields import P_MOD

    params = _params()
    n = params.n
    buf = io.BytesIO()
    srs.write_custom(params, buf, srs.SerdeFormat.RawBytes)
    data = bytearray(buf.getvalue())
    # y of point 7 of g plus one (Montgomery word + 1: off the curve), x of lagrange point 3 >= p
    off = 4 + 64 * 7 + 32
    data[off] ^= 1
    off2 = 4 + 64 * n + 64 * 3
    data[off2: off2 + 32] = (P_MOD + 1).to_bytes(32, "little")
    with pytest.raises(ValueError):
        srs.read_custom(io.BytesIO(bytes(data)), srs.SerdeFormat.RawBytes)
    loose = srs.read_custom(io.BytesIO(bytes(data)), srs.SerdeFormat.RawBytesUnchecked, fixed_base=False)
    got = dv.to_host(loose.g_device, cols=8)
    assert got[7, 4] == np.frombuffer(bytes(data[4 + 64 * 7 + 32: 4 + 64 * 7 + 40]), dtype=np.uint64)[0]
    # device validation counts and locates the bad points
    bad, first = srs.g1_validate_device(loose.g_device)
    assert (bad, first) == (1, 7)
    bad, first = srs.g1_validate_device(loose.g_lagrange_device)
    assert (bad, first) == (1, 3)
    assert srs.g1_validate_device(params.g_device) == (0, 2 ** 64 - 1)
    # Processed: x >= p, and an x with no point on the curve
    comp = srs.g1_compress_device(params.g_device).cpu().numpy().copy().reshape(-1, 32)
    comp[2] = np.frombuffer((P_MOD + 2).to_bytes(32, "little"), dtype=np.uint8)
    x = 1
    while pow((x ** 3 + 3) % P_MOD, (P_MOD - 1) // 2, P_MOD) == 1:
        x += 1
    comp[9] = np.frombuffer(x.to_bytes(32, "little"), dtype=np.uint8)
    import torch

    _, bad, first = srs.g1_decompress_device(torch.from_numpy(comp.reshape(-1)).to(params.g_device.device))
    assert (bad, first) == (2, 2)
