"""Serialized compressed columns (fls_column_serialize / fls_column_deserialize):
round trips of every encoding, the documented layout, and rejection of
truncated or corrupted images.  Host code only."""
import numpy as np
import pytest

import fastlanesgpu_amd as fls

RNG = np.random.default_rng(11)


def synthetic(encoding, n_rows):
    """A HostColumn with random arrays of the sizes its encoding carries."""
    nv = (n_rows + 1023) // 1024
    words = RNG.integers(0, 2**32, 3 * nv * 32, dtype=np.uint64).astype(np.uint32)
    vmeta = RNG.integers(0, 2**40, nv, dtype=np.uint64)
    bases = RNG.integers(0, 2**32, 32 * nv, dtype=np.uint64).astype(np.uint32)
    aux = RNG.integers(0, 2**63, nv, dtype=np.uint64)
    if encoding == fls.ENC_RLE:
        return fls.HostColumn(encoding, n_rows, words, vmeta=vmeta, bases=bases, aux=aux)
    return fls.HostColumn(encoding, n_rows, words, vmeta=vmeta, aux=aux)


def columns():
    x = RNG.integers(0, 2**11, 5000).astype(np.uint32)
    s = np.sort(RNG.integers(0, 2**30, 70_000)).astype(np.uint32)
    return {
        "bitpack": fls.bitpack_column(x, 11),
        "bitpack_for_transposed": fls.bitpack_column(x + 7_000_000, transposed=True, base=7_000_000),
        "bitpack_bw0": fls.bitpack_column(np.zeros(2048, np.uint32), 0),
        "delta": fls.delta_column(s),
        "rle": synthetic(fls.ENC_RLE, 9000),
        "mixed": synthetic(fls.ENC_MIXED, 3 * 1024 + 5),
    }


def same(a, b):
    assert (a.encoding, a.n_rows, a.bw, a.base, bool(a.transposed)) == (b.encoding, b.n_rows, b.bw, b.base,
                                                                        bool(b.transposed))
    for name in ("words", "vmeta", "bases", "aux"):
        u, v = getattr(a, name), getattr(b, name)
        if u is None or (u.size == 0 and name != "words"):
            assert v is None or v.size == 0, name
        else:
            np.testing.assert_array_equal(v, u, err_msg=name)


@pytest.mark.parametrize("name", ["bitpack", "bitpack_for_transposed", "bitpack_bw0", "delta", "rle", "mixed"])
def test_round_trip(name, tmp_path):
    c = columns()[name]
    img = c.to_bytes()
    same(c, fls.HostColumn.from_bytes(img))
    f = tmp_path / "col.fls"
    c.save(f)
    same(c, fls.HostColumn.load(f))


def test_layout():
    c = columns()["delta"]
    img = c.to_bytes()
    assert bytes(img[:8]) == b"FLSCOL01"
    h = img[:128].view(np.uint32)
    assert h[2] == 1 and h[3] == fls.ENC_DELTA  # version, encoding
    q = img[:128].view(np.uint64)
    n_rows, n_words, n_vmeta, n_bases, n_aux = q[4:9]
    offs = q[9:13]
    total = q[13]
    nv = (c.n_rows + 1023) // 1024
    assert (n_rows, n_words, n_vmeta, n_bases, n_aux) == (c.n_rows, c.words.size, nv, 32 * nv, 0)
    assert all(int(o) % 64 == 0 for o in offs) and int(offs[0]) == 128
    assert int(total) == img.size
    np.testing.assert_array_equal(img[int(offs[0]):int(offs[0]) + 4 * c.words.size].view(np.uint32), c.words)


def test_rejects_corruption():
    c = columns()["rle"]
    img = c.to_bytes()
    with pytest.raises(fls.FlsError):  # truncated
        fls.HostColumn.from_bytes(img[:-64])
    with pytest.raises(fls.FlsError):  # header only
        fls.HostColumn.from_bytes(img[:100])
    bad = img.copy()
    bad[0] ^= 1  # magic
    with pytest.raises(fls.FlsError):
        fls.HostColumn.from_bytes(bad)
    bad = img.copy()
    bad[200] ^= 0x40  # a payload byte: checksum
    with pytest.raises(fls.FlsError):
        fls.HostColumn.from_bytes(bad)
    bad = img.copy()
    bad[48:56] = np.frombuffer(np.uint64(12345).tobytes(), np.uint8)  # n_vmeta (header word q[6])
    with pytest.raises(fls.FlsError):
        fls.HostColumn.from_bytes(bad)


def test_serialize_rejects_short_fixed_width_column():
    c = fls.bitpack_column(RNG.integers(0, 2**9, 3000).astype(np.uint32), 9)
    short = fls.HostColumn(fls.ENC_BITPACK, c.n_rows, c.words[:-1], bw=9)
    with pytest.raises(fls.FlsError):
        short.to_bytes()
