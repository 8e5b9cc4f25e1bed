"""Serialized compressed columns (fls_column_serialize / fls_column_deserialize):
round trips of every encoding, the documented layout, and rejection of
truncated or corrupted images.  Host code only."""
import numpy as np
import pytest

import fastlanesgpu_amd as fls
from fastlanesgpu_amd import selector as sel

RNG = np.random.default_rng(11)


def rle_aux(vbase, n_runs, vbw, ibw):
    """fls_rle_meta {u32 vbase, u16 n_runs, u8 vbw, u8 ibw} as one uint64."""
    return np.uint64(vbase | (n_runs << 32) | (vbw << 48) | (ibw << 56))


def synthetic(encoding, n_rows):
    """A HostColumn with random payload words and valid random per-vector
    metadata of the sizes its encoding carries: every vector's payload lies
    inside the words."""
    nv = (n_rows + 1023) // 1024
    words = RNG.integers(0, 2**32, 80 * nv * 32, dtype=np.uint64).astype(np.uint32)
    bases = RNG.integers(0, 2**32, 32 * nv, dtype=np.uint64).astype(np.uint32)
    vmeta = np.zeros(nv, np.uint64)
    aux = np.zeros(nv, np.uint64)
    for v in range(nv):
        off = 80 * 32 * v
        if encoding == fls.ENC_RLE:
            ibw, n_runs, vbw = int(RNG.integers(0, 2)), int(RNG.integers(1, 1025)), int(RNG.integers(0, 33))
            vmeta[v] = (off << 8) | ibw
            aux[v] = rle_aux(int(RNG.integers(0, 2**32)), n_runs, vbw, ibw)
        else:
            enc, w = int(RNG.integers(0, 4)), int(RNG.integers(0, 33))
            vmeta[v] = (off << 8) | (enc << 6) | w
            aux[v] = RNG.integers(0, 2**32) if enc != sel.MIX_RLE else (
                int(RNG.integers(0, 2**32)) | (int(RNG.integers(1, 1025)) << 32) | (int(RNG.integers(0, 2)) << 48))
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


def corrupt_meta(c, v, vmeta=None, aux=None):
    vm, ax = c.vmeta.copy(), None if c.aux is None else c.aux.copy()
    if vmeta is not None:
        vm[v] = np.uint64(vmeta)
    if aux is not None:
        ax[v] = np.uint64(aux)
    return fls.HostColumn(c.encoding, c.n_rows, c.words, vmeta=vm, bases=c.bases, aux=ax)


def image_with_meta(c, v, vmeta=None, aux=None):
    """The image of c with vector v's metadata replaced after serialization
    (and the checksum recomputed), as a file or the wire could deliver it."""
    img = c.to_bytes().copy()
    q = img[:128].view(np.uint64)
    off_vmeta, off_aux = int(q[10]), int(q[12])
    if vmeta is not None:
        img[off_vmeta + 8 * v:off_vmeta + 8 * v + 8] = np.frombuffer(np.uint64(vmeta).tobytes(), np.uint8)
    if aux is not None:
        img[off_aux + 8 * v:off_aux + 8 * v + 8] = np.frombuffer(np.uint64(aux).tobytes(), np.uint8)
    h = 1469598103934665603
    n_words, n_vmeta, n_bases, n_aux = (int(x) for x in q[5:9])
    offs = [int(x) for x in q[9:13]]
    for o, nbytes in zip(offs, (4 * n_words, 8 * n_vmeta, 4 * n_bases, 8 * n_aux)):
        for b in img[o:o + nbytes].tobytes():
            h = ((h ^ b) * 1099511628211) & (2**64 - 1)
    img[120:128] = np.frombuffer(np.uint64(h).tobytes(), np.uint8)
    return img


def bad_metadata_cases():
    d = fls.delta_column(np.sort(RNG.integers(0, 2**30, 3000)).astype(np.uint32))
    rle = synthetic(fls.ENC_RLE, 3000)
    mix = synthetic(fls.ENC_MIXED, 3000)
    n = lambda c: int(c.words.size)  # noqa: E731
    return [
        ("delta width 33", d, dict(vmeta=(0 << 8) | 33)),
        ("delta payload past the words", d, dict(vmeta=((n(d) - 10) << 8) | 5)),
        ("delta offset past the words", d, dict(vmeta=((n(d) + 1) << 8) | 0)),
        ("rle 1025 runs", rle, dict(aux=rle_aux(0, 1025, 4, 0))),
        ("rle 0 runs", rle, dict(aux=rle_aux(0, 0, 4, 0))),
        ("rle value width 33", rle, dict(aux=rle_aux(0, 10, 33, 0))),
        ("rle run values past the words", rle, dict(vmeta=((n(rle) - 40) << 8) | 1, aux=rle_aux(0, 1024, 32, 1))),
        ("mixed width 40", mix, dict(vmeta=(0 << 8) | (1 << 6) | 40)),
        ("mixed delta past the words", mix, dict(vmeta=((n(mix) - 32) << 8) | (2 << 6) | 1)),
        ("mixed rle 2000 runs", mix, dict(vmeta=(3 << 6) | 8, aux=5 | (2000 << 32))),
        ("mixed rle index width 2", mix, dict(vmeta=(3 << 6) | 8, aux=5 | (10 << 32) | (2 << 48))),
    ]


@pytest.mark.parametrize("case", range(11))
def test_rejects_out_of_range_vector_metadata(case):
    name, c, kw = bad_metadata_cases()[case]
    c.to_bytes()  # the valid column round-trips
    with pytest.raises(fls.FlsError, match="malformed"):  # the image is refused on load
        fls.HostColumn.from_bytes(image_with_meta(c, 1, **kw))
    with pytest.raises(fls.FlsError, match="malformed"):  # and never written
        corrupt_meta(c, 1, **kw).to_bytes()
