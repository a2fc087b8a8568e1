"""Wire formats (via-protocol wire/format.rs) through the C ABI, on the CPU.

The cases restate the reference's own tests (format.rs:198-278): Uncompressed
round trips (empty; 0, 1, q1 limb 0, u64::MAX), a corrupted tag, an empty
buffer, PrgCompressed body decoding and its tag check, plus the byte layout
([tag][count u64 LE][coefficients u64 LE]) checked against struct packing and
truncated packets at every length.
"""
import struct

import numpy as np
import pytest


@pytest.fixture(scope="module")
def vg():
    import via_gpu

    return via_gpu


def test_uncompressed_round_trip_empty(vg):
    b = vg.wire_encode([])
    assert b == bytes([0]) + struct.pack("<Q", 0)
    assert vg.wire_decode(b).size == 0


def test_uncompressed_round_trip_small(vg):
    orig = np.array([0, 1, 137_438_822_401, 2**64 - 1], dtype=np.uint64)
    b = vg.wire_encode(orig)
    assert b[0] == 0x00
    assert b == bytes([0]) + struct.pack("<Q", 4) + struct.pack("<4Q", *[int(x) for x in orig])
    assert np.array_equal(vg.wire_decode(b), orig)


def test_uncompressed_wrong_tag(vg):
    b = bytearray(vg.wire_encode([1, 2]))
    b[0] = 0xFF
    with pytest.raises(vg.ViaError) as e:
        vg.wire_decode(bytes(b))
    assert e.value.status == vg.VIA_E_FORMAT_TAG and e.value.got == 0xFF


def test_uncompressed_buffer_too_short(vg):
    with pytest.raises(vg.ViaError) as e:
        vg.wire_decode(b"")
    assert e.value.status == vg.VIA_E_BUFFER_SHORT and (e.value.expected, e.value.got) == (1, 0)
    full = vg.wire_encode(np.arange(5, dtype=np.uint64))
    for cut in range(1, len(full)):
        with pytest.raises(vg.ViaError) as e:
            vg.wire_decode(full[:cut])
        want = 9 if cut < 9 else len(full)
        assert e.value.status == vg.VIA_E_BUFFER_SHORT and (e.value.expected, e.value.got) == (want, cut)
    # a count that cannot fit reports the bytes it would need
    huge = bytes([0]) + struct.pack("<Q", 2**40) + bytes(16)
    with pytest.raises(vg.ViaError) as e:
        vg.wire_decode(huge)
    assert (e.value.expected, e.value.got) == (9 + 8 * 2**40, len(huge))


def test_prg_compressed_deserialize_body_only(vg):
    seed = bytes(range(32))
    body = [42, 99]
    buf = bytes([0x01]) + seed + struct.pack("<Q", len(body)) + struct.pack("<2Q", *body)
    s, b = vg.wire_decode_prg(buf)
    assert s == seed and b.tolist() == body


def test_prg_compressed_wrong_tag_and_short(vg):
    buf = bytes([0x00]) + bytes([0xFF] * 49)
    with pytest.raises(vg.ViaError) as e:
        vg.wire_decode_prg(buf)
    assert e.value.status == vg.VIA_E_FORMAT_TAG and e.value.got == 0x00
    with pytest.raises(vg.ViaError) as e:
        vg.wire_decode_prg(bytes([0x01]) + bytes(10))
    assert (e.value.expected, e.value.got) == (41, 11)


def test_answer_bytes_layout(vg):
    mask = np.arange(512, dtype=np.uint64) * 3
    body = np.arange(512, dtype=np.uint64) + 7
    b = vg.answer_bytes(mask, body)
    assert len(b) == 2 * (9 + 8 * 512)
    assert np.array_equal(vg.wire_decode(b[: 9 + 8 * 512]), mask)
    assert np.array_equal(vg.wire_decode(b[9 + 8 * 512:]), body)
