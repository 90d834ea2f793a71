"""Pins of the oracle's exchange frame (oracle/kgo_frame.c) against the
reference's own checks (crates/krishiv-shuffle/src/tests.rs:73-126): the
CRC32C check value, a header assembled by hand from the documented layout
(frame.rs:6-24), and every refusal of a damaged frame."""
import struct

import pytest

import kgo


def hdr(kind=kgo.FRAME_DATA):
    return kgo.FrameHeader(id=0x0102030405060708, stage_attempt=3, map=7, attempt=1,
                           sequence=9, reduce=11, kind=kind)


def test_crc32c_check_values():
    assert kgo.crc32c(b"123456789") == 0xE3069283
    assert kgo.crc32c(b"") == 0
    assert kgo.crc32c(bytes(32)) == 0x8A9136AA  # RFC 3720 B.4: 32 bytes of zeros
    assert kgo.crc32c(b"\xff" * 32) == 0x62A8AB43  # RFC 3720 B.4: 32 bytes of ones


def test_header_layout_by_hand():
    body = b"an arrow ipc stream would be here"
    f = kgo.frame_encode(hdr(), body)
    want = struct.pack("<4sHBBQIIIIIII", b"KXFR", 1, 1, 0, 0x0102030405060708, 3, 7, 1, 9, 11,
                       len(body), kgo.crc32c(body))
    assert len(want) == 44 and f == want + body
    h, b = kgo.frame_decode(f)
    assert (h.kind, h.id, h.stage_attempt, h.map, h.attempt, h.sequence, h.reduce) == (
        1, 0x0102030405060708, 3, 7, 1, 9, 11) and b == body
    end = kgo.frame_encode(hdr(kgo.FRAME_END), b"")
    assert len(end) == 44 and end[40:44] == bytes(4)


def test_damaged_frames_are_refused():
    good = kgo.frame_encode(hdr(), b"0123456789abcdef")
    cases = [(good[:-2] + bytes([good[-2] ^ 1]) + good[-1:], "checksum"),
             (good[:4] + b"\x02\x00" + good[6:], "version"),
             (b"KXFX" + good[4:], "magic"),
             (good[:6] + b"\x04" + good[7:], "kind"),
             (good[:7] + b"\x01" + good[8:], "compression"),
             (good[:-1], "truncated"),
             (good[:43], "truncated"),
             (good + b"\0", "body")]
    for frame, why in cases:
        with pytest.raises(ValueError, match=why):
            kgo.frame_decode(frame)
