"""The device exchange frame writer (kg_exchange_frames) against the oracle's
frame decoder (oracle/kgo_frame.c: header layout and CRC32C,
krishiv-shuffle/src/frame.rs) and pyarrow's Arrow IPC reader (the body is an
IPC stream of one batch, frame.rs:205-221): every partition's frame decodes,
its checksum holds, and its batch is that partition's rows."""
import numpy as np
import pyarrow as pa
import pytest

from gpu_util import ctx, say
import kgo
import kgpu
from kgpu import Col, DeviceBuffer
from kgpu import abi as A

pytestmark = pytest.mark.gpu


def base_header():
    return A.FrameHeader(id=77, stage_attempt=2, map=5, attempt=1, sequence=4, reduce=0,
                         kind=A.FRAME_DATA)


def columns(rng, n):
    """(kgpu columns, pyarrow arrays of the same rows, names)."""
    i64 = rng.integers(-(1 << 62), 1 << 62, n)
    i32 = rng.integers(-(1 << 31), 1 << 31, n).astype(np.int32)
    f64 = rng.standard_normal(n)
    b = rng.random(n) > 0.5
    d64 = rng.integers(-10 ** 11, 10 ** 11, n)
    d128 = [int(x) * 10 ** 20 + int(x) % 977 for x in rng.integers(-10 ** 6, 10 ** 6, n)]
    dt = rng.integers(0, 40000, n).astype(np.int32)
    ts = rng.integers(0, 1 << 50, n)
    i8 = rng.integers(-128, 128, n).astype(np.int8)
    u16 = rng.integers(0, 1 << 16, n).astype(np.uint16)
    masks = [rng.random(n) > 0.1 for _ in range(10)]
    masks[2] = None  # one column without validity
    cols = [Col(A.T_INT64, i64, masks[0]), Col(A.T_INT32, i32, masks[1]), Col(A.T_FLOAT64, f64),
            Col(A.T_BOOL, b, masks[3]), Col(A.T_DECIMAL64, d64, masks[4], 12, 2),
            Col(A.T_DECIMAL128, d128, masks[5], 30, 4), Col(A.T_DATE32, dt, masks[6]),
            Col(A.T_TIMESTAMP, ts, masks[7]), Col(A.T_INT8, i8, masks[8]),
            Col(A.T_UINT16, u16, masks[9])]
    import decimal
    decimal.getcontext().prec = 40
    arrays = [pa.array(i64, mask=~masks[0]), pa.array(i32, mask=~masks[1]), pa.array(f64),
              pa.array(b, mask=~masks[3]),
              pa.array([decimal.Decimal(int(v)).scaleb(-2) for v in d64], pa.decimal64(12, 2),
                       mask=~masks[4]),
              pa.array([decimal.Decimal(v).scaleb(-4) for v in d128], pa.decimal128(30, 4),
                       mask=~masks[5]),
              pa.array(dt, pa.date32(), mask=~masks[6]),
              pa.array(ts, pa.timestamp("us"), mask=~masks[7]), pa.array(i8, mask=~masks[8]),
              pa.array(u16, mask=~masks[9])]
    names = ["k", "i32", "f", "flag", "price", "big", "day", "at", "tiny", "u16"]
    return cols, arrays, names


def check_frames(out, fo, total, offsets, arrays, names, hdr):
    data = out.to_numpy(np.uint8, total).tobytes()
    parts = len(offsets) - 1
    for p in range(parts):
        lo, hi = int(offsets[p]), int(offsets[p + 1])
        frame = data[int(fo[p]):int(fo[p + 1])]
        if hi == lo:
            assert frame == b"", p
            continue
        h, body = kgo.frame_decode(frame)  # magic, version, kind, length, CRC32C
        assert (h.kind, h.id, h.stage_attempt, h.map, h.attempt, h.sequence, h.reduce) == (
            1, hdr.id, hdr.stage_attempt, hdr.map, hdr.attempt, hdr.sequence, p)
        reader = pa.ipc.open_stream(body)
        batches = list(reader)
        assert len(batches) == 1, p
        rb = batches[0]
        rb.validate(full=True)
        assert rb.schema.names == names
        assert rb.num_rows == hi - lo
        for c, arr in enumerate(arrays):
            assert rb.column(c).equals(arr.slice(lo, hi - lo)), (p, names[c])
    assert int(fo[parts]) == total


@pytest.mark.parametrize("parts,n", [(1, 1000), (7, 5003), (64, 100_000), (1000, 300)])
def test_frames_decode_to_the_partitions_rows(parts, n):
    say(f"frames: {parts} partitions, {n} rows")
    rng = np.random.default_rng(parts)
    cols, arrays, names = columns(rng, n)
    cuts = np.sort(rng.integers(0, n + 1, parts - 1))
    offsets = np.concatenate([[0], cuts, [n]]).astype(np.int64)
    if parts > 2:
        offsets[2] = offsets[1]  # an empty partition
    hdr = base_header()
    out, fo, total = kgpu.exchange_frames(ctx(), cols, offsets, hdr, names)
    check_frames(out, fo, total, offsets, arrays, names, hdr)


def test_frames_after_repartition_on_the_device():
    """ExchangeSink's route + split (kg_repartition) then its frames."""
    say("frames: repartition then frames, 2M rows x 3 columns, 256 partitions")
    rng = np.random.default_rng(9)
    n, parts = 2_000_000, 256
    key = rng.integers(0, 1 << 40, n)
    v = rng.standard_normal(n)
    flag = rng.random(n) > 0.3
    ok = rng.random(n) > 0.05
    cols = [Col(A.T_INT64, key, ok), Col(A.T_FLOAT64, v), Col(A.T_BOOL, flag)]
    batch, offsets = kgpu.repartition(ctx(), cols, [0], parts)
    dcols = [batch.as_col(i) for i in range(3)]
    hdr = base_header()
    out, fo, total = kgpu.exchange_frames(ctx(), dcols, offsets, hdr, ["k", "v", "flag"])
    pids = kgo.route([kgo.Col(kgo.T_INT64, key, ok)], parts)
    order = np.argsort(pids, kind="stable")
    arrays = [pa.array(key[order], mask=~ok[order]), pa.array(v[order]), pa.array(flag[order])]
    check_frames(out, fo, total, offsets, arrays, ["k", "v", "flag"], hdr)


def test_end_frame_and_crc32c_on_the_device():
    say("frames: End frame, CRC32C of device bytes")
    h = A.FrameHeader(id=5, stage_attempt=1, map=2, attempt=3, sequence=4, reduce=6,
                      kind=A.FRAME_END)
    oh = kgo.FrameHeader(id=5, stage_attempt=1, map=2, attempt=3, sequence=4, reduce=6,
                         kind=kgo.FRAME_END)
    assert kgpu.frame_end(h) == kgo.frame_encode(oh, b"")
    rng = np.random.default_rng(3)
    for ln in (1, 3, 4, 255, 256, 257, 65535, 65536, 65537, 1_000_003, 40_000_000):
        data = rng.integers(0, 256, ln, dtype=np.uint8)
        dev = DeviceBuffer.from_numpy(ctx(), data)
        assert kgpu.crc32c(ctx(), dev, ln) == kgo.crc32c(data.tobytes()), ln
    dev = DeviceBuffer.from_numpy(ctx(), np.frombuffer(b"123456789", np.uint8))
    assert kgpu.crc32c(ctx(), dev, 9) == 0xE3069283
