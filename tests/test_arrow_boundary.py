"""The Arrow C Data Interface at the C ABI, host side (no GPU needed):
kg_column_from_arrow views every supported primitive format without a copy
-- type, precision and scale from the format string, the array's offset,
its validity only when it has NULLs -- and refuses the formats the kernels
do not take (strings, dictionaries, nested arrays)."""
import ctypes as C
import decimal

import numpy as np
import pyarrow as pa
import pytest

import kgpu
from kgpu import abi as A

CASES = [
    (pa.array([True, None, False]), A.T_BOOL, 0, 0),
    (pa.array([1, None, -3], pa.int8()), A.T_INT8, 0, 0),
    (pa.array([1, None, -3], pa.int16()), A.T_INT16, 0, 0),
    (pa.array([1, None, -3], pa.int32()), A.T_INT32, 0, 0),
    (pa.array([1, None, -3], pa.int64()), A.T_INT64, 0, 0),
    (pa.array([1, None, 3], pa.uint8()), A.T_UINT8, 0, 0),
    (pa.array([1, None, 3], pa.uint16()), A.T_UINT16, 0, 0),
    (pa.array([1, None, 3], pa.uint32()), A.T_UINT32, 0, 0),
    (pa.array([1, None, 3], pa.uint64()), A.T_UINT64, 0, 0),
    (pa.array([1.5, None, 3], pa.float32()), A.T_FLOAT32, 0, 0),
    (pa.array([1.5, None, 3], pa.float64()), A.T_FLOAT64, 0, 0),
    (pa.array([1, None, 3], pa.date32()), A.T_DATE32, 0, 0),
    (pa.array([1, None, 3], pa.timestamp("us", tz="UTC")), A.T_TIMESTAMP, 0, 0),
    (pa.array([1, None, 3], pa.date64()), A.T_TIMESTAMP, 0, 0),
    (pa.array([decimal.Decimal("1.25"), None], pa.decimal128(20, 2)), A.T_DECIMAL128, 20, 2),
    (pa.array([decimal.Decimal("1.25"), None], pa.decimal64(12, 2)), A.T_DECIMAL64, 12, 2),
]


@pytest.mark.parametrize("arr,t,p,s", CASES, ids=[str(c[0].type) for c in CASES])
def test_primitive_formats_are_viewed_in_place(arr, t, p, s):
    col = kgpu.arrow_col(arr)
    assert (col.type, col.precision, col.scale, col.length, col.offset) == (t, p, s, len(arr), 0)
    assert col.vptr == arr.buffers()[1].address
    assert col.bptr == arr.buffers()[0].address
    sl = arr.slice(1)
    c2 = kgpu.arrow_col(sl)
    assert (c2.offset, c2.length) == (1, len(arr) - 1)
    no_nulls = pa.array([1, 2, 3], arr.type) if t not in (A.T_BOOL, A.T_DECIMAL64,
                                                          A.T_DECIMAL128) else None
    if no_nulls is not None:
        assert kgpu.arrow_col(no_nulls).bptr is None


@pytest.mark.parametrize("arr", [pa.array(["a", "b"]), pa.array([1, 2]).dictionary_encode(),
                                 pa.array([[1], [2]]), pa.array([1, 2], pa.time32("s"))])
def test_other_formats_are_refused(arr):
    with pytest.raises(kgpu.KgError, match="UNSUPPORTED"):
        kgpu.arrow_col(arr)
