"""Column type ids: the values of arrow::Type::type, as in include/gazelle_amd.h."""
import numpy as np

NA = 0
BOOL = 1
UINT8 = 2
INT8 = 3
UINT16 = 4
INT16 = 5
UINT32 = 6
INT32 = 7
UINT64 = 8
INT64 = 9
FLOAT = 11
DOUBLE = 12
STRING = 13
BINARY = 14
DATE32 = 16
DATE64 = 17
TIMESTAMP = 18
DECIMAL128 = 23

_WIDTH = {UINT8: 1, INT8: 1, UINT16: 2, INT16: 2, UINT32: 4, INT32: 4, FLOAT: 4, DATE32: 4,
          UINT64: 8, INT64: 8, DOUBLE: 8, DATE64: 8, TIMESTAMP: 8, DECIMAL128: 16}

_NP = {UINT8: np.uint8, INT8: np.int8, UINT16: np.uint16, INT16: np.int16, UINT32: np.uint32,
       INT32: np.int32, FLOAT: np.float32, DATE32: np.int32, UINT64: np.uint64, INT64: np.int64,
       DOUBLE: np.float64, DATE64: np.int64, TIMESTAMP: np.int64}


def is_binary(t: int) -> bool:
    return t in (STRING, BINARY)


def is_fixed(t: int) -> bool:
    return t in _WIDTH


def width(t: int) -> int:
    return _WIDTH[t]


def numpy_dtype(t: int):
    if t == DECIMAL128:
        return np.dtype([("lo", np.uint64), ("hi", np.uint64)])
    return np.dtype(_NP[t])


def from_arrow_type(typ) -> int:
    import pyarrow as pa

    table = [
        (pa.types.is_null, NA), (pa.types.is_boolean, BOOL), (pa.types.is_uint8, UINT8),
        (pa.types.is_int8, INT8), (pa.types.is_uint16, UINT16), (pa.types.is_int16, INT16),
        (pa.types.is_uint32, UINT32), (pa.types.is_int32, INT32), (pa.types.is_uint64, UINT64),
        (pa.types.is_int64, INT64), (pa.types.is_float32, FLOAT), (pa.types.is_float64, DOUBLE),
        (pa.types.is_string, STRING), (pa.types.is_binary, BINARY), (pa.types.is_date32, DATE32),
        (pa.types.is_date64, DATE64), (pa.types.is_timestamp, TIMESTAMP),
        (pa.types.is_decimal128, DECIMAL128),
    ]
    for pred, tid in table:
        if pred(typ):
            return tid
    raise NotImplementedError(f"type {typ} is not supported by the split/aggregate/join path")


def to_arrow_type(t: int):
    import pyarrow as pa

    return {NA: pa.null(), BOOL: pa.bool_(), UINT8: pa.uint8(), INT8: pa.int8(),
            UINT16: pa.uint16(), INT16: pa.int16(), UINT32: pa.uint32(), INT32: pa.int32(),
            UINT64: pa.uint64(), INT64: pa.int64(), FLOAT: pa.float32(), DOUBLE: pa.float64(),
            STRING: pa.utf8(), BINARY: pa.binary(), DATE32: pa.date32(), DATE64: pa.date64(),
            TIMESTAMP: pa.timestamp("us"), DECIMAL128: pa.decimal128(38, 0)}[t]
