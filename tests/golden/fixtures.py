"""Loads the golden JSON fixtures into HostBatch objects."""
import json
import os
from decimal import Decimal

import numpy as np

from gazelle_plugin_amd import types as T
from gazelle_plugin_amd.batch import HostBatch, HostColumn, bitmap_from_bools

HERE = os.path.dirname(os.path.abspath(__file__))

_TYPES = {"null": T.NA, "bool": T.BOOL, "int8": T.INT8, "uint8": T.UINT8, "int16": T.INT16,
          "uint16": T.UINT16, "int32": T.INT32, "uint32": T.UINT32, "int64": T.INT64,
          "uint64": T.UINT64, "float": T.FLOAT, "double": T.DOUBLE, "utf8": T.STRING}


def type_id(name: str) -> int:
    if name.startswith("decimal128"):
        return T.DECIMAL128
    return _TYPES[name]


def decimal_scale(name: str) -> int:
    return int(name.split(",")[1].rstrip(")")) if name.startswith("decimal128") else 0


def column(name: str, values) -> HostColumn:
    t = type_id(name)
    n = len(values)
    valid = np.array([v is not None for v in values])
    vb = bitmap_from_bools(valid) if not valid.all() else None
    if t == T.NA:
        return HostColumn(t, None, None, None)
    if t == T.BOOL:
        return HostColumn(t, bitmap_from_bools(np.array([bool(v) for v in values])), vb)
    if t == T.STRING:
        enc = [(v or "").encode() for v in values]
        off = np.zeros(n + 1, np.int32)
        off[1:] = np.cumsum([len(e) for e in enc])
        data = np.frombuffer(b"".join(enc), np.uint8).copy()
        return HostColumn(t, data, vb, off)
    if t == T.DECIMAL128:
        scale = decimal_scale(name)
        raw = np.zeros(2 * n, np.int64)
        for i, v in enumerate(values):
            u = int(Decimal(v).scaleb(scale)) if v is not None else 0
            raw[2 * i] = np.int64(np.uint64(u & 0xFFFFFFFFFFFFFFFF))
            raw[2 * i + 1] = -1 if u < 0 else 0
        return HostColumn(t, raw.view(np.uint8), vb)
    dt = T.numpy_dtype(t)
    arr = np.array([0 if v is None else (float("nan") if v == "NaN" else v) for v in values]).astype(dt)
    return HostColumn(t, arr, vb)


def batch(schema, cols) -> HostBatch:
    n = len(cols[0])
    return HostBatch(n, [column(typ, vals) for (_, typ), vals in zip(schema, cols)])


def load(name: str):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)
