"""MPS in, MPS and the plain-text solution file out (include/samaya_io.h).

read_mps / read_mps_text follow samaya's read_mps / read_mps_from_string
(include/samaya/io.hpp:27-28, src/io/mps_reader.cpp): free MPS first, then fixed MPS, the same
conventions, and a MpsParseError carrying the line number where samaya throws ParseError.
write_solution writes the file of samaya's CLI --solution option (apps/cli/main.cpp:137-198).
The parsing and writing run in the native library.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from . import _native as N
from .lp import INFEASIBLE, UNBOUNDED, LpResult, Model, NativeModel, _p


class MpsParseError(ValueError):
    """Malformed MPS input (samaya's ParseError, include/samaya/io.hpp:11-21)."""

    def __init__(self, message: str, line: int):
        super().__init__(message)
        self.line = line


def _names(lib, h, which: int) -> list[str] | None:
    need = lib.smy_model_names(h, which, None, 0)
    if need <= 1:
        return None
    buf = C.create_string_buffer(need)
    lib.smy_model_names(h, which, buf, need)
    return buf.raw[:need - 1].decode(errors="surrogateescape").split("\n")[:-1]


def _to_model(h) -> Model:
    lib = N.load()
    md = Model._export(h)
    name = C.create_string_buffer(4096)
    lib.smy_model_name(h, name, len(name))
    md.name = name.value.decode(errors="surrogateescape")
    md.row_names = _names(lib, h, 0)
    md.col_names = _names(lib, h, 1)
    marks = np.zeros(md.n, np.uint8)
    lib.smy_model_integrality(h, _p(marks))
    md.col_integer = marks
    return md


def _native_from(lib, h, what: str, code: int, err, line: int) -> NativeModel:
    if code == N.SMY_INVALID:
        raise MpsParseError(err.value.decode(errors="replace"), line)
    if code != N.SMY_OK:
        raise OSError(err.value.decode(errors="replace") or what)
    return NativeModel._adopt(h, lib.smy_model_num_rows(h), lib.smy_model_num_cols(h))


def read_mps_native(path: str | os.PathLike) -> NativeModel:
    """The model of an MPS file, kept in the native library (no copy into Python arrays)."""
    lib = N.load()
    h, err = C.c_void_p(), C.create_string_buffer(1024)
    code = lib.smy_read_mps(os.fsencode(path), C.byref(h), err, len(err))
    # The message reads "line N: ..." when a line is at fault.
    found = re.match(r"line (\d+): ", err.value.decode(errors="replace"))
    return _native_from(lib, h, "read_mps", code, err, int(found.group(1)) if found else 0)


def read_mps_text_native(text: str | bytes, threads: int = 0) -> NativeModel:
    lib = N.load()
    data = text.encode() if isinstance(text, str) else bytes(text)
    h, err, ln = C.c_void_p(), C.create_string_buffer(1024), C.c_longlong()
    code = lib.smy_read_mps_text(data, len(data), threads, C.byref(h), err, len(err), C.byref(ln))
    return _native_from(lib, h, "read_mps_text", code, err, ln.value)


def read_mps(path: str | os.PathLike) -> Model:
    """read_mps (include/samaya/io.hpp:27): a Model with its name, row / column names and
    integrality. Raises MpsParseError on malformed input, OSError when the file cannot be read."""
    nm = read_mps_native(path)
    md = _to_model(nm.h)
    md._native = nm
    return md


def read_mps_text(text: str | bytes, threads: int = 0) -> Model:
    """read_mps_from_string (include/samaya/io.hpp:28); threads <= 0: up to 16 hardware threads."""
    nm = read_mps_text_native(text, threads)
    md = _to_model(nm.h)
    md._native = nm
    return md


def _native(model: Model | NativeModel) -> NativeModel:
    if isinstance(model, NativeModel):
        return model
    cached = getattr(model, "_native", None)
    return cached if cached is not None else model.native()


def mps_text(model: Model | NativeModel) -> str:
    """Free MPS of a model (names and integrality included)."""
    lib, nm = N.load(), _native(model)
    need = lib.smy_write_mps_text(nm.h, None, 0)
    N.check(0 if need > 0 else -need, "write_mps_text")
    buf = C.create_string_buffer(need)
    lib.smy_write_mps_text(nm.h, buf, need)
    return buf.raw[:need - 1].decode(errors="surrogateescape")


def write_mps(model: Model | NativeModel, path: str | os.PathLike) -> None:
    lib, nm = N.load(), _native(model)
    err = C.create_string_buffer(1024)
    if lib.smy_write_mps(nm.h, os.fsencode(path), err, len(err)) != N.SMY_OK:
        raise OSError(err.value.decode(errors="replace"))


def _solution_args(nm: NativeModel, result: LpResult, with_values: bool):
    res = N.SmyLpResult()
    res.status, res.verified, res.objective = result.status, int(result.verified), result.objective

    def arr(a, size: int):  # None for absent or empty vectors; the model's size otherwise
        if a is None or len(a) == 0:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (size,):
            raise ValueError(f"solution vector of length {len(a)} where {size} is needed")
        return a

    m, n = nm.m, nm.n
    keep = [arr(result.col_value, n) if with_values else None, arr(result.col_dual, n),
            arr(result.row_activity, m) if with_values else None, arr(result.row_dual, m)]
    cert = None
    if result.status in (INFEASIBLE, UNBOUNDED):
        cert = arr(result.certificate, m if result.status == INFEASIBLE else n)
    keep.append(cert)
    return res, keep


def solution_text(model: Model | NativeModel, result: LpResult, with_values: bool = True) -> str:
    """The --solution file's text for an LP answer (apps/cli/main.cpp:137-198)."""
    lib, nm = N.load(), _native(model)
    res, (x, d, act, y, cert) = _solution_args(nm, result, with_values)
    args = (nm.h, C.byref(res), _p(x), _p(d), _p(act), _p(y), _p(cert))
    need = lib.smy_solution_text(*args, None, 0)
    N.check(0 if need > 0 else -need, "solution_text")
    buf = C.create_string_buffer(need)
    lib.smy_solution_text(*args, buf, need)
    return buf.raw[:need - 1].decode(errors="surrogateescape")


def write_solution(model: Model | NativeModel, result: LpResult, path: str | os.PathLike,
                   with_values: bool = True) -> None:
    lib, nm = N.load(), _native(model)
    res, (x, d, act, y, cert) = _solution_args(nm, result, with_values)
    err = C.create_string_buffer(1024)
    code = lib.smy_write_solution(nm.h, os.fsencode(path), C.byref(res), _p(x), _p(d), _p(act),
                                  _p(y), _p(cert), err, len(err))
    if code != N.SMY_OK:
        raise OSError(err.value.decode(errors="replace"))
