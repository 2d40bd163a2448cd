"""On-disk formats (HAPR_fluxion_v4/src/parser/): the BLIF reader
(blif.rs:13-122) and the DEF writer (def.rs:12-41), both in the library's
host code (fx_blif_read, fx_def_write)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import check, lib, ptr
from .placer import Circuit, Placement

__all__ = ["parse_blif", "write_def"]


def parse_blif(path) -> Circuit:
    """blif::parse(&Path) -> Result<Circuit, String> (blif.rs:13-122).

    Nets come in a fixed order (driven nets by first drive, then
    readers-only nets by first read) where the reference iterates a HashMap;
    the nets themselves are the reference's."""
    h = C.c_void_p()
    check(lib.fx_blif_read(os.fsencode(os.fspath(path)), C.byref(h)))
    try:
        ng, ne = C.c_int32(), C.c_int32()
        npins = C.c_int64()
        w, hh = C.c_float(), C.c_float()
        check(lib.fx_circuit_info(h, C.byref(ng), C.byref(ne), C.byref(npins), C.byref(w),
                                  C.byref(hh)))
        net_ptr = np.zeros(ne.value + 1, np.int32)
        net_pins = np.zeros(max(npins.value, 1), np.int32)
        types = np.zeros(max(ng.value, 1), np.uint8)
        check(lib.fx_circuit_netlist(h, ptr(net_ptr), ptr(net_pins), ptr(types)))
        need = C.c_int64()
        check(lib.fx_circuit_net_names(h, None, 0, C.byref(need)))
        buf = C.create_string_buffer(max(need.value, 1))
        check(lib.fx_circuit_net_names(h, buf, need.value, C.byref(need)))
        names = buf.raw[:need.value].decode().split("\0")[:-1] if need.value else []
        name = lib.fx_circuit_name(h).decode()
    finally:
        lib.fx_circuit_free(h)
    return Circuit(ng.value, net_ptr, net_pins[:npins.value], w.value, hh.value, name=name,
                   gate_types=types[:ng.value], net_names=names)


def write_def(path, circuit: Circuit, placement: Placement) -> None:
    """def::write_def(&Path, &Circuit, &Placement) (def.rs:12-41)."""
    x = np.ascontiguousarray(placement.gate_x, np.float32)
    y = np.ascontiguousarray(placement.gate_y, np.float32)
    names = None
    if circuit.gate_names is not None:
        arr = (C.c_char_p * circuit.num_gates)(*[n.encode() for n in circuit.gate_names])
        names = C.cast(arr, C.c_void_p)
    check(lib.fx_def_write(os.fsencode(os.fspath(path)), circuit.name.encode(), circuit.num_gates,
                           names, ptr(x), ptr(y), float(circuit.chip_width),
                           float(circuit.chip_height)))
