"""CPU: the checkpoint file format, read by the library's host-only inspector.

Files here are written by an independent Python writer that follows the
format as documented in include/zsph.h (JSON header of top-level scalars
padded to 4096 bytes, then the raw sections, FNV-1a 64 checksum); the
library must read them, verify the checksum, and reject corrupt, truncated
or foreign files with ZSPH_E_IO.  The FNV-1a 64 restatement is pinned by the
published test vectors of the algorithm.  The device side (a restart that
continues bit for bit) is tests/test_gpu_checkpoint.py.
"""
import json

import numpy as np
import pytest

import zsph_amd as z
from zsph_amd import _abi as abi

MASK64 = (1 << 64) - 1


def fnv1a64(data, h=0xCBF29CE484222325):
    for b in bytes(data):
        h ^= b
        h = (h * 0x100000001B3) & MASK64
    return h


def test_fnv1a64_published_vectors():
    # FNV-1a 64-bit test vectors (Fowler/Noll/Vo reference test suite)
    assert fnv1a64(b"") == 0xCBF29CE484222325
    assert fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert fnv1a64(b"foobar") == 0x85944171F73967E8


def make_sections(n=37, seed=0):
    rng = np.random.default_rng(seed)
    return [("posr", rng.normal(size=(n, 4)).astype(np.float32).tobytes()),
            ("velp", rng.normal(size=(n, 4)).astype(np.float32).tobytes()),
            ("type", rng.integers(0, 3, n).astype(np.int32).tobytes()),
            ("body", bytes(range(200)))]


def write_ckpt(path, sections, step=1234, time=0.123456789012345678, **over):
    payload = b"".join(b for _, b in sections)
    hdr = {"format": "zsph-checkpoint", "version": 1, "endian": "little", "step": step, "time": time,
           "t0": time - 1e-4, "ptc_num": 37, "cells": 99, "scenes": 1, "dim": 3, "grid_x": 3, "grid_y": 3,
           "grid_z": 11, "killed": 2, "has_body": 1, "has_piston": 0, "closure": 0,
           "params": "%016x" % 0x0123456789ABCDEF, "sections": ",".join(n for n, _ in sections)}
    for n, b in sections:
        hdr[n + "_bytes"] = len(b)
    hdr["payload_bytes"] = len(payload)
    hdr["checksum"] = "%016x" % fnv1a64(payload)
    hdr.update(over)
    text = json.dumps(hdr)
    assert len(text) < 4096
    with open(path, "wb") as f:
        f.write((text.ljust(4095) + "\n").encode())
        f.write(payload)
    return payload


def code_of(path, verify=True):
    with pytest.raises(z.ZsphError) as e:
        z.checkpoint_info(str(path), verify=verify)
    return e.value.code


def test_inspect_reads_the_documented_format(tmp_path):
    p = tmp_path / "c.bin"
    secs = make_sections()
    payload = write_ckpt(p, secs)
    info = z.checkpoint_info(str(p))
    assert info["step"] == 1234 and info["time"] == 0.123456789012345678
    assert (info["ptc_num"], info["cells"], info["scenes"], info["dim"]) == (37, 99, 1, 3)
    assert info["grid"] == (3, 3, 11) and info["killed"] == 2
    assert (info["has_body"], info["has_piston"], info["closure"]) == (1, 0, 0)
    assert info["payload_bytes"] == len(payload)


def test_checksum_catches_one_flipped_bit(tmp_path):
    p = tmp_path / "c.bin"
    payload = write_ckpt(p, make_sections())
    raw = bytearray(p.read_bytes())
    raw[4096 + len(payload) // 2] ^= 0x10
    p.write_bytes(bytes(raw))
    assert code_of(p) == abi.E_IO
    assert z.checkpoint_info(str(p), verify=False)["step"] == 1234  # the header alone is intact


def test_truncated_and_foreign_files_are_rejected(tmp_path):
    p = tmp_path / "c.bin"
    write_ckpt(p, make_sections())
    full = p.read_bytes()
    p.write_bytes(full[:-1])  # one byte short
    assert code_of(p) == abi.E_IO
    p.write_bytes(full + b"\0")  # one byte long
    assert code_of(p) == abi.E_IO
    p.write_bytes(full[:1000])  # header cut
    assert code_of(p, verify=False) == abi.E_IO
    p.write_bytes(b'{"format": "something-else"}'.ljust(4095) + b"\n")
    assert code_of(p, verify=False) == abi.E_IO
    write_ckpt(p, make_sections(), version=2)
    assert code_of(p, verify=False) == abi.E_IO
    write_ckpt(p, make_sections(), payload_bytes=1)  # sections do not add up
    assert code_of(p, verify=False) == abi.E_IO
    write_ckpt(p, make_sections(), checksum="not-hex-digits!!")
    assert code_of(p, verify=False) == abi.E_IO
    assert code_of(tmp_path / "missing.bin", verify=False) == abi.E_IO


def test_checkpoint_info_struct_matches_the_library():
    assert abi.load(z.LIB_PATH).zsph_sizeof(11) == abi.C.sizeof(abi.CheckpointInfo)
