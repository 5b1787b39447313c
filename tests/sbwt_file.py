"""Test-side reader of the .sbwt plain-matrix v0.1 layout, with byte offsets.

Lets tests inspect the reference's fixture index and write copies with chosen
bits changed (the layout read by SbwtBuilder.cpp:58-132 of the reference:
variant and version strings, four sdsl bit vectors A, C, G, T, their four rank
structures, the suffix-group starts, then the C map and the remaining fields).
"""
import struct

import numpy as np


class PlainMatrix:
    def __init__(self, path):
        self.path = path
        self.raw = bytearray(open(path, "rb").read())
        off = 0

        def u64():
            nonlocal off
            v = struct.unpack_from("<Q", self.raw, off)[0]
            off += 8
            return v

        def string():
            nonlocal off
            n = u64()
            s = bytes(self.raw[off:off + n])
            off += n
            return s

        self.variant = string()
        self.version = string() if self.variant == b"plain-matrix" else self.variant
        self.vec_offsets = []  # byte offset of the words of A, C, G, T, starts
        self.num_bits = None
        for i in range(4):
            bits = u64()
            self.num_bits = bits if self.num_bits is None else self.num_bits
            self.vec_offsets.append(off)
            off += 8 * ((bits + 63) // 64)
        for i in range(4):  # rank structures, skipped
            bits = u64()
            off += 8 * ((bits + 63) // 64)
        bits = u64()
        assert bits == self.num_bits
        self.vec_offsets.append(off)

    @property
    def words(self):
        return (self.num_bits + 63) // 64

    def vector(self, i):
        """Bit vector i (0..3 = A, C, G, T, 4 = suffix-group starts) as u64 words."""
        return np.frombuffer(bytes(self.raw), np.uint64, self.words, self.vec_offsets[i]).copy()

    def bit(self, i, node):
        return int(self.vector(i)[node // 64] >> np.uint64(node % 64)) & 1

    def set_bit(self, i, node, value):
        at = self.vec_offsets[i] + 8 * (node // 64)
        w = struct.unpack_from("<Q", self.raw, at)[0]
        w = (w | (1 << (node % 64))) if value else (w & ~(1 << (node % 64)))
        struct.pack_into("<Q", self.raw, at, w)

    def write(self, path):
        open(path, "wb").write(bytes(self.raw))
