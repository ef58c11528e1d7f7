"""Assembled streams of tiny clouds, written out byte by byte from the
reference's format, for checking the oracle and the HIP path without either.

Every case uses the truncated-octahedron quantizer at the L2 bound
b = float32(sqrt(5) / 2): then 2b = 2.2360680 and float(double(2b)/sqrt(5))
= 1.0f exactly (truncated_octahedron_quantizer.hpp:29), so a point's scaled
coordinates q are its coordinates.  Every cloud contains a point with zero
coordinates on each axis, so shift_points (shifting.hpp:11-36) subtracts 0.

Fields (block_compressor.hpp:56-215, block_compressor_rle.hpp:120-314):
u64 num_points (N, or the unique cells for RLE), u64 params.size() = 0,
i32 range_x, range_y, range_z, u8 curve, u8 is_direct; per section
u64 data size, u64 meta size, meta, data; RLE counts; f32 offset x, y, z.
"""
import struct

import numpy as np

# float32 bound whose TO divisor is exactly 1.0f
B1 = float(np.float32(np.sqrt(5.0) / 2.0))


def u64(v):
    return struct.pack("<Q", v)


def i32(v):
    return struct.pack("<i", v)


def section(meta: bytes, data: bytes) -> bytes:
    return u64(len(data)) + u64(len(meta)) + meta + data


def header(n, rx, ry, rz, curve, direct):
    return u64(n) + u64(0) + i32(rx) + i32(ry) + i32(rz) + bytes([curve, direct])


TRAILER = struct.pack("<fff", 0.0, 0.0, 0.0)

# Golomb-Rice / BitPacker meta: u32 LE count, u8 k or width
# (golomb_rice_encoder.hpp:126-133, bitpacking_encoder.hpp:27-32).
def le_meta(n, last):
    return struct.pack("<I", n) + bytes([last])


# Huffman meta (huffman_encoder.hpp:126-146): u32 BE entry count, then per
# entry the symbol (u64 BE for 8-byte symbols, raw little-endian bytes for
# u8 / int), u8 length, u32 BE code bits.  A tree of one leaf codes it as
# {bits 0, length 1} (huffman_encoder.hpp:269-273).
def huf_meta_u64(entries):
    return struct.pack(">I", len(entries)) + b"".join(
        struct.pack(">Q", s) + bytes([ln]) + struct.pack(">I", bits) for s, ln, bits in entries)


def huf_meta_u8(entries):
    return struct.pack(">I", len(entries)) + b"".join(
        bytes([s, ln]) + struct.pack(">I", bits) for s, ln, bits in entries)


# Exp-Golomb meta (huffman_encoder.hpp:126-134): magic E1 C1, version 1,
# mode 1, k.
def eg_meta(k):
    return bytes([0xE1, 0xC1, 0x01, 0x01, k])


EMPTY = section(b"", b"")  # an empty patch list packs to empty meta and data


# ---------------------------------------------------------------------------
# Two points P = (2, 0, 0), Q = (0, 2, 0).
#   TO nodes: q = (2,0,0): E = round((2,0,0)*0.5)*2 = (2,0,0), dE = 0, so E;
#   x even: x >>= 1, y >>= 1 -> (1, 0, 0).  q = (0,2,0) -> E (0,2,0) -> (0, 1, 0).
#   Ranges (1, 1, 0).  Morton (z_order_curve.hpp:14-29): (1,0,0) -> 1,
#   (0,1,0) -> 2.
TWO = np.array([[2, 0, 0], [0, 2, 0]], np.float32)

# Direct mode (max range 1 < 1024), Z-order: sorted sfc [1, 2], deltas
# [1, 1]; PFD (p_for_delta_encoder.hpp:27-43): symbols v+1 = [2, 2], no
# patches.  Huffman over {2: 2}: one leaf, code {0, 1}; two 1-bit zero codes
# -> one byte 0x00 (the final byte left-aligned).
TWO_DIRECT = (header(2, 1, 1, 0, 0, 1) + EMPTY +
              section(huf_meta_u64([(2, 1, 0)]), b"\x00") + TRAILER)

# RLE counts [1, 1] (int): the first build is one Huffman leaf with meta
# 4 + 9 = 13 bytes against 1 data byte, 13 > 0.25 * 1, so the counts switch
# to Exp-Golomb and are built again (block_compressor_rle.hpp:290-303).
# select_best_exp_golomb_k over zigzag(1) = 2: k=0 costs 2*bitlen(3)-1 = 3
# bits per value (6), k=1 costs 3+1 = 4 (8), k=2 costs 1+2 = 3 (6, not
# less): k = 0.  Each value: one leading zero, then n1 = 3 as '11' ->
# '011011' -> 0b01101100 = 0x6C.
COUNTS_11 = section(eg_meta(0), b"\x6c")
TWO_DIRECT_RLE = (header(2, 1, 1, 0, 0, 1) + EMPTY +
                  section(huf_meta_u64([(2, 1, 0)]), b"\x00") + COUNTS_11 + TRAILER)

# Block mode (direct_threshold 0): nx = ny = nz = 1/128 + 1 = 1, both nodes
# in block id 0 with cx = cy = cz = 0, reid = Morton(dx, dy, dz) = 1 and 2.
# blkst [0] -> BitPacker [0]: width 1, one zero bit -> 0x00.  blkcnt [2]:
# width 2, value 2 LSB-first -> 0x02.  repos [1, 2-1] = [1, 1] -> PFD [2, 2]
# (no patches; Huffman leaf as above).  quads [0, 0] -> u8 PFD [1, 1], no
# patches; Huffman<u8> one leaf {1: code 0, length 1} -> 0x00.
TWO_BLOCK_BODY = (section(le_meta(1, 1), b"\x00") + section(le_meta(1, 2), b"\x02") + EMPTY +
                  section(huf_meta_u64([(2, 1, 0)]), b"\x00") + EMPTY +
                  section(huf_meta_u8([(1, 1, 0)]), b"\x00"))
TWO_BLOCK = header(2, 1, 1, 0, 0, 0) + TWO_BLOCK_BODY + TRAILER
TWO_BLOCK_RLE = header(2, 1, 1, 0, 0, 0) + TWO_BLOCK_BODY + COUNTS_11 + TRAILER

# ---------------------------------------------------------------------------
# Three points (16,0,0), (0,16,0), (32,0,0): nodes (8,0,0), (0,8,0),
# (16,0,0); ranges (16, 8, 0); Morton 512, 1024, 4096; one block (id 0).
# repos [512, 512, 3072], all above 254: PFD symbols [0, 0, 0], patches
# [512, 512, 3072].
# Golomb-Rice (golomb_rice_encoder.hpp:92-148): sum 4096 over 3 values,
# floor(log2(1365.3)) = 10 = k, M = 1024.  Bits, LSB-first into bytes:
#   512:  q 0 -> '0'; r 512 in 10 bits LSB-first -> bit 9 set: positions 0..10,
#         position 10 set
#   512:  positions 11..21, position 21 set
#   3072: q 3 -> '1110' at 22..25; r 0 -> 26..35 clear
# 36 bits -> bytes 0x00, 0x04 (pos 10), 0xE0 (pos 21, 22, 23), 0x01 (pos 24),
# 0x00.  Huffman over {0: 3}: one leaf, 3 bits -> 0x00.  blkcnt [3]: width 2
# -> 0x03.  quads as above.
THREE = np.array([[16, 0, 0], [0, 16, 0], [32, 0, 0]], np.float32)
THREE_BODY = (section(le_meta(1, 1), b"\x00") + section(le_meta(1, 2), b"\x03") +
              section(le_meta(3, 10), bytes([0x00, 0x04, 0xE0, 0x01, 0x00])) +
              section(huf_meta_u64([(0, 1, 0)]), b"\x00") + EMPTY +
              section(huf_meta_u8([(1, 1, 0)]), b"\x00"))
THREE_BLOCK = header(3, 16, 8, 0, 0, 0) + THREE_BODY + TRAILER

# RLE with the first point repeated: 3 cells, counts in cell order [2, 1, 1].
# First build: Huffman over {2: 1, 1: 2}, meta 4 + 2*9 = 22 bytes > 0.25 * 1
# -> Exp-Golomb.  k over zigzag [4, 2, 2]: k=0 costs 5+3+3 = 11, k=1 costs
# 4+4+4 = 12, k=2 costs 5+3+3 = 11 (not less), k=3: 12 -> k = 0.
# 4: n1 = 5 = '101' after two zeros -> '00101'; 2: n1 = 3 -> '011', twice:
# '00101011' '011' -> 0x2B, 0x60.
FOUR = np.array([[16, 0, 0], [0, 16, 0], [32, 0, 0], [16, 0, 0]], np.float32)
FOUR_BLOCK_RLE = (header(3, 16, 8, 0, 0, 0) + THREE_BODY +
                  section(eg_meta(0), bytes([0x2B, 0x60])) + TRAILER)

# (points, rle, direct_threshold, stream) -- all Z-order (curve byte 0)
CASES = {
    "two_direct": (TWO, False, 1024, TWO_DIRECT),
    "two_direct_rle": (TWO, True, 1024, TWO_DIRECT_RLE),
    "two_block": (TWO, False, 0, TWO_BLOCK),
    "two_block_rle": (TWO, True, 0, TWO_BLOCK_RLE),
    "three_block": (THREE, False, 0, THREE_BLOCK),
    "four_block_rle": (FOUR, True, 0, FOUR_BLOCK_RLE),
}


# ---------------------------------------------------------------------------
# Quantizer known answers at the same bound (q = coordinates).  Each row:
# point, TO node.  Derived from truncated_octahedron_quantizer.hpp:31-50:
#   (5, 0, 0): E_x = round(2.5)*2 = 6 (half away from zero), dE = 1;
#              O = (5, -1, -1) (round(-0.5) = -1), dO = 2 -> E -> (3, 0, 0).
#              (round-half-even would give E_x = 4 -> (2, 0, 0).)
#   (0.5, 0.5, 0.5): E = (0,0,0), O = (1,1,1), dE = dO = 0.75 -> E (dE <= dO)
#              -> (0, 0, 0).  (dE < dO would take O -> (1, 1, 1).)
#   (8388609, 0, 0) = 2^23 + 1: q*0.5 = 4194304.5 exactly, round -> 4194305,
#              E_x = 8388610, dE = 1; O_x = 8388609, O = (8388609, -1, -1),
#              dO = 2 -> E -> x even: (4194305, 0, 0).
#   (1, 1, 0.1): E = (round(0.5)*2, round(0.5)*2, 0) = (2, 2, 0), dE = 2.01;
#              O = (1, 1, round(-0.45)*2+1 = 1), dO = 0.81 -> O; x odd:
#              x = (1+1)>>1 = 1, y = 1 -> (1, 1, 1).
# The origin point fixes the shift at 0.
TO_KNOWN = [([0, 0, 0], [0, 0, 0]),
            ([5, 0, 0], [3, 0, 0]),
            ([0.5, 0.5, 0.5], [0, 0, 0]),
            ([8388609, 0, 0], [4194305, 0, 0]),
            ([1, 1, 0.1], [1, 1, 1])]

# Cube quantizer at b = float32(sqrt(3)/2): float(double(2b)/sqrt(3)) = 1.0f,
# nodes = floor(coordinates) (cube_quantizer.hpp:20-21): just below an
# integer stays below it.
B1_CUBE = float(np.float32(np.sqrt(3.0) / 2.0))
CUBE_KNOWN = [([0, 0, 0], [0, 0, 0]),
              ([float(np.nextafter(np.float32(3), np.float32(0))), 3.0, 0.5], [2, 3, 0]),
              ([float(np.nextafter(np.float32(1), np.float32(0))), 7.999999, 1e-7], [0, 7, 0])]
