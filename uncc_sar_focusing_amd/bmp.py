"""BMP output in the reference's file layout.

The reference writes its display image with loadbmp_encode_file
(include/third_party/loadbmp.h:182-245, called from
include/uncc_sar_globals.hpp:110-111): a 54-byte header (file size field
54 + 3*W*H, planes 1, 24 bits, no compression, other fields 0), then 24-bit
BGR rows bottom-up, each padded to a multiple of 4 bytes.  Grey pixels are
written to all three channels.  `rows` is the top-down 8-bit image that
Context.image_db8 (or the oracle) returns.
"""
from __future__ import annotations

import numpy as np


def encode_bmp(rows: np.ndarray) -> bytes:
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    if rows.ndim != 2 or rows.size == 0:
        raise ValueError("rows must be a non-empty (H, W) uint8 array")
    H, W = rows.shape
    size = 54 + W * H * 3
    file_hdr = bytearray(b"BM" + bytes(12))
    file_hdr[2:6] = (size & 0xFFFFFFFF).to_bytes(4, "little")
    file_hdr[10] = 54
    info = bytearray(40)
    info[0] = 40
    info[4:8] = W.to_bytes(4, "little")
    info[8:12] = H.to_bytes(4, "little")
    info[12] = 1
    info[14] = 24
    pad = (4 - (W * 3) % 4) % 4
    px = np.repeat(rows[::-1], 3, axis=1)  # bottom-up, grey into B, G, R
    if pad:
        px = np.concatenate([px, np.zeros((H, pad), dtype=np.uint8)], axis=1)
    return bytes(file_hdr) + bytes(info) + px.tobytes()


def write_bmp(path: str, rows) -> None:
    """Write a top-down (H, W) uint8 image (numpy or a torch tensor)."""
    if hasattr(rows, "cpu"):
        rows = rows.cpu().numpy()
    with open(path, "wb") as f:
        f.write(encode_bmp(np.asarray(rows)))
