"""BMP layout of the reference's writer (loadbmp.h:182-245): header fields,
24-bit grey pixels, bottom-up rows padded to 4 bytes."""
import numpy as np

from uncc_sar_focusing_amd.bmp import encode_bmp


def test_bmp_header_padding_and_row_order():
    rows = np.arange(2 * 5, dtype=np.uint8).reshape(2, 5)  # H = 2, W = 5 -> pad 1
    b = encode_bmp(rows)
    assert b[:2] == b"BM"
    assert int.from_bytes(b[2:6], "little") == 54 + 3 * 5 * 2  # size field ignores padding
    assert b[10] == 54 and b[14] == 40
    assert int.from_bytes(b[18:22], "little") == 5
    assert int.from_bytes(b[22:26], "little") == 2
    assert b[26] == 1 and b[28] == 24
    assert all(v == 0 for v in b[30:54])
    body = b[54:]
    assert len(body) == 2 * (15 + 1)
    # first stored row is the bottom display row (rows[1]), each grey pixel x3
    assert list(body[:15]) == [v for v in rows[1] for _ in range(3)]
    assert body[15] == 0
    assert list(body[16:31]) == [v for v in rows[0] for _ in range(3)]
