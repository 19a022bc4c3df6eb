"""Convert the reference's 64x64 16-bit blue-noise tile into a raw fixture.

Source: data/BlueNoise/64_64/HDR_L_0.png (CC0, Christoph Peters; licence text in
data/BlueNoise/LICENSE.txt). The reference uploads it as a GL_R16 texture
(src/Base/src/Textures.cpp:21-25) and reads it with texelFetch(pos & 0x3f), so a
texel value v decodes to v / 65535.

Output: tests/golden/blue_noise_64x64.u16 -- 4096 little-endian uint16, row-major
(x fastest, row 0 = first PNG scanline, which is GL texel row 0 because
glTextureSubImage2D takes rows bottom-up and stb_image returns them top-down in
memory order, i.e. the first memory row becomes texel row 0).

Run from the repository root:  python tests/golden/make_blue_noise.py
"""
import os
import struct
import sys
import zlib

import numpy as np

SRC = "/root/reference/data/BlueNoise/64_64/HDR_L_0.png"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "blue_noise_64x64.u16")


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def decode_gray16_png(blob: bytes) -> np.ndarray:
    assert blob[:8] == b"\x89PNG\r\n\x1a\n", "not a PNG"
    pos, idat, hdr = 8, b"", None
    while pos < len(blob):
        (ln,) = struct.unpack(">I", blob[pos:pos + 4])
        typ = blob[pos + 4:pos + 8]
        data = blob[pos + 8:pos + 8 + ln]
        pos += 12 + ln
        if typ == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", data)
        elif typ == b"IDAT":
            idat += data
    w, h, depth, ctype, _, _, interlace = hdr
    assert depth == 16 and ctype == 0 and interlace == 0, hdr
    raw = zlib.decompress(idat)
    bpp, stride = 2, w * 2
    out = bytearray(h * stride)
    prev = bytearray(stride)
    for y in range(h):
        ft = raw[y * (stride + 1)]
        line = bytearray(raw[y * (stride + 1) + 1:(y + 1) * (stride + 1)])
        for i in range(stride):
            a = line[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if ft == 1:
                line[i] = (line[i] + a) & 0xFF
            elif ft == 2:
                line[i] = (line[i] + b) & 0xFF
            elif ft == 3:
                line[i] = (line[i] + ((a + b) >> 1)) & 0xFF
            elif ft == 4:
                line[i] = (line[i] + _paeth(a, b, c)) & 0xFF
        out[y * stride:(y + 1) * stride] = line
        prev = line
    return np.frombuffer(bytes(out), dtype=">u2").reshape(h, w).astype(np.uint16)


def main():
    if not os.path.exists(SRC):
        sys.exit(f"{SRC} not found (the fixture is already committed)")
    img = decode_gray16_png(open(SRC, "rb").read())
    img.astype("<u2").tofile(DST)
    print(f"wrote {DST}: shape {img.shape}, min {img.min()}, max {img.max()}, mean {img.mean():.1f}")


if __name__ == "__main__":
    main()
