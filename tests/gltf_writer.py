"""Writers for the asset tests: glTF 2.0 / GLB files, PNG files (every
colour type and bit depth, all five row filters, Adam7 interlacing) and BMP
files laid out the way the reference's reader walks them.  Tests build
their inputs with these at run time, so every expected value follows from
the numbers written here and the file formats' specifications."""
import json
import struct
import zlib

import numpy as np

BYTE, UBYTE, SHORT, USHORT, UINT, FLOAT = 5120, 5121, 5122, 5123, 5125, 5126
_NP = {UBYTE: np.uint8, USHORT: np.uint16, UINT: np.uint32, FLOAT: np.float32, SHORT: np.int16,
       BYTE: np.int8}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}


class Gltf:
    """Accumulates one binary buffer and the JSON document around it."""

    def __init__(self):
        self.bin = bytearray()
        self.j = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": []}

    def _pad(self, n=4):
        while len(self.bin) % n:
            self.bin.append(0)

    def view(self, data, stride=None):
        self._pad()
        v = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        self.bin += data
        self.j["bufferViews"].append(v)
        return len(self.j["bufferViews"]) - 1

    def accessor(self, arr, ctype, typ, stride=None, pad_to=None):
        """arr (count, ncomp); with stride, each element starts `stride` bytes
        after the previous one (the gap filled with 0x5A bytes)."""
        a = np.ascontiguousarray(arr, _NP[ctype]).reshape(len(arr), -1)
        assert a.shape[1] == _NCOMP[typ]
        if stride:
            rows = [r.tobytes() + b"\x5a" * (stride - a.itemsize * a.shape[1]) for r in a]
            data = b"".join(rows)
        else:
            data = a.tobytes()
        if pad_to:
            data += b"\x00" * (pad_to - len(data))
        v = self.view(data, stride)
        self.j["accessors"].append({"bufferView": v, "componentType": ctype, "count": len(a),
                                    "type": typ})
        return len(self.j["accessors"]) - 1

    def add(self, key, obj):
        self.j.setdefault(key, []).append(obj)
        return len(self.j[key]) - 1

    def primitive(self, pos, nor=None, uv=None, idx=None, idx_type=USHORT, mode=None, material=None,
                  uv_stride=None):
        p = {"attributes": {"POSITION": self.accessor(pos, FLOAT, "VEC3")}}
        if nor is not None:
            p["attributes"]["NORMAL"] = self.accessor(nor, FLOAT, "VEC3")
        if uv is not None:
            p["attributes"]["TEXCOORD_0"] = self.accessor(uv, FLOAT, "VEC2", stride=uv_stride)
        if idx is not None:
            p["indices"] = self.accessor(np.asarray(idx).reshape(-1, 1), idx_type, "SCALAR",
                                         pad_to=(len(idx) * np.dtype(_NP[idx_type]).itemsize + 3) & ~3)
        if mode is not None:
            p["mode"] = mode
        if material is not None:
            p["material"] = material
        return p

    def png_image(self, png):
        v = self.view(png)
        return self.add("images", {"bufferView": v, "mimeType": "image/png"})

    def animation(self, channels):
        """channels: (node, path, times, values, interpolation or None)."""
        samplers, chans = [], []
        for node, path, times, values, interp in channels:
            typ = {"translation": "VEC3", "scale": "VEC3", "rotation": "VEC4", "weights": "SCALAR"}[path]
            s = {"input": self.accessor(np.asarray(times).reshape(-1, 1), FLOAT, "SCALAR"),
                 "output": self.accessor(values, FLOAT, typ)}
            if interp:
                s["interpolation"] = interp
            samplers.append(s)
            chans.append({"sampler": len(samplers) - 1, "target": {"node": node, "path": path}})
        return self.add("animations", {"samplers": samplers, "channels": chans})

    def write_gltf(self, path, bin_name):
        self._pad()
        self.j["buffers"] = [{"uri": bin_name, "byteLength": len(self.bin)}]
        with open(path, "w") as f:
            json.dump(self.j, f)
        with open(str(path).rsplit("/", 1)[0] + "/" + bin_name, "wb") as f:
            f.write(bytes(self.bin))

    def write_glb(self, path):
        self._pad()
        self.j["buffers"] = [{"byteLength": len(self.bin)}]
        js = json.dumps(self.j).encode()
        js += b" " * ((4 - len(js) % 4) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js
        body += struct.pack("<II", len(self.bin), 0x004E4942) + bytes(self.bin)
        with open(path, "wb") as f:
            f.write(struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body)


# ----------------------------------------------------------------- PNG ---
_ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
          (0, 1, 1, 2)]
_CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _pack_row(samples, depth):
    """samples: (w, channels) ints -> the row's bytes at `depth` bits."""
    flat = [int(v) for v in np.asarray(samples).reshape(-1)]
    if depth == 8:
        return bytes(flat)
    if depth == 16:
        return b"".join(struct.pack(">H", v) for v in flat)
    out, acc, nbits = bytearray(), 0, 0
    for v in flat:
        acc = (acc << depth) | v
        nbits += depth
        if nbits == 8:
            out.append(acc)
            acc, nbits = 0, 0
    if nbits:
        out.append(acc << (8 - nbits))
    return bytes(out)


def _filter(rows, bpp, first_filter):
    """Apply the five PNG filters in turn (row k uses (first_filter + k) % 5)."""
    out, prev = bytearray(), None
    for k, row in enumerate(rows):
        f = (first_filter + k) % 5
        cur = bytearray(row)
        up = prev if prev is not None else bytearray(len(row))
        res = bytearray(len(row))
        for i in range(len(row)):
            a = cur[i - bpp] if i >= bpp else 0
            b = up[i]
            c = up[i - bpp] if i >= bpp else 0
            if f == 0:
                pred = 0
            elif f == 1:
                pred = a
            elif f == 2:
                pred = b
            elif f == 3:
                pred = (a + b) >> 1
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pred = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            res[i] = (cur[i] - pred) & 255
        out.append(f)
        out += res
        prev = cur
    return bytes(out)


def _chunk(typ, data):
    return struct.pack(">I", len(data)) + typ + data + struct.pack(">I", zlib.crc32(typ + data) & 0xFFFFFFFF)


def png(samples, ctype, depth, interlace=False, plte=None, trns=None, first_filter=0):
    """samples: (h, w, channels) ints at `depth` bits (palette indices for
    colour type 3)."""
    s = np.asarray(samples)
    h, w = s.shape[:2]
    ch = _CHANNELS[ctype]
    s = s.reshape(h, w, ch)
    bpp = max(1, ch * depth // 8)
    raw = b""
    passes = _ADAM7 if interlace else [(0, 0, 1, 1)]
    for (x0, y0, dx, dy) in passes:
        sub = s[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        raw += _filter([_pack_row(r, depth) for r in sub], bpp, first_filter)
    out = b"\x89PNG\r\n\x1a\n"
    out += _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 1 if interlace else 0))
    if plte is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(plte, np.uint8).reshape(-1)))
    if trns is not None:
        out += _chunk(b"tRNS", bytes(trns))
    z = zlib.compress(raw, 9)
    out += _chunk(b"IDAT", z[: len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])   # split IDAT
    out += _chunk(b"IEND", b"")
    return out


# ----------------------------------------------------------------- BMP ---
def bmp_as_read(rgb):
    """rgb: (h, w, 3) uint8, top row first.  The 54-byte header the reader
    consults (width at 18, height at 22), then rows bottom up as b, g, r with
    (w * 3) % 4 bytes after each row."""
    a = np.asarray(rgb, np.uint8)
    h, w = a.shape[:2]
    pad = (w * 3) % 4
    hdr = bytearray(54)
    hdr[0:2] = b"BM"
    struct.pack_into("<iiHH", hdr, 18, w, h, 1, 24)
    body = bytearray()
    for row in a[::-1]:
        body += bytes(row[:, ::-1].reshape(-1)) + b"\xee" * pad
    struct.pack_into("<I", hdr, 2, 54 + len(body))
    struct.pack_into("<I", hdr, 10, 54)
    return bytes(hdr) + bytes(body)
