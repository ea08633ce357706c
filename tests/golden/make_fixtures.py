"""Copies the reference's input data into tests/golden/data and records what
independent decoders read from it (tests/golden/io_expected.json).

The reference ships inputs only (data/dem.tif, data/mask_500.tif,
data/samples.*, neighbor/*.nbr) and no expected outputs.  The expected
header fields and pixel checksums here come from Pillow's TIFF decoder and
from a plain struct/DBF parse written for this script, so the library's own
codecs are checked against code that shares nothing with them.

Run once, in a checkout next to the reference:
    python tests/golden/make_fixtures.py /path/to/reference
"""
import hashlib
import json
import os
import shutil
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "data")
FILES = [("data/dem.tif", "dem.tif"), ("data/mask_500.tif", "mask_500.tif"), ("data/samples.shp", "samples.shp"),
         ("data/samples.shx", "samples.shx"), ("data/samples.dbf", "samples.dbf"), ("data/samples.prj", "samples.prj"),
         ("neighbor/moore.nbr", "moore.nbr"), ("neighbor/local.nbr", "local.nbr"),
         ("neighbor/moore20.nbr", "moore20.nbr")]


def tiff_expect(path):
    from PIL import Image  # an independent TIFF decoder
    im = Image.open(path)
    a = np.asarray(im).astype(np.float64)
    tags = im.tag_v2
    scale, tie = tags.get(33550), tags.get(33922)
    nodata = tags.get(42113)
    return {"rows": a.shape[0], "cols": a.shape[1],
            "transform": [tie[3], scale[0], 0.0, tie[4], 0.0, -scale[1]],
            "nodata": float(nodata.strip("\x00")) if nodata else None,
            "sha256_f64": hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest(),
            "sum": float(a.sum())}


def shp_expect(base, field_index):
    shp = open(base + ".shp", "rb").read()
    dbf = open(base + ".dbf", "rb").read()
    nrec, hlen, rlen = struct.unpack("<IHH", dbf[4:12])
    fields, o, off = [], 32, 1
    while dbf[o] != 0x0D:
        fields.append((dbf[o:o + 11].split(b"\0")[0].decode(), off, dbf[o + 16]))
        off += dbf[o + 16]
        o += 32
    xs, ys, vs = [], [], []
    p = 100
    for r in range(nrec):
        clen = struct.unpack(">i", shp[p + 4:p + 8])[0] * 2
        st, x, y = struct.unpack("<idd", shp[p + 8:p + 28])
        p += 8 + clen
        rec = dbf[hlen + r * rlen:hlen + (r + 1) * rlen]
        if rec[:1] == b"*":
            continue
        assert st in (1, 11, 21)
        _, fo, fl = fields[field_index]
        xs.append(x)
        ys.append(y)
        vs.append(float(rec[fo:fo + fl]))
    arr = np.array([xs, ys, vs], dtype=np.float64)
    return {"n": len(xs), "field": fields[field_index][0], "fields": len(fields),
            "sha256_xyv": hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest(),
            "bbox": [min(xs), min(ys), max(xs), max(ys)], "value_sum": float(sum(vs))}


def main(ref):
    os.makedirs(DATA, exist_ok=True)
    for src, dst in FILES:
        shutil.copyfile(os.path.join(ref, src), os.path.join(DATA, dst))
    exp = {"dem.tif": tiff_expect(os.path.join(DATA, "dem.tif")),
           "mask_500.tif": tiff_expect(os.path.join(DATA, "mask_500.tif")),
           "samples.shp": shp_expect(os.path.join(DATA, "samples"), 3)}
    with open(os.path.join(HERE, "io_expected.json"), "w") as f:
        json.dump(exp, f, indent=1, sort_keys=True)
    print(json.dumps(exp, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
