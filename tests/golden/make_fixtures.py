"""Writes the golden rasters of the reference's Python test suite as .npy.

Source data (read-only reference tree, MIT licence, (c) the rusterize
authors): python/test/data/standard_output_sum.tif and
python/test/data/standard_output_sum_custom_shape.tif, produced by
python/test/test_many.py:228-235 and :329-346.  Only the pixel arrays are
kept (uint8, row 0 = north); the GeoTIFF tags are recorded in fixtures.json.
Run from the repository root: python tests/golden/make_fixtures.py
"""
import hashlib
import json
from pathlib import Path

import numpy as np
from PIL import Image

SRC = Path("/root/reference/python/test/data")
DST = Path(__file__).resolve().parent

meta = {}
for name in ["standard_output_sum", "standard_output_sum_custom_shape"]:
    im = Image.open(SRC / f"{name}.tif")
    arr = np.array(im)
    np.save(DST / f"{name}.npy", arr, allow_pickle=False)
    meta[name] = {
        "shape": list(arr.shape),
        "dtype": str(arr.dtype),
        "sum": int(arr.sum(dtype=np.int64)),
        "nonzero": int((arr > 0).sum()),
        "sha256": hashlib.sha256(arr.tobytes()).hexdigest(),
        "tiepoint": [float(v) for v in im.tag_v2.get(33922)],
        "pixel_scale": [float(v) for v in im.tag_v2.get(33550)],
    }
(DST / "fixtures.json").write_text(json.dumps(meta, indent=2) + "\n")
print(json.dumps(meta, indent=2))
