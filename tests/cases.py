"""Shared inputs of the test suite.

GEOMS / VALUES are the reference's test geometries (python/test/test_many.py:19-28);
EXPLODED is GDF.explode().explode() of them (test_many.py:38, 329-346): multi
parts and collection members become separate geometries, keeping the value of
their parent, in member order.
"""
import numpy as np

GEOMS = [
    "POLYGON ((-180 -20, -140 55, 10 0, -140 -60, -180 -20), (-150 -20, -100 -10, -110 20, -150 -20))",
    "POLYGON ((-10 0, 140 60, 160 0, 140 -55, -10 0))",
    "POLYGON ((-125 0, 0 60, 40 5, 15 -45, -125 0))",
    "MULTILINESTRING ((-180 -70, -140 -50), (-140 -50, -100 -70), (-100 -70, -60 -50), (-60 -50, -20 -70), (-20 -70, 20 -50), (20 -50, 60 -70), (60 -70, 100 -50), (100 -50, 140 -70), (140 -70, 180 -50))",
    "GEOMETRYCOLLECTION (POINT (50 -40), POLYGON ((75 -40, 75 -30, 100 -30, 100 -40, 75 -40)), LINESTRING (60 -40, 80 0), GEOMETRYCOLLECTION (POLYGON ((100 20, 100 30, 110 30, 110 20, 100 20))))",
]
VALUES = np.arange(1, 6)

_LINES = ["-180 -70, -140 -50", "-140 -50, -100 -70", "-100 -70, -60 -50", "-60 -50, -20 -70",
          "-20 -70, 20 -50", "20 -50, 60 -70", "60 -70, 100 -50", "100 -50, 140 -70", "140 -70, 180 -50"]
EXPLODED = (
    [(GEOMS[0], 1), (GEOMS[1], 2), (GEOMS[2], 3)]
    + [(f"LINESTRING ({s})", 4) for s in _LINES]
    + [("POINT (50 -40)", 5), ("POLYGON ((75 -40, 75 -30, 100 -30, 100 -40, 75 -40))", 5),
       ("LINESTRING (60 -40, 80 0)", 5), ("POLYGON ((100 20, 100 30, 110 30, 110 20, 100 20))", 5)]
)
EXPLODED_WKT = [w for w, _ in EXPLODED]
EXPLODED_VALUES = np.array([v for _, v in EXPLODED])


def sq(x0, y0, x1, y1):
    """The `sq` helper of R/rusterize/tests/testthat/helper-geoms.R:33-35."""
    return f"POLYGON (({x0} {y0}, {x1} {y0}, {x1} {y1}, {x0} {y1}, {x0} {y0}))"


# value ranges that keep every dtype's field exactly representable
FIELD_RANGE = {"uint8": (0, 256), "uint16": (0, 65536), "uint32": (0, 1 << 20), "uint64": (0, 1 << 20),
               "int8": (-128, 128), "int16": (-32768, 32768), "int32": (-(1 << 20), 1 << 20),
               "int64": (-(1 << 20), 1 << 20), "float32": (-1000, 1000), "float64": (-1000, 1000)}
