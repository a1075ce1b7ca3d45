"""JSON state I/O in the reference's serde format.

The reference serialises `Simulation` (src/simulation.rs:49-69) and
`SimulationGrid` (src/grid/mod.rs:85-125) with serde_json; ndarray arrays take
the form {"v": 1, "dim": [NX, NY], "data": [row-major]} and cells are
"Fluid" | {"Boundary": "NoSlip"} | {"Boundary": "Outflow"} |
{"Boundary": {"Inflow": {"velocity": [u, v]}}} (src/cell.rs:7-23).
This module adds {"Boundary": {"MovingWall": {"velocity": [u, v]}}} for the
moving lid, which the reference does not have.

Floats are written with Python's repr, i.e. shortest round-trip decimals, as
serde_json/ryu does, so a state written here reads back bit-for-bit.
"""
import json

import numpy as np

FLUID, NOSLIP, INFLOW, OUTFLOW, MOVING_WALL = 0, 1, 2, 3, 4

EDGE_NAMES = [None, "North", "NorthEast", "East", "SouthEast", "South", "SouthWest",
              "West", "NorthWest"]


def decode_cells(obj):
    """ndarray<Cell> JSON -> (codes uint8 [NX,NY], velocity float64 [NX,NY,2])."""
    nx, ny = obj["dim"]
    codes = np.zeros(nx * ny, dtype=np.uint8)
    vel = np.zeros((nx * ny, 2), dtype=np.float64)
    for i, c in enumerate(obj["data"]):
        if c == "Fluid":
            codes[i] = FLUID
            continue
        b = c["Boundary"]
        if b == "NoSlip":
            codes[i] = NOSLIP
        elif b == "Outflow":
            codes[i] = OUTFLOW
        elif isinstance(b, dict) and "Inflow" in b:
            codes[i] = INFLOW
            vel[i] = b["Inflow"]["velocity"]
        elif isinstance(b, dict) and "MovingWall" in b:
            codes[i] = MOVING_WALL
            vel[i] = b["MovingWall"]["velocity"]
        else:
            raise ValueError("unknown cell %r" % (c,))
    return codes.reshape(nx, ny), vel.reshape(nx, ny, 2)


def encode_cells(codes, vel):
    out = []
    for c, (vu, vv) in zip(codes.reshape(-1), vel.reshape(-1, 2)):
        c = int(c)
        if c == FLUID:
            out.append("Fluid")
        elif c == NOSLIP:
            out.append({"Boundary": "NoSlip"})
        elif c == OUTFLOW:
            out.append({"Boundary": "Outflow"})
        elif c == INFLOW:
            out.append({"Boundary": {"Inflow": {"velocity": [float(vu), float(vv)]}}})
        elif c == MOVING_WALL:
            out.append({"Boundary": {"MovingWall": {"velocity": [float(vu), float(vv)]}}})
        else:
            raise ValueError("unknown cell code %d" % c)
    return {"v": 1, "dim": list(codes.shape), "data": out}


def decode_array(obj):
    nx, ny = obj["dim"]
    return np.array(obj["data"], dtype=np.float64).reshape(nx, ny)


def encode_array(a):
    return {"v": 1, "dim": list(a.shape), "data": [float(x) for x in a.reshape(-1)]}


def dumps(obj):
    return json.dumps(obj, indent=2)


_U64_MAX = (1 << 64) - 1
_POW10 = [float("1e%d" % k) for k in range(309)]


def _f64_from_parts(positive, significand, exponent):
    f = float(significand)          # u64 -> f64, round half to even, as Rust's `as f64`
    while True:
        k = abs(exponent)
        if k < len(_POW10):
            f = f * _POW10[k] if exponent >= 0 else f / _POW10[k]
            break
        if f == 0.0:
            break
        if exponent >= 0:
            raise ValueError("number out of range")
        f /= 1e308
        exponent += 308
    return f if positive else -f


def serde_float(text):
    """Parse a JSON number the way serde_json 1.0.140 does without its
    `float_roundtrip` feature (the reference's Cargo.toml:18 enables none):
    digits accumulate into a u64 significand until the next one would
    overflow (later digits are dropped, integer-part digits still count in the
    exponent), then f = significand as f64, multiplied or divided once by the
    f64 power of ten.  This is not always correctly rounded: for 17-digit
    decimals it can differ from strtod by one ulp, and the reference's runs
    see the value computed here."""
    s = text
    i = 0
    positive = True
    if s[i] == "-":
        positive = False
        i += 1
    significand = 0
    exponent = 0
    overflowed = False
    n = len(s)
    while i < n and s[i].isdigit():
        d = ord(s[i]) - 48
        if overflowed or significand * 10 + d > _U64_MAX:
            overflowed = True
            exponent += 1
        else:
            significand = significand * 10 + d
        i += 1
    if i < n and s[i] == ".":
        i += 1
        while i < n and s[i].isdigit():
            d = ord(s[i]) - 48
            if not overflowed and significand * 10 + d <= _U64_MAX:
                significand = significand * 10 + d
                exponent -= 1
            else:
                overflowed = True
            i += 1
    if i < n and s[i] in "eE":
        i += 1
        esign = 1
        if s[i] in "+-":
            esign = -1 if s[i] == "-" else 1
            i += 1
        e = int(s[i:])
        exponent += esign * e
    return _f64_from_parts(positive, significand, exponent)


def load_reference_json(fh_or_text, exact=False):
    """Read JSON written for the reference (e.g. a `--sim-file`), parsing
    floats as the reference's serde_json does (exact=True: correctly rounded,
    so shortest round-trip decimals read back bit for bit)."""
    text = fh_or_text if isinstance(fh_or_text, str) else fh_or_text.read()
    return json.loads(text, parse_float=float if exact else serde_float)


# ------------------------------------------------------------ NaSt2D output

NAST2D_FLUID = 16   # NaSt2D's C_F flag; boundary cells carry only neighbour bits


def read_nast2d_out(data, int_bytes=4, int_type="i"):
    """A NaSt2D .out file (python/generate_test_data.py:97-154): two C ints
    imax, jmax, then U, V, P, T as (imax+2) x (jmax+2) C doubles each, x
    outer, then the flags as C ints.  Returns a dict of imax, jmax, the four
    float64 arrays [x][y] and the flags reduced to FLUID (16) or BOUNDARY (0),
    as the reference's Flag(IntFlag, boundary=CONFORM) keeps only the fluid
    bit.  `data` is the file's bytes or a binary file object."""
    raw = data if isinstance(data, (bytes, bytearray)) else data.read()
    # struct's int codes ("i" signed, "I" unsigned) at the given width, little-endian
    ity = np.dtype("<%s%d" % ("u" if int_type.isupper() else "i", int_bytes))
    imax, jmax = (int(v) for v in np.frombuffer(raw, ity, 2, 0))
    shape = (imax + 2, jmax + 2)
    n = shape[0] * shape[1]
    off = 2 * int_bytes
    fields = []
    for _ in range(4):
        fields.append(np.frombuffer(raw, "<f8", n, off).reshape(shape).copy())
        off += 8 * n
    flags = np.frombuffer(raw, ity, n, off).reshape(shape)
    flags = np.where(flags & NAST2D_FLUID, NAST2D_FLUID, 0).astype(np.int64)
    U, V, P, T = fields
    return {"imax": imax, "jmax": jmax, "U": U, "V": V, "P": P, "T": T, "flags": flags}


def nast2d_grid(out):
    """The reference's SimulationGrid JSON for a NaSt2D output
    (python/generate_test_data.py:42-76): NaSt2D does not record boundary
    kinds, so its presets are assumed -- west wall Inflow with the cell's own
    (U, V), east wall Outflow, every other boundary cell NoSlip (walls'
    corners included)."""
    imax, jmax = out["imax"], out["jmax"]
    size = [imax + 2, jmax + 2]
    U, V, flags = out["U"], out["V"], out["flags"]
    cells = []
    for x in range(size[0]):
        for y in range(size[1]):
            if flags[x, y] == NAST2D_FLUID:
                cells.append("Fluid")
            elif 0 < y <= jmax and x == 0:
                cells.append({"Boundary": {"Inflow": {"velocity": [float(U[x, y]),
                                                                   float(V[x, y])]}}})
            elif 0 < y <= jmax and x == imax + 1:
                cells.append({"Boundary": "Outflow"})
            else:
                cells.append({"Boundary": "NoSlip"})

    def flat(a):
        return {"v": 1, "dim": size, "data": [float(t) for t in a.reshape(-1)]}
    return {"size": size, "u": flat(U), "v": flat(V), "pressure": flat(out["P"]),
            "cell_type": {"v": 1, "dim": size, "data": cells}}
