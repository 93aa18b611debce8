"""Body CSV format of the reference (cmd/sim/fromcsv.go:15-112).

Columns: x, y, z, vx, vy, vz, mass, radius[, is_sun, collision_behavior, color,
fragmentation_factor, fragmentation_step].  '#' starts a comment line.  A row whose parse fails
is skipped silently (fromcsv.go:67-74), including rows with an empty optional column: is_sun goes
through strconv.ParseBool (fromcsv.go:127-133), which rejects "".  At most `body_count` rows are
read (fromcsv.go:59; the server's -b flag, default 1000).
"""
import csv
import io
import re

import numpy as np

from . import _lib as L
from .sim import Bodies

COLLISION_NAMES = ["none", "subsume", "elastic", "fragment"]  # globals.go:41-48
COLOR_NAMES = ["random", "black", "white", "darkgray", "gray", "lightgray", "red", "green",
               "blue", "yellow", "magenta", "cyan", "orange", "brown", "pink"]  # globals.go:65-73

_DEC = re.compile(r"^[+-]?((\d+(\.\d*)?)|(\.\d+))([eE][+-]?\d+)?$")
_SPECIAL = re.compile(r"^[+-]?(inf|infinity|nan)$", re.IGNORECASE)
_HEX = re.compile(r"^[+-]?0[xX]([0-9a-fA-F]+(\.[0-9a-fA-F]*)?|\.[0-9a-fA-F]+)[pP][+-]?\d+$")


def go_parse_float(s):
    """strconv.ParseFloat(s, 64) for the forms a body CSV uses; raises ValueError otherwise,
    including for a finite literal beyond the FP64 range (ParseFloat's ErrRange)."""
    if _SPECIAL.match(s):
        return float(s)
    if _DEC.match(s):
        v = float(s)
    elif _HEX.match(s):
        try:
            v = float.fromhex(s)
        except OverflowError:
            raise ValueError("out of range %r" % s) from None
    else:
        raise ValueError("invalid float %r" % s)
    if v in (float("inf"), float("-inf")):
        raise ValueError("out of range %r" % s)
    return v


def go_parse_bool(s):
    """strconv.ParseBool: 1 t T TRUE true True / 0 f F FALSE false False."""
    if s in ("1", "t", "T", "TRUE", "true", "True"):
        return True
    if s in ("0", "f", "F", "FALSE", "false", "False"):
        return False
    raise ValueError("invalid bool %r" % s)


def parse_collision(s):
    """globals.ParseCollisionBehavior: unknown names map to Elastic."""
    s = s.lower()
    return COLLISION_NAMES.index(s) if s in COLLISION_NAMES else L.NB_CB_ELASTIC


def parse_color(s):
    s = s.lower()
    return COLOR_NAMES.index(s) if s in COLOR_NAMES else 0


def read_csv(path=None, text=None, body_count=1000, default_collision=L.NB_CB_ELASTIC):
    """Returns (Bodies, colors, suns) for the first body_count parseable rows of the file at
    `path` (or of the string `text`)."""
    f = io.StringIO(text) if text is not None else open(path, newline="")
    rows = []
    with f:
        for rec in csv.reader(f):
            if len(rows) >= body_count:
                break
            if not rec or (rec[0].startswith("#")):
                continue
            try:
                v = [go_parse_float(rec[k].strip()) for k in range(8)]
                sun = go_parse_bool(rec[8].strip()) if len(rec) >= 9 else False
                cb = parse_collision(rec[9].strip()) if len(rec) >= 10 else default_collision
                color = parse_color(rec[10].strip()) if len(rec) >= 11 else 0
                ff = go_parse_float(rec[11].strip()) if len(rec) >= 12 else 0.0
                fs = go_parse_float(rec[12].strip()) if len(rec) >= 13 else 0.0
            except (ValueError, IndexError):
                continue
            rows.append((v, sun, cb, color, ff, fs))
    n = len(rows)
    a = np.array([r[0] for r in rows], np.float64).reshape(n, 8)
    flags = np.full(n, L.NB_F_EXISTS, np.uint8)
    suns = np.array([r[1] for r in rows], bool)
    flags[suns] |= L.NB_F_SUN
    b = Bodies(a[:, 0], a[:, 1], a[:, 2], a[:, 3], a[:, 4], a[:, 5], a[:, 6], a[:, 7],
               beh=np.array([r[2] for r in rows], np.int32), flags=flags,
               frag=np.array([r[4] for r in rows], np.float64),
               fstep=np.array([r[5] for r in rows], np.float64))
    colors = np.array([r[3] for r in rows], np.int32)
    return b, colors, suns


def write_csv(b: Bodies, path=None, colors=None):
    """Writes every body as x,y,z,vx,vy,vz,mass,radius,is_sun,collision,color,frag_factor,
    frag_step (repr() round-trips each float exactly)."""
    out = io.StringIO()
    for i in range(b.n):
        sun = "true" if (b.flags[i] & L.NB_F_SUN) else "false"
        cb = COLLISION_NAMES[int(b.beh[i])] if 0 <= b.beh[i] < 4 else "elastic"
        col = COLOR_NAMES[int(colors[i])] if colors is not None else "random"
        vals = [b.x[i], b.y[i], b.z[i], b.vx[i], b.vy[i], b.vz[i], b.m[i], b.r[i]]
        out.write(",".join(repr(float(v)) for v in vals))
        out.write(",%s,%s,%s,%r,%r\n" % (sun, cb, col, float(b.frag[i]), float(b.fstep[i])))
    s = out.getvalue()
    if path:
        with open(path, "w") as f:
            f.write(s)
    return s
