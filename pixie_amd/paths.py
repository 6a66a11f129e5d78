"""Path construction, mirroring pixie's Path builder API (paths.nim:24-652).

A Path is pixie's flat float32 command list: the command kind (paths.nim:18-22
ordinals) followed by its parameters.  The builders here append exactly the
commands pixie's builders append, computing coordinates in float32 (libm's
float32 sinf/cosf where pixie calls sin/cos on float32).  String paths are
parsed by the native library (pxh_parse_path), which restates parsePath
(paths.nim:119-262).
"""
import ctypes
import ctypes.util
import math

import numpy as np

f32 = np.float32

CLOSE, MOVE, LINE, HLINE, VLINE, CUBIC, SCUBIC, QUAD, TQUAD, ARC = range(10)
RMOVE, RLINE, RHLINE, RVLINE, RCUBIC, RSCUBIC, RQUAD, RTQUAD, RARC = range(10, 19)

PARAM_COUNT = {CLOSE: 0, MOVE: 2, LINE: 2, RMOVE: 2, RLINE: 2, TQUAD: 2, RTQUAD: 2,
               HLINE: 1, VLINE: 1, RHLINE: 1, RVLINE: 1, CUBIC: 6, RCUBIC: 6,
               SCUBIC: 4, RSCUBIC: 4, QUAD: 4, RQUAD: 4, ARC: 7, RARC: 7}

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
for _n in ("sinf", "cosf", "tanf", "acosf", "sqrtf"):
    getattr(_libm, _n).restype = ctypes.c_float
    getattr(_libm, _n).argtypes = [ctypes.c_float]

EPSILON = f32(0.0001 * math.pi)         # paths.nim:44
PI32 = f32(math.pi)
TAU32 = f32(2 * math.pi)
SPLINE_CIRCLE_K = 4.0 * (-1.0 + math.sqrt(2.0)) / 3  # paths.nim:529 (float64)


def sinf(x):
    return f32(_libm.sinf(float(x)))


def cosf(x):
    return f32(_libm.cosf(float(x)))


# ---- Mat3 (vmath layout: m[3*col + row]; translation in m[6], m[7]) ----

def mat3():
    return np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)


def translate(x, y):
    m = mat3()
    m[6], m[7] = x, y
    return m


def scale(x, y=None):
    m = mat3()
    m[0], m[4] = x, (x if y is None else y)
    return m


def mat_mul(a, b):
    """a * b for vmath column-major Mat3 (float32)."""
    A = np.asarray(a, np.float64).reshape(3, 3).T
    B = np.asarray(b, np.float64).reshape(3, 3).T
    return (A @ B).T.reshape(9).astype(np.float32)


class Path:
    """pixie's Path: flat float32 commands plus start/at bookkeeping."""

    def __init__(self, commands=None):
        self.commands = [] if commands is None else [f32(c) for c in commands]
        self.start = (f32(0), f32(0))
        self.at = (f32(0), f32(0))

    def array(self):
        return np.asarray(self.commands, np.float32)

    def __len__(self):
        return len(self.commands)

    def copy(self):
        p = Path(self.commands)
        p.start, p.at = self.start, self.at
        return p

    def add_path(self, other):
        self.commands.extend(other.commands)

    def close_path(self):
        self.commands.append(f32(CLOSE))
        self.at = self.start

    def move_to(self, x, y):
        x, y = f32(x), f32(y)
        self.commands += [f32(MOVE), x, y]
        self.start = (x, y)
        self.at = self.start

    def line_to(self, x, y):
        x, y = f32(x), f32(y)
        self.commands += [f32(LINE), x, y]
        self.at = (x, y)

    def bezier_curve_to(self, x1, y1, x2, y2, x3, y3):
        self.commands += [f32(CUBIC)] + [f32(v) for v in (x1, y1, x2, y2, x3, y3)]
        self.at = (f32(x3), f32(y3))

    def quadratic_curve_to(self, x1, y1, x2, y2):
        self.commands += [f32(QUAD)] + [f32(v) for v in (x1, y1, x2, y2)]
        self.at = (f32(x2), f32(y2))

    def elliptical_arc_to(self, rx, ry, rotation, large, sweep, x, y):
        self.commands += [f32(ARC), f32(rx), f32(ry), f32(rotation), f32(1 if large else 0),
                          f32(1 if sweep else 0), f32(x), f32(y)]
        self.at = (f32(x), f32(y))

    def rect(self, x, y, w, h, clockwise=True):
        """paths.nim:506-521"""
        x, y, w, h = f32(x), f32(y), f32(w), f32(h)
        self.move_to(x, y)
        if clockwise:
            self.line_to(x + w, y)
            self.line_to(x + w, y + h)
            self.line_to(x, y + h)
        else:
            self.line_to(x, y + h)
            self.line_to(x + w, y + h)
            self.line_to(x + w, y)
        self.close_path()

    def ellipse(self, cx, cy, rx, ry):
        """paths.nim:608-619 (splineCircleK is float64, offsets narrowed)."""
        cx, cy, rx, ry = f32(cx), f32(cy), f32(rx), f32(ry)
        mx = SPLINE_CIRCLE_K * float(rx)
        my = SPLINE_CIRCLE_K * float(ry)
        n = lambda v: f32(v)
        self.move_to(cx + rx, cy)
        self.bezier_curve_to(cx + rx, n(float(cy) + my), n(float(cx) + mx), cy + ry, cx, cy + ry)
        self.bezier_curve_to(n(float(cx) - mx), cy + ry, cx - rx, n(float(cy) + my), cx - rx, cy)
        self.bezier_curve_to(cx - rx, n(float(cy) - my), n(float(cx) - mx), cy - ry, cx, cy - ry)
        self.bezier_curve_to(n(float(cx) + mx), cy - ry, cx + rx, n(float(cy) - my), cx + rx, cy)
        self.close_path()

    def circle(self, cx, cy, r):
        self.ellipse(cx, cy, r, r)

    def polygon(self, x, y, size, sides):
        """paths.nim:633-646"""
        if sides <= 2:
            raise ValueError("Invalid polygon sides value")
        x, y, size = f32(x), f32(y), f32(size)
        self.move_to(f32(float(x) + float(size) * math.sin(0.0)),
                     f32(float(y) - float(size) * math.cos(0.0)))
        for side in range(1, sides):
            # float32 * PI promotes to float64; sin/cos and the sums are
            # float64, narrowed at the lineTo call.
            a = float(f32(side)) * 2.0 * math.pi / float(f32(sides))
            self.line_to(f32(float(x) + float(size) * math.sin(a)),
                         f32(float(y) - float(size) * math.cos(a)))
        self.close_path()

    def rounded_rect(self, x, y, w, h, nw, ne, se, sw, clockwise=True):
        """paths.nim:533-598"""
        x, y, w, h = f32(x), f32(y), f32(w), f32(h)
        mr = min(w / f32(2), h / f32(2))
        nw, ne, se, sw = [max(f32(0), min(f32(v), mr)) for v in (nw, ne, se, sw)]
        if nw == 0 and ne == 0 and se == 0 and sw == 0:
            self.rect(x, y, w, h, clockwise)
            return
        s = SPLINE_CIRCLE_K
        n = lambda v: f32(v)
        t1 = (x + nw, y); t2 = (x + w - ne, y)
        r1 = (x + w, y + ne); r2 = (x + w, y + h - se)
        b1 = (x + w - se, y + h); b2 = (x + sw, y + h)
        l1 = (x, y + h - sw); l2 = (x, y + nw)
        t1h = (n(float(t1[0]) - float(nw) * s), t1[1]); t2h = (n(float(t2[0]) + float(ne) * s), t2[1])
        r1h = (r1[0], n(float(r1[1]) - float(ne) * s)); r2h = (r2[0], n(float(r2[1]) + float(se) * s))
        b1h = (n(float(b1[0]) + float(se) * s), b1[1]); b2h = (n(float(b2[0]) - float(sw) * s), b2[1])
        l1h = (l1[0], n(float(l1[1]) + float(sw) * s)); l2h = (l2[0], n(float(l2[1]) - float(nw) * s))
        if clockwise:
            self.move_to(*t1); self.line_to(*t2); self.bezier_curve_to(*t2h, *r1h, *r1)
            self.line_to(*r2); self.bezier_curve_to(*r2h, *b1h, *b1)
            self.line_to(*b2); self.bezier_curve_to(*b2h, *l1h, *l1)
            self.line_to(*l2); self.bezier_curve_to(*l2h, *t1h, *t1)
        else:
            self.move_to(*t1); self.bezier_curve_to(*t1h, *l2h, *l2)
            self.line_to(*l1); self.bezier_curve_to(*l1h, *b2h, *b2)
            self.line_to(*b1); self.bezier_curve_to(*b1h, *r2h, *r2)
            self.line_to(*r1); self.bezier_curve_to(*r1h, *t2h, *t2)
            self.line_to(*t1)
        self.close_path()

    def arc(self, x, y, r, a0, a1, ccw=False):
        """paths.nim:416-453"""
        x, y, r, a0, a1 = f32(x), f32(y), f32(r), f32(a0), f32(a1)
        if r == 0:
            return
        if r < 0:
            raise ValueError("Invalid arc, negative radius: %s" % r)
        dx, dy = r * cosf(a0), r * sinf(a0)
        x0, y0 = x + dx, y + dy
        cw = not ccw
        if len(self.commands) == 0:
            self.move_to(x0, y0)
        elif abs(self.at[0] - x0) > EPSILON or abs(self.at[1] - y0) > EPSILON:
            self.line_to(x0, y0)
        angle = (a0 - a1) if ccw else (a1 - a0)
        if angle < 0:
            angle = f32(math.fmod(angle, TAU32)) + TAU32
        if angle > TAU32 - EPSILON:
            self.elliptical_arc_to(r, r, 0, True, cw, x - dx, y - dy)
            self.at = (x0, y0)
            self.elliptical_arc_to(r, r, 0, True, cw, x0, y0)
        elif angle > EPSILON:
            ax, ay = x + r * cosf(a1), y + r * sinf(a1)
            self.at = (ax, ay)
            self.elliptical_arc_to(r, r, 0, angle >= PI32, cw, ax, ay)


def parse_command_count(cmds):
    """Walk a command array and return the number of commands (validates)."""
    i, n = 0, 0
    cmds = list(cmds)
    while i < len(cmds):
        k = int(cmds[i])
        if k not in PARAM_COUNT:
            raise ValueError("bad command kind %r" % k)
        i += 1 + PARAM_COUNT[k]
        n += 1
    if i != len(cmds):
        raise ValueError("truncated command array")
    return n
