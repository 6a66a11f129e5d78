"""Oracle for SVG ingestion: pixie's parseSvg and newImage(svg)
(fileformats/svg.nim) restated in Python.

TEST INFRASTRUCTURE ONLY.  Only tests/ may import this module, and only as
the checker.  The XML is read by Python's expat-based ElementTree (an
independent parser), with the attribute-value whitespace of Nim's parsexml
applied on top (each run of blanks read as one space, a trailing run
dropped); CSS colour names come from Pillow's ImageColor table; paths are
parsed by the oracle's own parsePath and drawn by liboracle.  Everything
else restates svg.nim, each function citing the lines it follows, with
float64 values narrowed to float32 where pixie narrows them.
"""
import ctypes
import ctypes.util
import math
import re
import xml.etree.ElementTree as ET

import numpy as np

f32 = np.float32
NONZERO, EVENODD = 0, 1
NORMAL, OVERWRITE = 0, 17

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
for _n in ("sinf", "cosf"):
    getattr(_libm, _n).restype = ctypes.c_float
    getattr(_libm, _n).argtypes = [ctypes.c_float]


class SvgError(ValueError):
    pass


# ---- numbers and strings (Nim's strutils) ----

_FLOAT = re.compile(r"[+-]?(?:[0-9][0-9_]*)?(?:\.[0-9_]*)?(?:[eE][+-]?[0-9_]+)?")


def parse_float(s):
    """strutils.parseFloat: the whole string, '_' between digits allowed."""
    t = s.lower().lstrip("+-")
    if t in ("inf", "nan"):
        return float(s.lower().replace("+", ""))
    m = _FLOAT.fullmatch(s)
    mant = s.split("e")[0].split("E")[0]
    if not s or not m or not any(c.isdigit() for c in mant) or s.lstrip("+-").startswith("_"):
        raise SvgError("invalid float: " + s)
    return float(s.replace("_", ""))


def parse_int(s):
    if not re.fullmatch(r"[+-]?[0-9][0-9_]*", s):
        raise SvgError("invalid integer: " + s)
    return int(s.replace("_", ""))


def split_args(s):  # svg.nim:57-62
    return [e for e in s.replace(",", " ").split(" ") if e]


def _strip(s):
    return s.strip(" \t\r\n\v\f")


# ---- colours (chroma's parseHtmlColor, ColorRGBX conversions) ----

def parse_color(text):
    from PIL import ImageColor
    s = _strip(text)
    b = lambda v: float(f32(v) / f32(255))  # noqa: E731
    if not s:
        raise SvgError("Invalid color " + text)
    try:
        if s[0] == "#":
            h = s[1:]
            if len(h) == 3:
                return tuple(float(f32(int(c, 16)) / f32(15)) for c in h) + (1.0,)
            if len(h) == 6:
                return (b(int(h[0:2], 16)), b(int(h[2:4], 16)), b(int(h[4:6], 16)), 1.0)
            if len(h) == 8:
                return tuple(b(int(h[k:k + 2], 16)) for k in (0, 2, 4, 6))
            raise SvgError("Invalid color " + text)
        if s.startswith("rgba(") or s.startswith("rgb("):
            alpha = s[3] == "a"
            f = s[5 if alpha else 4:-1].split(",")
            if s[-1] != ")" or len(f) != (4 if alpha else 3):
                raise SvgError("Invalid color " + text)
            c = [b(parse_int(_strip(v))) for v in f[:3]] + [1.0]
            if alpha:
                c[3] = float(f32(parse_float(_strip(f[3]))))
            return tuple(c)
        hexv = ImageColor.colormap.get(s.lower())
        if hexv is None or not hexv.startswith("#") or len(hexv) != 7:
            raise SvgError("Invalid color " + text)
        return parse_color(hexv)
    except ValueError as e:
        raise SvgError(str(e))


def color_rgbx(c):
    """Color -> ColorRGBX: round(clamp(v) * 255) then (c * a + 127) // 255."""
    b = []
    for v in c:
        v = f32(min(max(f32(v), f32(0)), f32(1)))
        b.append(int(math.floor(float(f32(v * f32(255))) + 0.5)))
    if b[3] != 255:
        b = [(x * b[3] + 127) // 255 for x in b[:3]] + [b[3]]
    return tuple(b)


def rgbx_color(x):
    """ColorRGBX -> Color (straight bytes over 255)."""
    r, g, b, a = x
    if a not in (0, 255):
        mul = int(f32(f32(f32(255) / f32(a)) * f32(255)))
        r, g, b = [(v * mul + 127) // 255 for v in (r, g, b)]
    return tuple(float(f32(v) / f32(255)) for v in (r, g, b, a))


# ---- matrices (vmath, m[3 * col + row]) ----

def m3_mul(a, b):
    r = [f32(0)] * 9
    for i in range(3):
        for j in range(3):
            r[3 * i + j] = f32(f32(f32(f32(b[3 * i]) * f32(a[j])) + f32(f32(b[3 * i + 1]) * f32(a[3 + j])))
                               + f32(f32(b[3 * i + 2]) * f32(a[6 + j])))
    return r


def m3_apply(m, x, y):
    x, y = f32(x), f32(y)
    return (f32(f32(f32(m[0] * x) + f32(m[3] * y)) + m[6]),
            f32(f32(f32(m[1] * x) + f32(m[4] * y)) + m[7]))


IDENTITY = [f32(v) for v in (1, 0, 0, 0, 1, 0, 0, 0, 1)]


def parse_transform(t, m):  # svg.nim:236-293
    rem = t
    while rem:
        idx = rem.find(")")
        if idx == -1:
            raise SvgError("Unsupported SVG transform: " + t)
        f, rem = _strip(rem[:idx + 1]), rem[idx + 1:]
        if f.startswith("matrix("):
            a = split_args(f[7:-1])
            if len(a) != 6:
                raise SvgError("Unsupported SVG transform: " + t)
            v = [f32(parse_float(x)) for x in a]
            m = m3_mul(m, [v[0], v[1], 0, v[2], v[3], 0, v[4], v[5], 1])
        elif f.startswith("translate("):
            c = split_args(f[10:-1])
            tx = parse_float(c[0])
            ty = 0.0 if len(c) == 1 else parse_float(c[1])
            m = m3_mul(m, [1, 0, 0, 0, 1, 0, f32(tx), f32(ty), 1])
        elif f.startswith("rotate("):
            v = split_args(f[7:-1])
            angle = f32(parse_float(v[0]) * math.pi / 180)
            cx = f32(parse_float(v[1])) if len(v) > 1 else f32(0)
            cy = f32(parse_float(v[2])) if len(v) > 2 else f32(0)
            s, c = f32(_libm.sinf(float(angle))), f32(_libm.cosf(float(angle)))
            m = m3_mul(m, [1, 0, 0, 0, 1, 0, cx, cy, 1])
            m = m3_mul(m, [c, s, 0, -s, c, 0, 0, 0, 1])
            m = m3_mul(m, [1, 0, 0, 0, 1, 0, -cx, -cy, 1])
        elif f.startswith("scale("):
            v = split_args(f[6:-1])
            sx = f32(parse_float(v[0]))
            sy = f32(parse_float(v[1])) if len(v) > 1 else sx
            m = m3_mul(m, [sx, 0, 0, 0, sy, 0, 0, 0, 1])
        else:
            raise SvgError("Unsupported SVG transform: " + t)
    return m


# ---- properties (svg.nim:50-234) ----

def initial_props():
    return {"display": True, "fillRule": NONZERO, "fill": "black", "stroke": (0, 0, 0, 0),
            "strokeWidth": f32(1), "lineCap": 0, "lineJoin": 0, "miterLimit": f32(4),
            "dashes": [], "transform": list(IDENTITY), "opacity": f32(1),
            "fillOpacity": f32(1), "strokeOpacity": f32(1)}


STYLE_KEYS = {"fill-rule": "fill-rule", "fill": "fill", "stroke": "stroke",
              "stroke-linecap": "stroke-linecap", "stroke-linejoin": "stroke-linejoin",
              "stroke-width": "stroke-width", "stroke-miterlimit": "stroke-miterlimit",
              "stroke-dasharray": "stroke-dasharray", "display": "display",
              "opacity": "opacity", "fillOpacity": "fill-opacity",
              "strokeOpacity": "stroke-opacity"}


def parse_props(node, inherited):
    r = dict(inherited)
    r["dashes"] = list(inherited["dashes"])
    a = {k: node.attr(k) for k in ("fill-rule", "fill", "stroke", "stroke-width",
                                     "stroke-linecap", "stroke-linejoin", "stroke-miterlimit",
                                     "stroke-dasharray", "transform", "display", "opacity",
                                     "fill-opacity", "stroke-opacity")}
    for pair in node.attr("style").split(";"):
        parts = pair.split(":")
        if len(parts) == 2:
            k = STYLE_KEYS.get(_strip(parts[0]))
            if k is not None and a[k] == "":
                a[k] = _strip(parts[1])
    clamp = lambda v: f32(min(max(v, 0.0), 1.0))  # noqa: E731
    if a["display"]:
        r["display"] = _strip(a["display"]) != "none"
    if a["opacity"]:
        r["opacity"] = clamp(parse_float(a["opacity"]))
    fr = a["fill-rule"]
    if fr == "nonzero":
        r["fillRule"] = NONZERO
    elif fr == "evenodd":
        r["fillRule"] = EVENODD
    elif fr:
        raise SvgError("Invalid fill-rule value " + fr)
    r["fill"] = inherited["fill"] if a["fill"] in ("", "currentColor") else a["fill"]
    st = a["stroke"]
    if st == "currentColor":
        if r["stroke"] == (0, 0, 0, 0):
            r["stroke"] = (0, 0, 0, 255)
    elif st == "none":
        r["stroke"] = (0, 0, 0, 0)
    elif st:
        r["stroke"] = color_rgbx(parse_color(st))
    if a["fill-opacity"]:
        r["fillOpacity"] = clamp(parse_float(a["fill-opacity"]))
    if a["stroke-opacity"]:
        r["strokeOpacity"] = clamp(parse_float(a["stroke-opacity"]))
    sw = a["stroke-width"]
    if sw:
        if sw.endswith("px"):
            sw = sw[:-2]
        r["strokeWidth"] = f32(parse_float(sw))
        if r["stroke"] == (0, 0, 0, 0):
            r["stroke"] = (0, 0, 0, 255)
    caps = {"butt": 0, "round": 1, "square": 2}
    if a["stroke-linecap"] and a["stroke-linecap"] != "inherit":
        if a["stroke-linecap"] not in caps:
            raise SvgError("Invalid stroke-linecap value " + a["stroke-linecap"])
        r["lineCap"] = caps[a["stroke-linecap"]]
    joins = {"miter": 0, "round": 1, "bevel": 2}
    if a["stroke-linejoin"] and a["stroke-linejoin"] != "inherit":
        if a["stroke-linejoin"] not in joins:
            raise SvgError("Invalid stroke-linejoin value " + a["stroke-linejoin"])
        r["lineJoin"] = joins[a["stroke-linejoin"]]
    if a["stroke-miterlimit"]:
        r["miterLimit"] = f32(parse_float(a["stroke-miterlimit"]))
    if a["stroke-dasharray"]:
        r["dashes"] = r["dashes"] + [f32(parse_float(v)) for v in split_args(a["stroke-dasharray"])]
    if a["transform"]:
        r["transform"] = parse_transform(a["transform"], r["transform"])
    return r


# ---- XML ----

class Node:
    def __init__(self, el):
        tag = el.tag
        if tag.startswith("{http://www.w3.org/2000/svg}"):
            tag = tag.split("}", 1)[1]
        self.tag = tag
        # Nim's parsexml: each run of blanks in a value is one space, a
        # trailing run is dropped
        self.attrs = {k: re.sub(" +", " ", v).rstrip(" ") for k, v in el.attrib.items()}
        self.kids = [Node(k) for k in el]

    def attr(self, name):
        return self.attrs.get(name, "")


# ---- elements (svg.nim:295-500) ----

K = 4.0 * (-1.0 + math.sqrt(2.0)) / 3  # splineCircleK, float64


def _ellipse(cx, cy, rx, ry):  # paths.nim:608-619
    cx, cy, rx, ry = f32(cx), f32(cy), f32(rx), f32(ry)
    mx, my = K * float(rx), K * float(ry)
    n = f32
    return [1, f32(cx + rx), cy,
            5, f32(cx + rx), n(float(cy) + my), n(float(cx) + mx), f32(cy + ry), cx, f32(cy + ry),
            5, n(float(cx) - mx), f32(cy + ry), f32(cx - rx), n(float(cy) + my), f32(cx - rx), cy,
            5, f32(cx - rx), n(float(cy) - my), n(float(cx) - mx), f32(cy - ry), cx, f32(cy - ry),
            5, n(float(cx) + mx), f32(cy - ry), f32(cx + rx), n(float(cy) - my), f32(cx + rx), cy,
            0]


def parse_element(orc, node, svg, stack):
    tag = node.tag
    if tag in ("title", "desc", "defs", "radialGradient"):
        return
    num = lambda k, d: parse_float(node.attr(k) or d)  # noqa: E731
    if tag == "g":
        stack.append(parse_props(node, stack[-1]))
        for k in node.kids:
            parse_element(orc, k, svg, stack)
        stack.pop()
    elif tag == "path":
        p = parse_props(node, stack[-1])
        svg["elements"].append((orc.parse_path(node.attr("d")), p))
    elif tag == "line":
        p = parse_props(node, stack[-1])
        x1, y1, x2, y2 = num("x1", "0"), num("y1", "0"), num("x2", "0"), num("y2", "0")
        svg["elements"].append((np.array([1, x1, y1, 2, x2, y2], np.float32), p))
    elif tag in ("polyline", "polygon"):
        p = parse_props(node, stack[-1])
        pts = node.attr("points")
        v = []
        if "," in pts:
            for pair in pts.split(" "):
                parts = pair.split(",")
                if len(parts) != 2:
                    raise SvgError("Invalid SVG data")
                v.append((parse_float(parts[0]), parse_float(parts[1])))
        else:
            f = pts.split(" ")
            if len(f) % 2:
                raise SvgError("Invalid SVG data")
            v = [(parse_float(f[2 * k]), parse_float(f[2 * k + 1])) for k in range(len(f) // 2)]
        if not v:
            raise SvgError("Invalid SVG data")
        c = [1, v[0][0], v[0][1]]
        for x, y in v[1:]:
            c += [2, x, y]
        if tag == "polygon":
            c.append(0)
        svg["elements"].append((np.array(c, np.float32), p))
    elif tag == "rect":
        p = parse_props(node, stack[-1])
        x, y, w, h = num("x", "0"), num("y", "0"), num("width", "0"), num("height", "0")
        if w == 0 or h == 0:
            return
        rx, ry = max(num("rx", "0"), 0), max(num("ry", "0"), 0)
        if rx > 0 or ry > 0:
            if rx == 0:
                rx = ry
            elif ry == 0:
                ry = rx
            rx, ry = min(rx, w / 2), min(ry, h / 2)
            arc = lambda ex, ey: [9, rx, ry, 0, 0, 1, ex, ey]  # noqa: E731
            c = [1, x + rx, y, 2, x + w - rx, y] + arc(x + w, y + ry) + [2, x + w, y + h - ry]
            c += arc(x + w - rx, y + h) + [2, x + rx, y + h] + arc(x, y + h - ry)
            c += [2, x, y + ry] + arc(x + rx, y)
        else:
            x, y, w, h = f32(x), f32(y), f32(w), f32(h)
            c = [1, x, y, 2, f32(x + w), y, 2, f32(x + w), f32(y + h), 2, x, f32(y + h), 0]
        svg["elements"].append((np.array(c, np.float32), p))
    elif tag in ("circle", "ellipse"):
        p = parse_props(node, stack[-1])
        cx, cy = num("cx", "0"), num("cy", "0")
        if tag == "circle":
            rx = ry = f32(parse_float(node.attr("r")))
        else:
            rx, ry = f32(num("rx", "0")), f32(num("ry", "0"))
        svg["elements"].append((np.array(_ellipse(cx, cy, rx, ry), np.float32), p))
    elif tag == "linearGradient":
        parse_props(node, stack[-1])
        units, gt = node.attr("gradientUnits"), node.attr("gradientTransform")
        if units != "userSpaceOnUse":
            raise SvgError("Unsupported gradient units: " + units)
        if gt:
            raise SvgError("Unsupported gradient transform: " + gt)
        g = {"p": [f32(parse_float(node.attr(k))) for k in ("x1", "y1", "x2", "y2")], "stops": []}
        for k in node.kids:
            if k.tag != "stop":
                raise SvgError("Unexpected SVG tag: " + k.tag)
            color = k.attr("stop-color")
            if not color:
                for pair in k.attr("style").split(";"):
                    parts = pair.split(":")
                    if len(parts) == 2 and _strip(parts[0]) == "stop-color" and not color:
                        color = _strip(parts[1])
            if not color:
                raise SvgError("Invalid SVG gradient, missing stop-color")
            g["stops"].append((parse_color(color), float(f32(parse_float(k.attr("offset"))))))
        svg["gradients"][node.attr("id")] = g
    else:
        raise SvgError("Unsupported SVG tag: " + tag)


def parse_svg(orc, data, width=0, height=0):
    """parseSvg (svg.nim:502-555): {width, height, elements, gradients}."""
    try:
        root = Node(ET.fromstring(data))
    except ET.ParseError as e:
        raise SvgError("Invalid SVG data: " + str(e))
    if root.tag != "svg":
        raise SvgError("Invalid SVG data")
    box = root.attr("viewBox").split(" ")
    if len(box) < 4:
        raise SvgError("Invalid SVG data: viewBox")
    minx, miny, vw, vh = [parse_int(b) for b in box[:4]]
    props = parse_props(root, initial_props())
    if minx != 0 or miny != 0:
        props["transform"] = m3_mul(props["transform"], [1, 0, 0, 0, 1, 0, -minx, -miny, 1])
    svg = {"elements": [], "gradients": {}}
    if width == 0 and height == 0:
        svg["width"], svg["height"] = vw, vh
    else:
        svg["width"], svg["height"] = width, height
        sx, sy = f32(f32(width) / f32(vw)), f32(f32(height) / f32(vh))
        props["transform"] = m3_mul(props["transform"], [sx, 0, 0, 0, sy, 0, 0, 0, 1])
    stack = [props]
    for k in root.kids:
        parse_element(orc, k, svg, stack)
    return svg


def render(orc, svg):
    """newImage(svg) (svg.nim:557-602)."""
    img = orc.new_image(svg["width"], svg["height"])
    blend = OVERWRITE
    for cmds, p in svg["elements"]:
        if not (p["display"] and p["opacity"] > 0):
            continue
        mat = [float(v) for v in p["transform"]]
        if p["fill"] != "none":
            opacity = f32(p["fillOpacity"] * p["opacity"])
            if p["fill"].startswith("url("):
                close = p["fill"].find(")", 5)
                if close == -1:
                    raise SvgError("Malformed fill: " + p["fill"])
                g = svg["gradients"].get(p["fill"][5:close])
                if g is None:
                    raise SvgError("Missing SVG resource " + p["fill"][5:close])
                h0 = m3_apply(p["transform"], g["p"][0], g["p"][1])
                h1 = m3_apply(p["transform"], g["p"][2], g["p"][3])
                paint = orc.make_paint(orc.LINEAR, [h0, h1], g["stops"], float(opacity), blend)
                orc.fill_path_paint(img, cmds, paint, p["fillRule"], mat)
            else:
                c = rgbx_color(color_rgbx(parse_color(p["fill"])))
                o = min(max(opacity, f32(0)), f32(1))
                if o != 0 and (c[3] > 0 or blend == OVERWRITE):
                    rgbx = color_rgbx((c[0], c[1], c[2], float(f32(f32(c[3]) * o))))
                    orc.fill_path(img, cmds, rgbx, p["fillRule"], blend, mat)
        blend = NORMAL
        if p["stroke"] != (0, 0, 0, 0) and p["strokeWidth"] > 0:
            c = list(rgbx_color(p["stroke"]))
            c[3] = float(f32(f32(c[3]) * f32(p["opacity"] * p["strokeOpacity"])))
            if c[3] > 0:
                orc.stroke_path(img, cmds, color_rgbx(c), float(p["strokeWidth"]), p["lineCap"],
                                p["lineJoin"], float(p["miterLimit"]),
                                [float(d) for d in p["dashes"]], NORMAL, mat)
    return img
