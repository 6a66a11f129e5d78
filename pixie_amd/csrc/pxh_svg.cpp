// pxh_svg.cpp -- SVG ingestion for the fill path: parseSvg and newImage(svg)
// (fileformats/svg.nim).  The document is parsed on the host into pixie
// path commands with their properties; rendering records every fill and
// stroke, in document order, into one scene drawn in a single device pass
// (the first fill overwrites, as newImage does), with calls the scene form
// does not take (gradient fills) issued one by one in order between scene
// passes.
//
// Numbers follow the reference's types: attribute values are parsed as
// float64 (Nim's parseFloat) and narrowed to float32 where pixie stores or
// passes them as float32, after any float64 arithmetic pixie does first.

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/pixie_hip.h"
#include "pxh_common.h"
#include "pxh_host_geom.h"

namespace {

using namespace pxh;

struct SvgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

[[noreturn]] void invalid() { throw SvgError("Invalid SVG data"); }

// ---- XML (the subset of Nim's xmlparser an SVG document needs) ----

struct XNode {
  bool element = false;
  std::string tag;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<XNode> kids;
  const std::string& attr(const char* name) const {  // "" when absent
    static const std::string none;
    for (const auto& a : attrs)
      if (a.first == name) return a.second;
    return none;
  }
};

struct XmlParser {
  const char* s;
  size_t n, i = 0;
  [[noreturn]] void fail(const char* what) {
    throw SvgError(std::string("Invalid SVG data: XML ") + what + " at byte " + std::to_string(i));
  }
  bool at(const char* t) const { return strncmp(s + i, t, strlen(t)) == 0 && i + strlen(t) <= n; }
  static bool space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }
  static bool name_char(char c) {
    return !(space(c) || c == '/' || c == '>' || c == '=' || c == '<' || c == '\0' || c == '"' || c == '\'');
  }
  void skip_space() {
    while (i < n && space(s[i])) i++;
  }
  void skip_until(const char* end) {
    const char* e = strstr(s + i, end);
    if (!e || (size_t)(e - s) >= n) fail("unterminated construct");
    i = (size_t)(e - s) + strlen(end);
  }
  // <!DOCTYPE ...> with an optional internal subset in brackets
  void skip_doctype() {
    int depth = 0;
    while (i < n) {
      const char c = s[i++];
      if (c == '[') depth++;
      if (c == ']') depth--;
      if (c == '>' && depth <= 0) return;
    }
    fail("unterminated DOCTYPE");
  }
  std::string name() {
    const size_t b = i;
    while (i < n && name_char(s[i])) i++;
    if (i == b) fail("name expected");
    return std::string(s + b, i - b);
  }
  // &lt; &gt; &amp; &apos; &quot; &#n; &#xn; (others kept as written)
  void entity(std::string* out) {
    const size_t semi = i + 1;
    size_t e = semi;
    while (e < n && e - i < 12 && s[e] != ';') e++;
    if (e >= n || s[e] != ';') {
      out->push_back('&');
      i++;
      return;
    }
    const std::string nm(s + i + 1, e - i - 1);
    uint32_t c = 0;
    bool ok = true;
    if (nm == "lt") c = '<';
    else if (nm == "gt") c = '>';
    else if (nm == "amp") c = '&';
    else if (nm == "apos") c = '\'';
    else if (nm == "quot") c = '"';
    else if (nm.size() > 1 && nm[0] == '#') {
      char* end = nullptr;
      c = (nm[1] == 'x' || nm[1] == 'X') ? (uint32_t)strtoul(nm.c_str() + 2, &end, 16)
                                        : (uint32_t)strtoul(nm.c_str() + 1, &end, 10);
      ok = end && *end == 0;
    } else {
      ok = false;
    }
    if (!ok) {
      out->append(s + i, e + 1 - i);
      i = e + 1;
      return;
    }
    if (c < 0x80) {
      out->push_back((char)c);
    } else if (c < 0x800) {
      out->push_back((char)(0xC0 | (c >> 6)));
      out->push_back((char)(0x80 | (c & 63)));
    } else if (c < 0x10000) {
      out->push_back((char)(0xE0 | (c >> 12)));
      out->push_back((char)(0x80 | ((c >> 6) & 63)));
      out->push_back((char)(0x80 | (c & 63)));
    } else {
      out->push_back((char)(0xF0 | (c >> 18)));
      out->push_back((char)(0x80 | ((c >> 12) & 63)));
      out->push_back((char)(0x80 | ((c >> 6) & 63)));
      out->push_back((char)(0x80 | (c & 63)));
    }
    i = e + 1;
  }
  // An attribute value: entities replaced, each run of blanks and line
  // breaks read as one space, a trailing run dropped (as Nim's parsexml).
  std::string value() {
    if (i >= n || (s[i] != '"' && s[i] != '\'')) fail("quoted attribute value expected");
    const char q = s[i++];
    std::string v;
    bool pending = false;
    while (true) {
      if (i >= n) fail("unterminated attribute value");
      const char c = s[i];
      if (c == q) {
        i++;
        break;
      }
      if (space(c)) {
        pending = true;
        i++;
        continue;
      }
      if (pending) v.push_back(' ');
      pending = false;
      if (c == '&')
        entity(&v);
      else
        v.push_back(c), i++;
    }
    return v;
  }
  // Skips prolog items and reads the root element.
  XNode document() {
    if (n >= 3 && (uint8_t)s[0] == 0xEF && (uint8_t)s[1] == 0xBB && (uint8_t)s[2] == 0xBF) i = 3;
    while (true) {
      skip_space();
      if (i >= n) fail("root element expected");
      if (at("<?")) {
        skip_until("?>");
      } else if (at("<!--")) {
        skip_until("-->");
      } else if (at("<!")) {
        skip_doctype();
      } else if (s[i] == '<') {
        return element();
      } else {
        fail("root element expected");
      }
    }
  }
  int depth = 0;
  XNode element() {
    if (++depth > 512) fail("nesting too deep");
    XNode e;
    e.element = true;
    i++;  // '<'
    e.tag = name();
    while (true) {
      skip_space();
      if (i >= n) fail("unterminated tag");
      if (s[i] == '/') {
        if (i + 1 >= n || s[i + 1] != '>') fail("'>' expected");
        i += 2;
        depth--;
        return e;
      }
      if (s[i] == '>') {
        i++;
        break;
      }
      std::string k = name();
      skip_space();
      if (i >= n || s[i] != '=') fail("'=' expected");
      i++;
      skip_space();
      e.attrs.push_back({k, value()});
    }
    while (true) {  // content
      if (i >= n) fail("unterminated element");
      if (at("</")) {
        i += 2;
        const std::string closing = name();
        if (closing != e.tag) fail("mismatched closing tag");
        skip_space();
        if (i >= n || s[i] != '>') fail("'>' expected");
        i++;
        depth--;
        return e;
      }
      if (at("<!--")) {
        skip_until("-->");
      } else if (at("<![CDATA[")) {
        skip_until("]]>");
      } else if (at("<?")) {
        skip_until("?>");
      } else if (s[i] == '<') {
        e.kids.push_back(element());
      } else {
        while (i < n && s[i] != '<') i++;  // character data: not used
      }
    }
  }
};

// ---- numbers (Nim's strutils.parseFloat / parseInt) ----

// The whole string must be a number: optional sign, digits (with '_'
// separators), an optional fraction, an optional exponent, or inf / nan.
double parse_float(const std::string& t) {
  const char* s = t.c_str();
  size_t i = 0, n = t.size();
  std::string clean;
  if (i < n && (s[i] == '+' || s[i] == '-')) clean.push_back(s[i++]);
  auto word = [&](const char* w) {
    const size_t l = strlen(w);
    if (n - i != l) return false;
    for (size_t k = 0; k < l; k++)
      if ((s[i + k] | 0x20) != w[k]) return false;
    return true;
  };
  if (word("nan")) return NAN;
  if (word("inf")) return clean == "-" ? -INFINITY : INFINITY;
  int digits = 0;
  while (i < n && ((s[i] >= '0' && s[i] <= '9') || (s[i] == '_' && digits))) {
    if (s[i] != '_') clean.push_back(s[i]), digits++;
    i++;
  }
  if (i < n && s[i] == '.') {
    clean.push_back(s[i++]);
    while (i < n && ((s[i] >= '0' && s[i] <= '9') || s[i] == '_')) {
      if (s[i] != '_') clean.push_back(s[i]), digits++;
      i++;
    }
  }
  if (digits && i < n && (s[i] == 'e' || s[i] == 'E')) {
    clean.push_back(s[i++]);
    if (i < n && (s[i] == '+' || s[i] == '-')) clean.push_back(s[i++]);
    int ed = 0;
    while (i < n && ((s[i] >= '0' && s[i] <= '9') || s[i] == '_')) {
      if (s[i] != '_') clean.push_back(s[i]), ed++;
      i++;
    }
    if (!ed) throw SvgError("invalid float: " + t);
  }
  if (!digits || i != n) throw SvgError("invalid float: " + t);
  return strtod(clean.c_str(), nullptr);
}

int64_t parse_int(const std::string& t) {
  size_t i = 0;
  bool neg = false;
  if (i < t.size() && (t[i] == '+' || t[i] == '-')) neg = t[i++] == '-';
  int64_t v = 0;
  int digits = 0;
  for (; i < t.size(); i++) {
    if (t[i] == '_' && digits) continue;
    if (t[i] < '0' || t[i] > '9') break;
    v = v * 10 + (t[i] - '0');
    digits++;
  }
  if (!digits || i != t.size()) throw SvgError("invalid integer: " + t);
  return neg ? -v : v;
}

std::vector<std::string> split(const std::string& s, char sep) {  // strutils.split
  std::vector<std::string> out;
  size_t b = 0;
  for (size_t k = 0; k <= s.size(); k++)
    if (k == s.size() || s[k] == sep) {
      out.push_back(s.substr(b, k - b));
      b = k + 1;
    }
  return out;
}

std::string strip(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && (s[b] == ' ' || s[b] == '\t' || s[b] == '\r' || s[b] == '\n' || s[b] == '\v' || s[b] == '\f')) b++;
  while (e > b && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r' || s[e - 1] == '\n' ||
                   s[e - 1] == '\v' || s[e - 1] == '\f'))
    e--;
  return s.substr(b, e - b);
}

bool starts_with(const std::string& s, const char* p) { return s.compare(0, strlen(p), p) == 0; }
bool ends_with(const std::string& s, const char* p) {
  const size_t l = strlen(p);
  return s.size() >= l && s.compare(s.size() - l, l, p) == 0;
}

// "(1,1)", "(1 1)", "(1, 1)" and the like: the non-empty fields after commas
// become spaces (svg.nim:57-62).
std::vector<std::string> split_args(const std::string& s) {
  std::string t = s;
  for (char& c : t)
    if (c == ',') c = ' ';
  std::vector<std::string> out;
  for (auto& f : split(t, ' '))
    if (!f.empty()) out.push_back(f);
  return out;
}

// ---- colours (chroma's parseHtmlColor and its conversions) ----

struct Color {
  float r, g, b, a;
};

const std::unordered_map<std::string, uint32_t>& css_names() {
  // The CSS Color Module named colours (0xRRGGBB).
  static const std::unordered_map<std::string, uint32_t> m = {
      {"aliceblue", 0xF0F8FF}, {"antiquewhite", 0xFAEBD7}, {"aqua", 0x00FFFF},
      {"aquamarine", 0x7FFFD4}, {"azure", 0xF0FFFF}, {"beige", 0xF5F5DC},
      {"bisque", 0xFFE4C4}, {"black", 0x000000}, {"blanchedalmond", 0xFFEBCD},
      {"blue", 0x0000FF}, {"blueviolet", 0x8A2BE2}, {"brown", 0xA52A2A},
      {"burlywood", 0xDEB887}, {"cadetblue", 0x5F9EA0}, {"chartreuse", 0x7FFF00},
      {"chocolate", 0xD2691E}, {"coral", 0xFF7F50}, {"cornflowerblue", 0x6495ED},
      {"cornsilk", 0xFFF8DC}, {"crimson", 0xDC143C}, {"cyan", 0x00FFFF},
      {"darkblue", 0x00008B}, {"darkcyan", 0x008B8B}, {"darkgoldenrod", 0xB8860B},
      {"darkgray", 0xA9A9A9}, {"darkgreen", 0x006400}, {"darkgrey", 0xA9A9A9},
      {"darkkhaki", 0xBDB76B}, {"darkmagenta", 0x8B008B}, {"darkolivegreen", 0x556B2F},
      {"darkorange", 0xFF8C00}, {"darkorchid", 0x9932CC}, {"darkred", 0x8B0000},
      {"darksalmon", 0xE9967A}, {"darkseagreen", 0x8FBC8F}, {"darkslateblue", 0x483D8B},
      {"darkslategray", 0x2F4F4F}, {"darkslategrey", 0x2F4F4F}, {"darkturquoise", 0x00CED1},
      {"darkviolet", 0x9400D3}, {"deeppink", 0xFF1493}, {"deepskyblue", 0x00BFFF},
      {"dimgray", 0x696969}, {"dimgrey", 0x696969}, {"dodgerblue", 0x1E90FF},
      {"firebrick", 0xB22222}, {"floralwhite", 0xFFFAF0}, {"forestgreen", 0x228B22},
      {"fuchsia", 0xFF00FF}, {"gainsboro", 0xDCDCDC}, {"ghostwhite", 0xF8F8FF},
      {"gold", 0xFFD700}, {"goldenrod", 0xDAA520}, {"gray", 0x808080},
      {"grey", 0x808080}, {"green", 0x008000}, {"greenyellow", 0xADFF2F},
      {"honeydew", 0xF0FFF0}, {"hotpink", 0xFF69B4}, {"indianred", 0xCD5C5C},
      {"indigo", 0x4B0082}, {"ivory", 0xFFFFF0}, {"khaki", 0xF0E68C},
      {"lavender", 0xE6E6FA}, {"lavenderblush", 0xFFF0F5}, {"lawngreen", 0x7CFC00},
      {"lemonchiffon", 0xFFFACD}, {"lightblue", 0xADD8E6}, {"lightcoral", 0xF08080},
      {"lightcyan", 0xE0FFFF}, {"lightgoldenrodyellow", 0xFAFAD2}, {"lightgray", 0xD3D3D3},
      {"lightgreen", 0x90EE90}, {"lightgrey", 0xD3D3D3}, {"lightpink", 0xFFB6C1},
      {"lightsalmon", 0xFFA07A}, {"lightseagreen", 0x20B2AA}, {"lightskyblue", 0x87CEFA},
      {"lightslategray", 0x778899}, {"lightslategrey", 0x778899}, {"lightsteelblue", 0xB0C4DE},
      {"lightyellow", 0xFFFFE0}, {"lime", 0x00FF00}, {"limegreen", 0x32CD32},
      {"linen", 0xFAF0E6}, {"magenta", 0xFF00FF}, {"maroon", 0x800000},
      {"mediumaquamarine", 0x66CDAA}, {"mediumblue", 0x0000CD}, {"mediumorchid", 0xBA55D3},
      {"mediumpurple", 0x9370DB}, {"mediumseagreen", 0x3CB371}, {"mediumslateblue", 0x7B68EE},
      {"mediumspringgreen", 0x00FA9A}, {"mediumturquoise", 0x48D1CC}, {"mediumvioletred", 0xC71585},
      {"midnightblue", 0x191970}, {"mintcream", 0xF5FFFA}, {"mistyrose", 0xFFE4E1},
      {"moccasin", 0xFFE4B5}, {"navajowhite", 0xFFDEAD}, {"navy", 0x000080},
      {"oldlace", 0xFDF5E6}, {"olive", 0x808000}, {"olivedrab", 0x6B8E23},
      {"orange", 0xFFA500}, {"orangered", 0xFF4500}, {"orchid", 0xDA70D6},
      {"palegoldenrod", 0xEEE8AA}, {"palegreen", 0x98FB98}, {"paleturquoise", 0xAFEEEE},
      {"palevioletred", 0xDB7093}, {"papayawhip", 0xFFEFD5}, {"peachpuff", 0xFFDAB9},
      {"peru", 0xCD853F}, {"pink", 0xFFC0CB}, {"plum", 0xDDA0DD},
      {"powderblue", 0xB0E0E6}, {"purple", 0x800080}, {"rebeccapurple", 0x663399},
      {"red", 0xFF0000}, {"rosybrown", 0xBC8F8F}, {"royalblue", 0x4169E1},
      {"saddlebrown", 0x8B4513}, {"salmon", 0xFA8072}, {"sandybrown", 0xF4A460},
      {"seagreen", 0x2E8B57}, {"seashell", 0xFFF5EE}, {"sienna", 0xA0522D},
      {"silver", 0xC0C0C0}, {"skyblue", 0x87CEEB}, {"slateblue", 0x6A5ACD},
      {"slategray", 0x708090}, {"slategrey", 0x708090}, {"snow", 0xFFFAFA},
      {"springgreen", 0x00FF7F}, {"steelblue", 0x4682B4}, {"tan", 0xD2B48C},
      {"teal", 0x008080}, {"thistle", 0xD8BFD8}, {"tomato", 0xFF6347},
      {"turquoise", 0x40E0D0}, {"violet", 0xEE82EE}, {"wheat", 0xF5DEB3},
      {"white", 0xFFFFFF}, {"whitesmoke", 0xF5F5F5}, {"yellow", 0xFFFF00},
      {"yellowgreen", 0x9ACD32}};
  return m;
}

int hex_digit(char c) {
  if (c >= '0' && c <= '9') return c - '0';
  if (c >= 'a' && c <= 'f') return c - 'a' + 10;
  if (c >= 'A' && c <= 'F') return c - 'A' + 10;
  return -1;
}

int hex_byte(const std::string& s, size_t at) {
  const int a = hex_digit(s[at]), b = hex_digit(s[at + 1]);
  if (a < 0 || b < 0) throw SvgError("Invalid color " + s);
  return a * 16 + b;
}

// parseHtmlColor: "#rgb", "#rrggbb", "rgb(r, g, b)", "rgba(r, g, b, a)" or a
// CSS colour name (case-insensitive); channels byte / 255 in float32.
Color parse_color(const std::string& text) {
  const std::string s = strip(text);
  const auto byte = [](int v) { return (float)v / 255.0f; };
  if (s.empty()) throw SvgError("Invalid color " + text);
  if (s[0] == '#') {
    if (s.size() == 4) {
      int d[3];
      for (int k = 0; k < 3; k++)
        if ((d[k] = hex_digit(s[1 + k])) < 0) throw SvgError("Invalid color " + text);
      return {(float)d[0] / 15.0f, (float)d[1] / 15.0f, (float)d[2] / 15.0f, 1.0f};
    }
    if (s.size() == 7) return {byte(hex_byte(s, 1)), byte(hex_byte(s, 3)), byte(hex_byte(s, 5)), 1.0f};
    if (s.size() == 9)
      return {byte(hex_byte(s, 1)), byte(hex_byte(s, 3)), byte(hex_byte(s, 5)), byte(hex_byte(s, 7))};
    throw SvgError("Invalid color " + text);
  }
  if (starts_with(s, "rgba(") || starts_with(s, "rgb(")) {
    const bool alpha = s[3] == 'a';
    if (s.back() != ')') throw SvgError("Invalid color " + text);
    const auto f = split(s.substr(alpha ? 5 : 4, s.size() - (alpha ? 6 : 5)), ',');
    if (f.size() != (alpha ? 4u : 3u)) throw SvgError("Invalid color " + text);
    Color c{byte((int)parse_int(strip(f[0]))), byte((int)parse_int(strip(f[1]))),
            byte((int)parse_int(strip(f[2]))), 1.0f};
    if (alpha) c.a = (float)parse_float(strip(f[3]));
    return c;
  }
  std::string low = s;
  for (char& ch : low)
    if (ch >= 'A' && ch <= 'Z') ch = (char)(ch - 'A' + 'a');
  auto it = css_names().find(low);
  if (it == css_names().end()) throw SvgError("Invalid color " + text);
  const uint32_t v = it->second;
  return {byte((v >> 16) & 255), byte((v >> 8) & 255), byte(v & 255), 1.0f};
}

// Color -> ColorRGBX: each channel clamped, times 255 rounded half away from
// zero, then premultiplied as (c * a + 127) div 255.
uint32_t color_rgbx(Color c) {
  const float v[4] = {c.r, c.g, c.b, c.a};
  uint32_t b[4];
  for (int k = 0; k < 4; k++) {
    const float x = v[k] < 0.0f ? 0.0f : (v[k] > 1.0f ? 1.0f : v[k]);
    b[k] = (uint32_t)floor((double)(x * 255.0f) + 0.5);
  }
  if (b[3] != 255)
    for (int k = 0; k < 3; k++) b[k] = (b[k] * b[3] + 127) / 255;
  return pack(b[0], b[1], b[2], b[3]);
}

// ColorRGBX -> Color: straight alpha bytes (each premultiplied channel times
// uint32((255 / a) * 255), plus 127, div 255) over 255.
Color rgbx_color(uint32_t x) {
  uint32_t r = ch_r(x), g = ch_g(x), b = ch_b(x);
  const uint32_t a = ch_a(x);
  if (a != 0 && a != 255) {
    const uint32_t mul = nim_f2u32((255.0f / (float)a) * 255.0f);
    r = (r * mul + 127) / 255;
    g = (g * mul + 127) / 255;
    b = (b * mul + 127) / 255;
  }
  return {(float)r / 255.0f, (float)g / 255.0f, (float)b / 255.0f, (float)a / 255.0f};
}

// ---- properties (svg.nim:19-48, 50-293) ----

struct Props {
  bool display = true;
  int32_t fillRule = kNonZero;
  std::string fill = "black";
  uint32_t stroke = 0;  // ColorRGBX
  float strokeWidth = 1.0f;
  int32_t cap = 0, join = 0;
  float miterLimit = 4.0f;
  std::vector<float> dashes;
  float transform[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  float opacity = 1.0f, fillOpacity = 1.0f, strokeOpacity = 1.0f;
};

float clamp01(double v) { return (float)(v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v)); }

void mul_into(float* t, const float* m) {  // t = t * m
  float r[9];
  m3_mul(t, m, r);
  memcpy(t, r, sizeof(r));
}

void parse_transform(const std::string& transform, float* t) {
  auto bad = [&]() { throw SvgError("Unsupported SVG transform: " + transform); };
  std::string rem = transform;
  while (!rem.empty()) {
    const size_t idx = rem.find(')');
    if (idx == std::string::npos) bad();
    const std::string f = strip(rem.substr(0, idx + 1));
    rem = rem.substr(idx + 1);
    if (starts_with(f, "matrix(")) {
      const auto a = split_args(f.substr(7, f.size() - 8));
      if (a.size() != 6) bad();
      const float m[9] = {(float)parse_float(a[0]), (float)parse_float(a[1]), 0,
                          (float)parse_float(a[2]), (float)parse_float(a[3]), 0,
                          (float)parse_float(a[4]), (float)parse_float(a[5]), 1};
      mul_into(t, m);
    } else if (starts_with(f, "translate(")) {
      const auto c = split_args(f.substr(10, f.size() - 11));
      if (c.empty()) throw SvgError("index out of bounds");
      const double tx = parse_float(c[0]), ty = c.size() == 1 ? 0.0 : parse_float(c[1]);
      const float m[9] = {1, 0, 0, 0, 1, 0, (float)tx, (float)ty, 1};
      mul_into(t, m);
    } else if (starts_with(f, "rotate(")) {
      const auto v = split_args(f.substr(7, f.size() - 8));
      if (v.empty()) throw SvgError("index out of bounds");
      const float angle = (float)(parse_float(v[0]) * M_PI / 180.0);
      float cx = 0, cy = 0;
      if (v.size() > 1) cx = (float)parse_float(v[1]);
      if (v.size() > 2) cy = (float)parse_float(v[2]);
      const float s = sinf(angle), c = cosf(angle);
      const float tc[9] = {1, 0, 0, 0, 1, 0, cx, cy, 1};
      const float r[9] = {c, s, 0, -s, c, 0, 0, 0, 1};
      const float tn[9] = {1, 0, 0, 0, 1, 0, -cx, -cy, 1};
      mul_into(t, tc);
      mul_into(t, r);
      mul_into(t, tn);
    } else if (starts_with(f, "scale(")) {
      const auto v = split_args(f.substr(6, f.size() - 7));
      if (v.empty()) throw SvgError("index out of bounds");
      const float sx = (float)parse_float(v[0]);
      const float sy = v.size() > 1 ? (float)parse_float(v[1]) : sx;
      const float m[9] = {sx, 0, 0, 0, sy, 0, 0, 0, 1};
      mul_into(t, m);
    } else {
      bad();
    }
  }
}

Props parse_props(const XNode& node, const Props& inherited) {
  Props r = inherited;
  std::string fillRule = node.attr("fill-rule"), fill = node.attr("fill"), stroke = node.attr("stroke"),
              strokeWidth = node.attr("stroke-width"), cap = node.attr("stroke-linecap"),
              join = node.attr("stroke-linejoin"), miter = node.attr("stroke-miterlimit"),
              dash = node.attr("stroke-dasharray"), transform = node.attr("transform"),
              display = node.attr("display"), opacity = node.attr("opacity"),
              fillOpacity = node.attr("fill-opacity"), strokeOpacity = node.attr("stroke-opacity");
  const std::string style = node.attr("style");
  for (const std::string& pair : split(style, ';')) {  // element attributes win
    const auto parts = split(pair, ':');
    if (parts.size() != 2) continue;
    const std::string k = strip(parts[0]), v = strip(parts[1]);
    std::string* slot = nullptr;
    if (k == "fill-rule") slot = &fillRule;
    else if (k == "fill") slot = &fill;
    else if (k == "stroke") slot = &stroke;
    else if (k == "stroke-linecap") slot = &cap;
    else if (k == "stroke-linejoin") slot = &join;
    else if (k == "stroke-width") slot = &strokeWidth;
    else if (k == "stroke-miterlimit") slot = &miter;
    else if (k == "stroke-dasharray") slot = &dash;
    else if (k == "display") slot = &display;
    else if (k == "opacity") slot = &opacity;
    else if (k == "fillOpacity") slot = &fillOpacity;
    else if (k == "strokeOpacity") slot = &strokeOpacity;
    if (slot && slot->empty()) *slot = v;
  }
  if (!display.empty()) r.display = strip(display) != "none";
  if (!opacity.empty()) r.opacity = clamp01(parse_float(opacity));
  if (fillRule == "nonzero") r.fillRule = kNonZero;
  else if (fillRule == "evenodd") r.fillRule = kEvenOdd;
  else if (!fillRule.empty()) throw SvgError("Invalid fill-rule value " + fillRule);
  r.fill = (fill.empty() || fill == "currentColor") ? inherited.fill : fill;
  if (stroke == "currentColor") {
    if (r.stroke == 0) r.stroke = pack(0, 0, 0, 255);
  } else if (stroke == "none") {
    r.stroke = 0;
  } else if (!stroke.empty()) {
    r.stroke = color_rgbx(parse_color(stroke));
  }
  if (!fillOpacity.empty()) r.fillOpacity = clamp01(parse_float(fillOpacity));
  if (!strokeOpacity.empty()) r.strokeOpacity = clamp01(parse_float(strokeOpacity));
  if (!strokeWidth.empty()) {
    if (ends_with(strokeWidth, "px")) strokeWidth = strokeWidth.substr(0, strokeWidth.size() - 2);
    r.strokeWidth = (float)parse_float(strokeWidth);
    if (r.stroke == 0) r.stroke = pack(0, 0, 0, 255);
  }
  if (!cap.empty()) {
    if (cap == "butt") r.cap = 0;
    else if (cap == "round") r.cap = 1;
    else if (cap == "square") r.cap = 2;
    else if (cap != "inherit") throw SvgError("Invalid stroke-linecap value " + cap);
  }
  if (!join.empty()) {
    if (join == "miter") r.join = 0;
    else if (join == "round") r.join = 1;
    else if (join == "bevel") r.join = 2;
    else if (join != "inherit") throw SvgError("Invalid stroke-linejoin value " + join);
  }
  if (!miter.empty()) r.miterLimit = (float)parse_float(miter);
  if (!dash.empty())
    for (const auto& v : split_args(dash)) r.dashes.push_back((float)parse_float(v));
  if (!transform.empty()) parse_transform(transform, r.transform);
  return r;
}

// ---- path builders (paths.nim:335-631) ----

struct PathB {
  std::vector<float> c;
  void move(float x, float y) { c.insert(c.end(), {1.0f, x, y}); }
  void line(float x, float y) { c.insert(c.end(), {2.0f, x, y}); }
  void cubic(float a, float b, float d, float e, float f, float g) { c.insert(c.end(), {5.0f, a, b, d, e, f, g}); }
  void arc(float rx, float ry, float rot, bool large, bool sweep, float x, float y) {
    c.insert(c.end(), {9.0f, rx, ry, rot, large ? 1.0f : 0.0f, sweep ? 1.0f : 0.0f, x, y});
  }
  void close() { c.push_back(0.0f); }
  void rect(float x, float y, float w, float h) {
    move(x, y);
    line(x + w, y);
    line(x + w, y + h);
    line(x, y + h);
    close();
  }
  // splineCircleK * r stays float64; sums with it are narrowed to float32
  void ellipse(float cx, float cy, float rx, float ry) {
    const double k = 4.0 * (-1.0 + 1.4142135623730951) / 3.0;
    const double mx = k * (double)rx, my = k * (double)ry;
    auto n = [](double v) { return (float)v; };
    move(cx + rx, cy);
    cubic(cx + rx, n(cy + my), n(cx + mx), cy + ry, cx, cy + ry);
    cubic(n(cx - mx), cy + ry, cx - rx, n(cy + my), cx - rx, cy);
    cubic(cx - rx, n(cy - my), n(cx - mx), cy - ry, cx, cy - ry);
    cubic(n(cx + mx), cy - ry, cx + rx, n(cy - my), cx + rx, cy);
    close();
  }
};

struct LinearGradient {
  float x1, y1, x2, y2;
  std::vector<float> stops;  // r, g, b, a, position
};

struct Element {
  std::vector<float> cmds;
  Props props;
};

}  // namespace

struct pxh_svg {
  int32_t width = 0, height = 0;
  std::vector<Element> elements;
  std::unordered_map<std::string, LinearGradient> gradients;
  pxh_scene* scene = nullptr;
};

namespace {

void parse_element(const XNode& node, pxh_svg* svg, std::vector<Props>& stack) {  // svg.nim:295-500
  if (!node.element) return;
  const std::string& tag = node.tag;
  if (tag == "title" || tag == "desc" || tag == "defs" || tag == "radialGradient") return;
  auto num = [&](const char* name, const char* def) {
    const std::string& v = node.attr(name);
    return parse_float(v.empty() ? std::string(def) : v);
  };
  if (tag == "g") {
    stack.push_back(parse_props(node, stack.back()));
    for (const XNode& k : node.kids) parse_element(k, svg, stack);
    stack.pop_back();
    return;
  }
  if (tag == "path") {
    const std::string& d = node.attr("d");
    Props p = parse_props(node, stack.back());
    Element e;
    std::string err;
    if (parse_path(d.c_str(), &e.cmds, &err) != kOk) throw SvgError(err);
    e.props = std::move(p);
    svg->elements.push_back(std::move(e));
    return;
  }
  if (tag == "line") {
    Props p = parse_props(node, stack.back());
    const double x1 = num("x1", "0"), y1 = num("y1", "0"), x2 = num("x2", "0"), y2 = num("y2", "0");
    PathB b;
    b.move((float)x1, (float)y1);
    b.line((float)x2, (float)y2);
    svg->elements.push_back({std::move(b.c), std::move(p)});
    return;
  }
  if (tag == "polyline" || tag == "polygon") {
    Props p = parse_props(node, stack.back());
    const std::string& points = node.attr("points");
    std::vector<std::pair<float, float>> v;
    if (points.find(',') != std::string::npos) {
      for (const auto& pair : split(points, ' ')) {
        const auto parts = split(pair, ',');
        if (parts.size() != 2) invalid();
        v.push_back({(float)parse_float(parts[0]), (float)parse_float(parts[1])});
      }
    } else {
      const auto f = split(points, ' ');
      if (f.size() % 2 != 0) invalid();
      for (size_t k = 0; k < f.size() / 2; k++)
        v.push_back({(float)parse_float(f[2 * k]), (float)parse_float(f[2 * k + 1])});
    }
    if (v.empty()) invalid();
    PathB b;
    b.move(v[0].first, v[0].second);
    for (size_t k = 1; k < v.size(); k++) b.line(v[k].first, v[k].second);
    if (tag == "polygon") b.close();
    svg->elements.push_back({std::move(b.c), std::move(p)});
    return;
  }
  if (tag == "rect") {
    Props p = parse_props(node, stack.back());
    const double x = num("x", "0"), y = num("y", "0"), w = num("width", "0"), h = num("height", "0");
    if (w == 0 || h == 0) return;
    double rx = num("rx", "0"), ry = num("ry", "0");
    rx = rx > 0 ? rx : 0;  // max(v, 0)
    ry = ry > 0 ? ry : 0;
    PathB b;
    auto f = [](double v) { return (float)v; };
    if (rx > 0 || ry > 0) {
      if (rx == 0)
        rx = ry;
      else if (ry == 0)
        ry = rx;
      rx = rx <= w / 2 ? rx : w / 2;  // min
      ry = ry <= h / 2 ? ry : h / 2;
      b.move(f(x + rx), f(y));
      b.line(f(x + w - rx), f(y));
      b.arc(f(rx), f(ry), 0, false, true, f(x + w), f(y + ry));
      b.line(f(x + w), f(y + h - ry));
      b.arc(f(rx), f(ry), 0, false, true, f(x + w - rx), f(y + h));
      b.line(f(x + rx), f(y + h));
      b.arc(f(rx), f(ry), 0, false, true, f(x), f(y + h - ry));
      b.line(f(x), f(y + ry));
      b.arc(f(rx), f(ry), 0, false, true, f(x + rx), f(y));
    } else {
      b.rect(f(x), f(y), f(w), f(h));
    }
    svg->elements.push_back({std::move(b.c), std::move(p)});
    return;
  }
  if (tag == "circle" || tag == "ellipse") {
    Props p = parse_props(node, stack.back());
    const double cx = num("cx", "0"), cy = num("cy", "0");
    float rx, ry;
    if (tag == "circle") {
      rx = (float)parse_float(node.attr("r"));
      ry = rx;
    } else {
      rx = (float)num("rx", "0");
      ry = (float)num("ry", "0");
    }
    PathB b;
    b.ellipse((float)cx, (float)cy, rx, ry);
    svg->elements.push_back({std::move(b.c), std::move(p)});
    return;
  }
  if (tag == "linearGradient") {
    parse_props(node, stack.back());
    const std::string& units = node.attr("gradientUnits");
    const std::string& gt = node.attr("gradientTransform");
    if (units != "userSpaceOnUse") throw SvgError("Unsupported gradient units: " + units);
    if (!gt.empty()) throw SvgError("Unsupported gradient transform: " + gt);
    LinearGradient g;
    g.x1 = (float)parse_float(node.attr("x1"));
    g.y1 = (float)parse_float(node.attr("y1"));
    g.x2 = (float)parse_float(node.attr("x2"));
    g.y2 = (float)parse_float(node.attr("y2"));
    for (const XNode& k : node.kids) {
      if (!k.element) continue;
      if (k.tag != "stop") throw SvgError("Unexpected SVG tag: " + k.tag);
      std::string color = k.attr("stop-color");
      if (color.empty())
        for (const auto& pair : split(k.attr("style"), ';')) {
          const auto parts = split(pair, ':');
          if (parts.size() == 2 && strip(parts[0]) == "stop-color" && color.empty()) color = strip(parts[1]);
        }
      if (color.empty()) throw SvgError("Invalid SVG gradient, missing stop-color");
      const Color c = parse_color(color);
      g.stops.insert(g.stops.end(), {c.r, c.g, c.b, c.a, (float)parse_float(k.attr("offset"))});
    }
    svg->gradients[node.attr("id")] = std::move(g);
    return;
  }
  throw SvgError("Unsupported SVG tag: " + tag);
}

void parse_svg(pxh_svg* svg, const char* data, size_t len, int32_t width, int32_t height) {  // svg.nim:502-555
  XmlParser xp{data, len};
  const XNode root = xp.document();
  if (root.tag != "svg") invalid();
  const auto box = split(root.attr("viewBox"), ' ');
  if (box.size() < 4) throw SvgError("Invalid SVG data: viewBox");
  const int64_t minX = parse_int(box[0]), minY = parse_int(box[1]);
  const int64_t vbW = parse_int(box[2]), vbH = parse_int(box[3]);
  Props rootProps = parse_props(root, Props());
  if (minX != 0 || minY != 0) {
    const float t[9] = {1, 0, 0, 0, 1, 0, -(float)minX, -(float)minY, 1};
    mul_into(rootProps.transform, t);
  }
  if (width == 0 && height == 0) {
    svg->width = (int32_t)vbW;
    svg->height = (int32_t)vbH;
  } else {
    svg->width = width;
    svg->height = height;
    const float sx = (float)width / (float)vbW, sy = (float)height / (float)vbH;
    const float s[9] = {sx, 0, 0, 0, sy, 0, 0, 0, 1};
    mul_into(rootProps.transform, s);
  }
  std::vector<Props> stack{rootProps};
  for (const XNode& k : root.kids) parse_element(k, svg, stack);
}

}  // namespace

extern "C" {

pxh_svg* pxh_svg_parse(const char* data, int64_t len, int32_t width, int32_t height) {
  if (!data || len < 0 || width < 0 || height < 0) {
    set_last_error(kErrArg, "invalid SVG arguments");
    return nullptr;
  }
  pxh_svg* svg = new (std::nothrow) pxh_svg();
  if (!svg) {
    set_last_error(kErrNoMem, "out of host memory");
    return nullptr;
  }
  try {
    std::string copy(data, (size_t)len);  // NUL-terminated for the scanners
    parse_svg(svg, copy.c_str(), copy.size(), width, height);
  } catch (const SvgError& e) {
    set_last_error(kErrParse, e.what());
    delete svg;
    return nullptr;
  } catch (const std::bad_alloc&) {
    set_last_error(kErrNoMem, "out of host memory");
    delete svg;
    return nullptr;
  }
  return svg;
}

void pxh_svg_free(pxh_svg* svg) {
  if (!svg) return;
  if (svg->scene) pxh_scene_free(svg->scene);
  delete svg;
}

int pxh_svg_size(const pxh_svg* svg, int32_t* width, int32_t* height) {
  if (!svg) return set_last_error(kErrArg, "invalid svg");
  if (width) *width = svg->width;
  if (height) *height = svg->height;
  return 0;
}

int32_t pxh_svg_count(const pxh_svg* svg) { return svg ? (int32_t)svg->elements.size() : -1; }

int pxh_svg_element(const pxh_svg* svg, int32_t i, float** cmds, int32_t* ncmds, pxh_svg_props* props) {
  if (!svg || i < 0 || i >= (int32_t)svg->elements.size()) return set_last_error(kErrArg, "invalid element");
  const Element& e = svg->elements[i];
  if (cmds && ncmds) {
    *ncmds = (int32_t)e.cmds.size();
    *cmds = nullptr;
    if (!e.cmds.empty()) {
      *cmds = (float*)malloc(e.cmds.size() * sizeof(float));
      if (!*cmds) return set_last_error(kErrNoMem, "out of host memory");
      memcpy(*cmds, e.cmds.data(), e.cmds.size() * sizeof(float));
    }
  }
  if (props) {
    const Props& p = e.props;
    props->display = p.display;
    props->fill_rule = p.fillRule;
    snprintf(props->fill, sizeof(props->fill), "%s", p.fill.c_str());
    props->stroke_rgbx = p.stroke;
    props->stroke_width = p.strokeWidth;
    props->line_cap = p.cap;
    props->line_join = p.join;
    props->miter_limit = p.miterLimit;
    props->ndashes = (int32_t)p.dashes.size();
    props->dashes = p.dashes.empty() ? nullptr : p.dashes.data();
    memcpy(props->transform, p.transform, sizeof(props->transform));
    props->opacity = p.opacity;
    props->fill_opacity = p.fillOpacity;
    props->stroke_opacity = p.strokeOpacity;
  }
  return 0;
}

// newImage(svg) (svg.nim:557-602) onto img: cleared to transparent, then
// each displayed element's fill (the first one overwriting) and stroke in
// order.  Solid fills and strokes are recorded into one scene; a gradient
// fill flushes the scene and takes the per-call paint route.  A scene pass
// that runs out of a pool leaves the canvas untouched; its calls are then
// issued one by one.
int pxh_svg_render(pxh_svg* svg, pxh_image* img) {
  if (!svg) return set_last_error(kErrArg, "invalid svg");
  if (pxh_image_width(img) != svg->width || pxh_image_height(img) != svg->height ||
      pxh_image_format(img) != PXH_FORMAT_RGBA8)
    return set_last_error(kErrArg, "the image must be RGBA8 of the svg's size");
  if (int r = pxh_image_fill(img, 0)) return r;
  if (!svg->scene) {
    svg->scene = pxh_scene_new();
    if (!svg->scene) return set_last_error(kErrNoMem, "scene allocation failed");
  }
  pxh_scene* sc = svg->scene;
  pxh_scene_clear(sc);
  struct Call {
    const Element* e;
    bool stroke;
    uint32_t rgbx;
    int32_t blend;
  };
  std::vector<Call> pending;
  auto one_by_one = [&](const Call& c) -> int {
    const Props& p = c.e->props;
    const int32_t n = (int32_t)c.e->cmds.size();
    if (c.stroke)
      return pxh_stroke_path(img, c.e->cmds.data(), n, p.transform, c.rgbx, p.strokeWidth, p.cap, p.join,
                             p.miterLimit, p.dashes.data(), (int32_t)p.dashes.size(), c.blend);
    return pxh_fill_path(img, c.e->cmds.data(), n, p.transform, c.rgbx, p.fillRule, c.blend);
  };
  auto flush = [&]() -> int {
    const int32_t n = pxh_scene_count(sc);
    if (n == 0) return 0;
    int r = pxh_scene_render(sc, img);
    if (r == kErrCapacity) {
      for (const Call& c : pending)
        if (int q = one_by_one(c)) return q;
    } else if (r) {
      return r;
    } else {
      std::vector<int32_t> st((size_t)n);
      if (int q = pxh_scene_status(sc, st.data(), n)) return q;
      for (int32_t k = 0; k < n; k++)
        if (st[k]) return one_by_one(pending[k]);  // raises the call's own error
    }
    pxh_scene_clear(sc);
    pending.clear();
    return 0;
  };
  auto record = [&](const Call& c) -> int {
    const Props& p = c.e->props;
    const int32_t n = (int32_t)c.e->cmds.size();
    int r;
    if (c.stroke)
      r = pxh_scene_stroke_path(sc, c.e->cmds.data(), n, p.transform, c.rgbx, p.strokeWidth, p.cap, p.join,
                                p.miterLimit, p.dashes.data(), (int32_t)p.dashes.size(), c.blend);
    else
      r = pxh_scene_fill_path(sc, c.e->cmds.data(), n, p.transform, c.rgbx, p.fillRule, c.blend);
    if (r) return r;
    if (pxh_scene_count(sc) > (int32_t)pending.size()) pending.push_back(c);
    return 0;
  };
  try {
    int32_t blend = kOverwrite;
    for (const Element& e : svg->elements) {
      const Props& p = e.props;
      if (!(p.display && p.opacity > 0.0f)) continue;
      if (p.fill != "none") {
        const float opacity = p.fillOpacity * p.opacity;
        if (starts_with(p.fill, "url(")) {
          const size_t close = p.fill.find(')', 5);
          if (close == std::string::npos) throw SvgError("Malformed fill: " + p.fill);
          const std::string id = p.fill.substr(5, close - 5);
          auto it = svg->gradients.find(id);
          if (it == svg->gradients.end()) throw SvgError("Missing SVG resource " + id);
          const LinearGradient& g = it->second;
          pxh_paint paint = {};
          paint.kind = PXH_LINEAR_GRADIENT;
          paint.opacity = opacity;
          paint.blend_mode = blend;
          paint.nhandles = 2;
          m3_apply(p.transform, g.x1, g.y1, &paint.handles[0], &paint.handles[1]);
          m3_apply(p.transform, g.x2, g.y2, &paint.handles[2], &paint.handles[3]);
          paint.nstops = (int32_t)(g.stops.size() / 5);
          paint.stops = g.stops.data();
          if (int r = flush()) return r;
          if (int r = pxh_fill_path_paint(img, e.cmds.data(), (int32_t)e.cmds.size(), p.transform, &paint,
                                          p.fillRule))
            return r;
        } else {
          // the colour as the fill paint holds it (ColorRGBX -> Color), then
          // fillPath's solid route (paths.nim:2100-2113)
          const Color c = rgbx_color(color_rgbx(parse_color(p.fill)));
          float o = opacity;
          if (!(o > 0.0f)) o = 0.0f;
          if (o > 1.0f) o = 1.0f;
          if (o != 0.0f && (c.a > 0.0f || blend == kOverwrite))
            if (int r = record({&e, false, color_rgbx({c.r, c.g, c.b, c.a * o}), blend})) return r;
        }
      }
      blend = kNormal;
      if (p.stroke != 0 && p.strokeWidth > 0.0f) {
        Color c = rgbx_color(p.stroke);
        c.a *= p.opacity * p.strokeOpacity;
        if (c.a > 0.0f)
          if (int r = record({&e, true, color_rgbx(c), kNormal})) return r;
      }
    }
  } catch (const SvgError& e) {
    return set_last_error(kErrParse, e.what());
  }
  return flush();
}

}  // extern "C"
