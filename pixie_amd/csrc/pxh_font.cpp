// pxh_font.cpp -- text ingestion for the fill path: OpenType fonts with
// TrueType (glyf) outlines parsed into pixie path commands, a glyph path
// cache per typeface, typesetting, and the span paths that fillText /
// strokeText hand to fillPath / strokePath (fonts.nim,
// fontformats/opentype.nim).
//
// Everything here runs on the host: parsing and layout are byte and scalar
// work whose result is a path command array, which then takes the device
// path of pxh_fill_path_paint / pxh_stroke_path_paint (device geometry for
// large spans).  Float arithmetic is float32 in pixie's operation order
// (built with -ffp-contract=off); roundf rounds half away from zero as Nim's
// round does.

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/pixie_hip.h"
#include "pxh_common.h"
#include "pxh_geom.h"
#include "pxh_host_geom.h"

namespace {

using namespace pxh;

struct FontError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

[[noreturn]] void unsupported(const std::string& what) {  // opentype.nim:401-402
  throw FontError("Unsupported font " + what);
}

// Big-endian reads with pixie's EOF error (opentype.nim:397-399).
struct Buf {
  const uint8_t* p = nullptr;
  int64_t n = 0;
  void check(int64_t readTo) const {
    if (readTo > n || readTo < 0) throw FontError("Unexpected error reading font data, EOF");
  }
  uint8_t u8(int64_t i) const {
    check(i + 1);
    if (i < 0) check(-1);
    return p[i];
  }
  int8_t i8(int64_t i) const { return (int8_t)u8(i); }
  uint16_t u16(int64_t i) const {
    check(i + 2);
    if (i < 0) check(-1);
    return (uint16_t)((p[i] << 8) | p[i + 1]);
  }
  int16_t i16(int64_t i) const { return (int16_t)u16(i); }
  uint32_t u32(int64_t i) const {
    check(i + 4);
    if (i < 0) check(-1);
    return ((uint32_t)p[i] << 24) | ((uint32_t)p[i + 1] << 16) | ((uint32_t)p[i + 2] << 8) | p[i + 3];
  }
};

// readVersion16Dot16 (opentype.nim:422-431).
float read_version(const Buf& b, int64_t i) {
  const uint32_t d = b.u32(i);
  const uint32_t major = (d & 0x000F0000u) >> 16, minor = (d & 0x0000F000u) >> 12;
  if (major > 9 || minor > 9) unsupported("invalid version format");
  return (float)major + (float)minor / 10.0f;
}

// readFixed16 (opentype.nim:413-415): F2Dot14.
float read_f2dot14(const Buf& b, int64_t i) { return (float)b.i16(i) / 16384.0f; }

inline uint32_t pair_key(uint16_t a, uint16_t b) { return ((uint32_t)a << 16) | b; }

// One GPOS PairPos subtable with adjustments (opentype.nim:1952-2083).
struct PairPosTable {
  int32_t format = 0;
  std::vector<uint64_t> covered = std::vector<uint64_t>(1024, 0);  // coverage bit per glyph id
  std::unordered_map<uint32_t, int16_t> glyphPairs;                 // format 1
  std::unordered_map<uint16_t, uint16_t> class1, class2;            // format 2
  std::unordered_map<uint32_t, int16_t> classPairs;
  bool covers(uint16_t g) const { return (covered[g >> 6] >> (g & 63)) & 1; }
  void cover(uint16_t g) { covered[g >> 6] |= 1ull << (g & 63); }
};

constexpr int kMaxComponentDepth = 32;

}  // namespace

struct pxh_typeface {
  std::vector<uint8_t> data;
  Buf buf;
  std::unordered_map<uint32_t, uint16_t> cmap;  // runeToGlyphId
  uint16_t unitsPerEm = 0;
  int16_t indexToLocFormat = 0;
  int16_t ascender = 0, descender = 0, lineGap = 0;
  uint16_t numberOfHMetrics = 0, numGlyphs = 0;
  std::vector<uint16_t> advances;  // hMetrics advanceWidth
  std::vector<uint32_t> glyf;      // glyph data offsets (glyf table + loca)
  bool hasGlyf = false;
  // CFF outlines: charstring, global and local subroutine ranges [a, b) in
  // data; CID-keyed fonts pick their local subroutines per glyph (FDSelect)
  using Ranges = std::vector<std::pair<uint32_t, uint32_t>>;
  struct CffPrivate {
    Ranges subrs;
  };
  bool isCff = false;
  Ranges cffChars, cffGsubrs;
  CffPrivate cffTop;
  bool cffCid = false;
  std::vector<CffPrivate> cffFds;
  std::vector<uint8_t> cffFdSelect;
  bool hasKern = false;
  std::unordered_map<uint32_t, float> kernPairs;
  bool hasGpos = false;
  std::vector<PairPosTable> pairPos;
  int16_t underlinePosition = 0, underlineThickness = 0;
  int16_t strikeoutPosition = 0, strikeoutSize = 0, capHeight = 0;
  std::string fullName;
  std::unordered_map<uint32_t, std::vector<float>> glyphPaths;  // getGlyphPath cache
  std::vector<pxh_typeface*> fallbacks;
};

struct pxh_arrangement {
  std::vector<uint32_t> runes;
  std::vector<std::pair<int32_t, int32_t>> lines, spans;
  std::vector<const pxh_font*> fonts;
  std::vector<float> pos;    // x, y per rune
  std::vector<float> rects;  // x, y, w, h per rune
};

namespace {

// ---- tables (opentype.nim:433-860, 1768-2174) ----

void parse_cmap(pxh_typeface* t, int64_t offset) {  // opentype.nim:433-565
  const Buf& b = t->buf;
  int64_t i = offset;
  b.check(i + 4);
  const uint16_t numTables = b.u16(i + 2);
  i += 4;
  for (int j = 0; j < numTables; j++) {
    b.check(i + 8);
    const uint16_t platformID = b.u16(i);
    const uint32_t recOffset = b.u32(i + 4);
    i += 8;
    if (platformID != 3) continue;  // Windows records only
    int64_t k = offset + (int64_t)recOffset;
    b.check(k + 2);
    const uint16_t format = b.u16(k);
    if (format == 4) {
      b.check(k + 14);
      const int64_t segCount = b.u16(k + 6) / 2;
      k += 14;
      b.check(k + 2 + 4 * segCount * 2);
      const int64_t endPos = k, startPos = k + segCount * 2 + 2, deltaPos = startPos + segCount * 2,
                    rangePos = deltaPos + segCount * 2;
      for (int64_t s = 0; s < segCount; s++) {
        const uint32_t endCode = b.u16(endPos + 2 * s), startCode = b.u16(startPos + 2 * s);
        const int64_t idDelta = b.u16(deltaPos + 2 * s), idRangeOffset = b.u16(rangePos + 2 * s);
        for (uint32_t c = startCode; c <= endCode; c++) {
          int64_t glyphId;
          if (idRangeOffset != 0) {
            const int64_t at = rangePos + s * 2 + idRangeOffset + (int64_t)(c - startCode) * 2;
            b.check(at + 2);
            glyphId = b.u16(at);
            if (glyphId != 0) glyphId = (glyphId + idDelta) & 0xFFFF;
          } else {
            glyphId = ((int64_t)c + idDelta) & 0xFFFF;
          }
          if (c != 65535) t->cmap[c] = (uint16_t)glyphId;
        }
      }
    } else if (format == 12) {
      b.check(k + 16);
      const uint32_t numGroups = b.u32(k + 12);
      k += 16;
      b.check(k + (int64_t)numGroups * 12);
      for (uint32_t gi = 0; gi < numGroups; gi++) {
        const uint32_t startChar = b.u32(k), endChar = b.u32(k + 4), startGlyph = b.u32(k + 8);
        for (uint64_t c = startChar; c <= endChar; c++) {
          const uint64_t glyphId = startGlyph + (c - startChar);
          if (glyphId > 0xFFFF)
            throw FontError("Found glyph outside of uint16 range: " + std::to_string(glyphId));
          t->cmap[(uint32_t)c] = (uint16_t)glyphId;
        }
        k += 12;
      }
    }
  }
}

void parse_head(pxh_typeface* t, int64_t o) {  // opentype.nim:567-594
  const Buf& b = t->buf;
  b.check(o + 54);
  if (b.u16(o) != 1) unsupported("Head major version");
  if (b.u16(o + 2) != 0) unsupported("Head minor version");
  t->unitsPerEm = b.u16(o + 18);
  t->indexToLocFormat = b.i16(o + 50);
  if (b.i16(o + 52) != 0) unsupported("Head glyph data format");
}

void parse_hhea(pxh_typeface* t, int64_t o) {  // opentype.nim:596-623
  const Buf& b = t->buf;
  b.check(o + 36);
  if (b.u16(o) != 1) unsupported("Hhea major version");
  if (b.u16(o + 2) != 0) unsupported("Hhea minor version");
  t->ascender = b.i16(o + 4);
  t->descender = b.i16(o + 6);
  t->lineGap = b.i16(o + 8);
  if (b.i16(o + 32) != 0) unsupported("Hhea metric data format");
  t->numberOfHMetrics = b.u16(o + 34);
}

void parse_maxp(pxh_typeface* t, int64_t o) {  // opentype.nim:625-649
  const Buf& b = t->buf;
  b.check(o + 32);
  const float v = read_version(b, o);
  if (v != 0.5f && v != 1.0f) unsupported("Maxp version");
  t->numGlyphs = b.u16(o + 4);
}

void parse_hmtx(pxh_typeface* t, int64_t o) {  // opentype.nim:651-672
  const Buf& b = t->buf;
  const int64_t hMetricsSize = (int64_t)t->numberOfHMetrics * 4;
  const int64_t lsbSize = (int64_t)(uint16_t)(t->numGlyphs - t->numberOfHMetrics) * 2;
  b.check(o + hMetricsSize + lsbSize);
  for (int g = 0; g < t->numGlyphs && g < t->numberOfHMetrics; g++) t->advances.push_back(b.u16(o + 4 * g));
}

void parse_name(pxh_typeface* t, int64_t o) {  // opentype.nim:674-709
  const Buf& b = t->buf;
  b.check(o + 6);
  if (b.u16(o) != 0) unsupported("Name format");
  const uint16_t count = b.u16(o + 2), stringOffset = b.u16(o + 4);
  int64_t i = o + 6;
  b.check(i + (int64_t)count * 12);
  bool found = false;
  for (int j = 0; j < count; j++, i += 12) {
    const uint16_t platformID = b.u16(i), encodingID = b.u16(i + 2), languageID = b.u16(i + 4),
                   nameID = b.u16(i + 6), length = b.u16(i + 8), off = b.u16(i + 10);
    const int64_t s = o + stringOffset + off;
    if (s < 0 || s + length > b.n) throw FontError("Unexpected error reading font data, EOF");
    if (found || nameID != 6 || languageID != 1033) continue;
    std::string text((const char*)b.p + s, (size_t)length);
    if (platformID == 3 && encodingID == 1) {  // UTF-16BE -> UTF-8
      std::string u;
      for (size_t k = 0; k + 1 < text.size(); k += 2) {
        uint32_t c = ((uint8_t)text[k] << 8) | (uint8_t)text[k + 1];
        if (c >= 0xD800 && c < 0xDC00 && k + 3 < text.size()) {
          const uint32_t lo = ((uint8_t)text[k + 2] << 8) | (uint8_t)text[k + 3];
          if (lo >= 0xDC00 && lo < 0xE000) {
            c = 0x10000 + ((c - 0xD800) << 10) + (lo - 0xDC00);
            k += 2;
          }
        }
        if (c < 0x80) {
          u += (char)c;
        } else if (c < 0x800) {
          u += (char)(0xC0 | (c >> 6));
          u += (char)(0x80 | (c & 63));
        } else if (c < 0x10000) {
          u += (char)(0xE0 | (c >> 12));
          u += (char)(0x80 | ((c >> 6) & 63));
          u += (char)(0x80 | (c & 63));
        } else {
          u += (char)(0xF0 | (c >> 18));
          u += (char)(0x80 | ((c >> 12) & 63));
          u += (char)(0x80 | ((c >> 6) & 63));
          u += (char)(0x80 | (c & 63));
        }
      }
      text = u;
    }
    t->fullName = text;
    found = true;
  }
}

void parse_os2(pxh_typeface* t, int64_t o) {  // opentype.nim:711-771
  const Buf& b = t->buf;
  b.check(o + 78);
  const uint16_t version = b.u16(o);
  t->strikeoutSize = b.i16(o + 26);
  t->strikeoutPosition = b.i16(o + 28);
  int64_t i = o + 78;
  if (version >= 1) {
    b.check(i + 8);
    i += 8;
  }
  if (version >= 2) {
    b.check(i + 10);
    t->capHeight = b.i16(i + 2);
    i += 10;
  }
  if (version >= 5) b.check(i + 4);
}

// ---- CFF (opentype.nim:862-1030, 1299-1678) ----

// An INDEX at `start`: the [a, b) ranges of its objects; start moves past it
// (parseCFFIndexOffsets 862-906).
pxh_typeface::Ranges cff_index(const Buf& b, int64_t* start) {
  pxh_typeface::Ranges out;
  const int count = b.u16(*start);
  if (count == 0) {
    *start += 2;
    return out;
  }
  const int offSize = b.u8(*start + 2);
  const int64_t objects = *start + (int64_t)(count + 1) * offSize + 2;
  std::vector<int64_t> offs(count + 1);
  int64_t pos = *start + 3;
  for (int i = 0; i <= count; i++, pos += offSize) {
    int64_t v = 0;
    for (int k = 0; k < offSize; k++) v = (v << 8) | b.u8(pos + k);
    offs[i] = v;
  }
  for (int i = 0; i < count; i++) {
    const int64_t a = objects + offs[i], e = objects + offs[i + 1];
    if (a < 0 || e < a || e > b.n) throw FontError("Unexpected error reading font data, EOF");
    out.push_back({(uint32_t)a, (uint32_t)e});
  }
  *start = objects + offs[count];
  return out;
}

// A DICT (parseCFFDict 1322-1407): operator -> operands, for the operators
// the reference keeps (interpretDict: unknown ones are skipped, a repeated
// known one fails).
std::unordered_map<int, std::vector<double>> cff_dict(const Buf& b, int64_t a, int64_t e,
                                                      const std::vector<int>& known) {
  std::unordered_map<int, std::vector<double>> out;
  std::vector<double> operands;
  int64_t i = a;
  while (i < e) {
    int op = b.u8(i++);
    if (op <= 21) {
      if (op == 12) op = 1200 + b.u8(i++);
      if (std::find(known.begin(), known.end(), op) != known.end()) {
        if (out.count(op)) unsupported("CFF duplicate key");
        out[op] = operands;
      }
      operands.clear();
      continue;
    }
    double v;
    if (op == 28) {  // read unsigned, as the reference does
      v = (double)((b.u8(i) << 8) | b.u8(i + 1));
      i += 2;
    } else if (op == 29) {
      v = (double)(int32_t)b.u32(i);
      i += 4;
    } else if (op == 30) {  // real: nibbles
      static const char* const nib[15] = {"0", "1", "2", "3", "4", "5", "6", "7",
                                          "8", "9", ".", "E", "E-", "null", "-"};
      std::string str;
      for (;;) {
        const int byte = b.u8(i++);
        const int n1 = byte >> 4, n2 = byte & 15;
        if (n1 == 15) break;
        str += nib[n1];
        if (n2 == 15) break;
        str += nib[n2];
      }
      char* end = nullptr;
      v = strtod(str.c_str(), &end);
      if (str.empty() || *end) unsupported("float " + str);
    } else if (op >= 32 && op <= 246) {
      v = op - 139;
    } else if (op >= 247 && op <= 250) {
      v = (op - 247) * 256 + b.u8(i++) + 108;
    } else if (op >= 251 && op <= 254) {
      v = -(op - 251) * 256 - b.u8(i++) - 108;
    } else {
      unsupported("Invalid b0 " + std::to_string(op));
    }
    operands.push_back(v);
  }
  return out;
}

int dict_int(const std::unordered_map<int, std::vector<double>>& d, int key, int dflt, size_t k = 0) {
  auto it = d.find(key);
  if (it == d.end() || it->second.size() <= k) return dflt;
  return (int)it->second[k];
}

// A Top DICT's (or FD's) private DICT and its local subroutines
// (gatherCFFTopDicts 1528-1541): Subrs is relative to the private DICT.
pxh_typeface::CffPrivate cff_private(const Buf& b, int64_t cff, const std::unordered_map<int, std::vector<double>>& top) {
  pxh_typeface::CffPrivate pr;
  const int size = dict_int(top, 18, 0, 0), off = dict_int(top, 18, 0, 1);
  if (off != 0) {
    const auto priv = cff_dict(b, cff + off, cff + off + size, {19, 20, 21});
    const int subrs = dict_int(priv, 19, 0);
    if (subrs > 0) {
      int64_t at = cff + off + subrs;
      pr.subrs = cff_index(b, &at);
    }
  }
  return pr;
}

void parse_cff(pxh_typeface* t, int64_t offset) {  // parseCFFTable 1299-1678
  const Buf& b = t->buf;
  b.check(offset + 32);
  int64_t at = offset + 4;
  cff_index(b, &at);  // names
  const auto topIndex = cff_index(b, &at);
  const auto strings = cff_index(b, &at);
  t->cffGsubrs = cff_index(b, &at);
  if (topIndex.empty()) unsupported("CFF no topDict");
  if (topIndex.size() > 1) unsupported("CFF multiple topDict");
  const std::vector<int> topKeys = {0, 1, 2, 3, 4, 5, 13, 14, 15, 16, 17, 18, 1200, 1201, 1202, 1203, 1204,
                                    1205, 1206, 1207, 1208, 1230, 1231, 1232, 1233, 1234, 1235, 1236, 1237, 1238};
  const auto top = cff_dict(b, topIndex[0].first, topIndex[0].second, topKeys);
  t->cffTop = cff_private(b, offset, top);
  // CID-keyed when the registry and ordering strings are not empty
  auto sid_nonempty = [&](int sid) {
    if (sid <= 390) return true;  // standard strings
    return sid - 391 < (int)strings.size() && strings[sid - 391].second > strings[sid - 391].first;
  };
  auto ros = top.find(1230);
  if (ros != top.end() && ros->second.size() >= 2 && sid_nonempty((int)ros->second[0]) &&
      sid_nonempty((int)ros->second[1])) {
    t->cffCid = true;
    int64_t fdArray = offset + dict_int(top, 1236, 0), fdSelect = offset + dict_int(top, 1237, 0);
    if (fdArray == 0 || fdSelect == 0) unsupported("CFF CID tables are missing");
    for (const auto& fd : cff_index(b, &fdArray))
      t->cffFds.push_back(cff_private(b, offset, cff_dict(b, fd.first, fd.second, topKeys)));
    // FDSelect, format 3 only (parseCFFFDSelect 1595-1624)
    int64_t i = fdSelect;
    if (b.u8(i++) != 3) unsupported("CFF Table CID Font FDSelect format");
    const int nRanges = b.u16(i);
    int first = b.u16(i + 2), next = 0;
    i += 4;
    if (first != 0) unsupported("CFF Table CID Font FDSelect format 3 range has bad initial GID");
    for (int r = 0; r < nRanges; r++, i += 3) {
      const int fd = b.u8(i);
      next = b.u16(i + 1);
      if (fd >= (int)t->cffFds.size()) unsupported("CFF table CID Font FDSelect has bad FD index value");
      if (next > t->numGlyphs) unsupported("CFF Table CID Font FDSelect format 3 range has bad GID");
      for (; first < next; first++) t->cffFdSelect.push_back((uint8_t)fd);
      first = next;
    }
    if (next != t->numGlyphs) unsupported("CFF Table CID Font FDSelect format 3 range has bad final GID");
  }
  int64_t cs = offset + dict_int(top, 17, 0);
  t->cffChars = cff_index(b, &cs);
  t->isCff = true;
}

void parse_loca_glyf(pxh_typeface* t, int64_t loca, int64_t glyf) {  // opentype.nim:773-798
  const Buf& b = t->buf;
  if (t->indexToLocFormat == 0) {
    b.check(loca + (int64_t)t->numGlyphs * 2);
    for (int g = 0; g < t->numGlyphs; g++) t->glyf.push_back((uint32_t)(glyf + (uint32_t)b.u16(loca + 2 * g) * 2));
  } else {
    b.check(loca + (int64_t)t->numGlyphs * 4);
    for (int g = 0; g < t->numGlyphs; g++) t->glyf.push_back((uint32_t)(glyf + b.u32(loca + 4 * g)));
  }
  t->hasGlyf = true;
}

// parseKernTable (opentype.nim:800-860), including its accumulation: after
// each subtable is read, the pairs of every subtable read so far are applied
// again (non-override subtables add to what the map holds).
void parse_kern(pxh_typeface* t, int64_t o) {
  const Buf& b = t->buf;
  b.check(o + 2);
  const uint16_t version = b.u16(o);
  int64_t i = o + 2;
  if (version == 1) return;  // Mac format: no kern table
  if (version != 0) unsupported("Kern version");
  b.check(i + 2);
  const uint16_t nTables = b.u16(i);
  i += 2;
  struct Sub {
    uint16_t coverage;
    std::vector<std::pair<uint32_t, int16_t>> pairs;
  };
  std::vector<Sub> subs;
  t->hasKern = true;
  for (int s = 0; s < nTables; s++) {
    b.check(i + 14);
    if (b.u16(i) != 0) unsupported("Kern subtable version");
    Sub sub;
    sub.coverage = b.u16(i + 4);
    if ((sub.coverage >> 8) != 0) unsupported("Kern subtable coverage");
    const uint16_t nPairs = b.u16(i + 6);
    i += 14;
    for (int k = 0; k < nPairs; k++, i += 6) {
      b.check(i + 6);
      sub.pairs.push_back({pair_key(b.u16(i), b.u16(i + 2)), b.i16(i + 4)});
    }
    subs.push_back(std::move(sub));
    for (const Sub& tb : subs) {
      if ((tb.coverage & 1) == 0) continue;
      for (const auto& pr : tb.pairs) {
        if (pr.second == 0) continue;
        float value = (float)pr.second;
        if ((tb.coverage & 0x8) == 0) {
          auto it = t->kernPairs.find(pr.first);
          value += it == t->kernPairs.end() ? 0.0f : it->second;
        }
        t->kernPairs[pr.first] = value;
      }
    }
  }
}

int value_format_size(uint16_t f) {  // opentype.nim:1813-1821
  int n = 0;
  while (f) {
    f &= f - 1;
    n++;
  }
  return n * 2;
}

// parseValueRecord (opentype.nim:1823-1852): the xAdvance field, 0 if absent.
int16_t value_x_advance(const Buf& b, int64_t o, uint16_t f) {
  b.check(o + value_format_size(f));
  if ((f & 0x4) == 0) return 0;
  int64_t i = o;
  if (f & 0x1) i += 2;
  if (f & 0x2) i += 2;
  return b.i16(i);
}

struct Coverage {  // opentype.nim:1775-1811
  int32_t format = 0;
  std::vector<uint16_t> glyphs;                            // format 1
  std::vector<std::pair<std::pair<uint16_t, uint16_t>, uint16_t>> ranges;  // format 2
};

Coverage parse_coverage(const Buf& b, int64_t o) {
  Coverage c;
  b.check(o + 4);
  c.format = b.u16(o);
  int64_t i = o + 2;
  if (c.format == 1) {
    const uint16_t n = b.u16(i);
    i += 2;
    b.check(i + (int64_t)n * 2);
    for (int k = 0; k < n; k++) c.glyphs.push_back(b.u16(i + 2 * k));
  } else if (c.format == 2) {
    const uint16_t n = b.u16(i);
    i += 2;
    for (int k = 0; k < n; k++, i += 6) {
      b.check(i + 6);
      c.ranges.push_back({{b.u16(i), b.u16(i + 2)}, b.u16(i + 4)});
    }
  } else {
    unsupported("coverage format");
  }
  return c;
}

void add_coverage(PairPosTable* t, const Coverage& c) {
  for (uint16_t g : c.glyphs) t->cover(g);
  for (const auto& r : c.ranges)
    for (uint32_t g = r.first.first; g <= r.first.second; g++) t->cover((uint16_t)g);
}

// parseClassDef (opentype.nim:1920-1950) into the glyph -> class map of
// parsePairPos (opentype.nim:2041-2070).
void parse_class_def(const Buf& b, int64_t o, std::unordered_map<uint16_t, uint16_t>* m) {
  b.check(o + 2);
  const uint16_t format = b.u16(o);
  int64_t i = o + 2;
  if (format == 1) {
    b.check(i + 4);
    const uint16_t start = b.u16(i), count = b.u16(i + 2);
    i += 4;
    b.check(i + (int64_t)count * 2);
    for (uint32_t k = 0; k < count; k++) {
      const uint16_t v = b.u16(i + 2 * k);
      if (v != 0) (*m)[(uint16_t)(start + k)] = v;
    }
  } else if (format == 2) {
    b.check(i + 2);
    const uint16_t n = b.u16(i);
    i += 2;
    std::vector<uint16_t> rec;
    for (int k = 0; k < n; k++, i += 6) {
      b.check(i + 6);
      rec.push_back(b.u16(i));
      rec.push_back(b.u16(i + 2));
      rec.push_back(b.u16(i + 4));
    }
    for (size_t k = 0; k < rec.size(); k += 3) {
      if (rec[k] > rec[k + 1]) unsupported("glyph out of bounds");
      if (rec[k + 2] != 0)
        for (uint32_t g = rec[k]; g <= rec[k + 1]; g++) (*m)[(uint16_t)g] = rec[k + 2];
    }
  } else {
    unsupported("class format");
  }
}

PairPosTable parse_pair_pos(const Buf& b, int64_t o) {  // opentype.nim:1952-2083
  PairPosTable t;
  b.check(o + 4);
  t.format = b.u16(o);
  int64_t i = o + 2;
  if (t.format == 1) {
    b.check(i + 8);
    const uint16_t coverageOffset = b.u16(i), vf1 = b.u16(i + 2), vf2 = b.u16(i + 4),
                   pairSetCount = b.u16(i + 6);
    i += 8;
    b.check(i + 2 * (int64_t)pairSetCount);
    // pair sets: (second glyph, xAdvance of the first value record)
    std::vector<std::vector<std::pair<uint16_t, int16_t>>> sets(pairSetCount);
    const int64_t recSize = 2 + value_format_size(vf1) + value_format_size(vf2);
    for (int s = 0; s < pairSetCount; s++) {
      int64_t k = o + b.u16(i + 2 * s);
      b.check(k + 2);
      const uint16_t n = b.u16(k);
      k += 2;
      for (int r = 0; r < n; r++, k += recSize) {
        b.check(k + 2);
        const uint16_t second = b.u16(k);
        const int16_t xa = value_x_advance(b, k + 2, vf1);
        value_x_advance(b, k + 2 + value_format_size(vf1), vf2);
        sets[s].push_back({second, xa});
      }
    }
    const Coverage c = parse_coverage(b, o + coverageOffset);
    add_coverage(&t, c);
    if (vf1 & 0x4) {
      auto apply = [&](uint16_t g, int64_t ci) {
        if (ci < 0 || ci >= (int64_t)sets.size()) unsupported("pair set length");
        for (const auto& pv : sets[ci])
          if (pv.second != 0) t.glyphPairs[pair_key(g, pv.first)] = pv.second;
      };
      if (c.format == 1) {
        if (c.glyphs.size() != pairSetCount) unsupported("pair set count");
        for (size_t ci = 0; ci < c.glyphs.size(); ci++) apply(c.glyphs[ci], (int64_t)ci);
      } else {
        for (const auto& r : c.ranges) {
          int64_t ci = r.second;
          for (uint32_t g = r.first.first; g <= r.first.second; g++, ci++) apply((uint16_t)g, ci);
        }
      }
    }
  } else if (t.format == 2) {
    b.check(i + 14);
    const uint16_t coverageOffset = b.u16(i), vf1 = b.u16(i + 2), vf2 = b.u16(i + 4),
                   cd1 = b.u16(i + 6), cd2 = b.u16(i + 8), class1Count = b.u16(i + 10),
                   class2Count = b.u16(i + 12);
    i += 14;
    const int64_t recSize = value_format_size(vf1) + value_format_size(vf2);
    std::vector<int16_t> xa((size_t)class1Count * class2Count);
    for (int c1 = 0; c1 < class1Count; c1++)
      for (int c2 = 0; c2 < class2Count; c2++, i += recSize) {
        b.check(i + recSize);
        xa[(size_t)c1 * class2Count + c2] = value_x_advance(b, i, vf1);
        value_x_advance(b, i + value_format_size(vf1), vf2);
      }
    parse_class_def(b, o + cd1, &t.class1);
    parse_class_def(b, o + cd2, &t.class2);
    add_coverage(&t, parse_coverage(b, o + coverageOffset));
    if (vf1 & 0x4)
      for (int c1 = 0; c1 < class1Count; c1++)
        for (int c2 = 0; c2 < class2Count; c2++) {
          const int16_t v = xa[(size_t)c1 * class2Count + c2];
          if (v != 0) t.classPairs[pair_key((uint16_t)c1, (uint16_t)c2)] = v;
        }
  } else {
    unsupported("pair pos format");
  }
  return t;
}

void parse_gpos(pxh_typeface* t, int64_t o) {  // opentype.nim:2085-2164
  const Buf& b = t->buf;
  b.check(o + 10);
  const uint16_t major = b.u16(o), minor = b.u16(o + 2), lookupListOffset = b.u16(o + 8);
  if (major != 1) unsupported("Gpos major version");
  if (minor == 1)
    b.check(o + 14);
  else if (minor != 0)
    unsupported("Gpos minor version");
  const int64_t ll = o + lookupListOffset;
  b.check(ll + 2);
  const uint16_t lookupCount = b.u16(ll);
  b.check(ll + 2 + 2 * (int64_t)lookupCount);
  for (int l = 0; l < lookupCount; l++) {
    const int64_t lo = ll + b.u16(ll + 2 + 2 * l);
    b.check(lo + 6);
    const uint16_t type = b.u16(lo), flag = b.u16(lo + 2), subCount = b.u16(lo + 4);
    b.check(lo + 6 + 2 * (int64_t)subCount);
    if (flag & 0x0010) b.check(lo + 6 + 2 * (int64_t)subCount + 2);
    for (int s = 0; s < subCount; s++) {
      if (type != 2) continue;
      PairPosTable pp = parse_pair_pos(b, lo + b.u16(lo + 6 + 2 * s));
      if (!pp.glyphPairs.empty() || !pp.classPairs.empty()) t->pairPos.push_back(std::move(pp));
    }
  }
  t->hasGpos = true;
}

void parse_post(pxh_typeface* t, int64_t o) {  // opentype.nim:2166-2174
  const Buf& b = t->buf;
  b.check(o + 14);
  read_version(b, o);
  t->underlinePosition = b.i16(o + 8);
  t->underlineThickness = b.i16(o + 10);
}

// parseOpenType (opentype.nim:2531-2588).
void parse_open_type(pxh_typeface* t, int64_t start) {
  const Buf& b = t->buf;
  b.check(start + 12);
  const uint16_t numTables = b.u16(start + 4);
  int64_t i = start + 12;
  b.check(i + (int64_t)numTables * 16);
  std::unordered_map<std::string, uint32_t> tables;
  for (int j = 0; j < numTables; j++, i += 16)
    tables[std::string((const char*)b.p + i, 4)] = b.u32(i + 8);
  auto at = [&](const char* tag) -> int64_t {
    auto it = tables.find(tag);
    if (it == tables.end()) throw FontError(std::string("Missing required font table: key not found: ") + tag);
    return it->second;
  };
  parse_cmap(t, at("cmap"));
  parse_head(t, at("head"));
  parse_hhea(t, at("hhea"));
  parse_maxp(t, at("maxp"));
  parse_hmtx(t, at("hmtx"));
  parse_name(t, at("name"));
  parse_os2(t, at("OS/2"));
  if (tables.count("loca") && tables.count("glyf")) {
    parse_loca_glyf(t, tables["loca"], tables["glyf"]);
  } else if (tables.count("CFF ")) {
    parse_cff(t, tables["CFF "]);
  } else {
    unsupported("glyph outlines");
  }
  if (tables.count("kern")) parse_kern(t, tables["kern"]);
  if (tables.count("GPOS")) parse_gpos(t, tables["GPOS"]);
  parse_post(t, at("post"));
}

// ---- glyph paths (opentype.nim:2182-2463) ----

void path_move(std::vector<float>* p, float x, float y) { p->insert(p->end(), {1.0f, x, y}); }
void path_line(std::vector<float>* p, float x, float y) { p->insert(p->end(), {2.0f, x, y}); }
void path_quad(std::vector<float>* p, float cx, float cy, float x, float y) {
  p->insert(p->end(), {7.0f, cx, cy, x, y});
}
void path_close(std::vector<float>* p) { p->push_back(0.0f); }

bool is_identity(const float* m) {
  static const float id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int k = 0; k < 9; k++)
    if (m[k] != id[k]) return false;
  return true;
}

// Path.transform (paths.nim:264-333) of a command array.
void path_transform(std::vector<float>* p, const float* m) {
  if (is_identity(m)) return;
  std::vector<float>& c = *p;
  if (!c.empty() && c[0] == 10.0f) c[0] = 1.0f;  // a leading RMove becomes Move
  size_t i = 0;
  while (i < c.size()) {
    const int kind = (int)c[i++];
    float mm[9];
    memcpy(mm, m, sizeof(mm));
    if (kind >= 10) mm[6] = mm[7] = 0.0f;  // relative commands: no translation
    switch (kind) {
      case 0:
        break;
      case 1: case 2: case 10: case 11: case 8: case 17:
        m3_apply(mm, c[i], c[i + 1], &c[i], &c[i + 1]);
        i += 2;
        break;
      case 3: case 12: {
        float y;
        m3_apply(mm, c[i], 0.0f, &c[i], &y);
        i += 1;
        break;
      }
      case 4: case 13: {
        float x;
        m3_apply(mm, 0.0f, c[i], &x, &c[i]);
        i += 1;
        break;
      }
      case 5: case 14:
        for (int k = 0; k < 3; k++) m3_apply(mm, c[i + 2 * k], c[i + 2 * k + 1], &c[i + 2 * k], &c[i + 2 * k + 1]);
        i += 6;
        break;
      case 6: case 15: case 7: case 16:
        for (int k = 0; k < 2; k++) m3_apply(mm, c[i + 2 * k], c[i + 2 * k + 1], &c[i + 2 * k], &c[i + 2 * k + 1]);
        i += 4;
        break;
      default:  // arcs never occur in glyph paths
        throw FontError("glyph path holds an arc command");
    }
  }
}

// parseGlyphPath (opentype.nim:2186-2319): a simple glyph's contours, with
// the implied on-curve midpoints between consecutive off-curve points.
void parse_simple_glyph(const Buf& b, int64_t offset, int numberOfContours, std::vector<float>* out) {
  if (numberOfContours < 0) throw FontError("Glyph numberOfContours must be >= 0");
  if (numberOfContours == 0) return;
  int64_t i = offset;
  b.check(i + 2 * numberOfContours + 2);
  std::vector<uint16_t> endPts(numberOfContours);
  for (int k = 0; k < numberOfContours; k++) endPts[k] = b.u16(i + 2 * k);
  i += 2 * numberOfContours;
  const int64_t instructionLength = b.u16(i);
  i += 2;
  b.check(instructionLength);
  i += instructionLength;
  const int64_t numPoints = (int64_t)endPts.back() + 1;
  std::vector<uint8_t> flags;
  for (int64_t point = 0; point < numPoints;) {
    const uint8_t flag = b.u8(i++);
    flags.push_back(flag);
    point++;
    if (flag & 0x8) {
      const int repeat = b.u8(i++);
      for (int r = 0; r < repeat; r++, point++) flags.push_back(flag);
    }
  }
  struct Pt {
    float x = 0, y = 0;
    bool on = false;
  };
  std::vector<Pt> pts((size_t)numPoints);
  int64_t prev = 0;
  for (size_t p = 0; p < flags.size(); p++) {
    const uint8_t f = flags[p];
    int64_t d;
    if (f & 0x2) {
      d = b.u8(i++);
      if (!(f & 0x10)) d = -d;
    } else if (f & 0x10) {
      d = 0;
    } else {
      d = b.i16(i);
      i += 2;
    }
    prev += d;
    if ((int64_t)p >= numPoints) unsupported("point out of bounds");
    pts[p].x = (float)prev;
    pts[p].on = (f & 1) != 0;
  }
  prev = 0;
  for (size_t p = 0; p < flags.size(); p++) {
    const uint8_t f = flags[p];
    int64_t d;
    if (f & 0x4) {
      d = b.u8(i++);
      if (!(f & 0x20)) d = -d;
    } else if (f & 0x20) {
      d = 0;
    } else {
      d = b.i16(i);
      i += 2;
    }
    prev += d;
    pts[p].y = (float)prev;
  }
  int64_t startIdx = 0;
  for (int ci = 0; ci < numberOfContours; ci++) {
    const int64_t endIdx = endPts[ci];
    if (endIdx < startIdx || endIdx >= numPoints) unsupported("contour out of bounds");
    const Pt* c = pts.data() + startIdx;
    const int64_t n = endIdx - startIdx + 1;
    startIdx = endIdx + 1;
    Pt curr = c[n - 1], next = c[0];
    if (curr.on)
      path_move(out, curr.x, curr.y);
    else if (next.on)
      path_move(out, next.x, next.y);
    else
      path_move(out, (curr.x + next.x) / 2.0f, (curr.y + next.y) / 2.0f);
    for (int64_t p = 0; p < n; p++) {
      curr = next;
      next = c[(p + 1) % n];
      if (curr.on) {
        path_line(out, curr.x, curr.y);
      } else {
        Pt next2 = next;
        if (!next.on) {
          next2.x = (curr.x + next.x) / 2.0f;
          next2.y = (curr.y + next.y) / 2.0f;
        }
        path_quad(out, curr.x, curr.y, next2.x, next2.y);
      }
    }
    path_close(out);
  }
}

void parse_glyf_glyph(const pxh_typeface* t, uint32_t glyphId, std::vector<float>* out, int depth);

// parseCompositeGlyph (opentype.nim:2321-2414): components placed by their
// offsets and F2Dot14 scales.  Matched-point arguments place nothing (the
// reference reads and ignores them).
void parse_composite_glyph(const pxh_typeface* t, int64_t offset, std::vector<float>* out, int depth) {
  if (depth > kMaxComponentDepth) unsupported("composite glyph nesting");
  const Buf& b = t->buf;
  int64_t i = offset;
  for (bool more = true; more;) {
    b.check(i + 4);
    const uint16_t flags = b.u16(i);
    i += 2;
    const uint16_t glyphId = b.u16(i);
    i += 2;
    float xScale = 1, scale01 = 0, scale10 = 0, yScale = 1, dx = 0, dy = 0;
    if (flags & 1) {
      b.check(i + 4);
      if (flags & 2) {
        dx = (float)b.i16(i);
        dy = (float)b.i16(i + 2);
      }
      i += 4;
    } else {
      b.check(i + 2);
      if (flags & 2) {
        dx = (float)b.i8(i);
        dy = (float)b.i8(i + 1);
      }
      i += 2;
    }
    if (flags & 0x8) {
      b.check(i + 2);
      xScale = yScale = read_f2dot14(b, i);
      i += 2;
    } else if (flags & 0x40) {
      b.check(i + 4);
      xScale = read_f2dot14(b, i);
      yScale = read_f2dot14(b, i + 2);
      i += 4;
    } else if (flags & 0x80) {
      b.check(i + 8);
      xScale = read_f2dot14(b, i);
      scale10 = read_f2dot14(b, i + 2);
      scale01 = read_f2dot14(b, i + 4);
      yScale = read_f2dot14(b, i + 6);
      i += 8;
    }
    std::vector<float> sub;
    parse_glyf_glyph(t, glyphId, &sub, depth + 1);
    const float m[9] = {xScale, scale10, 0.0f, scale01, yScale, 0.0f, dx, dy, 1.0f};
    path_transform(&sub, m);
    out->insert(out->end(), sub.begin(), sub.end());
    more = (flags & 0x20) != 0;
  }
}

void parse_glyf_glyph(const pxh_typeface* t, uint32_t glyphId, std::vector<float>* out, int depth) {
  // opentype.nim:2416-2439
  if (glyphId >= t->glyf.size()) throw FontError("Invalid glyph ID " + std::to_string(glyphId));
  const uint32_t o = t->glyf[glyphId];
  if (glyphId + 1 < t->glyf.size() && o == t->glyf[glyphId + 1]) return;  // empty glyph
  t->buf.check((int64_t)o + 10);
  const int numberOfContours = t->buf.i16(o);
  if (numberOfContours < 0)
    parse_composite_glyph(t, (int64_t)o + 10, out, depth);
  else
    parse_simple_glyph(t->buf, (int64_t)o + 10, numberOfContours, out);
}

// calcCFFSubroutineBias (opentype.nim:1020-1026).
int cff_bias(const pxh_typeface::Ranges& subrs) {
  return subrs.size() < 1240 ? 107 : subrs.size() < 33900 ? 1131 : 32768;
}

// parseCFFCharstring (opentype.nim:1028-1297): the Type 2 charstring of glyph
// glyphId as moveTo / lineTo / bezierCurveTo commands in float32, with the
// reference's rules: the stack's front is shifted off as operands are used,
// a new contour starts with moveTo without closing the one before, endchar
// closes the path and does not stop the charstring, flex and other escaped
// operators are unsupported, and a 255 operand is read as an unsigned 32-bit
// number over 65536.
struct CffRun {
  const pxh_typeface* t;
  const pxh_typeface::Ranges* subrs;
  int subrsBias;
  std::vector<float>* out;
  std::vector<float> stack;
  size_t head = 0;  // stack[head..] is the stack
  int nStems = 0;
  bool haveWidth = false;
  float x = 0.0f, y = 0.0f;
  size_t len() const { return stack.size() - head; }
  float shift() {
    if (head >= stack.size()) throw FontError("Invalid CFF charstring: stack underflow");
    return stack[head++];
  }
  float pop() {
    if (head >= stack.size()) throw FontError("Invalid CFF charstring: stack underflow");
    const float v = stack.back();
    stack.pop_back();
    return v;
  }
  void clear() {
    stack.clear();
    head = 0;
  }
  void curve(float c1x, float c1y, float c2x, float c2y) {
    out->insert(out->end(), {5.0f, c1x, c1y, c2x, c2y, x, y});
  }
  void stems() {
    if (len() % 2 != 0 && !haveWidth) shift();  // the width
    nStems += (int)(len() >> 1);
    clear();
    haveWidth = true;
  }
  void width_if(size_t more) {
    if (len() > more && !haveWidth) {
      shift();
      haveWidth = true;
    }
  }
  void parse(uint32_t a, uint32_t e, int depth) {
    if (depth > 64) throw FontError("Invalid CFF charstring: subroutines nest too deep");
    const Buf& b = t->buf;
    for (uint32_t i = a; i < e;) {
      const int v = b.u8(i++);
      switch (v) {
        case 1: case 3: case 18: case 23:  // hstem vstem hstemhm vstemhm
          stems();
          break;
        case 4:  // vmoveto
          width_if(1);
          y += pop();
          path_move(out, x, y);
          break;
        case 5:  // rlineto
          while (len() > 0) {
            x += shift();
            y += shift();
            path_line(out, x, y);
          }
          break;
        case 6: case 7:  // hlineto vlineto
          for (bool horiz = v == 6; len() > 0; horiz = !horiz) {
            if (horiz)
              x += shift();
            else
              y += shift();
            path_line(out, x, y);
          }
          break;
        case 8:  // rrcurveto
          while (len() > 0) {
            const float c1x = x + shift(), c1y = y + shift(), c2x = c1x + shift(), c2y = c1y + shift();
            x = c2x + shift();
            y = c2y + shift();
            curve(c1x, c1y, c2x, c2y);
          }
          break;
        case 10: case 29: {  // callsubr callgsubr
          const pxh_typeface::Ranges& s = v == 10 ? *subrs : t->cffGsubrs;
          const int k = (int)pop() + (v == 10 ? subrsBias : cff_bias(t->cffGsubrs));
          if (k < 0 || k >= (int)s.size()) throw FontError("Invalid CFF charstring: subroutine index");
          if (s[k].second > s[k].first) parse(s[k].first, s[k].second, depth + 1);
          break;
        }
        case 11:  // return
          return;
        case 14:  // endchar
          width_if(2);
          path_close(out);
          break;
        case 19: case 20:  // hintmask cntrmask
          stems();
          i += (uint32_t)((nStems + 7) >> 3);
          break;
        case 21: {  // rmoveto
          width_if(2);
          y += pop();
          x += pop();
          path_move(out, x, y);
          break;
        }
        case 22:  // hmoveto
          width_if(1);
          x += pop();
          path_move(out, x, y);
          break;
        case 24: {  // rcurveline
          while (len() > 2) {
            const float c1x = x + shift(), c1y = y + shift(), c2x = c1x + shift(), c2y = c1y + shift();
            x = c2x + shift();
            y = c2y + shift();
            curve(c1x, c1y, c2x, c2y);
          }
          x += shift();
          y += shift();
          path_line(out, x, y);
          break;
        }
        case 25: {  // rlinecurve
          while (len() > 6) {
            x += shift();
            y += shift();
            path_line(out, x, y);
          }
          const float c1x = x + shift(), c1y = y + shift(), c2x = c1x + shift(), c2y = c1y + shift();
          x = c2x + shift();
          y = c2y + shift();
          curve(c1x, c1y, c2x, c2y);
          break;
        }
        case 26:  // vvcurveto
          if (len() % 2 != 0) x += shift();
          while (len() > 0) {
            const float c1x = x, c1y = y + shift(), c2x = c1x + shift(), c2y = c1y + shift();
            x = c2x;
            y = c2y + shift();
            curve(c1x, c1y, c2x, c2y);
          }
          break;
        case 27:  // hhcurveto
          if (len() % 2 != 0) y += shift();
          while (len() > 0) {
            const float c1x = x + shift(), c1y = y, c2x = c1x + shift(), c2y = c1y + shift();
            x = c2x + shift();
            y = c2y;
            curve(c1x, c1y, c2x, c2y);
          }
          break;
        case 28:  // shortint
          stack.push_back((float)b.i16(i));
          i += 2;
          break;
        case 30: case 31:  // vhcurveto hvcurveto: alternate vertical / horizontal starts
          for (bool vert = v == 30; len() > 0; vert = !vert) {
            if (vert) {
              const float c1x = x, c1y = y + shift(), c2x = c1x + shift(), c2y = c1y + shift();
              x = c2x + shift();
              y = c2y + (len() == 1 ? shift() : 0.0f);
              curve(c1x, c1y, c2x, c2y);
            } else {
              const float c1x = x + shift(), c1y = y, c2x = c1x + shift(), c2y = c1y + shift();
              y = c2y + shift();
              x = c2x + (len() == 1 ? shift() : 0.0f);
              curve(c1x, c1y, c2x, c2y);
            }
          }
          break;
        default:
          if (v < 32) {
            unsupported("CFF unknown operator: " + std::to_string(v));
          } else if (v < 247) {
            stack.push_back((float)(v - 139));
          } else if (v < 251) {
            stack.push_back((float)((v - 247) * 256 + b.u8(i) + 108));
            i += 1;
          } else if (v < 255) {
            stack.push_back((float)(-(v - 251) * 256 - b.u8(i) - 108));
            i += 1;
          } else {
            stack.push_back((float)(int64_t)b.u32(i) / 65536.0f);
            i += 4;
          }
      }
    }
  }
};

void parse_cff_glyph(const pxh_typeface* t, uint32_t glyphId, std::vector<float>* out) {
  if (glyphId >= t->cffChars.size()) throw FontError("Invalid glyph ID " + std::to_string(glyphId));
  const pxh_typeface::Ranges* subrs = &t->cffTop.subrs;
  if (t->cffCid) {
    if (glyphId >= t->cffFdSelect.size()) throw FontError("Invalid glyph ID " + std::to_string(glyphId));
    subrs = &t->cffFds[t->cffFdSelect[glyphId]].subrs;
  }
  CffRun run{t, subrs, cff_bias(*subrs), out};
  run.parse(t->cffChars[glyphId].first, t->cffChars[glyphId].second, 0);
}

uint16_t glyph_id(const pxh_typeface* t, uint32_t rune) {  // opentype.nim:2176-2177
  auto it = t->cmap.find(rune);
  return it == t->cmap.end() ? 0 : it->second;
}

// getGlyphPath (opentype.nim:2456-2463): parsed once, flipped to y down.
const std::vector<float>& ot_glyph_path(pxh_typeface* t, uint32_t rune) {
  auto it = t->glyphPaths.find(rune);
  if (it != t->glyphPaths.end()) return it->second;
  std::vector<float> p;
  if (t->isCff)
    parse_cff_glyph(t, glyph_id(t, rune), &p);
  else
    parse_glyf_glyph(t, glyph_id(t, rune), &p, 0);
  const float flip[9] = {1, 0, 0, 0, -1, 0, 0, 0, 1};
  path_transform(&p, flip);
  return t->glyphPaths.emplace(rune, std::move(p)).first->second;
}

float ot_advance(const pxh_typeface* t, uint32_t rune) {  // opentype.nim:2474-2479
  const uint16_t g = glyph_id(t, rune);
  if (t->advances.empty()) return 0.0f;
  return g < t->advances.size() ? (float)t->advances[g] : (float)t->advances.back();
}

float ot_kerning(const pxh_typeface* t, uint32_t left, uint32_t right) {  // opentype.nim:2481-2515
  auto l = t->cmap.find(left), r = t->cmap.find(right);
  if (l == t->cmap.end() || r == t->cmap.end()) return 0.0f;
  const uint16_t lg = l->second, rg = r->second;
  if (t->hasGpos) {
    for (const PairPosTable& pp : t->pairPos) {
      if (!pp.covers(lg)) continue;
      if (pp.format == 1) {
        auto it = pp.glyphPairs.find(pair_key(lg, rg));
        if (it != pp.glyphPairs.end()) return (float)it->second;
      } else if (pp.format == 2) {
        auto c1 = pp.class1.find(lg), c2 = pp.class2.find(rg);
        const uint16_t lc = c1 == pp.class1.end() ? 0 : c1->second;
        const uint16_t rc = c2 == pp.class2.end() ? 0 : c2->second;
        auto it = pp.classPairs.find(pair_key(lc, rc));
        if (it != pp.classPairs.end()) return (float)it->second;
      }
    }
    return 0.0f;
  }
  if (t->hasKern) {
    auto it = t->kernPairs.find(pair_key(lg, rg));
    return it == t->kernPairs.end() ? 0.0f : it->second;
  }
  return 0.0f;
}

// ---- typeface level, with fallbacks (fonts.nim:59-195) ----

float tf_scale(const pxh_typeface* t) { return (float)t->unitsPerEm; }
bool tf_has_glyph(const pxh_typeface* t, uint32_t rune) { return t->cmap.count(rune) != 0; }

pxh_typeface* tf_fallback(pxh_typeface* t, uint32_t rune, int depth = 0) {  // fonts.nim:124-131
  if (tf_has_glyph(t, rune)) return t;
  if (depth > 64) return nullptr;
  for (pxh_typeface* f : t->fallbacks)
    if (pxh_typeface* r = tf_fallback(f, rune, depth + 1)) return r;
  return nullptr;
}

float tf_advance(pxh_typeface* t, uint32_t rune) {  // fonts.nim:152-165
  pxh_typeface* t2 = tf_fallback(t, rune);
  if (!t2) t2 = t;  // tofu advance
  float r = ot_advance(t2, rune);
  r *= tf_scale(t) / tf_scale(t2);
  return r;
}

float tf_kerning(pxh_typeface* t, uint32_t left, uint32_t right) {  // fonts.nim:167-184
  pxh_typeface* tr = tf_fallback(t, right);
  pxh_typeface* tl = tf_fallback(t, left);
  if (!tr || !tl) return 0.0f;
  float r = 0.0f;
  if (tr == tl) {
    r = ot_kerning(tr, left, right);
    r *= tf_scale(t) / tf_scale(tr);
  }
  return r;
}

// ---- font level (fonts.nim:186-211) ----

float font_scale(const pxh_font* f) { return f->size / tf_scale(f->typeface); }

float default_line_height(const pxh_font* f) {
  const pxh_typeface* t = f->typeface;
  const float units = (float)t->ascender - (float)t->descender + (float)t->lineGap;
  return roundf(units * font_scale(f));
}

float font_line_height(const pxh_font* f) {
  return f->line_height >= 0.0f ? f->line_height : default_line_height(f);
}

float font_line_gap(const pxh_font* f) {
  const float lh = font_line_height(f);
  if (lh == default_line_height(f)) return (float)f->typeface->lineGap;
  return (lh / font_scale(f)) - (float)f->typeface->ascender + (float)f->typeface->descender;
}

// ---- runes ----

constexpr uint32_t kLF = 10, kSP = 32;

// fastRuneAt (Nim's unicode module): lead-byte forms up to six bytes,
// continuation bits taken as they come, U+FFFD when the text ends early.
std::vector<uint32_t> decode_utf8(const char* s, int64_t n) {
  std::vector<uint32_t> out;
  int64_t i = 0;
  auto cb = [&](int64_t k) { return (uint32_t)((uint8_t)s[i + k] & 0x3F); };
  while (i < n) {
    const uint32_t c = (uint8_t)s[i];
    int len = 1;
    uint32_t r = c;
    if (c <= 127) {
      len = 1;
    } else if ((c >> 5) == 0x6) {
      len = 2;
      if (i <= n - 2) r = ((c & 0x1F) << 6) | cb(1);
    } else if ((c >> 4) == 0xE) {
      len = 3;
      if (i <= n - 3) r = ((c & 0xF) << 12) | (cb(1) << 6) | cb(2);
    } else if ((c >> 3) == 0x1E) {
      len = 4;
      if (i <= n - 4) r = ((c & 0x7) << 18) | (cb(1) << 12) | (cb(2) << 6) | cb(3);
    } else if ((c >> 2) == 0x3E) {
      len = 5;
      if (i <= n - 5) r = ((c & 0x3) << 24) | (cb(1) << 18) | (cb(2) << 12) | (cb(3) << 6) | cb(4);
    } else if ((c >> 1) == 0x7E) {
      len = 6;
      if (i <= n - 6)
        r = ((c & 0x1) << 30) | (cb(1) << 24) | (cb(2) << 18) | (cb(3) << 12) | (cb(4) << 6) | cb(5);
    }
    if (i > n - len) {  // truncated sequence
      r = 0xFFFD;
      len = 1;
    }
    out.push_back(r);
    i += len;
  }
  return out;
}

// Nim's unicode isWhiteSpace.
bool is_white_space(uint32_t r) {
  return (r >= 9 && r <= 13) || r == 32 || r == 0x85 || r == 0xA0 || r == 0x1680 ||
         (r >= 0x2000 && r <= 0x200A) || r == 0x2028 || r == 0x2029 || r == 0x202F ||
         r == 0x205F || r == 0x3000;
}

bool can_wrap(uint32_t r) { return r == kSP || is_white_space(r); }  // fonts.nim:253-254

// toUpper / toLower for ASCII and Latin-1 letters; other runes unchanged.
uint32_t to_upper(uint32_t r) {
  if (r >= 'a' && r <= 'z') return r - 32;
  if (r >= 0xE0 && r <= 0xFE && r != 0xF7) return r - 32;
  if (r == 0xFF) return 0x178;
  if (r == 0xB5) return 0x39C;
  return r;
}
uint32_t to_lower(uint32_t r) {
  if (r >= 'A' && r <= 'Z') return r + 32;
  if (r >= 0xC0 && r <= 0xDE && r != 0xD7) return r + 32;
  return r;
}

void convert_case(std::vector<uint32_t>* runes, int32_t textCase) {  // fonts.nim:235-251
  if (textCase == 1) {
    for (uint32_t& r : *runes) r = to_upper(r);
  } else if (textCase == 2) {
    for (uint32_t& r : *runes) r = to_lower(r);
  } else if (textCase == 3) {
    uint32_t prev = kSP;
    for (uint32_t& r : *runes) {
      if (is_white_space(prev)) r = to_upper(r);
      prev = r;
    }
  }
}

// ---- typeset (fonts.nim:256-466) ----

float rune_advance(const pxh_font* f, const std::vector<uint32_t>& runes, size_t i) {
  float r = 0.0f;
  if (!f->no_kerning_adjustments && i + 1 < runes.size()) r += tf_kerning(f->typeface, runes[i], runes[i + 1]);
  r += tf_advance(f->typeface, runes[i]);
  r *= font_scale(f);
  return r;
}

void typeset(pxh_arrangement* a, const pxh_span* spans, int32_t nspans, float bx, float by, int32_t hAlign,
             int32_t vAlign, bool wrap) {
  int32_t start = 0;
  for (int32_t s = 0; s < nspans; s++) {
    std::vector<uint32_t> rs;
    for (uint32_t r : decode_utf8(spans[s].text, spans[s].len))
      if (r >= kSP || r == kLF) rs.push_back(r);
    if (rs.empty()) continue;
    convert_case(&rs, spans[s].font->text_case);
    a->runes.insert(a->runes.end(), rs.begin(), rs.end());
    a->spans.push_back({start, start + (int32_t)rs.size() - 1});
    a->fonts.push_back(spans[s].font);
    start += (int32_t)rs.size();
  }
  const int32_t n = (int32_t)a->runes.size();
  if (n == 0) return;
  a->pos.assign(2 * (size_t)n, 0.0f);
  a->rects.assign(4 * (size_t)n, 0.0f);
  auto& lines = a->lines;
  lines.push_back({0, 0});
  float* P = a->pos.data();
  float* R = a->rects.data();

  // horizontal arrangement with line breaks and wrapping
  float atx = 0.0f, aty = 0.0f;
  int32_t prevCanWrap = 0;
  for (size_t si = 0; si < a->spans.size(); si++) {
    const pxh_font* font = a->fonts[si];
    for (int32_t ri = a->spans[si].first; ri <= a->spans[si].second; ri++) {
      const uint32_t rune = a->runes[ri];
      if (rune == kLF) {
        const float adv = tf_advance(font->typeface, kSP) * font_scale(font);
        P[2 * ri] = atx;
        P[2 * ri + 1] = aty;
        R[4 * ri] = atx;
        R[4 * ri + 1] = aty;
        R[4 * ri + 2] = adv;
        R[4 * ri + 3] = 0.0f;
        atx = 0.0f;
        aty += 1.0f;
        prevCanWrap = 0;
        lines.back().second = ri;
        if (ri + 1 < n) lines.push_back({ri + 1, 0});
      } else {
        const float adv = rune_advance(font, a->runes, ri);
        if (wrap && rune != kSP && bx > 0.0f && atx + adv > bx) {
          atx = 0.0f;
          aty += 1.0f;
          int32_t lineStart = ri;
          if (prevCanWrap > 0 && prevCanWrap != ri) {
            for (int32_t i = prevCanWrap + 1; i < ri; i++) {
              P[2 * i] = atx;
              P[2 * i + 1] = aty;
              R[4 * i] = atx;
              R[4 * i + 1] = aty;
              atx += rune_advance(font, a->runes, i);
              lineStart--;
            }
            prevCanWrap = 0;
          }
          lines.back().second = lineStart - 1;
          lines.push_back({lineStart, 0});
        }
        if (can_wrap(rune)) prevCanWrap = ri;
        P[2 * ri] = atx;
        P[2 * ri + 1] = aty;
        R[4 * ri] = atx;
        R[4 * ri + 1] = aty;
        R[4 * ri + 2] = adv;
        R[4 * ri + 3] = 0.0f;
        atx += adv;
      }
    }
  }
  lines.back().second = n - 1;

  if (hAlign != 0) {
    for (const auto& ln : lines) {
      float furthestX = 0.0f;
      for (int32_t i = ln.second; i >= ln.first; i--)
        if (a->runes[i] != kSP && a->runes[i] != kLF) {
          furthestX = R[4 * i] + R[4 * i + 2];
          break;
        }
      float xAdj = 0.0f;
      if (hAlign == 1) xAdj = (bx - furthestX) / 2.0f;
      if (hAlign == 2) xAdj = bx - furthestX;
      if (xAdj != 0.0f)
        for (int32_t i = ln.first; i <= ln.second; i++) {
          P[2 * i] += xAdj;
          R[4 * i] += xAdj;
        }
    }
  }

  // vertical arrangement
  const int32_t L = (int32_t)lines.size();
  float initialY = 0.0f;
  for (size_t si = 0; si < a->spans.size(); si++) {
    const pxh_font* font = a->fonts[si];
    const float units = (float)font->typeface->ascender + font_line_gap(font) / 2.0f;
    initialY = nim_maxf(initialY, roundf(units * font_scale(font)));
    if (a->spans[si].second >= lines[0].second) break;
  }
  std::vector<float> lineHeights((size_t)L, 0.0f);
  {
    int32_t line = 0;
    for (size_t si = 0; si < a->spans.size(); si++) {
      const float flh = font_line_height(a->fonts[si]);
      lineHeights[line] = nim_maxf(lineHeights[line], flh);
      for (int32_t ri = a->spans[si].first; ri <= a->spans[si].second; ri++)
        if (line + 1 < L && ri == lines[line + 1].first) {
          line++;
          lineHeights[line] = nim_maxf(lineHeights[line], flh);
        }
      if (line + 1 < L && a->spans[si].second == lines[line].second) line++;
    }
  }
  {
    int32_t line = 0;
    float baseline = initialY;
    for (size_t si = 0; si < a->spans.size(); si++) {
      const pxh_font* font = a->fonts[si];
      const float lh = font_line_height(font);
      const float top = roundf(((float)font->typeface->ascender + font_line_gap(font) / 2.0f) * font_scale(font));
      for (int32_t ri = a->spans[si].first; ri <= a->spans[si].second; ri++) {
        if (line + 1 < L && ri == lines[line + 1].first) {
          line++;
          baseline += lineHeights[line];
        }
        P[2 * ri + 1] = baseline;
        R[4 * ri + 1] = baseline - top;
        R[4 * ri + 3] = lh;
      }
    }
  }
  if (vAlign != 0) {
    const float furthestY = R[4 * (n - 1) + 1] + R[4 * (n - 1) + 3];
    float yAdj = 0.0f;
    if (vAlign == 1) yAdj = roundf((by - furthestY) / 2.0f);
    if (vAlign == 2) yAdj = by - furthestY;
    if (yAdj != 0.0f)
      for (int32_t i = 0; i < n; i++) {
        P[2 * i + 1] += yAdj;
        R[4 * i + 1] += yAdj;
      }
  }
  // selection rects snapped to the pixel grid
  for (int32_t i = 0; i < n; i++) {
    float* r = R + 4 * i;
    const float minX = roundf(r[0]), maxX = roundf(r[0] + r[2]);
    const float minY = roundf(r[1]), maxY = roundf(r[1] + r[3]);
    r[0] = minX;
    r[2] = maxX - minX;
    r[1] = minY;
    r[3] = maxY - minY;
  }
}

// path.rect (paths.nim:506-521) for a clockwise-wound font.
void path_rect(std::vector<float>* p, float x, float y, float w, float h, bool clockwise) {
  path_move(p, x, y);
  if (clockwise) {
    path_line(p, x + w, y);
    path_line(p, x + w, y + h);
    path_line(p, x, y + h);
  } else {
    path_line(p, x, y + h);
    path_line(p, x + w, y + h);
    path_line(p, x + w, y);
  }
  path_close(p);
}

// computePaths (fonts.nim:517-564) for one span.  `line` carries the line
// counter across spans as the reference's loop does.
// Appends src transformed by m (Path.transform, paths.nim:264-333, applied
// to a copy) to out; an identity m copies the commands unchanged.
void append_transformed(std::vector<float>* out, const std::vector<float>& src, const float* m) {
  const size_t base = out->size();
  out->insert(out->end(), src.begin(), src.end());
  if (is_identity(m)) return;
  float* c = out->data() + base;
  const size_t n = src.size();
  if (n && c[0] == 10.0f) c[0] = 1.0f;  // a leading RMove becomes Move
  const float a = m[0], b = m[1], d = m[3], e = m[4], tx = m[6], ty = m[7];
  size_t i = 0;
  while (i < n) {
    const int kind = (int)c[i++];
    int pts;
    switch (kind) {
      case 0: pts = 0; break;
      case 1: case 2: case 10: case 11: case 8: case 17: pts = 1; break;
      case 6: case 15: case 7: case 16: pts = 2; break;
      case 5: case 14: pts = 3; break;
      default:  // glyph paths hold moves, lines, curves and closes only
        throw FontError("glyph path holds an unexpected command");
    }
    // m * (x, y) as m3_apply computes it; relative commands drop the
    // translation
    const float ox = kind >= 10 ? 0.0f : tx, oy = kind >= 10 ? 0.0f : ty;
    for (int k = 0; k < pts; k++, i += 2) {
      const float x = c[i], y = c[i + 1];
      c[i] = a * x + d * y + ox;
      c[i + 1] = b * x + e * y + oy;
    }
  }
}

// The glyph's path in t's units (getGlyphPath, fonts.nim:133-150): the
// cached outline of the typeface that has it, scaled by the units ratio when
// that is a fallback (into *tmp).  Empty when no typeface has the rune.
const std::vector<float>& glyph_ref(pxh_typeface* t, uint32_t rune, std::vector<float>* tmp) {
  static const std::vector<float> none;
  pxh_typeface* t2 = tf_fallback(t, rune);
  if (!t2) return none;
  const std::vector<float>& p = ot_glyph_path(t2, rune);
  const float ratio = tf_scale(t) / tf_scale(t2);
  if (ratio == 1.0f) return p;
  tmp->assign(p.begin(), p.end());
  const float m[9] = {ratio, 0, 0, 0, ratio, 0, 0, 0, 1};
  path_transform(tmp, m);
  return *tmp;
}

// computePaths (fonts.nim:517-564) for one span.  `line` carries the line
// counter across spans as the reference's loop does.
void span_path(const pxh_arrangement* a, size_t si, int32_t* line, std::vector<float>* out) {
  const pxh_font* font = a->fonts[si];
  pxh_typeface* t = font->typeface;
  const float scale = font_scale(font);
  const float ut = (float)t->underlineThickness * scale, up = (float)t->underlinePosition * scale;
  const float st = (float)t->strikeoutSize * scale, sp = (float)t->strikeoutPosition * scale;
  const int32_t L = (int32_t)a->lines.size();
  const float S[9] = {scale, 0, 0, 0, scale, 0, 0, 0, 1};
  std::vector<float> tmp;
  size_t need = out->size();
  for (int32_t ri = a->spans[si].first; ri <= a->spans[si].second; ri++)
    need += glyph_ref(t, a->runes[ri], &tmp).size() + ((font->underline ? 13 : 0) + (font->strikethrough ? 13 : 0));
  out->reserve(need);
  for (int32_t ri = a->spans[si].first; ri <= a->spans[si].second; ri++) {
    const float px = a->pos[2 * ri], py = a->pos[2 * ri + 1];
    const float T[9] = {1, 0, 0, 0, 1, 0, px, py, 1};
    float M[9];
    m3_mul(T, S, M);
    append_transformed(out, glyph_ref(t, a->runes[ri], &tmp), M);
    bool decorate = true;
    if (*line < L && ri == a->lines[*line].second) {
      (*line)++;
      if (a->runes[ri] == kSP) decorate = false;
    }
    if (decorate) {
      const float* r = a->rects.data() + 4 * ri;
      // isCCW (opentype.nim:2517-2521): clockwise rects unless the font is CFF
      const bool cw = !font->typeface->isCff;
      if (font->underline) path_rect(out, r[0], py - up + ut / 2.0f, r[2], ut, cw);
      if (font->strikethrough) path_rect(out, r[0], py - sp, r[2], st, cw);
    }
  }
}

// computePaths (fonts.nim:517-564) of span si as glyph instances for the
// device: each glyph's cached outline with its placement, each decoration
// rect as written by path.rect.
int span_instances(const pxh_arrangement* a, size_t si, int32_t* line, std::vector<TextInst>* out) {
  const pxh_font* font = a->fonts[si];
  pxh_typeface* t = font->typeface;
  const float scale = font_scale(font);
  const float ut = (float)t->underlineThickness * scale, up = (float)t->underlinePosition * scale;
  const float st = (float)t->strikeoutSize * scale, sp = (float)t->strikeoutPosition * scale;
  const int32_t L = (int32_t)a->lines.size();
  const float S[9] = {scale, 0, 0, 0, scale, 0, 0, 0, 1};
  std::vector<float> tmp;
  for (int32_t ri = a->spans[si].first; ri <= a->spans[si].second; ri++) {
    const float px = a->pos[2 * ri], py = a->pos[2 * ri + 1];
    const uint32_t rune = a->runes[ri];
    const std::vector<float>& gp = glyph_ref(t, rune, &tmp);
    if (!gp.empty()) {
      const float T[9] = {1, 0, 0, 0, 1, 0, px, py, 1};
      float M[9];
      m3_mul(T, S, M);
      TextInst gi;
      if (int r = text_glyph_ref(t, rune, gp.data(), (int32_t)gp.size(), &gi.ref)) return r;
      gi.identity = is_identity(M) ? 1 : 0;
      const float m6[6] = {M[0], M[1], M[3], M[4], M[6], M[7]};
      memcpy(gi.m, m6, sizeof(m6));
      out->push_back(gi);
    }
    bool decorate = true;
    if (*line < L && ri == a->lines[*line].second) {
      (*line)++;
      if (rune == kSP) decorate = false;
    }
    if (decorate) {
      const float* r = a->rects.data() + 4 * ri;
      // m[4] = 1: a counter-clockwise rect (isCCW, opentype.nim:2517-2521)
      const float ccw = font->typeface->isCff ? 1.0f : 0.0f;
      if (font->underline) out->push_back({-1, 0, {r[0], py - up + ut / 2.0f, r[2], ut, ccw, 0}});
      if (font->strikethrough) out->push_back({-1, 0, {r[0], py - sp, r[2], st, ccw, 0}});
    }
  }
  return 0;
}

// fillText / strokeText with solid paints only: every (span, paint) as a
// text call of one scene, drawn in one device pass.  0 when drawn, -1 when
// the caller must draw span by span, else an error status.
int text_through_scene(pxh_image* img, const pxh_arrangement* a, const float* mat9, bool stroke, float width,
                       int32_t cap, int32_t join, float miter, const float* dashes, int32_t ndashes) {
  for (const pxh_font* f : a->fonts)
    for (int32_t k = 0; k < f->npaints; k++)
      if (f->paints[k].kind != PXH_SOLID_PAINT) return -1;
  pxh_scene* sc = text_scene();
  if (!sc) return -1;
  int32_t line = 0;
  std::vector<TextInst> inst;
  for (size_t s = 0; s < a->spans.size(); s++) {
    inst.clear();
    if (int r = span_instances(a, s, &line, &inst)) return r;
    const pxh_font* f = a->fonts[s];
    for (int32_t k = 0; k < f->npaints; k++) {
      uint32_t rgbx;
      if (!paint_solid_rgbx(&f->paints[k], &rgbx)) continue;  // fillPath draws nothing
      if (int r = text_record(sc, inst, mat9, rgbx, kNonZero, f->paints[k].blend_mode, stroke, width, cap, join,
                              miter, dashes, ndashes))
        return r;
    }
  }
  return text_scene_render(img);
}

std::vector<std::vector<float>> compute_paths(const pxh_arrangement* a) {
  std::vector<std::vector<float>> paths(a->spans.size());
  int32_t line = 0;
  for (size_t si = 0; si < a->spans.size(); si++) span_path(a, si, &line, &paths[si]);
  return paths;
}

int check_font(const pxh_font* f) {
  if (!f) return set_last_error(kErrArg, "invalid font");
  if (!f->typeface) return set_last_error(kErrArg, "font has no typeface");
  if (f->npaints < 0 || (f->npaints > 0 && !f->paints)) return set_last_error(kErrArg, "invalid font paints");
  return 0;
}

int export_cmds(const std::vector<float>& p, float** cmds, int32_t* ncmds) {
  if (!cmds || !ncmds) return set_last_error(kErrArg, "null output");
  *cmds = nullptr;
  *ncmds = (int32_t)p.size();
  if (p.empty()) return 0;
  *cmds = (float*)malloc(p.size() * sizeof(float));
  if (!*cmds) return set_last_error(kErrNoMem, "out of host memory");
  memcpy(*cmds, p.data(), p.size() * sizeof(float));
  return 0;
}

pxh_typeface* parse_typeface(const uint8_t* data, int64_t len, int64_t start) {
  pxh_typeface* t = new (std::nothrow) pxh_typeface();
  if (!t) {
    set_last_error(kErrNoMem, "out of host memory");
    return nullptr;
  }
  try {
    t->data.assign(data, data + len);
    t->buf.p = t->data.data();
    t->buf.n = (int64_t)t->data.size();
    parse_open_type(t, start);
  } catch (const FontError& e) {
    set_last_error(kErrParse, e.what());
    delete t;
    return nullptr;
  } catch (const std::bad_alloc&) {
    set_last_error(kErrNoMem, "out of host memory");
    delete t;
    return nullptr;
  }
  return t;
}

}  // namespace

extern "C" {

pxh_typeface* pxh_typeface_parse(const uint8_t* data, int64_t len) {
  if (!data || len < 0) {
    set_last_error(kErrArg, "invalid font data");
    return nullptr;
  }
  return parse_typeface(data, len, 0);
}

pxh_typeface* pxh_typeface_parse_collection(const uint8_t* data, int64_t len, int32_t index, int32_t* count) {
  if (!data || len < 0) {
    set_last_error(kErrArg, "invalid font data");
    return nullptr;
  }
  try {  // parseOpenTypeCollection, opentype.nim:2591-2616
    Buf b{data, len};
    b.check(12);
    if (memcmp(data, "ttcf", 4) != 0) unsupported("invalid ttc file");
    const uint16_t major = b.u16(4), minor = b.u16(6);
    const uint32_t numFonts = b.u32(8);
    if ((major != 1 && major != 2) && minor != 0) unsupported("ttc version");
    std::vector<uint32_t> dirs;
    for (uint32_t k = 0; k < numFonts; k++) dirs.push_back(b.u32(12 + 4 * (int64_t)k));
    if (count) *count = (int32_t)numFonts;
    if (index < 0 || (uint32_t)index >= numFonts) {
      set_last_error(kErrArg, "font index out of range");
      return nullptr;
    }
    return parse_typeface(data, len, dirs[index]);
  } catch (const FontError& e) {
    set_last_error(kErrParse, e.what());
    return nullptr;
  }
}

void pxh_typeface_free(pxh_typeface* tf) {
  if (tf) text_glyphs_forget();  // its address may come back as another typeface
  delete tf;
}

int pxh_typeface_set_fallbacks(pxh_typeface* tf, pxh_typeface* const* fallbacks, int32_t n) {
  if (!tf || n < 0 || (n > 0 && !fallbacks)) return set_last_error(kErrArg, "invalid fallbacks");
  tf->fallbacks.assign(fallbacks, fallbacks + n);
  text_glyphs_forget();  // a rune may now resolve to another typeface's glyph
  return 0;
}

int pxh_typeface_metrics(const pxh_typeface* tf, float* out9) {
  if (!tf || !out9) return set_last_error(kErrArg, "invalid typeface");
  out9[0] = tf_scale(tf);
  out9[1] = (float)tf->ascender;
  out9[2] = (float)tf->descender;
  out9[3] = (float)tf->lineGap;
  out9[4] = (float)tf->capHeight;
  out9[5] = (float)tf->underlinePosition;
  out9[6] = (float)tf->underlineThickness;
  out9[7] = (float)tf->strikeoutPosition;
  out9[8] = (float)tf->strikeoutSize;
  return 0;
}

int32_t pxh_typeface_name(const pxh_typeface* tf, char* buf, int32_t cap) {
  if (!tf) return -1;
  const int32_t n = (int32_t)tf->fullName.size();
  if (buf && cap > 0) {
    const int32_t k = n < cap - 1 ? n : cap - 1;
    memcpy(buf, tf->fullName.data(), (size_t)k);
    buf[k] = 0;
  }
  return n;
}

int32_t pxh_typeface_has_glyph(const pxh_typeface* tf, uint32_t rune) {
  return tf && tf_has_glyph(tf, rune) ? 1 : 0;
}

float pxh_typeface_advance(const pxh_typeface* tf, uint32_t rune) {
  return tf ? tf_advance(const_cast<pxh_typeface*>(tf), rune) : 0.0f;
}

float pxh_typeface_kerning(const pxh_typeface* tf, uint32_t left, uint32_t right) {
  return tf ? tf_kerning(const_cast<pxh_typeface*>(tf), left, right) : 0.0f;
}

int pxh_typeface_glyph_path(pxh_typeface* tf, uint32_t rune, float** cmds, int32_t* ncmds) {
  if (!tf) return set_last_error(kErrArg, "invalid typeface");
  try {
    std::vector<float> tmp;
    return export_cmds(glyph_ref(tf, rune, &tmp), cmds, ncmds);
  } catch (const FontError& e) {
    return set_last_error(kErrParse, e.what());
  }
}

pxh_arrangement* pxh_typeset(const pxh_span* spans, int32_t nspans, float bounds_x, float bounds_y,
                             int32_t h_align, int32_t v_align, int32_t wrap) {
  if (nspans < 0 || (nspans > 0 && !spans)) {
    set_last_error(kErrArg, "invalid spans");
    return nullptr;
  }
  if (h_align < 0 || h_align > 2 || v_align < 0 || v_align > 2) {
    set_last_error(kErrArg, "invalid alignment");
    return nullptr;
  }
  for (int32_t s = 0; s < nspans; s++) {
    if (check_font(spans[s].font)) return nullptr;
    if (spans[s].len < 0 || (spans[s].len > 0 && !spans[s].text)) {
      set_last_error(kErrArg, "invalid span text");
      return nullptr;
    }
  }
  pxh_arrangement* a = new (std::nothrow) pxh_arrangement();
  if (!a) {
    set_last_error(kErrNoMem, "out of host memory");
    return nullptr;
  }
  try {
    typeset(a, spans, nspans, bounds_x, bounds_y, h_align, v_align, wrap != 0);
  } catch (const std::bad_alloc&) {
    set_last_error(kErrNoMem, "out of host memory");
    delete a;
    return nullptr;
  }
  return a;
}

void pxh_arrangement_free(pxh_arrangement* a) { delete a; }

int32_t pxh_arrangement_count(const pxh_arrangement* a, int32_t what) {
  if (!a) return -1;
  if (what == 0) return (int32_t)a->runes.size();
  if (what == 1) return (int32_t)a->lines.size();
  if (what == 2) return (int32_t)a->spans.size();
  return -1;
}

int pxh_arrangement_read(const pxh_arrangement* a, uint32_t* runes, float* positions, float* selection_rects,
                         int32_t* lines, int32_t* spans) {
  if (!a) return set_last_error(kErrArg, "invalid arrangement");
  const size_t n = a->runes.size();
  if (runes && n) memcpy(runes, a->runes.data(), n * 4);
  if (positions && n) memcpy(positions, a->pos.data(), n * 8);
  if (selection_rects && n) memcpy(selection_rects, a->rects.data(), n * 16);
  if (lines)
    for (size_t k = 0; k < a->lines.size(); k++) {
      lines[2 * k] = a->lines[k].first;
      lines[2 * k + 1] = a->lines[k].second;
    }
  if (spans)
    for (size_t k = 0; k < a->spans.size(); k++) {
      spans[2 * k] = a->spans[k].first;
      spans[2 * k + 1] = a->spans[k].second;
    }
  return 0;
}

int pxh_arrangement_layout_bounds(const pxh_arrangement* a, float* out2) {
  if (!a || !out2) return set_last_error(kErrArg, "invalid arrangement");
  out2[0] = out2[1] = 0.0f;
  const size_t n = a->runes.size();
  if (n == 0) return 0;
  for (size_t i = 0; i < n; i++)
    if (a->runes[i] != kLF) out2[0] = nim_maxf(out2[0], a->rects[4 * i] + a->rects[4 * i + 2]);
  const float* f = a->rects.data() + 4 * (n - 1);
  out2[1] = f[1] + f[3];
  if (a->runes[n - 1] == kLF) out2[1] += f[3];
  return 0;
}

int pxh_arrangement_span_path(const pxh_arrangement* a, int32_t span, float** cmds, int32_t* ncmds) {
  if (!a || span < 0 || span >= (int32_t)a->spans.size()) return set_last_error(kErrArg, "invalid span");
  for (const pxh_font* f : a->fonts)
    if (int r = check_font(f)) return r;
  try {
    std::vector<std::vector<float>> paths = compute_paths(a);
    return export_cmds(paths[span], cmds, ncmds);
  } catch (const FontError& e) {
    return set_last_error(kErrParse, e.what());
  }
}

int pxh_arrangement_compute_bounds(const pxh_arrangement* a, const float* mat9, float* out4) {
  if (!a || !out4) return set_last_error(kErrArg, "invalid arrangement");
  for (const pxh_font* f : a->fonts)
    if (int r = check_font(f)) return r;
  std::vector<float> full;
  try {
    for (const auto& p : compute_paths(a)) full.insert(full.end(), p.begin(), p.end());
  } catch (const FontError& e) {
    return set_last_error(kErrParse, e.what());
  }
  static const float id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (mat9) path_transform(&full, mat9);
  // Path.computeBounds (paths.nim:1119-1125) with the identity: the
  // flattened shapes' segments, horizontal ones kept, y quantized.
  Shapes sh;
  std::string err;
  const int32_t r = commands_to_shapes(full.data(), (int32_t)full.size(), true, pixel_scale(id), &sh, &err);
  if (r != kOk) return set_last_error(r, err);
  float xMin = 3.40282347e38f, xMax = -3.40282347e38f, yMin = 3.40282347e38f, yMax = -3.40282347e38f;
  bool any = false;
  for (int32_t s = 0; s < sh.count(); s++) {
    const int32_t a0 = sh.ofs[s], a1 = sh.ofs[s + 1];
    if (a1 - a0 < 2) continue;
    V2 v1 = sh.pts[a1 - 1];
    v1.y = quantize_y(v1.y);
    for (int32_t i = a0; i < a1; i++) {
      V2 v2 = sh.pts[i];
      v2.y = quantize_y(v2.y);
      if (i == a0 && v1.x == v2.x && v1.y == v2.y) continue;
      V2 at = v1, to = v2;
      v1 = v2;
      if (at.y > to.y) std::swap(at, to);
      xMin = nim_minf(xMin, nim_minf(at.x, to.x));
      xMax = nim_maxf(xMax, nim_maxf(at.x, to.x));
      yMin = nim_minf(yMin, at.y);
      yMax = nim_maxf(yMax, to.y);
      any = true;
    }
  }
  out4[0] = out4[1] = out4[2] = out4[3] = 0.0f;
  if (!any) {
    // no segments: the initial extremes give x = float32.high, w = low - high
    out4[0] = xMin;
    out4[1] = yMin;
    out4[2] = xMax - xMin;
    out4[3] = yMax - yMin;
    return 0;
  }
  if (xMin != xMin || xMax != xMax || yMin != yMin || yMax != yMax) return 0;
  out4[0] = xMin;
  out4[1] = yMin;
  out4[2] = xMax - xMin;
  out4[3] = yMax - yMin;
  return 0;
}

int pxh_fill_text(pxh_image* img, const pxh_arrangement* a, const float* mat9) {
  if (!a) return set_last_error(kErrArg, "invalid arrangement");
  for (const pxh_font* f : a->fonts)
    if (int r = check_font(f)) return r;
  try {
    const int r = text_through_scene(img, a, mat9, false, 0.0f, 0, 0, 4.0f, nullptr, 0);
    if (r >= 0) return r;
  } catch (const FontError& e) {
    return set_last_error(kErrParse, e.what());
  }
  std::vector<std::vector<float>> paths;
  try {
    paths = compute_paths(a);
  } catch (const FontError& e) {
    return set_last_error(kErrParse, e.what());
  }
  for (size_t s = 0; s < paths.size(); s++) {
    const pxh_font* f = a->fonts[s];
    for (int32_t k = 0; k < f->npaints; k++)
      if (int r = pxh_fill_path_paint(img, paths[s].data(), (int32_t)paths[s].size(), mat9, &f->paints[k],
                                      kNonZero))
        return r;
  }
  return 0;
}

int pxh_stroke_text(pxh_image* img, const pxh_arrangement* a, const float* mat9, float stroke_width,
                    int32_t line_cap, int32_t line_join, float miter_limit, const float* dashes,
                    int32_t ndashes) {
  if (!a) return set_last_error(kErrArg, "invalid arrangement");
  for (const pxh_font* f : a->fonts)
    if (int r = check_font(f)) return r;
  try {
    const int r = text_through_scene(img, a, mat9, true, stroke_width, line_cap, line_join, miter_limit, dashes,
                                     ndashes);
    if (r >= 0) return r;
  } catch (const FontError& e) {
    return set_last_error(kErrParse, e.what());
  }
  std::vector<std::vector<float>> paths;
  try {
    paths = compute_paths(a);
  } catch (const FontError& e) {
    return set_last_error(kErrParse, e.what());
  }
  for (size_t s = 0; s < paths.size(); s++) {
    const pxh_font* f = a->fonts[s];
    for (int32_t k = 0; k < f->npaints; k++)
      if (int r = pxh_stroke_path_paint(img, paths[s].data(), (int32_t)paths[s].size(), mat9, &f->paints[k],
                                        stroke_width, line_cap, line_join, miter_limit, dashes, ndashes))
        return r;
  }
  return 0;
}

}  // extern "C"
