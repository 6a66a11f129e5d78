"""Writes tests/golden/fonts/NotoEmoji-subset.otf: the reference's NotoEmoji.otf
(tests/fonts, 2.3 MB, CFF outlines) cut down with fontTools' subsetter to the
runes of the reference's emoji text test (test_fonts.nim:7-20) and .notdef
with its outline, so that the fixture stays small.  The kept glyphs'
charstrings, advances and vertical metrics are unchanged (checked by
tests/test_fonts_oracle.py against the oracle's reading of the same runes).

Run from the repository root with the reference checked out at
/root/reference: python tests/golden/make_emoji_font.py
"""
import os
import sys

from fontTools import subset

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import text_cases  # noqa: E402

SRC = "/root/reference/tests/fonts/NotoEmoji.otf"
OUT = os.path.join(ROOT, "tests", "golden", "fonts", "NotoEmoji-subset.otf")

if __name__ == "__main__":
    opts = subset.Options()
    opts.notdef_outline = True
    opts.name_IDs = ["*"]
    opts.layout_features = ["*"]
    opts.hinting = True
    opts.desubroutinize = False
    font = subset.load_font(SRC, opts)
    sub = subset.Subsetter(opts)
    sub.populate(unicodes=sorted({ord(c) for c in text_cases.EMOJI_TEXT if c != "\n"}))
    sub.subset(font)
    subset.save_font(font, OUT, opts)
    print(OUT, os.path.getsize(OUT))
