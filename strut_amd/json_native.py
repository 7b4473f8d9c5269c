"""The reference's case-document model and its STRUT_JSON_CACHE_V1 file,
for the host side of the process boundary.

A case file is parsed into the document the reference's solver binary
builds (json_native.mojo:114-197): nodes appended in post-order (a value's
children before the value), each with its tag (Null 0, Bool 1, Number 2,
String 3, Array 4, Object 5), its byte span in the UTF-8 text, the decoded
string or the literal number text, and for arrays / objects a run of
``array_items`` / ``object_entries`` appended when the container closes
(json_native.mojo:601-685).  The grammar and the error messages, with
their line / column / caret rendering, follow json_native.mojo:199-697.

The cache (strut_io.mojo:129-297) serialises that document so a later run
skips the parse; files written here and by the reference's binary are the
same bytes for the same input.
"""

from __future__ import annotations

import re
from dataclasses import dataclass
from pathlib import Path
from typing import List, Tuple

NULL, BOOL, NUMBER, STRING, ARRAY, OBJECT = range(6)
CACHE_MAGIC = "STRUT_JSON_CACHE_V1"


class JsonParseError(ValueError):
    pass


class CacheError(ValueError):
    pass


@dataclass
class JsonNode:
    tag: int = NULL
    start: int = 0
    end: int = 0
    bool_value: bool = False
    number_value: float = 0.0
    number_is_integer: bool = False
    text_value: str = ""
    child_offset: int = 0
    child_count: int = 0


@dataclass
class CaseSourceInfo:
    json_path: str = ""
    case_dir: str = ""


def case_source_from_path(path: str) -> CaseSourceInfo:
    """strut_io.mojo:118-121: the resolved path and its directory."""
    p = Path(path).resolve()
    return CaseSourceInfo(str(p), str(p.parent))


class JsonDocument:
    def __init__(self):
        self.nodes: List[JsonNode] = []
        self.object_entries: List[Tuple[str, int, int]] = []  # (key, value_index, key_offset)
        self.array_items: List[int] = []
        self.root_index = -1

    def add_node(self, node: JsonNode) -> int:
        self.nodes.append(node)
        return len(self.nodes) - 1

    def object_find(self, index: int, key: str) -> int:
        """First entry with this key (json_native.mojo:149-157), -1 if none."""
        node = self.nodes[index]
        if node.tag != OBJECT:
            raise ValueError("JSON node is not an object")
        for k, v, _ in self.object_entries[node.child_offset:node.child_offset + node.child_count]:
            if k == key:
                return v
        return -1

    def array_item(self, index: int, item: int) -> int:
        node = self.nodes[index]
        if node.tag != ARRAY:
            raise ValueError("JSON node is not an array")
        if item < 0 or item >= node.child_count:
            raise ValueError("JSON array index out of range")
        return self.array_items[node.child_offset + item]

    def to_python(self, index: int | None = None):
        """The value as Python objects: an object's first entry of a key
        wins, as object_find does; integer literals become int."""
        if index is None:
            index = self.root_index
        node = self.nodes[index]
        if node.tag == OBJECT:
            out = {}
            for k, v, _ in self.object_entries[node.child_offset:node.child_offset + node.child_count]:
                if k not in out:
                    out[k] = self.to_python(v)
            return out
        if node.tag == ARRAY:
            return [self.to_python(i) for i in self.array_items[node.child_offset:node.child_offset + node.child_count]]
        if node.tag == NUMBER:
            return int(node.text_value) if node.number_is_integer else node.number_value
        if node.tag == STRING:
            return node.text_value
        if node.tag == BOOL:
            return node.bool_value
        return None


# ------------------------------------------------------------------ parser

_WS = re.compile(rb"[ \n\r\t]*")
_NUM = re.compile(rb"-?(?:0|[1-9][0-9]*)(\.[0-9]+)?([eE][+-]?[0-9]+)?")
_PLAIN_STR = re.compile(rb'"([^"\\\x00-\x1f]*)"')
_DIGITS = b"0123456789"
_HEX = b"0123456789abcdefABCDEF"

# token tags (json_native.mojo:14-26)
_LBRACE, _RBRACE, _LBRACKET, _RBRACKET, _COLON, _COMMA, _STR, _NUMTOK, _TRUE, _FALSE, _NULLTOK, _EOF = range(1, 13)
_PUNCT = {ord("{"): _LBRACE, ord("}"): _RBRACE, ord("["): _LBRACKET, ord("]"): _RBRACKET, ord(":"): _COLON,
          ord(","): _COMMA}


class _Parser:
    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0
        # current token: tag, start, end, value_start, value_end, has_escape, number_is_integer
        self.tok = (0, 0, 0, 0, 0, False, False)

    def error(self, offset: int, message: str):
        """json_native.mojo:238-276: line, column (bytes), context line, caret."""
        data = self.data
        off = min(max(offset, 0), len(data))
        line = data.count(b"\n", 0, off) + 1
        line_start = data.rfind(b"\n", 0, off) + 1
        column = off - line_start + 1
        line_end = data.find(b"\n", line_start)
        if line_end < 0:
            line_end = len(data)
        context = data[line_start:line_end].decode("utf-8", errors="replace")
        text = f"JSON parse error at line {line}, column {column}: {message}"
        if context:
            text += "\n" + context + "\n" + " " * (column - 1) + "^"
        raise JsonParseError(text)

    # -- tokens (json_native.mojo:282-440)
    def advance(self):
        data = self.data
        pos = _WS.match(data, self.pos).end()
        self.pos = pos
        if pos >= len(data):
            self.tok = (_EOF, pos, pos, pos, pos, False, False)
            return
        ch = data[pos]
        tag = _PUNCT.get(ch)
        if tag is not None:
            self.pos = pos + 1
            self.tok = (tag, pos, pos + 1, pos, pos, False, False)
        elif ch == 0x22:
            self.tok = self.scan_string(pos)
        elif ch == 0x2D or 0x30 <= ch <= 0x39:
            self.tok = self.scan_number(pos)
        elif ch == 0x74:
            self.tok = self.scan_literal(pos, b"true", _TRUE)
        elif ch == 0x66:
            self.tok = self.scan_literal(pos, b"false", _FALSE)
        elif ch == 0x6E:
            self.tok = self.scan_literal(pos, b"null", _NULLTOK)
        else:
            self.error(pos, "unexpected token")

    def scan_literal(self, pos, literal, tag):
        data = self.data
        for i in range(len(literal)):
            if pos + i >= len(data) or data[pos + i] != literal[i]:
                self.error(pos, "invalid literal")
        end = pos + len(literal)
        self.pos = end
        return (tag, pos, end, pos, end, False, False)

    def scan_string(self, pos):
        data = self.data
        m = _PLAIN_STR.match(data, pos)
        if m:
            self.pos = m.end()
            return (_STR, pos, m.end(), pos + 1, m.end() - 1, False, False)
        has_escape = False
        i = pos + 1
        n = len(data)
        while i < n:
            ch = data[i]
            if ch == 0x22:
                self.pos = i + 1
                return (_STR, pos, i + 1, pos + 1, i, has_escape, False)
            if ch == 0x5C:
                has_escape = True
                i += 1
                if i >= n:
                    self.error(pos, "unterminated string escape")
                esc = data[i]
                if esc in b'"\\/bfnrt':
                    i += 1
                    continue
                if esc == 0x75:
                    for j in range(4):
                        k = i + 1 + j
                        if k >= n or data[k] not in _HEX:
                            self.error(i, "invalid unicode escape")
                    i += 5
                    continue
                self.error(i, "invalid escape sequence")
            if ch < 0x20:
                self.error(i, "unescaped control character in string")
            i += 1
        self.error(pos, "unterminated string")

    def scan_number(self, pos):
        data = self.data
        n = len(data)
        m = _NUM.match(data, pos)
        if m and (m.end() >= n or data[m.end()] not in b"0123456789.eE+-"):
            end = m.end()
            self.pos = end
            return (_NUMTOK, pos, end, pos, end, False, m.group(1) is None and m.group(2) is None)
        # the exact scan, for its error positions (json_native.mojo:392-436)
        is_int = True
        i = pos
        if data[i] == 0x2D:
            i += 1
            if i >= n:
                self.error(pos, "number is missing digits")
        if data[i] == 0x30:
            i += 1
            if i < n and data[i] in _DIGITS:
                self.error(i, "numbers cannot have leading zeroes")
        elif data[i] in _DIGITS:
            while i < n and data[i] in _DIGITS:
                i += 1
        else:
            self.error(i, "number is missing digits")
        if i < n and data[i] == 0x2E:
            is_int = False
            i += 1
            if i >= n or data[i] not in _DIGITS:
                self.error(i, "fractional number is missing digits")
            while i < n and data[i] in _DIGITS:
                i += 1
        if i < n and data[i] in b"eE":
            is_int = False
            i += 1
            if i < n and data[i] in b"+-":
                i += 1
            if i >= n or data[i] not in _DIGITS:
                self.error(i, "exponent is missing digits")
            while i < n and data[i] in _DIGITS:
                i += 1
        self.pos = i
        return (_NUMTOK, pos, i, pos, i, False, is_int)

    def decode_string(self, tok) -> str:
        """json_native.mojo:488-530, with the surrogate-pair rules of :458-486."""
        data = self.data
        vs, ve, has_escape = tok[3], tok[4], tok[5]
        if not has_escape:
            return data[vs:ve].decode("utf-8")
        out = bytearray()
        i = seg = vs
        simple = {0x22: b'"', 0x5C: b"\\", 0x2F: b"/", 0x62: b"\b", 0x66: b"\f", 0x6E: b"\n", 0x72: b"\r",
                  0x74: b"\t"}
        while i < ve:
            if data[i] != 0x5C:
                i += 1
                continue
            out += data[seg:i]
            esc = data[i + 1]
            if esc in simple:
                out += simple[esc]
                i += 2
            else:
                start = i + 2
                value = int(data[start:start + 4], 16)
                nxt = start + 4
                if 0xD800 <= value <= 0xDBFF:
                    if nxt + 6 > ve or data[nxt] != 0x5C or data[nxt + 1] != 0x75:
                        self.error(start - 2, "high surrogate must be followed by low surrogate")
                    low = int(data[nxt + 2:nxt + 6], 16)
                    if low < 0xDC00 or low > 0xDFFF:
                        self.error(nxt + 2, "invalid low surrogate")
                    value = 0x10000 + ((value - 0xD800) << 10) + (low - 0xDC00)
                    nxt += 6
                elif 0xDC00 <= value <= 0xDFFF:
                    self.error(start - 2, "low surrogate without preceding high surrogate")
                out += chr(value).encode("utf-8")
                i = nxt
            seg = i
        out += data[seg:ve]
        return out.decode("utf-8")

    # -- values (json_native.mojo:545-685)
    def parse_value(self, doc: JsonDocument) -> int:
        tok = self.tok
        tag = tok[0]
        if tag == _STR:
            self.advance()
            return doc.add_node(JsonNode(STRING, tok[1], tok[2], text_value=self.decode_string(tok)))
        if tag == _NUMTOK:
            self.advance()
            text = self.data[tok[3]:tok[4]].decode("ascii")
            return doc.add_node(JsonNode(NUMBER, tok[1], tok[2], number_value=float(text),
                                         number_is_integer=tok[6], text_value=text))
        if tag in (_TRUE, _FALSE):
            self.advance()
            return doc.add_node(JsonNode(BOOL, tok[1], tok[2], bool_value=(tag == _TRUE)))
        if tag == _NULLTOK:
            self.advance()
            return doc.add_node(JsonNode(NULL, tok[1], tok[2]))
        if tag == _LBRACKET:
            return self.parse_array(doc)
        if tag == _LBRACE:
            return self.parse_object(doc)
        self.error(tok[1], "expected JSON value")

    def parse_array(self, doc: JsonDocument) -> int:
        start = self.tok[1]
        self.advance()
        if self.tok[0] == _RBRACKET:
            end = self.tok[2]
            self.advance()
            return doc.add_node(JsonNode(ARRAY, start, end, child_offset=len(doc.array_items)))
        items = []
        while True:
            items.append(self.parse_value(doc))
            if self.tok[0] == _COMMA:
                self.advance()
                if self.tok[0] == _RBRACKET:
                    self.error(self.tok[1], "trailing comma in array")
                continue
            if self.tok[0] == _RBRACKET:
                end = self.tok[2]
                self.advance()
                off = len(doc.array_items)
                doc.array_items.extend(items)
                return doc.add_node(JsonNode(ARRAY, start, end, child_offset=off, child_count=len(items)))
            self.error(self.tok[1], "expected ',' or ']' after array element")

    def parse_object(self, doc: JsonDocument) -> int:
        start = self.tok[1]
        self.advance()
        if self.tok[0] == _RBRACE:
            end = self.tok[2]
            self.advance()
            return doc.add_node(JsonNode(OBJECT, start, end, child_offset=len(doc.object_entries)))
        entries = []
        while True:
            if self.tok[0] != _STR:
                self.error(self.tok[1], "expected object key string")
            key_tok = self.tok
            key = self.decode_string(key_tok)
            self.advance()
            if self.tok[0] != _COLON:
                self.error(self.tok[1], "expected ':' after object key")
            self.advance()
            value = self.parse_value(doc)
            entries.append((key, value, key_tok[1]))
            if self.tok[0] == _COMMA:
                self.advance()
                if self.tok[0] == _RBRACE:
                    self.error(self.tok[1], "trailing comma in object")
                continue
            if self.tok[0] == _RBRACE:
                end = self.tok[2]
                self.advance()
                off = len(doc.object_entries)
                doc.object_entries.extend(entries)
                return doc.add_node(JsonNode(OBJECT, start, end, child_offset=off, child_count=len(entries)))
            self.error(self.tok[1], "expected ',' or '}' after object entry")


def parse_json_native(text) -> JsonDocument:
    """Parse a JSON text (str or UTF-8 bytes) into the reference's document."""
    data = text.encode("utf-8") if isinstance(text, str) else bytes(text)
    p = _Parser(data)
    doc = JsonDocument()
    p.advance()
    doc.root_index = p.parse_value(doc)
    if p.tok[0] != _EOF:
        p.error(p.tok[1], "unexpected trailing content")
    return doc


def load_json_native(path: str) -> JsonDocument:
    return parse_json_native(Path(path).read_bytes())


# ------------------------------------------------------------------ cache

def _block(value: str) -> bytes:
    b = value.encode("utf-8")
    return str(len(b)).encode() + b"\n" + b + b"\n"


def serialize_json_cache(doc: JsonDocument, source_info: CaseSourceInfo) -> bytes:
    """strut_io.mojo:129-179."""
    parts = [CACHE_MAGIC.encode() + b"\n", _block(source_info.json_path), _block(source_info.case_dir),
             f"{doc.root_index}\n{len(doc.nodes)}\n".encode()]
    for nd in doc.nodes:
        t = nd.text_value.encode("utf-8")
        parts.append(f"{nd.tag}\t{nd.start}\t{nd.end}\t{int(nd.bool_value)}\t{int(nd.number_is_integer)}\t"
                     f"{nd.child_offset}\t{nd.child_count}\t{len(t)}\n".encode())
        parts.append(t + b"\n")
    parts.append(f"{len(doc.object_entries)}\n".encode())
    for key, value_index, key_offset in doc.object_entries:
        k = key.encode("utf-8")
        parts.append(f"{value_index}\t{key_offset}\t{len(k)}\n".encode() + k + b"\n")
    parts.append(f"{len(doc.array_items)}\n".encode())
    parts.append(b"".join(f"{i}\n".encode() for i in doc.array_items))
    return b"".join(parts)


def write_json_cache(path: str, doc: JsonDocument, source_info: CaseSourceInfo) -> None:
    Path(path).write_bytes(serialize_json_cache(doc, source_info))


class _CacheReader:
    """strut_io.mojo:70-110: newline-ended lines and length-prefixed blocks."""

    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0

    def line(self) -> str:
        end = self.data.find(b"\n", self.pos)
        if end < 0:
            raise CacheError("invalid native input cache: truncated line")
        out = self.data[self.pos:end].decode("utf-8")
        self.pos = end + 1
        return out

    def block(self, length: int) -> str:
        if length < 0:
            raise CacheError("invalid native input cache: negative block length")
        end = self.pos + length
        if end > len(self.data):
            raise CacheError("invalid native input cache: truncated block")
        if end >= len(self.data) or self.data[end] != 0x0A:
            raise CacheError("invalid native input cache: missing block terminator")
        out = self.data[self.pos:end].decode("utf-8")
        self.pos = end + 1
        return out


def _field_int(fields: List[str], k: int) -> int:
    try:
        return int(fields[k])
    except (IndexError, ValueError):
        raise CacheError(f"invalid native input cache: field {k} of a record is not an integer") from None


def load_json_cache(path: str) -> Tuple[JsonDocument, CaseSourceInfo]:
    """strut_io.mojo:182-297."""
    r = _CacheReader(Path(path).read_bytes())
    if r.line() != CACHE_MAGIC:
        raise CacheError("invalid native input cache magic")
    json_path = r.block(int(r.line()))
    case_dir = r.block(int(r.line()))
    info = CaseSourceInfo(json_path, case_dir)
    if info.json_path == "":
        info = case_source_from_path(path)
    elif info.case_dir == "":
        info.case_dir = case_source_from_path(info.json_path).case_dir
    doc = JsonDocument()
    doc.root_index = int(r.line())
    for _ in range(int(r.line())):
        f = r.line().split("\t")
        text = r.block(_field_int(f, 7))
        node = JsonNode(tag=_field_int(f, 0), start=_field_int(f, 1), end=_field_int(f, 2),
                        bool_value=_field_int(f, 3) != 0, number_is_integer=_field_int(f, 4) != 0,
                        child_offset=_field_int(f, 5), child_count=_field_int(f, 6), text_value=text)
        if node.tag == NUMBER and text:
            node.number_value = float(text)
        doc.nodes.append(node)
    for _ in range(int(r.line())):
        f = r.line().split("\t")
        key = r.block(_field_int(f, 2))
        doc.object_entries.append((key, _field_int(f, 0), _field_int(f, 1)))
    for _ in range(int(r.line())):
        doc.array_items.append(int(r.line()))
    return doc, info


def load_case_document(input_path: str, input_bin_path: str, write_input_bin_path: str):
    """strut_io.mojo:300-323: one of --input / --input-bin, the cache written
    on request.  Returns (document, source info)."""
    if input_path and input_bin_path:
        raise ValueError("use only one of --input or --input-bin")
    if not input_path and not input_bin_path:
        raise ValueError("missing --input or --input-bin")
    if input_bin_path:
        doc, info = load_json_cache(input_bin_path)
    else:
        info = case_source_from_path(input_path)
        doc = load_json_native(input_path)
    if write_input_bin_path:
        write_json_cache(write_input_bin_path, doc, info)
    return doc, info
