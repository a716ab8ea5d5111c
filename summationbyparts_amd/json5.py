"""A JSON5 reader (https://spec.json5.org, version 1.0.0): the configuration
format of the multigrid front end (multigrid/src/main.rs:57 reads it with the
`json5` crate, version 0.4.1). No JSON5 module exists in this image, so the
grammar is read here by hand.

    loads(text) -> dict / list / str / int / float / bool / None

Objects keep their key order (the grids of a configuration are an ordered
map: a grid's position is its block index, multigrid/src/parsing.rs:205).
Beyond JSON: // and /* */ comments, unquoted (identifier) keys, single-quoted
strings, escaped line breaks inside strings, trailing commas, hexadecimal
integers, leading or trailing decimal points, explicit plus signs, Infinity
and NaN. Integers without a fraction or exponent stay Python ints (the
configuration's `steps` are unsigned integers); other numbers are floats.
"""

__all__ = ["loads", "load", "JSON5Error"]


class JSON5Error(ValueError):
    """A syntax error, with the line and column where it was found."""

    def __init__(self, msg, text, pos):
        line = text.count("\n", 0, pos) + 1
        col = pos - (text.rfind("\n", 0, pos) + 1) + 1
        super().__init__(f"{msg} at line {line}, column {col}")
        self.line, self.column = line, col


# ECMAScript white space and line terminators (spec 7.2, 7.3)
_WS = set(" \t\n\r\v\f\u00a0\u2028\u2029\ufeff\u1680\u2000\u2001\u2002\u2003\u2004\u2005\u2006\u2007"
          "\u2008\u2009\u200a\u202f\u205f\u3000")
_LINE_END = set("\n\r\u2028\u2029")
_ESCAPES = {"b": "\b", "f": "\f", "n": "\n", "r": "\r", "t": "\t", "v": "\v", "0": "\0", "'": "'", '"': '"',
            "\\": "\\", "/": "/"}


def _id_start(c):
    return c.isalpha() or c in "$_" or (ord(c) > 127 and c.isidentifier())


def _id_part(c):
    return _id_start(c) or c.isdigit() or c in "\u200c\u200d" or (ord(c) > 127 and ("a" + c).isidentifier())


class _Reader:
    def __init__(self, text):
        self.t = text
        self.i = 0
        self.n = len(text)

    def fail(self, msg, pos=None):
        raise JSON5Error(msg, self.t, self.i if pos is None else pos)

    def skip(self):
        """White space and comments."""
        t, n = self.t, self.n
        while self.i < n:
            c = t[self.i]
            if c in _WS:
                self.i += 1
            elif t.startswith("//", self.i):
                while self.i < n and t[self.i] not in _LINE_END:
                    self.i += 1
            elif t.startswith("/*", self.i):
                end = t.find("*/", self.i + 2)
                if end < 0:
                    self.fail("unterminated comment")
                self.i = end + 2
            else:
                break

    def peek(self):
        return self.t[self.i] if self.i < self.n else ""

    def value(self):
        self.skip()
        c = self.peek()
        if c == "{":
            return self.obj()
        if c == "[":
            return self.arr()
        if c in "\"'":
            return self.string()
        if c == "" :
            self.fail("unexpected end of input")
        if c.isdigit() or c in "+-.":
            return self.number()
        for word, val in (("true", True), ("false", False), ("null", None), ("Infinity", float("inf")),
                          ("NaN", float("nan"))):
            if self.t.startswith(word, self.i) and not (self.i + len(word) < self.n and
                                                        _id_part(self.t[self.i + len(word)])):
                self.i += len(word)
                return val
        self.fail(f"unexpected character {c!r}")

    def obj(self):
        self.i += 1
        out = {}
        while True:
            self.skip()
            if self.peek() == "}":
                self.i += 1
                return out
            key = self.key()
            self.skip()
            if self.peek() != ":":
                self.fail("expected ':' after an object key")
            self.i += 1
            out[key] = self.value()
            self.skip()
            c = self.peek()
            if c == ",":
                self.i += 1
            elif c == "}":
                self.i += 1
                return out
            else:
                self.fail("expected ',' or '}' in an object")

    def key(self):
        c = self.peek()
        if c in "\"'":
            return self.string()
        start = self.i
        if c == "\\" or (c and _id_start(c)):
            out = []
            while self.i < self.n:
                c = self.t[self.i]
                if c == "\\":
                    if not self.t.startswith("\\u", self.i):
                        self.fail("only \\u escapes are allowed in identifiers")
                    out.append(self.unicode_escape(self.i + 2))
                    self.i += 6
                elif (_id_start(c) if not out else _id_part(c)):
                    out.append(c)
                    self.i += 1
                else:
                    break
            return "".join(out)
        self.fail("expected an object key", start)

    def unicode_escape(self, pos):
        h = self.t[pos:pos + 4]
        if len(h) != 4 or any(ch not in "0123456789abcdefABCDEF" for ch in h):
            self.fail("bad \\u escape", pos)
        return chr(int(h, 16))

    def arr(self):
        self.i += 1
        out = []
        while True:
            self.skip()
            if self.peek() == "]":
                self.i += 1
                return out
            out.append(self.value())
            self.skip()
            c = self.peek()
            if c == ",":
                self.i += 1
            elif c == "]":
                self.i += 1
                return out
            else:
                self.fail("expected ',' or ']' in an array")

    def string(self):
        q = self.t[self.i]
        self.i += 1
        out = []
        t, n = self.t, self.n
        while True:
            if self.i >= n:
                self.fail("unterminated string")
            c = t[self.i]
            if c == q:
                self.i += 1
                return "".join(out)
            if c == "\\":
                self.i += 1
                e = t[self.i] if self.i < n else ""
                if e in _LINE_END:  # line continuation
                    self.i += 2 if t.startswith("\r\n", self.i) else 1
                    continue
                if e == "x":
                    h = t[self.i + 1:self.i + 3]
                    if len(h) != 2 or any(ch not in "0123456789abcdefABCDEF" for ch in h):
                        self.fail("bad \\x escape")
                    out.append(chr(int(h, 16)))
                    self.i += 3
                    continue
                if e == "u":
                    out.append(self.unicode_escape(self.i + 1))
                    self.i += 5
                    continue
                if e == "0" and self.i + 1 < n and t[self.i + 1].isdigit():
                    self.fail("octal escapes are not allowed")
                if e.isdigit() and e != "0":
                    self.fail("digit escapes are not allowed")
                out.append(_ESCAPES.get(e, e))
                self.i += 1
                continue
            if c in "\n\r":
                self.fail("line break inside a string")
            out.append(c)
            self.i += 1

    def number(self):
        t, start = self.t, self.i
        sign = 1
        if t[self.i] in "+-":
            sign = -1 if t[self.i] == "-" else 1
            self.i += 1
        for word, val in (("Infinity", float("inf")), ("NaN", float("nan"))):
            if t.startswith(word, self.i):
                self.i += len(word)
                return sign * val
        if t.startswith(("0x", "0X"), self.i):
            j = self.i + 2
            while j < self.n and t[j] in "0123456789abcdefABCDEF":
                j += 1
            if j == self.i + 2:
                self.fail("bad hexadecimal number")
            v = int(t[self.i + 2:j], 16)
            self.i = j
            return sign * v
        j = self.i
        while j < self.n and t[j].isdigit():
            j += 1
        int_digits = t[self.i:j]
        if len(int_digits) > 1 and int_digits[0] == "0":
            self.fail("leading zeros are not allowed", start)
        frac = exp = ""
        if j < self.n and t[j] == ".":
            k = j + 1
            while k < self.n and t[k].isdigit():
                k += 1
            frac = t[j:k]
            j = k
        if not int_digits and len(frac) <= 1:
            self.fail("bad number", start)
        if j < self.n and t[j] in "eE":
            k = j + 1
            if k < self.n and t[k] in "+-":
                k += 1
            d0 = k
            while k < self.n and t[k].isdigit():
                k += 1
            if k == d0:
                self.fail("bad exponent", j)
            exp = t[j:k]
            j = k
        if j < self.n and _id_part(t[j]):
            self.fail("bad number", start)
        self.i = j
        if not frac and not exp:
            return sign * int(int_digits)
        return sign * float((int_digits or "0") + (frac if frac != "." else "") + exp)


def loads(text):
    """The JSON5 value of `text` (JSON5Error on a syntax error)."""
    if isinstance(text, (bytes, bytearray)):
        text = text.decode("utf-8")
    r = _Reader(text)
    v = r.value()
    r.skip()
    if r.i != r.n:
        r.fail("trailing characters after the value")
    return v


def load(fp):
    return loads(fp.read())
