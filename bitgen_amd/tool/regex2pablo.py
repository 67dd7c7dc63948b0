"""Regex group -> pablo bitstream program (the IR producer).

The reference obtains pablo from an external icgrep-1.0 binary
(``bitgen/tool/pcre2pablo.py:33-48``: ``icgrep -f group.txt -print-pablo=...
-c -if-insertion-gap=2147483647 -disable-matchstar
-disable-Unicode-linebreak``), which is not vendored.  This module is this
project's own producer of programs in the same dialect and conventions:

* basis streams ``basis0..7`` with basis0 = most significant bit of each byte
  (``bitgen/template/template.py:135-145``);
* character classes named ``CC__<hex ranges>`` (by default built from 32
  shared nibble masks in the group header, ``cc_mode="nibble"``; or as a
  decision diagram over the basis bits, ``cc_mode="bdd"``), loop variables ``test*`` /
  ``accum*`` (the names ``bitgen/passes/remove_alias.py:24`` and
  ``bitgen/passes/cc_advance.py:106`` key on), result in ``output.matches``;
* markers sit one past the consumed character: consuming class C from marker
  stream m is ``pablo.Advance(m & C, 1)`` (``README.md:10``);
* line mode (``-c`` counts matching lines, ``-disable-Unicode-linebreak`` makes
  LF the only line break): no class matches LF, ``^`` is "after LF or at
  position 0", ``$`` is "at an LF", and ``output.matches`` marks the LF that
  ends every line containing a match of any regex of the group.  A match on a
  final line that has no terminating LF is not marked;
* unbounded repetition of a single class is ``pablo.MatchStar`` (star_mode
  ``"matchstar"``) or the fixed-point ``while`` loop icgrep emits under
  ``-disable-matchstar`` (star_mode ``"while"``); a star directly followed by
  a class disjoint from the starred one becomes ``ScanThru(m & C, C) | m``;
  repetition of anything else is always a ``while`` loop; bounded repetition
  is unrolled.  ``ScanThru(M, C) = (M + C) & ~C`` is only emitted with M inside
  C-runs (M & C), where no carry can land on another marker.

Supported syntax: literals (bytes 0..255), ``\\xHH``, ``\\d \\D \\w \\W \\s
\\S \\t \\n \\r \\f \\v \\e \\a \\0`` and escaped punctuation, ``.``, classes
with ranges/negation/escapes, ``( )``, ``(?: )``, ``|``, ``* + ? {n} {n,}
{n,m}`` with optional lazy ``?`` / possessive ``+`` suffix (all match the same
lines), ``^`` and ``$``.
"""
from dataclasses import dataclass, field
from typing import List, Optional

ALL = (1 << 256) - 1
LF = 0x0A


class RegexSyntaxError(ValueError):
    pass


def _bit(b):
    return 1 << b


def _range(lo, hi):
    return ((1 << (hi + 1)) - 1) ^ ((1 << lo) - 1)


DIGIT = _range(0x30, 0x39)
WORD = _range(0x30, 0x39) | _range(0x41, 0x5A) | _range(0x61, 0x7A) | _bit(0x5F)
SPACE = _bit(0x20) | _bit(0x09) | _bit(0x0A) | _bit(0x0B) | _bit(0x0C) | _bit(0x0D)
_SIMPLE_ESC = {"t": 0x09, "n": 0x0A, "r": 0x0D, "f": 0x0C, "v": 0x0B, "e": 0x1B, "a": 0x07, "0": 0x00}
_CLASS_ESC = {"d": DIGIT, "D": ALL ^ DIGIT, "w": WORD, "W": ALL ^ WORD, "s": SPACE, "S": ALL ^ SPACE}


# --------------------------------------------------------------------------
# AST
@dataclass
class CC:
    mask: int


@dataclass
class Seq:
    items: List[object] = field(default_factory=list)


@dataclass
class Alt:
    items: List[object] = field(default_factory=list)


@dataclass
class Rep:
    node: object
    lo: int
    hi: Optional[int]  # None = unbounded


@dataclass
class Start:
    pass


@dataclass
class End:
    pass


class _RegexParser:
    def __init__(self, text):
        self.s = text
        self.i = 0

    def peek(self):
        return self.s[self.i] if self.i < len(self.s) else None

    def take(self):
        c = self.peek()
        if c is None:
            raise RegexSyntaxError(f"unexpected end of regex {self.s!r}")
        self.i += 1
        return c

    def parse(self):
        node = self.alt()
        if self.i != len(self.s):
            raise RegexSyntaxError(f"unexpected {self.peek()!r} at {self.i} in {self.s!r}")
        return node

    def alt(self):
        items = [self.seq()]
        while self.peek() == "|":
            self.i += 1
            items.append(self.seq())
        return items[0] if len(items) == 1 else Alt(items)

    def seq(self):
        items = []
        while self.peek() is not None and self.peek() not in "|)":
            items.append(self.quantified())
        return Seq(items)

    def quantified(self):
        atom = self.atom()
        while True:
            c = self.peek()
            if c == "*":
                self.i += 1
                lo, hi = 0, None
            elif c == "+":
                self.i += 1
                lo, hi = 1, None
            elif c == "?":
                self.i += 1
                lo, hi = 0, 1
            elif c == "{" and self._is_brace_quant():
                lo, hi = self.brace()
            else:
                return atom
            if isinstance(atom, (Start, End)):
                raise RegexSyntaxError("quantified anchor")
            if self.peek() in ("?", "+"):
                self.i += 1  # lazy / possessive: same matched lines
            atom = Rep(atom, lo, hi)

    def _is_brace_quant(self):
        j = self.s.find("}", self.i)
        if j < 0:
            return False
        body = self.s[self.i + 1:j]
        parts = body.split(",")
        return 1 <= len(parts) <= 2 and parts[0].isdigit() and all(p == "" or p.isdigit() for p in parts[1:])

    def brace(self):
        j = self.s.find("}", self.i)
        body = self.s[self.i + 1:j]
        self.i = j + 1
        if "," not in body:
            n = int(body)
            return n, n
        a, b = body.split(",")
        lo = int(a)
        hi = None if b == "" else int(b)
        if hi is not None and hi < lo:
            raise RegexSyntaxError(f"bad repeat {{{body}}}")
        return lo, hi

    def atom(self):
        c = self.take()
        if c == "(":
            if self.peek() == "?":
                self.i += 1
                if self.peek() != ":":
                    raise RegexSyntaxError("only (?: ) groups are supported")
                self.i += 1
            node = self.alt()
            if self.take() != ")":
                raise RegexSyntaxError("missing )")
            return node
        if c == "[":
            return CC(self.klass())
        if c == ".":
            return CC(ALL ^ _bit(LF))
        if c == "^":
            return Start()
        if c == "$":
            return End()
        if c == "\\":
            return CC(self.escape(in_class=False))
        if c in ")*+?":
            raise RegexSyntaxError(f"unexpected {c!r} in {self.s!r}")
        return CC(_bit(self._byte(c)))

    @staticmethod
    def _byte(c):
        o = ord(c)
        if o > 255:
            raise RegexSyntaxError(f"non-byte character {c!r}")
        return o

    def escape(self, in_class):
        c = self.take()
        if c == "x":
            h = self.s[self.i:self.i + 2]
            if len(h) != 2 or any(ch not in "0123456789abcdefABCDEF" for ch in h):
                raise RegexSyntaxError(f"bad \\x escape in {self.s!r}")
            self.i += 2
            return _bit(int(h, 16))
        if c in _CLASS_ESC:
            return _CLASS_ESC[c]
        if c in _SIMPLE_ESC:
            return _bit(_SIMPLE_ESC[c])
        if c.isalnum():
            raise RegexSyntaxError(f"unsupported escape \\{c}")
        return _bit(self._byte(c))

    def _class_atom(self):
        """One class member: returns (mask, byte or None) -- byte if usable in a range."""
        c = self.take()
        if c == "\\":
            m = self.escape(in_class=True)
            single = m.bit_length() - 1 if m and (m & (m - 1)) == 0 else None
            return m, single
        if c == "[" and self.peek() == ":":
            j = self.s.find(":]", self.i)
            if j < 0:
                raise RegexSyntaxError("bad POSIX class")
            name = self.s[self.i + 1:j]
            self.i = j + 2
            posix = {
                "alpha": _range(0x41, 0x5A) | _range(0x61, 0x7A),
                "digit": DIGIT,
                "alnum": DIGIT | _range(0x41, 0x5A) | _range(0x61, 0x7A),
                "space": SPACE,
                "upper": _range(0x41, 0x5A),
                "lower": _range(0x61, 0x7A),
                "xdigit": DIGIT | _range(0x41, 0x46) | _range(0x61, 0x66),
                "print": _range(0x20, 0x7E),
                "punct": _range(0x21, 0x2F) | _range(0x3A, 0x40) | _range(0x5B, 0x60) | _range(0x7B, 0x7E),
            }
            if name not in posix:
                raise RegexSyntaxError(f"unknown POSIX class {name}")
            return posix[name], None
        b = self._byte(c)
        return _bit(b), b

    def klass(self):
        neg = False
        if self.peek() == "^":
            neg = True
            self.i += 1
        mask = 0
        first = True
        while True:
            c = self.peek()
            if c is None:
                raise RegexSyntaxError("unterminated class")
            if c == "]" and not first:
                self.i += 1
                break
            first = False
            m, lo = self._class_atom()
            if self.peek() == "-" and self.i + 1 < len(self.s) and self.s[self.i + 1] != "]" and lo is not None:
                self.i += 1
                m2, hi = self._class_atom()
                if hi is None or hi < lo:
                    raise RegexSyntaxError("bad class range")
                m = _range(lo, hi)
            mask |= m
        return (ALL ^ mask) if neg else mask


def parse_regex(text):
    return _RegexParser(text).parse()


# --------------------------------------------------------------------------
# Emission
def cc_name(mask):
    """``CC__`` + hex ranges, icgrep style (e.g. ``CC__41_5a_61_7a``)."""
    parts = []
    b = 0
    while b < 256:
        if mask >> b & 1:
            e = b
            while e + 1 < 256 and mask >> (e + 1) & 1:
                e += 1
            parts.append(f"{b:02x}" if b == e else f"{b:02x}_{e:02x}")
            b = e + 1
        else:
            b += 1
    return "CC__" + "_".join(parts) if parts else "CC__none"


class PabloEmitter:
    """Builds one pablo program for a group of regexes."""

    def __init__(self, star_mode="matchstar", line_mode=True, cc_mode="nibble"):
        if star_mode not in ("matchstar", "while"):
            raise ValueError(star_mode)
        if cc_mode not in ("nibble", "bdd"):
            raise ValueError(cc_mode)
        self.star_mode = star_mode
        self.line_mode = line_mode
        self.cc_mode = cc_mode
        self.nib_hi = None
        self.nib_lo = None
        self.lines = []
        self.indent = ""
        self.counter = {}
        self.used_names = set()
        self.scopes = [{}]
        self.group_memo = {}

    # -- naming / emission ---------------------------------------------------
    def fresh(self, prefix):
        n = self.counter.get(prefix, 0)
        while True:
            name = prefix if n == 0 else f"{prefix}_{n}"
            n += 1
            if name not in self.used_names:
                break
        self.counter[prefix] = n
        self.used_names.add(name)
        return name

    def emit(self, name, expr):
        self.lines.append(f"{self.indent}{name} = {expr}")
        return name

    def memo_get(self, key):
        for scope in reversed(self.scopes):
            if key in scope:
                return scope[key]
        return None

    def memo_put(self, key, value):
        self.scopes[-1][key] = value
        return value

    # -- bitwise helpers with constant folding --------------------------------
    def AND(self, a, b):
        if a == "0" or b == "0":
            return "0"
        if a == "1":
            return b
        if b == "1" or a == b:
            return a
        key = ("and",) + tuple(sorted((a, b)))
        got = self.memo_get(key)
        return got or self.memo_put(key, self.emit(self.fresh("and"), f"({a} & {b})"))

    def OR(self, a, b):
        if a == "1" or b == "1":
            return "1"
        if a == "0":
            return b
        if b == "0" or a == b:
            return a
        key = ("or",) + tuple(sorted((a, b)))
        got = self.memo_get(key)
        return got or self.memo_put(key, self.emit(self.fresh("or"), f"({a} | {b})"))

    def NOT(self, a):
        if a == "0":
            return "1"
        if a == "1":
            return "0"
        key = ("not", a)
        got = self.memo_get(key)
        return got or self.memo_put(key, self.emit(self.fresh("not"), f"(~{a})"))

    def ADV(self, a, n=1):
        if a == "0":
            return "0"
        key = ("adv", a, n)
        got = self.memo_get(key)
        return got or self.memo_put(key, self.emit(self.fresh("adv"), f"pablo.Advance({a}, {n})"))

    # -- character classes ----------------------------------------------------
    def _cc_fn(self, level, mask):
        """Boolean function over basis{level..7} selecting the 2**(8-level) values in mask."""
        width = 1 << (8 - level)
        full = (1 << width) - 1
        if mask == 0:
            return "0"
        if mask == full:
            return "1"
        key = ("ccfn", level, mask)
        got = self.memo_get(key)
        if got:
            return got
        half = width >> 1
        lo = mask & ((1 << half) - 1)       # values with basis{level} = 0
        hi = mask >> half                   # values with basis{level} = 1
        b = f"basis{level}"
        if lo == hi:
            res = self._cc_fn(level + 1, lo)
        else:
            fh = self._cc_fn(level + 1, hi)
            fl = self._cc_fn(level + 1, lo)
            if fl == "0":
                res = self.AND(b, fh)
            elif fh == "0":
                res = self.AND(self.NOT(b), fl)
            elif fh == "1":
                res = self.OR(b, fl)
            elif fl == "1":
                res = self.OR(self.NOT(b), fh)
            else:
                res = self.OR(self.AND(b, fh), self.AND(self.NOT(b), fl))
        return self.memo_put(key, res)

    def cc(self, mask, keep_lf=False):
        if self.line_mode and not keep_lf:
            mask &= ~_bit(LF)
        if mask == 0:
            return "0"
        if mask == ALL:
            return "1"
        key = ("cc", mask)
        got = self.memo_get(key)
        if got:
            return got
        if self.cc_mode == "nibble" and self.nib_hi is not None:
            fn = self._cc_nibble(mask)
        else:
            # Values are indexed with basis0 (the MSB) as the top split: value
            # v's bit within the level-0 mask is v itself.
            fn = self._cc_fn(0, mask)
        if fn in ("0", "1"):
            return fn
        name = self.emit(self.fresh(cc_name(mask)), fn)
        return self.memo_put(key, name)

    # -- nibble-factored classes -------------------------------------------------
    def emit_nibble_masks(self):
        """Group header: H[h] = (high nibble == h), L[l] = (low nibble == l),
        built from 2-bit products of the basis streams (56 operations), so a
        single-byte class is one AND: CC(b) = H[b >> 4] & L[b & 15]."""
        nb = [self.NOT(f"basis{i}") for i in range(8)]

        def pair(i, k):
            a = f"basis{i}" if (k >> 1) & 1 else nb[i]
            b = f"basis{i + 1}" if k & 1 else nb[i + 1]
            return self.AND(a, b)

        p01 = [pair(0, k) for k in range(4)]
        p23 = [pair(2, k) for k in range(4)]
        p45 = [pair(4, k) for k in range(4)]
        p67 = [pair(6, k) for k in range(4)]
        self.nib_hi = [self.emit(self.fresh(f"nib_hi_{h:x}"), f"({p01[h >> 2]} & {p23[h & 3]})") for h in range(16)]
        self.nib_lo = [self.emit(self.fresh(f"nib_lo_{l:x}"), f"({p45[l >> 2]} & {p67[l & 3]})") for l in range(16)]

    def _low_fn(self, lowmask):
        """Function of basis4..7 selecting the low nibbles in the 16-bit mask."""
        n = bin(lowmask).count("1")
        if n == 16:
            return "1"
        if n <= 2:
            out = "0"
            for l in range(16):
                if lowmask >> l & 1:
                    out = self.OR(out, self.nib_lo[l])
            return out
        if n >= 14:
            out = "0"
            for l in range(16):
                if not lowmask >> l & 1:
                    out = self.OR(out, self.nib_lo[l])
            return self.NOT(out)
        return self._cc_fn(4, lowmask)

    def _cc_nibble(self, mask):
        if bin(mask).count("1") > 128:
            return self.NOT(self._cc_nibble(ALL & ~mask))
        rows = {}
        for h in range(16):
            row = (mask >> (16 * h)) & 0xFFFF
            if row:
                rows.setdefault(row, []).append(h)
        out = "0"
        for row, hs in sorted(rows.items(), key=lambda kv: kv[1][0]):
            hi = "0"
            for h in hs:
                hi = self.OR(hi, self.nib_hi[h])
            out = self.OR(out, self.AND(hi, self._low_fn(row)))
        return out

    # -- group-level streams --------------------------------------------------
    def lf(self):
        if "lf" not in self.group_memo:
            saved = self.scopes
            self.scopes = [self.group_memo.setdefault("_scope", {})]
            self.group_memo["lf"] = self.cc(_bit(LF), keep_lf=True)
            self.scopes = saved
        return self.group_memo["lf"]

    def sol(self):
        if "sol" not in self.group_memo:
            lf = self.lf()
            saved = self.scopes
            self.scopes = [self.group_memo["_scope"]]
            self.group_memo["sol"] = self.NOT(self.ADV(self.NOT(lf), 1))
            self.scopes = saved
        return self.group_memo["sol"]

    # -- markers --------------------------------------------------------------
    def consume(self, m, mask):
        c = self.cc(mask)
        return self.ADV(self.AND(m, c), 1)

    def compile(self, node, m):
        if m == "0":
            return "0"
        if isinstance(node, CC):
            return self.consume(m, node.mask)
        if isinstance(node, Seq):
            items = node.items
            i = 0
            while i < len(items):
                it = items[i]
                nxt = items[i + 1] if i + 1 < len(items) else None
                if (isinstance(it, Rep) and it.hi is None and isinstance(it.node, CC)
                        and isinstance(nxt, CC)):
                    c = it.node.mask & (~_bit(LF) if self.line_mode else ALL)
                    d = nxt.mask & (~_bit(LF) if self.line_mode else ALL)
                    if c & d == 0:
                        # star of c followed by a class disjoint from c: only
                        # the landing positions of c-runs (and m itself) can
                        # consume d, so MatchStar(m, c) & d ==
                        # (ScanThru(m & c, c) | m) & d.  Markers of m & c lie
                        # inside c-runs, so no carry lands on a marker.
                        for _ in range(it.lo):
                            m = self.consume(m, it.node.mask)
                        m = self.scan_star(m, it.node.mask)
                        i += 1
                        continue
                m = self.compile(it, m)
                if m == "0":
                    return "0"
                i += 1
            return m
        if isinstance(node, Alt):
            out = "0"
            for br in node.items:
                out = self.OR(out, self.compile(br, m))
            return out
        if isinstance(node, Rep):
            for _ in range(node.lo):
                m = self.compile(node.node, m)
            if node.hi is None:
                return self.star(node.node, m)
            acc = m
            for _ in range(node.hi - node.lo):
                m = self.compile(node.node, m)
                acc = self.OR(acc, m)
            return acc
        if isinstance(node, Start):
            return self.AND(m, self.sol())
        if isinstance(node, End):
            return self.AND(m, self.lf())
        raise TypeError(node)

    def scan_star(self, m, mask):
        c = self.cc(mask)
        if c == "0" or m in ("0", "1"):
            return m
        t = self.AND(m, c)
        s = self.emit(self.fresh("scan"), f"pablo.ScanThru({t}, {c})")
        return self.OR(s, m)

    def star(self, node, m):
        if isinstance(node, CC) and self.star_mode == "matchstar":
            c = self.cc(node.mask)
            if c == "0" or m == "1":
                return m
            return self.emit(self.fresh("star"), f"pablo.MatchStar({m}, {c})")
        if m == "1":
            return "1"
        if isinstance(node, CC):
            self.cc(node.mask)  # class computed once, outside the loop
        # Fixed-point loop: accum collects every marker reachable by zero or
        # more repetitions; test holds the markers new in this iteration.
        test = self.fresh("test")
        accum = self.fresh("accum")
        self.emit(test, m)
        self.emit(accum, m)
        self.lines.append(f"{self.indent}while {test}:")
        saved = self.indent
        self.indent += "  "
        self.scopes.append({})
        nxt = self.compile(node, test)
        if nxt == "0":
            nxt = self.emit(self.fresh("zero"), "0")
        nacc = self.emit(self.fresh("not"), f"(~{accum})")
        new = self.emit(self.fresh("new"), f"({nxt} & {nacc})")
        self.emit(accum, f"({accum} | {nxt})")
        self.emit(test, new)
        self.scopes.pop()
        self.indent = saved
        # accum is reassigned in the loop: forget memo entries naming it
        return accum

    # -- group ----------------------------------------------------------------
    def emit_group(self, regexes):
        asts = [parse_regex(r) for r in regexes]
        group_scope = self.group_memo.setdefault("_scope", {})
        # group header: shared streams are defined before any regex code
        self.scopes = [group_scope]
        if self.cc_mode == "nibble":
            self.emit_nibble_masks()
        if self.line_mode:
            self.lf()
        if any(_has_start(a) for a in asts):
            self.sol()
        acc = "0"
        for ast in asts:
            # common subexpressions are shared inside one regex only, so every
            # regex's classes are computed next to their use
            self.scopes = [group_scope, {}]
            acc = self.OR(acc, self.compile(ast, "1"))
        self.scopes = [group_scope, {}]
        if self.line_mode and acc != "0":
            lf = self.lf()
            if acc == "1":
                res = lf
            else:
                # MatchStar, not ScanThru: a marker sitting on an LF while
                # another marker of the same line scans toward it must not
                # push the carry into the next line.
                nlf = self.NOT(lf)
                ends = self.emit(self.fresh("lineends"), f"pablo.MatchStar({acc}, {nlf})")
                res = self.emit(self.fresh("matchedlines"), f"({ends} & {lf})")
        else:
            res = acc
        self.emit("output.matches", res)
        return "\n".join(self.lines) + "\n"


def _has_start(node):
    if isinstance(node, Start):
        return True
    if isinstance(node, (Seq, Alt)):
        return any(_has_start(i) for i in node.items)
    if isinstance(node, Rep):
        return _has_start(node.node)
    return False


def regexes_to_pablo(regexes, star_mode="matchstar", line_mode=True, cc_mode="nibble"):
    """Pablo text for one group of regexes (strings without slashes)."""
    if isinstance(regexes, str):
        regexes = [regexes]
    return PabloEmitter(star_mode=star_mode, line_mode=line_mode, cc_mode=cc_mode).emit_group(list(regexes))
