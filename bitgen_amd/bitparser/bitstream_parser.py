"""Pablo text -> bitstream IR.

Accepts the grammar of the reference parser
(``bitgen/bitparser/bitstream_parser.py:9-279``): one statement per line,
``#`` comments, indentation-delimited ``while c:`` and ``if c:`` / ``else:``
blocks, and the expressions

    a            (a)          ~a            (~a)          Next(a)
    (a & b)      (a | b)      (a ^ b)       (c ? a : b)
    pablo.Advance(a, n)       pablo.MatchStar(m, c)       pablo.ScanThru(m, c)

Differences from the reference, all additive: ``pablo.MatchStar`` and
``pablo.ScanThru`` are accepted (the reference raises NotImplementedError at
:206 and :211), ``if`` without ``else`` works (the reference crashes at :145),
and blank lines inside a block do not end it.

For the ``cuda`` and ``hip`` backends the parser prepends eight
``BsStreamLoad`` of ``basis0..7`` and appends a ``BsStreamStore`` of the last
statement's result into ``bs_result_stream`` (reference :251-276), and it
splits ``c ? a : b`` into not/and/and/or (reference :218-227).
"""
import re

from ..inst import (
    BsAnd, BsOr, BsXor, BsNot, BsAssign, BsAdvance, BsMatchStar, BsScanThru,
    BsSel, BsWhile, BsIf, BsStreamLoad, BsStreamStore, BsInst,
)
from .. import config as cfg

_FUNC_RE = re.compile(r"^\s*([\w\.]+)\s*\(\s*(.*?)\s*\)\s*$")
_BINOPS = {"&": BsAnd, "|": BsOr, "^": BsXor}
_STREAM_BACKENDS = ("cuda", "hip")


class PabloSyntaxError(ValueError):
    pass


class BitstreamParser:
    def __init__(self, backend=None):
        self.var_name_map = {}
        self.name_var_map = {}
        self.backend = backend

    # -- variables ---------------------------------------------------------
    def add_var(self, var_name):
        if var_name not in self.name_var_map:
            var_id = len(self.name_var_map)
            self.name_var_map[var_name] = var_id
            self.var_name_map[var_id] = var_name
        return self.name_var_map[var_name]

    def get_var(self, var_name):
        if var_name in self.name_var_map:
            return self.name_var_map[var_name]
        raise PabloSyntaxError("Invalid variable name: " + var_name)

    def _backend(self):
        return self.backend if self.backend is not None else cfg.get_config("backend")

    # -- expressions -------------------------------------------------------
    def parse_expr(self, expr):
        """Return (kind, payload) for the right-hand side of a statement."""
        expr = expr.strip()
        if not expr:
            raise PabloSyntaxError("empty expression")
        if expr[0] == "(" and expr[-1] == ")":
            expr = expr[1:-1].strip()
        m = _FUNC_RE.match(expr)
        if m and not expr.startswith("("):
            fname = m.group(1)
            args = [a.strip() for a in m.group(2).split(",")]
            if fname == "Next":
                return "var", args[0]
            if fname in ("pablo.Advance", "pablo.MatchStar", "pablo.ScanThru"):
                if len(args) != 2:
                    raise PabloSyntaxError(f"{fname} takes 2 arguments: {expr}")
                return fname, args
            raise PabloSyntaxError("Invalid function call: " + fname)
        tokens = expr.split()
        if len(tokens) == 5 and tokens[1] == "?" and tokens[3] == ":":
            return "sel", (tokens[0], tokens[2], tokens[4])
        if len(tokens) == 3 and tokens[1] in _BINOPS:
            return "bin", (tokens[1], tokens[0], tokens[2])
        if len(tokens) == 2 and tokens[0] == "~":
            return "not", tokens[1]
        if len(tokens) == 1:
            if tokens[0].startswith("~"):
                return "not", tokens[0][1:]
            return "var", tokens[0]
        raise PabloSyntaxError("Invalid expression: " + expr)

    # -- blocks ------------------------------------------------------------
    @staticmethod
    def get_indent(line):
        return len(line) - len(line.lstrip(" "))

    @staticmethod
    def _strip(line):
        i = line.find("#")
        if i != -1:
            line = line[:i]
        return line.rstrip()

    def _block_end(self, codes, start, indent_out):
        i = start
        while i < len(codes):
            s = self._strip(codes[i])
            if s.strip() and self.get_indent(s) <= indent_out:
                return i
            i += 1
        return len(codes)

    def _cond(self, line, kw):
        body = line.strip()[len(kw):].strip()
        if not body.endswith(":"):
            raise PabloSyntaxError(f"missing ':' after {kw}: {line!r}")
        return body[:-1].strip()

    def parse_branch_while(self, codes, line_id):
        line = self._strip(codes[line_id])
        cond = self.add_var(self._cond(line, "while"))
        indent_out = self.get_indent(line)
        end = self._block_end(codes, line_id + 1, indent_out)
        body = self.parse_lines(codes[line_id + 1:end])
        if not body:
            raise PabloSyntaxError(f"empty while body at line {line_id}")
        return end, BsWhile(body, cond)

    def parse_branch_if(self, codes, line_id):
        line = self._strip(codes[line_id])
        cond = self.add_var(self._cond(line, "if"))
        indent_out = self.get_indent(line)
        end1 = self._block_end(codes, line_id + 1, indent_out)
        body1 = self.parse_lines(codes[line_id + 1:end1])
        body2 = None
        end = end1
        if end1 < len(codes):
            nxt = self._strip(codes[end1])
            if nxt.strip().startswith("else") and self.get_indent(nxt) == indent_out:
                end = self._block_end(codes, end1 + 1, indent_out)
                body2 = self.parse_lines(codes[end1 + 1:end])
        return end, BsIf(body1, body2, cond)

    # -- statements --------------------------------------------------------
    def parse_line(self, codes, line_id):
        line = self._strip(codes[line_id]).strip()
        if not line:
            return line_id + 1, None
        if re.match(r"^while\b", line):
            return self.parse_branch_while(codes, line_id)
        if re.match(r"^if\b", line):
            return self.parse_branch_if(codes, line_id)
        if "=" not in line:
            raise PabloSyntaxError(f"not a statement: {line!r}")
        lhs, rhs = line.split("=", 1)
        kind, payload = self.parse_expr(lhs)
        if kind != "var":
            raise PabloSyntaxError(f"left side must be a variable: {line!r}")
        kind, payload = self.parse_expr(rhs)
        # operands are resolved before the target is declared, so ``x = (x & y)``
        # reads the old x.
        if kind == "var":
            src = self.get_var(payload)
            return line_id + 1, BsAssign([src], self.add_var(lhs.strip()))
        if kind == "not":
            src = self.get_var(payload)
            return line_id + 1, BsNot([src], self.add_var(lhs.strip()))
        if kind == "bin":
            op, a, b = payload
            va, vb = self.get_var(a), self.get_var(b)
            return line_id + 1, _BINOPS[op]([va, vb], self.add_var(lhs.strip()))
        if kind == "pablo.Advance":
            src = self.get_var(payload[0])
            try:
                n = int(payload[1])
            except ValueError:
                raise PabloSyntaxError(f"Advance amount must be an integer: {line!r}")
            if n < 0:
                raise PabloSyntaxError(f"negative Advance: {line!r}")
            return line_id + 1, BsAdvance([src, n], self.add_var(lhs.strip()))
        if kind in ("pablo.MatchStar", "pablo.ScanThru"):
            m, c = self.get_var(payload[0]), self.get_var(payload[1])
            cls = BsMatchStar if kind == "pablo.MatchStar" else BsScanThru
            return line_id + 1, cls([m, c], self.add_var(lhs.strip()))
        if kind == "sel":
            c, a, b = payload
            vc, va, vb = self.get_var(c), self.get_var(a), self.get_var(b)
            ret = self.add_var(lhs.strip())
            if self._backend() in _STREAM_BACKENDS:
                nc = self.add_var("sel_not_" + c)
                ma = self.add_var("sel_mask_" + a)
                mb = self.add_var("sel_mask_" + b)
                return line_id + 1, [
                    BsNot([vc], nc),
                    BsAnd([vc, va], ma),
                    BsAnd([nc, vb], mb),
                    BsOr([ma, mb], ret),
                ]
            return line_id + 1, BsSel([va, vb], vc, ret)
        raise PabloSyntaxError(f"unsupported statement: {line!r}")

    def parse_lines(self, codes):
        insts = []
        i = 0
        while i < len(codes):
            i, inst = self.parse_line(codes, i)
            if inst is None:
                continue
            if isinstance(inst, list):
                insts.extend(inst)
            elif isinstance(inst, BsInst):
                insts.append(inst)
            else:
                raise PabloSyntaxError("Invalid instruction")
        return insts

    def parse_text(self, text):
        insts = []
        stream_mode = self._backend() in _STREAM_BACKENDS
        if stream_mode:
            basis_stream = self.add_var("basic_stream")
            for i in range(8):
                insts.append(BsStreamLoad([basis_stream], self.add_var(f"basis{i}"), i))
        else:
            for i in range(8):
                self.add_var(f"basis{i}")
        self.add_var("0")
        self.add_var("1")
        body = self.parse_lines(text.split("\n"))
        if not body:
            raise PabloSyntaxError("empty pablo program")
        insts.extend(body)
        if stream_mode:
            last = insts[-1].ret
            if last is None:
                raise PabloSyntaxError("program must end with an assignment")
            store = BsStreamStore([last], self.add_var("bs_result_stream"))
            store.use_or = True
            insts.append(store)
        return insts

    def parse_file(self, file_path):
        with open(file_path, "r") as f:
            return self.parse_text(f.read())
