"""Bitstream IR: instruction node types.

Same node vocabulary as the reference (``bitgen/inst/inst.py:8-31``): a flat
list of ``BsInst`` objects whose operands and results are integer variable ids
resolved through a ``var_name_map`` (id -> name).  Structured control flow
(``BsWhile``, ``BsIf``) holds nested instruction lists.

Semantics (whole-stream, position p = byte offset; bit p of word p//64,
LSB-first):

* ``BsAnd/BsOr/BsXor/BsNot/BsAssign`` -- bitwise, per position.
* ``BsAdvance(x, n)`` -- shift toward higher positions: y[p] = x[p-n], zero
  fill (``bitgen/template/template.py:63-70``).  With ``operation`` containing
  ``"Left"`` it is the cc-advance look-ahead form: y[p] = x[p+n].
* ``BsSel(c ? a : b)`` -- (c & a) | (~c & b) (``bitgen/inst/inst_sel.py:42-57``).
* ``BsMatchStar(M, C)`` = (((M & C) + C) ^ C) | M and ``BsScanThru(M, C)`` =
  (M + C) & ~C, with ``+`` an addition whose carries flow toward higher
  positions (``bitgen/template/template.py:96-102``,
  ``bitgen/template/template_torch.py:377-389``).
* ``BsWhile(cond, body)`` -- repeat body while cond has any set bit
  (``bitgen/inst/inst_block_while.py:61-68``).
* ``BsIf(cond, then, else)`` -- run ``then`` if cond has any set bit.
* ``BsSkipZero(v, label)`` -- zero-block skipping: when v has no set bit in the
  unit of execution, jump to ``BsLabel(label)``.  The pass that inserts it
  guarantees every value defined in the skipped range and read afterwards is
  zero-initialised and provably zero whenever v is zero, so the jump changes no
  result; it is the typed form of the reference's
  ``if (block_all_zeros(v, &zero_flag)) goto LABELk;`` string
  (``bitgen/passes/short_circuit.py:318-326``).
"""
from enum import Enum
import copy


class BsInstType(Enum):
    AND = 0
    OR = 1
    NOT = 2
    XOR = 3
    ADD = 4
    MATCHSTAR = 5
    SCANTHRU = 6
    ASSIGN = 7
    SEL = 8
    IF = 9
    WHILE = 10
    ADVANCE = 11
    TERNARY = 12
    MATCH = 13
    CALL = 14
    BLOCK = 15
    STATE = 16
    GRAPHBREAK = 17
    STREAMSTORE = 18
    STREAMLOAD = 19
    STREAMDEFINE = 20
    SCANTHRUSTREAM = 21
    STR = 22
    SKIPZERO = 23
    LABEL = 24


class BsInst:
    n_operand = 0

    def __init__(self, inst_type, operands, ret):
        self.type = inst_type
        self.operands = list(operands) if operands is not None else []
        self.ret = ret
        self.comment = ""

    # reference-style accessors (bitgen/inst/*.py use operand1/operand2)
    @property
    def operand1(self):
        return self.operands[0]

    @operand1.setter
    def operand1(self, v):
        self.operands[0] = v

    @property
    def operand2(self):
        return self.operands[1]

    @operand2.setter
    def operand2(self, v):
        self.operands[1] = v

    def uses(self):
        """Variable ids read by this instruction (not counting nested bodies)."""
        return list(self.operands[: self.n_operand])

    def defs(self):
        return [] if self.ret is None else [self.ret]

    def replace_uses(self, mapping):
        for i in range(self.n_operand):
            op = self.operands[i]
            if op in mapping:
                self.operands[i] = mapping[op]

    def deepcopy(self):
        return copy.deepcopy(self)

    def __repr__(self):
        return f"{self.__class__.__name__}({self.operands}->{self.ret})"


class BsAnd(BsInst):
    n_operand = 2

    def __init__(self, operands, ret):
        super().__init__(BsInstType.AND, operands, ret)


class BsOr(BsInst):
    n_operand = 2

    def __init__(self, operands, ret):
        super().__init__(BsInstType.OR, operands, ret)


class BsXor(BsInst):
    n_operand = 2

    def __init__(self, operands, ret):
        super().__init__(BsInstType.XOR, operands, ret)


class BsNot(BsInst):
    n_operand = 1

    def __init__(self, operands, ret):
        super().__init__(BsInstType.NOT, operands, ret)


class BsAssign(BsInst):
    n_operand = 1

    def __init__(self, operands, ret):
        super().__init__(BsInstType.ASSIGN, operands, ret)


ADV_RIGHT = "BSAdvanceRightFunctionSync"
ADV_LEFT = "BSAdvanceLeftFunction"


class BsAdvance(BsInst):
    """Advance by a constant; ``operands = [var, n]``.

    ``operation`` keeps the reference's spelling
    (``bitgen/inst/inst_advance.py:4-12``, ``bitgen/passes/cc_advance.py:156-286``):
    names containing ``Left`` shift toward lower positions (look-ahead).
    """

    n_operand = 1

    def __init__(self, operands, ret, operation=ADV_RIGHT):
        super().__init__(BsInstType.ADVANCE, [operands[0], int(operands[1])], ret)
        self.operation = operation

    @property
    def amount(self):
        return int(self.operands[1])

    @property
    def is_left(self):
        return "Left" in self.operation

    def __repr__(self):
        d = "L" if self.is_left else "R"
        return f"BsAdvance{d}({self.operands[0]},{self.amount}->{self.ret})"


class BsMatchStar(BsInst):
    n_operand = 2

    def __init__(self, operands, ret):
        super().__init__(BsInstType.MATCHSTAR, operands, ret)


class BsScanThru(BsInst):
    n_operand = 2

    def __init__(self, operands, ret):
        super().__init__(BsInstType.SCANTHRU, operands, ret)


class BsSel(BsInst):
    """``ret = condition ? operands[0] : operands[1]``."""

    n_operand = 2

    def __init__(self, operands, condition, ret):
        super().__init__(BsInstType.SEL, operands, ret)
        self.condition = condition

    def uses(self):
        return [self.condition] + list(self.operands[:2])

    def replace_uses(self, mapping):
        super().replace_uses(mapping)
        self.condition = mapping.get(self.condition, self.condition)


class BsBlock(BsInst):
    def __init__(self, body, ret=None):
        super().__init__(BsInstType.BLOCK, [], ret)
        self.body = list(body)


class BsWhile(BsInst):
    def __init__(self, body, condition):
        ret = body[-1].ret if body else None
        super().__init__(BsInstType.WHILE, [], ret)
        self.body = list(body)
        self.condition = condition

    def uses(self):
        return [self.condition]

    def defs(self):
        return []

    def replace_uses(self, mapping):
        self.condition = mapping.get(self.condition, self.condition)

    def __repr__(self):
        return f"While[{self.condition}]{self.body}"


class BsIf(BsInst):
    def __init__(self, body_1, body_2, condition):
        super().__init__(BsInstType.IF, [], None)
        self.body_1 = list(body_1)
        self.body_2 = list(body_2) if body_2 is not None else None
        self.condition = condition

    def uses(self):
        return [self.condition]

    def defs(self):
        return []

    def replace_uses(self, mapping):
        self.condition = mapping.get(self.condition, self.condition)

    def __repr__(self):
        return f"If[{self.condition}]{self.body_1} else {self.body_2}"


class BsStreamLoad(BsInst):
    """Load basis stream ``stream_index`` (0..7, basis0 = most significant bit)."""

    n_operand = 0

    def __init__(self, operands, ret, stream_index=0):
        super().__init__(BsInstType.STREAMLOAD, operands, ret)
        self.stream_index = stream_index


class BsStreamStore(BsInst):
    """Store ``operands[0]`` to the group's result stream (``bs_result_stream``)."""

    n_operand = 1

    def __init__(self, operands, ret):
        super().__init__(BsInstType.STREAMSTORE, operands, ret)
        self.use_or = False

    def defs(self):
        return []


class BsGraphBreak(BsInst):
    def __init__(self):
        super().__init__(BsInstType.GRAPHBREAK, [], None)


class BsStr(BsInst):
    def __init__(self, operands):
        super().__init__(BsInstType.STR, list(operands), None)

    @property
    def str(self):
        return self.operands[0]

    def uses(self):
        return []


class BsSkipZero(BsInst):
    def __init__(self, var, label):
        super().__init__(BsInstType.SKIPZERO, [var], None)
        self.label = label

    n_operand = 1

    def __repr__(self):
        return f"SkipZero({self.operands[0]}->L{self.label})"


class BsLabel(BsInst):
    def __init__(self, label):
        super().__init__(BsInstType.LABEL, [], None)
        self.label = label

    def __repr__(self):
        return f"Label{self.label}"


def walk(insts):
    """Yield every instruction, descending into loop and branch bodies."""
    for inst in insts:
        yield inst
        if inst.type == BsInstType.WHILE:
            yield from walk(inst.body)
        elif inst.type == BsInstType.IF:
            yield from walk(inst.body_1)
            if inst.body_2:
                yield from walk(inst.body_2)


def count_insts(insts):
    return sum(1 for _ in walk(insts))
