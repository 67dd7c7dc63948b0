from .inst import (  # noqa: F401
    BsInstType, BsInst, BsAnd, BsOr, BsXor, BsNot, BsAssign, BsAdvance,
    BsMatchStar, BsScanThru, BsSel, BsBlock, BsWhile, BsIf, BsStreamLoad,
    BsStreamStore, BsGraphBreak, BsStr, BsSkipZero, BsLabel, ADV_LEFT,
    ADV_RIGHT, walk, count_insts,
)
