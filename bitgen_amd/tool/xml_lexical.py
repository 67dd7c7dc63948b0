"""The XML lexical program of ``--xml`` (``bitgen/from_icgrep.py:390-425``).

The reference ships this program as pablo text
(``bitgen/bitparser/xml_lexical_parsing.pablo``).  It is restated here from
what it computes, with this package's class emitter:

* name characters N = [A-Za-z], and the markers '<' and '>';
* start markers: the position after each '<' whose character is not a name
  character (``</``, ``<?``, ``<!``, ``< ``);
* a loop that moves the markers forward over name characters one position
  per iteration, keeping in ``accum`` the markers of the iteration before
  the last (the loop ends when no marker anywhere in the stream is left:
  a whole-stream condition);
* the result: those markers that do not sit on a '>'.

The loop's result depends on the longest run anywhere in the input, so the
hip backend runs it at stream level (``codegen/gen_stream.py``).
"""
from .regex2pablo import PabloEmitter, _bit, _range

XML_NAME = "xml_lexical_parsing.pablo"


def xml_lexical_pablo():
    e = PabloEmitter(line_mode=False, cc_mode="bdd")
    name = e.cc(_range(ord("A"), ord("Z")) | _range(ord("a"), ord("z")))
    lt = e.cc(_bit(ord("<")))
    gt = e.cc(_bit(ord(">")))
    start = e.AND(e.ADV(lt, 1), e.NOT(name))
    test, accum = e.fresh("test"), e.fresh("accum")
    e.emit(test, start)
    e.emit(accum, start)
    e.lines.append(f"while {test}:")
    e.indent = "  "
    e.scopes.append({})
    e.emit(accum, test)
    step = e.emit(e.fresh("and"), f"({test} & {name})")
    e.emit(test, f"pablo.Advance({step}, 1)")
    e.scopes.pop()
    e.indent = ""
    not_gt = e.emit(e.fresh("not"), f"(~{gt})")
    e.emit("output.matches", e.emit(e.fresh("and"), f"({accum} & {not_gt})"))
    return "\n".join(e.lines) + "\n"
