"""Circuit / operator construction from strings (reference
src/qforte/helper/io_helper.py:4-94)."""

from .. import _core


def build_circuit(Inputstr):
    """``"H_0 cX_1_0 Rz_2_0.5"`` -> Circuit (io_helper.py:52-75).

    ``G_t`` is a one-qubit gate on t; ``G_t_c`` a two-qubit gate with target t
    and control c, unless the gate name contains "R", in which case the third
    field is the rotation parameter.
    """
    circ = _core.Circuit()
    for token in Inputstr.split():
        parts = token.split("_")
        if len(parts) == 2:
            circ.add(_core.gate(parts[0], int(parts[1]), int(parts[1])))
        elif "R" in parts[0]:
            circ.add(_core.gate(parts[0], int(parts[1]), int(parts[1]), float(parts[2])))
        else:
            circ.add(_core.gate(parts[0], int(parts[1]), int(parts[2])))
    return circ


def build_operator(Inputstr):
    """``"c1, circ1; c2, circ2"`` -> QubitOperator (io_helper.py:87-94)."""
    ops = _core.QubitOperator()
    for term in Inputstr.split(";"):
        coeff, circ = term.split(",")
        ops.add(complex(coeff), build_circuit(circ))
    return ops


def smart_print(Inputobj):
    """Print a QubitOperator, Circuit or Computer (io_helper.py:4-40)."""
    if isinstance(Inputobj, _core.QubitOperator):
        print("\n Quantum operator:")
        for k, term in enumerate(Inputobj.terms()):
            if k:
                print("+", end=" ")
            print(term[0], end=" ")
            print("(" + " ".join(term[1].str()[1:-1].split(" ")) + ")", "|Ψ>")
    if isinstance(Inputobj, _core.Circuit):
        print("\n Quantum circuit:")
        print("(" + " ".join(Inputobj.str()[1:-1].split(" ")) + ")", "|Ψ>")
    if isinstance(Inputobj, _core.Computer):
        print("\n Quantum Computer:")
        print(Inputobj.str())
