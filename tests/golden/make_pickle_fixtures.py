"""Convert the reference's ridge fixtures ridges.pkl / ordered_ridges.pkl
into plain .npz data (ridges.npz, ordered_ridges.npz) WITHOUT unpickling.

The two files are lists of float64 numpy arrays written by pickle protocol 4
(tests/conftest.py:134-168 of the reference loads them with pickle.load).
torch's weights-only loader refuses them (SHORT_BINBYTES / STACK_GLOBAL), and
a full unpickler would execute whatever the stream names.  Instead this
script walks the opcode stream with pickletools.genops (a parser: it never
imports or calls anything) and interprets the small opcode subset numpy's
ndarray.__reduce__ emits on a tiny stack machine of its own:

  * globals are accepted only from ALLOWED below (numpy's array reconstruct
    triple) and are kept as inert name tokens, never resolved;
  * REDUCE of _reconstruct / dtype and BUILD of their states are evaluated
    by hand: the dtype string must be a little-endian float64 ('f8' with
    byte order '<'), the array shape a tuple of ints, and the payload bytes
    are turned into an array by numpy.frombuffer;
  * any other opcode or global raises.

Run once in a container that holds the reference checkout (path given as
the first argument, default /root/reference).  Output arrays are stored as
r0, r1, ... in list order.
"""
import os
import pickletools
import sys

import numpy as np

ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

ALLOWED = {
    ("numpy.core.multiarray", "_reconstruct"),
    ("numpy._core.multiarray", "_reconstruct"),  # the same, as numpy >= 2 names it
    ("numpy", "ndarray"),
    ("numpy", "dtype"),
}


class _Global:
    def __init__(self, module, name):
        if (module, name) not in ALLOWED:
            raise ValueError(f"global {module}.{name} not allowed")
        self.key = (module, name)


class _Reduced:
    """Result of REDUCE: (callable token, args), state filled by BUILD."""

    def __init__(self, fn, args):
        self.fn, self.args, self.state = fn, args, None


_MARK = object()


def _pop_mark(stack):
    items = []
    while True:
        v = stack.pop()
        if v is _MARK:
            return items[::-1]
        items.append(v)


def _finish(obj):
    """Turn the reduced/built tokens into numpy arrays."""
    if isinstance(obj, list):
        return [_finish(o) for o in obj]
    if not isinstance(obj, _Reduced) or obj.fn.key[1] != "_reconstruct":
        raise ValueError("top-level items must be reconstructed arrays")
    cls, shape0, _ = obj.args
    if cls.key != ("numpy", "ndarray") or shape0 != (0,):
        raise ValueError("unexpected ndarray reconstruct arguments")
    version, shape, dt, fortran, payload = obj.state
    if version != 1 or not all(isinstance(s, int) for s in shape) or fortran:
        raise ValueError("unexpected ndarray state")
    if dt.fn.key != ("numpy", "dtype") or dt.args[0] != "f8":
        raise ValueError("only float64 arrays are accepted")
    if dt.state[1] != "<":
        raise ValueError("only little-endian arrays are accepted")
    arr = np.frombuffer(bytes(payload), dtype="<f8")
    if arr.size != int(np.prod(shape)):
        raise ValueError("payload size does not match the shape")
    return arr.reshape(shape).copy()


def read_array_list(path):
    data = open(path, "rb").read()
    stack, memo = [], []
    result = None
    for op, arg, _pos in pickletools.genops(data):
        name = op.name
        if name in ("PROTO", "FRAME"):
            continue
        if name == "EMPTY_LIST":
            stack.append([])
        elif name == "MARK":
            stack.append(_MARK)
        elif name == "MEMOIZE":
            memo.append(stack[-1])
        elif name in ("BINGET", "LONG_BINGET"):
            stack.append(memo[arg])
        elif name in ("SHORT_BINUNICODE", "BINUNICODE"):
            stack.append(arg)
        elif name in ("SHORT_BINBYTES", "BINBYTES"):
            stack.append(bytes(arg))
        elif name in ("BININT1", "BININT2", "BININT"):
            stack.append(int(arg))
        elif name == "NONE":
            stack.append(None)
        elif name == "NEWTRUE":
            stack.append(True)
        elif name == "NEWFALSE":
            stack.append(False)
        elif name == "STACK_GLOBAL":
            nm = stack.pop()
            mod = stack.pop()
            stack.append(_Global(mod, nm))
        elif name == "TUPLE1":
            stack.append((stack.pop(),))
        elif name == "TUPLE2":
            b = stack.pop()
            a = stack.pop()
            stack.append((a, b))
        elif name == "TUPLE3":
            c = stack.pop()
            b = stack.pop()
            a = stack.pop()
            stack.append((a, b, c))
        elif name == "TUPLE":
            stack.append(tuple(_pop_mark(stack)))
        elif name == "REDUCE":
            args = stack.pop()
            fn = stack.pop()
            if not isinstance(fn, _Global):
                raise ValueError("REDUCE of a non-global")
            stack.append(_Reduced(fn, args))
        elif name == "BUILD":
            state = stack.pop()
            target = stack[-1]
            if not isinstance(target, _Reduced) or target.state is not None:
                raise ValueError("BUILD on an unexpected object")
            target.state = state
        elif name == "APPENDS":
            items = _pop_mark(stack)
            stack[-1].extend(items)
        elif name == "APPEND":
            v = stack.pop()
            stack[-1].append(v)
        elif name == "STOP":
            result = stack.pop()
            break
        else:
            raise ValueError(f"opcode {name} not supported")
    if not isinstance(result, list):
        raise ValueError("top-level object is not a list")
    return _finish(result)


def main():
    src = os.path.join(ROOT, "tests", "testing_data")
    for stem in ("ridges", "ordered_ridges"):
        arrs = read_array_list(os.path.join(src, stem + ".pkl"))
        out = os.path.join(HERE, stem + ".npz")
        np.savez(out, **{f"r{k}": a for k, a in enumerate(arrs)})
        print(stem, [a.shape for a in arrs], "->", out)


if __name__ == "__main__":
    main()
