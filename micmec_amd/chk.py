"""Reader and writer for the CHK text checkpoint format.

MicMec stores systems in molmod's CHK format (micmec/system.py:213-259, 298-335). Each record
is one header line, ``<key padded to 40> kind=<kind> <shape or scalar>``, followed, for arrays,
by the values four per line (``%22.15e`` for floats, ``%22i`` for integers). Kinds seen in
the reference fixtures (data/*_micmec.chk) are ``int``, ``flt``, ``str``, ``intar`` and
``fltar``; ``bln`` and ``blnar`` are handled too. Nothing in a file is executed: every value
is parsed from text.
"""

import numpy as np

__all__ = ["load_chk", "dump_chk"]


def _parse_scalar(kind, text):
    if kind == "int":
        return int(text)
    if kind == "flt":
        return float(text)
    if kind == "str":
        return text
    if kind == "bln":
        return text.strip() == "True"
    raise IOError("Unknown CHK scalar kind %r" % kind)


def load_chk(filename):
    """Load a CHK file into a dict of Python scalars, strings and numpy arrays."""
    result = {}
    with open(filename) as f:
        lines = f.read().split("\n")
    i = 0
    nlines = len(lines)
    while i < nlines:
        line = lines[i]
        i += 1
        if not line.strip():
            continue
        if "kind=" not in line:
            raise IOError("Malformed CHK header line %i in %s" % (i, filename))
        key = line[:40].strip()
        rest = line[line.index("kind=") + 5:]
        kind = rest[:5].strip()
        arg = rest[5:].strip() if kind in ("intar", "fltar", "blnar", "strar") else rest[3:].strip()
        if kind in ("int", "flt", "bln"):
            result[key] = _parse_scalar(kind, arg)
        elif kind == "str":
            # strings keep their internal spacing; molmod writes ``kind=str   <value>``
            result[key] = line[line.index("kind=") + 11:]
        elif kind in ("intar", "fltar", "blnar"):
            shape = tuple(int(s) for s in arg.split(",")) if arg else (0,)
            size = int(np.prod(shape)) if len(shape) else 1
            words = []
            while len(words) < size:
                if i >= nlines:
                    raise IOError("Truncated array %r in %s" % (key, filename))
                words.extend(lines[i].split())
                i += 1
            if kind == "intar":
                arr = np.array([int(w) for w in words[:size]], dtype=np.int64)
            elif kind == "fltar":
                arr = np.array([float(w) for w in words[:size]], dtype=np.float64)
            else:
                arr = np.array([w == "True" for w in words[:size]], dtype=bool)
            result[key] = arr.reshape(shape)
        else:
            raise IOError("Unsupported CHK kind %r for key %r" % (kind, key))
    return result


def _fmt_line(values, fmt):
    return " ".join(fmt % v for v in values)


def dump_chk(filename, data):
    """Write a dict to a CHK file, keys sorted, in the layout :func:`load_chk` reads."""
    with open(filename, "w") as f:
        for key in sorted(data):
            value = data[key]
            if len(key) > 40:
                raise ValueError("CHK keys are at most 40 characters: %r" % key)
            head = key.ljust(40)
            if isinstance(value, str):
                if "\n" in value:
                    raise ValueError("CHK strings may not contain newlines")
                f.write("%s kind=str   %s\n" % (head, value))
                continue
            if isinstance(value, (bool, np.bool_)):
                f.write("%s kind=bln   %s\n" % (head, bool(value)))
                continue
            if isinstance(value, (int, np.integer)):
                f.write("%s kind=int   %i\n" % (head, int(value)))
                continue
            if isinstance(value, (float, np.floating)):
                f.write("%s kind=flt   %22.15e\n" % (head, float(value)))
                continue
            arr = np.asarray(value)
            if arr.dtype == bool:
                kind, fmt, vals = "blnar", "%s", [str(bool(v)) for v in arr.ravel()]
            elif np.issubdtype(arr.dtype, np.integer):
                kind, fmt, vals = "intar", "%22i", [int(v) for v in arr.ravel()]
            elif np.issubdtype(arr.dtype, np.floating):
                kind, fmt, vals = "fltar", "%22.15e", [float(v) for v in arr.ravel()]
            else:
                raise TypeError("Cannot store %r of dtype %s in a CHK file" % (key, arr.dtype))
            f.write("%s kind=%s %s\n" % (head, kind, ",".join(str(s) for s in arr.shape)))
            for j in range(0, len(vals), 4):
                f.write(_fmt_line(vals[j:j + 4], fmt) + "\n")
