"""The model's state dumps (the event format of the Fortran module `pydyn_dump`), read into NumPy
or straight into device arrays, and written (for fixtures).

The format (pydyn/dump.py:1-8, 113-232): one event is `<dir>/e<seq:05d>_<tag>_r<region>_p<rank:03d>.meta`,
text lines `key = value` (a value is T/F, an int, a float, a tuple of ints or text), an array
declared by `array NAME = lb1 ub1 lb2 ub2 ...` (8-byte reals) or `iarray NAME = ...` (4-byte
ints), and one raw file `<prefix>_<NAME>.bin` per array in Fortran order.  Every header carries
tag, seq, region, im, jm, lm.

Fortran order is i fastest, the layout the device arrays of this library use, so
`Event.device(name)` uploads a dumped array's bytes as they lie: a (lb..ub, ...) array becomes a
device tensor of that logical shape with stride 1 along i, without a transpose on the host.
`Event.hview(name)` is its `tm5_hview` (the Fortran lower bounds kept), the form the zoom entry
points take.
"""

import re
from pathlib import Path

import numpy as np

__all__ = ["DumpError", "Event", "FArray", "INTERIOR", "interior_section", "load_events", "to_state", "write_event"]

_NAME = re.compile(r"^e(\d{5})_(.+)_r(\d+)_p(\d{3})\.meta$")
REQUIRED = ("tag", "seq", "region", "im", "jm", "lm")

# the pydyn-layout interior of each array a dump carries: halo cells dropped, the fluxes'
# boundary faces kept (pydyn/layout.py:38-48); {im1} is im + 1 and so on
INTERIOR = {
    **dict.fromkeys(("m", "rm", "rxm", "rym", "rzm"), "1:{im}, 1:{jm}, 1:{lm}"),
    "am": "0:{im}, 1:{jm}, 1:{lm}",
    "bm": "1:{im}, 1:{jm1}, 1:{lm}",
    **dict.fromkeys(("cm", "mfw1", "mfw2", "pw1", "pw2"), "1:{im}, 1:{jm}, 0:{lm}"),
    **dict.fromkeys(("pu", "pu1", "pu2", "mfu1", "mfu2"), "0:{im}, 1:{jm}, 1:{lm}"),
    **dict.fromkeys(("pv", "pv1", "pv2", "mfv1", "mfv2"), "1:{im}, 1:{jm1}, 1:{lm}"),
    **dict.fromkeys(("sp", "sp1", "sp2"), "1:{im}, 1:{jm}, 1"),
    "phlb": "1:{im}, 1:{jm}, 1:{lm1}",
    "temper": "1:{im}, 1:{jm}, 1:{lm}",
}


class DumpError(ValueError):
    """A file that is not what the dump module writes: a truncated or foreign header."""


def interior_section(name, im, jm, lm):
    """The inclusive Fortran section (slices and ints) of `name` that is its interior."""
    text = INTERIOR[name].format(im=im, im1=im + 1, jm=jm, jm1=jm + 1, lm=lm, lm1=lm + 1)
    out = []
    for part in text.split(","):
        lo, colon, hi = part.strip().partition(":")
        out.append(slice(int(lo), int(hi)) if colon else int(lo))
    return tuple(out)


class FArray:
    """A NumPy array addressed by the inclusive Fortran bounds it was declared with:
    `a[0:1, 2]` is `a(0:1, 2)`; axes left out are taken whole."""

    def __init__(self, data, lower):
        self.data = data
        self.lower = tuple(int(x) for x in lower)

    @property
    def upper(self):
        return tuple(lo + n - 1 for lo, n in zip(self.lower, self.data.shape))

    def _ix(self, key):
        key = key if isinstance(key, tuple) else (key,)
        if len(key) > len(self.lower):
            raise IndexError(f"{len(key)} indices for a {len(self.lower)}-d array")
        ix = []
        for lo, k in zip(self.lower, key):
            if isinstance(k, slice):
                if k.step not in (None, 1):
                    raise IndexError("Fortran sections have unit stride")
                ix.append(slice(None if k.start is None else k.start - lo, None if k.stop is None else k.stop - lo + 1))
            else:
                ix.append(int(k) - lo)
        return tuple(ix)

    def __getitem__(self, key):
        return np.ascontiguousarray(self.data[self._ix(key)])

    def __repr__(self):
        return "FArray(" + ", ".join(f"{lo}:{hi}" for lo, hi in zip(self.lower, self.upper)) + ")"


def _value(text):
    words = text.split()
    if len(words) == 1:
        w = words[0]
        if w in ("T", "F"):
            return w == "T"
        for kind in (int, float):
            try:
                return kind(w)
            except ValueError:
                pass
        return w
    try:
        return tuple(int(w) for w in words)
    except ValueError:
        return text.strip()


class Event:
    """One dump event: `meta` (the header's scalars and tuples), `tag`, `seq`, `region`, `rank`,
    and its arrays by name, read from their .bin files on first use."""

    def __init__(self, meta_path):
        self.path = Path(meta_path)
        self.prefix = self.path.with_suffix("")
        self.meta, self.bounds, self.dtypes, self._cache = {}, {}, {}, {}
        try:
            text = self.path.read_text()
        except (OSError, UnicodeDecodeError) as e:
            raise DumpError(f"{self.path}: not readable as a dump header ({e})") from e
        for n, line in enumerate(text.splitlines(), 1):
            key, eq, value = line.partition("=")
            if not eq:
                continue
            key = key.strip()
            if key.startswith(("array ", "iarray ")):
                kind, _, name = key.partition(" ")
                name = name.strip()
                words = value.split()
                if not words or len(words) % 2 or not all(re.fullmatch(r"-?\d+", w) for w in words):
                    raise DumpError(f"{self.path}:{n}: not an array line: {line.strip()[:80]}")
                pairs = [(int(words[k]), int(words[k + 1])) for k in range(0, len(words), 2)]
                if not name.isidentifier() or any(hi < lo - 1 for lo, hi in pairs):
                    raise DumpError(f"{self.path}:{n}: bad array bounds or name: {line.strip()[:80]}")
                self.bounds[name] = pairs
                self.dtypes[name] = np.dtype("<i4") if kind == "iarray" else np.dtype("<f8")
            else:
                self.meta[key] = _value(value)
        missing = [k for k in REQUIRED if k not in self.meta]
        if missing:
            raise DumpError(f"{self.path}: no {', '.join(missing)} in the header (a truncated dump?)")
        self.tag, self.seq, self.region = self.meta["tag"], self.meta["seq"], self.meta["region"]
        self.rank = self.meta.get("rank", 0)

    @property
    def arrays(self):
        return tuple(self.bounds)

    def __contains__(self, name):
        return name in self.bounds

    def shape(self, name):
        return tuple(hi - lo + 1 for lo, hi in self.bounds[name])

    def lower(self, name):
        return tuple(lo for lo, _ in self.bounds[name])

    def raw(self, name):
        """The array's values in file order (Fortran order), checked against its bounds."""
        raw = np.fromfile(f"{self.prefix}_{name}.bin", dtype=self.dtypes[name])
        if raw.size != int(np.prod(self.shape(name))):
            raise DumpError(f"{self.prefix}_{name}.bin: {raw.size} values for bounds {self.bounds[name]}")
        return raw

    def __getitem__(self, name):
        if name not in self._cache:
            self._cache[name] = FArray(self.raw(name).reshape(self.shape(name), order="F"), self.lower(name))
        return self._cache[name]

    def interior(self, name):
        """The pydyn-layout interior of a state or flux array: the section INTERIOR[name]."""
        return self[name][interior_section(name, self.meta["im"], self.meta["jm"], self.meta["lm"])]

    def device(self, name):
        """The whole array as a device tensor of its declared shape, i fastest (the file's
        bytes uploaded as they lie)."""
        from . import device as D

        torch = D._torch()
        shape = self.shape(name)
        dtype = torch.float64 if self.dtypes[name].itemsize == 8 else torch.int32
        flat = torch.from_numpy(self.raw(name)).to("cuda")
        return flat.view(tuple(reversed(shape))).permute(*reversed(range(len(shape)))).to(dtype)

    def hview(self, name, t=None):
        """tm5_hview of the array on the device (t: a tensor from `device`, kept alive by the caller)."""
        from . import device as D

        return D.hview(self.device(name) if t is None else t, self.lower(name))

    def release(self):
        """Forget the arrays read so far (a dump set holds tens of GB)."""
        self._cache.clear()

    def __repr__(self):
        return f"Event(seq={self.seq}, tag={self.tag!r}, region={self.region}, rank={self.rank})"


def load_events(dump_dir, tags=None, region=None, rank=None):
    """The events of a dump directory in sequence order (then rank), optionally only some
    tags, one region, one rank."""
    out = []
    for p in sorted(Path(dump_dir).iterdir()):
        mo = _NAME.match(p.name)
        if mo is None:
            continue
        if tags is not None and mo.group(2) not in tags:
            continue
        if region is not None and int(mo.group(3)) != region:
            continue
        if rank is not None and int(mo.group(4)) != rank:
            continue
        out.append(Event(p))
    out.sort(key=lambda e: (e.seq, e.rank))
    return out


def _text(v):
    if isinstance(v, bool):
        return "T" if v else "F"
    if isinstance(v, (tuple, list)):
        return " ".join(str(int(x)) for x in v)
    if isinstance(v, float):
        return repr(v)
    return str(v)


def write_event(dump_dir, seq, tag, region, meta, arrays, rank=0):
    """Write one event as the dump module does: `arrays` maps a name to (FArray-like data,
    lower bounds) or an FArray; float arrays as 8-byte reals, integer arrays as 4-byte ints.
    Returns the .meta path."""
    d = Path(dump_dir)
    d.mkdir(parents=True, exist_ok=True)
    prefix = d / f"e{seq:05d}_{tag}_r{region}_p{rank:03d}"
    head = {"tag": tag, "seq": seq, "region": region, **({"rank": rank} if rank else {}), **meta}
    lines = [f"{k} = {_text(v)}" for k, v in head.items()]
    for name, a in arrays.items():
        data, lower = (a.data, a.lower) if isinstance(a, FArray) else a
        data = np.asarray(data)
        integer = np.issubdtype(data.dtype, np.integer)
        bounds = " ".join(f"{lo} {lo + n - 1}" for lo, n in zip(lower, data.shape))
        lines.append(f"{'iarray' if integer else 'array'} {name} = {bounds}")
        out = data.astype("<i4" if integer else "<f8")
        out.reshape(-1, order="F").tofile(f"{prefix}_{name}.bin")
    path = prefix.with_suffix(".meta")
    path.write_text("\n".join(lines) + "\n")
    return path


def to_state(event, st, names=None):
    """Copy the event's arrays into its region of a zoom state (pydyn_amd.zoom.state.ZoomState):
    each named array (default: every one the state holds) into the Field of that name, whose
    declared bounds it must have."""
    r = event.region
    fields = st[r]
    for name in names if names is not None else [n for n in event.arrays if n in fields]:
        f = fields[name]
        want = tuple((lo, lo + n - 1) for lo, n in zip(f.lower, f.t.shape))
        got = tuple(event.bounds[name])
        if len(got) == len(want) - 1 and want[-1] == (1, 1):  # a one-tracer array dumped 3-d
            got = got + ((1, 1),)
        if got != want:
            raise DumpError(f"{event!r} {name}: bounds {got}, region {r} declares {want}")
        f.t.copy_(event.device(name).reshape(f.t.shape))
