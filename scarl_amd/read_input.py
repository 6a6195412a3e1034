"""Text decks of ScaRL: ``config_ScaRL.txt``, ``extra_config_ScaRL.txt`` and the sample deck.

Mirrors the reference's ``read_input`` module (SRC_F/read_input.f90):

* ``getLine`` (:565-585) returns the next line that, once left-adjusted, does not start with the
  comment tag. It gives up after 200 lines, and returns "eof_gl" at end of file.
* Each value line is read with a list-directed READ(buffer, *) of the items the deck expects
  (``list_read``). The read stops after those items, so a trailing "#..." note is ignored.
  Items are separated by blanks or a comma, and a slash ends the record early. An item may
  carry a repeat count (r*c).
* ``read_config_ScaRL`` (:26-105): generation method, four 0/1 switches and the deck's path.
  A missing file gives the reference's defaults.
* ``read_extra_config_ScaRL`` (:110-147): an optional topology (subdomains per axis).
* ``read_input_ScaRL`` (:152-348): output folder, number of samples and twelve lines per
  sample. The SEM path taken when ``random.spec`` exists (:198-216) is not supported.

The parser keeps the reference's checks, so a malformed deck fails where the reference's READ
would.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

FFT, FFT_MPI = 1, 2            # SRC_F/constants_ScaRL.f90:12-13
cm_GAUSSIAN = 1                # :15
fom_GAUSSIAN, fom_LOGNORMAL = 1, 2  # :20-21
LINE_LEN = 1024                # character(len=1024) line buffer


class DeckError(ValueError):
    """A deck that the reference's READ statements would reject."""


class LineSource:
    """Sequential reads of a text file's records (Fortran unit semantics)."""

    def __init__(self, path):
        with open(path, "r", encoding="latin-1", newline=None) as fh:
            self.lines = fh.read().split("\n")
        if self.lines and self.lines[-1] == "":
            self.lines.pop()
        self.pos = 0

    def read(self) -> Optional[str]:
        if self.pos >= len(self.lines):
            return None
        s = self.lines[self.pos]
        self.pos += 1
        return s


def getLine(src: LineSource, comment_Tag: str = "#") -> str:
    """SRC_F/read_input.f90:565-585."""
    nextLine = ""
    for _ in range(200):
        raw = src.read()
        if raw is None:
            return "eof_gl"
        nextLine = raw[:LINE_LEN].rstrip("\r").lstrip(" ")
        if nextLine[:1] != comment_Tag:
            break
    return nextLine


# ---- list-directed READ ------------------------------------------------------------------------

_REAL = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eEdDqQ][+-]?\d+|[+-]\d+)?$")
_INT = re.compile(r"^[+-]?\d+$")
_SPECIAL = re.compile(r"^[+-]?(inf|infinity|nan)$", re.IGNORECASE)


def _values(buffer: str):
    """The value items of one record: strings keep their quotes; a slash ends the list."""
    out, i, n = [], 0, len(buffer)
    expect_value = True
    while i < n:
        c = buffer[i]
        if c in " \t":
            i += 1
            continue
        if c == ",":
            if expect_value:
                out.append(None)  # null value
            expect_value = True
            i += 1
            continue
        if c == "/":
            break
        if c in "'\"":
            q, j, s = c, i + 1, []
            while True:
                if j >= n:
                    raise DeckError(f"unterminated character constant in {buffer!r}")
                if buffer[j] == q:
                    if j + 1 < n and buffer[j + 1] == q:
                        s.append(q)
                        j += 2
                        continue
                    break
                s.append(buffer[j])
                j += 1
            out.append(("str", "".join(s)))
            i = j + 1
        else:
            j = i
            while j < n and buffer[j] not in " \t,/":
                j += 1
            tok = buffer[i:j]
            m = re.match(r"^(\d+)\*(.*)$", tok)
            if m:
                r, rest = int(m.group(1)), m.group(2)
                out.extend([("tok", rest) if rest else None] * r)
            else:
                out.append(("tok", tok))
            i = j
        expect_value = False
        # a value is followed by blanks, then an optional comma
        while i < n and buffer[i] in " \t":
            i += 1
        if i < n and buffer[i] == ",":
            i += 1
            expect_value = True
    return out


def _convert(item, kind):
    if item is None:
        raise DeckError("null value where the deck needs one")
    tag, s = item
    if kind == "a":
        return s
    if tag == "str":
        raise DeckError(f"character constant {s!r} where a number is expected")
    if kind == "i":
        if not _INT.match(s):
            raise DeckError(f"bad integer {s!r}")
        return int(s)
    if _SPECIAL.match(s):
        return float(s.lower().replace("infinity", "inf"))
    if not _REAL.match(s):
        raise DeckError(f"bad real number {s!r}")
    t = s.replace("d", "e").replace("D", "e").replace("q", "e").replace("Q", "e")
    m = re.match(r"^([+-]?(?:\d+\.?\d*|\.\d+))([+-]\d+)$", t)
    if m:  # exponent without a letter: 1.0+3
        t = f"{m.group(1)}e{m.group(2)}"
    return float(t)


def list_read(buffer: str, kinds: str):
    """READ(buffer, *) of len(kinds) items ('i' integer, 'd' double, 'a' character)."""
    if buffer == "eof_gl":
        raise DeckError("end of file before every value was read")
    vals = _values(buffer)
    if len(vals) < len(kinds):
        raise DeckError(f"end of record: {len(kinds)} values expected in {buffer.strip()!r}")
    return [_convert(v, k) for v, k in zip(vals, kinds)]


# ---- the three decks -----------------------------------------------------------------------------

@dataclass
class Config:
    gen_meth: int = FFT
    c_info_file: bool = False
    d_sample: bool = False
    one_file: bool = True
    one_hdf_ds: bool = True
    input_file: str = "./input_ScaRL.txt"
    found: bool = False


def read_config_ScaRL(config_file) -> Config:
    """SRC_F/read_input.f90:26-105."""
    if not os.path.isfile(config_file):
        return Config()
    src = LineSource(config_file)
    c = Config(found=True)
    (c.gen_meth,) = list_read(getLine(src), "i")
    flags = []
    for _ in range(4):
        (v,) = list_read(getLine(src), "i")
        flags.append(v == 1)
    c.c_info_file, c.d_sample, c.one_file, c.one_hdf_ds = flags
    (c.input_file,) = list_read(getLine(src), "a")
    return c


def read_extra_config_ScaRL(x_config_file):
    """SRC_F/read_input.f90:110-147: (extra_config, topo_shape_in)."""
    if not os.path.isfile(x_config_file):
        return False, [-1, -1, -1]
    src = LineSource(x_config_file)
    return True, list_read(getLine(src), "iii")


@dataclass
class Sample:
    output_name: str
    xMinGlob: List[float]
    xMaxGlob: List[float]
    corrL: List[float]
    overlap: List[float]
    pointsPerCorrL: List[int]
    corrMod: int
    margiFirst: int
    avg: float
    std_dev: float
    seedBase: int


@dataclass
class InputDeck:
    output_folder: str
    nSamples: int
    samples: List[Sample] = field(default_factory=list)


def read_input_ScaRL(file_path_in, cur_folder: Optional[str] = None) -> InputDeck:
    """SRC_F/read_input.f90:152-348. The deck is cur_folder // "/" // file_path_in."""
    cur_folder = os.getcwd() if cur_folder is None else cur_folder
    if os.path.isfile(os.path.join(cur_folder, "random.spec")):
        raise NotImplementedError("'random.spec' found: the SEM input path is not supported")
    file_path = f"{cur_folder}/{file_path_in}"
    if not os.path.isfile(file_path):
        raise FileNotFoundError(f"'input_Scarl.txt' NOT FOUND ({file_path})")
    src = LineSource(file_path)
    (output_folder,) = list_read(getLine(src), "a")
    (nSamples,) = list_read(getLine(src), "i")
    deck = InputDeck(output_folder, nSamples)
    for _ in range(nSamples):
        (name,) = list_read(getLine(src), "a")
        xMin = list_read(getLine(src), "ddd")
        xMax = list_read(getLine(src), "ddd")
        corrL = list_read(getLine(src), "ddd")
        overlap = list_read(getLine(src), "ddd")
        ppc = list_read(getLine(src), "iii")
        (corrMod,) = list_read(getLine(src), "i")
        (margi,) = list_read(getLine(src), "i")
        (avg,) = list_read(getLine(src), "d")
        (sd,) = list_read(getLine(src), "d")
        (seed,) = list_read(getLine(src), "i")
        deck.samples.append(Sample(name, xMin, xMax, corrL, overlap, ppc, corrMod, margi, avg, sd, seed))
    return deck


def verify_inputs(s: Sample, gen_meth: int):
    """verify_inputs (SRC_F/main_ScaRL.f90:286-333): raises DeckError where the reference stops.
    Returns the warnings it prints."""
    warn = []
    if any(a >= b for a, b in zip(s.xMinGlob, s.xMaxGlob)):
        raise DeckError("ERROR: xMinGlob >= xMaxGlob")
    if any(c < 0.0 for c in s.corrL):
        raise DeckError("ERROR: corrL < 0d0")
    if any(o < 0.0 for o in s.overlap):
        warn.append("WARNING: overlap < 0d0, the mesh will be discontinuous")
    if s.corrMod != cm_GAUSSIAN:
        raise DeckError("ERROR: Correlation model not implemented")
    if s.margiFirst not in (fom_GAUSSIAN, fom_LOGNORMAL):
        raise DeckError("ERROR: First-order marginal not implemented")
    if gen_meth not in (FFT, FFT_MPI):
        raise DeckError("ERROR: Generation Method  not implemented")
    if s.avg <= 0.0 and s.margiFirst == fom_LOGNORMAL:
        raise DeckError("ERROR: When using lognormal first-order marginal density the average should be greater than 0")
    if s.std_dev < 0.0:
        raise DeckError("ERROR: Standard deviation should not be smaller than 0d0")
    if any(p < 2 for p in s.pointsPerCorrL):
        raise DeckError("ERROR: PointsPerCorrL < 2")
    return warn
