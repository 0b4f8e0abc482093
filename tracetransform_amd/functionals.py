"""Functional enums, wrappers and their string grammar.

Mirrors c++/src/sinogram.hpp:22-48 and c++/src/circus.hpp:27-57 (enum
ordinals, wrapper fields) and the ``operator>>`` parsers of
c++/src/sinogram.cpp:34-65 and c++/src/circus.cpp:33-66: "0".."7" name the
T-functionals Radon, T1..T7; "1".."3" name P1..P3 and "H<n>" the Hermite
P-functional of order n.  Anything else raises ValueError (the reference
throws boost::program_options::validation_error).
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, field
from typing import Optional


class TFunctional(enum.IntEnum):
    Radon = 0
    T1 = 1
    T2 = 2
    T3 = 3
    T4 = 4
    T5 = 5
    T6 = 6
    T7 = 7


class PFunctional(enum.IntEnum):
    Hermite = 0
    P1 = 1
    P2 = 2
    P3 = 3


@dataclass
class TFunctionalArguments:
    pass


@dataclass
class PFunctionalArguments:
    order: Optional[int] = None
    center: Optional[int] = None


@dataclass
class TFunctionalWrapper:
    name: str
    functional: TFunctional
    arguments: TFunctionalArguments = field(default_factory=TFunctionalArguments)

    @classmethod
    def parse(cls, text: str) -> "TFunctionalWrapper":
        names = {"0": "Radon", "1": "T1", "2": "T2", "3": "T3", "4": "T4",
                 "5": "T5", "6": "T6", "7": "T7"}
        if text not in names:
            raise ValueError(f"invalid T-functional {text!r}")
        return cls(names[text], TFunctional(int(text)))


@dataclass
class PFunctionalWrapper:
    name: str
    functional: PFunctional
    arguments: PFunctionalArguments = field(default_factory=PFunctionalArguments)

    @classmethod
    def parse(cls, text: str) -> "PFunctionalWrapper":
        if text in ("1", "2", "3"):
            return cls("P" + text, PFunctional(int(text)))
        if text[:1] == "H":
            if len(text) < 2:
                raise ValueError("missing order parameter for Hermite P-functional")
            digits = text[1:]
            if not digits.isdigit():
                raise ValueError(f"unparseable order parameter {digits!r} for "
                                 "Hermite P-functional")
            return cls(text, PFunctional.Hermite,
                       PFunctionalArguments(order=int(digits)))
        raise ValueError(f"invalid P-functional {text!r}")


def tcodes_of(tfunctionals) -> list:
    """T codes of a list of wrappers, enums, ints or strings."""
    out = []
    for t in tfunctionals:
        if isinstance(t, TFunctionalWrapper):
            out.append(int(t.functional))
        elif isinstance(t, str):
            out.append(int(TFunctionalWrapper.parse(t).functional))
        else:
            out.append(int(t))
    return out


def pfuncs_of(pfunctionals) -> list:
    """(code, order, center) triples; center -1 when unset."""
    out = []
    for p in pfunctionals:
        if isinstance(p, str):
            p = PFunctionalWrapper.parse(p)
        if isinstance(p, PFunctionalWrapper):
            a = p.arguments
            out.append((int(p.functional),
                        -1 if a.order is None else int(a.order),
                        -1 if a.center is None else int(a.center)))
        elif isinstance(p, tuple):
            code = int(p[0])
            order = int(p[1]) if len(p) > 1 else -1
            center = int(p[2]) if len(p) > 2 else -1
            out.append((code, order, center))
        else:
            out.append((int(p), -1, -1))
    return out
