"""Helpers for the GPU tests: build spade_amd objects from the shared cases."""
import numpy as np

import golden_io as gi
import spade_amd as sa
from spade_amd import _abi

T = gi.TWO_PI_REF


def grid(spec, periodic=(True, True, True), maps=(None, None, None)):
    return sa.CartesianGrid(spec["nc"], spec["nb"], ((0.0, T),) * 3, num_exch=spec["ng"],
                            periodic=periodic, maps=maps)


def flux_desc(terms, gamma, R):
    d = _abi.FluxDesc()
    d.nterms = len(terms)
    for i, t in enumerate(terms):
        for k, v in t.items():
            setattr(d.terms[i], k, v)
    d.gamma, d.R = gamma, R
    return d


class RawFlux:
    """A flux object accepted by spade_amd.flux_div built from raw term dicts."""

    def __init__(self, terms, gamma, R):
        gas = sa.IdealGas(gamma, R)
        self._terms = []
        for t in terms:
            fields = dict(t)
            self._terms.append(sa._Term(gas, **fields))

    def terms(self):
        return self._terms


def rel_err(a, b):
    return gi.rel_maxnorm(a, b)
