"""CPU checks of the Python mirror's host logic (no GPU call is made):

* every raw pointer handed to the library is checked first: a wrong dtype, a
  non-contiguous view, a short tensor or a host tensor raises before any device call
  (the kernels reinterpret the memory, so these would read or write out of bounds or
  compute from the wrong elements);
* limitBoxShrinking (domain/include/cstone/sfc/box.hpp:397-414), which the steps apply to
  open dimensions from their second call on (domain/assignment.hpp:104-106), equals the
  oracle's and the compiled reference's bit for bit, and its defining cases hold.
"""
import numpy as np
import pytest
import torch

import oracle
from sphexa_amd import hydro
from sphexa_amd._lib import BC_OPEN, BC_PERIODIC, FP64, MIXED

CPU = oracle.Cpu("oracle")
REF = oracle.Cpu("ref") if oracle.available("ref") else None


def test_checked_pointer_rejects_bad_tensors():
    good = torch.zeros(8, dtype=torch.float64)
    with pytest.raises(TypeError, match="dtype"):
        hydro._dev(good.float(), torch.float64, 8, "x")
    with pytest.raises(ValueError, match="contiguous"):
        hydro._dev(torch.zeros(8, 3, dtype=torch.float64)[:, 0], torch.float64, 8, "x")
    with pytest.raises(ValueError, match="at least"):
        hydro._dev(good, torch.float64, 9, "x")
    with pytest.raises(ValueError, match="GPU"):
        hydro._dev(good, torch.float64, 8, "x")
    with pytest.raises(ValueError, match="missing"):
        hydro._dev(None, torch.float64, 8, "x")
    assert hydro._dev(None, torch.float64, 8, "x", optional=True) is None
    with pytest.raises(TypeError):
        hydro._dev(np.zeros(8), torch.float64, 8, "x")


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_particles_struct_checks_every_field(precision):
    """A float64 h in a MIXED dataset (or float32 in FP64), or float32 positions, would make
    the kernels read or write past the field: the view refuses them."""
    n = 16
    d = hydro.ParticlesData(n, precision)
    d.x, d.y, d.z = (torch.zeros(n, dtype=torch.float64) for _ in range(3))
    d.h = torch.zeros(n, dtype=torch.float32 if precision == FP64 else torch.float64)
    with pytest.raises(TypeError, match="ParticlesData.h"):
        d.struct()
    d.h = torch.zeros(n, dtype=d.dtype)
    d.x = torch.zeros(n, dtype=torch.float32)
    with pytest.raises(TypeError, match="ParticlesData.x"):
        d.struct()
    d.x = torch.zeros(2 * n, dtype=torch.float64)[::2]
    with pytest.raises(ValueError, match="contiguous"):
        d.struct()
    d.x = torch.zeros(n - 1, dtype=torch.float64)
    with pytest.raises(ValueError, match="at least"):
        d.struct()


def _box(lo, hi, bc):
    return oracle.Box(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], bc[0], bc[1], bc[2], 0)


def _as_tuple(b):
    return (b.xmin, b.xmax, b.ymin, b.ymax, b.zmin, b.zmax, b.bx, b.by, b.bz)


def test_limit_box_shrinking_cases():
    prev = _box([0.0, -1.0, 2.0], [1.0, 1.0, 6.0], [BC_OPEN, BC_OPEN, BC_PERIODIC])
    # a fitted box inside the previous one by more than 5 % per side: each side stops at 5 %
    fit = _box([0.3, -0.5, 2.0], [0.6, 0.2, 6.0], [BC_OPEN] * 3)
    r = hydro.limit_box_shrinking(fit, prev)
    assert _as_tuple(r) == (0.0 + 0.05 * 1.0, 1.0 - 0.05 * 1.0, -1.0 + 0.05 * 2.0, 1.0 - 0.05 * 2.0, 2.0, 6.0,
                            BC_OPEN, BC_OPEN, BC_PERIODIC)
    # growth and shrinking within the limit pass through unchanged
    fit = _box([-0.5, -0.99, 2.0], [0.97, 1.5, 6.0], [BC_OPEN] * 3)
    r = hydro.limit_box_shrinking(fit, prev)
    assert _as_tuple(r)[:6] == (-0.5, 0.97, -0.99, 1.5, 2.0, 6.0)


def test_limit_box_shrinking_matches_oracle_and_reference():
    rng = np.random.default_rng(3)
    libs = [CPU] + ([REF] if REF is not None else [])
    for _ in range(500):
        lo = rng.normal(0, 1, 3)
        hi = lo + rng.uniform(0.1, 3, 3)
        prev = _box(lo, hi, rng.integers(0, 3, 3))
        flo = lo + rng.uniform(-0.3, 0.3, 3) * (hi - lo)
        fhi = hi + rng.uniform(-0.3, 0.3, 3) * (hi - lo)
        fit = _box(flo, np.maximum(fhi, flo + 1e-3), [BC_OPEN] * 3)
        mine = _as_tuple(hydro.limit_box_shrinking(fit, prev))
        for lib in libs:
            assert mine == _as_tuple(lib.limit_box_shrinking(fit, prev)), lib.pre


def test_deferred_min_dt_bookkeeping():
    """The Python side of the deferred Courant time step, with a stand-in for the library
    (no GPU): a call whose sphx_particles view comes back with NaN leaves its ParticlesData
    pending on the context; reading minDtCourant asks the context once; a call for another
    ParticlesData reads the pending one's value first (the library keeps only the last);
    a call that read the value back sets it at once; assigning clears the pending state;
    the view a call receives holds NaN while a value is pending."""
    library = {"value": None}  # the value sphx_min_dt_courant would return
    asked = []

    ctx = object.__new__(hydro.Context)  # no device: only the bookkeeping is exercised
    ctx.h = None
    ctx._min_dt_owner = None

    def min_dt_courant():
        asked.append(library["value"])
        return library["value"]

    ctx.min_dt_courant = min_dt_courant

    def deferred_call(d, v):
        def call():
            s = d.struct()
            library["value"] = v
            s.minDtCourant = float("nan")
            return s
        return call

    def immediate_call(d, v):
        def call():
            s = d.struct()
            s.minDtCourant = v
            return s
        return call

    d1, d2 = hydro.ParticlesData(n=4), hydro.ParticlesData(n=4)
    assert d1.minDtCourant == float("inf")
    ctx._min_dt_call(d1, deferred_call(d1, 0.25))
    assert d1._min_dt_ctx is ctx and asked == []
    assert np.isnan(d1.struct().minDtCourant)
    ctx._min_dt_call(d1, deferred_call(d1, 0.5))  # the same ParticlesData: the newer value wins
    assert asked == []
    ctx._min_dt_call(d2, deferred_call(d2, 0.125))  # d1 is read before d2's call replaces it
    assert asked == [0.5] and d1._min_dt_ctx is None and d1.minDtCourant == 0.5
    assert d2.minDtCourant == 0.125 and d2.minDtCourant == 0.125 and asked == [0.5, 0.125]
    ctx._min_dt_call(d1, immediate_call(d1, 0.75))
    assert d1.minDtCourant == 0.75 and d1._min_dt_ctx is None and d1.struct().minDtCourant == 0.75
    ctx._min_dt_call(d2, deferred_call(d2, 1.5))
    d2.minDtCourant = 3.0  # an assignment replaces the pending value
    assert d2.minDtCourant == 3.0 and asked == [0.5, 0.125]
    ctx._min_dt_call(d1, deferred_call(d1, 2.0))
    ctx._resolve_min_dt()  # what switching the option off and closing do first
    assert d1._min_dt_ctx is None and d1.minDtCourant == 2.0
