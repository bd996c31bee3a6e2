"""Bitwise pins of the oracle against the reference's own expressions.

Two things of the growth loop that tolerance-based checks cannot pin:

* the precedence of the two front lists: SCORE lets every completed Inside cell
  infect before any completed Border cell (SCORE_Kernel.cpp:6642-6647), so a
  deformed cell that both can reach goes to the Inside parent;
* the association order of the floating-point chain that sets rxFrac and the
  time step: G(T) and G b^2/2 (SCORE_GeneralPhysics.cpp:65-78,
  SCORE_Kernel.cpp:5340-5352), the mobility (:5355-5368, :5459-5462), the
  increment `v = m*p; v *= shape; rxFrac + (v*dt)*(1/dx)` (:4077-4092, with
  1/dx premultiplied at :5758), the carries `(rxFrac-1)*{0.0682,0.0542,0.0484}`
  (:4568-4599) and `dt = (maxfill*dx)/vmax` (:5523-5531).

The expected values are computed here in Python, whose float operations are
IEEE double with one rounding each and no fused multiply-add, as the
reference's GCC build without -march rounds them; math.exp is the host libm,
as in the reference.  Each alternative association that a restatement could
plausibly have written is computed too, and the test asserts that it gives
different bits at these parameters, so that the pin would catch it."""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from score_amd import abi
from score_amd.synth import DomainArrays, ParamsSpec

KB, EC = 1.3806488e-23, 1.602176565e-19  # SCORE_Defs_Physics.h:51,53

# the 26 attempts in call order (SCORE_Kernel.cpp:4611-4638): offset and infector code
ORDER = [((-1, -1, 0), 16), ((0, -1, 0), 15), ((1, -1, 0), 14), ((-1, 0, 0), 13), ((1, 0, 0), 12),
         ((-1, 1, 0), 11), ((0, 1, 0), 10), ((1, 1, 0), 9),
         ((-1, 1, -1), 19), ((0, 1, -1), 18), ((1, 1, -1), 17), ((-1, -1, -1), 25), ((0, -1, -1), 24),
         ((1, -1, -1), 23), ((-1, 0, -1), 22), ((0, 0, -1), 21), ((1, 0, -1), 20),
         ((-1, -1, 1), 8), ((0, -1, 1), 7), ((1, -1, 1), 6), ((-1, 0, 1), 5), ((0, 0, 1), 4), ((1, 0, 1), 3),
         ((-1, 1, 1), 2), ((0, 1, 1), 1), ((1, 1, 1), 0)]
SHAPE = {1: 1.0, 2: 0.7957, 3: 0.6494}        # FGEOFACE/EDGE/DIAG by |dx|+|dy|+|dz|
OVERFAC = {1: 0.0682, 2: 0.0542, 3: 0.0484}   # overFac1..3, SCORE_Kernel.cpp:4568-4570


def qaxis(axis, deg):
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    h = math.radians(deg) / 2
    return np.array([math.cos(h), *(math.sin(h) * a)])


def lin(x, y, z, nx, ny):
    return x + y * nx + z * nx * ny


def bits(x: float) -> int:
    return int(np.float64(x).view(np.uint64))


# ---------------------------------------------------------------- precedence

def test_inside_parent_beats_border_parent_on_a_shared_target():
    """Nucleus A sits at (2,5,5): strictly interior, so it is an Inside-list record
    (infect_JuvenileNucleation, SCORE_GrowthMachine.cpp:124-189), and more than
    one cell from every face, so its targets go to the Inside list
    (InfectIntoBorder, SCORE_Kernel.cpp:4606-4609).  Nucleus B sits at (0,5,5) on
    the x face: a Border-list record whose x-1 targets wrap to x = n-1
    (:4362-4377).  Both hold rxFrac = 1 and complete in step 0.  The nine cells
    (1, 4..6, 4..6) neighbour both.  With the Inside list first (:6642-6647) A
    claims them; B then finds them CURRENTLY_INFECTED and rejects them (:4397).
    So, slot by slot:
      Inside slot 0        A, retired; slots 1..26 A's 26 children in call order
                           (appended: no slot was INACTIVE in this step's sweep);
      Border slot 0        B, retired; slots 1..17 B's children outside x = 1, in
                           call order (every one is on the x face, so Border).
    Had the Border list gone first, B would own the nine shared cells (they are
    strictly interior, so they would go to the Inside list, ahead of A's)."""
    n = 12
    a_xyz, b_xyz = (2, 5, 5), (0, 5, 5)
    N = n ** 3
    q = [qaxis([0, 0, 1], 0.0), qaxis([0, 0, 1], 45.0), qaxis([1, 1, 1], 60.0)]
    dom = DomainArrays(n, n, n, 0.5e-6, np.zeros(N, np.uint32), np.array(q), np.array([0], np.uint32),
                       np.array([1e15]), np.array([1, 2], np.uint32),
                       np.array([lin(*a_xyz, n, n), lin(*b_xyz, n, n)], np.uint32))
    o = Oracle(dom)
    fi, fb = o.front(abi.LIST_INSIDE), o.front(abi.LIST_BORDER)
    assert list(fi["rxg"]) == [0] and list(fb["rxg"]) == [1]
    o.step(1)
    fi, fb = o.front(abi.LIST_INSIDE), o.front(abi.LIST_BORDER)
    ax, ay, az = a_xyz
    bx, by, bz = b_xyz
    # Inside: A retired in slot 0, its 26 children in call order, nothing else
    assert len(fi["activity"]) == 27 and fi["activity"][0] == 0 and fi["activity"][1:].all()
    for k, ((dx, dy, dz), code) in enumerate(ORDER, start=1):
        assert fi["cell"][k] == lin(ax + dx, ay + dy, az + dz, n, n)
        assert fi["infector"][k] == code and fi["rxg"][k] == 0
    # Border: B retired in slot 0, then its children that A did not take, in call order
    want = []
    for (dx, dy, dz), code in ORDER:
        x, y, z = (bx + dx) % n, (by + dy) % n, bz + dz
        if x == 1:
            continue                      # a shared cell, claimed by A
        want.append((lin(x, y, z, n, n), code))
    assert len(want) == 17
    assert len(fb["activity"]) == 18 and fb["activity"][0] == 0 and fb["activity"][1:].all()
    assert [(int(c), int(i)) for c, i in zip(fb["cell"][1:], fb["infector"][1:])] == want
    assert set(fb["rxg"][1:].tolist()) == {1}
    # the shared cells: infected, held by A's records, with A's infector codes
    g = o.grainmap(abi.MAP_CELLGRID)
    shared = [lin(1, y, z, n, n) for z in (4, 5, 6) for y in (4, 5, 6)]
    assert all(g[c] == abi.CURRENTLY_INFECTED for c in shared)
    owner = {int(c): (int(r), int(i)) for c, r, i in zip(fi["cell"][1:], fi["rxg"][1:], fi["infector"][1:])}
    codes = {lin(ax + dx, ay + dy, az + dz, n, n): code for (dx, dy, dz), code in ORDER}
    for c in shared:
        assert owner[c] == (0, codes[c])
    tr = o.trajectory()[0]
    assert tr["nfull_inside"] == 1 and tr["nfull_border"] == 1 and tr["ninfected"] == 26 + 17
    o.close()


# ------------------------------------------------------- association order

# parameters at which every alternative below changes at least one compared bit
T_K, M0, H_EV, RHO, DX, MAXFILL = 1012.5, 0.1, 2.8, 5.1e14, 1.3e-6, 0.1


def reference_chain(alt: str | None = None) -> dict:
    """The first two steps of one nucleus in one deformed grain, in the
    reference's association order (alt=None) or with one alternative:
      'gbb'    Gb^2/2 as ((0.5*G)*b)*b instead of (0.5*G)*SQR(b)
      'horner' G(T) in Horner form instead of the reference's sum of powers
      'mps'    v = m*(p*shape) instead of v = m*p; v *= shape
      'vdt'    rxFrac + v*(dt*inv_dx) instead of rxFrac + (v*dt)*inv_dx
      'div'    (v*dt)/dx instead of multiplying by the premultiplied 1/dx
      'dt'     maxfill*(dx/vmax) instead of (maxfill*dx)/vmax"""
    T = T_K
    if alt == "horner":
        G = 6.92e10 * (((-1.38e-9 * T + 1.21e-6) * T - 4.58e-4) * T + 1.00)
    else:
        G = 6.92e10 * (-1.38e-9 * (T * T * T) + 1.21e-6 * (T * T) - 4.58e-4 * T + 1.00)
    b = 2.48e-10
    gb = ((0.5 * G) * b) * b if alt == "gbb" else (0.5 * G) * (b * b)
    mRH = M0 * math.exp(-1.0 * (H_EV * EC) / (KB * T))
    inv_dx = 1.0 / DX
    P = 1.0                                   # 45 deg <001>: far above the LAGB/HAGB transition
    p = (gb * (RHO - 1e11)) - 0.0

    def inc(shape, dt):
        m = P * mRH
        if alt == "mps":
            v = m * (p * shape)
        else:
            v = m * p
            v *= shape
        if alt == "vdt":
            return v * (dt * inv_dx)
        if alt == "div":
            return (v * dt) / DX
        return (v * dt) * inv_dx

    dt0 = 1e-6                                # temperature rule at t = 0 (SCORE_Kernel.cpp:5563-5571)
    f0 = 1.0 + inc(1.0, dt0)                  # the nucleus, infector 26, shape 1.0
    vmax = mRH * ((gb * RHO) - 0.0)
    dt1 = maxfill_dt(vmax, alt)
    out = {"dt0": dt0, "dt1": dt1, "f0": f0, "carry": {}, "f1": {}}
    for cls in (1, 2, 3):
        carry = (f0 - 1.0) * OVERFAC[cls]
        out["carry"][cls] = carry
        out["f1"][cls] = carry + inc(SHAPE[cls], dt1)
    return out


def maxfill_dt(vmax, alt):
    return MAXFILL * (DX / vmax) if alt == "dt" else (MAXFILL * DX) / vmax


def flat(ch: dict) -> list:
    return [ch["dt1"], ch["f0"]] + [ch["carry"][c] for c in (1, 2, 3)] + [ch["f1"][c] for c in (1, 2, 3)]


@pytest.mark.parametrize("alt", ["gbb", "horner", "mps", "vdt", "div", "dt"])
def test_each_alternative_association_changes_bits(alt):
    """the pin below can fail: every alternative order gives different bits here"""
    ref, other = flat(reference_chain()), flat(reference_chain(alt))
    assert any(bits(x) != bits(y) for x, y in zip(ref, other)), alt


def test_rxfrac_carries_and_dt_bitwise_first_two_steps():
    """rxFrac of the nucleus after step 0, the carries of its 26 children, dt of
    step 1 and the children's rxFrac after step 1, all bit for bit."""
    n = 12
    N = n ** 3
    c = (6, 6, 6)
    q = [qaxis([0, 0, 1], 0.0), qaxis([0, 0, 1], 45.0)]
    dom = DomainArrays(n, n, n, DX, np.zeros(N, np.uint32), np.array(q), np.array([0], np.uint32),
                       np.array([RHO]), np.array([1], np.uint32), np.array([lin(*c, n, n)], np.uint32))
    prm = ParamsSpec(rh_hagb_m0=M0, rh_hagb_hact_ev=H_EV, sched_temp=(T_K, T_K), maxfill=MAXFILL)
    ref = reference_chain()
    o = Oracle(dom, prm)
    assert o.front(abi.LIST_INSIDE)["P"][0] == 1.0
    assert o.dt() == ref["dt0"]
    o.step(1)
    f = o.front(abi.LIST_INSIDE)
    for k, ((dx, dy, dz), _) in enumerate(ORDER, start=1):
        cls = abs(dx) + abs(dy) + abs(dz)
        assert bits(f["rxfrac"][k]) == bits(ref["carry"][cls]), (k, cls)
    assert bits(o.dt()) == bits(ref["dt1"])
    o.step(1)
    f = o.front(abi.LIST_INSIDE)
    for k, ((dx, dy, dz), _) in enumerate(ORDER, start=1):
        cls = abs(dx) + abs(dy) + abs(dz)
        assert bits(f["rxfrac"][k]) == bits(ref["f1"][cls]), (k, cls)
    tr = o.trajectory()
    assert bits(tr["dt"][0]) == bits(ref["dt0"]) and bits(tr["dt"][1]) == bits(ref["dt1"])
    assert tr["t"][1] == ref["dt0"]
    # the nucleus' own value after step 0 follows from the carries' common factor
    assert bits((ref["f0"] - 1.0) * OVERFAC[1]) == bits(ref["carry"][1])
    o.close()
