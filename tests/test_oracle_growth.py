"""Pins of the oracle's growth loop (SCORE_Kernel.cpp:3587-4774, :5311-5723, :6497-6792,
SCORE_GrowthMachine.cpp:50-189): hand-derived first steps (direction order, infector
codes, overshoot carries, slot policy, list choice, boundary rules), the time-step
schedule, invariants of complete runs, and the statistics of the reference's own
shipped 200^3 run (example/jmak)."""
import json
import math
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, OracleError
from score_amd import abi, synth
from score_amd.synth import DomainArrays, ParamsSpec

HERE = os.path.dirname(os.path.abspath(__file__))
KB, EC = 1.3806488e-23, 1.602176565e-19

# call order of the 26 attempts and their infector codes, as listed in
# SCORE_Kernel.cpp:4611-4638 (offset dx,dy,dz -> code)
ORDER = [((-1, -1, 0), 16), ((0, -1, 0), 15), ((1, -1, 0), 14), ((-1, 0, 0), 13), ((1, 0, 0), 12),
         ((-1, 1, 0), 11), ((0, 1, 0), 10), ((1, 1, 0), 9),
         ((-1, 1, -1), 19), ((0, 1, -1), 18), ((1, 1, -1), 17), ((-1, -1, -1), 25), ((0, -1, -1), 24),
         ((1, -1, -1), 23), ((-1, 0, -1), 22), ((0, 0, -1), 21), ((1, 0, -1), 20),
         ((-1, -1, 1), 8), ((0, -1, 1), 7), ((1, -1, 1), 6), ((-1, 0, 1), 5), ((0, 0, 1), 4), ((1, 0, 1), 3),
         ((-1, 1, 1), 2), ((0, 1, 1), 1), ((1, 1, 1), 0)]
CARRY = {1: 0.0682, 2: 0.0542, 3: 0.0484}


def qaxis(axis, deg):
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    h = math.radians(deg) / 2
    return np.array([math.cos(h), *(math.sin(h) * a)])


def one_grain(n, sites, rho=1e15, nuc_deg=45.0, shape=None):
    nx, ny, nz = shape or (n, n, n)
    N = nx * ny * nz
    q = [qaxis([0, 0, 1], 0.0), qaxis([0, 0, 1], nuc_deg)]
    return DomainArrays(nx, ny, nz, 0.5e-6, np.zeros(N, np.uint32), np.array(q), np.array([0], np.uint32),
                        np.array([rho]), np.ones(len(sites), np.uint32), np.array(sites, np.uint32))


def lin(x, y, z, nx, ny):
    return x + y * nx + z * nx * ny


def test_list_capacities_follow_reference():
    dom = one_grain(20, [lin(10, 10, 10, 20, 20)])
    caps = Oracle(dom).capacities()
    N = 8000
    assert caps[0] == int(0.2 * N)                         # Inside slots
    assert caps[1] == int(0.2 * N * 0.1 * 3.0)             # Inside recycling cap
    assert caps[2] == 2 * (400 + 400 + 400)                # Border slots
    assert caps[3] == int(2 * 1200 * 0.1 * 3.0)            # Border recycling cap


def test_first_infection_order_codes_carries_and_slots():
    n = 20
    c = (10, 10, 10)
    dom = one_grain(n, [lin(*c, n, n)])
    o = Oracle(dom)
    fi = o.front(abi.LIST_INSIDE)
    assert len(fi["activity"]) == 1 and fi["rxfrac"][0] == 1.0 and fi["infector"][0] == 26
    assert o.dt() == 1e-6                                  # temperature rule at t = 0
    o.step(1)
    fi = o.front(abi.LIST_INSIDE)
    fb = o.front(abi.LIST_BORDER)
    assert len(fb["activity"]) == 0
    # the nucleus retired, its 26 neighbours appended in call order into slots 1..26
    assert fi["activity"][0] == 0 and len(fi["activity"]) == 27
    g = o.grainmap()
    assert g[lin(*c, n, n)] == 1 + 0                       # ndefg + rxg id
    tr = o.trajectory()[0]
    frac_nucleus = 1.0 + tr["sv"] * 0 + 0.0
    overshoot = None
    for k, ((dx, dy, dz), code) in enumerate(ORDER, start=1):
        assert fi["cell"][k] == lin(c[0] + dx, c[1] + dy, c[2] + dz, n, n)
        assert fi["infector"][k] == code
        cls = abs(dx) + abs(dy) + abs(dz)
        if overshoot is None:
            overshoot = fi["rxfrac"][k] / CARRY[cls]
        assert fi["rxfrac"][k] == pytest.approx(overshoot * CARRY[cls], rel=1e-12)
        assert fi["P"][k] == 1.0 and fi["rxg"][k] == 0 and fi["defg"][k] == 0
        assert g[fi["cell"][k]] == abi.CURRENTLY_INFECTED
    # overshoot of the nucleus in step 0: v * 1e-6 s / dx with v = m * G b^2/2 (rho - 1e11)
    T = 1000.0
    G = 6.92e10 * (-1.38e-9 * T ** 3 + 1.21e-6 * T ** 2 - 4.58e-4 * T + 1.0)
    v = 0.1 * math.exp(-2.8 * EC / (KB * T)) * 0.5 * G * 2.48e-10 ** 2 * (1e15 - 1e11)
    assert overshoot == pytest.approx(v * 1e-6 / 0.5e-6, rel=1e-9)
    assert tr["sv"] == 1 and tr["nfull_inside"] == 1 and tr["ninfected"] == 26


def test_second_dt_and_increment_closed_form():
    """after t = 0 the step is maxfill * dx / vmax, so a face cell of the fastest grain
    advances by maxfill * (rho - 1e11) / rho per step (SCORE_Kernel.cpp:5523-5531)"""
    n = 20
    dom = one_grain(n, [lin(10, 10, 10, n, n)])
    o = Oracle(dom)
    o.step(1)
    before = o.front(abi.LIST_INSIDE)["rxfrac"].copy()
    T = 1000.0
    G = 6.92e10 * (-1.38e-9 * T ** 3 + 1.21e-6 * T ** 2 - 4.58e-4 * T + 1.0)
    vmax = 0.1 * math.exp(-2.8 * EC / (KB * T)) * 0.5 * G * 2.48e-10 ** 2 * 1e15
    assert o.dt() == pytest.approx(0.1 * 0.5e-6 / vmax, rel=1e-12)
    o.step(1)
    after = o.front(abi.LIST_INSIDE)["rxfrac"]
    inc = after[1:27] - before[1:27]
    shape = {1: 1.0, 2: 0.7957, 3: 0.6494}
    for k, ((dx, dy, dz), code) in enumerate(ORDER):
        cls = abs(dx) + abs(dy) + abs(dz)
        assert inc[k] == pytest.approx(0.1 * shape[cls] * (1e15 - 1e11) / 1e15, rel=1e-9)


def test_zener_drag_constant_closed_form():
    """constant drag p_z = alpha*gamma*f/r (get_zener, SCORE_Kernel.cpp:5448-5456): the
    first step runs without it (SCORE's prologue does not update the microchemistry),
    then the time step is maxfill*dx/(m*(Gb^2/2*rhomax - p_z)) and a face cell advances
    by maxfill*(Gb^2/2*(rho - 1e11) - p_z)/(Gb^2/2*rho - p_z) per step"""
    n = 20
    dom = one_grain(n, [lin(10, 10, 10, n, n)])
    fr0 = 5.0e5
    prm = ParamsSpec(zener_mode=1, zener_alpha=1.5, zener_gamma=0.324, disp_time=(0.0,), disp_fr=(fr0,))
    o, o0 = Oracle(dom, prm), Oracle(dom)
    o.step(1)
    o0.step(1)
    assert np.array_equal(o.front(abi.LIST_INSIDE)["rxfrac"], o0.front(abi.LIST_INSIDE)["rxfrac"])
    before = o.front(abi.LIST_INSIDE)["rxfrac"].copy()
    T = 1000.0
    G = 6.92e10 * (-1.38e-9 * T ** 3 + 1.21e-6 * T ** 2 - 4.58e-4 * T + 1.0)
    gb = 0.5 * G * 2.48e-10 ** 2
    pz = 1.5 * 0.324 * fr0
    m = 0.1 * math.exp(-2.8 * EC / (KB * T))
    assert o.dt() == pytest.approx(0.1 * 0.5e-6 / (m * (gb * 1e15 - pz)), rel=1e-12)
    o.step(1)
    inc = o.front(abi.LIST_INSIDE)["rxfrac"][1:27] - before[1:27]
    shape = {1: 1.0, 2: 0.7957, 3: 0.6494}
    for k, ((dx, dy, dz), code) in enumerate(ORDER):
        cls = abs(dx) + abs(dy) + abs(dz)
        assert inc[k] == pytest.approx(0.1 * shape[cls] * (gb * (1e15 - 1e11) - pz) / (gb * 1e15 - pz), rel=1e-9)
    o.close()
    o0.close()


def test_zener_drag_time_dependent_dt_rule():
    """f/r linear in time (update_microchemistry, SCORE_Kernel.cpp:5370-5405): the time
    step is SMALL_ZENERFORCE / (alpha*gamma*|d(f/r)/dt|) while inside the table
    (get_dtmax_instantslope_zener, :5639-5666); past its end f/r keeps its last
    interpolated value and the drag rule is INFINITE"""
    n = 20
    dom = one_grain(n, [lin(10, 10, 10, n, n)])
    a, g, slope = 1.5, 0.324, 1.0e6 / 100.0
    prm = ParamsSpec(zener_mode=2, zener_alpha=a, zener_gamma=g, disp_time=(0.0, 0.05), disp_fr=(0.0, 0.05 * slope))
    o = Oracle(dom, prm)
    o.step(4)
    tr = o.trajectory()
    want = 100.0 / (a * g * (abs(0.05 * slope - 0.0) / (0.05 - 0.0)))
    # step 0 ran with the initial 1e-6 s; then the drag rule sets the step while t < 0.05
    assert tr["dt"][0] == 1e-6 and all(d == want for d, t in zip(tr["dt"][1:], tr["t"][1:]) if t < 0.05)
    # past the table: the cells rule with the last f/r reached inside it
    while o.trajectory()["t"][-1] < 0.05:
        o.step(1)
    tr = o.trajectory()
    t_in = max(t for t in tr["t"] if t <= 0.05)
    fr_last = 0.0 + ((0.05 * slope - 0.0) / (0.05 - 0.0)) * (t_in - 0.0)
    T = 1000.0
    G = 6.92e10 * (-1.38e-9 * T ** 3 + 1.21e-6 * T ** 2 - 4.58e-4 * T + 1.0)
    gb = 0.5 * G * 2.48e-10 ** 2
    m = 0.1 * math.exp(-2.8 * EC / (KB * T))
    o.step(1)
    assert o.dt() == pytest.approx(0.1 * 0.5e-6 / (m * (gb * 1e15 - a * g * fr_last)), rel=1e-9)
    o.close()


def _atomistic(T=1000.0):
    G = 6.92e10 * (-1.38e-9 * T ** 3 + 1.21e-6 * T ** 2 - 4.58e-4 * T + 1.0)
    return G, 2.48e-10, 0.5 * G * 2.48e-10 ** 2


def _mp_factor(t):
    """Michalak-Paxton recovery of SCORE_RecoveryModels.cpp:89-125: two relaxations
    with time constants 0.3605 min and 5.2824 min, weights 1/4 and 3/4"""
    return 0.25 / (1.0 + t / (0.3605 * 60.0)) + 0.75 / math.sqrt(1.0 + t / (5.2824 * 60.0))


def test_recovery_michalak_paxton_closed_form():
    """rho = rho0 * factor(t)^2 after each step (update_deformedsubstructure,
    SCORE_Kernel.cpp:5408-5437), applied after the master section: rhomax of step j is
    rho0 * factor(t_j)^2, the time step lags by one step (the cells rule with that
    rhomax, no recovery rule for this model), and a face cell of step j advances by
    maxfill * (rho_j - 1e11) / rhomax_(j-1)"""
    n = 20
    dom = one_grain(n, [lin(10, 10, 10, n, n)])
    o = Oracle(dom, ParamsSpec(recovery_mode=3, rh_hagb_m0=40.0))
    o.step(2)
    before = o.front(abi.LIST_INSIDE)["rxfrac"].copy()
    o.step(1)
    after = o.front(abi.LIST_INSIDE)["rxfrac"]
    o.step(5)
    tr = o.trajectory()
    assert tr["rhomax"][0] == 1e15
    for j in range(1, len(tr)):
        assert tr["rhomax"][j] == pytest.approx(1e15 * _mp_factor(tr["t"][j]) ** 2, rel=1e-13)
    G, b, gb = _atomistic()
    m = 40.0 * math.exp(-2.8 * EC / (KB * 1000.0))
    for j in range(1, len(tr)):
        assert tr["dt"][j] == pytest.approx(0.1 * 0.5e-6 / (m * gb * tr["rhomax"][j - 1]), rel=1e-12)
    # the recovery is visible: rho fell by a fifth within the first steps
    assert tr["rhomax"][3] < 0.99 * tr["rhomax"][1]
    inc = after[1:27] - before[1:27]
    shape = {1: 1.0, 2: 0.7957, 3: 0.6494}
    for k, ((dx, dy, dz), code) in enumerate(ORDER):
        cls = abs(dx) + abs(dy) + abs(dz)
        assert inc[k] == pytest.approx(0.1 * shape[cls] * (tr["rhomax"][2] - 1e11) / tr["rhomax"][1], rel=1e-9)
    o.close()


def _nes_drdt(D0, rho, T=1000.0, c3=1.0, c5=0.65, kappa2=0.1, alpha3=1.0, H=2.9 * EC):
    """dr/dt of the Nes vacancy-core-diffusion network growth, r = rho^-1/2:
    pi C3 C5 D0 exp(-H/kT) / (kappa2 r) * exp(alpha3 kappa2 G b^3 / kT)
    (SCORE_RecoveryModels.cpp:50-80)"""
    G, b, _ = _atomistic(T)
    r = rho ** -0.5
    return math.pi * c3 * c5 * D0 * math.exp(-H / (KB * T)) / (kappa2 * r) * math.exp(alpha3 * kappa2 * G * b ** 3 / (KB * T))


def test_recovery_vacancy_core_diffusion_euler_step_and_dt_rule():
    """r = rho^-1/2 of every grain takes one forward-Euler step dr = dr/dt * dt with the
    time step just done; the next time step is the smaller of the cells rule and
    (1/sqrt(SMALL_RHO)) / max(dr/dt at rhomax, 1/sqrt(SMALL_RECOVERYRATE))
    (get_dtmax_instantslope_rho, SCORE_Kernel.cpp:5594-5636)"""
    n = 20
    dom = one_grain(n, [lin(10, 10, 10, n, n)])
    D0 = 2e-5
    o = Oracle(dom, ParamsSpec(recovery_mode=1, rec_vacancy_d0=D0, rh_hagb_m0=20.0))
    o.step(40)
    tr = o.trajectory()
    G, b, gb = _atomistic()
    m = 20.0 * math.exp(-2.8 * EC / (KB * 1000.0))
    floor = math.sqrt(2.77e10 / 1e11)
    used = set()
    for j in range(len(tr) - 1):
        r = tr["rhomax"][j] ** -0.5 + _nes_drdt(D0, tr["rhomax"][j]) * tr["dt"][j]
        assert tr["rhomax"][j + 1] == pytest.approx(r ** -2, rel=1e-12)
        cells = 0.1 * 0.5e-6 / (m * gb * tr["rhomax"][j])
        rule = 1e11 ** -0.5 / max(_nes_drdt(D0, tr["rhomax"][j]), 2.77e10 ** -0.5)
        assert rule == pytest.approx(floor, rel=1e-12)
        assert tr["dt"][j + 1] == pytest.approx(min(cells, rule), rel=1e-12)
        used.add(cells < rule)
    assert used == {True, False}           # both rules set the step in this run
    assert tr["rhomax"][-1] < 0.8e15
    o.close()
    # a fast recovery: the rule's rate is above its floor and sets the second step
    o = Oracle(dom, ParamsSpec(recovery_mode=1, rec_vacancy_d0=1.0))
    o.step(3)
    tr = o.trajectory()
    assert _nes_drdt(1.0, tr["rhomax"][0]) > 2.77e10 ** -0.5
    assert tr["dt"][1] == pytest.approx(1e11 ** -0.5 / _nes_drdt(1.0, tr["rhomax"][0]), rel=1e-12)
    o.close()


def test_recovery_solute_drag_dt_rule():
    """the solute-drag model changes no rho; its rule divides 1/sqrt(SMALL_RHO) by
    C3 b D0 exp(-H/kT) exp(alpha3 ls G b^3 / kT * rhomax^2), ls = LsPropFactor c^(-2/3),
    floored at 1/sqrt(SMALL_RECOVERYRATE).  With alpha3 > 0 the exponent overflows
    to infinity for any realistic rhomax, the time step is 0 from the start and the
    domain never grows (SCORE_Kernel.cpp:5618-5630)"""
    n = 20
    dom = one_grain(n, [lin(10, 10, 10, n, n)])
    G, b, gb = _atomistic()
    for D0, want in ((1.0, math.sqrt(2.77e10 / 1e11)),
                     (1e14, 1e11 ** -0.5 / (1.0 * b * 1e14 * math.exp(-1.6 * EC / (KB * 1000.0))))):
        o = Oracle(dom, ParamsSpec(recovery_mode=2, rec_alpha3=0.0, rec_solute_d0=D0))
        o.step(4)
        tr = o.trajectory()
        assert np.all(tr["rhomax"] == 1e15)
        # the first step is the initial 1e-6 s of the temperature rule at t = 0
        assert tr["dt"][0] == 1e-6 and np.all(tr["dt"][1:] == pytest.approx(want, rel=1e-12))
        o.close()
    o = Oracle(dom, ParamsSpec(recovery_mode=2, nmax=30))
    r = o.solve_rxgrowth()
    tr = o.trajectory()
    assert r["t"] == 0.0 and np.all(tr["dt"] == 0.0) and r["step"] == 31 and r["xcells"] == 1
    o.close()


def test_recycled_slot_reused_first():
    """slot 0 (retired nucleus) is the first recycled slot: the next infection takes it"""
    n = 20
    dom = one_grain(n, [lin(10, 10, 10, n, n)])
    o = Oracle(dom)
    while True:
        o.step(1)
        tr = o.trajectory()
        if tr["ninfected"][-1] > 0 and len(tr) > 1:
            break
    fi = o.front(abi.LIST_INSIDE)
    # the six face neighbours complete together; in slot order the first is (10,9,10)
    # (slot 2, offset (0,-1,0)); its first free target (-1,-1,0) -> (9,8,10) takes
    # the recycled slot 0, the following ones append after slot 26 in call order
    assert tr["nfull_inside"][-1] == 6
    assert fi["activity"][0] == 1 and fi["cell"][0] == lin(9, 8, 10, n, n) and fi["infector"][0] == 16
    assert [int(c) for c in fi["cell"][27:30]] == [lin(10, 8, 10, n, n), lin(11, 8, 10, n, n), lin(9, 8, 9, n, n)]
    assert list(fi["infector"][27:30]) == [15, 14, 25]


def test_border_nucleus_wraps_xy_and_drops_z():
    n = 12
    corner = lin(0, 0, 0, n, n)
    dom = one_grain(n, [corner])
    o = Oracle(dom)
    assert len(o.front(abi.LIST_BORDER)["activity"]) == 1
    o.step(1)
    g = o.grainmap()
    inf = np.nonzero(g == abi.CURRENTLY_INFECTED)[0]
    # z = -1 targets are dropped: 17 of 26 remain; x and y wrap periodically
    assert len(inf) == 17
    xs, ys, zs = inf % n, (inf // n) % n, inf // (n * n)
    assert set(xs) == {0, 1, n - 1} and set(ys) == {0, 1, n - 1} and set(zs) == {0, 1}
    fi, fb = o.front(abi.LIST_INSIDE), o.front(abi.LIST_BORDER)
    # the only strictly interior target (1,1,1) goes to the Inside list
    assert list(fi["cell"][fi["activity"] == 1]) == [lin(1, 1, 1, n, n)]
    assert (fb["activity"] == 1).sum() == 16


def test_inside_parent_near_face_feeds_border_list():
    n = 12
    dom = one_grain(n, [lin(1, 6, 6, n, n)])   # interior, but within two cells of the x face
    o = Oracle(dom)
    assert len(o.front(abi.LIST_INSIDE)["activity"]) == 1
    o.step(1)
    fi, fb = o.front(abi.LIST_INSIDE), o.front(abi.LIST_BORDER)
    assert (fi["activity"] == 1).sum() == 0 and (fb["activity"] == 1).sum() == 26
    # in call order, appended to the Border list
    assert list(fb["infector"][:26]) == [c for _, c in ORDER]


def test_duplicate_site_rejected_in_pool_order():
    n = 12
    s = lin(6, 6, 6, n, n)
    dom = one_grain(n, [s, s, lin(3, 3, 3, n, n)])
    o = Oracle(dom)
    assert o.result()["nuclei_placed"] == 2
    fi = o.front(abi.LIST_INSIDE)
    assert list(fi["rxg"]) == [0, 2]


def _check_complete_run(dom, prm):
    o = Oracle(dom, prm)
    r = o.solve_rxgrowth()
    tr = o.trajectory()
    g = o.grainmap()
    dc, rc = o.counts()
    nd = dom.n_defg
    assert r["completed"] == 1
    # cell conservation: counts + infected = N; the grid agrees with the counts
    ninf = int((g == abi.CURRENTLY_INFECTED).sum())
    assert int(dc.sum()) + int(rc.sum()) == dom.N
    assert np.array_equal(np.bincount(g[g < nd], minlength=nd) + 0, dc - np.bincount(
        dom.cellgrid[g == abi.CURRENTLY_INFECTED], minlength=nd))
    rx = g[(g >= nd) & (g != abi.CURRENTLY_INFECTED)] - nd
    assert np.array_equal(np.bincount(rx, minlength=dom.n_rxg), rc)
    # X is an exact cell count, monotone
    assert np.all(np.diff(tr["xcells"]) >= 0) and tr["xcells"][-1] == r["xcells"] == rc.sum()
    assert r["X"] == r["xcells"] / dom.N
    # the time axis adds the recorded steps
    assert np.all(tr["t"][1:] == tr["t"][:-1] + tr["dt"][:-1])
    return o, g, ninf


def _components_26(mask, nx, ny, nz):
    """26-connected components with periodic x and y (numpy flood fill)"""
    lab = -np.ones(mask.shape, np.int64)
    idx = np.argwhere(mask)
    cur = 0
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if (a, b, c) != (0, 0, 0)]
    for z, y, x in idx:
        if lab[z, y, x] >= 0:
            continue
        stack = [(z, y, x)]
        lab[z, y, x] = cur
        while stack:
            cz, cy, cx = stack.pop()
            for dz, dy, dx in offs:
                zz, yy, xx = cz + dz, (cy + dy) % ny, (cx + dx) % nx
                if 0 <= zz < nz and mask[zz, yy, xx] and lab[zz, yy, xx] < 0:
                    lab[zz, yy, xx] = cur
                    stack.append((zz, yy, xx))
        cur += 1
    return cur


def test_complete_run_invariants_and_connectivity():
    dom = synth.make_voronoi_domain(24, 24, 24, n_grains=5, n_nuclei=25, seed=21)
    o, g, _ = _check_complete_run(dom, ParamsSpec(xmax=1.0))
    assert o.result()["xcells"] == dom.N
    gg = g.reshape(24, 24, 24)
    for r in range(dom.n_rxg):
        m = gg == dom.n_defg + r
        if m.any():
            assert _components_26(m, 24, 24, 24) == 1
            assert m.ravel()[dom.rxg_site[r]]


def test_voronoi100_invariants_and_determinism():
    dom = synth.workload("voronoi100", seed=1, scale=0.6)
    o1, g1, _ = _check_complete_run(dom, ParamsSpec())
    o2 = Oracle(dom)
    o2.solve_rxgrowth()
    assert np.array_equal(g1, o2.grainmap())
    assert o1.trajectory().tobytes() == o2.trajectory().tobytes()
    assert o1.result()["X"] > 0.99


def test_avrami_exponent_site_saturated():
    """homogeneous site-saturated growth at constant rate: JMAK exponent n = 3"""
    n = 80
    rng = np.random.default_rng(3)
    sites = rng.choice(n ** 3, 120, replace=False)
    o = Oracle(one_grain(n, sites), ParamsSpec(xmax=0.999))
    o.solve_rxgrowth()
    tr = o.trajectory()
    t, X = tr["t"] + tr["dt"], tr["X"]
    sel = (X > 0.05) & (X < 0.7)
    slope = np.polyfit(np.log(t[sel]), np.log(-np.log(1 - X[sel])), 1)[0]
    assert 2.6 < slope < 3.4


def test_jmak_reference_statistics():
    """Statistics of the reference's shipped 200^3 run: 1 grain (rho 1e15), 947 nuclei,
    Rollett-Holm m0=300, 2.45 eV, 500 C, maxfill 0.1 (example/jmak)."""
    ref = json.load(open(os.path.join(HERE, "golden", "jmak_stats.json")))
    s = ref["setup"]
    n, N = s["nx"], s["nx"] ** 3
    sites = synth.distinct_sites(synth.substream(1, 3), s["n_nuclei"], N)
    dom = one_grain(n, sites, rho=s["rho0"])
    dom.cellsize = s["cellsize_um"] * 1e-6
    T = s["T_C"] + 273.15
    prm = ParamsSpec(rh_hagb_m0=s["RHModelHAGBm0"], rh_hagb_hact_ev=s["RHModelHAGBHact_eV"], G0=s["G0"], b0=s["b0"],
                     sched_time=(0.0, 3600.0, 36000.0), sched_temp=(T, T, T), xmax=1.0, tmax=36000.0)
    o = Oracle(dom, prm)
    o.solve_rxgrowth()
    tr = o.trajectory()
    _, rc = o.counts()
    # total front work (sum of Sv) within 1% of the reference's 124,189,713
    assert int(tr["sv"].sum()) == pytest.approx(sum(ref["Sv"]), rel=0.01)
    # the kinetics X(step) agree step by step
    Xr = np.array(ref["X"])
    m = min(len(Xr), len(tr))
    assert np.abs(tr["X"][:m] - Xr[:m]).max() < 0.05
    # steps to completion within 20% (the tail depends on the largest void)
    assert 0.8 * ref["iterations"] <= len(tr) <= 1.2 * ref["iterations"]
    # first step: 947 nuclei fill and infect up to 26 neighbours each
    assert tr["sv"][0] == ref["Sv"][0] and abs(int(tr["sv"][1]) - ref["Sv"][1]) < 0.02 * ref["Sv"][1]
    # final grain-size distribution
    assert rc.sum() == N and rc.mean() == pytest.approx(ref["gsd_summary"]["mean"], rel=1e-6)
    assert np.median(rc) == pytest.approx(ref["gsd_summary"]["median"], rel=0.06)
    assert rc.std() == pytest.approx(ref["gsd_summary"]["std"], rel=0.12)


def test_border_overflow_is_an_error():
    dom = one_grain(30, [(3 * i + 1) + 30 * (3 * j + 1) + 900 for j in range(10) for i in range(10)],
                    shape=(30, 30, 3))
    o = Oracle(dom, ParamsSpec(xmax=1.0))
    with pytest.raises(OracleError) as e:
        o.solve_rxgrowth()
    assert e.value.code == abi.ERR_CAPACITY


def _records(f, sel):
    return list(zip(f["cell"][sel].tolist(), f["defg"][sel].tolist(), f["rxg"][sel].tolist(),
                    f["infector"][sel].tolist(), f["rxfrac"][sel].view(np.uint64).tolist(),
                    f["P"][sel].view(np.uint64).tolist()))


def test_defragmentation_is_a_stable_compaction():
    """ompshar_defragmentation (SCORE_Kernel.cpp:3877-3934) keeps the slots before
    the first INACTIVE one, moves every later ACTIVE record down in slot order and
    sets nextSlotNeverActive to the end of the compacted block."""
    dom = synth.make_voronoi_domain(30, 28, 26, n_grains=6, n_nuclei=12, seed=1)
    o = Oracle(dom, ParamsSpec())
    while o.result()["X"] < 0.3:
        assert not o.step(1)
    before = {lst: o.front(lst) for lst in (abi.LIST_INSIDE, abi.LIST_BORDER)}
    o.defragment()
    for lst, b in before.items():
        a = o.front(lst)
        act = b["activity"].astype(bool)
        assert not act.all(), "the case needs INACTIVE slots to close"
        first = int(np.argmin(act))
        assert len(a["activity"]) == first + int(act[first + 1:].sum())
        assert a["activity"].all()
        assert _records(a, slice(None)) == _records(b, act)
        assert _records(a, slice(0, first)) == _records(b, slice(0, first))
    o.close()


def test_defragmentation_revives_records_left_behind():
    """The reference's search for the first INACTIVE slot is unbounded
    (SCORE_Kernel.cpp:3888-3891) and the slots behind the compacted block keep
    their records.  After a defragmentation in a step without completed cells the
    next one finds no INACTIVE slot below nextSlotNeverActive, runs on into the
    records left behind and takes them back: those cells are then held twice,
    retire twice, and a deformed grain's uint32 count wraps below zero."""
    dom = synth.make_voronoi_domain(30, 28, 26, n_grains=6, n_nuclei=12, seed=1)
    o = Oracle(dom, ParamsSpec(defrag_at=(0.0, 0.0, 0.0)))
    r = o.solve_rxgrowth()
    g = o.grainmap(abi.MAP_CELLGRID)
    rx_cells = int(((g >= dom.n_defg) & (g < abi.CURRENTLY_INFECTED)).sum())
    dc, rc = o.counts()
    assert int(rc.sum()) == r["xcells"] > rx_cells
    assert int(dc.max()) > dom.N
    o.close()


def test_status_vector_follows_front_and_threshold():
    """ompcrit_clarify_status_infectedcells (SCORE_Io_Characterization.cpp:925-997):
    one entry per ACTIVE record, Inside then Border; the nucleus id (offset by
    n_defg) once rxFrac >= threshold, else the deformed grain.  Checked against
    the front lists themselves at three thresholds, one that puts every cell on
    the deformed side and one that puts them all on the nucleus side."""
    dom = synth.make_voronoi_domain(30, 28, 26, n_grains=6, n_nuclei=12, seed=2)
    o = Oracle(dom, ParamsSpec())
    o.step(25)
    fr = [o.front(lst) for lst in (abi.LIST_INSIDE, abi.LIST_BORDER)]
    act = [f["activity"].astype(bool) for f in fr]
    cells = np.concatenate([f["cell"][a] for f, a in zip(fr, act)])
    frac = np.concatenate([f["rxfrac"][a] for f, a in zip(fr, act)])
    defg = np.concatenate([f["defg"][a] for f, a in zip(fr, act)])
    rxg = np.concatenate([f["rxg"][a] for f, a in zip(fr, act)])
    assert 0.0 < frac.min() < 0.5 < frac.max()
    for thr in (0.5, 2.0, 0.0):
        ixyz, gid = o.status(thr)
        assert np.array_equal(ixyz, cells)
        assert np.array_equal(gid, np.where(frac < thr, defg, dom.n_defg + rxg))
    assert np.array_equal(o.status(2.0)[1], defg)
    assert np.array_equal(o.status(0.0)[1], dom.n_defg + rxg)
    o.close()
