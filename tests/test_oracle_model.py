"""Pin the oracle's model equations with properties the mathematics fixes.

The Eigen-dependent reference headers (GW_utils.hpp, SSA.hpp, GW.hpp) cannot be
compiled here, so their formulas are pinned by properties a transcription error
would break: detailed balance (Kolmogorov cycle criterion) of the published
channel models, the quasi-steady-state reduction behind the lumped RyR rates,
generator row sums, reversal potentials, exchanger thermodynamics, whole-cell
calcium conservation, and SSA transition-rate estimates against the rate table.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle.oracle import GwGlobals, GwParams, OracleModel

F_RT = 96.5 / (8.314 * 310.0)
RT_F = 1.0 / F_RT

# new LCC state for (state, transition index), the map of sample_LCC (src/includes/SSA.tpp:103-197)
LCC_NEXT = {1: (2, 7, 7), 2: (1, 3, 8), 3: (2, 4, 9), 4: (3, 5, 10), 5: (4, 6, 11), 6: (5, 5, 5),
            7: (1, 8, 8), 8: (2, 7, 9), 9: (3, 8, 10), 10: (4, 9, 11), 11: (5, 10, 12), 12: (11, 11, 11)}
# (from, to) of the 12 RyR transitions in storage order (src/includes/GW_utils.tpp:228-251)
RYR_EDGES = [(1, 2), (2, 3), (2, 5), (3, 4), (4, 5), (5, 6), (2, 1), (3, 2), (4, 3), (5, 2), (5, 4), (6, 5)]


def subunit(L, p, V, LCC, y, ryr, clch, CaSS):
    lcc3 = np.zeros(3)
    ryr12 = np.zeros(12)
    inact = C.c_double()
    cl = C.c_double()
    r6 = np.asarray(ryr, dtype=np.int32)
    tot = L.gwo_subunit_rates(C.byref(p), V, LCC, y, r6.ctypes.data, clch, CaSS, lcc3.ctypes.data,
                              C.byref(inact), ryr12.ctypes.data, C.byref(cl))
    return tot, lcc3, inact.value, ryr12, cl.value


def lcc_generator(L, p, V, CaSS):
    Q = np.zeros((13, 13))
    for s in range(1, 13):
        _, r, _, _, _ = subunit(L, p, V, s, 1, [5, 0, 0, 0, 0, 0], 0, CaSS)
        for k in range(3):
            if r[k] != 0.0:
                Q[s, LCC_NEXT[s][k]] += r[k]
    return Q


@pytest.mark.parametrize("V,CaSS", [(-91.382, 1.45e-4), (-40.0, 1e-3), (5.0, 0.02), (30.0, 0.2)])
def test_lcc_detailed_balance(oracle_lib, V, CaSS):
    Q = lcc_generator(oracle_lib, GwParams.default(), V, CaSS)
    # the squares (l, l+1, l+7, l+6) generate the cycle space of the 12-state graph
    for l in range(1, 5):
        fwd = Q[l, l + 1] * Q[l + 1, l + 7] * Q[l + 7, l + 6] * Q[l + 6, l]
        bwd = Q[l, l + 6] * Q[l + 6, l + 7] * Q[l + 7, l + 1] * Q[l + 1, l]
        assert fwd > 0 and fwd == pytest.approx(bwd, rel=1e-12), l
    # open states are leaves reached only from 5 and 11
    assert set(np.nonzero(Q[:, 6])[0]) == {5} and set(np.nonzero(Q[:, 12])[0]) == {11}


def test_lcc_rates_voltage_dependence(oracle_lib):
    # activation rises with depolarisation, deactivation falls (GW_utils.hpp:365-369)
    p = GwParams.default()
    q1 = lcc_generator(oracle_lib, p, -60.0, 1e-4)
    q2 = lcc_generator(oracle_lib, p, 20.0, 1e-4)
    assert q2[1, 2] > q1[1, 2] and q2[2, 1] < q1[2, 1]
    # Ca-dependent inactivation is linear in CaSS
    qa = lcc_generator(oracle_lib, p, 0.5, 1e-4)
    qb = lcc_generator(oracle_lib, p, 0.5, 2e-4)
    for s in range(1, 6):
        assert qb[s, s + 6] == pytest.approx(2 * qa[s, s + 6], rel=1e-14)


def test_ryr_detailed_balance_published_parameters(oracle_lib):
    p = GwParams.default()
    CaSS = 1e-4  # below both lumping thresholds: every transition is simulated
    _, _, _, r, _ = subunit(oracle_lib, p, -80.0, 1, 1, [1, 1, 1, 1, 1, 1], 0, CaSS)
    q = {e: r[k] for k, e in enumerate(RYR_EDGES)}
    fwd = q[(2, 3)] * q[(3, 4)] * q[(4, 5)] * q[(5, 2)]
    bwd = q[(2, 5)] * q[(5, 4)] * q[(4, 3)] * q[(3, 2)]
    # holds to the rounding of the published constants (k23 k34 k45 k52 vs k25 k54 k43 k32)
    assert fwd == pytest.approx(bwd, rel=1e-3)
    # count-weighted: rates scale with the occupancy of the source state
    _, _, _, r2, _ = subunit(oracle_lib, p, -80.0, 1, 1, [2, 3, 1, 4, 2, 5], 0, CaSS)
    for k, (a, _) in enumerate(RYR_EDGES):
        assert r2[k] == pytest.approx(r[k] * [2, 3, 1, 4, 2, 5][a - 1], rel=1e-14)


def test_ryr_lumped_rates_are_the_quasi_steady_state_reduction(oracle_lib):
    p = GwParams.default()
    # per-channel rates at a reference CaSS below the thresholds
    _, _, _, r, _ = subunit(oracle_lib, p, -80.0, 1, 1, [1, 1, 1, 1, 1, 1], 0, 1e-5)
    k34 = r[3] / 1e-10
    k43, k45 = r[8], r[4]
    k56 = r[5] / 1e-10
    k65, k52 = r[11], r[9]
    k54 = r[10] / 1e-10
    k32 = r[7]
    for CaSS in (0.05, 0.3):
        n = [0, 1, 2, 1, 2, 0]  # n34 = 3, n56 = 2
        _, _, _, R, _ = subunit(oracle_lib, p, -80.0, 1, 1, n, 0, CaSS)
        a34 = k34 * CaSS**2
        frac4 = a34 / (a34 + k43)      # rapid equilibrium of 3 <-> 4
        frac5 = k65 / (k56 * CaSS**2 + k65)  # rapid equilibrium of 5 <-> 6
        assert R[3] == 0 and R[8] == 0 and R[5] == 0 and R[11] == 0
        assert R[4] == pytest.approx(3 * k45 * frac4, rel=1e-12)
        assert R[7] == pytest.approx(3 * k32 * (1 - frac4), rel=1e-12)
        assert R[9] == pytest.approx(2 * k52 * frac5, rel=1e-12)
        assert R[10] == pytest.approx(2 * k54 * CaSS**2 * frac5, rel=1e-12)
    # between the thresholds only 5 <-> 6 is lumped
    _, _, _, R, _ = subunit(oracle_lib, p, -80.0, 1, 1, [0, 1, 2, 1, 2, 1], 0, 1e-3)
    assert R[3] > 0 and R[8] > 0 and R[5] == 0 and R[11] == 0


def test_subunit_total_and_side_rates(oracle_lib):
    p = GwParams.default()
    V, CaSS = -10.0, 2e-3
    tot, lcc, inact, ryr, cl = subunit(oracle_lib, p, V, 3, 0, [2, 1, 0, 0, 1, 1], 0, CaSS)
    assert tot == pytest.approx(lcc.sum() + inact + ryr.sum() + cl, rel=1e-14)
    yinf = oracle_lib.gwo_yinfLCC(V)
    tau = oracle_lib.gwo_tauLCC(V)
    assert inact == pytest.approx(yinf / tau, rel=1e-15)   # y = 0 -> 1
    _, _, inact1, _, cl1 = subunit(oracle_lib, p, V, 3, 1, [2, 1, 0, 0, 1, 1], 1, CaSS)
    assert inact1 == pytest.approx((1 - yinf) / tau, rel=1e-15)  # stationary P(y = 1) = yinf
    assert cl == pytest.approx(p.kfClCh * CaSS, rel=1e-15) and cl1 == p.kbClCh


@pytest.mark.parametrize("V", [-91.382, -40.0, 0.5, 30.0])
def test_potassium_chain_generators(oracle_lib, V):
    p = GwParams.default()
    QKr, Q14, Q43 = np.zeros(25), np.zeros(100), np.zeros(100)
    oracle_lib.gwo_Q_matrices(C.byref(p), V, QKr.ctypes.data, Q14.ctypes.data, Q43.ctypes.data)
    QKr, Q14, Q43 = QKr.reshape(5, 5), Q14.reshape(10, 10), Q43.reshape(10, 10)
    for Q in (QKr, Q14, Q43):
        off = Q - np.diag(np.diag(Q))
        assert np.all(off >= 0)
        assert np.all(np.abs(Q.sum(axis=1)) <= 1e-15 * np.abs(off).sum(axis=1) + 1e-300)
    # IKr: the 2-3-4 cycle obeys detailed balance by construction of Q(4,2) (GW.tpp:42)
    assert QKr[2, 3] * QKr[3, 4] * QKr[4, 2] == pytest.approx(QKr[2, 4] * QKr[4, 3] * QKr[3, 2], rel=1e-13)
    assert QKr[1, 2] == p.Kf and QKr[2, 1] == p.Kb
    # Kv1.4 / Kv4.3: the four squares (i, i+1, i+6, i+5) of the two-row chain
    for Q in (Q14, Q43):
        for i in range(4):
            fwd = Q[i, i + 1] * Q[i + 1, i + 6] * Q[i + 6, i + 5] * Q[i + 5, i]
            bwd = Q[i, i + 5] * Q[i + 5, i + 6] * Q[i + 6, i + 1] * Q[i + 1, i]
            if fwd == 0 and bwd == 0:
                continue
            assert fwd == pytest.approx(bwd, rel=1e-13), i


def _rhs(L, g, p=None, N=1250, Istim=0.0, sums=(0.0, 0.0, 0.0, 0.0)):
    p = p or GwParams.default()
    out, cur = np.zeros(36), np.zeros(16)
    rc = L.gwo_global_rhs(C.byref(p), N, C.byref(g), Istim, *sums, out.ctypes.data, cur.ctypes.data)
    names = "INa INab INaCa INaK IKr IKs Ito1 Ito2 IK1 IKp ICaL ICab IpCa Jup Jtr Jxfer".split()
    return rc, out, dict(zip(names, cur))


def test_reversal_potentials_and_gating_indices(oracle_lib):
    g = GwGlobals.default()
    g.Nai, g.Ki = 10.3, 131.84
    ENa = RT_F * math.log(138.0 / g.Nai)
    EK = RT_F * math.log(4.0 / g.Ki)
    g.m, g.h, g.j = 0.3, 0.5, 0.6
    g.V = ENa
    _, _, c = _rhs(oracle_lib, g)
    assert abs(c["INa"]) < 1e-12 and abs(c["INab"]) < 1e-13
    g.V = EK
    for k in range(10):
        g.Kv14[k] = 0.0
    _, _, c = _rhs(oracle_lib, g)
    for name in ("IKr", "IK1", "IKp", "Ito1"):
        assert abs(c[name]) < 1e-13, name
    # ICab reverses at the K-based "ECa" (the reference computes ECa from Ki/Ko, GW.tpp:176)
    g.V = 0.5 * RT_F * math.log(4.0 / g.Ki)
    _, _, c = _rhs(oracle_lib, g)
    assert abs(c["ICab"]) < 1e-15
    # IKs reverses at EKs and is quadratic in xKs
    g.V = RT_F * math.log((4.0 + 0.01833 * 138.0) / (g.Ki + 0.01833 * g.Nai))
    _, _, c = _rhs(oracle_lib, g)
    assert abs(c["IKs"]) < 1e-15
    g.V = 10.0
    g.xKs = 0.1
    _, _, c1 = _rhs(oracle_lib, g)
    g.xKs = 0.2
    _, _, c2 = _rhs(oracle_lib, g)
    assert c2["IKs"] == pytest.approx(4 * c1["IKs"], rel=1e-14)
    # IKr flows only through Kr[3]; IKv43 only through Kv43[4]
    g.V = 0.0 + 10.0
    for k in range(5):
        g.Kr[k] = 0.2 if k != 3 else 0.0
    _, _, c = _rhs(oracle_lib, g)
    assert c["IKr"] == 0.0
    g.Kr[3] = 0.2
    _, _, c = _rhs(oracle_lib, g)
    assert c["IKr"] > 0.0
    for k in range(10):
        g.Kv43[k] = 0.1 if k != 4 else 0.0
        g.Kv14[k] = 0.0
    _, _, c = _rhs(oracle_lib, g)
    assert c["Ito1"] == 0.0


def test_ghk_zero_flux_and_exchanger_equilibrium(oracle_lib):
    g = GwGlobals.default()
    # IKv14 (GHK, GW_utils.hpp:299-303) vanishes when Ki = Ko e^{-VF/RT} and Nai = Nao e^{-VF/RT}
    V = 20.0
    e = math.exp(-V * F_RT)
    g.V, g.Ki, g.Nai = V, 4.0 * e, 138.0 * e
    for k in range(10):
        g.Kv43[k] = 0.0
        g.Kv14[k] = 0.1
    _, _, c = _rhs(oracle_lib, g)
    assert abs(c["Ito1"]) < 1e-9
    # INaCa (3 Na : 1 Ca) vanishes at V = RT/F ln(Nao^3 Cai / (Nai^3 Cao))
    g = GwGlobals.default()
    g.Nai, g.Cai = 10.0, 1e-4
    g.V = RT_F * math.log(138.0**3 * g.Cai / (g.Nai**3 * 2.0))
    _, _, c = _rhs(oracle_lib, g)
    assert abs(c["INaCa"]) < 1e-12
    g.V += 5.0
    _, _, c2 = _rhs(oracle_lib, g)
    assert c2["INaCa"] > 0  # outward above the reversal potential (forward mode)


def test_flux_sum_currents_are_linear_with_the_published_constants(oracle_lib):
    p = GwParams.default()
    g = GwGlobals.default()
    g.V = 10.0
    N = 1250
    _, _, c1 = _rhs(oracle_lib, g, p, N, sums=(1000.0, 0.0, 0.0, 7.0))
    _, _, c2 = _rhs(oracle_lib, g, p, N, sums=(2000.0, 0.0, 0.0, 14.0))
    factor = p.NCaRU / N
    assert c1["ICaL"] == pytest.approx(-1000.0 * (2 * 96.5 * p.VSS) * factor / p.CSA * 1000.0, rel=1e-14)
    assert c2["ICaL"] == pytest.approx(2 * c1["ICaL"], rel=1e-14) and c1["ICaL"] < 0
    assert c2["Ito2"] == pytest.approx(2 * c1["Ito2"], rel=1e-14)
    # Ito2 reverses at the chloride equilibrium potential RT/F ln(Clcyto/Clo)
    g.V = RT_F * math.log(p.Clcyto / p.Clo)
    _, _, c3 = _rhs(oracle_lib, g, p, N, sums=(0.0, 0.0, 0.0, 7.0))
    assert abs(c3["Ito2"]) < 1e-14
    # at V == 0 the analytic limit replaces 0/0
    g.V = 0.0
    rc, d, c4 = _rhs(oracle_lib, g, p, N, sums=(0.0, 0.0, 0.0, 7.0))
    assert rc == 0 and np.all(np.isfinite(d))
    lim = 7.0 * (1e9 * p.Pto2 * 96.5 * factor / p.CSA) * (p.Clo - p.Clcyto)
    assert c4["Ito2"] == pytest.approx(lim, rel=1e-14)
    # IpCa is half-maximal at Cai = KmpCa; Jup vanishes when uptake balances leak
    g = GwGlobals.default()
    g.Cai = p.KmpCa
    _, _, c5 = _rhs(oracle_lib, g, p, N)
    assert c5["IpCa"] == pytest.approx(p.IpCamax / 2, rel=1e-15)
    g.Cai = p.Kmf * (g.CaNSR / p.Kmr)  # Hf == Hr and Vmaxf == Vmaxr
    _, _, c6 = _rhs(oracle_lib, g, p, N)
    assert abs(c6["Jup"]) < 1e-18


def test_rest_state_is_near_equilibrium(oracle_lib):
    g = GwGlobals.default()
    N = 12500
    sums = (0.0, N * 4 * 200.0 * (1.45370e-4 - 1.45273e-4), N * 0.333 * (0.908882 - 0.908408), N * 4 * 0.002)
    rc, d, c = _rhs(oracle_lib, g, N=N, sums=sums)
    assert rc == 0
    currents = np.array([c[k] for k in ("INa", "INab", "INaCa", "INaK", "IKr", "IKs", "Ito1", "Ito2", "IK1",
                                        "IKp", "ICaL", "ICab", "IpCa")])
    # the published rest state balances currents of ~0.4 pA/pF to within 0.03 pA/pF
    assert abs(d[0]) < 0.1 * np.abs(currents).max()
    assert abs(d[1]) < 1e-5 * g.Nai and abs(d[2]) < 1e-6 * g.Ki


def total_calcium(p, N, g, s):
    f = p.NCaRU / N
    cyto = p.Vcyto * (g.Cai + p.CMDNT * g.Cai / (p.KCMDN + g.Cai) + g.CaLTRPN + g.CaHTRPN)
    nsr = p.VNSR * g.CaNSR
    jsr = p.VJSR * (s["CaJSR"] + p.CSQNT * s["CaJSR"] / (p.KCSQN + s["CaJSR"])).sum()
    ss = s["CaSS"]
    dyad = p.VSS * (ss + p.BSRT * ss / (p.KBSR + ss) + p.BSLT * ss / (p.KBSL + ss)).sum()
    return cyto + nsr + f * (jsr + dyad)


def test_calcium_mass_balance_during_release():
    """With every sarcolemmal Ca flux off, release, uptake, transfer and buffering
    only move calcium between compartments: total Ca is conserved up to the
    integrator's error. Wrong volume ratios, CRU scaling or flux signs break this."""
    p = GwParams.default()
    p.PCaL = 0.0
    p.kNaCa = 0.0
    p.GCab = 0.0
    p.IpCamax = 0.0
    N = 40
    m = OracleModel(N, params=p, seed=7)
    m.sample_rest_state()
    s0 = m.state()
    ryr = s0["RyR"].copy()
    ryr[:, 0, :] = [2, 0, 3, 0, 0, 0]  # three open RyRs in subunit 0 of every CRU trigger release
    m.set_state(globals_=GwGlobals.default(), RyR=ryr)
    g0 = m.globals()
    tot0 = total_calcium(p, N, g0, m.state())
    m.run(1e-3, 20000)
    g1 = m.globals()
    s1 = m.state()
    tot1 = total_calcium(p, N, g1, s1)
    released = p.NCaRU / N * p.VJSR * (s0["CaJSR"].sum() - s1["CaJSR"].sum())
    assert abs(released) > 0.05 * tot0 * 0.01  # something moved
    assert abs(tot1 - tot0) < 2e-4 * tot0, (tot0, tot1)


def test_ssa_realises_the_rate_table():
    """Frozen globals, CaSS pinned at Cai (no LCC flux, no open RyR): every channel
    is an independent CTMC. Transition counts over dwell time estimate each rate,
    which must match the rate table within 4 standard errors, and no transition
    outside the model's graph may occur."""
    p = GwParams.default()
    p.PCaL = 0.0
    V, Ca = 5.0, 1.0e-4
    N, steps, dt = 300, 12000, 1e-3
    m = OracleModel(N, params=p, seed=11)
    g = GwGlobals.default()
    g.V, g.Cai = V, Ca
    rng = np.random.default_rng(5)
    LCC = rng.choice(np.arange(1, 13), size=(N, 4)).astype(np.int32)
    y = rng.integers(0, 2, size=(N, 4)).astype(np.int32)
    clch = rng.integers(0, 2, size=(N, 4)).astype(np.int32)
    ryr = np.zeros((N, 4, 6), np.int32)
    ryr[..., 0] = 5
    m.set_state(globals_=g, CaSS=np.full((N, 4), Ca), CaJSR=np.full(N, g.CaNSR), LCC=LCC, LCC_inactivation=y,
                RyR=ryr, ClCh=clch)
    g.CaNSR = g.CaNSR
    m.set_frozen(True)
    dwell = np.zeros(13)
    counts = np.zeros((13, 13))
    ydw, ycnt = np.zeros(2), np.zeros(2)
    cdw, ccnt = np.zeros(2), np.zeros(2)
    prev = m.state()
    for _ in range(steps):
        assert m.step(dt, 0.0) == 0
        cur = m.state()
        a, b = prev["LCC"].ravel(), cur["LCC"].ravel()
        np.add.at(dwell, a, dt)
        ch = a != b
        np.add.at(counts, (a[ch], b[ch]), 1)
        ya, yb = prev["LCC_inactivation"].ravel(), cur["LCC_inactivation"].ravel()
        np.add.at(ydw, ya, dt)
        np.add.at(ycnt, ya[ya != yb], 1)
        ca, cb = prev["ClCh"].ravel(), cur["ClCh"].ravel()
        np.add.at(cdw, ca, dt)
        np.add.at(ccnt, ca[ca != cb], 1)
        prev = cur
    assert np.all(cur["CaSS"] == Ca)
    from oracle.oracle import lib

    Q = lcc_generator(lib(), p, V, Ca)
    Q2 = Q @ Q  # two transitions chained inside one step happen at ~ (Q^2)_ab dt / 2
    checked = 0
    for a in range(1, 13):
        for b in range(1, 13):
            if a == b:
                continue
            if Q[a, b] == 0:
                e2 = Q2[a, b] * dt / 2 * dwell[a]
                assert counts[a, b] <= e2 + 4 * math.sqrt(e2) + 2, (a, b, counts[a, b], e2)
                continue
            expect = Q[a, b] * dwell[a]
            if expect < 30:
                continue
            checked += 1
            assert abs(counts[a, b] - expect) < 4 * math.sqrt(expect) + 0.01 * expect, (a, b, counts[a, b], expect)
    assert checked >= 10
    yinf, tau = lib().gwo_yinfLCC(V), lib().gwo_tauLCC(V)
    for state, rate in ((0, yinf / tau), (1, (1 - yinf) / tau)):
        expect = rate * ydw[state]
        assert abs(ycnt[state] - expect) < 4 * math.sqrt(expect) + 2, (state, ycnt[state], expect)
    for state, rate in ((0, p.kfClCh * Ca), (1, p.kbClCh)):
        expect = rate * cdw[state]
        assert abs(ccnt[state] - expect) < 4 * math.sqrt(expect) + 2, (state, ccnt[state], expect)


@pytest.mark.parametrize("n,steps", [(500, 3000), (4200, 1500)], ids=["small", "sums-per-thread"])
def test_parallel_restatement_matches_the_serial_one(n, steps):
    """The OpenMP build of the restatement (per-CRU streams; the exact flux sums
    of a cell of 4,096 CRUs or more summed per thread) gives the serial build's
    bits through the paced upstroke and the start of release."""
    ist = np.zeros(steps)
    ist[:1000] = 100.0
    runs = []
    for kind in ("strict", "par"):
        m = OracleModel(n, seed=2024, kind=kind)
        m.sample_rest_state()
        m.run(1e-3, steps, ist)
        runs.append((m.globals().as_array(), m.state(), m.sums(), m.digest(), m.iterations()))
    (g0, s0, u0, d0, i0), (g1, s1, u1, d1, i1) = runs
    assert np.array_equal(g0, g1)
    for k in s0:
        assert np.array_equal(s0[k], s1[k]), k
    assert np.array_equal(u0, u1)
    assert np.array_equal(d0[0], d1[0]) and np.array_equal(d0[1], d1[1])
    assert i0 == i1 and i0 > n * steps   # some iterations carried events
