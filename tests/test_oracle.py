"""Pins the oracle (oracle/bolt_oracle.cpp) before anything is checked against it.

Each check is chosen so that a plausible slip in the restatement -- a wrong opposite
index, an even/odd swap, a sign in a linear term, a dropped quadratic term, a forcing
shift divided by the wrong omega, a TRT half with the wrong sign -- fails:

* velocity tables equal the fixture parsed from the reference's velocities.rs / lib.cl;
* equilibrium moments equal their closed forms (including the D3Q7 0.95 quirk);
* omega = 0: two AA steps are an exact permutation equal to textbook streaming with
  full-way bounce-back (bit-exact);
* the full update matches an independent float64 push/pull LBM (tests/lbm_textbook.py)
  for BGK and TRT, D2Q9 and D3Q27, with walls, fixed cells, gravity and tracers;
* TRT(w, w) == BGK(w);
* the reference's own integration asserts (test_core.py:57-127) hold on the oracle.
"""

import json
import os

import numpy as np
import pytest

import oracle
import lbm_textbook as tb

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "velocity_sets.json")


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()


@pytest.mark.parametrize("d,q,name", [(2, 9, "D2Q9"), (2, 5, "D2Q5"), (3, 7, "D3Q7"),
                                      (3, 27, "D3Q27")])
def test_velocity_tables_match_reference(d, q, name):
    from fractions import Fraction
    gold = json.load(open(GOLDEN))
    ref = gold["velocities_rs"][name]
    c, w, opp = oracle.velocity_set(d, q)
    assert c.tolist() == ref["c"]
    np.testing.assert_array_equal(w, np.array([float(Fraction(x)) for x in ref["w"]],
                                              np.float32))
    # the same ordering is hard-coded in every lib.cl kernel of this set
    for kname, qs in gold["lib_cl_qs"].items():
        if kname.startswith(f"update_d{d}q{q}_"):
            assert qs == ref["c"]
    # opp pairs (0), (1,2), (3,4), ... as the AA addressing relies on
    assert opp[0] == 0
    for k in range(1, q, 2):
        assert opp[k] == k + 1 and opp[k + 1] == k
    # textbook table agrees too
    assert tb.C[(d, q)] == [tuple(r) for r in ref["c"]]
    np.testing.assert_allclose(tb.weights(d, q), [float(Fraction(x)) for x in ref["w"]])


@pytest.mark.parametrize("d,q", [(2, 9), (3, 27), (2, 5), (3, 7)])
def test_equilibrium_moments(d, q):
    rng = np.random.default_rng(1)
    n = 64
    rho = (1 + 0.1 * rng.standard_normal(n)).astype(np.float32)
    u = (0.05 * rng.standard_normal((n, d))).astype(np.float32)
    f = oracle.equilibrate(q, rho, u).astype(np.float64)
    c = np.array(tb.C[(d, q)], np.float64)
    r64, u64 = rho.astype(np.float64), u.astype(np.float64)
    uu = np.sum(u64 * u64, -1)
    if q in (9, 27, 5):
        np.testing.assert_allclose(f.sum(-1), r64, rtol=1e-6)
        np.testing.assert_allclose(f @ c, r64[:, None] * u64, atol=1e-7)
    else:  # D3Q7: sum = C (1 - 0.075 u^2), first moment 0.95 C u
        np.testing.assert_allclose(f.sum(-1), r64 * (1 - 0.075 * uu), rtol=1e-6)
        np.testing.assert_allclose(f @ c, 0.95 * r64[:, None] * u64, atol=1e-7)
    if q in (9, 27):  # second moment rho (I/3 + u u)
        P = np.einsum("nq,qa,qb->nab", f, c, c)
        want = r64[:, None, None] * (np.eye(d) / 3 + u64[:, :, None] * u64[:, None, :])
        np.testing.assert_allclose(P, want, atol=2e-7)
    # and against the textbook equilibrium directly
    np.testing.assert_allclose(f, tb.feq(d, q, r64, u64), atol=1e-7)


def _flags_mixed(counts, rng, fixed=True):
    flags = np.zeros(counts, np.int32)
    wallfrac = rng.random(counts) < 0.12
    flags[wallfrac] = 1
    if fixed:
        fx = (rng.random(counts) < 0.06) & ~wallfrac
        flags[fx] = rng.choice([2, 4, 6], size=int(fx.sum()))
    # higher bits are ignored by the kernels
    flags |= (rng.random(counts) < 0.1).astype(np.int32) * 16
    return flags


@pytest.mark.parametrize("counts", [(12, 10), (7, 6, 5)])
def test_omega_zero_is_streaming_with_bounceback(counts):
    d = len(counts)
    q = 9 if d == 2 else 27
    rng = np.random.default_rng(2)
    flags = _flags_mixed(counts, rng, fixed=False)
    F0 = rng.random(counts + (q,)).astype(np.float32) + 0.5
    f = F0.copy()
    rho = np.ones(counts, np.float32)
    vel = np.zeros(counts + (d,), np.float32)
    for step in range(4):
        oracle.update_ns(counts, q, ("bgk", 0.0), 1 - step % 2, f, rho, vel, flags)
    # textbook: omega = 0 leaves pre-collision values in place; walls reverse
    sim = tb.TextbookSim(counts, q, ("bgk", 0.0), flags, np.ones(counts),
                         np.zeros(counts + (d,)))
    sim.F = F0.astype(np.float64)
    for _ in range(4):
        sim.step()
    m = tb.reachable_mask(flags, d, q)
    np.testing.assert_array_equal(f.astype(np.float64)[m], sim.F[m])
    # the map is a permutation: the multiset of values is unchanged
    np.testing.assert_array_equal(np.sort(f.ravel()), np.sort(F0.ravel()))


def _setup(counts, rng, q, wall_fixed=True):
    d = len(counts)
    flags = _flags_mixed(counts, rng, fixed=wall_fixed)
    rho = (1 + 0.02 * rng.standard_normal(counts)).astype(np.float32)
    vel = (0.04 * rng.standard_normal(counts + (d,))).astype(np.float32)
    return flags, rho, vel


CASES = [
    ((16, 12), 9, ("bgk", 1 / 0.6)),
    ((16, 12), 9, ("trt", 1 / 0.55, 1 / (0.25 / 0.05 + 0.5))),
    ((8, 6, 7), 27, ("bgk", 1 / 0.7)),
    ((8, 6, 7), 27, ("trt", 1 / 0.6, 1 / (0.25 / 0.1 + 0.5))),
]


@pytest.mark.parametrize("counts,q,omega", CASES)
@pytest.mark.parametrize("gravity", [False, True])
def test_update_matches_textbook(counts, q, omega, gravity):
    d = len(counts)
    rng = np.random.default_rng(3)
    flags, rho, vel = _setup(counts, rng, q)
    omega = (omega[0],) + tuple(float(np.float32(x)) for x in omega[1:])
    sim = oracle.OracleSimulation(counts, omega, q=q)
    sim.flags[...] = flags
    sim.rho[...] = rho
    sim.vel[...] = vel
    g = np.array([1e-4, -2e-4, 3e-5][:d], np.float32)
    acc = None
    if gravity:
        sim.set_gravity(g)
        acc = np.broadcast_to(g.astype(np.float64), counts + (d,))
    ref = tb.TextbookSim(counts, q, omega, flags, rho, vel, acc=acc)
    for _ in range(3):
        sim.iterate(2)
        ref.step()
        ref.step()
        m = tb.reachable_mask(flags, d, q)
        np.testing.assert_allclose(sim.f.astype(np.float64)[m], ref.F[m], rtol=0, atol=2e-6)
        nw = (flags & 1) == 0
        np.testing.assert_allclose(sim.rho[nw], ref.rho[nw], atol=2e-6)
        np.testing.assert_allclose(sim.vel[nw], ref.vel[nw], atol=2e-6)


@pytest.mark.parametrize("counts,q", [((16, 12), 5), ((8, 6, 7), 7)])
def test_tracer_matches_textbook(counts, q):
    d = len(counts)
    nsq = 9 if d == 2 else 27
    rng = np.random.default_rng(4)
    flags, rho, vel = _setup(counts, rng, nsq)
    flags[rng.random(counts) < 0.05] |= 8
    om = float(np.float32(1 / 0.8))
    sim = oracle.OracleSimulation(counts, ("bgk", float(np.float32(1 / 0.6))), q=nsq)
    sim.flags[...] = flags
    sim.rho[...] = rho
    sim.vel[...] = vel
    tr = sim.add_tracer("T", om, q)
    Cv = rng.random(counts).astype(np.float32)
    tr.C[...] = Cv
    ref = tb.TextbookSim(counts, nsq, sim.omega, flags, rho, vel)
    ref.add_tracer(q, om, Cv)
    for _ in range(3):
        sim.iterate(2)
        ref.step()
        ref.step()
        m = tb.reachable_mask(flags, d, q)
        np.testing.assert_allclose(tr.g.astype(np.float64)[m], ref.tracers[0][2][m], atol=2e-6)
        nw = (flags & 1) == 0
        np.testing.assert_allclose(tr.C[nw], ref.tracers[0][3][nw], atol=2e-6)


@pytest.mark.parametrize("counts,q", [((16, 12), 9), ((6, 7, 8), 27)])
def test_trt_equal_rates_is_bgk(counts, q):
    rng = np.random.default_rng(5)
    flags, rho, vel = _setup(counts, rng, q)
    w = float(np.float32(1 / 0.7))
    out = []
    for om in (("bgk", w), ("trt", w, w)):
        s = oracle.OracleSimulation(counts, om, q=q)
        s.flags[...] = flags
        s.rho[...] = rho
        s.vel[...] = vel
        s.iterate(6)
        out.append(s.f.copy())
    np.testing.assert_allclose(out[0], out[1], rtol=0, atol=5e-7)


def test_f64_path_matches_textbook_tightly():
    counts, q = (10, 8, 6), 27
    rng = np.random.default_rng(6)
    flags, rho, vel = _setup(counts, rng, q)
    om = ("trt", 1 / 0.6, 1 / 2.0)
    s = oracle.OracleSimulation(counts, om, q=q, dtype=np.float64)
    s.flags[...] = flags
    s.rho[...] = rho
    s.vel[...] = vel
    g = np.array([1e-4, 0, -1e-4])
    s.set_gravity(g)
    ref = tb.TextbookSim(counts, q, om, flags, rho, vel,
                         acc=np.broadcast_to(g.astype(np.float32).astype(np.float64),
                                             counts + (3,)))
    s.iterate(4)
    for _ in range(4):
        ref.step()
    m = tb.reachable_mask(flags, 3, q)
    np.testing.assert_allclose(s.f[m], ref.F[m], rtol=0, atol=1e-13)


@pytest.mark.parametrize("counts,q", [((16, 12), 5), ((8, 6, 7), 7)])
def test_f64_tracer_matches_textbook_tightly(counts, q):
    # orc_update_ad_f64 on its own pin: a moving fluid (|u| ~ 0.05), fixed-scalar cells and
    # walls, against the float64 push/pull textbook update
    d = len(counts)
    nsq = 9 if d == 2 else 27
    rng = np.random.default_rng(7)
    flags, rho, vel = _setup(counts, rng, nsq)
    flags[rng.random(counts) < 0.05] |= 8
    om = ("bgk", 1 / 0.6)
    s = oracle.OracleSimulation(counts, om, q=nsq, dtype=np.float64)
    s.flags[...] = flags
    s.rho[...] = rho
    s.vel[...] = vel
    Cv = rng.random(counts)
    tr = s.add_tracer("T", 1 / 0.8, q)
    tr.C[...] = Cv
    ref = tb.TextbookSim(counts, nsq, om, flags, rho, vel)
    ref.add_tracer(q, 1 / 0.8, Cv)
    assert np.abs(vel).max() > 0.02
    for _ in range(3):
        s.iterate(2)
        ref.step()
        ref.step()
        m = tb.reachable_mask(flags, d, q)
        np.testing.assert_allclose(tr.g[m], ref.tracers[0][2][m], rtol=0, atol=1e-13)
        nw = (flags & 1) == 0
        np.testing.assert_allclose(tr.C[nw], ref.tracers[0][3][nw], rtol=0, atol=1e-13)


# ---- the reference's own integration asserts (python/tests/test_core.py:57-127) -----

def test_reference_rest_state():  # test_core.py:57-79
    sim = oracle.OracleSimulation((100, 100), ("bgk", float(np.float32(1 / 0.51))))
    assert sim.vel.shape == (100, 100, 2) and sim.rho.shape == (100, 100)
    sim.iterate(10)
    assert abs(sim.vel.mean()) < 1e-4
    assert abs(sim.rho.mean() - 1) < 1e-4


def test_reference_tracer_conservation():  # test_core.py:82-93
    sim = oracle.OracleSimulation((100, 100), ("bgk", float(np.float32(1 / 0.51))))
    tr = sim.add_tracer("tracer", float(np.float32(1 / 0.51)))
    tr.C[10:20, 30:40] = 1.0
    sim.iterate(10)
    assert abs(tr.C.sum() - 10 ** 2) < 1e-4


def test_reference_gravity_kat():  # test_core.py:96-103
    sim = oracle.OracleSimulation((100, 100), ("bgk", 1.0))
    sim.set_gravity(np.array([0, 0.1], np.float32))
    sim.iterate(10)
    assert abs(sim.vel[..., 1].mean() - 0.1 * 10) < 1e-4


@pytest.mark.parametrize("omega", [("bgk", 1.0), ("trt", 1.0, 0.7), ("trt", 1.0, 1.0)])
def test_gravity_kat_3d(omega):
    # uniform state: momentum relaxes with the antisymmetric rate, so with omega+ = 1
    # the stored (shifted) velocity after n steps is (n - 1) * omega- * a + a
    sim = oracle.OracleSimulation((6, 6, 8), omega)
    a = np.float32(0.05)
    sim.set_gravity(np.array([0, 0, a], np.float32))
    sim.iterate(6)
    wm = omega[2] if omega[0] == "trt" else omega[1]
    assert abs(sim.vel[..., 2].mean() - (5 * wm * a + a)) < 1e-5


def test_boussinesq_last_coupling_wins_and_overrides_gravity():
    # simulation.rs:326-373: gravity written once, each coupling overwrites acc per step
    counts = (12, 10)
    base = oracle.OracleSimulation(counts, ("bgk", 1.0))
    t = base.add_tracer("a", 1.0)
    t.C[...] = np.linspace(0, 1, 120, dtype=np.float32).reshape(counts)
    base.add_tracer("b", 1.0).C[...] = 0.25
    base.set_gravity(np.array([0.3, 0.3], np.float32))
    base.add_boussinesq_coupling("a", 0.1, 0.5, np.array([0, -1], np.float32))
    base.add_boussinesq_coupling("b", 0.2, 0.0, np.array([0, 1], np.float32))
    base.iterate(2)
    # only the last coupling acts: uniform acc = (0, 0.2*0.25) at every step (b stays 0.25
    # while u_y grows uniformly), omega = 1 -> vel_y = 2 * 0.05 after two steps
    np.testing.assert_allclose(base.vel[..., 0], 0.0, atol=1e-7)
    np.testing.assert_allclose(base.vel[..., 1], 0.1, atol=1e-6)


def test_mass_conservation_periodic_walls():
    counts, q = (24, 20), 9
    rng = np.random.default_rng(7)
    flags, rho, vel = _setup(counts, rng, q, wall_fixed=False)
    sim = oracle.OracleSimulation(counts, ("trt", float(np.float32(1 / 0.55)),
                                           float(np.float32(1 / 3.0))))
    sim.flags[...] = flags
    sim.rho[...] = rho
    sim.vel[...] = vel
    sim.iterate(2)
    m0 = sim.f.astype(np.float64).sum()
    sim.iterate(40)
    assert abs(sim.f.astype(np.float64).sum() - m0) / m0 < 1e-6


def test_odd_iterate_rejected():
    sim = oracle.OracleSimulation((8, 8), ("bgk", 1.0))
    with pytest.raises(ValueError):
        sim.iterate(3)


# ---- physics closed forms (tests/closed_forms.py) ---------------------------------------
import closed_forms as cf  # noqa: E402


@pytest.mark.parametrize("coll", ["bgk", "trt"])
def test_shear_wave_viscosity(coll):
    got, want = cf.shear_wave(cf.OracleBackend, coll)
    assert abs(got / want - 1) < 2e-3, (got, want)


@pytest.mark.parametrize("coll,magic", [("bgk", 0.25), ("trt", 0.25), ("trt", 3 / 16)])
def test_poiseuille_with_reference_force_semantics(coll, magic):
    # a forcing shift divided by omega- (instead of the reference's omega+) would leave
    # TRT at tau = 0.8, magic 1/4 about 40 % off this profile
    assert cf.poiseuille(cf.OracleBackend, coll, magic=magic) < 2e-3


@pytest.mark.parametrize("d", [2, 3])
def test_gaussian_tracer_diffusivity(d):
    got, want = cf.gaussian_diffusion(cf.OracleBackend, d)
    np.testing.assert_allclose(got, want, rtol=1e-4)


@pytest.mark.parametrize("coll", ["bgk", "trt"])
def test_taylor_green_energy_decay(coll):
    got, want = cf.taylor_green(cf.OracleBackend, coll, n=32, dt=200)
    assert abs(got / want - 1) < 1e-2, (got, want)


def test_oracle_equilibrium_is_the_float32_formula_in_order():
    # orc_equilibrate is velocities.rs:43-52 evaluated in float32 left to right:
    # rho * w * (1 + 3 vq + 4.5 vq vq - 1.5 vv), vq and vv accumulated from 0 in component
    # order; numpy float32 scalars replay that order operation by operation
    rng = np.random.default_rng(0)
    rho = (1 + 0.1 * rng.standard_normal(50)).astype(np.float32)
    u = (0.1 * rng.standard_normal((50, 3))).astype(np.float32)
    f = oracle.equilibrate(27, rho, u)
    c, w, _ = oracle.velocity_set(3, 27)
    one, three, k45, k15 = (np.float32(x) for x in (1.0, 3.0, 4.5, 1.5))
    for n in range(50):
        vv = np.float32(0)
        for a in range(3):
            vv = np.float32(vv + u[n, a] * u[n, a])
        for i in range(27):
            vq = np.float32(0)
            for a in range(3):
                vq = np.float32(vq + u[n, a] * np.float32(c[i, a]))
            val = np.float32(np.float32(rho[n] * w[i]) *
                             np.float32(np.float32(np.float32(one + np.float32(three * vq)) +
                                                   np.float32(np.float32(k45 * vq) * vq)) -
                                        np.float32(k15 * vv)))
            assert f[n, i] == val
