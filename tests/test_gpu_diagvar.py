"""Diagnostic variables and history output on the GPU.

fedg_atmos_calc_diagvar against the oracle, bit for bit, for every name it
diagnoses, dry (no tracers) and moist (three water species), at p = 3 and p = 7
over terrain; its error behaviour; and the history writer over a HEVE and a
HEVI run: the records in the file equal the oracle's diagnostics of the state
the stepper held at each output step.  The oracle is pinned by
tests/test_oracle_diagvar.py."""
import numpy as np
import pytest

import fedg
import fedg_oracle as O
from fedg import cases, history
from test_oracle_diagvar import CVQ, ENGI0, MASS

pytestmark = pytest.mark.gpu

NAMES = ("DENS", "U", "V", "W", "Umet", "Vmet", "PRES_diff", "T", "T_diff", "PT", "PT_diff", "QDRY", "ENGK", "ENGP",
         "ENGI", "ENGT")


@pytest.fixture(autouse=True)
def portable_pow():
    with O.pow_mode(O.POW_PORTABLE):
        yield


def dev(a):
    return fedg.DevField(a.size, np.ascontiguousarray(a))


@pytest.mark.parametrize("p,shape", [(3, (4, 3, 3)), (7, (2, 2, 2))], ids=["p3", "p7"])
@pytest.mark.parametrize("moist", [False, True], ids=["dry", "moist"])
def test_diagvar_bitwise(p, shape, moist):
    print(f"diagvar p={p} {shape} moist={moist}", flush=True)
    NeX, NeY, NeZ = shape
    mesh = fedg.BoxMesh(p, NeX, NeY, NeZ, (0.0, 2e3 * NeX), (0.0, 2e3 * NeY), (0.0, 1.5e3 * NeZ),
                        periodic=(True, True, False))
    mesh.set_topography(lambda x, y: 300.0 * np.exp(-((x - 2e3) / 1.5e3) ** 2))
    bs = cases.base_state(mesh, "constBVFreq")
    om = O.OracleMesh(mesh)
    rng = np.random.default_rng(7 * p + moist)
    n = mesh.Np * mesh.Ne
    DD = 1e-3 * rng.standard_normal(n)
    MX, MY, MZ = (5.0 * rng.standard_normal(n) for _ in range(3))
    Dh, Ph = (np.ascontiguousarray(a[:n]) for a in (bs.DENS_hyd, bs.PRES_hyd))
    P = Ph * (1.0 + 1e-3 * rng.standard_normal(n))
    Q = [0.01 * rng.uniform(0.5, 1.0, n), 1e-4 * rng.uniform(0.0, 1.0, n), 1e-5 * rng.uniform(0.0, 1.0, n)] \
        if moist else []
    R = np.full(n, 287.04) * (1.0 + (0.6 * Q[0] if moist else 0.0))
    CV = np.full(n, 717.6)
    CP = np.full(n, 1004.64) * (1.0 + (0.9 * Q[0] if moist else 0.0))
    zlev = np.ascontiguousarray(mesh.height[:n])
    pos_z = np.ascontiguousarray(mesh.z[:n])
    tr = dict(mass=MASS, cv=CVQ, engi0=ENGI0) if moist else {}
    dm = fedg.DeviceMesh(mesh)
    g = {k: dev(a) for k, a in dict(DD=DD, MX=MX, MY=MY, MZ=MZ, P=P, Dh=Dh, Ph=Ph, R=R, CV=CV, CP=CP, zlev=zlev,
                                    pos_z=pos_z).items()}
    gQ = [dev(a) for a in Q]
    out = fedg.DevField(n)
    for name in NAMES:
        want = om.atmos_calc_diagvar(name, DD, MX, MY, MZ, P, Dh, Ph, R, CP, Q=Q, zlev=zlev, pos_z=pos_z, **tr)
        fedg.atmos_calc_diagvar(dm, name, out, g["DD"], g["MX"], g["MY"], g["MZ"], g["P"], g["Dh"], g["Ph"], g["R"],
                                g["CV"], g["CP"], Q=gQ, zlev=g["zlev"], pos_z=g["pos_z"], **tr)
        got = out.download()
        assert np.array_equal(got, want), f"{name}: max abs diff {np.max(np.abs(got - want))}"
    # "PRES" leaves the output as it is; RH and unknown names are errors; a missing input is an error
    out.upload(np.full(n, 7.0))
    fedg.atmos_calc_diagvar(dm, "PRES", out, g["DD"], PRES=g["P"])
    assert np.all(out.download() == 7.0)
    for name in ("RH", "RHL", "NOPE"):
        with pytest.raises(fedg.FedgError):
            fedg.atmos_calc_diagvar(dm, name, out, g["DD"], g["MX"], g["MY"], g["MZ"], g["P"], g["Dh"], g["Ph"],
                                    g["R"], g["CV"], g["CP"])
    with pytest.raises(fedg.FedgError):
        fedg.atmos_calc_diagvar(dm, "ENGP", out, g["DD"], DENS_hyd=g["Dh"])  # no zlev
    with pytest.raises(fedg.FedgError):
        fedg.atmos_calc_diagvar(dm, "U", out, g["DD"], DENS_hyd=g["Dh"])  # no MOMX
    assert np.all(out.download() == 7.0)


ITEMS = ("DDENS", "THERM", "MOMZ", "U", "W", "T", "PT", "PT_diff", "PRES", "PRES_diff", "ENGK", "ENGP", "ENGI",
         "ENGT", "DENS_hyd", "QDRY")


def _history_run(tmp_path, mesh, bs, q, make, nsteps, tint, dt):
    dm = fedg.DeviceMesh(mesh)
    Dx, Dy = fedg.calc_phyd_hgrad(dm, bs.PRES_hyd, bs.PRES_hyd_ref)
    st = make(dm, bs, Dx, Dy, dt)
    st.set_state(q)
    hf = history.HistoryFile(mesh, dt, basename=str(tmp_path / "history"), tinterval=tint * dt, datatype="REAL8")
    ah = history.AtmosHistory(dm, bs, hf, ITEMS)
    states = {}
    for step in range(nsteps + 1):
        if step:
            st.run(1)
        s = st.get_state()
        if ah.history(step, s):
            states[step] = s
    ah.close()
    hf.close()
    st.close()
    return hf.path, states


def _check_history(path, mesh, bs, states, dt):
    om = O.OracleMesh(mesh)
    n = mesh.Np * mesh.Ne
    time, _, data = history.read_history(path, mesh)
    assert np.array_equal(time, [s * dt for s in sorted(states)])
    Dh, Ph = (np.ascontiguousarray(a[:n]) for a in (bs.DENS_hyd, bs.PRES_hyd))
    R, CP = (np.ascontiguousarray(a[:n]) for a in (bs.Rtot, bs.CPtot))
    zlev, pos_z = np.ascontiguousarray(mesh.height[:n]), np.ascontiguousarray(mesh.z[:n])
    for k, step in enumerate(sorted(states)):
        DD, DR, MZ, MX, MY = states[step]
        DRf = mesh.field()
        DRf[:n] = DR[:n]
        P = np.ascontiguousarray(om.drhot2pres(DRf, bs)[0][:n])
        want = {"DDENS": DD[:n], "THERM": DR[:n], "MOMZ": MZ[:n], "PRES": P, "DENS_hyd": Dh}
        for name in ITEMS:
            if name not in want:
                want[name] = om.atmos_calc_diagvar(name, DD[:n], MX[:n], MY[:n], MZ[:n], P, Dh, Ph, R, CP,
                                                   zlev=zlev, pos_z=pos_z)
            assert np.array_equal(data[name][k], want[name]), f"step {step} {name}"


def test_history_of_a_heve_run(tmp_path):
    print("history HEVE p=3 6x2x4", flush=True)
    mesh = fedg.BoxMesh(3, 6, 2, 4, (0.0, 1.5e4), (0.0, 5e3), (0.0, 2.5e3))
    bs = cases.base_state(mesh, "constBVFreq")
    q = cases.warm_bubble_state(mesh, bs, [(2.0, 2.0e3, (7.0e3, 2.5e3, 1.2e3))])
    path, states = _history_run(tmp_path, mesh, bs, q, fedg.HeveStepper, 6, 2, 0.05)
    assert sorted(states) == [0, 2, 4, 6]
    _check_history(path, mesh, bs, states, 0.05)


def test_history_of_a_hevi_run_over_terrain(tmp_path):
    print("history HEVI p=3 8x2x4 terrain", flush=True)
    mesh = fedg.BoxMesh(3, 8, 2, 4, (0.0, 8.0e3), (0.0, 2.0e3), (0.0, 4.0e3))
    mesh.set_topography(lambda x, y: 200.0 * np.exp(-((x - 4e3) / 1.5e3) ** 2))
    bs = cases.base_state(mesh, "constBVFreq")
    q = cases.igw_state(mesh, bs, xc=4.0e3, a=1.0e3)
    path, states = _history_run(tmp_path, mesh, bs, q, fedg.HeviStepper, 3, 3, 0.5)
    assert sorted(states) == [0, 3]
    _check_history(path, mesh, bs, states, 0.5)


def test_history_with_tracers(tmp_path):
    """Moist history: the tracers' own items, QDRY and the energies from the
    mixture's specific heats (fedg_tracer_specific_heat), against the oracle."""
    print("history HEVE p=3 4x2x3 moist", flush=True)
    from test_oracle_negfix import THERMO
    mesh = fedg.BoxMesh(3, 4, 2, 3, (0.0, 1e4), (0.0, 5e3), (0.0, 3e3), periodic=(True, True, False))
    bs = cases.base_state(mesh, "constBVFreq")
    q = cases.warm_bubble_state(mesh, bs, [(2.0, 1.5e3, (5e3, 2.5e3, 1.2e3))])
    n = mesh.Np * mesh.Ne
    rng = np.random.default_rng(5)
    Qh = [0.01 * rng.uniform(0.5, 1.0, n), 1e-4 * rng.uniform(0.0, 1.0, n), 1e-5 * rng.uniform(0.0, 1.0, n)]
    names = ("QV", "QC", "QI")
    dm = fedg.DeviceMesh(mesh)
    Dx, Dy = fedg.calc_phyd_hgrad(dm, bs.PRES_hyd, bs.PRES_hyd_ref)
    st = fedg.HeveStepper(dm, bs, Dx, Dy, 0.05)
    st.set_state(q)
    gQ = [dev(a) for a in Qh]
    tracers = {nm: (g, MASS[i], THERMO[1][i], CVQ[i], THERMO[3][i], ENGI0[i], nm, "kg/kg")
               for i, (nm, g) in enumerate(zip(names, gQ))}
    items = ("QV", "QI", "QDRY", "ENGI", "ENGT", "T", "RTOT", "CPTOT")
    hf = history.HistoryFile(mesh, 0.05, basename=str(tmp_path / "hmoist"), tinterval=0.1, datatype="REAL8")
    ah = history.AtmosHistory(dm, bs, hf, items, tracers=tracers)
    states = {}
    for step in range(3):
        if step:
            st.run(1)
        s = st.get_state()
        if ah.history(step, s):
            states[step] = s
    ah.close()
    hf.close()
    st.close()
    assert sorted(states) == [0, 2]
    om = O.OracleMesh(mesh)
    R, CV, CP = mesh.field(), mesh.field(), mesh.field()
    om.tracer_specific_heat([np.concatenate([a, np.zeros(mesh.field().size - n)]) for a in Qh], THERMO, R, CV, CP)
    bsm = cases.BaseState(DENS_hyd=bs.DENS_hyd, PRES_hyd=bs.PRES_hyd, THERM_hyd=bs.THERM_hyd,
                          PRES_hyd_ref=bs.PRES_hyd_ref, Rtot=R, CVtot=CV, CPtot=CP, CORIOLIS=bs.CORIOLIS,
                          PT_hyd=bs.PT_hyd)
    time, _, data = history.read_history(hf.path, mesh)
    Dh, Ph = (np.ascontiguousarray(a[:n]) for a in (bs.DENS_hyd, bs.PRES_hyd))
    zlev, pos_z = np.ascontiguousarray(mesh.height[:n]), np.ascontiguousarray(mesh.z[:n])
    for k, step in enumerate(sorted(states)):
        DD, DR, MZ, MX, MY = states[step]
        DRf = mesh.field()
        DRf[:n] = DR[:n]
        P = np.ascontiguousarray(om.drhot2pres(DRf, bsm)[0][:n])
        want = {"QV": Qh[0], "QI": Qh[2], "RTOT": R[:n], "CPTOT": CP[:n]}
        for name in ("QDRY", "ENGI", "ENGT", "T"):
            want[name] = om.atmos_calc_diagvar(name, DD[:n], MX[:n], MY[:n], MZ[:n], P, Dh, Ph,
                                               np.ascontiguousarray(R[:n]), np.ascontiguousarray(CP[:n]), Q=Qh,
                                               mass=MASS, cv=CVQ, engi0=ENGI0, zlev=zlev, pos_z=pos_z)
        for name in items:
            assert np.array_equal(data[name][k], want[name]), f"step {step} {name}"
