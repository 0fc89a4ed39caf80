"""The explicit HEVI tendency in FP32 perturbation form (fedg_hevi_set_precision
32: flat box, p = 7, dry air; the state, the IMEX combination and the Newton
solve stay FP64) against the FP64 HEVI stepper, whose explicit tendency on the
same p = 7 layout is bit for bit the reference's (tests/test_gpu_hevi.py,
tests/test_gpu_general_inputs.py), on a piece of the inertia-gravity-wave
channel (8x2x8 elements of 1.25 km, dt 0.05 s, 6 IMEX_ARK324 steps).

Without mean wind the FP32 form holds every prognostic to the bound below,
relative to its own field.  With the channel's 20 m/s the momentum is
dominated by the mean flow (MOMX ~ 24, RHOT u ~ 7000) and the FP32 roundoff of
the fluxes' divergence is as large as the wave's own tendency: the error is
then bounded only against the full state (base state plus perturbation), and
the measured errors relative to the perturbations are printed.
"""
import numpy as np
import pytest

import fedg
from fedg import cases

pytestmark = pytest.mark.gpu

# U = 0: relative L2 of each prognostic against its own field after 6 steps
# (measured; the bound leaves a factor ~10)
BOUND_NOWIND = 2e-5
# U = 20 m/s: relative L2 against the full state's scale
BOUND_FULL = 1e-6


def run_pair(U, max_it, rtol):
    case = cases.scaled_case(cases.config4(), 8, 2, 8)
    mesh, bs, _ = cases.build(case)
    q = cases.igw_state(mesh, bs, U=U, xc=0.5 * case.Lx)
    dm = fedg.DeviceMesh(mesh)
    Dx, Dy = fedg.calc_phyd_hgrad(dm, bs.PRES_hyd, bs.PRES_hyd_ref)
    out = {}
    for bits in (64, 32):
        st = fedg.HeviStepper(dm, bs, Dx, Dy, case.dt, scheme=case.scheme, max_it=max_it, rtol=rtol)
        st.set_precision(bits)
        st.set_state(q)
        st.run(6)
        out[bits] = st.get_state()
        st.close()
    dm.close()
    return mesh, bs, out


@pytest.mark.parametrize("max_it,rtol", [(1, 0.0), (8, 1e-10)], ids=["one-iteration", "newton"])
def test_hevi_fp32_tendency_without_mean_wind(max_it, rtol):
    print(f"HEVI FP32 tendency vs FP64, U = 0, max_it={max_it}", flush=True)
    mesh, bs, out = run_pair(0.0, max_it, rtol)
    errs = {}
    for name, a, b in zip(fedg.VARS, out[32], out[64]):
        errs[name] = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
        print(f"  {name}: rel L2 {errs[name]:.2e}", flush=True)
    for name in ("DDENS", "DRHOT", "MOMZ", "MOMX"):  # MOMY stays at roundoff level in both runs
        assert errs[name] <= BOUND_NOWIND, name


def test_hevi_fp32_tendency_with_mean_wind():
    print("HEVI FP32 tendency vs FP64, U = 20 m/s", flush=True)
    mesh, bs, out = run_pair(20.0, 1, 0.0)
    n = mesh.Np * mesh.Ne
    full = [out[64][0] + bs.DENS_hyd[:n], out[64][1] + bs.THERM_hyd[:n]] + [out[64][3]] * 3
    for name, a, b, f in zip(fedg.VARS, out[32], out[64], full):
        e_own = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
        e_full = np.linalg.norm(a - b) / np.linalg.norm(f)
        print(f"  {name}: rel L2 vs its own field {e_own:.2e}, vs the full state {e_full:.2e}", flush=True)
        assert np.all(np.isfinite(a))
        assert e_full <= BOUND_FULL, name


def test_hevi_fp32_refused_off_its_domain():
    case = cases.scaled_case(cases.config4(), 4, 2, 4, porder=3)
    mesh, bs, _ = cases.build(case)
    dm = fedg.DeviceMesh(mesh)
    Dx, Dy = fedg.calc_phyd_hgrad(dm, bs.PRES_hyd, bs.PRES_hyd_ref)
    st = fedg.HeviStepper(dm, bs, Dx, Dy, case.dt, scheme=case.scheme)
    with pytest.raises(RuntimeError):
        st.set_precision(32)
    st.set_precision(64)
    st.close()
    dm.close()
