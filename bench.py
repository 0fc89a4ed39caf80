#!/usr/bin/env python3
"""Benchmark: FElib regional nonhydrostatic DG HEVE step on one MI355X.

Workload: 160 km x 160 km x 10 km box, 64x64x16 hexahedra, p=7 (512 LGL
nodes per element, 33.5M nodes, 167.8M DOF), dry isothermal base state,
three 2 K warm bubbles, classic RK4 (ERK_4s4o), dt = 0.04 s, plain FP64.
A step is one RK4 step = 4 fused stage kernels over all elements, state
resident in HBM.  value = DOF*stage/s = 5 Np Ne * 4 * K / wall time of K steps.

roofline: the fused stage kernel against 8 TB/s, with the algorithmic
traffic of 36.8 B per DOF*stage (q^n, stage state and RK accumulator reads and
writes plus three base-state fields per stage); its duration is the mean of
HIP events recorded around every stage launch on the stepper's stream during
the timed steps.  traffic: HBM bytes per launch from rocprofv3 FETCH_SIZE /
WRITE_SIZE (profiles/heve_stage_flat_p7_64x64x16.md), scaled by a copy kernel.

cpu_baseline: the CPU oracle (C restatement of the reference Fortran, OpenMP,
FP64, the C library's pow) timed on this host over whole RK4 steps of the
same mesh (at least 10 s), with the CPU model, nproc, the affinity mask and
the thread pinning recorded.

forms: the same step with the FP32 perturbation-form tendency (FP64 state and
RK), and the HEVI step of the 240x16x8 p=7 inertia-gravity-wave channel
(IMEX_ARK324, dt = 0.05 s) with the reference's single Newton iteration and
with Newton to a relative residual of 1e-10; and the metric's second form,
the FP32-tendency HEVI step with the learned Newton guess (trained on the GPU
first) against plain FP64 at the same Newton tolerance, on the channel and on
its wide-element variant (5 km elements, dt = 0.2 s), with the errors against
the FP64 run.

forms also times one tracer's advection tendency and FCT coefficient
(atm_dyn_dgm_trcadvect3d_heve) on the same 64x64x16 p=7 mesh.

--workload selects what is timed alone (for profiling): heve (default),
heve-fp32, hevi (the HEVI step with the reference's single iteration) or
trcadv (the tracer tendency and FCT coefficient, --steps calls each).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fe-project_amd"))

BYTES_PER_DOF_STAGE = 36.8
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "heve_stage_flat_p7_64x64x16.md")


def read_traffic():
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                line = line.strip()
                if line.startswith("hbm_bytes_per_launch="):
                    return float(line.split("=")[1])
    except OSError:
        return None
    return None


def host_info():
    """CPU model, logical CPUs of the host and of this process's affinity mask."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return model, os.cpu_count(), affinity


def cpu_baseline(mesh, bs, q, dt, threads, min_seconds=10.0):
    """The CPU oracle (OpenMP, FP64, the C library's pow) over RK4 steps of the
    same mesh until at least min_seconds of wall time; threads pinned one per
    core (OMP_PROC_BIND=close, OMP_PLACES=cores unless the environment sets them)."""
    # the process's affinity mask before the OpenMP runtime binds this thread to one core
    model, ncpu, affinity = host_info()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import fedg_oracle as O
    O.set_pow_mode(O.POW_LIBM)
    om = O.OracleMesh(mesh)
    Dx, Dy = om.calc_phyd_hgrad(bs)
    qq = [a.copy() for a in q]
    steps = 0
    t0 = time.perf_counter()
    while True:
        om.heve_update(qq, bs, Dx, Dy, dt, 1)
        steps += 1
        sec = time.perf_counter() - t0
        if sec >= min_seconds:
            break
    dof_stage = 5 * mesh.Np * mesh.Ne * 4 * steps
    return {"value": dof_stage / sec, "unit": "DOF·stage/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{steps} RK4 steps ({4 * steps} stages) of the same 64x64x16 p=7 mesh, "
                      f"{dof_stage:.3e} DOF·stage, {sec:.2f} s",
            "cpu_model": model, "nproc": ncpu, "affinity_cpus": affinity,
            "pinning": f"OMP_PROC_BIND={os.environ['OMP_PROC_BIND']} OMP_PLACES={os.environ['OMP_PLACES']}"}


def time_heve(fedg, cases, bits, steps, warmup):
    case = cases.config2()
    mesh, bs, q = cases.build(case)
    dm = fedg.DeviceMesh(mesh)
    Dx, Dy = fedg.calc_phyd_hgrad(dm, bs.PRES_hyd, bs.PRES_hyd_ref)
    st = fedg.HeveStepper(dm, bs, Dx, Dy, case.dt, scheme=case.scheme)
    st.set_precision(bits)
    st.set_state(q)
    st.run(warmup)
    fedg.check(fedg.LIB.fedg_device_sync(), "sync")
    t0 = time.perf_counter()
    ms_events, ms_kernel = st.run_timed(steps)
    fedg.check(fedg.LIB.fedg_device_sync(), "sync")
    wall = time.perf_counter() - t0
    flat = st.flat
    st.close()
    return dict(case=case, mesh=mesh, bs=bs, q=q, dm=dm, wall=wall, ms_kernel=ms_kernel, flat=flat)


def time_hevi(fedg, cases, steps, warmup, max_it, rtol, case=None, dm=None):
    case = case or cases.config4()
    mesh, bs, q = cases.build(case)
    dm = dm or fedg.DeviceMesh(mesh)
    Dx, Dy = fedg.calc_phyd_hgrad(dm, bs.PRES_hyd, bs.PRES_hyd_ref)
    st = fedg.HeviStepper(dm, bs, Dx, Dy, case.dt, scheme=case.scheme, max_it=max_it, rtol=rtol)
    st.set_state(q)
    st.run(warmup)
    it0, calls0 = st.newton_stats()
    fedg.check(fedg.LIB.fedg_device_sync(), "sync")
    t0 = time.perf_counter()
    tot, im, ex = st.run_timed(steps)
    wall = time.perf_counter() - t0
    it, calls = st.newton_stats()
    st.close()
    dof = 5 * mesh.Np * mesh.Ne
    return {"ms_per_step": wall * 1e3 / steps, "value": dof * 4 * steps / wall,
            "ms_implicit_per_stage": im, "ms_explicit_per_stage": ex,
            "newton_iterations_per_solve": (it - it0) / max(calls - calls0, 1), "max_it": max_it, "rtol": rtol}, dm


def time_hevi_learned(fedg, cases, steps, dx=1.25e3, dt=None):
    """The metric's second form on the 240x16x8 p=7 channel: the FP32-tendency
    HEVI step with the learned Newton guess, against plain FP64, both with
    Newton to 1e-10 from the same state over the same steps.  The guess network
    is trained first on the GPU (harvest from 8x2x8-element pieces of the
    channel with the same element size and step, seeded wave amplitudes,
    centres and winds; 1500 Adam steps), as tests/test_gpu_learned_guess.py
    does.  dx: horizontal element size (1.25 km: the channel; 5 km: wide
    elements, dt 0.2 s)."""
    import numpy as np
    from fedg import learned
    rtol, max_it = 1e-10, 8
    base = cases.config4()
    piece = cases.scaled_case(base, 8, 2, 8)
    piece.Lx, piece.Ly = 8 * dx, 2 * dx
    if dt is not None:
        piece.dt = dt
    mesh_p, bs_p, _ = cases.build(piece)
    dm_p = fedg.DeviceMesh(mesh_p)
    Dx_p, Dy_p = fedg.calc_phyd_hgrad(dm_p, bs_p.PRES_hyd, bs_p.PRES_hyd_ref)
    states = []
    for sd in (11, 12, 13, 14):
        rng = np.random.default_rng(sd)
        states.append(cases.igw_state(mesh_p, bs_p, U=float(rng.uniform(10.0, 30.0)),
                                      dtheta=float(rng.uniform(0.005, 0.05)), xc=float(rng.uniform(0.2, 0.8) * piece.Lx)))
    t0 = time.perf_counter()
    X, Y = learned.harvest(dm_p, bs_p, Dx_p, Dy_p, piece.dt, piece.scheme, states, nsteps=6, per_stage=256,
                           rtol=rtol, warm=2)
    L = learned.line_length(mesh_p)
    scale = learned.scales_of(X, Y, L)
    Xn, Yn = learned.normalise(X, Y, scale, L)
    net = fedg.Mlp(10 * L + 1, 256, 3, 5 * L, seed=5)
    losses = learned.train(net, Xn, Yn, steps=1500, batch=512, lr=1e-3)
    t_train = time.perf_counter() - t0
    dm_p.close()

    case = cases.config4(dt=dt)
    case.Lx, case.Ly = 240 * dx, 16 * dx
    mesh, bs, _ = cases.build(case)
    q = cases.igw_state(mesh, bs, U=20.0, xc=case.Lx / 3.0)
    dm = fedg.DeviceMesh(mesh)
    Dx, Dy = fedg.calc_phyd_hgrad(dm, bs.PRES_hyd, bs.PRES_hyd_ref)
    runs = {}
    for label, bits, guess in (("fp64_plain", 64, False), ("fp64_learned", 64, True), ("fp32_learned", 32, True)):
        st = fedg.HeviStepper(dm, bs, Dx, Dy, case.dt, scheme=case.scheme, max_it=max_it, rtol=rtol)
        st.set_precision(bits)
        st.set_state(q)
        if guess:
            st.set_guess(net, scale)
        st.run(1)
        it0, c0 = st.newton_stats()
        fedg.check(fedg.LIB.fedg_device_sync(), "sync")
        ta = time.perf_counter()
        st.run(steps)
        wall = time.perf_counter() - ta
        it, c = st.newton_stats()
        lines, fb = st.guess_stats()
        rp, rs = st.guess_residuals()
        runs[label] = dict(state=st.get_state(), ms=wall * 1e3 / steps, it=(it - it0) / max(c - c0, 1),
                           fallback=fb / max(lines, 1), resid=(rs / rp if rp > 0 else None))
        st.close()
    dm.close()
    net.close()
    ref = runs["fp64_plain"]["state"]
    # the full state's scale of every prognostic (MOMY of the y-uniform channel
    # is roundoff in both runs, so its error relative to itself is noise)
    n = mesh.Np * mesh.Ne
    full = [float(np.linalg.norm(ref[0] + bs.DENS_hyd[:n])), float(np.linalg.norm(ref[1] + bs.THERM_hyd[:n]))]
    full += [float(np.linalg.norm(ref[3]))] * 3
    out = {"workload": f"240x16x8 p=7 channel, {dx / 1e3:g} km elements, dt={case.dt} s, U=20 m/s, "
                       f"IMEX_ARK324, Newton to {rtol}", "steps": steps,
           "guess_training": {"rows": int(X.shape[0]), "adam_steps": 1500, "seconds": t_train,
                              "loss_first_last": [float(np.mean(losses[:20])), float(np.mean(losses[-20:]))]}}
    for label, r in runs.items():
        e = {}
        for name, a, b, fs in zip(fedg.VARS, r["state"], ref, full):
            nb = float(np.linalg.norm(b))
            e[name] = {"rel_l2": float(np.linalg.norm(a - b)) / nb if nb > 0 else None,
                       "rel_linf": float(np.max(np.abs(a - b)) / np.max(np.abs(b))) if nb > 0 else None,
                       "rel_l2_full_state": float(np.linalg.norm(a - b)) / fs}
        out[label] = {"ms_per_step": r["ms"], "newton_iterations_per_solve": r["it"], "fallback_rate": r["fallback"],
                      "start_residual_over_plain": r["resid"], "error_vs_fp64_plain": e}
    out["speedup_fp32_learned_over_fp64_plain"] = runs["fp64_plain"]["ms"] / runs["fp32_learned"]["ms"]
    return out


# tracer advection tendency, per node at p=7.  General metrics: 13 node fields
# read (QTRC_, MOMX/Y/Z, fct_coef, RHOQ_tp, J, Gsqrt, G13, G23, E11, E22, E33)
# and one written, plus per face node (0.75 per node) alphDENS_M/_P, the
# normal, Fscale and the two vmap indices.  Flat box (metrics constant): six
# node fields read (J only at face nodes) and one written, plus per face node
# alphDENS_M/_P, J and the two vmap indices.
TRC_BYTES_PER_NODE = {False: 14 * 8 + 0.75 * (6 * 8 + 2 * 4), True: 7 * 8 + 0.75 * (3 * 8 + 2 * 4)}


def time_trcadv(fedg, mesh, bs, q, dm, reps):
    """One tracer's advection tendency and FCT coefficient on the stage mesh,
    host-timed per call (each call ends with a stream synchronisation), and
    whole steps of the tracer-only advection driver (SSP RK3)."""
    import numpy as np
    n, nA, nf = mesh.Np * mesh.Ne, mesh.Np * mesh.NeA, mesh.NfpTot * mesh.Ne
    dens = bs.DENS_hyd + q[0]
    x = np.zeros(nA)
    x[:n] = mesh.x
    Q = 1.0e-3 * (1.0 + 0.5 * np.sin(2.0 * np.pi * x / 160.0e3))
    f = lambda a: fedg.DevField(a.size, a)
    d = dict(Q=f(Q), MX=f(dens * 10.0), MY=f(dens * 5.0), MZ=f(dens * 0.5), R=f(np.ones(nA)), tp=f(np.zeros(nA)),
             J=f(mesh.J), Dh=f(bs.DENS_hyd), DD=f(q[0]), aM=f(np.full(nf, 3.0)), aP=f(np.full(nf, 3.0)),
             out=fedg.DevField(nA))
    del dens, x, Q

    def tend():
        fedg.trcadvect_cal_tend(dm, d["out"], d["Q"], d["MX"], d["MY"], d["MZ"], d["aM"], d["aP"], d["R"], d["tp"],
                                d["J"])

    def fct():
        fedg.trcadvect_calc_fct_coef(dm, d["out"], d["Q"], d["MX"], d["MY"], d["MZ"], d["tp"], d["aM"], d["aP"],
                                     d["Dh"], d["DD"], d["DD"], 1.0, 0.04, d["J"])

    res = {}
    for name, fn in (("cal_tend", tend), ("calc_fct_coef", fct)):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        res["ms_" + name] = (time.perf_counter() - t0) * 1e3 / reps
    Qh = d["Q"].download()
    for v in d.values():
        v.free()
    # whole tracer steps of the tracer-only advection driver (SSP RK3, limiter
    # and TMAR on): per stage two halo exchanges, the FCT coefficient, the
    # tendency and the density-weighted update
    from fedg.tracer import TracerAdvection
    dens = bs.DENS_hyd + q[0]
    st = TracerAdvection(dm, dens * 10.0, dens * 5.0, dens * 0.5, q[0], bs.DENS_hyd, mesh.J, 0.04,
                         scheme="ERK_SSP_3s3o")
    del dens
    gQ = fedg.DevField(nA, Qh)
    st.step(gQ, 1)
    nstep = max(1, reps // 3)
    t0 = time.perf_counter()
    st.step(gQ, nstep)
    res["ms_tracer_step_ssp3"] = (time.perf_counter() - t0) * 1e3 / nstep
    gQ.free()
    st.close()
    bpn = TRC_BYTES_PER_NODE[dm.flat]
    achieved = bpn * n / (res["ms_cal_tend"] * 1e-3) / 1e9
    res.update({"nodes": n, "flat_metrics": dm.flat, "bytes_per_node": bpn, "cal_tend_GBs": achieved,
                "cal_tend_frac_of_hbm_peak": achieved / HBM_PEAK_GBS, "reps": reps,
                "timing": "host wall per call, stream synchronised"})
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("heve", "heve-fp32", "hevi", "trcadv"), default="heve")
    ap.add_argument("--no-forms", action="store_true", help="skip the FP32 and HEVI lines of the default run")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed steps, run one device copy of Np*Ne doubles (copy_kernel), the "
                         "known-size reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--no-probe", action="store_true",
                    help="skip the access-pattern probe (the stage's loads and stores without arithmetic)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU sharding is a design "
              "(horizontal tiles with per-stage face exchange), not an implemented path.", file=sys.stderr)
        sys.exit(2)

    import fedg
    from fedg import cases

    if args.workload == "hevi":
        r, _ = time_hevi(fedg, cases, args.steps, args.warmup, 1, 0.0)
        print(json.dumps({"metric": "DOF·stage/s", "workload": "HEVI inertia-gravity wave 240x16x8 p=7, "
                          "IMEX_ARK324 dt=0.05 s, one Newton iteration", **r}))
        return
    if args.workload == "trcadv":
        case = cases.config2()
        mesh, bs, q = cases.build(case)
        dm = fedg.DeviceMesh(mesh)
        r = time_trcadv(fedg, mesh, bs, q, dm, args.steps)
        dm.close()
        print(json.dumps({"metric": "ms per call", "workload": "tracer advection of one tracer, 64x64x16 p=7", **r}))
        return
    bits = 32 if args.workload == "heve-fp32" else 64
    h = time_heve(fedg, cases, bits, args.steps, args.warmup)
    case, mesh, bs, q, dm = h["case"], h["mesh"], h["bs"], h["q"], h["dm"]
    wall, ms_kernel = h["wall"], h["ms_kernel"]

    dof = 5 * mesh.Np * mesh.Ne
    nstage = 4
    value = dof * nstage * args.steps / wall
    achieved = BYTES_PER_DOF_STAGE * dof / (ms_kernel * 1e-3) / 1e9
    out = {
        "metric": "DOF·stage/s at p=7 (plain FP64 vs FP32+learned-HEVI, equal err); %HBM peak",
        "value": value,
        "unit": "DOF·stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if bits == 64 else "f32 tendency, f64 state",
        "data": "synthetic",
        "config": {"workload": ("plain FP64" if bits == 64 else "FP32 perturbation-form") +
                   " HEVE warm bubbles, 64x64x16 elements, p=7, RK4 dt=0.04 s",
                   "elements": mesh.Ne, "nodes": mesh.Np * mesh.Ne, "dof": dof, "kernel": "fused stage, "
                   + ("flat-terrain" if h["flat"] else "general-metric")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic()},
    }
    if not args.no_probe:
        # the same loads and stores without the arithmetic: the practical HBM
        # ceiling of this traffic pattern (first, middle x2, last RK4 stage)
        pm = [fedg.stage_probe(md, mesh.NeX, mesh.NeY, mesh.NeZ, 10) for md in (0, 1, 2)]
        probe = (pm[0] + 2 * pm[1] + pm[2]) / 4
        out["access_pattern_probe"] = {
            "ms_per_stage": probe, "ms_first_middle_last": pm, "kernel_ms_per_stage": ms_kernel,
            "kernel_over_probe": ms_kernel / probe,
            "probe_frac_of_peak": BYTES_PER_DOF_STAGE * dof / (probe * 1e-3) / 1e9 / HBM_PEAK_GBS}
    if args.copy_probe:
        t = fedg.TimeIntRK(dm, "ERK_4s4o", case.dt)
        src = fedg.DevField(mesh.Np * mesh.NeA)
        t.store_var0(src, 1)
        out["copy_probe_bytes"] = {"read": 8 * mesh.Np * mesh.Ne, "write": 8 * mesh.Np * mesh.Ne}
        src.free()
        t.close()
    if args.workload == "heve" and not args.no_forms:
        f32 = time_heve(fedg, cases, 32, args.steps, args.warmup)
        f32["dm"].close()
        hevi1, hdm = time_hevi(fedg, cases, 5, 1, 1, 0.0)
        hevin, _ = time_hevi(fedg, cases, 5, 1, 4, 1e-10, dm=hdm)
        hdm.close()
        out["forms"] = {
            "heve_fp32": {"ms_per_step": f32["wall"] * 1e3 / args.steps, "value": dof * 4 * args.steps / f32["wall"],
                          "ms_stage_kernel": f32["ms_kernel"]},
            "hevi_240x16x8_p7": {"one_iteration": hevi1, "newton_1e-10": hevin},
            "hevi_fp32_learned": time_hevi_learned(fedg, cases, 3),
            "hevi_fp32_learned_wide_elements": time_hevi_learned(fedg, cases, 3, dx=5.0e3, dt=0.2),
            "tracer_advection_64x64x16_p7": time_trcadv(fedg, mesh, bs, q, dm, 10),
        }
    if not args.no_cpu_baseline:
        # the host cores this GPU may use: OMP_NUM_THREADS when the environment
        # sets it (the GPU hosts here give each GPU a 16-core share, whatever
        # nproc shows), else the process's affinity mask
        ncore = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = int(os.environ.get("OMP_NUM_THREADS", str(ncore)))
        out["cpu_baseline"] = cpu_baseline(mesh, bs, q, case.dt, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
