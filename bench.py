"""Benchmark: KBE-GF2 timesteps/s at nao=64 (one MI355X).

Workload (default): nao=64, naux=256, 128 GLL nodes, dt=0.01, k=5, GF2,
spin-restricted -- the nao=64 configuration -- timed at mid-run, T = 1000 of
nt = 2000, where the per-step cost equals the whole-run average (the Sigma cost
grows linearly with T).  The initial state is the equilibrium GF2 solution
(ned_matsubara_gf2: G^M, Fock matrix, mu).  The history t <= T0 is the
non-interacting solution of that Fock matrix (ned_g0_from_h0) with its
self-energy slice continued time-translation invariantly (ned_tti_extend);
the timed steps T0+W+1 .. T0+W+K then run the full timestep -- predictor,
and `corr` corrector passes of {HF, Sigma, Dyson} -- on the quenched
Hamiltonian.

The same steps are re-run with FP64 Sigma contractions from the same history;
the line reports the mixed-precision rate as `value`, the FP64 rate, their
ratio, and the max-abs difference of every G block written by the steps.

`roofline` is for the dominant kernel, the exchange GEMM of Sigma (at
nao = 64 two launches of kern_nt_bf16x3, the parts of rows (item, plane, aux)
x U' symmetric and antisymmetric in the contracted pair), timed with HIP
events on the library's stream around both launches.  `cpu_baseline` times the CPU oracle (the restated
reference algorithm) on one lesser, one retarded and one mixing time point of
Sigma at the same sizes on one thread, and on a small slice with OpenMP over
time points on all host threads, and scales to a whole timestep.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "KBE timesteps/s at nao=64 GF2 vs plain FP64 at ≤1e-6 G error; % MFMA peak"
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16x3": 2500.0 / 3}
# Coulomb-factor spread used for the nao=64 workload.  B[Q,i,j] ~ N(0, 0.1)
# makes this 64-orbital system strongly correlated (the GF2 equilibrium mu
# lies 18 Hartree below the Hartree-Fock one), and on the 1000-step synthetic
# history below even 0.02 grows the corrector update to O(100); at 0.01 the
# window stays smooth (corrector update ~1e-3, particle number constant to
# 1e-7).  The work per step does not depend on the value.
STABLE_B_SD = {2: 0.01}
# Coulomb-factor spread of the propagated precision check (the whole run from
# the GF2 equilibrium through the quench).  At the configuration's 0.1 the
# equal-time retarded self-energy reaches |Sigma^R(0,0)_ij| = 2.8e3, so
# dt |Sigma| = 28 at dt = 0.01: the FP64 corrector update grows from 3e4 to
# 1e100 in four passes of the first step after the bootstrap (0.07: N leaves
# 32 by 0.24 and the run overflows at step 41).  0.05 propagates.
TRAJ_B_SD = {2: 0.05}
# Coulomb-factor spread of the learned-predictor runs (nao = 48, dt = 0.02),
# for the same reason: the propagated run must stay a solution.
PRED_B_SD = 0.05


def gf2_equilibrium(ned, p, log):
    """Replace the Hartree-Fock initial state of problem p by the equilibrium
    GF2 one (ned_matsubara_gf2): G^M on the GLL nodes, its Fock matrix and mu."""
    t0 = time.perf_counter()
    with ned.Context(p.nao, p.naux, p.k, p.ntau, p.k) as c:
        c.set_coulomb(p.B)
        c.write_blocks(ned.NED_G, ned.NED_MAT, 0, p.Gmat)
        mu, fock, it, err = c.matsubara_gf2(p.h1, p.beta, nelec=float(p.nelec), mu=p.mu, tol=1e-10)
        p.Gmat = c.read_blocks(ned.NED_G, ned.NED_MAT, 0, p.ntau + 1)
    log(f"[setup] GF2 equilibrium: mu {p.mu:.6f} -> {mu:.6f}, {it} iterations, last update {err:.1e} "
        f"({time.perf_counter() - t0:.1f} s)")
    p.fock, p.mu = fock, mu


def ref_sigma_flops(n, a, T, ntau1):
    """Reference decomposition (mol2bSolver_decomp_impl.h), 4/8 flops per
    real/complex MAC: (T+1)(36 n^4 a + 128 n^3 a + 44 n^2 a^2) + Ntau(12 n^4 a + 36 n^3 a + 12 n^2 a^2)."""
    return (T + 1) * (36 * n ** 4 * a + 128 * n ** 3 * a + 44 * n ** 2 * a ** 2) + \
        ntau1 * (12 * n ** 4 * a + 36 * n ** 3 * a + 12 * n ** 2 * a ** 2)


def dyson_flops(n, T, ntau1):
    return 16 * T ** 2 * n ** 3 + 16 * T * ntau1 * n ** 3 + 4 * ntau1 ** 2 * n ** 3


def cpu_baseline(p, T, corr, log):
    """CPU oracle timed on this host: one lesser, one retarded and one mixing
    Sigma time point single-threaded, then one small slice with OpenMP over its
    time points on every host thread; both scaled to a timestep at T."""
    from oracle import oracle as O
    from nedyson_amd.green import GreenFunc
    n, a = p.nao, p.naux
    rng = np.random.default_rng(1)

    def rand_green(nt, ntau):
        g = GreenFunc(nt, ntau, n)
        for arr in (g.ret, g.les, g.tv):
            arr[:] = 0.1 * (rng.uniform(-1, 1, arr.shape) + 1j * rng.uniform(-1, 1, arr.shape))
        return g

    G = rand_green(2, 3)
    S = GreenFunc(2, 3, n)
    times = {}
    for comp, name in ((0, "les"), (1, "ret"), (2, "tv")):
        t0 = time.perf_counter()
        O.sigma_point(p.V, comp, 2, 1, S, G)
        times[name] = time.perf_counter() - t0
        log(f"[cpu] Sigma {name} point, 1 thread: {times[name]:.2f} s")
    les_flops = 12 * n ** 4 * a + 36 * n ** 3 * a + 12 * n ** 2 * a ** 2
    rate = les_flops / times["les"]
    t_sigma = (T + 1) * (times["les"] + times["ret"]) + (p.ntau + 1) * times["tv"]
    # one oracle Dyson step (dyson_step.h:329-382 restated) at the same nao and
    # nodes with a history of Td steps, single-threaded, scaled to T by the
    # step's flop count 16 T^2 n^3 + 16 T Ntau n^3 + 4 Ntau^2 n^3
    Td = int(os.environ.get("NED_CPU_DYSON_T", 8))
    Gd, Sd = rand_green(Td, p.ntau), rand_green(Td, p.ntau)
    for arr in (Sd.ret, Sd.les, Sd.tv):
        arr *= 0.01
    Gd.mat[:] = p.Gmat
    hmf = np.repeat(p.fock[None].astype(complex), Td + 1, axis=0)
    D = O.Dyson(Td, p.ntau, n, p.k)
    t0 = time.perf_counter()
    D.dyson_step(Td, Gd, Sd, hmf, p.mu, p.beta, p.dt)
    t_dyson_Td = time.perf_counter() - t0
    t_dyson = t_dyson_Td * dyson_flops(n, T, p.ntau + 1) / dyson_flops(n, Td, p.ntau + 1)
    log(f"[cpu] Dyson step at T={Td}, 1 thread: {t_dyson_Td:.2f} s; scaled to T={T}: {t_dyson:.1f} s")
    t_step1 = corr * (t_sigma + t_dyson)

    # OpenMP over the time points of a slice with about one point per thread
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    Ts, ntau_s = max(1, threads // 4 - 1), max(1, threads // 2 - 1)
    Gs, Ss = rand_green(Ts, ntau_s), GreenFunc(Ts, ntau_s, n)
    t0 = time.perf_counter()
    O.sigma_solve(p.V, Ts, Ss, Gs, nthreads=threads)
    t_omp = time.perf_counter() - t0
    serial = (Ts + 1) * (times["les"] + times["ret"]) + (ntau_s + 1) * times["tv"]
    speedup = serial / t_omp
    log(f"[cpu] Sigma slice T={Ts}, ntau+1={ntau_s + 1} ({2 * Ts + 2 + ntau_s + 1} points), {threads} threads: "
        f"{t_omp:.2f} s (x{speedup:.2f} over 1 thread)")
    # Sigma spread over the threads as measured; the oracle's Dyson step runs on one
    t_stepN = corr * (t_sigma / speedup + t_dyson)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": 1.0 / t_stepN, "unit": "timesteps/s", "cores": threads, "kind": "port",
            "cpu_model": model, "nproc": os.cpu_count(),
            "sample": f"CPU oracle (reference decomposition restated, C++): one lesser, one retarded and one mixing "
                      f"Sigma time point at nao={n}, naux={a} on 1 thread ({sum(times.values()):.1f} s), then a "
                      f"slice of {2 * Ts + 2 + ntau_s + 1} time points with OpenMP over points on {threads} threads "
                      f"({t_omp:.1f} s, x{speedup:.2f}); one Dyson step at T={Td} on 1 thread ({t_dyson_Td:.1f} s); "
                      f"scaled to a timestep at T={T} with {corr} corrector passes as (T+1)(t_les+t_ret)+Ntau*t_tv "
                      f"(Sigma, over the threads) plus the Dyson step scaled by its flop count (T^2 history terms)",
            "value_1thread": 1.0 / t_step1, "sec_per_step": t_stepN, "gflops_1thread": rate / 1e9,
            "dyson_s_at_T": t_dyson}


def precision_trajectory(ned, p, nsteps, every, corr, log, precs=("f32", "bf16x3"), probe_passes=0):
    """The whole run from the equilibrium GF2 state (ned_matsubara_gf2) through
    the quench, stepped in lockstep in FP64 and in each mixed precision for
    nsteps time steps (bootstrap included, each in its own precision).  After
    every step the new slice T of G (retarded row, lesser column, mixing row)
    is compared with the FP64 run's, and the particle number and total energy
    E_kin + E_pot of each run are recorded; every `every` steps (and at the
    end) a checkpoint is logged.  Returns the summary for the benchmark line."""
    t0 = time.perf_counter()
    ctxs = {}
    with ned.Context(p.nao, p.naux, p.k, p.ntau, p.k) as c:
        c.set_coulomb(p.B)
        try:
            mu, fock, it, err = c.matsubara_gf2(p.h1, p.beta, nelec=float(p.nelec), mu=p.mu, tol=1e-7)
        except ned.NedError as e:
            raise RuntimeError(f"equilibrium GF2 solve failed: {e}")
        Gmat = c.read_blocks(ned.NED_G, ned.NED_MAT, 0, p.ntau + 1)
    log(f"[traj] GF2 equilibrium (sigma_B={p.b_sd}): mu {mu:.6f}, {it} iterations, last update {err:.1e} "
        f"({time.perf_counter() - t0:.1f} s)")
    runs = ("f64",) + tuple(precs)
    boot = {}
    for pr in runs:
        c = ned.Context(p.nao, p.naux, nsteps, p.ntau, p.k)
        c.set_coulomb(p.B)
        c.write_blocks(ned.NED_G, ned.NED_MAT, 0, Gmat)
        c.set_matsubara(p.beta)
        c.g0_from_h0(p.k, fock, mu, p.beta, p.dt)
        try:
            boot[pr] = c.do_boot(fock, mu, p.beta, p.dt, pr, 60, 1e-11)
        except ned.NedError as e:
            if e.code != 5:
                raise
            boot[pr] = (60, float("nan"))
            log(f"[traj] {pr} bootstrap stopped at its iteration limit: {e}")
        ctxs[pr] = c
    n, np1 = p.nao, p.ntau + 1
    if probe_passes:
        # the corrector passes of the first step after the bootstrap, one by
        # one, with their update norms (FP64)
        c, T, passes = ctxs["f64"], p.k + 1, []
        c.extrapolate(T)
        for q in range(probe_passes):
            c.hf(T, p.h0)
            c.sigma(T, "f64")
            try:
                d = c.dyson_step(T, mu, p.beta, p.dt)
            except ned.NedError as ex:
                log(f"[probe] T={T} pass {q + 1}: {ex}")
                passes.append(None)
                break
            passes.append([float(x) for x in d])
            log(f"[probe] T={T} pass {q + 1}: update {d.sum():.3e} (ret {d[0]:.2e} tv {d[1]:.2e} les {d[2]:.2e})")
        for cc in ctxs.values():
            cc.close()
        return {"sigma_B": p.b_sd, "dt": p.dt, "mu": mu, "probe_T": T, "pass_updates": passes}

    def slice_of(c, T):
        s = T * (T + 1) // 2
        return (c.read_blocks(ned.NED_G, ned.NED_RET, s, T + 1), c.read_blocks(ned.NED_G, ned.NED_LES, s, T + 1),
                c.read_blocks(ned.NED_G, ned.NED_TV, T * np1, np1))

    obs = {pr: [] for pr in runs}
    err_by_T = {pr: [] for pr in precs}
    step_ms = {pr: [] for pr in runs}
    iters = {pr: [] for pr in runs}
    checkpoints = []
    run_max = {pr: 0.0 for pr in precs}
    diverged = None
    for T in range(p.k + 1, nsteps + 1):
        sl = {}
        for pr in runs:
            c = ctxs[pr]
            c.sync()
            ts = time.perf_counter()
            try:
                itn, e = c.do_tstp(T, p.h0, mu, p.beta, p.dt, pr, corr, 0.0)
            except ned.NedError as ex:
                diverged = {"prec": pr, "T": T, "error": str(ex)}
                log(f"[traj] {pr} stopped at T={T}: {ex}")
                break
            c.sync()
            step_ms[pr].append(1e3 * (time.perf_counter() - ts))
            iters[pr].append(itn)
            en = c.energy(T, p.h0, p.beta, p.dt)
            obs[pr].append((T, en[2], en[0] + en[1], e))
            sl[pr] = slice_of(c, T)
        if diverged:  # keep the steps that every run completed
            n_ok = min(len(obs[pr]) for pr in runs)
            for pr in runs:
                del obs[pr][n_ok:]
            break
        for pr in precs:
            d = max(float(np.max(np.abs(a - b))) for a, b in zip(sl[pr], sl["f64"]))
            err_by_T[pr].append(d)
            run_max[pr] = max(run_max[pr], d)
        if (T - p.k) % every == 0 or T == nsteps:
            row = {"T": T}
            N0, E0 = obs["f64"][0][1], obs["f64"][0][2]
            row["f64"] = {"N_drift": obs["f64"][-1][1] - N0, "E_drift": obs["f64"][-1][2] - E0,
                          "corr_update": obs["f64"][-1][3], "ms_per_step": step_ms["f64"][-1]}
            for pr in precs:
                row[pr] = {"G_err_at_T": err_by_T[pr][-1], "G_err_max_so_far": run_max[pr],
                           "N_diff_vs_f64": obs[pr][-1][1] - obs["f64"][-1][1],
                           "E_diff_vs_f64": obs[pr][-1][2] - obs["f64"][-1][2],
                           "N_drift": obs[pr][-1][1] - obs[pr][0][1], "E_drift": obs[pr][-1][2] - obs[pr][0][2],
                           "corr_update": obs[pr][-1][3], "ms_per_step": step_ms[pr][-1]}
            checkpoints.append(row)
            log(f"[traj] T={T}: " + "  ".join(
                f"{pr} dG={row[pr]['G_err_at_T']:.2e} (max {row[pr]['G_err_max_so_far']:.2e}) "
                f"dN={row[pr]['N_diff_vs_f64']:.1e} dE={row[pr]['E_diff_vs_f64']:.1e}" for pr in precs) +
                f"  f64 N drift {row['f64']['N_drift']:.1e} E drift {row['f64']['E_drift']:.1e} "
                f"{row['f64']['ms_per_step']:.0f} ms/step")
    for c in ctxs.values():
        c.close()
    if not obs["f64"]:
        raise RuntimeError(f"precision check: no step completed ({diverged})")
    out = {"sigma_B": p.b_sd, "steps": obs["f64"][-1][0], "diverged": diverged, "corrector_passes": corr, "initial": "GF2 equilibrium + quench",
           "mu": mu, "bootstrap_iters": {pr: boot[pr][0] for pr in runs},
           "max_abs_G_err_vs_f64": {pr: run_max[pr] for pr in precs},
           "max_abs_N_diff_vs_f64": {pr: max(abs(a[1] - b[1]) for a, b in zip(obs[pr], obs["f64"])) for pr in precs},
           "max_abs_E_diff_vs_f64": {pr: max(abs(a[2] - b[2]) for a, b in zip(obs[pr], obs["f64"])) for pr in precs},
           "max_abs_N_drift": {pr: max(abs(a[1] - obs[pr][0][1]) for a in obs[pr]) for pr in runs},
           "max_abs_E_drift": {pr: max(abs(a[2] - obs[pr][0][2]) for a in obs[pr]) for pr in runs},
           "ms_per_step_avg": {pr: float(np.mean(step_ms[pr])) for pr in runs},
           "checkpoints": checkpoints}
    log(f"[traj] done in {time.perf_counter() - t0:.1f} s")
    return out


STORAGE = {"whole": 0, "f32": 1, "ring": 2, "f32ring": 3}  # ned_ctx_create_ex flags


def storage_bytes(p, nt, storage):
    """Device bytes of the context's time histories for a storage mode: the
    G^R / G^< triangles (or their ring of k + 3 rows), the Sigma^R / Sigma^<
    triangles in FP64 or FP32, and G^rceil / Sigma^rceil."""
    es, ntri = p.nao * p.nao, (nt + 1) * (nt + 2) // 2
    f = STORAGE[storage]
    g = (min(nt + 1, p.k + 3) * (nt + 1) if f & 2 else ntri) * es * 16
    sg = ntri * es * (8 if f & 1 else 16)
    tv = (nt + 1) * (p.ntau + 1) * es * 16
    return 2 * g + 2 * sg + 2 * tv


def window_context(ned, p, T0, nt, storage="whole"):
    """A context holding the benchmark history t <= T0: the free G of the
    equilibrium Fock matrix with Sigma(T0) continued time-translation
    invariantly."""
    c = ned.Context(p.nao, p.naux, nt, p.ntau, p.k, storage=STORAGE[storage])
    c.set_coulomb(p.B)
    c.write_blocks(ned.NED_G, ned.NED_MAT, 0, p.Gmat)
    c.set_matsubara(p.beta)
    c.g0_from_h0(T0, p.fock, p.mu, p.beta, p.dt)
    c.write_hmf(np.repeat(p.fock[None].astype(complex), T0 + 1, axis=0))
    c.sigma(T0, "f64")
    c.tti_extend(ned.NED_SIGMA, T0)
    c.sync()
    return c


def batch_bench(args, log):
    """The batched configuration: 16 systems (problem.batch_systems), each
    from its GF2 equilibrium, timed on a window of steps after a history of T0
    steps (free G of its Fock matrix, Sigma continued time-translation
    invariantly), FP32 Sigma and FP64 Dyson, 3 corrector passes: the grouped
    batch (one launch per link of the Dyson recurrences for all systems), the
    batch of one host thread and stream per system, and the systems one after
    another."""
    import nedyson_amd as ned
    from nedyson_amd import problem
    T0 = args.T0 if args.T0 is not None else 750
    W, K = args.warmup, args.steps
    nt = T0 + W + K
    # as for the nao=64 window: the synthetic history is no solution of the
    # interacting equations, and at 0.05 the nao=48 system's corrector
    # overflows on it; the work per step does not depend on the values
    bsd = args.b_sd if args.b_sd is not None else 0.01
    ps = problem.batch_systems(b_sd=bsd, nt=nt)
    for p in ps:
        gf2_equilibrium(ned, p, log)
    log(f"[batch] {len(ps)} systems, nao {[p.nao for p in ps]}, history T0={T0}")
    rates = {"grouped": None, "threads": None, "sequential": None}
    for mode in args.batch_modes.split(","):
        ctxs = [window_context(ned, p, T0, nt) for p in ps]
        b = ned.Batch(ctxs, grouped=(mode == "grouped")) if mode != "sequential" else None
        times = []
        for i, T in enumerate(range(T0 + 1, nt + 1)):
            for c in ctxs:
                c.sync()
            t0 = time.perf_counter()
            if b is not None:
                it, err = b.do_tstp(T, [p.h0 for p in ps], [p.mu for p in ps], [p.beta for p in ps],
                                    [p.dt for p in ps], "f32", args.corr)
            else:
                for c, p in zip(ctxs, ps):
                    c.do_tstp(T, p.h0, p.mu, p.beta, p.dt, "f32", args.corr)
            for c in ctxs:
                c.sync()
            if i >= W:
                times.append(time.perf_counter() - t0)
        if b is not None:
            b.close()
        for c in ctxs:
            c.close()
        rates[mode] = len(ps) * len(times) / sum(times)
        log(f"[batch] {mode}: {1e3 * np.mean(times):.1f} ms per batch step, {rates[mode]:.2f} system-steps/s")
    print(json.dumps({"metric": "batched KBE system-steps/s (16 systems of uneven nao)", "value": rates["grouped"],
                      "unit": "system-steps/s", "threads_value": rates["threads"],
                      "sequential_value": rates["sequential"], "steps": K, "warmup": W,
                      "config": {"workload": f"16 systems nao {[p.nao for p in ps]}, naux 4 nao, ntau+1=64, "
                                             f"dt=0.02, T={T0 + W + 1}..{nt}, {args.corr} corrector passes, "
                                             f"FP32 Sigma / FP64 Dyson", "sigma_B": bsd, "T0": T0}}), flush=True)


def predictor_bench(args, log):
    """The learned-predictor configuration (problem.predictor_system: nao=48,
    naux=192, beta=20, 96 GLL nodes, dt=0.02, quench norms 0.05..0.4), at
    reduced counts: `pred_train` training systems and `pred_test` test
    systems (the last ones of the 64), each propagated from its GF2
    equilibrium to step `pred_nt`.  The training runs iterate the FP64
    corrector to 1e-10 and record the polynomial extrapolation's error at the
    predictor points; the network (width 512) trains on the GPU; each test
    system then runs three times from the same equilibrium, the corrector
    iterated to 1e-8 -- polynomial guess, learned guess with FP32 inference,
    learned guess with BF16 inference -- and the corrector passes and wall
    time over the steps where the network acts are compared."""
    import nedyson_amd as ned
    from nedyson_amd import predictor, problem
    nt, bsd = args.pred_nt, args.b_sd if args.b_sd is not None else PRED_B_SD
    t_all = time.perf_counter()

    def system(i):
        p = problem.predictor_system(i, nt=nt, b_sd=bsd)
        gf2_equilibrium(ned, p, log)
        return p

    def booted(p):
        c = ned.Context(p.nao, p.naux, nt, p.ntau, p.k)
        c.set_coulomb(p.B)
        c.write_blocks(ned.NED_G, ned.NED_MAT, 0, p.Gmat)
        c.set_matsubara(p.beta)
        c.g0_from_h0(p.k, p.fock, p.mu, p.beta, p.dt)
        c.do_boot(p.fock, p.mu, p.beta, p.dt, "f64", 60, 1e-10)
        return c

    X, Y = [], []
    for i in range(args.pred_train):
        p = system(i)
        t0 = time.perf_counter()
        with booted(p) as c:
            r = []
            for T in range(p.k + 1, nt + 1):
                c.do_tstp(T, p.h0, p.mu, p.beta, p.dt, "f64", 20, 1e-10)
                if T >= p.k + 33:
                    r.append(c.predictor_residual(T))
        x, y, _ = predictor.dataset(r)
        X.append(x)
        Y.append(y)
        log(f"[pred] training system {i}: quench norm {p.quench_norm:.3f}, {len(r)} steps recorded "
            f"({time.perf_counter() - t0:.1f} s)")
    X, Y = predictor.subsample(np.concatenate(X), np.concatenate(Y), args.pred_rows)
    t0 = time.perf_counter()
    net, losses = predictor.train(X, Y, width=512, epochs=args.pred_epochs, batch=256, lr=1e-3, seed=0)
    t_train = time.perf_counter() - t0
    log(f"[pred] trained on {X.shape[0]} rows: loss {losses[0]:.3e} -> {losses[-1]:.3e} ({t_train:.1f} s)")
    modes = {"polynomial": None, "learned_fp32": False, "learned_bf16": True}
    tot = {m: [0, 0.0, 0] for m in modes}  # passes, seconds, fallbacks
    t1 = None
    for j in range(args.pred_test):
        i = problem.PREDICTOR_NSYS - args.pred_test + j
        p = system(i)
        t1 = predictor.first_learned_step(p.k)
        for m, bf16 in modes.items():
            with booted(p) as c:
                if bf16 is not None:
                    net.set_bf16(bf16)
                    c.set_predictor(net)
                steps = [s for s in predictor.run_to_tolerance(c, p, p.k + 1, nt) if s[0] >= t1]
                fb = c.predictor_stats()[1] if bf16 is not None else 0
                c.set_predictor(None)
            n_pass, secs = sum(s[1] for s in steps), sum(s[2] for s in steps)
            tot[m][0] += n_pass
            tot[m][1] += secs
            tot[m][2] += fb
            log(f"[pred] test system {i} ({m}): {n_pass} passes over {len(steps)} steps, {secs:.2f} s, "
                f"{fb} fallbacks; per step {[s[1] for s in steps]}")
    nsteps = args.pred_test * (nt - t1 + 1)
    out = {m: {"passes_per_step": v[0] / nsteps, "ms_per_step": 1e3 * v[1] / nsteps, "fallbacks": v[2]}
           for m, v in tot.items()}
    print(json.dumps({"metric": "corrector passes and time to residual 1e-8, learned vs polynomial guess",
                      "value": out["polynomial"]["ms_per_step"] / out["learned_bf16"]["ms_per_step"],
                      "unit": "x (time to 1e-8, polynomial / learned BF16)", "modes": out,
                      "train_seconds": t_train, "total_seconds": time.perf_counter() - t_all,
                      "config": {"workload": f"nao=48 naux=192 beta=20 ntau+1=96 dt=0.02, {args.pred_train} "
                                             f"training and {args.pred_test} test systems, steps {t1}..{nt} "
                                             f"compared, FP64 Sigma, corrector to 1e-8",
                                 "sigma_B": bsd, "rows": int(X.shape[0]), "epochs": args.pred_epochs}}),
          flush=True)


def sweep(args, p, log):
    """Per-step time at several history lengths T, and the whole-run average
    over T = k+1..nt from a quadratic fit t(T) = a + b T + c T^2 (the Sigma
    cost grows linearly and the Dyson cost quadratically in T)."""
    import nedyson_amd as ned
    Ts = [int(x) for x in args.sweep.split(",")]
    rows = []
    for T0 in Ts:
        W, K = args.warmup, args.steps
        c = window_context(ned, p, T0, T0 + W + K, args.storage)
        times, split = [], []
        for i, T in enumerate(range(T0 + 1, T0 + W + K + 1)):
            if i >= W:
                c.profile(True)
            c.sync()
            t0 = time.perf_counter()
            c.do_tstp(T, p.h0, p.mu, p.beta, p.dt, args.prec, args.corr, 0.0)
            c.sync()
            if i >= W:
                times.append(time.perf_counter() - t0)
            c.profile(False)
        pr = c.profile_read()
        c.close()
        ms = 1e3 * float(np.mean(times))
        gb = storage_bytes(p, T0 + W + K, args.storage) / 1e9
        rows.append({"T": T0 + W + 1, "ms_per_step": ms, "sigma_ms_per_eval": pr["sigma_ms"] / max(pr["sigma_n"], 1),
                     "dyson_ms_per_step": pr["dyson_ms"] / max(pr["dyson_n"], 1), "history_GB": gb})
        log(f"[sweep] T={T0 + W + 1}: {ms:.1f} ms per step ({args.storage} storage, histories {gb:.1f} GB)")
    out = {"metric": "per-step time vs history length", "prec": args.prec, "nao": p.nao, "naux": p.naux,
           "nt_run": p.nt, "storage": args.storage, "points": rows}
    if len(rows) >= 3:
        T = np.array([r["T"] for r in rows], dtype=float)
        y = np.array([r["ms_per_step"] for r in rows])
        coef = np.polyfit(T, y, 2)
        steps = np.arange(p.k + 1, p.nt + 1, dtype=float)
        avg = float(np.mean(np.polyval(coef, steps)))
        out["fit_ms"] = {"c2": coef[0], "c1": coef[1], "c0": coef[2]}
        out["whole_run_avg_ms_per_step"] = avg
        out["whole_run_timesteps_per_s"] = 1e3 / avg
        out["extrapolated_ms_per_step"] = {str(int(t)): float(np.polyval(coef, t)) for t in (1500, 2000)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--T0", type=int, default=None, help="last history step (default nt/2)")
    ap.add_argument("--prec", default="bf16x3", choices=["f32", "bf16x3", "f64"],
                    help="Sigma arithmetic of the timed run (f64: for --sweep)")
    ap.add_argument("--corr", type=int, default=3, help="corrector passes per step")
    ap.add_argument("--f64-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--b-sd", type=float, default=None,
                    help="sigma_B of the Coulomb factor (default: %s for the nao=64 workload)" % STABLE_B_SD)
    ap.add_argument("--dt", type=float, default=None)
    ap.add_argument("--initial", default="gf2", choices=["gf2", "hf"],
                    help="initial state: equilibrium GF2 (default) or Hartree-Fock")
    ap.add_argument("--sweep", default=None,
                    help="comma-separated history lengths T0: print per-step times at each and the fitted "
                         "whole-run average instead of the benchmark line")
    ap.add_argument("--batch-modes", default="grouped,threads,sequential",
                    help="batched configuration (--config 3): the modes to time, comma-separated")
    ap.add_argument("--storage", default="whole", choices=sorted(STORAGE),
                    help="--sweep: storage of the histories (FP32 Sigma triangles, ring of G rows, or both)")
    ap.add_argument("--pred-train", type=int, default=6, help="learned predictor (--config 4): training systems")
    ap.add_argument("--pred-test", type=int, default=2, help="learned predictor: test systems")
    ap.add_argument("--pred-nt", type=int, default=100, help="learned predictor: steps per system")
    ap.add_argument("--pred-epochs", type=int, default=40, help="learned predictor: training epochs")
    ap.add_argument("--pred-rows", type=int, default=200000, help="learned predictor: training rows drawn")
    ap.add_argument("--traj-steps", type=int, default=100,
                    help="also run the precision check: the configuration's own system (its sigma_B) from the GF2 "
                         "equilibrium through the quench for this many steps in FP64, FP32 and bf16x3 in lockstep")
    ap.add_argument("--traj-every", type=int, default=50, help="checkpoint interval of the precision check")
    ap.add_argument("--traj-b-sd", type=float, default=None, help="sigma_B of the precision check's system")
    ap.add_argument("--traj-dt", type=float, default=None, help="time step of the precision check's run")
    ap.add_argument("--traj-quench-norm", type=float, default=None,
                    help="quench norm of the precision check's run (0: the equilibrium propagated unperturbed)")
    ap.add_argument("--traj-f64-only", action="store_true", help="precision check: the FP64 run alone")
    ap.add_argument("--probe-passes", type=int, default=0,
                    help="instead of the precision check, log the FP64 corrector update of each of this many "
                         "passes of the first step after the bootstrap")
    ap.add_argument("--traj-only", action="store_true", help="print only the precision check's JSON line")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain device-to-device copy of this many bytes (the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the timestep runs on one GPU; --gpus {args.gpus} is not supported "
              f"(sharding the history over GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.config == 3:
        batch_bench(args, log)
        return
    if args.config == 4:
        predictor_bench(args, log)
        return

    if args.copy_bytes:
        import torch
        n = args.copy_bytes // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        log(f"[copy] {4 * n} bytes copied")
        return

    import nedyson_amd as ned
    from nedyson_amd import problem

    traj = None
    if args.traj_steps > 0:
        ts = time.perf_counter()
        bsd = args.traj_b_sd if args.traj_b_sd is not None else TRAJ_B_SD.get(args.config)
        over = {"b_sd": bsd} if bsd is not None else {}
        if args.traj_dt is not None:
            over["dt"] = args.traj_dt
        if args.traj_quench_norm is not None:
            over["quench_norm"] = args.traj_quench_norm
        pt = problem.config(args.config, nt=args.traj_steps, **over)
        log(f"[traj] problem c{args.config}: nao={pt.nao} naux={pt.naux} sigma_B={pt.b_sd} "
            f"({time.perf_counter() - ts:.1f} s)")
        traj = precision_trajectory(ned, pt, args.traj_steps, args.traj_every, args.corr, log,
                                    precs=() if args.traj_f64_only else ("f32", "bf16x3"),
                                    probe_passes=args.probe_passes)
        if args.traj_only:
            print(json.dumps({"metric": "max-abs G error vs FP64 along a propagated run", "config": args.config,
                              "nao": pt.nao, "naux": pt.naux, "nL": pt.ntau + 1, "dt": pt.dt, **traj}), flush=True)
            return

    t_setup = time.perf_counter()
    over = {}
    if args.b_sd is not None:
        over["b_sd"] = args.b_sd
    elif args.config in STABLE_B_SD:
        over["b_sd"] = STABLE_B_SD[args.config]
    if args.dt is not None:
        over["dt"] = args.dt
    p = problem.config(args.config, **over)
    log(f"[setup] problem c{args.config}: nao={p.nao} naux={p.naux} ntau+1={p.ntau + 1} mu={p.mu:.6f} "
        f"({time.perf_counter() - t_setup:.1f} s)")
    if args.initial == "gf2":
        gf2_equilibrium(ned, p, log)
    if args.sweep:
        sweep(args, p, log)
        return
    T0 = p.nt // 2 if args.T0 is None else args.T0
    W, K = args.warmup, args.steps
    nt = T0 + W + K
    n, a = p.nao, p.naux
    c = window_context(ned, p, T0, nt)
    log(f"[setup] history t<=T0={T0} ready ({time.perf_counter() - t_setup:.1f} s)")

    def run(prec, steps, profile):
        t_list, iters, obs = [], [], []
        for i, T in enumerate(range(T0 + 1, nt + 1)):
            timed = i >= W
            if timed and profile:
                c.profile(True)
            c.sync()
            t0 = time.perf_counter()
            it, err = c.do_tstp(T, p.h0, p.mu, p.beta, p.dt, prec, args.corr, 0.0)
            c.sync()
            dt_ = time.perf_counter() - t0
            c.profile(False)
            e = c.energy(T, p.h0, p.beta, p.dt)  # outside the timed region
            log(f"[{prec}] step T={T} iters={it} err={err:.3e} {dt_ * 1e3:.1f} ms  N={e[2]:.10f} E={e[0] + e[1]:.10f}")
            if timed:
                t_list.append(dt_)
                iters.append(it)
                obs.append((e[2], e[0] + e[1]))
        prof = c.profile_read() if profile else None
        return t_list, iters, prof, np.array(obs)

    def slices():
        out = []
        for T in range(T0 + 1, nt + 1):
            s = T * (T + 1) // 2
            out.append(c.read_blocks(ned.NED_G, ned.NED_RET, s, T + 1))
            out.append(c.read_blocks(ned.NED_G, ned.NED_LES, s, T + 1))
            out.append(c.read_blocks(ned.NED_G, ned.NED_TV, T * (p.ntau + 1), p.ntau + 1))
        return out

    t_mixed, iters, prof, obs_mixed = run(args.prec, K, True)
    g_mixed = slices()
    elapsed = sum(t_mixed)
    value = K / elapsed
    items = 2 * (T0 + W + 1) + p.ntau + 1  # items of a slice near the timed steps
    f64_value, max_err, dN, dE, iters_f64 = None, None, None, None, None
    if not args.no_f64:
        t_f64, iters_f64, _, obs_f64 = run("f64", K, False)
        g_f64 = slices()
        f64_value = K / sum(t_f64)
        max_err = float(max(np.max(np.abs(x - y)) for x, y in zip(g_mixed, g_f64)))
        dN = float(np.max(np.abs(obs_mixed[:, 0] - obs_f64[:, 0])))
        dE = float(np.max(np.abs(obs_mixed[:, 1] - obs_f64[:, 1])))
    c.close()

    ugemm_ms = prof["ugemm_ms"] / max(prof["ugemm_n"], 1)
    # HBM bytes per exchange (both GEMM launches) from the rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs recorded next to this file (profiles/), when they were
    # taken on this workload's shape
    traffic, traffic_note = None, "no counter run recorded for this shape"
    tfile = os.path.join(ROOT, "profiles", "kern_nt_bf16x3_nao64_T1000_traffic.json")
    # (the counter runs used the default chunk of 464 items, M = 237568 rows;
    # the library reads NED_SIGMA_CHUNK with atoi, so a non-number means default)
    try:
        forced_chunk = int(os.environ.get("NED_SIGMA_CHUNK", "0").strip() or 0)
    except ValueError:
        forced_chunk = 0
    if (args.prec == "bf16x3" and os.path.exists(tfile) and p.nao == 64 and p.naux == 256
            and forced_chunk <= 0):
        t = json.load(open(tfile))
        traffic = t["traffic_bytes_per_launch"]
        traffic_note = (f"bytes per full-chunk exchange ({t['shape']}) from rocprofv3 "
                        f"FETCH_SIZE x {t['fetch_scale']:.3f} + WRITE_SIZE x {t['write_scale']:.3f}, scaled by a 1 GiB "
                        f"copy; algorithmic {t['algorithmic_bytes_per_launch']:.3e} B")
    n_sigma = max(prof["sigma_n"], 1)
    # the library splits the exchange by the pair symmetry of U' (NED_SYMX=0
    # keeps the full contraction, as atoi reads it): the split-bf16 GEMMs do so
    # at nao = 64 only and contract all pairs at the other shapes they take
    # (nao^2 and 2 naux multiples of 128, naux of 64); every other exchange
    # splits from nao = 8 on
    symx_env = os.environ.get("NED_SYMX")
    try:
        symx_off = symx_env is not None and int(symx_env.strip() or 0) == 0
    except ValueError:
        symx_off = True
    uaug = (p.nao * p.nao) % 128 == 0 and (2 * p.naux) % 128 == 0 and p.naux % 64 == 0
    if args.prec == "bf16x3" and uaug:
        split_exchange = p.nao == 64 and not symx_off
    else:
        split_exchange = p.nao >= 8 and not symx_off
    f64_split = p.nao >= 8 and not symx_off
    if split_exchange == f64_split:
        speedup_basis = ("FP64 run uses the same exchange form ("
                         + ("split by pair symmetry" if f64_split else "full contraction")
                         + "; precision-only ratio)")
    else:
        speedup_basis = ("FP64 run splits the exchange by pair symmetry, the mixed run contracts all pairs"
                         if f64_split else "mixed run splits the exchange by pair symmetry, the FP64 run contracts all pairs")
    # algorithmic flops of one exchange with M = 2 naux P rows (P time points
    # of the slice): 2 M (ns (ns + naux) + na^2) over the ns = nao(nao+1)/2
    # and na = nao(nao-1)/2 pairs where the exchange is split (padding not
    # counted), 2 M n^2 (n^2 + naux) for the split-bf16 full contraction (Pi
    # folded in), 2 M n^4 for the other full contraction; averaged over
    # exchanges
    flops_per_launch = prof["ugemm_flops"] / max(prof["ugemm_n"], 1)
    achieved = flops_per_launch / (ugemm_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.prec]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.prec} Sigma contractions, f64 Dyson",
        "data": f"synthetic (seeded h1, B[Q,i,j] ~ N(0, {p.b_sd}); {args.initial.upper()} equilibrium; timed window: "
                "history = free G of its Fock matrix; precision_run: whole run from the GF2 equilibrium of the "
                "same generators at its own sigma_B)",
        "config": {"workload": f"GF2 KBE timestep at T={T0 + W + 1}..{nt}, nao={n}, naux={a}, "
                               f"ntau+1={p.ntau + 1}, dt={p.dt}, k={p.k}, {args.corr} corrector passes",
                   "nao": n, "naux": a, "nL": p.ntau + 1, "T": T0 + W + 1, "nt_run": p.nt,
                   "corrector_passes": args.corr, "sigma_prec": args.prec, "window_sigma_B": p.b_sd,
                   "precision_run_sigma_B": traj["sigma_B"] if traj else None,
                   "precision_run_steps": traj["steps"] if traj else None,
                   "speedup_basis": speedup_basis},
        "f64_value": f64_value,
        "speedup_vs_f64": (value / f64_value) if f64_value else None,
        "max_abs_G_err_vs_f64": max_err,
        "max_abs_N_diff_vs_f64": dN,
        "max_abs_E_diff_vs_f64": dE,
        "corrector_iters_per_step": iters,
        "corrector_iters_per_step_f64": iters_f64,
        "sigma_ms_per_eval": prof["sigma_ms"] / n_sigma,
        "dyson_ms_per_step": prof["dyson_ms"] / max(prof["dyson_n"], 1),
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "traffic_note": traffic_note,
                     "kernel": f"Sigma exchange GEMM ({args.prec}), {flops_per_launch:.3e} algorithmic flop "
                               f"per exchange, avg {ugemm_ms:.3f} ms over {prof['ugemm_n']} exchanges ("
                               + ("each the symmetric and antisymmetric launch)"
                                  if split_exchange else "each one full-contraction launch)")},
    }
    if traj is not None:
        line["precision_run"] = traj
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(p, T0 + W + 1, args.corr, log)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
