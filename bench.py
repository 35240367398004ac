"""Benchmark of the CSRST PIC time step on one MI355X.

Workload (default, c3): 32768 x 4096 window at 5 um, 20 deg bend of radius
10 m, seeded Gaussian bunch gamma 1000, 1 nC, sigma_s = sigma_x = 50 um,
linear energy chirp -10/m, 1e8 macro-particles, second-order shape
function, cell sort every 20 steps.  The timed slice starts with the bunch
half-way round the bend (10 deg into the arc, moving along its tangent, the
bend field acting on every particle).  The FP32 split-precision step is
timed there, then the same run in FP64, and the CSR wake of both runs is
compared; the FP32 step with the bunch 5 mm before the bend (moving along
the beam axis) is timed beside it.

Prints one JSON line (the benchmark record).  A step = one full TE/TM time
step (two pushes with deposits, the field sweeps, the window move).
"""
import argparse
import json
import os
import subprocess
import sys
import time

T_START = time.perf_counter()
ROOT = os.path.dirname(os.path.abspath(__file__))
# the whole default run (set-up, both precisions, both slices, the CPU sample)
# is meant to finish well inside a driver's few-minute window
DRIVER_WINDOW_S = 600.0
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes.  Per particle per push launch: the particle state
# (int32 cell x/y, offset x/y, momentum x/y) read once and written once:
# FP32 2*(8 + 16) = 48 B, FP64 2*(8 + 32) = 80 B.  Per cell per step: k_heh
# reads Ex, Ey, Hz and the class byte and writes Ex', Hz'; k_ey reads Ey, Hz
# and the class byte and writes Ey (currents only inside the bunch footprint):
# FP32 13 + 8 + 9 + 4 = 34 B, FP64 25 + 16 + 17 + 8 = 66 B.
B_PART = {True: 48, False: 80}
B_CELL = {True: 34, False: 66}


def run_mw(cfg, fp32, steps, warmup, R, P, Q, stat):
    from csrst_amd.mw import MovingWindowStep
    m = MovingWindowStep(cfg["W"], cfg["H"], cfg["dx"], cfg["dy"], cfg["lenBefore"], cfg["lenAfter"], cfg["width"],
                         cfg["phiDeg"], cfg["ekinMeV"], cfg["bz"], order=cfg["order"], fp32=fp32,
                         sortEvery=cfg["sortEvery"], statModulus=stat, cfl=cfg["cfl"], pml=cfg["pml"])
    m.init_bunch(R, P, Q)
    if warmup:
        m.step(warmup)
    t0 = time.perf_counter()
    m.step(steps)          # synchronises the library's stream before returning
    t1 = time.perf_counter()
    timing = m.timing()
    wake = m.wake()
    info = m.info()
    m.close()
    return (t1 - t0), timing, wake, info


def _oracle_sample(seconds):
    """The FP64 CPU restatement of the reference step (oracle) on one core, on
    the reference's CPU-runnable workload: global grid of the 10 deg bend at
    ~50 um (4694 x 595), 1e5 macro-particles, IntCIC; as many steps as fit in
    the given seconds.  Returns (particles, cells, steps, seconds)."""
    from oracle import oracle as O
    from csrst_amd.bunch import gaussian_bunch
    Nx, Ny = 4694, 595
    geo = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)
    s = O.Sim(Nx, Ny, order=0, statModulus=100000, **geo)
    n = 100000
    R, P, Q = gaussian_bunch(n, (0.03, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=1)
    s.init(R, P, Q)
    s.step(1)
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        s.step(2)
        steps += 2
    return n, Nx * Ny, steps, time.perf_counter() - t0


def parity_mode_gpu(steps=200, warmup=10):
    """The reference's own two-grid FP64 step (csrst_ctx, bit-identical to the
    oracle) on the GPU, on the workload of the CPU baseline: the 4694 x 595
    global grid of the 10 deg bend plus the bunch window, 1e5 particles,
    IntCIC.  The like-for-like GPU figure beside the CPU one."""
    from csrst_amd import Commands, TwoGridStep
    from csrst_amd.bunch import gaussian_bunch
    Nx, Ny = 4694, 595
    cmd = Commands(numgridpx=Nx, numgridpy=Ny, lengthbefore=0.05, length=0.01, width=0.0128, phi=10.0,
                   Ekin=50.59, Bz=0.1704, order=0, statmodulus=100000)
    n = 100000
    R, P, Q = gaussian_bunch(n, (0.03, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=1)
    g = TwoGridStep(cmd)
    g.init_bunch(R, P, Q)
    g.step(warmup)
    t0 = time.perf_counter()
    g.step(steps)          # synchronises before returning
    t = time.perf_counter() - t0
    inf = g.info()
    g.close()
    return dict(particle_steps_per_s=n * steps / t, cell_updates_per_s=(Nx * Ny + inf["W"] * inf["H"]) * steps / t,
                ms_per_step=1e3 * t / steps, steps=steps,
                workload="two-grid FP64 (parity mode), 4694x595 global grid + %dx%d window, 1e5 particles IntCIC"
                         % (inf["W"], inf["H"]))


UNET_LEVELS = [0, 0, 1, 1, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0]   # resolution level of each 3x3 convolution


def warmstart_bench(H=1024, W=2048, B=4, warmup=2, reps=5):
    """Inference of the BF16-MFMA warm-start U-Net on the 2048 x 1024 charge
    densities of the C4 workload, batch B: FLOPs of the 15 convolutions
    (2 * 9 * cin * cout per output pixel) and wall time per batch including
    the host copies of the input and output."""
    import numpy as np
    from csrst_amd.warmstart import WarmStart, layers
    L, _ = layers()
    per_layer = [2 * 9 * l["cin"] * l["cout"] * (H >> lv) * (W >> lv) * B for l, lv in zip(L, UNET_LEVELS)]
    u = WarmStart(H, W, max_batch=B)
    rho = np.random.default_rng(1).normal(size=(B, H, W)).astype(np.float32)
    for _ in range(warmup):
        u.infer(rho)
    t0 = time.perf_counter()
    for _ in range(reps):
        u.infer(rho)
    t = (time.perf_counter() - t0) / reps
    # training: forward, input gradients and weight gradients (about three
    # times the forward FLOPs) and the Adam update
    phi = np.random.default_rng(2).normal(size=(B, H, W)).astype(np.float32)
    u.train_step(rho, phi, lr=1e-4)
    t0 = time.perf_counter()
    for _ in range(reps):
        u.train_step(rho, phi, lr=1e-4)
    tt = (time.perf_counter() - t0) / reps
    u.close()
    return dict(grid="%dx%d" % (W, H), batch=B, reps=reps, flops_per_batch=float(sum(per_layer)),
                conv_flops_by_layer=per_layer, wall_ms_per_batch=1e3 * t,
                wall_tflops=sum(per_layer) / t / 1e12, train_ms_per_batch=1e3 * tt,
                train_pairs_per_s=B / tt, train_wall_tflops=3 * sum(per_layer) / tt / 1e12)


def warmstart_density(rng, H, W, h):
    """One density of the learned-start workload: 1-4 Gaussians, sigma
    20-500 um, centres in the middle 40 % of each axis, a chirp (the x offset
    scaled by 1 + c (y - y0) / H, |c| <= 0.3), amplitudes summing to one; the
    grid edges are the grounded chamber wall."""
    import numpy as np
    rho = np.zeros((H, W))
    k = int(rng.integers(1, 5))
    amps = rng.random(k)
    amps /= amps.sum()
    for a in amps:
        s = max(rng.uniform(20e-6, 500e-6) / h, 0.5)
        cx, cy = rng.uniform(0.3 * W, 0.7 * W), rng.uniform(0.3 * H, 0.7 * H)
        chirp = rng.uniform(-0.3, 0.3)
        r = int(np.ceil(7 * s * 1.3)) + 2
        x0, x1 = max(int(cx) - r, 1), min(int(cx) + r, W - 1)
        y0, y1 = max(int(cy) - r, 1), min(int(cy) + r, H - 1)
        y, x = np.mgrid[y0:y1, x0:x1]
        xs = (x - cx) * (1 + chirp * (y - cy) / H)
        rho[y0:y1, x0:x1] += a * np.exp(-(xs ** 2 + (y - cy) ** 2) / (2 * s * s)) / (2 * np.pi * s * s)
    return rho


def warmstart_experiment(H=1024, W=2048, h=10e-6, train_seconds=300.0, test=1000, kmax=12, nrhs=2, seed=11,
                         load=None, save=None, oneshot=100):
    """The learned start on 2048 x 1024 densities, network between V-cycles:
    (1) train on (residual, error) pairs of the multigrid's own iterates, the
    first half of the time on plain V-cycle iterates, the second on the
    iterates of the learned loop; (2) on `test` held-out densities solve to a
    relative residual of 1e-10 plainly from zero, with the network's step
    before every V-cycle, with the step over the network's direction and the
    previous change, and over the previous change alone; on the first
    `oneshot` of them also with the network before the first V-cycle only;
    report V-cycles, final residuals and wall time per solve (host copies
    included).  load:
    a weight file (numpy .npy) to start from; save: where to write the
    trained weights."""
    import numpy as np
    from csrst_amd import poisson as PS
    from csrst_amd.warmstart import WarmStart
    cls = np.zeros((H, W), np.uint8)
    cls[1:-1, 1:-1] = PS.UNKNOWN
    u = WarmStart(H, W, max_batch=4)
    if load:
        u.set_weights(np.load(load))
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    losses, pairs, calls = [], 0, 0
    while time.perf_counter() - t0 < train_seconds:
        el = (time.perf_counter() - t0) / train_seconds
        b = np.stack([warmstart_density(rng, H, W, h) for _ in range(nrhs)])
        lr = 2e-4 if el < 0.6 else 1e-4 if el < 0.85 else 4e-5
        loss, n = u.train_iterates(h, h, cls, b, kmax=kmax, inloop=el >= 0.5, lr=lr)
        losses.append(loss)
        pairs += n
        calls += 1
        if calls % 20 == 0:
            print("warmstart: %.0f s, %d pairs, loss %.4f (last 20 calls)" % (
                time.perf_counter() - t0, pairs, float(np.mean(losses[-20:]))), file=sys.stderr, flush=True)
    ttrain = time.perf_counter() - t0
    if save:
        np.save(save, u.weights())
    trng = np.random.default_rng(seed + 1000003)
    res = {m: [] for m in range(5)}
    for i in range(test):
        b = warmstart_density(trng, H, W, h)
        for mode in (0, 1, 3, 4, 2) if i < oneshot else (0, 1, 3, 4):
            t1 = time.perf_counter()
            _, c, r = u.solve(h, h, cls, b, tol=1e-10, maxit=100, mode=mode)
            res[mode].append((c, r, time.perf_counter() - t1))
        if (i + 1) % 50 == 0:
            print("warmstart: %d held-out densities, mean V-cycles: plain %.2f, network %.2f, network + previous "
                  "change %.2f, previous change %.2f" % (i + 1, *[np.mean([x[0] for x in res[m]]) for m in (0, 1, 3, 4)]),
                  file=sys.stderr, flush=True)
    u.close()

    def summary(v):
        c = np.array([x[0] for x in v], float)
        r = np.array([x[1] for x in v], float)
        t = np.array([x[2] for x in v], float)
        return dict(vcycles_mean=float(c.mean()), vcycles_median=float(np.median(c)), vcycles_max=int(c.max()),
                    resid_max=float(r.max()), resid_median=float(np.median(r)), ms_per_solve=1e3 * float(t.mean()))
    out = dict(grid="%dx%d" % (W, H), h_m=h, tol=1e-10, start=load or "He initialisation", train_seconds=ttrain,
               train_calls=calls, train_pairs=pairs, loss_first=float(np.mean(losses[:20])) if losses else None,
               loss_last=float(np.mean(losses[-20:])) if losses else None, test_densities=test,
               plain=summary(res[0]), inloop=summary(res[1]), inloop_prev=summary(res[3]), prev_only=summary(res[4]),
               oneshot=summary(res[2]), oneshot_densities=len(res[2]))
    for k in ("inloop", "inloop_prev", "prev_only", "oneshot"):
        out["vcycle_ratio_" + k] = out["plain"]["vcycles_mean"] / out[k]["vcycles_mean"]
        out["time_ratio_" + k] = out["plain"]["ms_per_solve"] / out[k]["ms_per_solve"]
    return out


def poisson_bench(H=1024, W=2048, h=10e-6, n=20, seed=5):
    """The FP64 multigrid on the c4 grid (2048 x 1024 nodes, grounded chamber)
    and c4 densities: V-cycles to a relative residual of 1e-10 and device time
    per solve and per V-cycle, plain and with the previous-change step
    (csrst_warmstart_solve modes 0 and 4; the hierarchy stays on the device).
    Wall times include the solve's host copies of b and phi."""
    import numpy as np
    from csrst_amd import poisson as PS
    from csrst_amd.warmstart import WarmStart
    cls = np.zeros((H, W), np.uint8)
    cls[1:-1, 1:-1] = PS.UNKNOWN
    u = WarmStart(H, W, max_batch=1)
    rng = np.random.default_rng(seed)
    dens = [warmstart_density(rng, H, W, h) for _ in range(n + 1)]
    out = {}
    for mode, name in ((0, "plain"), (4, "previous_change")):
        u.solve(h, h, cls, dens[0], tol=1e-10, maxit=100, mode=mode)      # warm-up: hierarchy, code
        cyc, t = [], []
        for b in dens[1:]:
            t0 = time.perf_counter()
            _, c, r = u.solve(h, h, cls, b, tol=1e-10, maxit=100, mode=mode)
            t.append(time.perf_counter() - t0)
            cyc.append(c)
        out[name] = dict(vcycles_mean=float(np.mean(cyc)), ms_per_solve=1e3 * float(np.mean(t)),
                         ms_per_vcycle=1e3 * float(np.sum(t)) / float(np.sum(cyc)))
    u.close()
    return dict(grid="%dx%d" % (W, H), h_m=h, densities=n, tol=1e-10, **out)


def _oracle_worker(seconds):
    """Child process of cpu_baseline's all-cores leg: one sample, as JSON."""
    print(json.dumps(_oracle_sample(seconds)), flush=True)


def host_cpu():
    model = "?"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def cpu_baseline(seconds=12.0, max_procs=16):
    """The oracle timed on the host: one core (the reference's single-rank
    semantics), then one independent sample per core in as many child
    processes (at most max_procs, the host share of one GPU), summed."""
    n, cells, steps, t = _oracle_sample(seconds)
    model, ncpu = host_cpu()
    k = max(1, min(ncpu, max_procs))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    procs = [subprocess.Popen([sys.executable, "-c", "import bench; bench._oracle_worker(%r)" % seconds],
                              cwd=ROOT, env=env, stdout=subprocess.PIPE, text=True) for _ in range(k)]
    agg_p = agg_c = 0.0
    for p in procs:
        out, _ = p.communicate(timeout=10 * seconds + 120)
        if p.returncode != 0:
            raise RuntimeError("cpu baseline worker failed")
        pn, pc, ps, pt = json.loads(out.strip().splitlines()[-1])
        agg_p += pn * ps / pt
        agg_c += pc * ps / pt
    return dict(value=n * steps / t, unit="particle-steps/s", cores=1, kind="port",
                cell_updates_per_s=cells * steps / t,
                sample="oracle two-grid step, 4694x595 global grid + bunch window, 1e5 particles IntCIC, "
                       "%d steps in %.1f s on one host core" % (steps, t),
                host=dict(cpu=model, cores_available=ncpu),
                all_cores=dict(value=agg_p, cell_updates_per_s=agg_c, cores=k, processes=k,
                               sample="the same sample in %d independent processes at once, one per core, "
                                      "throughputs summed" % k))


def copy_bandwidth(mib=1024, reps=5):
    """Attainable HBM bandwidth: the best of `reps` device-to-device copies of
    mib MiB (the runtime's copy kernel), timed with HIP events; read plus
    written bytes over the time, GB/s."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = C.c_size_t(mib << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    for q in (src, dst):
        if hip.hipMalloc(C.byref(q), nbytes) != 0:
            raise RuntimeError("copy bandwidth: hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, nbytes) != 0 or hip.hipEventCreate(C.byref(e0)) != 0 or \
                hip.hipEventCreate(C.byref(e1)) != 0:
            raise RuntimeError("copy bandwidth: set-up failed")
        best = float("inf")
        for _ in range(reps + 1):           # the first copy warms up
            hip.hipEventRecord(e0, None)
            hip.hipMemcpyAsync(dst, src, nbytes, 3, None)
            hip.hipEventRecord(e1, None)
            if hip.hipEventSynchronize(e1) != 0:
                raise RuntimeError("copy bandwidth: copy failed")
            ms = C.c_float(0)
            hip.hipEventElapsedTime(C.byref(ms), e0, e1)
            best = min(best, ms.value)
        return 2.0 * (mib << 20) / (best * 1e-3) / 1e9
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
        if e0:
            hip.hipEventDestroy(e0)
        if e1:
            hip.hipEventDestroy(e1)


def copy_probe(mib):
    """One device-to-device hipMemcpy of mib MiB (the runtime's copyBuffer
    kernel): reads and writes mib MiB, a known traffic for the HBM counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = C.c_size_t(mib << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), nbytes) != 0:
            raise RuntimeError("copy probe: hipMalloc failed")
    if hip.hipMemset(src, 1, nbytes) != 0 or hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("copy probe: hipMemset failed")
    if hip.hipMemcpy(dst, src, nbytes, 3) != 0 or hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("copy probe: hipMemcpy failed")
    hip.hipFree(src)
    hip.hipFree(dst)
    print("bench: copy probe %d MiB read + %d MiB written" % (mib, mib), file=sys.stderr, flush=True)


def measured_traffic(workload):
    """HBM bytes per k_push launch from the committed rocprofv3 --pmc summary
    (profiles/k_push_*_traffic.txt) measured on this workload; None if none was."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "k_push_*_traffic.txt"))):
        kv = {}
        with open(path) as f:
            for line in f:
                if ":" in line:
                    k, v = line.split(":", 1)
                    kv[k.strip()] = v.strip()
        if kv.get("workload") == workload and "hbm_bytes_per_launch" in kv:
            return float(kv["hbm_bytes_per_launch"])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--slice", default="bend", choices=("bend", "straight"),
                    help="bend: bunch half-way round the arc (headline); straight: 5 mm before the bend")
    ap.add_argument("--particles", type=int, default=0, help="override the particle count")
    ap.add_argument("--angle", type=float, default=0.0,
                    help="bunch direction in degrees from the beam axis (a bend's later part); not the headline")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-straight", action="store_true", help="skip the straight-section slice")
    ap.add_argument("--stat", type=int, default=20, help="wake sampling period (steps)")
    ap.add_argument("--warmstart", action="store_true",
                    help="only time the warm-start U-Net inference on 2048x1024 densities and print it")
    ap.add_argument("--poisson", action="store_true",
                    help="only time the FP64 multigrid on 2048x1024 c4 densities and print it")
    ap.add_argument("--warmstart-experiment", action="store_true",
                    help="only run the learned-start experiment (network between V-cycles, 2048x1024) and print it")
    ap.add_argument("--ws-train-seconds", type=float, default=300.0)
    ap.add_argument("--ws-test", type=int, default=1000, help="held-out densities of the experiment")
    ap.add_argument("--ws-grid", default="2048x1024", help="W x H of the experiment")
    ap.add_argument("--ws-load", default=None, help="weights (.npy) to start the experiment from")
    ap.add_argument("--ws-save", default=None, help="write the trained weights (.npy) here")
    ap.add_argument("--parity-mode", action="store_true",
                    help="only time the two-grid FP64 step on the CPU baseline's workload and print it")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run one plain device copy of MIB MiB, to calibrate HBM counters")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the PIC step on one GPU only; --gpus %d is not supported" % a.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from csrst_amd import _lib
    from csrst_amd.configs import config, bunch_for, bunch_in_bend
    if not _lib.has_gpu():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if a.copy_probe:
        copy_probe(a.copy_probe)
    if a.parity_mode:
        print(json.dumps(parity_mode_gpu(steps=a.steps, warmup=a.warmup)), flush=True)
        return
    if a.warmstart:
        print(json.dumps(warmstart_bench()), flush=True)
        return
    if a.poisson:
        print(json.dumps(poisson_bench()), flush=True)
        return
    if a.warmstart_experiment:
        gw, gh = (int(v) for v in a.ws_grid.split("x"))
        print(json.dumps(warmstart_experiment(H=gh, W=gw, h=10e-6 * 2048 / gw, train_seconds=a.ws_train_seconds,
                                              test=a.ws_test, load=a.ws_load, save=a.ws_save)), flush=True)
        return
    cfg = config(a.config)
    n = a.particles or cfg["n"]
    print("bench: config %s, %d particles, %dx%d cells, %d+%d steps, slice %s" % (
        a.config, n, cfg["W"], cfg["H"], a.warmup, a.steps, a.slice), file=sys.stderr, flush=True)
    if a.slice == "bend" and not a.angle:
        R, P, Q = bunch_in_bend(cfg, n=n, seed=1, arc_fraction=0.5)
    else:
        R, P, Q = bunch_for(cfg, n=n, seed=1, angle_deg=a.angle)
    t32, tim32, wake32, info32 = run_mw(cfg, True, a.steps, a.warmup, R, P, Q, a.stat)
    print("bench: fp32 %.3f s, timing %s" % (t32, tim32), file=sys.stderr, flush=True)
    cells = cfg["W"] * cfg["H"]
    ps32 = n * a.steps / t32
    cu32 = cells * a.steps / t32
    rec64 = None
    if not a.no_fp64:
        t64, tim64, wake64, info64 = run_mw(cfg, False, a.steps, a.warmup, R, P, Q, a.stat)
        print("bench: fp64 %.3f s, timing %s" % (t64, tim64), file=sys.stderr, flush=True)
        scale = np.abs(wake64).max()
        werr = float(np.abs(wake32 - wake64).max() / scale) if scale > 0 else None
        push64_ms = tim64["push_ms"] / 2.0
        rec64 = dict(particle_steps_per_s=n * a.steps / t64, cell_updates_per_s=cells * a.steps / t64,
                     ms_per_step=1e3 * t64 / a.steps, speedup_fp32_over_fp64=t64 / t32, wake_rel_err=werr,
                     window_moves_match=info64["pos"] == info32["pos"], timing_ms=tim64,
                     push_launch_ms=push64_ms,
                     push_roofline_frac=n * B_PART[False] / (push64_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     step_roofline_frac=(cells * B_CELL[False] + n * B_PART[False]) * a.steps / t64 / 1e9
                     / HBM_PEAK_GBS,
                     step_roofline_frac_96B_cells=(cells * 96 + n * B_PART[False]) * a.steps / t64 / 1e9
                     / HBM_PEAK_GBS)
    other = None
    if a.slice == "bend" and not a.angle and not a.no_straight:
        del R, P, Q
        Rs, Ps, Qs = bunch_for(cfg, n=n, seed=1)
        ts, tims, _, _ = run_mw(cfg, True, a.steps, a.warmup, Rs, Ps, Qs, a.stat)
        del Rs, Ps, Qs
        print("bench: fp32 straight %.3f s, timing %s" % (ts, tims), file=sys.stderr, flush=True)
        other = dict(slice="straight: bunch centred 5 mm before the bend entry, moving along the beam axis",
                     particle_steps_per_s=n * a.steps / ts, cell_updates_per_s=cells * a.steps / ts,
                     ms_per_step=1e3 * ts / a.steps, push_launch_ms=tims["push_ms"] / 2.0,
                     sweep_ms_per_step=tims["sweep_ms"])
    # roofline of the dominant kernel (the particle push + deposit, two launches per step)
    push_launch_ms = tim32["push_ms"] / 2.0
    sweep_ms = tim32["sweep_ms"]
    achieved = n * B_PART[True] / (push_launch_ms * 1e-3) / 1e9
    frac_step = (cells * B_CELL[True] + n * B_PART[True]) * a.steps / t32 / 1e9 / HBM_PEAK_GBS
    frac_step48 = (cells * 48 + n * B_PART[True]) * a.steps / t32 / 1e9 / HBM_PEAK_GBS
    workload = ("%s: %dx%d window dx=%g um, %.0f deg bend R=%g m, gamma %g, %g nC, %d particles, shape S%d, "
                "sort every %d" % (a.config, cfg["W"], cfg["H"], cfg["dx"] * 1e6, cfg["phiDeg"], cfg["R"],
                                   cfg["gamma"], cfg["charge_nC"], n, cfg["order"], cfg["sortEvery"]))
    if a.angle:
        workload += ", bunch at %g deg" % a.angle
    elif a.slice == "bend":
        workload += ", bunch half-way round the bend"
    attainable = copy_bandwidth()
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline()
        cpu["gpu_same_workload"] = parity_mode_gpu()
    rec = {
        "metric": "particle-steps/s + cell-updates/s on 1 MI355X, FP32 vs FP64 at wake err <=1e-4",
        "value": ps32,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t32 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic seeded Gaussian bunch (Philox), analytic bend chamber",
        "config": {"workload": workload, "cells": cells, "particles": n, "pml": cfg["pml"], "cfl": cfg["cfl"],
                   "positions": "int32 cell + f32 offset"},
        "cell_updates_per_s": cu32,
        "step_roofline_frac": frac_step,
        "step_roofline_frac_48B_cells": frac_step48,
        "slice": "bend: bunch centred half-way round the arc, moving along its tangent" if a.slice == "bend"
                 else "straight: bunch centred 5 mm before the bend entry",
        "other_slice": other,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(workload),
                     "kernel": "k_push (push + deposit, %d B per particle per launch)" % B_PART[True],
                     "launch_ms": push_launch_ms, "sweep_ms_per_step": sweep_ms,
                     "attainable_copy_gbs": attainable},
        "fp64": rec64,
        "cpu_baseline": cpu,
    }
    rec["run_wall_s"] = time.perf_counter() - T_START
    rec["fits_in_driver_run"] = rec["run_wall_s"] < DRIVER_WINDOW_S
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
