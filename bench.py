"""SOAP-GAP E+F+virial throughput on one MI355X.

Workload: si4096 -- an 8x8x8 diamond Si supercell (a = 5.43 A), Gaussian
displacements sigma = 0.15 A, 1% random vacancies (4055 atoms); SOAP cutoff
5.0, cutoff_transition_width 1.0, atom_sigma 0.5, n_max 8, l_max 6 (d = 253),
2000 sparse points, zeta 4, delta 1, alpha ~ N(0, 1e-3).  A step is one full
evaluation of energy, forces and virial (neighbour list included) with the
positions already resident in HBM, through gapx_calc_device.

Reported beside the FP64 plain form: the bf16x3 and fp16x3 split-precision
forms (throughput and error against FP64), a per-stage breakdown from HIP
events on the library's stream, the roofline of the dominant kernel, the HBM
GB/s and FP64-VALU fraction of each descriptor stage (per-launch bytes and flops
from the committed counter profile over this run's stage times), the batching
form on batch512 and on si4096x8 (eight si4096-sized structures), the mixed/FP64
ratio on each workload measured in both forms, and the oracle's FP64 CPU
throughput on this host's cores (si4096, with bounded samples of si64, sic32768
and batch512, and the host's lscpu figures) -- a reported baseline.
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from quip_amd import _lib as L  # noqa: E402
from quip_amd import configs  # noqa: E402
from quip_amd.gap_xml import GAPModel, GPCoordinate, write_gap_xml  # noqa: E402
from quip_amd.potential import Potential  # noqa: E402

MFMA_FP64_PEAK = 78.6e12
MFMA_BF16_PEAK = 2.5e15
VALU_FP64_PEAK = 78.6e12
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_descriptor_fn(tmpdir):
    cache = {}

    def fn(d, at):
        key = d.string
        if key not in cache:
            m = GAPModel("D", 1527075646, np.zeros(119),
                         [GPCoordinate("D1", d, 2, 1.0, 0.0, 1.0, np.zeros((1, d.dim)), np.zeros(1), np.ones(1))])
            path = os.path.join(tmpdir, "desc%d.xml" % len(cache))
            write_gap_xml(path, m)
            cache[key] = Potential(param_filename=path)
        return cache[key].descriptors(at)[0]
    return fn


class DeviceArrays:
    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        if self.hip.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 8))) != 0:
            raise RuntimeError("hipMalloc failed")
        self.ptrs.append(p)
        return p

    def h2d(self, dst, arr):
        if self.hip.hipMemcpy(dst, L.ptr(arr), C.c_size_t(arr.nbytes), 1) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    def d2h(self, arr, src):
        if self.hip.hipMemcpy(L.ptr(arr), src, C.c_size_t(arr.nbytes), 2) != 0:
            raise RuntimeError("hipMemcpy D2H failed")

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _stage_pass(pot, step, steps):
    """Stage times (ms per call) from HIP events on the library's stream over `steps`
    more calls, after the timed ones: the events sit between the stages in the stream,
    so the timed calls run without them."""
    pot.set_timing(True)
    for _ in range(steps):
        step()
    tot, n = pot.timing_totals()
    pot.set_timing(False)
    return {k: tot[k] / max(n, 1) for k in tot}


def run_form(pot, at, dev, bufs, steps, warmup, graphs=False):
    """FP64/split plain form on one structure.  The timed steps run as a caller runs
    them (no stage events); eager runs then take the stage times from a second pass
    of as many steps.  With graphs the launch sequence, neighbour list included, is
    replayed as a HIP graph and no stage times are taken."""
    lib = L.lib()
    N = len(at)
    Z, lat, pbc = L.i32(at.numbers), L.f64(at.cell), L.i32(at.pbc)
    want = L.WANT_E | L.WANT_F | L.WANT_V

    def step():
        L.check(lib.gapx_calc_device(pot.handle, N, bufs["pos"], L.ptr(Z), L.ptr(lat), L.ptr(pbc), 0.0, want,
                                     bufs["e"], None, bufs["f"], bufs["v"], None), "gapx_calc_device")

    pot.set_graphs(graphs)
    for _ in range(warmup):
        step()
    L.check(lib.gapx_sync(pot.handle), "sync")
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    L.check(lib.gapx_sync(pot.handle), "sync")
    dt = time.perf_counter() - t0
    e, f, v = np.zeros(1), np.zeros((N, 3)), np.zeros(9)
    dev.d2h(e, bufs["e"])
    dev.d2h(f, bufs["f"])
    dev.d2h(v, bufs["v"])
    stages = _stage_pass(pot, step, steps) if not graphs else {}
    return dt, dict(energy=e[0], force=f, virial=v.reshape(3, 3).T), stages


def run_batch(pot, frames, dev, steps, warmup, graphs=False):
    """Batching form: every structure of `frames` in one pass (gapx_calc_batch_device)."""
    lib = L.lib()
    S = len(frames)
    off = np.zeros(S + 1, np.int32)
    off[1:] = np.cumsum([len(a) for a in frames])
    N = int(off[-1])
    Z = L.i32(np.concatenate([a.numbers for a in frames]))
    lats = L.f64(np.stack([a.cell for a in frames]))
    pbcs = L.i32(np.stack([a.pbc for a in frames]))
    pos = L.f64(np.vstack([a.positions for a in frames]))
    b = dict(pos=dev.alloc(N * 24), e=dev.alloc(S * 8), f=dev.alloc(N * 24), v=dev.alloc(S * 72))
    dev.h2d(b["pos"], pos)
    want = L.WANT_E | L.WANT_F | L.WANT_V

    def step():
        L.check(lib.gapx_calc_batch_device(pot.handle, S, L.ptr(off), b["pos"], L.ptr(Z), L.ptr(lats), L.ptr(pbcs),
                                           0.0, want, b["e"], None, b["f"], b["v"], None), "gapx_calc_batch_device")

    pot.set_graphs(graphs)
    for _ in range(warmup):
        step()
    L.check(lib.gapx_sync(pot.handle), "sync")
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    L.check(lib.gapx_sync(pot.handle), "sync")
    dt = time.perf_counter() - t0
    E, F, V = np.zeros(S), np.zeros((N, 3)), np.zeros(S * 9)
    dev.d2h(E, b["e"])
    dev.d2h(F, b["f"])
    dev.d2h(V, b["v"])
    stages = _stage_pass(pot, step, steps) if not graphs else {}
    return dt, N, dict(energy=E, force=F, natoms=np.diff(off)), stages


def host_cpu():
    """lscpu model / sockets / cores / threads and the cores this process may use."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    info["nproc"] = len(os.sched_getaffinity(0))
    return info


def _time_oracle(fn, budget_s, max_reps=20):
    t0 = time.perf_counter()
    r = fn()                                           # warm-up
    first = time.perf_counter() - t0
    reps = max(1, min(max_reps, int(budget_s / max(first, 1e-3))))
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), reps, r


def cpu_baseline(model, at, budget_s=12.0):
    """The oracle's FP64 E+F+virial (OpenMP over centres) on the si4096 workload."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    med, reps, r = _time_oracle(lambda: O.gap_calc(model, at, n_threads=threads), budget_s)
    return dict(value=len(at) / med, unit="atoms/s", cores=threads, kind="port",
                sample="si4096 (4055 atoms) full E+F+virial, median of %d runs after 1 warm-up, "
                       "oracle/liboracle.so FP64 with OpenMP" % reps), r


def cpu_baseline_configs(si64_model, si4096_model, sic_model, frames, threads, sic_centres=128,
                         batch_structs=64):
    """Bounded samples of the other workloads on the same host cores: si64 whole (model with
    500 sparse points and a distance_2b term), sic32768 with the first `sic_centres` atoms as
    the only centres (atom_mask: the per-centre work of the full structure), and the first
    `batch_structs` structures of batch512 (the si4096 model)."""
    from oracle import oracle as O
    out = {}
    at0 = configs.si64()
    med, reps, _ = _time_oracle(lambda: O.gap_calc(si64_model, at0, n_threads=threads), 1.5, 40)
    out["si64"] = dict(atoms_per_s=len(at0) / med, sample="si64 (64 atoms) whole, median of %d runs" % reps)
    if sic_model is not None:
        sat = configs.sic32768()
        mask = np.zeros(len(sat), bool)
        mask[:sic_centres] = True
        med, reps, _ = _time_oracle(lambda: O.gap_calc_ex(sic_model, sat, atom_mask=mask, n_threads=threads),
                                    4.0, 3)
        out["sic32768"] = dict(atoms_per_s=sic_centres / med,
                               sample="sic32768 (32768 atoms), the first %d atoms as centres (atom_mask), median "
                                      "of %d runs" % (sic_centres, reps))
    sub = frames[:batch_structs]
    n = sum(len(a) for a in sub)

    def run():
        for a in sub:
            O.gap_calc(si4096_model, a, n_threads=threads)
    med, reps, _ = _time_oracle(run, 3.0, 3)
    out["batch512"] = dict(atoms_per_s=n / med, sample="the first %d structures of batch512 (%d atoms) one after "
                                                        "another, median of %d runs" % (len(sub), n, reps))
    return out


def run_serving(path, prec, n_pool, duration, fracs, n_slots, max_batch_atoms):
    """Serving form: saturation throughput (the whole pool submitted at once), then
    Poisson arrivals at fractions of it; latency is submission -> completion."""
    import threading
    from quip_amd.serve import Server
    pool = configs.sic_requests(n_pool)
    pot = Potential(param_filename=path, precision=prec)
    out = {}
    with Server(pot, n_slots=n_slots, hbm_budget=96 << 30, max_batch_atoms=max_batch_atoms) as srv:
        for _ in range(2):   # warm-up: workspaces grow to their working size
            for t in [srv.submit(a) for a in pool]:
                srv.wait(t)
        t0 = time.perf_counter()
        tickets = [srv.submit(a) for a in pool]
        lat = [srv.wait(t)["latency_s"] for t in tickets]
        dt = time.perf_counter() - t0
        atoms = sum(len(a) for a in pool)
        sat = len(pool) / dt
        out["saturation"] = dict(requests_per_s=sat, atoms_per_s=atoms / dt, requests=len(pool),
                                 mean_atoms=atoms / len(pool))
        rng = np.random.default_rng(44)
        for frac in fracs:
            rate = frac * sat
            n = max(50, int(rate * duration))
            gaps = rng.exponential(1.0 / rate, n)
            picks = rng.integers(0, len(pool), n)
            tick = [None] * n
            lats = [0.0] * n
            ready = threading.Semaphore(0)

            def waiter():
                for k in range(n):
                    ready.acquire()
                    lats[k] = srv.wait(tick[k])["latency_s"]

            th = threading.Thread(target=waiter)
            th.start()
            st0 = srv.stats()
            t0 = time.perf_counter()
            t_next = t0
            for k in range(n):
                t_next += gaps[k]
                while True:
                    now = time.perf_counter()
                    if now >= t_next:
                        break
                    time.sleep(min(t_next - now, 5e-4))
                tick[k] = srv.submit(pool[picks[k]])
                ready.release()
            th.join()
            dt = time.perf_counter() - t0
            st1 = srv.stats()
            la = np.array(lats) * 1e3
            out["rate_%.2f" % frac] = dict(
                offered_requests_per_s=rate, requests=n, atoms_per_s=(st1["atoms"] - st0["atoms"]) / dt,
                p50_ms=float(np.percentile(la, 50)), p99_ms=float(np.percentile(la, 99)),
                mean_batch=(st1["requests"] - st0["requests"]) / max(1, st1["batches"] - st0["batches"]))
            log("serving %s rate %.0f/s: p50 %.2f ms p99 %.2f ms" % (prec, rate, out["rate_%.2f" % frac]["p50_ms"],
                                                                    out["rate_%.2f" % frac]["p99_ms"]))
    return out


def kernel_traffic(name):
    """HBM bytes per launch from a committed counter summary (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration it records), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch"):
            return float(line.split()[1])
    return None


def fp64_gp_choice(pot):
    """The FP64 GP form the autotuner chose for coordinate 0: ("fused", parts) or ("gemms", None)."""
    n = C.c_int()
    b, g1, g2, sp = (np.zeros(8, np.int32) for _ in range(4))
    m1, m2 = np.zeros(8), np.zeros(8)
    L.check(L.lib().gapx_model_tuning(pot.handle, 0, 8, L.ptr(b), L.ptr(g1), L.ptr(g2), L.ptr(sp), L.ptr(m1),
                                      L.ptr(m2), C.byref(n)), "gapx_model_tuning")
    k = n.value - 1
    if k >= 0 and g1[k] == -1:
        return "fused", int(sp[k])
    return "gemms", None


def stage_counters():
    """Per-launch HBM bytes (FETCH_SIZE + WRITE_SIZE scaled by a device copy), FP64 VALU
    flops and kernel-trace time of each stage's dominant kernel, from the committed
    counter summary: {(workload, stage): (kernel, bytes, flops, avg_us)}."""
    path = os.path.join(ROOT, "profiles", "descriptor_stages_si4096_batch512.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            f = line.split()
            if f and f[0] == "stage_counter":
                out[(f[1], f[2])] = (f[3], float(f[5]), float(f[6]), float(f[7]))
    return out


SINGLE_KERNEL_STAGES = ("soap_forward", "soap_adjoint", "gemm1")


def stage_roofline(forms, model, n_si, nc_batch):
    """Achieved HBM GB/s and FP64-VALU fraction per descriptor stage.  Stages that are one
    kernel (forward, adjoint, the GP kernel) divide the committed per-launch bytes and
    flops by the stage time measured in this run with HIP events on the library's stream;
    the neighbour-list and gather stages span several kernels, so their dominant kernel
    is priced at its kernel-trace time and the live stage time is given beside it.  The
    GP kernels also get their MFMA fraction (algorithmic 4 nc M d; the split form runs
    three passes)."""
    cnt = stage_counters()
    if not cnt:
        return None
    live = {"si4096_fp64": forms.get("fp64", {}).get("stage_ms", {}),
            "batch512_bf16x3": forms.get("batch512_bf16x3", {}).get("stage_ms", {})}
    centres = {"si4096_fp64": n_si, "batch512_bf16x3": nc_batch}
    out = {}
    for (wl, stage), (kernel, nbytes, flops, avg_us) in sorted(cnt.items()):
        st_ms = live.get(wl, {}).get(stage)
        if stage in SINGLE_KERNEL_STAGES and st_ms:
            t, src = st_ms * 1e-3, "HIP events, this run"
        else:
            t, src = avg_us * 1e-6, "kernel trace of the committed profile"
        e = dict(kernel=kernel, hbm_bytes_per_launch=nbytes, fp64_valu_flops_per_launch=flops,
                 kernel_ms=round(t * 1e3, 4), time_source=src, stage_ms_live=st_ms,
                 hbm_GBps=round(nbytes / t / 1e9, 1), hbm_frac=round(nbytes / t / HBM_PEAK, 4),
                 fp64_valu_TFps=round(flops / t / 1e12, 3), fp64_valu_frac=round(flops / t / VALU_FP64_PEAK, 4))
        if stage == "gemm1":   # the GP kernel: MFMA is its roof
            c = model.coords[0]
            alg = 4.0 * centres[wl] * c.n_sparse * c.descriptor.dim
            if "split" in kernel:
                e.update(bound="mfma", algorithmic_flops=alg, mfma_flops_executed=3 * alg,
                         mfma_frac_algorithmic=round(alg / t / MFMA_BF16_PEAK, 4),
                         mfma_frac_executed=round(3 * alg / t / MFMA_BF16_PEAK, 4), mfma_peak="bf16 2.5 PFLOP/s")
            else:
                e.update(bound="mfma", algorithmic_flops=alg, mfma_frac=round(alg / t / MFMA_FP64_PEAK, 4),
                         mfma_peak="FP64 78.6 TFLOP/s")
        out.setdefault(wl, {})[stage] = e
    out["source"] = ("bytes and flops per launch: profiles/descriptor_stages_si4096_batch512.txt (rocprofv3 "
                     "FETCH_SIZE, WRITE_SIZE, SQ_INSTS_VALU_*_F64); peaks HBM 8 TB/s, FP64 VALU 78.6 TFLOP/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,bf16x3,fp16x3", help="comma list of GEMM precision forms to time")
    ap.add_argument("--batch", default="fp64,bf16x3",
                    help="precision forms of the 512-structure batching form ('' to skip)")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--big-batch", default="fp64,bf16x3",
                    help="forms to time on si4096x8 (eight si4096-sized structures in one batch)")
    ap.add_argument("--big-batch-steps", type=int, default=5)
    ap.add_argument("--sic", default="fp64,bf16x3,fp16x3",
                    help="precision forms timed on the 32768-atom SiC structure ('' to skip)")
    ap.add_argument("--sic-steps", type=int, default=5)
    ap.add_argument("--serve", default="bf16x3,fp64",
                    help="precision forms of the serving form ('' to skip); the first gets the three arrival rates")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="length of each Poisson stream")
    ap.add_argument("--serve-pool", type=int, default=256, help="distinct requests the streams draw from")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: --gpus %d is not supported (the SOAP-GAP path shards over centre "
              "atoms; a multi-GPU run is not part of this build)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    tmp = tempfile.mkdtemp()
    log("building si4096 model from GPU descriptors")
    model = configs.model_si4096(gpu_descriptor_fn(tmp))
    path = os.path.join(tmp, "si4096.xml")
    write_gap_xml(path, model)
    at = configs.si4096()
    N = len(at)
    dev = DeviceArrays()
    bufs = dict(pos=dev.alloc(N * 24), e=dev.alloc(8), f=dev.alloc(N * 24), v=dev.alloc(72))
    dev.h2d(bufs["pos"], L.f64(at.positions))
    forms = {}
    results = {}
    form_list = [f for f in args.forms.split(",") if f]
    if "fp64" not in form_list:
        form_list.insert(0, "fp64")
    if args.copy_probe_mb > 0:
        nb = args.copy_probe_mb << 20
        src, dst = dev.alloc(nb), dev.alloc(nb)
        dev.hip.hipMemset(src, 1, C.c_size_t(nb))
        dev.hip.hipMemcpy(dst, src, C.c_size_t(nb), 3)
        dev.hip.hipDeviceSynchronize()
        log("copy probe: %d bytes device-to-device" % nb)
    gp_choice = ("gemms", None)
    for prec in form_list:
        pot = Potential(param_filename=path, precision=prec)
        log("timing", prec)
        dt, res, stages = run_form(pot, at, dev, bufs, args.steps, args.warmup)
        if prec == "fp64":
            gp_choice = fp64_gp_choice(pot)
        forms[prec] = dict(atoms_per_s=N * args.steps / dt, ms_per_step=1e3 * dt / args.steps,
                           stage_ms={k: round(v, 5) for k, v in stages.items()})
        results[prec] = res
        log("  %s energy %.12f, |F| sum %.12f" % (prec, res["energy"], float(np.abs(res["force"]).sum())))
        del pot
    for prec in [f for f in form_list if f in ("fp64", "bf16x3")]:
        pot = Potential(param_filename=path, precision=prec)
        log("timing", prec, "with graph replay")
        dt, res, stages = run_form(pot, at, dev, bufs, args.steps, args.warmup, graphs=True)
        forms[prec + "_graph"] = dict(atoms_per_s=N * args.steps / dt, ms_per_step=1e3 * dt / args.steps,
                                      max_dF_vs_eager=float(np.abs(res["force"] - results[prec]["force"]).max()))
        del pot
    ref = results["fp64"]
    for prec in [f for f in form_list if f != "fp64"]:
        r = results[prec]
        forms[prec]["dE_per_atom_vs_fp64"] = abs(r["energy"] - ref["energy"]) / N
        forms[prec]["force_rmse_vs_fp64"] = float(np.sqrt(np.mean((r["force"] - ref["force"]) ** 2)))
        forms[prec]["speedup_vs_fp64"] = forms[prec]["atoms_per_s"] / forms["fp64"]["atoms_per_s"]
    batch_list = [f for f in args.batch.split(",") if f]
    frames = configs.batch512()
    if batch_list:
        if "fp64" not in batch_list:
            batch_list.insert(0, "fp64")
        bres = {}
        for prec in batch_list:
            pot = Potential(param_filename=path, precision=prec)
            log("timing batch", prec)
            dt, NB, r, stages = run_batch(pot, frames, dev, args.batch_steps, 2)
            bres[prec] = r
            forms["batch512_" + prec] = dict(
                atoms_per_s=NB * args.batch_steps / dt, ms_per_step=1e3 * dt / args.batch_steps,
                structures=len(frames), atoms=NB, stage_ms={k: round(v, 4) for k, v in stages.items()},
                speedup_vs_fp64_plain=NB * args.batch_steps / dt / forms["fp64"]["atoms_per_s"])
            del pot
        if "bf16x3" in batch_list:
            pot = Potential(param_filename=path, precision="bf16x3")
            log("timing batch bf16x3 with graph replay")
            dt, NB, r, stages = run_batch(pot, frames, dev, args.batch_steps, 2, graphs=True)
            forms["batch512_bf16x3_graph"] = dict(
                atoms_per_s=NB * args.batch_steps / dt, ms_per_step=1e3 * dt / args.batch_steps,
                speedup_vs_fp64_plain=NB * args.batch_steps / dt / forms["fp64"]["atoms_per_s"],
                max_dF_vs_eager=float(np.abs(r["force"] - bres["bf16x3"]["force"]).max()))
            del pot
        rb = bres["fp64"]
        for prec in batch_list[1:]:
            r = bres[prec]
            forms["batch512_" + prec]["max_dE_per_atom_vs_fp64"] = float(
                np.max(np.abs(r["energy"] - rb["energy"]) / rb["natoms"]))
            forms["batch512_" + prec]["force_rmse_vs_fp64"] = float(np.sqrt(np.mean((r["force"] - rb["force"]) ** 2)))
    big_list = [f for f in args.big_batch.split(",") if f]
    if big_list:   # the batching form on structures of the headline's size: a like-for-like ratio
        if "fp64" not in big_list:
            big_list.insert(0, "fp64")
        big = configs.si4096_batch()
        gres = {}
        for prec in big_list:
            pot = Potential(param_filename=path, precision=prec)
            log("timing si4096x8 batch", prec)
            dt, NG, r, stages = run_batch(pot, big, dev, args.big_batch_steps, 2)
            gres[prec] = r
            forms["si4096x8_" + prec] = dict(
                atoms_per_s=NG * args.big_batch_steps / dt, ms_per_step=1e3 * dt / args.big_batch_steps,
                structures=len(big), atoms=NG, stage_ms={k: round(v, 4) for k, v in stages.items()})
            del pot
        rb = gres["fp64"]
        n0 = len(big[0])   # the first structure is si4096: the batched FP64 forces against the plain form's
        forms["si4096x8_fp64"]["max_dF_first_vs_plain"] = float(
            np.abs(rb["force"][:n0] - results["fp64"]["force"]).max())
        for prec in big_list[1:]:
            r = gres[prec]
            forms["si4096x8_" + prec]["max_dE_per_atom_vs_fp64"] = float(
                np.max(np.abs(r["energy"] - rb["energy"]) / rb["natoms"]))
            forms["si4096x8_" + prec]["force_rmse_vs_fp64"] = float(np.sqrt(np.mean((r["force"] - rb["force"]) ** 2)))
    serving = None
    serve_list = [f for f in args.serve.split(",") if f]
    sic_list = [f for f in args.sic.split(",") if f]
    smodel = None
    if serve_list or sic_list or not args.no_cpu:
        log("building the two-species SiC model from GPU descriptors")
        smodel = configs.model_sic(gpu_descriptor_fn(tmp))
        spath = os.path.join(tmp, "sic.xml")
        write_gap_xml(spath, smodel)
    if sic_list:
        if "fp64" not in sic_list:
            sic_list.insert(0, "fp64")
        sat = configs.sic32768()
        NS = len(sat)
        sb = dict(pos=dev.alloc(NS * 24), e=dev.alloc(8), f=dev.alloc(NS * 24), v=dev.alloc(72))
        dev.h2d(sb["pos"], L.f64(sat.positions))
        sres = {}
        for prec in sic_list:
            pot = Potential(param_filename=spath, precision=prec)
            log("timing sic32768", prec)
            dt, res, stages = run_form(pot, sat, dev, sb, args.sic_steps, 1)
            sres[prec] = res
            forms["sic32768_" + prec] = dict(atoms_per_s=NS * args.sic_steps / dt,
                                             ms_per_step=1e3 * dt / args.sic_steps,
                                             stage_ms={k: round(v, 4) for k, v in stages.items()})
            del pot
        for prec in sic_list[1:]:
            r, r0 = sres[prec], sres["fp64"]
            f = forms["sic32768_" + prec]
            f["dE_per_atom_vs_fp64"] = abs(r["energy"] - r0["energy"]) / NS
            f["force_rmse_vs_fp64"] = float(np.sqrt(np.mean((r["force"] - r0["force"]) ** 2)))
            f["speedup_vs_fp64"] = f["atoms_per_s"] / forms["sic32768_fp64"]["atoms_per_s"]
    if serve_list:
        serving = {"model": "Si/C two coordinates (centre Si, centre C), SOAP n_max 12 l_max 8 cutoff 5.0 (d=%d), "
                            "2 x 2000 sparse points, zeta 4" % smodel.coords[0].descriptor.dim,
                   "requests": "Si/C zincblende, N log-uniform on [16, 4096]; bulk / (111) slab / vacancies",
                   "slots": 2, "max_batch_atoms": 16384}
        for k, prec in enumerate(serve_list):
            log("serving", prec)
            serving[prec] = run_serving(spath, prec, args.serve_pool, args.serve_seconds,
                                        (0.5, 0.8, 0.95) if k == 0 else (), 2, 16384)
    st = forms["fp64"]["stage_ms"]
    nc, M, d = N, model.coords[0].n_sparse, model.coords[0].descriptor.dim
    if gp_choice[0] == "fused":   # K = X Xs^T, W and G = W Xs in one kernel: both products
        gp_flops = 4.0 * nc * M * d
        roofline = dict(bound="mfma", kernel="k_gp_fused_f64 (K = X Xs^T, energy and W = w zeta K^(zeta-1), "
                        "G = W Xs in one pass, W in registers; FP64 MFMA 16x16x4; %d sparse-point parts)"
                        % gp_choice[1], traffic=kernel_traffic("k_gp_fused_f64_si4096_traffic.txt"),
                        traffic_source="HBM bytes per launch: FETCH_SIZE + WRITE_SIZE of this kernel scaled by a "
                                       "copy-kernel calibration, profiles/k_gp_fused_f64_si4096_traffic.txt")
    else:
        gp_flops = 2.0 * nc * M * d
        roofline = dict(bound="mfma", kernel="k_gemm_f64 GEMM 1 (K = X Xs^T, fused pow/alpha epilogue; FP64 MFMA "
                        "16x16x4)", traffic=kernel_traffic("k_gemm_f64_gemm1_si4096_traffic.txt"),
                        traffic_source="HBM bytes per launch: FETCH_SIZE + WRITE_SIZE of this kernel scaled by a "
                                       "copy-kernel calibration, profiles/k_gemm_f64_gemm1_si4096_traffic.txt")
    achieved = gp_flops / (st["gemm1"] * 1e-3)
    roofline.update(achieved=achieved / 1e12, peak=MFMA_FP64_PEAK / 1e12, unit="TFLOP/s",
                    frac=achieved / MFMA_FP64_PEAK, flops_per_launch=gp_flops, avg_launch_ms=st["gemm1"],
                    time_source="HIP events around the kernel on the library's stream, this run")
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu, ro = cpu_baseline(model, at)
        forms["fp64"]["dE_per_atom_vs_oracle"] = abs(ref["energy"] - ro["energy"]) / N
        forms["fp64"]["max_dF_vs_oracle"] = float(np.abs(ref["force"] - ro["forces"]).max())
        log("cpu baseline, other workloads")
        m64 = configs.model_si64(gpu_descriptor_fn(tmp))
        cpu["other_workloads"] = cpu_baseline_configs(m64, model, smodel, frames, cpu["cores"])
        cpu["host"] = host_cpu()
    same = {}
    for key, a, b in (("si4096_eager", "bf16x3", "fp64"), ("si4096_graph", "bf16x3_graph", "fp64_graph"),
                      ("batch512_eager", "batch512_bf16x3", "batch512_fp64"),
                      ("si4096x8_eager", "si4096x8_bf16x3", "si4096x8_fp64"),
                      ("sic32768_eager", "sic32768_bf16x3", "sic32768_fp64")):
        if a in forms and b in forms:
            same[key] = forms[a]["atoms_per_s"] / forms[b]["atoms_per_s"]
    nc_batch = sum(len(a) for a in frames)
    stages_roof = stage_roofline(forms, model, N, nc_batch)
    dev.free()
    fp = forms["fp64"]
    mixed = [k for k in forms if k.startswith("batch512_") and "fp64" not in k]
    line = {
        "metric": "SOAP-GAP atoms/s (E+F+virial), mixed vs FP64 GPU; p50/p99 served latency",
        "value": fp["atoms_per_s"], "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": fp["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded si4096 structure, model written as GAP XML from the seed)",
        "config": {"workload": "si4096: 8x8x8 diamond Si, sigma 0.15, 1% vacancies (4055 atoms); SOAP n_max 8 "
                               "l_max 6 cutoff 5.0 (d=253), 2000 sparse points, zeta 4; FP64 plain form",
                   "atoms": N, "sparse_points": M, "descriptor_dim": d},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stage_roofline": stages_roof,
        "mixed_batched_vs_fp64_plain": max([forms[k]["speedup_vs_fp64_plain"] for k in mixed], default=None),
        "mixed_vs_fp64_same_workload": same,
        "forms": forms,
        "serving": serving,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
