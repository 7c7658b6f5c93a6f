#!/usr/bin/env python
"""Benchmark of the FDTD field-update hot path on one MI355X.

Workload (headline): 3-D full-vector FDTD on 256^3 cells -- background
eps = 1, 40 seed-placed dielectric spheres (radius 4-20 cells, eps in
[1.5, 12]), a 12-cell CPML on all faces, a Gaussian-pulse Ex point source
at the centre, Courant 0.5 -- run in FP32 (the value) and in FP64 (the
reference's precision).  A "step" is one full leapfrog step (E and H half
steps with ghosts, CPML, source) over the whole grid; fields stay resident
in HBM.  Metric: Mcell-updates/s = cells x steps / s.

Also measured, as extra fields of the same line:

* ``batched``: 64 heterogeneous FP32 instances (32 2-D TMz of 128-1024
  cells per side, 32 3-D of 64-192 cells per side; dielectric, Drude and
  Lorentz objects; UPML or CPML) advanced together for 1000 steps by one
  launch per kernel kind, beside the same instances stepped one at a time.
* ``dispersive``: the 384 x 384 x 192 dispersive Bloch cell (Drude blocks,
  Lorentz spheres, a DCP slab, CPML on z, TFSF plane wave) in complex FP32,
  timed over a few hundred of its 3000 steps.
* ``served``: a Poisson stream (20 requests/s, 200 requests) of
  photonic-crystal-slab solves (64^3-256^3 cells, complex FP32, TFSF plane
  wave, Bloch x/y, CPML on z, a 50-frequency running-DFT flux plane,
  1500-4000 steps) through the continuous-batching request loop; p50/p99
  latency (completion minus arrival) beside aggregate Mcell-updates/s.
  Admission takes the waiting request of least work first
  (``--serve-policy arrival`` for arrival order).
* ``air2d``: the plain FP64 method, 2-D TMz 200 x 200 in air with a CPML
  shell run to t = 10 (286 steps), in time per step on the GPU and with
  the reference's own updaters on the host.

Printed as one JSON line on stdout; diagnostics go to stderr.
"""

import argparse
import contextlib
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(sim, electric):
    """Algorithmic HBM bytes of one main-kernel launch: every field word of
    the updated type read and written once, every word of the other type
    read once, the two CPML auxiliary words of each CPML cell and the one
    UPML auxiliary word of each UPML cell read and written once.  Material
    ids and coefficient tables are not counted."""
    from gmes_amd import constant as const
    from gmes_amd._abi import CLS_CPML, CLS_UPML, NUMPY_DTYPE
    s = NUMPY_DTYPE[sim.dtype_code]().itemsize
    own = const.E_COMPONENTS if electric else const.H_COMPONENTS
    other = const.H_COMPONENTS if electric else const.E_COMPONENTS
    n_own = sum(sim._arrays[c].size for c in own if c in sim._active)
    n_other = sum(sim._arrays[c].size for c in other if c in sim._active)
    n_cpml = sum(sim._tiles.class_counts[c].get(CLS_CPML, 0) for c in own if c in sim._active)
    n_upml = sum(sim._tiles.class_counts[c].get(CLS_UPML, 0) for c in own if c in sim._active)
    return s * (2 * n_own + n_other) + 4 * s * n_cpml + 2 * s * n_upml


def measured_traffic(n, fam):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-kernel scaled), for the
    256^3 FP32 workload only; None otherwise."""
    path = os.path.join(ROOT, "profiles", "k_main3d_spheres256_fp32.md")
    if n != 256 or not os.path.exists(path):
        return None
    key = "traffic_bytes_per_launch_%s:" % fam
    for line in open(path):
        if line.startswith(key):
            return int(line.split(":")[1])
    return None


def measured_traffic_batch(fam):
    """HBM bytes per launch of the batched 3-D main kernel (E or H) from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-kernel scaled);
    None if there is none."""
    path = os.path.join(ROOT, "profiles", "k_main3d_batch64_fp32.md")
    if not os.path.exists(path):
        return None
    key = "traffic_bytes_per_launch_%s:" % ("E" if fam.startswith("e") else "H")
    for line in open(path):
        if line.startswith(key):
            return int(line.split(":")[1])
    return None


def time_steps(sim, lib, steps):
    lib.sync(sim._sim)
    t0 = time.perf_counter()
    sim._advance(steps)
    lib.sync(sim._sim)
    return time.perf_counter() - t0


def fp32_vs_fp64(sim32, sim64, lib, horizon):
    """Relative L2 error of the FP32 fields against the FP64 fields of the
    same workload, both advanced (untimed) to the same step ``horizon``."""
    from gmes_amd import constant as const
    for s in (sim32, sim64):
        done = lib.get_step(s._sim)
        s._advance(horizon - done)  # raises if a sim is already past the horizon
        lib.sync(s._sim)
    steps = {lib.get_step(s._sim) for s in (sim32, sim64)}
    if steps != {horizon}:
        raise RuntimeError("FP32 and FP64 sims are at steps %s, not %d" % (sorted(steps), horizon))
    num = den = 0.0
    for c in const.COMPONENTS:
        a = np.asarray(sim32.field[c], np.float64)
        b = np.asarray(sim64.field[c], np.float64)
        num += float(np.sum((a - b) ** 2))
        den += float(np.sum(b ** 2))
    err = dict(steps=horizon, rel_l2_fp32_vs_fp64=(num / den) ** 0.5 if den else 0.0, bound=1e-5)
    log("fp32 vs fp64 after %d steps: %.3g" % (horizon, err["rel_l2_fp32_vs_fp64"]))
    return err


def host_cpu():
    """CPU model and topology (lscpu) and the CPUs this process may use."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass

    def num(key):
        try:
            return int(info.get(key, ""))
        except ValueError:
            return None
    sockets, cores = num("Socket(s)"), num("Core(s) per socket")
    return dict(model=info.get("Model name"), sockets=sockets, cores_per_socket=cores,
                threads_per_core=num("Thread(s) per core"),
                physical_cores=sockets * cores if sockets and cores else None,
                usable_cpus=len(os.sched_getaffinity(0)))


def set_omp_threads(n):
    """Team size of later OpenMP regions of this thread in the reference
    build (it runs on the system libgomp)."""
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def median_s_per_step(sim, warm, reps, steps):
    """benchmarks/field_updates.py:122-156 of the reference: warm-up steps,
    then the median over repeats of the time per step."""
    sim._advance(warm)
    per = []
    for _ in range(reps):
        t0 = time.perf_counter()
        sim._advance(steps)
        per.append((time.perf_counter() - t0) / steps)
    return float(np.median(per))


def cpu_baseline(n, reps=5, steps=20, warm=10, reps1=3, steps1=2, warm1=1):
    """The reference's own updaters (oracle/_ref/libgmes_ref.so, compiled
    from the reference's src/pw_*.hh) in the reference's step order, FP64
    (complex128 for the served slab), OpenMP over the host's usable cores and
    on one thread.  Bounded samples (a few seconds each) of: the 256^3
    workload, 8 of the 64 batched instances run one after another, and the
    median-size served slab."""
    from gmes_amd import workloads
    from gmes_amd._abi import Library
    path = os.path.join(ROOT, "oracle", "_ref", "libgmes_ref.so")
    kind = "reference"
    if not os.path.exists(path):
        path = os.path.join(ROOT, "oracle", "libgmes_oracle.so")
        kind = "port"
    lib = Library(path)
    cpu = host_cpu()
    threads = cpu["usable_cpus"]
    if cpu["physical_cores"]:
        threads = min(threads, cpu["physical_cores"] * (cpu["threads_per_core"] or 1))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    t_all = time.perf_counter()
    # 256^3 spheres + CPML, all usable cores
    set_omp_threads(threads)
    sim = workloads.spheres3d(n, precision="fp64", _engine=lib)
    t_n = median_s_per_step(sim, warm, reps, steps)
    done = warm + reps * steps
    # the same workload on one thread (fewer steps: ~0.5 s each)
    set_omp_threads(1)
    sim1 = workloads.spheres3d(n, precision="fp64", _engine=lib)
    t_1 = median_s_per_step(sim1, warm1, reps1, steps1)
    sim1.close()
    # batched workload, one instance after another: every 8th instance
    set_omp_threads(threads)
    pick = set(range(0, 64, 8))
    inst = workloads.batch_instances(precision="fp64", _engine=lib, only=pick)
    cells4 = sum(int(np.prod(s.space.whole_field_size)) for s in inst)
    for s in inst:
        s._advance(1)
    t0 = time.perf_counter()
    for s in inst:
        s._advance(2)
    t4 = (time.perf_counter() - t0) / 2
    for s in inst:
        s.close()
    # the served stream's median-size slab, complex128
    sides = sorted(r[1] for r in workloads.serve_requests(200))
    side = sides[len(sides) // 2]
    slab, _ = workloads.pc_slab(side, 0.3, precision="fp64", _engine=lib)
    t5 = median_s_per_step(slab, 1, 3, 3)
    slab.close()
    secs = time.perf_counter() - t_all
    value = n ** 3 / t_n / 1e6
    out = dict(value=round(value, 2), unit="Mcell-updates/s", cores=threads, kind=kind,
               sample="%d^3 spheres+CPML workload, FP64, OpenMP on %d threads (OMP_PROC_BIND=close): "
                      "%d warm-up steps, median of %d x %d steps" % (n, threads, warm, reps, steps),
               cpu=cpu,
               one_thread=dict(value=round(n ** 3 / t_1 / 1e6, 2), unit="Mcell-updates/s",
                               sample="%d warm-up step(s), median of %d x %d steps" % (warm1, reps1, steps1)),
               batched_sequential=dict(value=round(cells4 / t4 / 1e6, 2), unit="Mcell-updates/s",
                                       instances=sorted(pick), cells=cells4, steps=2, threads=threads),
               served_median_slab=dict(side=side, s_per_step=round(t5, 5),
                                       value=round(side ** 3 / t5 / 1e6, 2), unit="Mcell-updates/s",
                                       dtype="c128", threads=threads),
               wall_s=round(secs, 1))
    return sim, done, out


def batched_leg(steps, warm=5, alone_steps=50, prof_steps=50):
    """The 64-instance batched workload: Mcell-updates/s summed over the
    instances, batched and one instance at a time."""
    from gmes_amd import workloads
    from gmes_amd.batch import Batch
    t = time.perf_counter()
    sims = workloads.batch_instances(precision="fp32")
    setup = time.perf_counter() - t
    cells = sum(int(np.prod(s.space.whole_field_size)) for s in sims)
    nbytes = sum(kernel_bytes(s, True) + kernel_bytes(s, False) for s in sims)
    log("batch: %d instances, %.1f M cells, setup %.1f s" % (len(sims), cells / 1e6, setup))
    # one at a time (each on its own stream, synchronised per instance)
    lib = sims[0]._lib
    t0 = time.perf_counter()
    for s in sims:
        s._advance(alone_steps)
        lib.sync(s._sim)
    alone = cells * alone_steps / (time.perf_counter() - t0) / 1e6
    b = Batch(sims)
    b.advance(warm)
    b.sync()
    t0 = time.perf_counter()
    b.advance(steps)
    b.sync()
    secs = time.perf_counter() - t0
    # roofline of the dominant launch (the 3-D main kernel of all 3-D
    # members), HIP events on the batch's stream
    b.profile(True)
    b.advance(prof_steps)
    b.sync()
    ms, launches = b.read_profile()
    b.profile(False)
    b.close()
    three = [s for s in sims if len(s._active) == 6]
    fam = {}
    for k, name, sel, el in ((0, "e_main3d", three, True), (1, "h_main3d", three, False),
                             (2, "e_main2d", [s for s in sims if s not in three], True),
                             (3, "h_main2d", [s for s in sims if s not in three], False)):
        avg = ms[k] / max(1, launches[k])
        nb = sum(kernel_bytes(s, el) for s in sel)
        fam[name] = dict(avg_launch_ms=round(avg, 5), bytes_per_launch=int(nb),
                         gbps=round(nb / (avg * 1e-3) / 1e9, 1) if avg > 0 else None)
    split = {k: round(ms[i] / max(1, launches[i]), 4) for i, k in enumerate(
        ("e_main3d", "h_main3d", "e_main2d", "h_main2d", "ade", "rest"))}
    dom = max(("e_main3d", "h_main3d"), key=lambda k: fam[k]["avg_launch_ms"])
    roof = dict(bound="hbm", achieved=fam[dom]["gbps"], peak=HBM_PEAK_GBPS, unit="GB/s",
                frac=round(fam[dom]["gbps"] / HBM_PEAK_GBPS, 4),
                traffic=measured_traffic_batch(dom),
                kernel="k_main3d_batch<%s, float> (%d 3-D members, one launch per profile-row kind)" % (
                    "true" if dom == "e_main3d" else "false", len(three)),
                bytes_per_launch=fam[dom]["bytes_per_launch"], avg_launch_ms=fam[dom]["avg_launch_ms"],
                families=fam)
    for s in sims:
        s.close()
    value = cells * steps / secs / 1e6
    log("batch: %d steps in %.3f s -> %.1f Mcell/s (one at a time %.1f); ms/launch %s; %s %.0f GB/s"
        % (steps, secs, value, alone, split, dom, fam[dom]["gbps"] or 0))
    return dict(instances=len(sims), cells=cells, steps=steps, value=round(value, 1),
                unit="Mcell-updates/s", ms_per_step=round(secs / steps * 1e3, 4),
                hbm_gbps_algorithmic=round(nbytes * steps / secs / 1e9, 1),
                hbm_pct_step=round(100 * nbytes * steps / secs / 1e9 / HBM_PEAK_GBPS, 1),
                roofline=roof, ms_per_launch=split,
                one_at_a_time_value=round(alone, 1), one_at_a_time_steps=alone_steps,
                setup_s=round(setup, 1), dtype="f32")


def dispersive_leg(steps, warm=10):
    """The dispersive Bloch cell in complex FP32."""
    from gmes_amd import workloads
    t = time.perf_counter()
    sim = workloads.dispersive3d(precision="fp32")
    setup = time.perf_counter() - t
    lib = sim._lib
    cells = int(np.prod(sim.space.whole_field_size))
    lib.set_graph(sim._sim, 1)
    sim._advance(warm)
    secs = time_steps(sim, lib, steps)
    lib.profile(sim._sim, 1)
    sim._advance(50)
    lib.sync(sim._sim)
    ms, launches = lib.read_profile(sim._sim)
    lib.profile(sim._sim, 0)
    sim.close()
    value = cells * steps / secs / 1e6
    split = {k: round(ms[i] / max(1, launches[i]), 4) for i, k in enumerate(
        ("e_main", "h_main", "ade", "sources", "ghost", "aux_grid"))}
    log("dispersive: %d steps in %.3f s -> %.1f Mcell/s (setup %.1f s), ms/launch %s"
        % (steps, secs, value, setup, split))
    return dict(cells=cells, steps=steps, value=round(value, 1), unit="Mcell-updates/s",
                ms_per_step=round(secs / steps * 1e3, 4), dtype="c64", setup_s=round(setup, 1),
                ms_per_launch=split)


def air2d_leg():
    """The plain FP64 method: 2-D TMz in air, 200 x 200 cells, run to t = 10
    (286 steps, examples/air2d.py of the reference), on the GPU and with the
    reference's own updaters on the host.  The grid is cache resident, so
    this reports time per step rather than a bandwidth."""
    from gmes_amd import workloads
    from gmes_amd._abi import Library
    out = {}
    for name, lib in (("gpu", None), ("cpu", "reference")):
        engine = None
        if lib:
            path = os.path.join(ROOT, "oracle", "_ref", "libgmes_ref.so")
            if not os.path.exists(path):
                path = os.path.join(ROOT, "oracle", "libgmes_oracle.so")
                lib = "port"
            engine = Library(path)
        sim = workloads.air2d(precision="fp64", _engine=engine)
        if engine is None:
            sim._lib.set_graph(sim._sim, 1)
        cells = int(np.prod(sim.space.whole_field_size))
        t0 = time.perf_counter()
        with contextlib.redirect_stdout(sys.stderr):  # its progress lines
            sim.step_until_t(10)
        sim._lib.sync(sim._sim)
        secs = time.perf_counter() - t0
        steps = int(sim.time_step.n)
        out[name] = dict(us_per_step=round(secs / steps * 1e6, 2), steps=steps,
                         value=round(cells * steps / secs / 1e6, 1), unit="Mcell-updates/s")
        if lib:
            out[name]["kind"] = lib
        sim.close()
        if engine is None:
            # the same run driven the way scripts usually drive the
            # reference: one step() call per step (waveforms are uploaded
            # in windows, so the calls only enqueue work)
            sim = workloads.air2d(precision="fp64")
            sim._lib.set_graph(sim._sim, 1)
            t0 = time.perf_counter()
            for _ in range(steps):
                sim.step()
            sim._lib.sync(sim._sim)
            loop = time.perf_counter() - t0
            out[name]["step_loop_us_per_step"] = round(loop / steps * 1e6, 2)
            sim.close()
    log("air2d: gpu %.1f us/step (step() loop %.1f), cpu %.1f us/step (%d steps)"
        % (out["gpu"]["us_per_step"], out["gpu"]["step_loop_us_per_step"], out["cpu"]["us_per_step"],
           out["gpu"]["steps"]))
    return out


def served_leg(nreq, budget_gb=64.0, max_active=16, chunk=250, policy="shortest", prof_steps=20):
    """The served photonic-crystal-slab stream through gmes_amd.serve, then
    the dominant kernel timed on one batch of 16 slabs drawn from the same
    stream."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor
    from gmes_amd import workloads
    from gmes_amd.batch import Batch
    from gmes_amd.serve import Server, latency_stats
    # material tables are built in worker processes (host only, no GPU)
    pool = ProcessPoolExecutor(max_workers=8, mp_context=multiprocessing.get_context("spawn"))
    reqs = workloads.serve_request_list(n=nreq, host_pool=pool)
    work = sum(float(r.side) ** 3 * r.steps for r in reqs)
    step_bytes = {}

    def counted(r):
        build = r.build

        def wrapped():
            sim, fn = build()
            step_bytes[r.rid] = kernel_bytes(sim, True) + kernel_bytes(sim, False)
            return sim, fn
        return wrapped
    for r in reqs:
        r.build = counted(r)
    srv = Server(hbm_budget=int(budget_gb * 2 ** 30), max_active=max_active, chunk=chunk, workers=8,
                 policy=policy)
    try:
        done = srv.run(reqs)
        # the served form's dominant kernel: one batch of the stream's first
        # 16 slabs, timed with HIP events on the batch stream
        sample = workloads.serve_request_list(n=16, host_pool=pool)
        built = [r.build() for r in sample]
    finally:
        pool.shutdown(wait=True)
    ok = [r for r in done if r.error is None]
    st = latency_stats(done)
    value = work / srv.elapsed / 1e6
    moved = sum(step_bytes[r.rid] * r.steps for r in ok)
    sims = [sim for sim, _ in built]
    b = Batch(sims)
    b.advance(5)
    b.sync()
    b.profile(True)
    b.advance(prof_steps)
    b.sync()
    ms, launches = b.read_profile()
    b.profile(False)
    b.close()
    fam = 0 if ms[0] >= ms[1] else 1
    avg = ms[fam] / max(1, launches[fam])
    nb = sum(kernel_bytes(sm, fam == 0) for sm in sims)
    for sm in sims:
        sm.close()
    gbps = nb / (avg * 1e-3) / 1e9
    roof = dict(bound="hbm", achieved=round(gbps, 1), peak=HBM_PEAK_GBPS, unit="GB/s",
                frac=round(gbps / HBM_PEAK_GBPS, 4), traffic=None,
                kernel="k_main3d_batch<%s, cplx<float>> on the stream's first 16 slabs (sides %s)"
                       % ("true" if fam == 0 else "false", sorted(r.side for r in sample)),
                bytes_per_launch=int(nb), avg_launch_ms=round(avg, 5))
    log("served: %d requests (%d ok) in %.1f s, p50 %.2f s p99 %.2f s, %.1f Mcell/s, peak active %d; "
        "stream %.0f GB/s algorithmic; batch-16 main kernel %.0f GB/s"
        % (len(done), len(ok), srv.elapsed, st["p50"] or -1, st["p99"] or -1, value, srv.peak_active,
           moved / srv.elapsed / 1e9, gbps))
    return dict(requests=len(done), completed=len(ok), rate_per_s=20.0,
                arrival_window_s=round(max(r.arrival for r in reqs), 3),
                p50_latency_s=round(st["p50"], 3) if st["p50"] is not None else None,
                p99_latency_s=round(st["p99"], 3) if st["p99"] is not None else None,
                mean_latency_s=round(st["mean"], 3) if st["mean"] is not None else None,
                value=round(value, 1), unit="Mcell-updates/s", elapsed_s=round(srv.elapsed, 2),
                cell_updates=work, hbm_gbps_algorithmic_stream=round(moved / srv.elapsed / 1e9, 1),
                roofline=roof, peak_active=srv.peak_active, rounds=srv.rounds,
                hbm_budget_gb=budget_gb, dtype="c64", policy=policy,
                host_s=dict(idle=round(srv.t_idle, 2), batching=round(srv.t_batch, 2),
                            advance=round(srv.t_advance, 2), sync=round(srv.t_sync, 2)),
                note="latency includes each solve's host set-up (8 table-building processes); "
                     "stream GB/s counts the main kernels' algorithmic bytes over the whole elapsed time")


def copy_gbps(reps=20):
    """Read-plus-write rate of a device-to-device hipMemcpyAsync of 256 MiB
    into 256 MiB (HIP events on the null stream; the HIP runtime the engine
    links): the rate a plain stream reaches on this part, beside the 8 TB/s
    nominal peak.  Plain HIP calls keep it usable under a kernel tracer."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    n = ctypes.c_size_t(256 << 20)
    a, b, e0, e1 = (ctypes.c_void_p() for _ in range(4))

    def ok(rc):
        if rc != 0:
            raise RuntimeError("HIP error %d in the copy probe" % rc)

    ok(hip.hipMalloc(ctypes.byref(a), n))
    ok(hip.hipMalloc(ctypes.byref(b), n))
    try:
        ok(hip.hipMemset(a, 1, n))
        ok(hip.hipEventCreate(ctypes.byref(e0)))
        ok(hip.hipEventCreate(ctypes.byref(e1)))
        for _ in range(3):
            ok(hip.hipMemcpyAsync(b, a, n, 3, None))  # 3: hipMemcpyDeviceToDevice
        ok(hip.hipEventRecord(e0, None))
        for _ in range(reps):
            ok(hip.hipMemcpyAsync(b, a, n, 3, None))
        ok(hip.hipEventRecord(e1, None))
        ok(hip.hipEventSynchronize(e1))
        ms = ctypes.c_float()
        ok(hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1))
        ok(hip.hipEventDestroy(e0))
        ok(hip.hipEventDestroy(e1))
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return 2 * n.value / (ms.value / reps * 1e-3) / 1e9


def copy_probe():
    """A plain device-to-device copy of 256 MiB in, 256 MiB out, five times."""
    import torch
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("copy probe: 5 x %d bytes" % (src.numel() * 4))


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256, help="cells per axis")
    ap.add_argument("--fp64-steps", type=int, default=500)
    ap.add_argument("--error-steps", type=int, default=2000,
                    help="step at which the FP32 fields are compared with the FP64 fields")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-ghost-fusion", action="store_true",
                    help="keep the separate E ghost-plane copy in the 256^3 legs")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--dispersive-steps", type=int, default=400)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-policy", default="shortest", choices=("arrival", "shortest"),
                    help="admission order of the served stream")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy 5 times (scale reference for "
                         "the FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py: this engine runs one simulation per GPU on one-GPU boxes; "
              "--gpus %d is not supported (use --gpus 1)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from gmes_amd import _abi, workloads
    lib = _abi.engine()
    n = args.n
    cells = n ** 3

    # ---- FP32 headline ---------------------------------------------------
    t = time.perf_counter()
    sim = workloads.spheres3d(n, precision="fp32")
    log("fp32 setup %.1f s, footprint %.2f GB" % (time.perf_counter() - t,
                                                  lib.footprint(sim._sim) / 1e9))
    lib.set_graph(sim._sim, 0 if args.no_graph else 1)
    fused = lib.set_ghost_fusion(sim._sim, 0 if args.no_ghost_fusion else 1)
    sim._advance(args.warmup)
    secs = time_steps(sim, lib, args.steps)
    value = cells * args.steps / secs / 1e6
    log("fp32: %d steps in %.3f s -> %.1f Mcell/s" % (args.steps, secs, value))

    # ---- roofline of the dominant kernel (HIP events on the sim's stream)
    prof_steps = min(args.steps, 200)
    lib.profile(sim._sim, 1)
    sim._advance(prof_steps)
    lib.sync(sim._sim)
    ms, launches = lib.read_profile(sim._sim)
    lib.profile(sim._sim, 0)
    fam = 0 if ms[0] >= ms[1] else 1
    avg_ms = ms[fam] / max(1, launches[fam])
    nbytes = kernel_bytes(sim, electric=(fam == 0))
    achieved = nbytes / (avg_ms * 1e-3) / 1e9
    other = 1 - fam
    other_avg = ms[other] / max(1, launches[other])
    other_gbps = kernel_bytes(sim, electric=(other == 0)) / (other_avg * 1e-3) / 1e9
    log("kernel ms/launch: E %.4f H %.4f ghost %.4f src %.4f" % (
        ms[0] / max(1, launches[0]), ms[1] / max(1, launches[1]),
        ms[4] / max(1, launches[4]), ms[3] / max(1, launches[3])))
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBPS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBPS, 4),
                    traffic=measured_traffic(n, "E" if fam == 0 else "H"),
                    kernel="k_main3d<%s, float, 1, false, true, true, false>" % ("true" if fam == 0 else "false"),
                    bytes_per_launch=int(nbytes), avg_launch_ms=round(avg_ms, 5),
                    other_kernel_gbps=round(other_gbps, 1))
    cg = copy_gbps()
    roofline["copy_gbps"] = round(cg, 1)
    if roofline["traffic"]:
        roofline["traffic_gbps"] = round(roofline["traffic"] / (avg_ms * 1e-3) / 1e9, 1)
    log("device copy (256 MiB): %.0f GB/s; kernel algorithmic %.0f GB/s" % (cg, achieved))
    fp32_steps_done = args.warmup + args.steps + prof_steps

    # ---- FP64 (the reference's precision) --------------------------------
    fp64 = None
    fp32_err = None
    if args.fp64_steps > 0:
        sim64 = workloads.spheres3d(n, precision="fp64")
        lib.set_graph(sim64._sim, 0 if args.no_graph else 1)
        lib.set_ghost_fusion(sim64._sim, 0 if args.no_ghost_fusion else 1)
        warm64 = min(args.warmup, 10)
        sim64._advance(warm64)
        s64 = time_steps(sim64, lib, args.fp64_steps)
        fp64 = dict(value=round(cells * args.fp64_steps / s64 / 1e6, 1), unit="Mcell-updates/s",
                    steps=args.fp64_steps, ms_per_step=round(s64 / args.fp64_steps * 1e3, 4))
        log("fp64: %d steps in %.3f s -> %.1f Mcell/s" % (args.fp64_steps, s64, fp64["value"]))
        fp32_err = fp32_vs_fp64(sim, sim64, lib, max(args.error_steps, fp32_steps_done,
                                                     warm64 + args.fp64_steps))
        sim64.close()
    sim.close()

    # ---- batched instances -----------------------------------------------
    batched = None
    if not args.no_batch and args.batch_steps > 0:
        batched = batched_leg(args.batch_steps)

    # ---- dispersive Bloch cell --------------------------------------------
    air2d = None if args.no_cpu else air2d_leg()
    dispersive = None
    if args.dispersive_steps > 0:
        dispersive = dispersive_leg(args.dispersive_steps)

    # ---- served stream ---------------------------------------------------
    served = None
    if not args.no_serve and args.serve_requests > 0:
        served = served_leg(args.serve_requests, policy=args.serve_policy)

    # ---- CPU baseline ----------------------------------------------------
    cpu = None
    parity = None
    if not args.no_cpu:
        ref_sim, ref_steps, cpu = cpu_baseline(n)
        log("cpu baseline:", cpu)
        # the same steps on the GPU in FP64 must reproduce the reference
        from gmes_amd import constant as const
        g64 = workloads.spheres3d(n, precision="fp64")
        g64._advance(ref_steps)
        num = den = 0.0
        for c in const.COMPONENTS:
            a = np.asarray(g64.field[c], np.float64)
            b = np.asarray(ref_sim.field[c], np.float64)
            num += float(np.sum((a - b) ** 2))
            den += float(np.sum(b ** 2))
        parity = dict(steps=ref_steps, rel_l2_fp64_gpu_vs_reference=(num / den) ** 0.5 if den else 0.0)
        log("parity:", parity)
        g64.close()

    out = {
        "metric": "Mcell-updates/s and % HBM peak (FP32 batched vs FP64 plain); p50/p99 solve latency",
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3-D FDTD %d^3, 40 dielectric spheres, 12-cell CPML, Bandpass Ex "
                               "source, Courant 0.5, FP32" % n, "cells": cells,
                   "graph": not args.no_graph, "fused_e_ghosts": fused},
        "hbm_pct": round(100 * achieved / HBM_PEAK_GBPS, 1),
        "fp64": fp64,
        "batched": batched,
        "dispersive": dispersive,
        "air2d": air2d,
        "served": served,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "fp32_error": fp32_err,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
