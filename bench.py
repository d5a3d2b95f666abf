"""Benchmark: HIF-DE factor + PCG solve to 1e-10 at n = 128 on one MI355X.

Workload (one step): the mixed-precision factorization of the n = 128 grid
(127^3 unknowns; lognormal coefficient a = exp(g), g a unit-variance Gaussian
random field with correlation length 0.1, seed 2; symmetric 7-point stencil;
leaf P = 4; ID tolerance 1e-4) followed by one FP64 PCG solve to a 1e-10
relative residual with the factorization as preconditioner (RHS iid N(0,1),
seed 2).  The operator and RHS are resident in HBM before the timed region.

value = steps per second (factor + solve).  The same run also times the FP64
GPU path once and the CPU oracle on a bounded sample (see cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "factor s + PCG solves/s at 128^3 to 1e-10 resid: mixed vs FP64 GPU vs CPU"
PEAK = {"hbm_GBs": 8000.0, "fp64_TF": 78.6, "fp32_TF": 157.3}
PMC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_n128.json")
PMC_UTIL_FILE = os.path.join(ROOT, "profiles", "pmc_mfma_n128.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_problem(n, seed, corr):
    from hifde3d_amd import problems as pb
    a = pb.lognormal_coeff(n, seed, corr)
    V = pb.zeros_coeff(n)
    b = pb.rhs_normal(n, seed)
    return a, V, b


def short_kernel_name(name):
    """rocprofv3's kernel name without return type, namespaces and arguments."""
    for pre in ("void ", "hifde::(anonymous namespace)::", "hifde::"):
        if name.startswith(pre):
            name = name[len(pre):]
    return name.split("(")[0]


def kernel_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload, scaled by the
    copy-kernel calibration; None if not available."""
    if not os.path.exists(PMC_FILE):
        return None
    pm = json.load(open(PMC_FILE))
    e = pm["kernels"].get(kernel)
    if not e or e["dispatches"] <= 0:
        return None
    return 1024.0 * (pm["fetch_scale"] * e["fetch_kib"] + pm["write_scale"] * e["write_kib"]) / \
        e["dispatches"]


def roofline_from_stats(stats):
    """Roofline of the kernel with the most device time, from the per-launch
    HIP-event records (families named kernel:<name>)."""
    kern = {k[len("kernel:"):]: v for k, v in stats.items() if k.startswith("kernel:")}
    name = max(kern, key=lambda k: kern[k]["sec"])
    st = kern[name]
    avg = st["sec"] / max(st["launches"], 1)
    traffic = kernel_traffic(name)
    if "gemm" in name:
        peak = PEAK["fp32_TF"] if name.split("<")[1].startswith("float") else PEAK["fp64_TF"]
        achieved = st["flops"] / st["sec"] / 1e12
        return name, avg, {"bound": "mfma", "achieved": achieved, "peak": peak,
                           "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic}
    achieved = st["bytes"] / st["sec"] / 1e9
    return name, avg, {"bound": "hbm", "achieved": achieved, "peak": PEAK["hbm_GBs"],
                       "unit": "GB/s", "frac": achieved / PEAK["hbm_GBs"], "traffic": traffic}


def pmc_json(fetch_csv, write_csv, calfetch_csv, calwrite_csv, workload):
    """Per-kernel FETCH_SIZE / WRITE_SIZE totals of two rocprofv3 --pmc runs,
    with the scale factors that make a 1 GiB device copy read and write 1 GiB."""
    import csv

    def load(path):
        out = {}
        for r in csv.DictReader(open(path)):
            k = short_kernel_name(r["Kernel_Name"])
            e = out.setdefault(k, [0, 0.0])
            e[0] += 1
            e[1] += float(r["Counter_Value"])
        return out

    f, w = load(fetch_csv), load(write_csv)
    cf = load(calfetch_csv)["__amd_rocclr_copyBuffer"]
    cw = load(calwrite_csv)["__amd_rocclr_copyBuffer"]
    gib_kib = float(1 << 20)
    fetch_scale, write_scale = gib_kib / (cf[1] / cf[0]), gib_kib / (cw[1] / cw[0])
    kernels = {k: {"dispatches": f[k][0], "fetch_kib": f[k][1], "write_kib": w.get(k, [0, 0.0])[1]}
               for k in f}
    return {"workload": workload, "fetch_scale": round(fetch_scale, 4),
            "write_scale": round(write_scale, 4),
            "calibration": (f"1 GiB device copy (__amd_rocclr_copyBuffer): FETCH_SIZE "
                            f"{cf[1] / cf[0]:.0f} KiB, WRITE_SIZE {cw[1] / cw[0]:.0f} KiB per dispatch"),
            "kernels": kernels}


def pmc_util_json(mfma_csv, tcc_csv, workload):
    """Per-kernel matrix-core utilisation and L2 hit rate of two rocprofv3
    --pmc runs: (1) MfmaUtil (rocprofv3's derived counter: MFMA busy cycles
    over the active cycles of all SIMDs) with the SQ_INSTS_VALU_MFMA_MOPS_F32
    and _F64 op counts, (2) TCC_HIT_sum and TCC_MISS_sum.  A kernel's
    utilisation is the duration-weighted mean over its dispatches."""
    import csv

    def load(path):
        out = {}
        for r in csv.DictReader(open(path)):
            k = short_kernel_name(r["Kernel_Name"])
            did = r.get("Dispatch_Id") or r.get("Correlation_Id")
            d = out.setdefault(k, {}).setdefault(did, {})
            if r.get("Start_Timestamp") and r.get("End_Timestamp"):
                d["_ns"] = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        return out

    m, t = load(mfma_csv), load(tcc_csv)
    kernels = {}
    for k, disp in m.items():
        f32 = 512.0 * sum(d.get("SQ_INSTS_VALU_MFMA_MOPS_F32", 0.0) for d in disp.values())
        f64 = 512.0 * sum(d.get("SQ_INSTS_VALU_MFMA_MOPS_F64", 0.0) for d in disp.values())
        if f32 + f64 == 0:
            continue
        ns = sum(d.get("_ns", 0.0) for d in disp.values())
        e = {"dispatches": len(disp), "mfma_flops_f32": f32, "mfma_flops_f64": f64}
        if ns > 0:
            e["avg_us"] = ns / 1e3 / len(disp)
            e["mfma_tflops"] = (f32 + f64) / ns / 1e3
            e["mfma_util_pct"] = sum(d.get("MfmaUtil", 0.0) * d.get("_ns", 0.0)
                                     for d in disp.values()) / ns
        tc = t.get(k)
        if tc:
            hit = sum(d.get("TCC_HIT_sum", 0.0) for d in tc.values())
            miss = sum(d.get("TCC_MISS_sum", 0.0) for d in tc.values())
            if hit + miss > 0:
                e["l2_hit_pct"] = 100.0 * hit / (hit + miss)
        kernels[k] = e
    return {"workload": workload, "kernels": kernels}


def copy_calibration():
    """1 GiB device-to-device copy, three times: the reference kernel for
    scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": x.numel() * 8, "copies": 3}))


def _oracle_boxes(o):
    """Per level and kind (0 cells, 1 faces): list of (r, s) of the oracle's boxes."""
    import ctypes
    from oracle import hifo
    L = hifo.lib()
    sz = np.zeros(4, np.int32)
    out = []
    for lv in range(o.num_levels):
        per = []
        for kind in (0, 1):
            nb = L.hifo_num_boxes(o._h, lv, kind)
            boxes = []
            for b in range(max(nb, 0)):
                L.hifo_box_sizes(o._h, lv, kind, b, sz.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
                boxes.append((int(sz[1]), int(sz[2])))
            per.append(boxes)
        out.append(per)
    return out


def _gpu_boxes(f):
    """Per level and kind: list of (r, s) of a factor object's boxes."""
    import ctypes
    import hifde3d_amd as H
    L = H.lib()
    sz = np.zeros(2, np.int32)
    out = []
    for lv in range(f.num_levels):
        per = []
        for kind in (0, 1):
            boxes = []
            for b in range(f.num_boxes(lv, kind)):
                L.hifde_box_sizes(f._h, lv, kind, b, sz.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
                boxes.append((int(sz[0]), int(sz[1])))
            per.append(boxes)
        out.append(per)
    return out


def _work(boxes):
    """Dense-work proxies of one level: cells sum(r^3 + r^2 s + r s^2) (LU,
    explicit inverse, X and S), faces sum(n^3) with n = r + s (pivoted QR and
    the transform), the level-matrix update sum(s^2) (entries added); and the
    apply's sum(r^2 + 2 r s) (A22^{-1} once, X or W twice)."""
    cells, faces = boxes
    wc = sum(r ** 3 + r * r * s + r * s * s for r, s in cells)
    wf = sum((r + s) ** 3 for r, s in faces)
    wu = sum(s * s for r, s in cells) + sum((r + s) ** 2 for r, s in faces)
    wa = sum(r * r + 2 * r * s for r, s in cells + faces)
    return wc, wf, wu, wa, len(cells), len(faces)


def _oracle_pcg(o, A, b, tol, maxit=300):
    x = np.zeros_like(b)
    rr = b.copy()
    z = o.apply(rr, True)
    p = z.copy()
    rz = rr @ z
    bn = np.linalg.norm(b)
    it = 0
    while np.linalg.norm(rr) / bn > tol and it < maxit:
        it += 1
        q = A @ p
        al = rz / (p @ q)
        x += al * p
        rr -= al * q
        if np.linalg.norm(rr) / bn <= tol:
            break
        z = o.apply(rr, True)
        rz2 = rr @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return it


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_run(n, seed, corr, eps, P, tol, threads):
    """The oracle (FP64 CPU restatement, its own dense kernels, OpenMP over
    the boxes of a level) on the same workload family at grid n: factor and
    oracle-preconditioned CG to tol, with per-level times and box sizes."""
    import scipy.sparse as sp
    from oracle import hifo
    a, V, b = make_problem(n, seed, corr)
    N = n - 1
    r, c, v = hifo.stencil_coo(N, a, V, True)
    A = sp.csr_matrix((v, (r, c)), shape=(n ** 3,) * 2)
    o = hifo.Oracle(N, P, eps)
    o.set_matrix_coo(r, c, v)
    t0 = time.perf_counter()
    o.factor(threads)
    t_f = time.perf_counter() - t0
    t0 = time.perf_counter()
    it = _oracle_pcg(o, A, b, tol)
    t_s = time.perf_counter() - t0
    return {"n": n, "threads": threads, "factor_s": t_f, "solve_s": t_s, "pcg_iters": it,
            "level_s": [list(t) for t in o.level_times()], "boxes": _oracle_boxes(o)}


def estimate_from(run, target_boxes, target_iters, threads):
    """Level-by-level estimate of the oracle's factor and solve time on the
    target grid from a measured run: each target level takes the times of
    the sample level with the same box width (the sample's top level for the
    levels above it), scaled by the ratio of dense-work proxies and of the
    boxes that run in parallel (OpenMP over boxes: min(boxes, threads)); the
    level-matrix update is serial.  The solve is the target's PCG iteration
    count times the sample's time per iteration scaled by the apply's work."""
    sb = run["boxes"]
    Ls = len(sb)
    sw = [_work(b_) for b_ in sb]

    def src(lt, idx):
        # the sample level of the same box width, or the highest below it
        # that has work of this kind (the root has no faces and no update)
        for ls in range(min(lt, Ls - 1), -1, -1):
            if sw[ls][idx] > 0:
                return ls
        return None

    est_levels = []
    for lt, tb in enumerate(target_boxes):
        wc_t, wf_t, wu_t, _, nc_t, nf_t = _work(tb)
        e = 0.0
        ls = src(lt, 0)
        if wc_t and ls is not None:
            e += run["level_s"][ls][0] * (wc_t / sw[ls][0]) * \
                (min(sw[ls][4], threads) / max(min(nc_t, threads), 1))
        ls = src(lt, 1)
        if wf_t and ls is not None:
            e += run["level_s"][ls][1] * (wf_t / sw[ls][1]) * \
                (min(sw[ls][5], threads) / max(min(nf_t, threads), 1))
        ls = src(lt, 2)
        if wu_t and ls is not None:
            e += run["level_s"][ls][2] * (wu_t / sw[ls][2])
        est_levels.append(e)
    wa_t = sum(_work(tb)[3] for tb in target_boxes)
    wa_s = sum(_work(b_)[3] for b_ in sb)
    per_it = run["solve_s"] / max(run["pcg_iters"], 1)
    est_solve = target_iters * per_it * wa_t / max(wa_s, 1)
    return sum(est_levels), est_solve, est_levels


def cpu_baseline(args, target_boxes, target_iters, n_target, seed, corr, eps, P, tol):
    """The CPU baseline: the oracle timed on this host on a bounded sample of
    the workload (grid --cpu-n, one thread and all cores), and the n_target
    time estimated level by level from each run (estimate_from)."""
    threads_all = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    runs = [cpu_run(args.cpu_n, seed, corr, eps, P, tol, 1)]
    if threads_all > 1:
        runs.append(cpu_run(args.cpu_n, seed, corr, eps, P, tol, threads_all))
    out_runs = []
    for r_ in runs:
        ef, es, el = estimate_from(r_, target_boxes, target_iters, r_["threads"])
        out_runs.append({"n": r_["n"], "threads": r_["threads"], "factor_s": r_["factor_s"],
                         "solve_s": r_["solve_s"], "pcg_iters": r_["pcg_iters"],
                         "level_s": r_["level_s"],
                         "estimate_n_target": {"factor_s": ef, "solve_s": es,
                                               "level_factor_s": el}})
    best = min(out_runs, key=lambda x: x["estimate_n_target"]["factor_s"] +
               x["estimate_n_target"]["solve_s"])
    e = best["estimate_n_target"]
    step = e["factor_s"] + e["solve_s"]
    return {"value": 1.0 / step, "unit": "factor+solve/s", "cores": best["threads"],
            "kind": "port",
            "sample": (f"oracle (FP64 CPU restatement with its own dense kernels, OpenMP over "
                       f"the boxes of a level) on this host ({_cpu_model()}), measured at "
                       f"n={args.cpu_n} with 1 and {threads_all} threads: factor + "
                       f"oracle-preconditioned CG to {tol:g}; value is the n={n_target} "
                       f"ESTIMATE built level by level from the {best['threads']}-thread run "
                       f"(per-level times scaled by dense-work proxies of the measured box "
                       f"sizes and by box parallelism; solve = {target_iters} iterations x the "
                       f"sample's time per iteration scaled by the apply's work)"),
            "cpu_model": _cpu_model(), "threads_all": threads_all,
            "estimate": True, "runs": out_runs}


def cpu_validate(args):
    """Checks the level-by-level estimate against a measurement: the oracle
    at grid --cpu-n and at grid --cpu-validate (all cores), the larger time
    estimated from the smaller run with the larger run's own box sizes."""
    seed, corr = 2, 0.1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    small = cpu_run(args.cpu_n, seed, corr, args.eps, args.P, args.tol, threads)
    big = cpu_run(args.cpu_validate, seed, corr, args.eps, args.P, args.tol, threads)
    ef, es, el = estimate_from(small, big["boxes"], big["pcg_iters"], threads)
    out = {"cpu_model": _cpu_model(), "threads": threads,
           "sample": {k: small[k] for k in ("n", "factor_s", "solve_s", "pcg_iters", "level_s")},
           "measured": {k: big[k] for k in ("n", "factor_s", "solve_s", "pcg_iters", "level_s")},
           "estimate": {"factor_s": ef, "solve_s": es, "level_factor_s": el},
           "ratio_estimate_over_measured": {"factor": ef / big["factor_s"],
                                            "solve": es / big["solve_s"]}}
    print(json.dumps(out), flush=True)


def batch_workload(args):
    """48 independent problems, NC drawn from {24,...,96} (seed 3), a = 10^s with
    s ~ U[-2, 2], 16 right-hand sides each: mixed-precision factor and a block
    of 16 PCG solves to 1e-10 per problem.

    --batch-mode packed (default): the problems of each grid size are packed,
    up to eight at a time, as the octants of one factor object
    (hifde_create_batch), so each level of their factorizations runs in shared
    launches, and their 16 right-hand sides each go through one block of
    PCG solves with per-(problem, column) recurrences.  --batch-mode separate:
    one factor object per problem.  Either way --batch-threads host threads
    take the objects (largest first, to the least loaded thread), each object
    on its own stream.  Coefficients, operators and right-hand sides are
    generated and assembled on the GPU before the timed region; a step is one
    pass over all 48 problems."""
    import threading
    import torch
    import hifde3d_amd as H
    from hifde3d_amd import problems as pb

    inst = pb.batch_instances()
    R = 16
    units = []  # (object, B, X, nprob, NC)
    if args.batch_mode == "packed":
        by_nc = {}
        for p in inst:
            by_nc.setdefault(p["NC"], []).append(p)
        for NC, ps in sorted(by_nc.items()):
            for c0 in range(0, len(ps), 8):
                chunk = ps[c0:c0 + 8]
                coeffs = [H.device_problem(NC, "pow10", p["coeff_seed"], lo=-2.0, hi=2.0)
                          for p in chunk]
                f = H.HIFBatch(NC - 1, chunk[0]["P"], 1e-4, len(chunk), precision="mixed")
                f.set_stencil_batch_device([a.data_ptr() for a, _ in coeffs],
                                           [V.data_ptr() for _, V in coeffs], True)
                del coeffs
                Bs = torch.empty((len(chunk) * R, NC ** 3), dtype=torch.float64, device="cuda")
                for o, p in enumerate(chunk):
                    H.gen_rhs_normal(Bs[o * R].data_ptr(), NC, p["rhs_seed"], nrhs=R)
                Bp = torch.empty((R, f.size), dtype=torch.float64, device="cuda")
                f.pack(Bs.data_ptr(), Bp.data_ptr(), R)
                units.append((f, Bp, torch.zeros_like(Bp), len(chunk), NC))
    else:
        for p in inst:
            NC = p["NC"]
            a, V = H.device_problem(NC, "pow10", p["coeff_seed"], lo=-2.0, hi=2.0)
            f = H.HIFFactor(NC - 1, p["P"], 1e-4, precision="mixed")
            f.set_stencil_device(a.data_ptr(), V.data_ptr(), True)
            B = torch.empty((R, NC ** 3), dtype=torch.float64, device="cuda")
            H.gen_rhs_normal(B.data_ptr(), NC, p["rhs_seed"], nrhs=R)
            units.append((f, B, torch.zeros_like(B), 1, NC))
    nthr = max(1, min(args.batch_threads, len(units)))
    for f, *_ in units:
        f.set_option("host_threads", max(1, 16 // nthr))
    groups = [[] for _ in range(nthr)]
    load = [0] * nthr
    for i in sorted(range(len(units)), key=lambda i: -units[i][3] * units[i][4] ** 3):
        g = load.index(min(load))
        groups[g].append(i)
        load[g] += units[i][3] * units[i][4] ** 3
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream() for _ in range(nthr)]

    def run_group(gi, out):
        with torch.cuda.stream(streams[gi]):
            for i in groups[gi]:
                f, b, x, _, _ = units[i]
                f.Factor()
                x.zero_()
                it, rr = f.pcg_multi_device(b.data_ptr(), x.data_ptr(), R, tol=args.tol,
                                            maxit=500)
                out.append((float(np.max(rr)), int(np.max(it))))
            streams[gi].synchronize()

    def one_pass():
        outs = [[] for _ in groups]
        th = [threading.Thread(target=run_group, args=(g, outs[g])) for g in range(nthr)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        torch.cuda.synchronize()
        res = [r for o in outs for r in o]
        if len(res) != len(units):
            raise RuntimeError("batch: a worker thread failed")
        return max(r[0] for r in res), [r[1] for r in res]

    for w in range(args.warmup):
        t0 = time.perf_counter()
        worst, its = one_pass()
        log(f"[bench] batch warmup {w}: {time.perf_counter() - t0:.3f}s worst relres {worst:.2e}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        worst, its = one_pass()
        log(f"[bench] batch step {k}: worst relres {worst:.2e} max iters {max(its)}")
    elapsed = time.perf_counter() - t0
    dofs = sum((p["NC"] - 1) ** 3 for p in inst)
    out = {
        "metric": "batched uneven instances: factor + 16-RHS PCG to 1e-10, problems/s",
        "value": len(inst) * args.steps / elapsed, "unit": "problems/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (generated on the GPU)",
        "config": {"workload": ("48 problems, NC from {24,32,40,48,64,80,96} (seed 3), P so that "
                                "NC/P is a power of two, a = 10^U[-2,2], eps=1e-4, 16 RHS "
                                "N(0,1) each, mixed factor + block PCG to 1e-10"),
                   "problems": len(inst), "nrhs": R, "tol": args.tol,
                   "mode": args.batch_mode, "factor_objects": len(units),
                   "host_threads": nthr},
        "unknowns": dofs, "rhs_per_s": R * len(inst) * args.steps / elapsed,
        "worst_relres": worst, "max_pcg_iters": max(its),
    }
    for f, *_ in units:
        f.close()
    print(json.dumps(out), flush=True)


def c5_workload(args):
    """The large-leaf config at full size: NC=256 (16.6M unknowns), P=8,
    eps=1e-3, lognormal a (corr 0.1, seed 4), reaction term V ~ U[0,1]
    (seed 40), mixed precision; a block of --c5-rhs N(0,1) right-hand sides
    solved together by PCG to 1e-10.  The first factorization sizes the
    workspaces and is not timed; the second is.  The workspaces (larger than
    the factor at this size) are then released so that the block PCG fits.
    A step is one factorization plus one block solve."""
    import torch
    import hifde3d_amd as H
    NC, P, eps, R = 256, 8, 1e-3, args.c5_rhs
    t0 = time.perf_counter()
    # coefficients generated and the operator assembled on the GPU
    a, V = H.device_problem(NC, "lognormal", 4, V="uniform", v_seed=40)
    f = H.HIFFactor(NC - 1, P, eps, precision="mixed")
    if args.bf16x3:
        f.set_option("bf16x3", 1)
    f.set_option("keep_a22", args.keep_a22)
    f.set_stencil_device(a.data_ptr(), V.data_ptr(), True)
    torch.cuda.synchronize()
    log(f"[bench] c5 problem set up on the GPU in {time.perf_counter() - t0:.1f}s")
    assert f.Factor() == 0
    f.set_stencil_device(a.data_ptr(), V.data_ptr(), True)
    del a, V
    torch.cuda.synchronize()
    t = time.perf_counter()
    assert f.Factor() == 0
    torch.cuda.synchronize()
    factor_s = time.perf_counter() - t
    counts = f.level_counts()
    fbytes = f.factor_bytes()
    f.set_option("keep_workspaces", 0)
    log(f"[bench] c5 factor {factor_s:.3f}s, factor {fbytes / 2**30:.1f} GiB, "
        f"free {torch.cuda.mem_get_info()[0] / 2**30:.1f} GiB")
    B = torch.empty((R, NC ** 3), dtype=torch.float64, device="cuda")
    H.gen_rhs_normal(B.data_ptr(), NC, 41, nrhs=R)
    X = torch.zeros_like(B)
    solve = []
    for rep in range(1 + args.steps):
        X.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        its, rr = f.pcg_multi_device(B.data_ptr(), X.data_ptr(), R, tol=args.tol, maxit=500)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        log(f"[bench] c5 block PCG {rep}: {dt:.3f}s, max iters {max(its)}, max relres {max(rr):.2e}")
        if rep:
            solve.append(dt)
    solve_s = sum(solve) / len(solve)
    out = {
        "metric": f"NC=256 P=8 factor + {R}-RHS block PCG to 1e-10, steps/s",
        "value": 1.0 / (factor_s + solve_s), "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": 1e3 * (factor_s + solve_s),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3+f32+f64" if args.bf16x3 else "f32+f64",
        "data": "synthetic (generated on the GPU)",
        "config": {"workload": ("NC=256 (255^3 unknowns), P=8, eps=1e-3, lognormal a (corr 0.1, "
                                "seed 4), V ~ U[0,1] (seed 40), symmetric 7-pt stencil, mixed "
                                f"precision{' with bf16x3 Schur products' if args.bf16x3 else ''}, "
                                f"{R} RHS N(0,1) (seed 41), block PCG to 1e-10"),
                   "NC": NC, "P": P, "eps": eps, "nrhs": R, "tol": args.tol},
        "factor_s": factor_s, "solve_s": solve_s, "rhs_per_s": R / solve_s,
        "pcg_iters_max": int(max(its)), "relres_max": float(max(rr)),
        "factor_bytes": fbytes, "level_counts": counts,
    }
    f.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--P", type=int, default=4)
    ap.add_argument("--eps", type=float, default=1e-4)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gemm-tile64", type=int, default=0, choices=(0, 1, 2),
                    help="factorization option gemm_tile64 (64 x 64 matrix-core tiles)")
    ap.add_argument("--host-threads", type=int, default=0,
                    help="OpenMP threads of the factorization's host planning (0: the default)")
    ap.add_argument("--bf16x3", action="store_true",
                    help="FP32 levels' Schur products as bf16x3 split products")
    ap.add_argument("--cpu-n", type=int, default=32)
    ap.add_argument("--cpu-validate", type=int, default=0, metavar="N",
                    help="time the oracle at --cpu-n and at N (all cores) and compare the "
                         "level-by-level estimate of N's time with the measurement (CPU only)")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--batch-threads", type=int, default=8,
                    help="batch workload: host threads driving the factor objects concurrently")
    ap.add_argument("--batch-mode", choices=("packed", "separate"), default="packed",
                    help="batch workload: up to 8 problems of a grid size per factor object "
                         "(packed, shared launches) or one object per problem")
    ap.add_argument("--workload", choices=("n128", "batch", "c5"), default="n128",
                    help="n128: the headline factor + PCG at n=128 (default); batch: the 48 "
                         "independent uneven problems with 16 right-hand sides each; c5: "
                         "NC=256, P=8, eps=1e-3 with a reaction term and a block of RHS")
    ap.add_argument("--c5-rhs", type=int, default=64, help="c5 workload: right-hand sides")
    ap.add_argument("--keep-a22", type=int, default=0, choices=(0, 1),
                    help="factor option keep_a22: 0 (default) stores only what F^-1 needs "
                         "(the solves never read A22); 1 keeps A22 beside its inverse")
    ap.add_argument("--pmc-util", nargs=2, metavar=("MFMA", "TCC"),
                    help="write the per-kernel matrix-core utilisation and L2 hit rate of two "
                         "rocprofv3 --pmc counter_collection.csv files to "
                         "profiles/pmc_mfma_n128.json")
    ap.add_argument("--pmc-json", nargs=4, metavar=("FETCH", "WRITE", "CALFETCH", "CALWRITE"),
                    help="write the per-kernel HBM traffic summary of four rocprofv3 --pmc "
                         "counter_collection.csv files to profiles/pmc_traffic_n128.json")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.pmc_json:
        out = pmc_json(*args.pmc_json, workload="bench.py --steps 1 --warmup 0 --no-cpu --no-fp64 "
                       "(one factor + one PCG solve, n=128 mixed)")
        json.dump(out, open(PMC_FILE, "w"), indent=1, sort_keys=True)
        print(f"wrote {PMC_FILE}: {len(out['kernels'])} kernels")
        return
    if args.pmc_util:
        out = pmc_util_json(*args.pmc_util, workload="bench.py --steps 1 --warmup 0 --no-cpu "
                            "--no-fp64 (one factor + one PCG solve, n=128 mixed)")
        json.dump(out, open(PMC_UTIL_FILE, "w"), indent=1, sort_keys=True)
        print(f"wrote {PMC_UTIL_FILE}: {len(out['kernels'])} kernels")
        return
    if args.cpu_validate:
        cpu_validate(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.workload == "batch":
        batch_workload(args)
        return
    if args.workload == "c5":
        c5_workload(args)
        return

    import torch
    import hifde3d_amd as H

    n, P, eps, tol = args.n, args.P, args.eps, args.tol
    seed, corr = 2, 0.1
    N = n - 1
    log(f"[bench] problem n={n} P={P} eps={eps}")
    a, V, b = make_problem(n, seed, corr)
    dev = torch.device("cuda:0")
    d_b = torch.from_numpy(b).to(dev)
    d_x = torch.zeros_like(d_b)

    f = H.HIFFactor(N, P, eps, precision="mixed")
    f.set_option("keep_a22", args.keep_a22)
    if args.bf16x3:
        f.set_option("bf16x3", 1)
    if args.host_threads:
        f.set_option("host_threads", args.host_threads)
    if args.gemm_tile64:
        f.set_option("gemm_tile64", args.gemm_tile64)
    f.set_stencil(a, V, symmetric=True)
    torch.cuda.synchronize()

    def step(fac):
        t0 = time.perf_counter()
        info = fac.Factor()
        t1 = time.perf_counter()
        d_x.zero_()
        torch.cuda.synchronize()
        it, rr = fac.pcg_device(d_b.data_ptr(), d_x.data_ptr(), tol=tol, maxit=500)
        t2 = time.perf_counter()
        return info, t1 - t0, t2 - t1, it, rr

    # The first warmup step is the cold start (it also builds the level-0
    # index, which depends on the grid only).  After the warmup one extra
    # step runs fully profiled (every kernel family and launch) and names the
    # dominant kernel; the timed steps then time only that kernel's
    # launches, which keeps the event overhead out of the timed region while
    # its roofline is still measured there.
    cold = None
    for w in range(args.warmup):
        info, tf, ts, it, rr = step(f)
        if w == 0:
            cold = {"factor_s": tf, "solve_s": ts}
        log(f"[bench] warmup {w}: factor {tf:.3f}s solve {ts:.3f}s iters {it} relres {rr:.2e}")
    fam_stats = None
    if args.warmup > 0:
        f.set_profiling(True)
        step(f)
        fam_stats = f.kernel_stats()
        kname, _, _ = roofline_from_stats(fam_stats)
        f.set_profiling(True, kernel=kname)
    else:
        f.set_profiling(True)
    f.reset_kernel_stats()
    torch.cuda.synchronize()
    tf_l, ts_l, it_l, rr_l = [], [], [], []
    t_start = time.perf_counter()
    for k in range(args.steps):
        info, tf, ts, it, rr = step(f)
        tf_l.append(tf); ts_l.append(ts); it_l.append(it); rr_l.append(rr)
        log(f"[bench] step {k}: factor {tf:.3f}s solve {ts:.3f}s iters {it} relres {rr:.2e}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    stats = f.kernel_stats()
    if fam_stats is None:
        fam_stats = stats
    stages = f.stage_times()
    level_stages = f.level_stage_times()
    counts = f.level_counts()
    x_mixed = d_x.cpu().numpy()
    kname, avg_launch, roof = roofline_from_stats(stats)
    log(f"[bench] dominant kernel {kname}: avg launch {avg_launch * 1e3:.3f} ms")
    log("[bench] kernel families of the profiled step:")
    for k_, v_ in sorted(fam_stats.items(), key=lambda kv: -kv[1]["sec"]):
        sec_ = max(v_["sec"], 1e-12)
        log(f"[bench]   {k_:40s} launches {v_['launches']:6d} dev {v_['sec']:.4f}s "
            f"{v_['flops'] / sec_ / 1e12:7.2f} TF/s {v_['bytes'] / sec_ / 1e9:8.1f} GB/s")
    fbytes_mixed = f.factor_bytes()
    target_boxes = _gpu_boxes(f) if not args.no_cpu else None
    f.close()

    fp64 = None
    if not args.no_fp64:
        g = H.HIFFactor(N, P, eps, precision="fp64")
        g.set_option("keep_a22", args.keep_a22)
        g.set_stencil(a, V, symmetric=True)
        for w in range(args.warmup):
            step(g)
        runs = [step(g) for _ in range(args.steps)]
        tf64 = sum(r[1] for r in runs) / len(runs)
        ts64 = sum(r[2] for r in runs) / len(runs)
        it64, rr64 = runs[-1][3], runs[-1][4]
        x64 = d_x.cpu().numpy()
        err = float(np.linalg.norm(x_mixed - x64) / np.linalg.norm(x64))
        fp64 = {"factor_s": tf64, "solve_s": ts64, "pcg_iters": it64, "relres": rr64,
                "factor_bytes": g.factor_bytes(), "mixed_vs_fp64_solution_relerr": err,
                "level_counts": g.level_counts()}
        log(f"[bench] fp64: factor {tf64:.3f}s solve {ts64:.3f}s iters {it64} err {err:.2e}")
        g.close()

    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline sample n={args.cpu_n}")
        it_target = fp64["pcg_iters"] if fp64 else it_l[-1]
        cpu = cpu_baseline(args, target_boxes, it_target, n, seed, corr, eps, P, tol)
        for r_ in cpu["runs"]:
            e = r_["estimate_n_target"]
            log(f"[bench] cpu n={r_['n']} {r_['threads']} thread(s): factor {r_['factor_s']:.2f}s "
                f"solve {r_['solve_s']:.2f}s ({r_['pcg_iters']} it) -> n={n} estimate factor "
                f"{e['factor_s']:.1f}s solve {e['solve_s']:.1f}s")

    ms = 1e3 * elapsed / args.steps
    value = args.steps / elapsed
    out = {
        "metric": METRIC, "value": value, "unit": "factor+solve/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3+f32+f64" if args.bf16x3 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"HIF-DE factor (mixed: FP32 Schur/LU"
                                f"{' with bf16x3 Schur products' if args.bf16x3 else ''}, "
                                f"FP64 ID pivots) + FP64 "
                                f"PCG to {tol:g}, n={n} ({N}^3 unknowns), lognormal GRF a "
                                f"(corr {corr}, seed {seed}), symmetric 7-pt stencil, P={P}, "
                                f"eps={eps:g}, 1 RHS N(0,1)"),
                   "n": n, "P": P, "eps": eps, "tol": tol, "nrhs": 1,
                   "keep_a22": args.keep_a22},
        "factor_s": float(np.mean(tf_l)), "solve_s": float(np.mean(ts_l)),
        "solves_per_s": float(1.0 / np.mean(ts_l)), "pcg_iters": it_l, "relres": rr_l,
        "dof_per_s": float(N ** 3 / np.mean(np.array(tf_l) + np.array(ts_l))),
        "factor_bytes": fbytes_mixed, "level_counts": counts,
        "stage_s": stages,
        "level_stage_s": level_stages,
        "kernel_families": {k_: {"launches": v_["launches"], "sec": v_["sec"],
                                 "avg_ms": 1e3 * v_["sec"] / max(v_["launches"], 1)}
                            for k_, v_ in fam_stats.items()},
        "kernel_families_from": ("one profiled step after the warmup" if args.warmup > 0
                                 else "timed steps (profiled)"),
        "cold_start": cold,
        "roofline": dict(roof, kernel=kname, avg_launch_ms=avg_launch * 1e3,
                         launches=stats["kernel:" + kname]["launches"]),
        "fp64_gpu": fp64,
        "cpu_baseline": cpu,
    }
    if fp64:
        out["speedup_mixed_vs_fp64"] = (fp64["factor_s"] + fp64["solve_s"]) / (
            out["factor_s"] + out["solve_s"])
    if cpu:
        out["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
