"""Benchmark of the MI355X mesh-optimisation inner loop.

Headline (value): FP64 AMIPS tet-evals/s -- energy + gradient + Hessian
w.r.t. vertex 0 for every tet of a jittered Kuhn lattice of 119^3 cubes
(10,110,954 tets, 1,728,000 vertices, jitter N(0, 0.15 h), seed 0), inputs
resident in HBM, one launch of amips_eval_kernel<2> per step.

Secondary measurements in the same JSON line ("secondary"):
  amips_mixed    the same lattice through the mixed FP32/FP64 kernel
  amips_packed_hessian  the same with the Hessian returned as its upper
                 triangle (TWG_AMIPS_EGH_SYM), FP64 and mixed
  envelope       isFaceOutEnvelop on 20M query triangles against a 1,003,520-facet
                 bumpy sphere (eps = sampling_dist = 1e-3), FP64 and mixed
  mixed_slivers  20M tets with aspect ratio 10^U[0,4] (mixed vs FP64 energies:
                 fallback rate, max error, decision mismatches) and 20M points
                 within +-1% of eps of the surface (point envelope checks)
  smoothing      200k one-ring Newton problems (ring sizes 4..200, median 24,
                 20% surface vertices on the bumpy sphere), autotuned tiles,
                 one persistent launch replayed from a hipGraph, FP64 and mixed,
                 against one launch per problem
  collapse       collapseAnEdge's checks for 1M directed-edge candidates of a
                 jittered 80^3 Kuhn lattice carrying an open surface sheet,
                 autotuned tiles, the batch replayed from a hipGraph, FP64 and
                 mixed
  config1        the end-to-end workload's icosphere (2562 v / 5120 f, eps 1e-3
                 diag) in a tet mesh through the C++ mirrors: smoothing,
                 collapse and removal checks, reductions, envelope
  reduce         the energy reductions (twg_energy_stats) over 10.1M stored
                 qualities, the bulk quality pass (twg_tet_qualities, FP64
                 and mixed) and isFlip (twg_tet_flip) over the 10.1M lattice
                 tets, and the configuration check (twg_tet_groups) over
                 2.5M groups of four tets
  amips_sequence (--only sequence; not in the default run) per-launch times
                 of the headline kernel in back-to-back sequences started on
                 an idle GPU, beside torch's copy and fill kernels
Each carries its CPU baseline (the reference's compiled AMIPS functions, or
the CPU restatement, over every host thread OpenMP is given, with the
one-core rate beside it; bounded samples), its parity counts against that
restatement and, for the envelope and the smoothing line search, the counts
of decisions resting on a near-tie.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--quick]
                  [--only amips,envelope,slivers,smoothing,collapse,reduce,config1,sequence]

Prints ONE JSON line.  The roofline object covers the headline kernel:
algorithmic bytes per launch / average launch time from HIP events on the
launch stream over the timed steps; roofline.sequence holds the median launch
of a further 300 back-to-back launches (the steady state: on an idle GPU the
first ~70 launches of this kernel run up to 40 % slower).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "AMIPS tet-evals/s + envelope checks/s, mixed vs FP64, decisions matched"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# FLOPs per tet-eval of the closed form as implemented (energy 86 incl. the
# four Newton steps of the cube root, gradient 32, Hessian 46)
FLOPS_PER_EVAL = 164
# bytes per tet-eval: 16 (ids) + 8 (E) + 24 (grad) + 72 (Hessian); + 24 per vertex
BYTES_PER_TET = 120
BYTES_PER_VERTEX = 24
# envelope: 72 B query + 1 B result; 48 B per box and 72 B per facet tested
ENV_QUERY_BYTES, ENV_BOX_BYTES, ENV_LEAF_BYTES = 73, 48, 72


def host_cpu():
    """the host threads the CPU baseline uses (OMP_NUM_THREADS when set, else
    every CPU this process may run on) and the CPU model"""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(n, int(omp)) if omp.isdigit() and int(omp) > 0 else n
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": threads, "nproc": n, "model": model}


def cpu_entry(value, value_1core, unit, kind, sample, host):
    """a cpu_baseline object: all host threads, with the one-core rate beside it"""
    return {"value": value, "unit": unit, "cores": host["threads"], "kind": kind,
            "sample": sample, "value_1core": value_1core, "nproc": host["nproc"],
            "cpu_model": host["model"]}


L2_PEAK_GBS = 8 * 2048 * 2.4  # 8 XCDs x 2 KiB per clock (CDNA3 figure) x 2.4 GHz
L2_REQUEST_BYTES = 128        # an L2 (TCC) request moves at most one 128-byte line


def counters(kernel):
    """per-launch counter values of `kernel` measured by rocprofv3 --pmc on
    this workload (profiles/kernel_counters.json, written from the runs that
    profiles/kernel_counters.md lists), or None"""
    path = os.path.join(ROOT, "profiles", "kernel_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("kernels", {}).get(kernel)


def counted_roofline(c, seconds):
    """rates of a launch of `seconds` from its counted work: issued FP64 /
    FP32 lane-FLOPs (VALU FLOP counters x 64 lanes: inactive lanes of an
    issued instruction count too), HBM-side traffic (FETCH_SIZE + WRITE_SIZE,
    calibrated on a plain copy) and L2 requests"""
    if c is None:
        return {}
    return {
        "counted_fp64_flops_per_launch": c["fp64_flops"],
        "counted_fp64_tflops": c["fp64_flops"] / seconds / 1e12,
        "counted_fp64_frac": c["fp64_flops"] / seconds / 1e12 / FP64_PEAK_TFLOPS,
        "counted_fp32_tflops": c["fp32_flops"] / seconds / 1e12,
        "valu_lane_utilisation": c["valu_lane_util"],
        "traffic": c["fabric_bytes"],
        "hbm_side_gbs": c["fabric_bytes"] / seconds / 1e9,
        "l2_gbs": c["l2_requests"] * L2_REQUEST_BYTES / seconds / 1e9,
        "l2_peak_gbs": L2_PEAK_GBS,
        "l2_hit_rate": c["l2_hit_rate"],
        "counters": "profiles/kernel_counters.json",
    }


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def events_time(fn, steps, warmup):
    """(wall seconds per step, GPU seconds per step from HIP events on the
    current stream); barrier + synchronize on both sides of the timed region"""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        fn()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall / steps, ev0.elapsed_time(ev1) / 1e3 / steps


def launch_sequence(fn, reps):
    """GPU seconds of each of `reps` back-to-back launches of fn (HIP events
    around every launch on the current stream)"""
    import numpy as np
    import torch
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    torch.cuda.synchronize()
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    return np.array([a.elapsed_time(b) / 1e3 for a, b in ev])


def per_call_rate(call, calls, warm=20):
    """calls/s of call(i), one C-ABI call for item i, with the device
    synchronised after each: a caller that needs each result before its next
    call, as the reference's per-element functions are used"""
    import torch
    for i in range(warm):
        call(i)
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(calls):
        call(i)
        torch.cuda.synchronize()
    return calls / (time.perf_counter() - t0)


def bits_mismatch(a, b):
    import numpy as np
    return int(np.count_nonzero(np.ascontiguousarray(a).view(np.int64) !=
                                np.ascontiguousarray(b).view(np.int64)))


# ------------------------------------------------------------- (a) --------
def bench_amips(ctx, args, result):
    import numpy as np
    import torch
    from tetwild_amd import workloads as W

    dev = ctx.device
    t0 = time.time()
    x, y, z, tets, swapped = W.kuhn_lattice(args.cubes, jitter=0.15, seed=0)
    n, nv = len(tets), len(x)
    log(f"[amips] {n} tets, {nv} vertices, {swapped} swapped ({time.time() - t0:.1f}s)")
    tx, ty, tz = (torch.from_numpy(a).to(dev) for a in (x, y, z))
    tt = torch.from_numpy(tets).to(dev)
    E = torch.empty(n, dtype=torch.float64, device=dev)
    G = torch.empty((n, 3), dtype=torch.float64, device=dev)
    H = torch.empty((n, 9), dtype=torch.float64, device=dev)

    launch = ctx.amips_launcher(tt, tx, ty, tz, 2, (E, G, H))
    wall, kern = events_time(launch, args.steps, args.warmup)
    alg_bytes = n * BYTES_PER_TET + nv * BYTES_PER_VERTEX
    achieved = alg_bytes / kern / 1e9
    roofline = {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": None,
        "kernel": "amips_eval_kernel<2>", "avg_launch_ms": kern * 1e3,
        "alg_bytes_per_launch": alg_bytes,
        "fp64_tflops": n * FLOPS_PER_EVAL / kern / 1e12,
        "fp64_frac": n * FLOPS_PER_EVAL / kern / 1e12 / FP64_PEAK_TFLOPS,
    }
    prof = os.path.join(ROOT, "profiles", "amips_egh_fp64_kuhn119_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            roofline["traffic"] = json.load(f).get("hbm_bytes_per_launch")
    c = counters("amips_eval_kernel<2>")
    if c is not None:
        roofline["counted_fp64_flops_per_launch"] = c["fp64_flops"]
        roofline["counted_fp64_tflops"] = c["fp64_flops"] / kern / 1e12
        roofline["counted_fp64_frac"] = c["fp64_flops"] / kern / 1e12 / FP64_PEAK_TFLOPS
    # A longer back-to-back sequence: some 2 ms after memory-heavy work starts
    # on an idle GPU, launches slow down by up to 40 % and recover over the
    # next ~15 ms; the median launch of the sequence is the steady state.
    seq = launch_sequence(launch, 300)
    med = float(np.median(seq))
    roofline["sequence"] = {
        "launches": len(seq), "median_launch_ms": med * 1e3,
        "mean_launch_ms": float(seq.mean()) * 1e3,
        "slowest_launch_ms": float(seq.max()) * 1e3,
        "mean_last_100_ms": float(seq[-100:].mean()) * 1e3,
        "median_achieved_gbs": alg_bytes / med / 1e9,
        "median_frac": alg_bytes / med / 1e9 / HBM_PEAK_GBS,
    }
    result.update({
        "value": n / wall, "unit": "tet-evals/s", "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "config": {"workload": f"amips_egh_fp64_kuhn{args.cubes}", "tets": n, "vertices": nv,
                   "jitter": "N(0,0.15h)", "seed": 0, "swapped_2_3": swapped},
        "roofline": roofline,
    })

    # mixed kernel on the same lattice
    Em = torch.empty(n, dtype=torch.float64, device=dev)
    Gm = torch.empty((n, 3), dtype=torch.float32, device=dev)
    Hm = torch.empty((n, 9), dtype=torch.float32, device=dev)
    fb = torch.empty(n, dtype=torch.uint8, device=dev)
    tol = 2.0**-16
    wall_m, kern_m = events_time(
        lambda: ctx.amips_eval_mixed(tt, tx, ty, tz, mode=2, tol=tol, out=(Em, Gm, Hm, fb)),
        args.steps, args.warmup)
    rel = ((Em - E).abs() / E.abs()).max().item()
    result["secondary"]["amips_mixed"] = {
        "tet_evals_per_s": n / wall_m, "avg_launch_ms": kern_m * 1e3, "tol": tol,
        "fallback_rate": fb.float().mean().item(), "max_rel_err_E_vs_fp64": rel,
        "speedup_vs_fp64": wall / wall_m,
        "alg_bytes_per_launch": n * (16 + 8 + 12 + 36 + 1) + nv * 24,
        "hbm_gbs": (n * (16 + 8 + 12 + 36 + 1) + nv * 24) / kern_m / 1e9,
    }
    # at the energy tolerance 1e-9: FP32 (unit roundoff 6e-8) certifies no
    # tet, so every tet falls back to FP64
    tol9 = 1e-9
    wall_9, kern_9 = events_time(
        lambda: ctx.amips_eval_mixed(tt, tx, ty, tz, mode=2, tol=tol9, out=(Em, Gm, Hm, fb)),
        max(2, args.steps // 4), 1)
    result["secondary"]["amips_mixed"]["at_tol_1e-9"] = {
        "tol": tol9, "tet_evals_per_s": n / wall_9, "fallback_rate": fb.float().mean().item(),
        "max_rel_err_E_vs_fp64": ((Em - E).abs() / E.abs()).max().item(),
        "speedup_vs_fp64": wall / wall_9}

    # the plain drop-in form: one tet per call (comformalAMIPS*_new are
    # per-tet functions), synchronised per call
    ctx._bind()
    L, hc = ctx.L, ctx.h
    pt, px, py, pz = tt.data_ptr(), tx.data_ptr(), ty.data_ptr(), tz.data_ptr()
    pE, pG, pH = E.data_ptr(), G.data_ptr(), H.data_ptr()

    def one_tet(i):
        r = L.twg_amips_eval(hc, pt + 16 * i, 1, px, py, pz, nv, 2, pE + 8 * i, pG + 24 * i,
                             pH + 72 * i)
        if r:
            raise RuntimeError("twg_amips_eval failed")
    per_call = per_call_rate(one_tet, 2000)
    result["secondary"]["amips_mixed"]["one_tet_per_call_fp64"] = {
        "calls": 2000, "tet_evals_per_s": per_call,
        "mixed_batched_vs_it": (n / wall_m) / per_call,
        "fp64_batched_vs_it": (n / wall) / per_call}

    # the packed Hessian (TWG_AMIPS_EGH_SYM: its upper triangle, 96 instead of
    # 120 output bytes per tet), FP64 and mixed, against the full layout
    from tetwild_amd import _capi as C
    H6 = torch.empty((n, 6), dtype=torch.float64, device=dev)
    E6 = torch.empty(n, dtype=torch.float64, device=dev)
    G6 = torch.empty((n, 3), dtype=torch.float64, device=dev)
    launch6 = ctx.amips_launcher(tt, tx, ty, tz, C.TWG_AMIPS_EGH_SYM, (E6, G6, H6))
    wall6, kern6 = events_time(launch6, args.steps, args.warmup)
    seq6 = launch_sequence(launch6, 300)
    upper = torch.tensor([0, 1, 2, 4, 5, 8], device=dev)
    mism = int((H6.view(torch.int64) != H[:, upper].view(torch.int64)).sum().item() +
               (E6.view(torch.int64) != E.view(torch.int64)).sum().item() +
               (G6.view(torch.int64) != G.view(torch.int64)).sum().item())
    alg6 = n * (BYTES_PER_TET - 24) + nv * BYTES_PER_VERTEX
    Hm6 = torch.empty((n, 6), dtype=torch.float32, device=dev)
    wallm6, kernm6 = events_time(
        lambda: ctx.amips_eval_mixed(tt, tx, ty, tz, mode=C.TWG_AMIPS_EGH_SYM, tol=tol,
                                     out=(Em, Gm, Hm6, fb)), args.steps, args.warmup)
    algm6 = n * (16 + 8 + 12 + 24 + 1) + nv * 24
    result["secondary"]["amips_packed_hessian"] = {
        "layout": "H as its upper triangle (xx, xy, xz, yy, yz, zz), TWG_AMIPS_EGH_SYM",
        "tet_evals_per_s_fp64": n / wall6, "avg_launch_ms_fp64": kern6 * 1e3,
        "sequence_median_launch_ms_fp64": float(np.median(seq6)) * 1e3,
        "alg_bytes_per_launch_fp64": alg6,
        "sequence_median_gbs_fp64": alg6 / float(np.median(seq6)) / 1e9,
        "bit_mismatch_vs_full_layout": mism,
        "tet_evals_per_s_mixed": n / wallm6, "avg_launch_ms_mixed": kernm6 * 1e3,
        "alg_bytes_per_launch_mixed": algm6,
        "speedup_fp64_vs_full_layout": wall / wall6,
    }
    del H6, E6, G6, Hm6

    if not args.no_cpu:
        log("[cpu] timing the CPU baseline sample")
        torch.cuda.synchronize()
        cpu, parity = cpu_baseline_amips(x, y, z, tets, E.cpu().numpy(), H.cpu().numpy())
        result["cpu_baseline"] = cpu
        result["parity"] = parity


def cpu_baseline_amips(x, y, z, tets, E_gpu, H_gpu, budget_s=12.0):
    """The reference's own AMIPS functions (or the CPU restatement where that
    library is absent) on one host core over a bounded sample of the same
    lattice; bit mismatches of the GPU result against the restatement."""
    import numpy as np
    from oracle import oracle as O

    n = min(len(tets), 2_000_000)
    t = tets[:n]
    T = np.ascontiguousarray(np.stack([np.stack([x[t[:, j]], y[t[:, j]], z[t[:, j]]], 1)
                                       for j in range(4)], 1).reshape(-1, 12))
    kind = "reference" if O.ref() is not None else "port"
    E = np.empty(n); G = np.empty((n, 3)); H = np.empty((n, 9))
    reps = 0
    t0 = time.perf_counter()
    while True:
        if kind == "reference":
            O.ref().ref_amips_batch_ejh(O._p(T), n, O._p(E), O._p(G), O._p(H))
        else:
            O.lib().orc_amips_batch(O._p(np.ascontiguousarray(t)), n, O._p(x), O._p(y), O._p(z),
                                    2, O._p(E), O._p(G), O._p(H))
        reps += 1
        el = time.perf_counter() - t0
        if el > budget_s or reps >= 20:
            break
    one_core = n * reps / el
    # all host threads over the same sample
    host = host_cpu()
    reps_mt = 0
    t0 = time.perf_counter()
    while True:
        if kind == "reference":
            O.ref().ref_amips_batch_ejh_mt(O._p(T), n, O._p(E), O._p(G), O._p(H), host["threads"])
        else:
            O.lib().orc_amips_batch_mt(O._p(np.ascontiguousarray(t)), n, O._p(x), O._p(y),
                                       O._p(z), 2, O._p(E), O._p(G), O._p(H), host["threads"])
        reps_mt += 1
        el_mt = time.perf_counter() - t0
        if el_mt > budget_s / 2 or reps_mt >= 40:
            break
    Eo, Go, Ho = O.amips_batch(t, x, y, z, 2)
    ref_rel = None
    if kind == "reference":
        ok = np.isfinite(E)
        ref_rel = float(np.max(np.abs(E_gpu[:n][ok] - E[ok]) / np.abs(E[ok])))
    return cpu_entry(
        n * reps_mt / el_mt, one_core, "tet-evals/s", kind,
        f"first {n} lattice tets x {reps_mt} passes ({reps} on one core), E+grad+Hessian, "
        + ("reference comformalAMIPS*_new (g++ -O3)" if kind == "reference" else "CPU restatement"),
        host), {"bit_mismatch_E": bits_mismatch(E_gpu[:n], Eo),
        "bit_mismatch_H": bits_mismatch(H_gpu[:n], Ho), "sample_tets": n,
        "max_rel_err_vs_reference_E": ref_rel}


# ------------------------------------------------------------- (b) --------
def bench_envelope(ctx, args, result, surf_cache):
    import numpy as np
    import torch
    from tetwild_amd import workloads as W

    dev = ctx.device
    nq = 2_000_000 if args.quick else 20_000_000
    eps = 1e-3
    t0 = time.time()
    V, F = surf_cache["VF"]
    Q = W.query_triangles(V, F, nq, eps=eps, edge=0.002, seed=0)
    surf = surf_cache["surf"]
    log(f"[envelope] {nq} queries on {len(F)} facets (gen {time.time() - t0:.1f}s)")
    tq = torch.from_numpy(Q).to(dev)
    out = {}
    rates = {}
    stats = None
    for prec, name in ((0, "fp64"), (1, "mixed")):
        o, st = ctx.envelope_faces(surf, tq, eps * eps, eps, precision=prec, stats=True)
        if stats is None:
            stats = st.cpu().numpy()
        buf = torch.empty(nq, dtype=torch.uint8, device=dev)
        wall, kern = events_time(
            lambda: ctx.envelope_faces(surf, tq, eps * eps, eps, precision=prec, out=buf),
            max(2, args.steps // 4), 1)
        out[name] = o.cpu().numpy()
        rates[name] = (nq / wall, kern)
    # the plain drop-in form: one triangle per call (isFaceOutEnvelop is a
    # per-triangle function), synchronised per call
    ctx._bind()
    L, hc, hs = ctx.L, ctx.h, surf.h
    pq, po = tq.data_ptr(), buf.data_ptr()

    def one_tri(i):
        r = L.twg_envelope_faces(hc, hs, pq + 72 * i, 1, eps * eps, eps, 0, po + i, None)
        if r:
            raise RuntimeError("twg_envelope_faces failed")
    per_call = per_call_rate(one_tri, 1000)
    samples, trav, boxes, leaves = (float(v) / nq for v in stats)
    alg = ENV_QUERY_BYTES + boxes * ENV_BOX_BYTES + leaves * ENV_LEAF_BYTES
    sec = {
        "queries": nq, "facets": len(F), "eps": eps, "sampling_dist": eps,
        "checks_per_s_fp64": rates["fp64"][0], "checks_per_s_mixed": rates["mixed"][0],
        "avg_launch_ms_fp64": rates["fp64"][1] * 1e3, "avg_launch_ms_mixed": rates["mixed"][1] * 1e3,
        "mixed_speedup": rates["mixed"][0] / rates["fp64"][0],
        "decision_mismatch_mixed_vs_fp64": int(np.count_nonzero(out["fp64"] != out["mixed"])),
        "out_rate": float(out["fp64"].mean()),
        "one_triangle_per_call_fp64": {"calls": 1000, "checks_per_s": per_call,
                                       "mixed_batched_vs_it": rates["mixed"][0] / per_call,
                                       "fp64_batched_vs_it": rates["fp64"][0] / per_call},
        "per_check": {"samples_tested": samples, "full_traversals": trav, "boxes": boxes,
                      "facets": leaves},
        "alg_bytes_per_check": alg,
        "alg_gbs_fp64": alg * nq / rates["fp64"][1] / 1e9,
    }
    # the dominant kernel (faces_list_kernel): HBM-side and L2 rates from its
    # counted traffic over the HIP-event time of the whole call (Morton
    # ordering included); its FP64 lane-FLOP rate
    kern64 = rates["fp64"][1]
    rl = {"bound": "latency (dependent box / facet reads)", "kernel": "faces_list_kernel<false, 80>",
          "avg_launch_ms": kern64 * 1e3, "unit": "GB/s", "peak": HBM_PEAK_GBS}
    rl.update(counted_roofline(counters("faces_list_kernel<false, 80>"), kern64))
    if "hbm_side_gbs" in rl:
        rl["achieved"] = rl["hbm_side_gbs"]
        rl["frac"] = rl["achieved"] / HBM_PEAK_GBS
    sec["roofline"] = rl
    if not args.no_cpu:
        log("[cpu] timing the CPU baseline sample")
        from oracle import oracle as O
        S = O.Surface(V, F)
        host = host_cpu()
        k = 20000 if args.quick else 200000
        t0 = time.perf_counter()
        ref, _ = S.face_out(Q[:k], eps * eps, eps)
        el = time.perf_counter() - t0
        km = 10 * k
        t0 = time.perf_counter()
        ref_mt = S.face_out_mt(Q[:km], eps * eps, eps, host["threads"])
        el_mt = time.perf_counter() - t0
        sec["cpu_baseline"] = cpu_entry(
            km / el_mt, k / el, "checks/s", "port",
            f"first {km} queries on all threads ({k} on one core), CPU restatement (tree)", host)
        sec["decision_mismatch_vs_cpu"] = int(np.count_nonzero(ref_mt != out["fp64"][:km]))
        # decisions resting on a sample within 4 ulp(eps^2) of eps^2 (the
        # reference could decide those the other way)
        band = 4 * float(np.spacing(eps * eps))
        r = S.face_risk(Q[:k], eps * eps, eps, band, host["threads"])
        sec["near_tie"] = {"queries": k, "band": band, "samples": r[0], "samples_in_band": r[1],
                           "queries_with_band_sample": r[2], "queries_decided_in_band": r[3]}
    result["secondary"]["envelope"] = sec


# --------------------------------------------------------- (a)+(b) mixed --
def bench_slivers(ctx, args, result, surf_cache):
    import numpy as np
    import torch
    from tetwild_amd import workloads as W

    dev = ctx.device
    n = 2_000_000 if args.quick else 20_000_000
    t0 = time.time()
    x, y, z, tets = W.aspect_tets(n, seed=0)
    log(f"[slivers] {n} aspect-ratio tets (gen {time.time() - t0:.1f}s)")
    tx, ty, tz = (torch.from_numpy(a).to(dev) for a in (x, y, z))
    tt = torch.from_numpy(tets).to(dev)
    E = torch.empty(n, dtype=torch.float64, device=dev)
    G = torch.empty((n, 3), dtype=torch.float64, device=dev)
    H = torch.empty((n, 9), dtype=torch.float64, device=dev)
    Em = torch.empty(n, dtype=torch.float64, device=dev)
    Gm = torch.empty((n, 3), dtype=torch.float32, device=dev)
    Hm = torch.empty((n, 9), dtype=torch.float32, device=dev)
    fb = torch.empty(n, dtype=torch.uint8, device=dev)
    tol = 2.0**-16
    steps = max(2, args.steps // 4)
    w64, k64 = events_time(lambda: ctx.amips_eval(tt, tx, ty, tz, 2, out=(E, G, H)), steps, 1)
    wm, km = events_time(lambda: ctx.amips_eval_mixed(tt, tx, ty, tz, 2, tol, out=(Em, Gm, Hm, fb)),
                         steps, 1)
    fbm = fb.bool()
    rel = ((Em - E).abs() / E.abs())
    q64, _ = ctx.tet_qualities(tt, tx, ty, tz, precision=0)
    qm, _ = ctx.tet_qualities(tt, tx, ty, tz, precision=1, tol=tol)
    sent_mm = int(((q64 == 1e50) != (qm == 1e50)).sum().item())
    sec = {
        "tets": n, "tol": tol,
        "tet_evals_per_s_fp64": n / w64, "tet_evals_per_s_mixed": n / wm,
        "mixed_speedup": w64 / wm, "fallback_rate": fbm.float().mean().item(),
        "max_rel_err_fp32_part": rel[~fbm].max().item() if (~fbm).any() else 0.0,
        "max_rel_err_fallback_part": rel[fbm].max().item() if fbm.any() else 0.0,
        "quality_sentinel_mismatch": sent_mm,
    }
    w9, _ = events_time(lambda: ctx.amips_eval_mixed(tt, tx, ty, tz, 2, 1e-9, out=(Em, Gm, Hm, fb)),
                        steps, 1)
    sec["at_tol_1e-9"] = {"tol": 1e-9, "tet_evals_per_s": n / w9, "speedup_vs_fp64": w64 / w9,
                          "fallback_rate": fb.float().mean().item(),
                          "max_rel_err_E_vs_fp64":
                              ((Em - E).abs() / E.abs())[torch.isfinite(E)].max().item()}
    del E, G, H, Em, Gm, Hm, tt, tx, ty, tz
    # point envelope samples clustered within +-1% of eps
    V, F = surf_cache["VF"]
    surf = surf_cache["surf"]
    eps = 1e-3
    P = W.envelope_samples(V, F, n, eps=eps, band=0.01, seed=0)
    tp = torch.from_numpy(P).to(dev)
    buf = torch.empty(n, dtype=torch.uint8, device=dev)
    o64 = ctx.envelope_points(surf, tp, eps * eps, precision=0).cpu().numpy()
    om = ctx.envelope_points(surf, tp, eps * eps, precision=1).cpu().numpy()
    wp64, kp64 = events_time(lambda: ctx.envelope_points(surf, tp, eps * eps, 0), steps, 1)
    wpm, kpm = events_time(lambda: ctx.envelope_points(surf, tp, eps * eps, 1), steps, 1)
    sec.update({
        "points": n, "point_checks_per_s_fp64": n / wp64, "point_checks_per_s_mixed": n / wpm,
        "point_mixed_speedup": wp64 / wpm, "point_out_rate": float(o64.mean()),
        "point_decision_mismatch_mixed_vs_fp64": int(np.count_nonzero(o64 != om)),
    })
    if not args.no_cpu:
        log("[cpu] timing the CPU baseline sample")
        from oracle import oracle as O
        S = O.Surface(V, F)
        host = host_cpu()
        k = 2_000_000
        ref = S.point_out_mt(P[:k], eps * eps, host["threads"])
        sec["point_decision_mismatch_vs_cpu"] = int(np.count_nonzero(ref != o64[:k]))
        sec["point_cpu_checked"] = k
        band = 4 * float(np.spacing(eps * eps))
        r = S.point_risk(P[:k], eps * eps, band, host["threads"])
        sec["point_near_tie"] = {"points": r[0], "band": band, "points_in_band": r[1]}
    result["secondary"]["mixed_slivers"] = sec


# ------------------------------------------------------------- (c) --------
def bench_smoothing(ctx, args, result, surf_cache):
    import numpy as np
    import torch
    from tetwild_amd import _capi as C
    from tetwild_amd import workloads as W
    from tetwild_amd.ops import SmoothBatch

    n = 20000 if args.quick else 200_000
    V, F = surf_cache["VF"]
    surf = surf_cache["surf"]
    eps = 1e-3
    t0 = time.time()
    B = W.smoothing_batch(n, seed=0, surface_frac=0.2, V=V, F=F)
    q, _ = ctx.tet_qualities(*(torch.from_numpy(a).to(ctx.device) for a in
                               (B["ring_tets"], B["x"], B["y"], B["z"])))
    old_q = q.cpu().numpy()
    log(f"[smoothing] {n} one-rings, {len(B['ring_tets'])} ring tets (gen {time.time() - t0:.1f}s)")
    batch = SmoothBatch(ctx, B, sf=surf, bt=surf, eps2=eps * eps, sampling_dist=eps, old_q=old_q)
    sec = {"problems": n, "ring_tets": int(len(B["ring_tets"])),
           "surface_fraction": float((B["kind"] > 0).mean())}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        tile, chunk, ms = batch.autotune(precision=prec)
        plan = batch.plan(precision=prec)
        wall, kern = events_time(plan.run, max(2, args.steps // 4), 1)
        split = tile == C.TWG_SMOOTH_TILE_SPLIT
        sec[f"problems_per_s_{name}"] = n / wall
        sec[f"tile_{name}"] = "split: 64 lanes per surface problem, 32 per other" if split else tile
        sec[f"chunk_{name}"] = chunk
        sec[f"graph_ms_{name}"] = kern * 1e3
        if prec == 0:
            kname = "smooth_split_kernel<false>" if split else f"smooth_kernel<{tile}, false>"
            rl = {"bound": "fp64 valu", "kernel": kname,
                  "avg_launch_ms": kern * 1e3, "unit": "TFLOP/s", "peak": FP64_PEAK_TFLOPS}
            rl.update(counted_roofline(counters(kname), kern))
            if "counted_fp64_tflops" in rl:
                rl["achieved"] = rl["counted_fp64_tflops"]
                rl["frac"] = rl["counted_fp64_frac"]
            sec["roofline"] = rl
        plan.close()
    torch.cuda.synchronize()
    st = batch.status.cpu().numpy()
    p_gpu = batch.p_out.cpu().numpy()
    sec["status_counts"] = {s: int(c) for s, c in
                            zip(["moved", "invalid", "not_moved", "flip_end", "reject_flip",
                                 "reject_quality", "reject_envelope", "bad_input"],
                                np.bincount(st, minlength=8))}
    # plain path: one launch per problem, as a per-vertex drop-in would issue
    k = 2000
    sub = {kk: (v[:k] if kk in ("center", "kind") else v) for kk, v in B.items()}
    sub["ring_off"] = B["ring_off"][:k + 1]
    sub["tri_off"] = B["tri_off"][:k + 1]
    one = [batch.problem_desc(p) for p in range(500)]
    for d in one[:20]:
        batch.run(precision=0, tile=64, chunk=1, desc=d)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for d in one:
        batch.run(precision=0, tile=64, chunk=1, desc=d)
    torch.cuda.synchronize()
    per_call = (time.perf_counter() - t0) / len(one)
    sec["problems_per_s_one_launch_each"] = 1.0 / per_call
    sec["batched_mixed_vs_one_launch_each"] = sec["problems_per_s_mixed"] * per_call
    sec["batched_fp64_vs_one_launch_each"] = sec["problems_per_s_fp64"] * per_call
    if not args.no_cpu:
        log("[cpu] timing the CPU baseline sample")
        from oracle import oracle as O
        S = O.Surface(V, F[surf.order()])
        host = host_cpu()
        t0 = time.perf_counter()
        O.smooth_batch_mt(sub, old_q, S, S, eps * eps, eps, nthreads=1)
        el = time.perf_counter() - t0
        km = 20000 if args.quick else 100_000
        sub = {kk: (v[:km] if kk in ("center", "kind") else v) for kk, v in B.items()}
        sub["ring_off"] = B["ring_off"][:km + 1]
        sub["tri_off"] = B["tri_off"][:km + 1]
        t0 = time.perf_counter()
        rp, rs, rq, risk = O.smooth_batch_mt(sub, old_q, S, S, eps * eps, eps, host["threads"])
        el_mt = time.perf_counter() - t0
        sec["cpu_baseline"] = cpu_entry(
            km / el_mt, k / el, "problems/s", "port",
            f"first {km} problems on all threads ({k} on one core), CPU restatement", host)
        sec["status_mismatch_vs_cpu"] = int(np.count_nonzero(rs != st[:km]))
        sec["position_bit_mismatch_vs_cpu"] = bits_mismatch(p_gpu[:km], rp)
        # line-search comparisons within the summation band of a tie (the
        # reference sums rings in unordered_set order)
        sec["near_tie"] = {"problems": km, "energy_compares": risk[0],
                           "energy_compares_in_band": risk[1], "threshold_tests": risk[2],
                           "threshold_tests_in_band": risk[3]}
    result["secondary"]["smoothing"] = sec


def bench_sequence(ctx, args, result):
    """Per-launch times of amips_eval_kernel<2> on the 119^3 lattice in
    back-to-back sequences started on an idle GPU (200 ms host sleep before
    each), after a 10 ms spin kernel, with a 2 ms host pause between launches,
    and of torch's copy (1.25 GB moved) and fill (1.05 GB written) kernels.
    Not part of the default run (--only sequence)."""
    import numpy as np
    import torch
    from tetwild_amd import workloads as W

    dev = ctx.device
    x, y, z, tets, _ = W.kuhn_lattice(args.cubes, jitter=0.15, seed=0)
    n, nv = len(tets), len(x)
    tx, ty, tz = (torch.from_numpy(a).to(dev) for a in (x, y, z))
    tt = torch.from_numpy(tets).to(dev)
    out = (torch.empty(n, dtype=torch.float64, device=dev),
           torch.empty((n, 3), dtype=torch.float64, device=dev),
           torch.empty((n, 9), dtype=torch.float64, device=dev))
    launch = ctx.amips_launcher(tt, tx, ty, tz, 2, out)
    alg = n * BYTES_PER_TET + nv * BYTES_PER_VERTEX
    half = alg // 16
    src = torch.ones(half, dtype=torch.float64, device=dev)
    dst = torch.empty(half, dtype=torch.float64, device=dev)
    wr = torch.empty(n * 13, dtype=torch.float64, device=dev)

    def idle():
        torch.cuda.synchronize()
        time.sleep(0.2)

    def paused(fn, reps):
        t = []
        for _ in range(reps):
            t.append(launch_sequence(fn, 1)[0])
            time.sleep(0.002)
        return np.array(t)

    for _ in range(3):
        launch()
    seqs = {}
    for r in range(3):
        idle()
        seqs[f"amips_after_idle_{r}"] = (launch_sequence(launch, 150), alg)
    idle()
    torch.cuda._sleep(20_000_000)
    seqs["amips_after_spin"] = (launch_sequence(launch, 150), alg)
    idle()
    seqs["amips_2ms_pauses"] = (paused(launch, 60), alg)
    idle()
    seqs["copy_after_idle"] = (launch_sequence(lambda: dst.copy_(src), 150), 2 * half * 8)
    idle()
    seqs["fill_after_idle"] = (launch_sequence(lambda: wr.fill_(2.0), 150), n * 104)
    res = {}
    for k, (t, nbytes) in seqs.items():
        us = np.round(t * 1e6, 1)
        res[k] = {"launches": len(t), "bytes": int(nbytes), "median_us": float(np.median(us)),
                  "mean_us": float(us.mean()), "max_us": float(us.max()),
                  "mean_launches_6_to_25_us": float(us[5:25].mean()),
                  "median_gbs": nbytes / float(np.median(t)) / 1e9, "us": us.tolist()}
    result["secondary"]["amips_sequence"] = res


# ------------------------------------------------------------ (f) --------
def bench_collapse(ctx, args, result):
    """collapseAnEdge's checks for 1M directed-edge candidates of a jittered
    80^3 Kuhn lattice (3.07M tets) carrying an open surface sheet at mid-height (eps =
    sampling_dist = 0.05 h): autotuned, one hipGraph replay of the batch per step"""
    import numpy as np
    import torch
    from tetwild_amd import workloads as W
    from tetwild_amd.ops import CollapseBatch

    cubes = 20 if args.quick else 80
    n = 100_000 if args.quick else 1_000_000
    t0 = time.time()
    M = W.sheet_mesh(cubes, seed=0)
    tx, ty, tz = (torch.from_numpy(M[k]).to(ctx.device) for k in ("x", "y", "z"))
    tt = torch.from_numpy(M["tets"]).to(ctx.device)
    q = ctx.tet_qualities(tt, tx, ty, tz)[0].cpu().numpy()
    B = W.collapse_batch(M, q, n_cand=n, seed=0)
    n = len(B["v1"])
    eps = 0.05 / cubes
    sf = ctx.surface(M["sf_V"], M["sf_F"])
    bt = ctx.surface(M["b_V"], M["b_F"])
    log(f"[collapse] {n} candidates on {len(M['tets'])} tets, {len(B['ring_tets'])} ring tets, "
        f"{len(B['tri_ids'])} epsilon triangles (gen {time.time() - t0:.1f}s)")
    batch = CollapseBatch(ctx, B, sf, bt, eps * eps, eps)
    sec = {"candidates": n, "tets": int(len(M["tets"])), "ring_tets": int(len(B["ring_tets"])),
           "eps_triangles": int(len(B["tri_ids"])),
           "v1_surface_fraction": float((B["flags"] & 1).astype(bool).mean())}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        tile, chunk, _ = batch.autotune(precision=prec)
        plan = batch.plan(precision=prec)
        wall, kern = events_time(plan.run, max(2, args.steps // 4), 1)
        sec[f"candidates_per_s_{name}"] = n / wall
        sec[f"ring_tets_per_s_{name}"] = len(B["ring_tets"]) / wall
        sec[f"graph_ms_{name}"] = kern * 1e3
        sec[f"tile_{name}"] = tile
        sec[f"chunk_{name}"] = chunk
        plan.close()
        if prec == 0:
            st0 = batch.status[:n].clone()
            nq0 = batch.new_q[:len(B["ring_tets"])].clone()
    torch.cuda.synchronize()
    st = batch.status[:n].cpu().numpy()
    nq = batch.new_q[:len(B["ring_tets"])].cpu().numpy()
    sec["status_mismatch_mixed_vs_fp64"] = int(np.count_nonzero(st != st0.cpu().numpy()))
    sec["status_counts"] = {s: int(c) for s, c in
                            zip(["success", "flip", "quality", "envelop", "envelop_suc", "-", "-",
                                 "bad_input"], np.bincount(st, minlength=8)) if s != "-"}
    del nq0
    if not args.no_cpu:
        log("[cpu] timing the CPU baseline sample")
        from oracle import oracle as O
        host = host_cpu()
        S, Sb = O.Surface(M["sf_V"], M["sf_F"]), O.Surface(M["b_V"], M["b_F"])

        def first(k):
            sub = dict(B)
            for key in ("v1", "v2", "flags"):
                sub[key] = B[key][:k]
            sub["ring_off"] = B["ring_off"][:k + 1]
            sub["tri_off"] = B["tri_off"][:k + 1]
            return sub
        k1 = 20000 if args.quick else 200_000
        t0 = time.perf_counter()
        O.collapse_batch_mt(first(k1), S, Sb, eps * eps, eps, nthreads=1)
        el = time.perf_counter() - t0
        k = n
        t0 = time.perf_counter()
        so, nqo = O.collapse_batch_mt(first(k), S, Sb, eps * eps, eps, nthreads=host["threads"])
        el_mt = time.perf_counter() - t0
        sec["cpu_baseline"] = cpu_entry(
            k / el_mt, k1 / el, "candidates/s", "port",
            f"all {k} candidates on all threads (first {k1} on one core), CPU restatement", host)
        sec["status_mismatch_vs_cpu"] = int(np.count_nonzero(so != st[:k]))
        ok = np.isin(so, [0, 4])
        rows = np.repeat(ok, np.diff(B["ring_off"][:k + 1]))
        r1 = int(B["ring_off"][k])
        sec["new_q_bit_mismatch_vs_cpu"] = bits_mismatch(nq[:r1][rows], nqo[:r1][rows])
    sf.close()
    bt.close()
    result["secondary"]["collapse"] = sec


def bench_config1(ctx, args, result):
    """The end-to-end CPU workload's surface -- icosphere subdivided 4x (2562
    vertices, 5120 triangles, radius 1) jittered radially by N(0, 0.002),
    eps = 1e-3 bbox diagonal, stage-1 eps = eps (1 - 1/sqrt 3), sampling
    distance eps, ideal edge diagonal / 20 -- embedded in a tet mesh
    (workloads.icosphere_tetmesh), through the C++ mirrors of the in-scope
    operators of the refinement loop: smooth() in vertex order and the
    batched colour-class variant, EdgeCollapser checks on every directed edge,
    EdgeRemover decisions on every interior edge of 3..5 tets, the energy
    reductions, plus the envelope test of every surface-vertex face and every
    vertex.  The reference pipeline itself (CGAL, geogram, Eigen, libigl)
    cannot be built here, so its own numbers are not measured."""
    import numpy as np
    import torch
    from oracle import oracle as O
    from tetwild_amd import driver as D
    from tetwild_amd import workloads as W

    M = W.icosphere_tetmesh(levels=4, sigma=0.002, seed=0)
    P = np.stack([M["x"], M["y"], M["z"]], 1)
    tets = M["tets"]
    tt = torch.from_numpy(tets).to(ctx.device)
    tx, ty, tz = (torch.from_numpy(M[k]).to(ctx.device) for k in "xyz")
    q = ctx.tet_qualities(tt, tx, ty, tz)[0].cpu().numpy()
    flags = (M["on_surface"] * 1 + M["on_bbox"] * 2).astype(np.uint8)
    sf = M["sf_V"][M["sf_F"]].reshape(-1, 9)
    bf = M["b_V"][M["b_F"]].reshape(-1, 9)
    d = D.tempdir()
    mesh = os.path.join(d, "config1.bin")
    D.write_mesh(mesh, P, tets, M["is_surface_fs"], flags, q, sf, bf, M["eps"], M["eps_2"],
                 M["sampling_dist"], closed=True)
    sec = {"surface_vertices": int(len(M["sf_V"])), "surface_triangles": int(len(M["sf_F"])),
           "vertices": int(len(P)), "tets": int(len(tets)), "bbox_diag": M["diag"],
           "eps_input": M["eps_input"], "eps": M["eps"], "sampling_dist": M["sampling_dist"],
           "ideal_edge": M["edge_len"], "energy_before": {"max": float(q.max()),
                                                          "avg": float(q.mean())},
           "reference": "not measurable: the reference pipeline needs CGAL, geogram, Eigen "
                        "and libigl, absent here"}
    nv, nt = len(P), len(tets)
    # seq: the reference's vertex order, one launch per dependency level;
    # serial: the same results one launch per vertex; batch: colour classes
    raws = {}
    for mode in ("seq", "serial", "batch"):
        info, out, el = D.run(mesh, mode)
        raw = np.fromfile(out, dtype=np.float64)
        raws[mode] = raw
        qn = raw[3 * nv:3 * nv + nt]
        Pn = raw[:3 * nv].reshape(nv, 3)
        sec[f"smooth_{mode}"] = {"counter": info["counter"], "suc_counter": info["suc_counter"],
                                 "moved": int((Pn != P).any(1).sum()),
                                 "energy_after": {"max": float(qn.max()), "avg": float(qn.mean())},
                                 "seconds": el, "smooth_seconds": info["smooth_seconds"]}
        if mode == "seq":
            sec["smooth_seq"]["levels"] = [info["levels_single"], info["levels_surface"]]
    sec["smooth_seq_vs_serial_bit_mismatch"] = bits_mismatch(raws["seq"], raws["serial"])
    info, out, el = D.run(mesh, "collapse")
    codes = np.fromfile(out, dtype=np.int32, count=info["candidates"])
    sec["collapse"] = {"candidates": info["candidates"], "seconds": el,
                       "codes": dict(zip(["success", "flip", "quality", "envelop", "envelop_suc"],
                                         np.bincount(codes, minlength=5).tolist()))}
    info, out, el = D.run(mesh, "remove")
    sec["remove"] = {"edges": info["edges"], "removable": info["removable"], "seconds": el}
    info, _, el = D.run(mesh, "energy")
    sec["energy_stats_tets_t_mod_7_ne_3"] = info
    # envelope decisions: faces with three surface vertices, every vertex
    surf = ctx.surface(M["sf_V"], M["sf_F"])
    on = M["on_surface"]
    faces = set()
    for j in range(4):
        f = np.sort(np.stack([tets[:, (j + k) % 4] for k in (1, 2, 3)], 1), 1)
        faces |= set(map(tuple, f[on[f].all(1)]))
    faces = np.array(sorted(faces), dtype=np.int64)
    Q = np.ascontiguousarray(P[faces].reshape(-1, 9))
    fo = ctx.envelope_faces(surf, torch.from_numpy(Q).to(ctx.device), M["eps_2"],
                            M["sampling_dist"]).cpu().numpy()
    po = ctx.envelope_points(surf, torch.from_numpy(np.ascontiguousarray(P)).to(ctx.device),
                             M["eps_2"]).cpu().numpy()
    surf.close()
    sec["envelope"] = {"faces": int(len(Q)), "faces_out": int(fo.sum()), "points": int(nv),
                       "points_out": int(po.sum())}
    if not args.no_cpu:
        S = O.Surface(M["sf_V"], M["sf_F"])
        fr, _ = S.face_out(Q, M["eps_2"], M["sampling_dist"])
        sec["envelope"]["decision_mismatch_vs_cpu"] = int(np.count_nonzero(fr != fo)) + \
            int(np.count_nonzero(S.point_out(P, M["eps_2"]) != po))
        sec["quality_bit_mismatch_vs_cpu"] = bits_mismatch(q, O.tet_qualities(tets, M["x"], M["y"],
                                                                               M["z"]))
    result["secondary"]["config1"] = sec


def bench_reduce(ctx, args, result):
    """twg_energy_stats over the 119^3 lattice's qualities (10.1M tets, 10%
    removed, 1% of vertices locked) and twg_tet_groups over the same tets in
    groups of four"""
    import numpy as np
    import torch
    from tetwild_amd import workloads as W

    x, y, z, tets, _ = W.kuhn_lattice(args.cubes, jitter=0.15, seed=0)
    n, nv = len(tets), len(x)
    tx, ty, tz, tt = (torch.from_numpy(a).to(ctx.device) for a in (x, y, z, tets))
    q = ctx.tet_qualities(tt, tx, ty, tz)[0]
    st = W.Stream(9, 0)
    removed = torch.from_numpy((st.uniform(n) < 0.1).astype(np.uint8)).to(ctx.device)
    locked = torch.from_numpy((st.uniform(nv) < 0.01).astype(np.uint8)).to(ctx.device)
    sec = {"tets": n}
    k = max(4, args.steps // 4)
    for _ in range(2):
        ctx.energy_stats(q, removed, tt, locked)
    t0 = time.perf_counter()
    for _ in range(k):
        s = ctx.energy_stats(q, removed, tt, locked)
    el = (time.perf_counter() - t0) / k
    sec["energy_stats_ms"] = el * 1e3
    sec["energy_stats_gbs"] = (n * (8 + 1 + 16) + nv) / el / 1e9
    sec["energy_stats"] = {"sum": s[0], "max": s[1], "second_max": s[2], "count": s[3],
                           "buckets": s[4]}
    if not args.no_cpu:
        log("[cpu] timing the CPU baseline sample")
        from oracle import oracle as O
        t0 = time.perf_counter()
        so = O.energy_stats(q.cpu().numpy(), removed.cpu().numpy(), tets, locked.cpu().numpy())
        el = time.perf_counter() - t0
        sec["cpu_baseline"] = {"value": n / el, "unit": "tets/s", "cores": 1, "kind": "port",
                               "sample": f"all {n} tets, CPU restatement (one core: the sum "
                                         "runs in index order, as the reference's loop)"}
        sec["max_bucket_count_mismatch_vs_cpu"] = int(
            (s[1] != so[1]) + (s[2] != so[2]) + (s[3] != so[3]) + int(list(s[4]) != list(so[4])))
        sec["sum_rel_diff_vs_cpu"] = abs(s[0] - so[0]) / abs(so[0])
    # the bulk quality pass (calTetQualities over every tet) and isFlip over
    # every tet, on the same lattice
    for prec, name in ((0, "fp64"), (1, "mixed")):
        wall, kern = events_time(lambda: ctx.tet_qualities(tt, tx, ty, tz, precision=prec), k, 2)
        sec[f"qualities_per_s_{name}"] = n / wall
        sec[f"qualities_gbs_{name}"] = (n * (16 + 8 + 1) + 24 * nv) / wall / 1e9
    wall, kern = events_time(lambda: ctx.tet_flip(tt, tx, ty, tz), k, 2)
    sec["flip_checks_per_s"] = n / wall
    sec["flip_gbs"] = (n * (16 + 1) + 24 * nv) / wall / 1e9
    q64 = ctx.tet_qualities(tt, tx, ty, tz, precision=0)[0]
    qm = ctx.tet_qualities(tt, tx, ty, tz, precision=1)[0]
    # the mixed pass returns FP64 values within tol (relative) and the same
    # MAX_ENERGY sentinels
    finite = q64 < 1e50
    sec["qualities_mixed_max_rel_err"] = ((qm - q64).abs() / q64.abs())[finite].max().item()
    sec["qualities_sentinel_mismatch_mixed_vs_fp64"] = int(((q64 == 1e50) != (qm == 1e50)).sum().item())
    if not args.no_cpu:
        log("[cpu] timing the CPU baseline sample")
        from oracle import oracle as O
        host = host_cpu()
        kq = 2_000_000
        t0 = time.perf_counter()
        O.tet_qualities(tets[:kq], x, y, z)
        el = time.perf_counter() - t0
        qo = np.empty(n)
        t0 = time.perf_counter()
        O.lib().orc_tet_qualities_mt(O._p(tets), n, O._p(x), O._p(y), O._p(z), O._p(qo),
                                     host["threads"])
        el_mt = time.perf_counter() - t0
        sec["qualities_cpu_baseline"] = cpu_entry(
            n / el_mt, kq / el, "tets/s", "port",
            f"all {n} lattice tets on all threads (first {kq} on one core), CPU restatement", host)
        sec["qualities_bit_mismatch_vs_cpu"] = bits_mismatch(q64.cpu().numpy(), qo)
    g = n // 4
    off = torch.arange(0, 4 * g + 1, 4, dtype=torch.int32, device=ctx.device)
    groups = tt[:4 * g].contiguous()
    wall, kern = events_time(lambda: ctx.tet_groups(groups, off, tx, ty, tz), k, 2)
    fl = ctx.tet_groups(groups, off, tx, ty, tz)[0]
    torch.cuda.synchronize()
    sec["groups"] = g
    sec["groups_per_s"] = g / wall
    sec["groups_flipped"] = int(fl.sum().item())
    result["secondary"]["reduce"] = sec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="smaller secondary workloads")
    ap.add_argument("--cubes", type=int, default=119)
    ap.add_argument("--only", default="amips,envelope,slivers,smoothing,collapse,reduce,config1")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain 1 GiB device copy (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path's "
              "multi-GPU sharding is a design note, not an implemented mode.", file=sys.stderr)
        sys.exit(2)

    import torch

    from tetwild_amd import workloads as W
    from tetwild_amd.ops import Context

    only = set(args.only.split(","))
    ctx = Context(0)
    if args.copy_calib:
        src = torch.empty(1 << 27, dtype=torch.float64, device=ctx.device).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    result = {
        "metric": METRIC, "value": None, "unit": "tet-evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic", "config": {}, "secondary": {},
    }
    t_all = time.time()

    def stage(name, fn, *a):
        t0 = time.time()
        fn(ctx, args, result, *a)
        torch.cuda.empty_cache()
        log(f"[{name}] done in {time.time() - t0:.1f}s (total {time.time() - t_all:.1f}s)")

    if "amips" in only:
        stage("amips", bench_amips)
    surf_cache = {}
    if only & {"envelope", "slivers", "smoothing"}:
        t0 = time.time()
        V, F = W.bumpy_sphere(freq=224, seed=0)
        surf_cache["VF"] = (V, F)
        surf_cache["surf"] = ctx.surface(V, F)
        log(f"[surface] {len(F)} facets, tree depth {surf_cache['surf'].depth} "
            f"({time.time() - t0:.1f}s)")
    if "envelope" in only:
        stage("envelope", bench_envelope, surf_cache)
    if "slivers" in only:
        stage("slivers", bench_slivers, surf_cache)
    if "smoothing" in only:
        stage("smoothing", bench_smoothing, surf_cache)
    if "surf" in surf_cache:
        surf_cache["surf"].close()
    if "collapse" in only:
        stage("collapse", bench_collapse)
    if "reduce" in only:
        stage("reduce", bench_reduce)
    if "config1" in only:
        stage("config1", bench_config1)
    if "sequence" in only:
        stage("sequence", bench_sequence)
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
