"""Benchmark of the equiprop path on one MI355X (one JSON line on stdout).

Workload: dim-64 Heisenberg chain with 3 controls, 65,536 steps, dt 5e-3, Simpson
quadrature (32,767 slices), synthetic seeded inputs (parament_amd.workloads.heisenberg6).
A step of this benchmark is one Parament_equipropDevice call over that whole batch, with
controls and result resident in HBM.  Three arithmetic paths are timed: the plain FP64
path, the mixed-precision path (complex128 interface) and the plain FP32 path.  The
headline ``value`` is time steps per second (pts / time per call) of the fastest path
whose propagator is within 1e-5 relative Frobenius error of the FP64 path's.  Each path
also reports ``host_arrays_*``: the same call through the reference's host-array entry
point (Parament_equiprop via the Parament class), control upload and result download
included; that rate is never ``value``.

``roofline`` describes that path's Chebyshev slice kernel (the dominant kernel), timed
with HIP events on the stream it runs on; ``cpu_baseline`` times the numpy FP64
restatement of the reference algorithm (oracle/) on a bounded sample of the same
workload on the host cores.  ``grape_grads_per_s`` is the rate of complete FP64 gradients
(F and every dF/dc) on the GRAPE workload (dim 32, 4 controls, 4096 steps); ``grape`` holds
both precisions, both memory layouts and the gradient / forward cost ratio;
``precision_variant_dim256`` times the dim-256 Ising workload in FP64, the three
mixed-precision product modes and complex64, with each one's error against FP64;
``ensemble`` times 512 independent systems of dims 4-128 as one tuned graph and as the
per-instance loop.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "mixed": 2500.0 / 6.0}  # mixed: bf16 dense / 6 split products

# HBM bytes per launch of the slice kernel on this workload, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes over `bench.py --only mixed`, each scaled by the counter's reading for a
# 1 GiB device copy in the same run (FETCH_SIZE x 2.00, WRITE_SIZE x 1.00): 18.5 MB + 306.7 MB.
MEASURED_TRAFFIC_BYTES = {"mixed": 3.25e8}


def fused_levels(path, n):
    """Tree levels the mixed slice kernel folds in (the library's rule for N slices)."""
    if path != "mixed":
        return 0
    return 3 if n >= 4096 else 2 if n >= 1024 else 1 if n >= 256 else 0


def hamiltonian_norm(h0, hk):
    """The library's Hnorm: the sum of the operators' one-norms (largest absolute row sum)."""
    return sum(float(np.abs(np.asarray(m)).sum(axis=1).max()) for m in (h0, *hk))


def products_per_slice(path, m, tau_hnorm):
    """Complex d^3 products per slice of the path's slice evaluation: the Clenshaw recurrence
    takes M - 1 (the reference's first step multiplies by a_M I, a scaling); the mixed path
    evaluates the same degree-M polynomial in powers of tau X when tau Hnorm <= 1, with
    (M + 1) / 2 products."""
    if path == "mixed" and tau_hnorm <= 1.0:
        return (m + 1) // 2
    return m - 1


def slice_flops(d, products, ncoef, n, levels=0):
    """Algorithmic flops of the slice kernel: ``products`` complex d^3 GEMMs per slice, the
    assembly, and one d^3 product per node of the tree levels folded into the kernel; 8 flops
    per complex MAC."""
    nodes = sum((n + (1 << l) - 1) >> l for l in range(1, levels + 1))
    return n * (8.0 * d ** 3 * products + 8.0 * d * d * ncoef) + nodes * 8.0 * d ** 3


def run_path(name, prec, mixed, h0, hk, c, dt, quad, steps, warmup):
    import torch
    import parament_amd as parament
    dev = torch.device("cuda:0")
    ctype = torch.complex128 if prec == "fp64" else torch.complex64
    d = h0.shape[0]
    amps, pts = c.shape
    ctx = parament.Parament(prec)
    if mixed:
        ctx.set_precision_mode("mixed")
    ctx.set_hamiltonian(h0, *hk, quadrature_mode=quad)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    carr = torch.tensor(np.ascontiguousarray(c), dtype=ctype, device=dev).contiguous()
    out = torch.empty(d * d, dtype=ctype, device=dev)
    torch.cuda.synchronize()
    call = lambda: ctx.equiprop_device(dt, carr.data_ptr(), pts, amps, out.data_ptr())
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    u = out.cpu().numpy().reshape(d, d).astype(np.complex128)
    ctx.set_timing(True)
    ks, kt = [], []
    for _ in range(5):
        call()
        a, b = ctx.kernel_times()
        ks.append(a)
        kt.append(b)
    m = ctx.iteration_cycles
    ctx.set_timing(False)
    # the reference's boundary: host arrays in and out, so the control upload, the result download
    # and the Python argument conversion are included
    for _ in range(2):
        ctx.equiprop(dt, *c)
    t0 = time.perf_counter()
    for _ in range(5):
        ctx.equiprop(dt, *c)
    th = (time.perf_counter() - t0) / 5
    ctx.destroy()
    return {"name": name, "s_per_call": t, "u": u, "M": m, "slice_ms": float(np.median(ks)),
            "tree_ms": float(np.median(kt)), "host_s_per_call": th}


def grape_bench(steps, warmup):
    """GRAPE workload: dim 32, 4 GUE controls, 4096 steps, dt 2e-3, Haar target.  Per precision:
    wall time of one equipropDevice call (forward) and of one gateFidelityGradientDevice call
    (F and all 4 x 4096 derivatives), with the whole product tree kept (default) and with
    sqrt(N) checkpointing."""
    import torch
    import parament_amd as parament
    from parament_amd import workloads
    h0, hk, c, dt, quad, target = workloads.gue32_grape()
    amps, pts = c.shape
    d = h0.shape[0]
    dev = torch.device("cuda:0")
    res = {}
    grads = {}
    for prec, ct, rt in (("fp64", torch.complex128, torch.float64), ("fp32", torch.complex64, torch.float32)):
        carr = torch.tensor(c, dtype=ct, device=dev).contiguous()
        tgt = torch.tensor(target, dtype=ct, device=dev).contiguous()
        out = torch.empty(d * d, dtype=ct, device=dev)
        g = torch.empty((amps, pts), dtype=rt, device=dev)
        f = torch.empty(1, dtype=rt, device=dev)
        ctx = parament.Parament(prec)
        ctx.set_hamiltonian(h0, *hk, quadrature_mode=quad)
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)

        def timed(call):
            for _ in range(warmup):
                call()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                call()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / steps

        t_fwd = timed(lambda: ctx.equiprop_device(dt, carr.data_ptr(), pts, amps, out.data_ptr()))
        entry = {"forward_ms": t_fwd * 1e3, "M": ctx.iteration_cycles}
        for label, seg in (("full_tree", 0), ("sqrt_checkpoint", 1)):
            ctx.set_gradient_checkpointing(seg, 0)
            t_grad = timed(lambda: ctx.gate_fidelity_gradient_device(dt, carr.data_ptr(), pts, amps, tgt.data_ptr(),
                                                                     f.data_ptr(), g.data_ptr()))
            entry[label] = {"grad_ms": t_grad * 1e3, "grads_per_s": 1.0 / t_grad, "grad_over_forward": t_grad / t_fwd,
                            "workspace_mb": ctx.gradient_workspace / 2 ** 20}
        grads[prec] = g.double().cpu().numpy()
        entry["fidelity"] = float(f.item())
        res[prec] = entry
        ctx.destroy()
    res["fp32_grad_err_vs_fp64"] = float(np.abs(grads["fp32"] - grads["fp64"]).max() / np.abs(grads["fp64"]).max())
    res["workload"] = "gue32_grape: dim 32, 4 controls, 4096 steps, dt 2e-3, no quadrature, Haar target"
    return res


def precision_variant(calls):
    """Dim-256 precision study: 8-qubit transverse-field Ising chain (J = 1, h ~ U[0, 2]) with 8
    local sigma_x controls, Gaussian-smoothed N(0, 1) controls, 16,384 steps, dt 2e-3, no
    quadrature.  Time per equipropDevice call and Frobenius error against the FP64 path, for
    FP64, the three mixed-precision product modes (FP64 first-order terms and upper tree
    levels) and plain complex64."""
    import torch
    import parament_amd as parament
    from parament_amd import workloads
    h0, hk, c, dt, quad = workloads.ising8()
    amps, pts = c.shape
    d = h0.shape[0]
    dev = torch.device("cuda:0")
    out, ref = {}, None
    for name, prec, mode in (("fp64", "fp64", "native"), ("mixed_bf16x3", "fp64", "mixed"),
                             ("mixed_fp32", "fp64", "mixed_fp32"), ("mixed_bf16x2", "fp64", "mixed_bf16x2"),
                             ("fp32", "fp32", None)):
        ct = torch.complex128 if prec == "fp64" else torch.complex64
        carr = torch.tensor(c, dtype=ct, device=dev).contiguous()
        u = torch.empty((d, d), dtype=ct, device=dev)
        ctx = parament.Parament(prec)
        if mode:
            ctx.set_precision_mode(mode)
        ctx.set_hamiltonian(h0, *hk, quadrature_mode=quad)
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        call = lambda: ctx.equiprop_device(dt, carr.data_ptr(), pts, amps, u.data_ptr())
        call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            call()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / calls
        un = u.cpu().numpy().astype(np.complex128)
        ref = un if ref is None else ref
        out[name] = {"ms_per_call": t * 1e3, "steps_per_s": pts / t, "M": ctx.iteration_cycles,
                     "rel_err_vs_fp64": float(np.linalg.norm(un - ref) / np.linalg.norm(ref))}
        ctx.destroy()
    out["workload"] = "ising8: dim 256, 8 controls, 16384 steps, dt 2e-3, no quadrature (16384 slices)"
    return out


def small_dim_bench(steps, warmup):
    """Dim 8, one GUE control, no quadrature (the reference's CPU-sized configuration, 4096 steps,
    and the same system over 1,048,576 steps): time per equipropDevice call in complex64 and
    complex128, the slice kernel (with its folded tree levels) and tree times, and the algorithmic
    HBM traffic -- the controls read once and each slice propagator written once and read once by
    the tree, sizeof(c) * (A P + 2 N d^2) -- over the whole call."""
    from parament_amd import workloads
    res = {}
    for pts in (4096, 1 << 20):
        h0, hk, c, dt, quad = workloads.dim8_gue(pts=pts)
        d, (amps, _) = h0.shape[0], c.shape
        for prec, csize in (("fp32", 8), ("fp64", 16)):
            r = run_path(prec, prec, False, h0, hk, c, dt, quad, steps, warmup)
            n = pts   # no quadrature: one slice per step
            call_bytes = csize * (amps * pts + 2 * n * d * d)
            res[f"{prec}_{pts}"] = {
                "ms_per_call": r["s_per_call"] * 1e3, "steps_per_s": pts / r["s_per_call"], "M": r["M"],
                "slice_kernel_ms": r["slice_ms"], "tree_ms": r["tree_ms"],
                "algorithmic_bytes": call_bytes,
                "call_GB_per_s": call_bytes / r["s_per_call"] / 1e9}
    res["workload"] = "dim8_gue: dim 8, 1 GUE control, dt 1e-3, no quadrature; 4096 and 1,048,576 steps"
    return res


def ensemble_bench(calls):
    """512 independent complex64 systems (dims round(exp(U[ln 4, ln 128])), 1-4 controls, 256-8192
    steps, dt 1e-2): time per solve of the whole ensemble as one tuned hipGraph and as the
    per-instance loop, with the total steps per second of each."""
    import torch
    from parament_amd.ensemble import Ensemble
    from parament_amd import workloads
    dims, amps, pts, dt, hs, cs = workloads.ensemble512()
    e = Ensemble(dims, amps, pts, dt)
    e.set_hamiltonians(hs)
    carr = torch.from_numpy(e._pack_controls(cs)).to("cuda:0")
    out = torch.empty(int(np.sum(dims.astype(np.int64) ** 2)), dtype=torch.complex64, device="cuda:0")
    res = {}
    for mode in ("loop", "graph"):
        e.set_execution(mode)
        t0 = time.perf_counter()
        e.equiprop_device(carr.data_ptr(), out.data_ptr())   # graph: tunes and captures
        e.synchronize()
        first = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(calls):
            e.equiprop_device(carr.data_ptr(), out.data_ptr())
        e.synchronize()
        t = (time.perf_counter() - t0) / calls
        res[mode] = {"ms_per_solve": t * 1e3, "steps_per_s": float(pts.sum()) / t, "first_call_s": first}
    res["graph_speedup_vs_loop"] = res["loop"]["ms_per_solve"] / res["graph"]["ms_per_solve"]
    # forward plus gradient of every instance's gate fidelity (Haar targets)
    rng = np.random.default_rng(20260006)
    tgt = torch.from_numpy(np.concatenate([np.ravel(workloads.haar_unitary(rng, int(d))) for d in dims])
                           .astype(np.complex64)).to("cuda:0")
    fid = torch.empty(len(dims), dtype=torch.float32, device="cuda:0")
    grad = torch.empty(int((amps * pts).sum()), dtype=torch.float32, device="cuda:0")
    gres = {}
    for mode, reps in (("loop", 1), ("graph", calls)):
        e.set_execution(mode)
        t0 = time.perf_counter()
        e.gate_fidelity_gradient_device(carr.data_ptr(), tgt.data_ptr(), fid.data_ptr(), grad.data_ptr())
        e.synchronize()
        first = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(reps):
            e.gate_fidelity_gradient_device(carr.data_ptr(), tgt.data_ptr(), fid.data_ptr(), grad.data_ptr())
        e.synchronize()
        t = (time.perf_counter() - t0) / reps
        gres[mode] = {"ms_per_gradient": t * 1e3, "first_call_s": first}
    gres["graph_speedup_vs_loop"] = gres["loop"]["ms_per_gradient"] / gres["graph"]["ms_per_gradient"]
    gres["ratio_to_forward"] = gres["graph"]["ms_per_gradient"] / res["graph"]["ms_per_solve"]
    gres["mean_fidelity"] = float(fid.mean().item())
    res["gradient"] = gres
    res["tuning"] = {str(k): [int(x) for x in v] if isinstance(v, tuple) else v for k, v in e.tuning.items()}
    res["workload"] = (f"ensemble512: {len(dims)} systems, dims 4-128 ({int((dims > 64).sum())} above 64), "
                       f"{int(pts.sum())} steps in total, complex64")
    e.destroy()
    return res


def cpu_baseline(h0, hk, c, dt, quad, sample_pts):
    from oracle import equiprop_oracle as eo
    try:
        from threadpoolctl import threadpool_info
        cores = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        cores = os.cpu_count() or 1
    cs = c[:, :sample_pts]
    t0 = time.perf_counter()
    u = eo.equiprop(h0, hk, cs, dt, mode=quad, precision="fp64")
    t = time.perf_counter() - t0
    return u, {"value": sample_pts / t, "unit": "steps/s", "cores": int(cores), "kind": "port",
               "sample": f"first {sample_pts} of the 65,536 steps ({(sample_pts - 1) // 2} slices), numpy FP64 "
                         f"restatement of the reference algorithm (M={eo.select_iteration_cycles_fp64(eo.hamiltonian_norm(h0, hk), 2 * dt)}), "
                         f"{t:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample-pts", type=int, default=49153)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["all", "fp64", "mixed", "fp32", "grape", "precision", "ensemble", "smalldim"],
                    default="all",
                    help="run one part only (for profiling); prints that part's measurements")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB (a known-size kernel for calibrating "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build's equiprop path runs on a single device "
              f"(--gpus {args.gpus} is not supported; the multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mb:
        import torch
        x = torch.empty(args.copy_probe_mb << 18, dtype=torch.float32, device="cuda:0").fill_(1.0)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y

    from parament_amd import workloads
    h0, hk, c, dt, quad = workloads.heisenberg6()
    d, amps, pts = h0.shape[0], len(hk), c.shape[1]
    n_slices = (pts - 1) // 2
    tau_hnorm = 2 * dt * hamiltonian_norm(h0, hk)   # Simpson: one slice spans two steps
    if args.only != "all":
        part = {}
        if args.only in ("fp64", "mixed", "fp32"):
            r = run_path(args.only, "fp32" if args.only == "fp32" else "fp64", args.only == "mixed", h0, hk, c, dt, quad,
                         args.steps, args.warmup)
            part = {k: v for k, v in r.items() if k != "u"}
        elif args.only == "grape":
            part = grape_bench(args.steps, args.warmup)
        elif args.only == "precision":
            part = precision_variant(args.steps)
        elif args.only == "smalldim":
            part = small_dim_bench(args.steps, args.warmup)
        else:
            part = ensemble_bench(args.steps)
        print(json.dumps({"only": args.only, **part}), flush=True)
        return

    runs = [run_path(*p, h0, hk, c, dt, quad, args.steps, args.warmup)
            for p in (("fp64", "fp64", False), ("mixed", "fp64", True), ("fp32", "fp32", False))]
    u64 = runs[0]["u"]
    paths = {}
    for r in runs:
        err = float(np.linalg.norm(r["u"] - u64) / np.linalg.norm(u64))
        fl = slice_flops(d, products_per_slice(r["name"], r["M"], tau_hnorm), amps, n_slices,
                         fused_levels(r["name"], n_slices))
        ach = fl / (r["slice_ms"] * 1e-3) / 1e12
        paths[r["name"]] = {"steps_per_s": pts / r["s_per_call"], "ms_per_call": r["s_per_call"] * 1e3,
                            "rel_err_vs_fp64": err, "M": r["M"], "slice_kernel_ms": r["slice_ms"],
                            "tree_ms": r["tree_ms"], "slice_kernel_tflops": ach,
                            "slice_kernel_frac_of_peak": ach / PEAK_TFLOPS[r["name"]],
                            "host_arrays_ms_per_call": r["host_s_per_call"] * 1e3,
                            "host_arrays_steps_per_s": pts / r["host_s_per_call"]}
    ok = [k for k in ("mixed", "fp32", "fp64") if paths[k]["rel_err_vs_fp64"] <= 1e-5]
    best = max(ok, key=lambda k: paths[k]["steps_per_s"])
    p = paths[best]
    fl = slice_flops(d, products_per_slice(best, p["M"], tau_hnorm), amps, n_slices, fused_levels(best, n_slices))
    line = {
        "metric": "steps/s at dim 64 @<=1e-5 vs FP64 (mixed vs plain FP64); GRAPE grads/s",
        "value": p["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": p["ms_per_call"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "complex128", "mixed": "bf16x3(fp32 acc)+fp64", "fp32": "complex64"}[best],
        "data": "synthetic",
        "config": {"workload": "heisenberg6: dim 64, 3 controls, 65536 steps, dt 5e-3, simpson (32767 slices)",
                   "path": best, "dim": d, "controls": amps, "pts": pts, "slices": n_slices},
        "roofline": {"bound": "mfma", "achieved": p["slice_kernel_tflops"], "peak": PEAK_TFLOPS[best],
                     "unit": "TFLOP/s", "frac": p["slice_kernel_frac_of_peak"],
                     "traffic": MEASURED_TRAFFIC_BYTES.get(best),
                     "kernel": "Chebyshev slice kernel", "flops_per_launch": fl},
        "speedup_mixed_vs_fp64": paths["mixed"]["steps_per_s"] / paths["fp64"]["steps_per_s"],
        "paths": paths,
    }
    grape = grape_bench(args.steps, args.warmup)
    line["grape_grads_per_s"] = grape["fp64"]["full_tree"]["grads_per_s"]
    line["grape"] = grape
    line["precision_variant_dim256"] = precision_variant(3)
    line["ensemble"] = ensemble_bench(3)
    line["small_dim"] = small_dim_bench(args.steps, args.warmup)
    if not args.no_cpu:
        ucpu, cb = cpu_baseline(h0, hk, c, dt, quad, args.cpu_sample_pts)
        line["cpu_baseline"] = cb
        r = run_path("fp64", "fp64", False, h0, hk, c[:, :args.cpu_sample_pts], dt, quad, 1, 1)
        line["fp64_vs_oracle_on_sample"] = float(np.linalg.norm(r["u"] - ucpu) / np.linalg.norm(ucpu))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
