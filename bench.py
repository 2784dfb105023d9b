"""Benchmark: SMAC projection step, cell-steps/s at 512^3 with MG-PCG to 1e-8.

Workload (one GPU): triply periodic [0,2pi)^3 box, 512^3 MAC grid, Taylor-Green
velocity plus a seeded random-phase solenoidal k^-5/3 perturbation (|k| <= 32,
5% of the kinetic energy), Re = 1600, dt fixed from CFL 0.3 of the initial
field. One step = predictor -> RHS -> MG-PCG (RBGS V(2,2) preconditioner,
||r||_2/||b||_2 < 1e-8) -> correction. Inputs are resident in HBM before the
timed region; each form is timed over --steps steps after --warmup steps.

Prints ONE JSON line. The headline (value, ms_per_step) is timed with kernel
profiling off. A second pass of --prof-steps further steps records HIP events
around every launch on the context stream (the stream the kernels run on);
`roofline` is the dominant kernel of that pass (largest total device time) and
`roofline_kernels` the six largest, each with algorithmic bytes per launch from
the byte model below. `cpu_baseline` is the CPU FP64 oracle (a restatement of
the reference's GPU algorithm; the reference's own CPU SMACSolver has no step),
compiled on this host with -O3 -march=native -ffp-contract=off and timed here
on a bounded sample: the reference algorithm (REF mode: reference GMG-PCG,
weighted-Jacobi V(4,4) + 30 coarse sweeps, 4 levels) and the build's algorithm
(NS mode: RBGS-MG-PCG V(2,2)), each 1 warm-up step then the median of 3 steps,
at --cpu-n^3 of the same workload.

Canonical byte model (bytes per interior cell, FP64 word = 8 B; W_s = storage
width of the V-cycle):
  step   B_step = 160 + (17.43*W_s + 32) + N_it * B_it,
         B_it   = 104 + 17.43*W_s (+16 when the PCG is flexible)
  kernels (per launch, per cell of the level it runs on; W_z = width of z,
  8 in FP64, W_s in the mixed forms):
         spmv_dot 16, update 24 (r, Ad; +W_s: the mixed forms also write
         (T) r), update_zs the same + 0.5*W_s (it also does the next
         V-cycle's zero-start red half sweep, e = f/diag), dot2 8 + W_z (+8 flexible), xpay 32 + W_z (z, d, dp: the
         dp += alpha d of the previous update rides on the direction update),
         dir_spmv 40 + W_z (xpay fused into the next SpMV: z, d_old, dp read,
         dp, d, Ad written), first_spmv 24 + W_z (the first SpMV of a
         solve with d = z and r.d: z, r read, d, Ad written), dp_flush 24 (that dp update after the last iteration; a step
         leaves it to the correction, which then also reads d: +8 there),
         mg_rbgs_l0 1.5*W_s (a half sweep: other colour, own rhs, own
         colour written), mg_rbgs_l0_zs 1.0*W_s (the first red half sweep of
         a V-cycle starts from zero: rhs read, red cells written),
         mg_rbgs_dot 1.5*W_s + 8 (+8 flexible) (the last red half sweep of a
         V-cycle with dot2's products: r and Ad read, z not read again; the
         FP32 and FP64 forms),
         mg_restrict_l0 2.125*W_s
         (residual + restriction), mg_prolong_l0 1.125*W_s (red cells),
         mg_fill_l0 W_s, predict 56, rhs 32, correct 80 (u* and dp read, u
         written, p += dp, then dp cleared)

`--copy-calib BYTES` instead times a plain device copy of BYTES bytes (the
calibration of the FETCH_SIZE / WRITE_SIZE counters) and prints its rate.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

KERNEL_BYTES = {  # bytes per cell of the level, per launch; W = storage width (< 8: a mixed form)
    "spmv_dot": lambda W, flex: 16.0,
    "update": lambda W, flex: 24.0 + (W if W < 8.0 else 0.0),
    "update_zs": lambda W, flex: 24.0 + (W if W < 8.0 else 0.0) + 0.5 * W,
    "dot2": lambda W, flex: 8.0 + W + (8.0 if flex else 0.0),
    "xpay": lambda W, flex: 32.0 + W,
    "dir_spmv": lambda W, flex: 40.0 + W,
    "first_spmv": lambda W, flex: 24.0 + W,
    "dp_flush": lambda W, flex: 24.0,
    "mg_rbgs_l0": lambda W, flex: 1.5 * W,
    "mg_rbgs_l0_zs": lambda W, flex: 1.0 * W,
    "mg_rbgs_dot": lambda W, flex: 1.5 * W + 8.0 + (8.0 if flex else 0.0),
    "mg_restrict_l0": lambda W, flex: 2.125 * W,
    "mg_prolong_l0": lambda W, flex: 1.125 * W,
    "mg_fill_l0": lambda W, flex: 1.0 * W,
    "predict": lambda W, flex: 56.0,
    "rhs": lambda W, flex: 32.0,
    "correct": lambda W, flex: 88.0,
}

FORMS = {
    "fp64": dict(precision="fp64", W=8.0, flexible=0),
    "mixed_fp32": dict(precision="fp32", W=4.0, flexible=1),
    "mixed_bf16": dict(precision="bf16", W=2.0, flexible=1),
}


def step_bytes(W, n_it_per_step, flexible):
    b_it = 104.0 + 17.43 * W + (16.0 if flexible else 0.0)
    return 160.0 + (17.43 * W + 32.0) + n_it_per_step * b_it


def run_form(name, n, steps, warmup, prof_steps, fields, re, cfl, tol, levels):
    import torch

    import mydemcfd_amd
    from mydemcfd_amd import cases

    spec = FORMS[name]
    desc = cases.periodic_desc(n, re=re)
    ctx = mydemcfd_amd.SMACContext(**desc)
    torch.cuda.synchronize()  # the context's stream is not ordered with torch's
    for fname, t in zip(("u", "v", "w"), fields):
        ctx.upload_device(fname, t.data_ptr(), t.numel())
    ctx.sync()
    dt = cases.cfl_dt(ctx, cfl)
    opts = ctx.opts(precond="mg", precision=spec["precision"], stop_rule="rel_l2", tol=tol,
                    flexible=spec["flexible"], mg_levels=levels)
    for _ in range(warmup):
        st = ctx.step(dt, opts)
        if not st.converged:
            raise RuntimeError(f"{name}: warmup step did not converge: {st.as_dict()}")
    ctx.set_profiling(False)
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    res = []
    for _ in range(steps):
        st = ctx.step(dt, opts)
        iters.append(st.iters)
        res.append(st.rel_res)
        if not st.converged:
            raise RuntimeError(f"{name}: step did not converge: {st.as_dict()}")
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    # kernel breakdown: a second pass with HIP events around every launch
    ctx.set_profiling(True)
    ctx.reset_kernel_times()
    prof_iters = []
    t2 = time.perf_counter()
    for _ in range(prof_steps):
        st = ctx.step(dt, opts)
        prof_iters.append(st.iters)
    ctx.sync()
    t3 = time.perf_counter()
    kt = {}
    for k in ctx.kernel_names():
        cnt, ms = ctx.kernel_time(k)
        kt[k] = (cnt, ms)
    ctx.set_profiling(False)
    out = {"form": name, "dt": dt, "wall_s": wall, "ms_per_step": 1e3 * wall / steps,
           "cell_steps_per_s": n ** 3 * steps / wall, "iters": iters, "rel_res_max": max(res),
           "kernels": kt, "prof_steps": prof_steps, "prof_iters": prof_iters,
           "prof_ms_per_step": 1e3 * (t3 - t2) / max(prof_steps, 1), "W": spec["W"], "flexible": spec["flexible"]}
    n_it = sum(iters) / len(iters)
    bstep = step_bytes(spec["W"], n_it, spec["flexible"])
    out["bytes_per_cell_step"] = bstep
    out["step_model_gbps"] = bstep * n ** 3 * steps / wall / 1e9
    out["step_roofline_frac"] = out["step_model_gbps"] / HBM_PEAK_GBS
    # fields for the error report of the mixed forms
    out["_ctx"] = ctx
    return out


def run_c5(steps, warmup, tol, adjoint_steps):
    """The variable-coefficient workload: the VoF dam break on 256x128x128
    ([0,2]x[0,1]x[0,1], heavy column x < 0.25, z < 0.5, rho 1000:1, mu 100:1,
    gravity -z, no-slip walls), dt = 1e-3, each loop step = the split THINC/WLIC
    transport (two substeps) + the SMAC step with mixed FP32 RBGS-MG-PCG to
    rel-res tol. Timed: `steps` loop steps after `warmup`; then the adjoint of
    the column-front position x_f = (h_x/n_y) sum alpha over the floor layer
    w.r.t. the initial VoF field over `adjoint_steps` steps (forward with
    per-step checkpoints + backward). The byte model is the library's
    bytes_model (variable coefficients: +48 B fixed, +24 B per iteration, +(15 x
    8/7) W per V-cycle); the transport has no term in it, so the roofline
    fraction is taken over the step time without the transport kernels
    (measured in a profiled pass)."""
    import torch

    import mydemcfd_amd
    from mydemcfd_amd import autograd, cases

    n = (256, 128, 128)
    cells = n[0] * n[1] * n[2]
    dt, sub = 1e-3, 2
    desc = cases.dam_break_desc(n)
    ctx = mydemcfd_amd.SMACContext(**desc)
    alpha0 = cases.dam_break_alpha(n)
    ctx.upload("alpha", alpha0)
    ctx.update_properties(True)
    opts = ctx.opts(precond="mg", precision="fp32", flexible=1, stop_rule="rel_l2", tol=tol, max_iter=3000)
    k = 0
    for _ in range(warmup):
        ctx.transport_alpha(dt, sub, k)
        ctx.step(dt, opts)
        k += 1
    ctx.set_profiling(False)
    ctx.sync()
    t0 = time.perf_counter()
    iters, bmodel = [], 0.0
    for _ in range(steps):
        ctx.transport_alpha(dt, sub, k)
        st = ctx.step(dt, opts)
        if not st.converged:
            raise RuntimeError(f"c5: step did not converge: {st.as_dict()}")
        iters.append(st.iters)
        bmodel += st.bytes_model
        k += 1
    ctx.sync()
    wall = time.perf_counter() - t0
    # transport share: a profiled pass of three loop steps
    ctx.set_profiling(True)
    ctx.reset_kernel_times()
    for _ in range(3):
        ctx.transport_alpha(dt, sub, k)
        ctx.step(dt, opts)
        k += 1
    ctx.sync()
    kt = {name: ctx.kernel_time(name) for name in ctx.kernel_names()}
    ctx.set_profiling(False)
    total_ms = sum(ms for _, ms in kt.values())
    vof_ms = sum(ms for name, (_, ms) in kt.items() if "vof" in name)
    vof_share = vof_ms / total_ms if total_ms > 0 else 0.0
    step_s = wall * (1.0 - vof_share)
    achieved = bmodel / step_s / 1e9
    out = {"workload": "VoF dam break 256x128x128 ([0,2]x[0,1]x[0,1]), rho 1000:1, mu 100:1, gravity -z, "
                       "dt 1e-3, transport (2 substeps) + step, mixed FP32 RBGS-MG-PCG V(2,2) to rel-res "
                       f"{tol:g}",
           "steps": steps, "warmup": warmup, "ms_per_step": 1e3 * wall / steps,
           "cell_steps_per_s": cells * steps / wall, "iters": iters,
           "transport_share": round(vof_share, 4),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "bytes_model_per_step": bmodel / steps,
                        "basis": "bytes_model of the steps / (loop time x (1 - transport share))"},
           "kernels_ms": {name: round(ms / 3, 4) for name, (_, ms) in kt.items()}}
    del ctx
    torch.cuda.empty_cache()
    if adjoint_steps > 0:
        actx = mydemcfd_amd.SMACContext(**desc)
        leaves = {"alpha": torch.tensor(alpha0, dtype=torch.float64, device="cuda").requires_grad_()}
        for f in ("u", "v", "w", "p"):
            leaves[f] = torch.zeros(actx.shape(f), dtype=torch.float64, device="cuda", requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs = autograd.vof_rollout(actx, leaves["alpha"], leaves["u"], leaves["v"], leaves["w"], leaves["p"], dt,
                                    adjoint_steps, sub, opts)
        front = outs[0][1, 1:n[1] + 1, 1:n[0] + 1].sum() * ((2.0 / n[0]) / n[1])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        front.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        g = leaves["alpha"].grad
        out["adjoint"] = {"steps": adjoint_steps, "x_front": float(front.detach()),
                          "forward_ms_per_step": 1e3 * (t1 - t0) / adjoint_steps,
                          "backward_ms_per_step": 1e3 * (t2 - t1) / adjoint_steps,
                          "grad_norm": float(torch.linalg.vector_norm(g)), "grad_finite": bool(torch.isfinite(g).all())}
        del outs, leaves, actx
        torch.cuda.empty_cache()
    return out


def field_errors(ctx_a, ctx_b, n):
    """max|a-b|/max|b| per field on the GPU (p modulo its interior mean)."""
    import torch

    from mydemcfd_amd import _abi

    out = {}
    for f in ("u", "v", "w", "p"):
        shape = _abi.field_shape((n, n, n), _abi.FIELD_NAMES[f])
        ta = torch.empty(shape, dtype=torch.float64, device="cuda")
        tb = torch.empty_like(ta)
        torch.cuda.synchronize()
        ctx_a.download_device(f, ta.data_ptr(), ta.numel())
        ctx_b.download_device(f, tb.data_ptr(), tb.numel())
        ia = ta[1:n + 1, 1:n + 1, 1:n + 1]
        ib = tb[1:n + 1, 1:n + 1, 1:n + 1]
        if f == "p":
            ia = ia - ia.mean()
            ib = ib - ib.mean()
        out[f] = float((ia - ib).abs().max() / ib.abs().max())
        del ta, tb
    return out


TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic.md")


def measured_traffic(n, form, kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (scaled by the plain-copy calibration), from the table in
    profiles/hbm_traffic.md: rows | n | form | kernel | ... | bytes per launch |."""
    try:
        with open(TRAFFIC_TABLE) as fh:
            for line in fh:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) >= 4 and cells[0] == str(n) and cells[1] == form and cells[2] == kernel:
                    return int(float(cells[-1]))
    except (OSError, ValueError):
        return None
    return None


def kernel_roofline(form, n, k):
    cnt, ms = form["kernels"][k]
    per_launch = KERNEL_BYTES[k](form["W"], form["flexible"]) * n ** 3
    avg_s = ms / cnt / 1e3
    achieved = per_launch / avg_s / 1e9
    return {"bound": "hbm", "kernel": k, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(n, form["form"], k),
            "launches": cnt, "avg_us": round(1e6 * avg_s, 2), "alg_bytes_per_launch": per_launch,
            "share_of_step": round(ms / (form["prof_ms_per_step"] * form["prof_steps"]), 4)}


def top_kernels(form, count):
    """The `count` kernels of the byte model with the largest total device time."""
    ks = [k for k, (cnt, ms) in form["kernels"].items() if cnt > 0 and k in KERNEL_BYTES]
    ks.sort(key=lambda k: -form["kernels"][k][1])
    return ks[:count]


def dominant_roofline(form, n):
    ks = top_kernels(form, 1)
    return kernel_roofline(form, n, ks[0]) if ks else None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n_cpu, re, cfl, tol):
    """The CPU FP64 oracle, compiled here with -O3 -march=native
    -ffp-contract=off, on n_cpu^3 of the workload: 1 warm-up step, then the
    median of 3 steps, in REF mode (the reference algorithm) and NS mode."""
    import statistics
    import tempfile

    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from mydemcfd_amd import cases
    from oracle import oracle

    tmp = tempfile.mkdtemp(prefix="smac_cpu_baseline_")
    lib = oracle.native_lib(tmp)
    fields = cases.tg_fields(n_cpu, seed=0x5EED0003, pert="k53")
    modes = {"ref": dict(precond="ref_gmg", mg_levels=4, stop_rule="rel_l2", tol=tol),
             "ns": dict(precond="mg", stop_rule="rel_l2", tol=tol)}
    out = {}
    for mode, kw in modes.items():
        ctx = oracle.OracleContext(lib_=lib, **cases.periodic_desc(n_cpu, re=re))
        for f, a in zip(("u", "v", "w"), fields):
            ctx.upload(f, a)
        dt = cases.cfl_dt(ctx, cfl)
        ctx.step(dt, **kw)  # warm-up (multigrid set-up included)
        ts, its = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            st = ctx.step(dt, **kw)
            ts.append(time.perf_counter() - t0)
            its.append(st.iters)
        med = statistics.median(ts)
        out[mode] = {"value": n_cpu ** 3 / med, "median_s": med, "runs_s": ts, "iters": its}
        ctx.close()
        print(f"# cpu baseline {mode}: {n_cpu}^3 median {med:.2f} s/step, iters {its}", file=sys.stderr, flush=True)
    threads = oracle.threads(lib)
    return {"value": out["ref"]["value"], "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n_cpu}^3 periodic TG+k^-5/3 (same workload type), REF mode = the reference algorithm "
                      f"(reference GMG-PCG, 4 levels, to rel-res {tol:g}): 1 warm-up step + median of 3 steps "
                      f"({out['ref']['median_s']:.2f} s/step, iterations {out['ref']['iters']})",
            "ns_mode": {"value": out["ns"]["value"], "median_s": out["ns"]["median_s"],
                        "iters": out["ns"]["iters"], "what": "the build's RBGS-MG-PCG V(2,2) on the CPU"},
            "ref_mode_runs_s": out["ref"]["runs_s"], "ns_mode_runs_s": out["ns"]["runs_s"],
            "cpu_model": cpu_model(), "threads": threads,
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
            "build": " ".join(["gcc"] + oracle.NATIVE_FLAGS)}


def protocol(tol):
    """The CPU-baseline protocol in full, beside the GPU on the same steps (a
    few minutes; not part of the default run): C1 the 64^3 lid-driven cavity
    (Re 100, dt 1e-3, 100 steps, Jacobi-PCG), C2 the 256^3 periodic TG + curl
    perturbation (seed 0x5EED0002, RMS 1e-3, Re 1600, CFL 0.3, 20 steps), C3
    the 512^3 TG + k^-5/3 field (2 steps on the CPU). CPU: the oracle built
    here with -O3 -march=native in REF mode (the reference algorithm: Jacobi-PCG
    for C1, the reference GMG-PCG for C2 and C3) and NS mode (RBGS-MG-PCG
    V(2,2)), one warm-up step, then every step of the run timed (C3: the
    median of its two). GPU: the same steps in FP64 and the mixed forms, with
    the byte model of the library (bytes_model) for the roofline fraction.
    All solves to rel-res tol. Prints one JSON line."""
    import statistics
    import tempfile

    import torch

    import mydemcfd_amd
    from mydemcfd_amd import cases
    from oracle import oracle

    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    lib = oracle.native_lib(tempfile.mkdtemp(prefix="smac_protocol_"))
    mydemcfd_amd.load()
    c2_fields = None

    def setup(ctx, cfg, n):
        if cfg == "C1":
            return 1e-3
        fields = c2_fields if cfg == "C2" else cases.tg_fields(n, seed=0x5EED0003, pert="k53")
        for f, a in zip(("u", "v", "w"), fields):
            ctx.upload(f, a)
        return cases.cfl_dt(ctx, 0.3)

    def run(make, cfg, n, steps, kw, gpu):
        ctx = make()
        dt = setup(ctx, cfg, n)
        ctx.step(dt, **kw)  # warm-up (multigrid set-up included)
        if gpu:
            ctx.sync()
        ts, its, by = [], [], 0.0
        for _ in range(steps):
            t0 = time.perf_counter()
            st = ctx.step(dt, **kw)
            if gpu:
                ctx.sync()
            ts.append(time.perf_counter() - t0)
            its.append(st.iters)
            by += st.bytes_model
        ctx.close()
        cells = n ** 3
        t = statistics.median(ts) * steps if cfg == "C3" and not gpu else sum(ts)
        out = {"cell_steps_per_s": cells * steps / t, "s": t, "iters_mean": sum(its) / len(its)}
        if gpu:
            out["roofline_frac"] = by / t / 1e9 / HBM_PEAK_GBS
        print(f"# protocol {cfg} {'gpu' if gpu else 'cpu'} {kw}: {out}", file=sys.stderr, flush=True)
        return out

    rows = {}
    cfgs = [("C1", 64, 100), ("C2", 256, 20), ("C3", 512, 2)]
    for cfg, n, steps in cfgs:
        if cfg == "C2":
            c2_fields = cases.tg_fields(256, seed=0x5EED0002, pert="white", pert_rms=1e-3)
        desc = cases.cavity_desc(n) if cfg == "C1" else cases.periodic_desc(n, re=1600.0)
        ref = dict(precond="jacobi", tol=tol, max_iter=5000) if cfg == "C1" else \
            dict(precond="ref_gmg", mg_levels=4, tol=tol)
        ns = dict(precond="mg", tol=tol)
        row = {"n": n, "steps": steps}
        row["cpu_ref"] = run(lambda: oracle.OracleContext(lib_=lib, **desc), cfg, n, steps, ref, False)
        if cfg != "C1":
            row["cpu_ns"] = run(lambda: oracle.OracleContext(lib_=lib, **desc), cfg, n, steps, ns, False)
        gsteps = steps if cfg != "C3" else 10
        if cfg == "C1":
            row["gpu_fp64"] = run(lambda: mydemcfd_amd.SMACContext(**desc), cfg, n, gsteps, ref, True)
        else:
            for form, spec in FORMS.items():
                kw = dict(precond="mg", precision=spec["precision"], flexible=spec["flexible"], tol=tol)
                row[f"gpu_{form}"] = run(lambda: mydemcfd_amd.SMACContext(**desc), cfg, n, gsteps, kw, True)
        rows[cfg] = row
        torch.cuda.empty_cache()
    line = {"protocol": "CPU baseline in full beside the GPU", "cpu_model": cpu_model(),
            "threads": oracle.threads(lib), "build": " ".join(["gcc"] + oracle.NATIVE_FLAGS),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
            "tol": tol, "configs": rows}
    print(json.dumps(line), flush=True)


def copy_calibration(nbytes, reps=4):
    """A plain device-to-device copy of nbytes (reads nbytes, writes nbytes),
    the reference point for the HBM byte counters."""
    import torch

    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes": nbytes, "launches": reps + 1, "GBps_read_plus_write": 2 * nbytes / dt / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--forms", default="mixed_fp32,mixed_bf16,fp64")
    ap.add_argument("--headline", default=None, help="form reported as value (default: first form)")
    ap.add_argument("--re", type=float, default=1600.0)
    ap.add_argument("--cfl", type=float, default=0.3)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--levels", type=int, default=0)
    ap.add_argument("--prof-steps", type=int, default=3, help="steps of the profiled (kernel breakdown) pass")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="BYTES")
    ap.add_argument("--protocol", action="store_true",
                    help="run the CPU-baseline protocol in full beside the GPU instead (minutes)")
    ap.add_argument("--c5-steps", type=int, default=50, help="steps of the VoF dam-break workload (0: skip)")
    ap.add_argument("--c5-adjoint-steps", type=int, default=50, help="steps of its adjoint (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU (z-slab) decomposition "
              f"is a design, not built here.", file=sys.stderr)
        sys.exit(2)
    import torch

    if args.copy_calib:
        copy_calibration(args.copy_calib)
        return
    if args.protocol:
        protocol(args.tol)
        return

    import mydemcfd_amd
    from mydemcfd_amd import cases

    mydemcfd_amd.load()
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    headline = args.headline or forms[0]
    t_ic = time.perf_counter()
    fields = cases.tg_fields_torch(args.n, device="cuda")
    torch.cuda.synchronize()
    t_ic = time.perf_counter() - t_ic
    results = {}
    for f in forms:
        results[f] = run_form(f, args.n, args.steps, args.warmup, args.prof_steps, fields, args.re, args.cfl,
                              args.tol, args.levels)
        print(f"# {f}: {results[f]['ms_per_step']:.2f} ms/step, iters {results[f]['iters']}", file=sys.stderr,
              flush=True)
    del fields
    torch.cuda.empty_cache()
    h = results[headline]
    line = {
        "metric": "cell-steps/s at 512^3 (MG-PCG to 1e-8 resid) + % HBM roofline, mixed vs FP64",
        "value": h["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed_fp32": "f32+f64", "mixed_bf16": "bf16+f32+f64"}[headline],
        "data": "synthetic",
        "config": {"workload": f"periodic TG + k^-5/3 (k<=32, 5% energy) {args.n}^3, Re={args.re:g}, "
                               f"CFL {args.cfl:g}, RBGS-MG-PCG V(2,2) to rel-res {args.tol:g}, form {headline}",
                   "n": args.n, "form": headline},
        "roofline": dominant_roofline(h, args.n),
        "roofline_kernels": [kernel_roofline(h, args.n, k) for k in top_kernels(h, 6)],
        "kernel_pass": {"steps": h["prof_steps"], "iters": h["prof_iters"],
                        "ms_per_step_with_events": h["prof_ms_per_step"]},
        "step_model_gbps": h["step_model_gbps"],
        "step_roofline_frac": h["step_roofline_frac"],
        "bytes_per_cell_step": h["bytes_per_cell_step"],
        "iters_per_step": h["iters"],
        "forms": {f: {"cell_steps_per_s": r["cell_steps_per_s"], "ms_per_step": r["ms_per_step"],
                      "iters": r["iters"], "step_model_gbps": r["step_model_gbps"],
                      "step_roofline_frac": r["step_roofline_frac"], "rel_res_max": r["rel_res_max"],
                      "kernels_ms": {k: round(v[1], 3) for k, v in r["kernels"].items()}}
                  for f, r in results.items()},
        "ic_seconds": t_ic,
    }
    if "fp64" in results:
        for f, r in results.items():
            if f != "fp64":
                line["forms"][f]["error_vs_fp64"] = field_errors(r["_ctx"], results["fp64"]["_ctx"], args.n)
        if headline != "fp64":
            line["error_vs_fp64"] = line["forms"][headline]["error_vs_fp64"]
            line["speedup_vs_fp64"] = results[headline]["cell_steps_per_s"] / results["fp64"]["cell_steps_per_s"]
    for r in results.values():
        r.pop("_ctx", None)
    if args.c5_steps > 0:
        line["c5"] = run_c5(args.c5_steps, 2, args.tol, args.c5_adjoint_steps)
        print(f"# c5: {line['c5']['ms_per_step']:.2f} ms/step, iters {line['c5']['iters'][:5]}...", file=sys.stderr,
              flush=True)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_n, args.re, args.cfl, args.tol)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
