"""cholqr2 throughput on one MI355X: rows/s of CholeskyQR2 on a 10,077,696 x 512
ComplexF64 basis (the 216^3 periodic Helmholtz stencil applied to a seeded
GhostPanels matrix, panels of 32 columns), resident in HBM, FP64, plus the
mixed-precision first pass against FP64-only.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one cholqr2 of the whole basis. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "funicular.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = "cholqr2! rows/s at N=1e7,k=512; mixed vs FP64 speedup at ‖QᴴQ−I‖≤1e-13"
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class KernelTimer:
    """HIP timing events around every Gram and triangular-solve call (the
    solve of a sweep's first block includes folding the factor, the others
    reuse it), recorded on the stream the kernels are launched on."""

    def __init__(self, backend):
        import ctypes
        from funicular_hip import _lib
        from funicular_hip.device_api import current_stream
        self.backend = backend
        self.ctypes = ctypes
        self._lib = _lib
        self.current_stream = current_stream
        self.records = {"gram": [], "trsm": []}
        self.active = False
        self._orig_gram = backend.gram_accumulate
        self._orig_trsm = backend.rdiv_upper
        self._orig_folded = backend.rdiv_upper_folded
        timer = self

        def gram(C, A, B, alpha, beta, work=None, compute=None):
            ev = timer._start()
            r = timer._orig_gram(C, A, B, alpha, beta, work, compute)
            timer._stop("gram", ev, (A.rows, A.cols, compute))
            return r

        def trsm(Y, R, work=None, compute=None):
            ev = timer._start()
            r = timer._orig_trsm(Y, R, work, compute)
            timer._stop("trsm", ev, (Y.rows, Y.cols, compute))
            return r
        def folded(Y, work, compute=None, adjoint=False):
            ev = timer._start()
            r = timer._orig_folded(Y, work, compute, adjoint)
            timer._stop("trsm", ev, (Y.rows, Y.cols, compute))
            return r
        backend.gram_accumulate = gram
        backend.rdiv_upper = trsm
        backend.rdiv_upper_folded = folded

    def _event(self):
        e = self.ctypes.c_void_p()
        self._lib.call("fhip_event_create", self.ctypes.byref(e), 0)
        return e.value

    def _start(self):
        if not self.active:
            return None
        e = self._event()
        self._lib.call("fhip_event_record", e, self.current_stream())
        return e

    def _stop(self, kind, start, shape):
        if start is None:
            return
        e = self._event()
        self._lib.call("fhip_event_record", e, self.current_stream())
        self.records[kind].append((start, e, shape))

    def collect(self):
        """(kind, compute) -> list of (ms, rows, k)."""
        out = {}
        for kind, recs in self.records.items():
            for start, stop, (rows, k, compute) in recs:
                self._lib.call("fhip_event_sync", stop)
                ms = self.ctypes.c_float()
                self._lib.call("fhip_event_elapsed", self.ctypes.byref(ms), start, stop)
                out.setdefault((kind, compute is not None), []).append((ms.value, rows, k))
                self._lib.call("fhip_event_destroy", start)
                self._lib.call("fhip_event_destroy", stop)
        self.records = {"gram": [], "trsm": []}
        return out


def measured_traffic(kind):
    """HBM bytes per launch of the 629,856 x 512 row-block kernels, as the
    FETCH_SIZE / WRITE_SIZE counter runs measured them (scaled by a plain copy
    kernel); the summary, with its commands, is in
    profiles/hbm_traffic_rowblock_629856x512.md."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_rowblock_629856x512.md")
    try:
        for line in open(path):
            if line.startswith("machine-readable:"):
                fields = dict(tok.split("=") for tok in line.split(":", 1)[1].split())
                return int(fields[f"{kind}_traffic_bytes"])
    except (OSError, KeyError, ValueError):
        pass
    return None


def gram_flops(rows, k):
    """LAPACK Working Note 41 count for a Hermitian k x k Gram over `rows` rows."""
    return 4.0 * rows * k * (k + 1)


def trsm_flops(rows, k):
    return 4.0 * rows * k * k + 2.0 * rows * k


def summarize(samples, flops_fn):
    ms = np.array([s[0] for s in samples])
    flops = np.array([flops_fn(s[1], s[2]) for s in samples])
    avg_ms = float(ms.mean())
    tflops = float(flops.sum() / (ms.sum() * 1e-3) / 1e12)
    rows = float(sum(s[1] for s in samples))
    return {"launches": len(samples), "avg_ms": avg_ms, "tflops": tflops,
            "flops_per_launch": float(flops.mean()), "ms_per_row": float(ms.sum()) / rows}


def cpu_baseline(k, w, rows, repeats=3, full_rows=10077696):
    """The oracle's cholqr2 (the reference CPU backend's algorithm: row-block
    Grams, host Cholesky, row-block triangular solves, FP64 LAPACK/BLAS) timed
    on this host's cores over a bounded sample of rows of the same workload:
    one untimed call, then the best of `repeats`, end to end and per operation
    (panelmul!, gram, the Cholesky and rdiv_rows!, as the reference's
    benchmarks report them, benchmark/common.jl:12-15). project is timed
    once: its p (p + 1) steps take several times a whole cholqr2 on the CPU.
    rows/s is the sample's rows over its time; the work per row is the same
    at any N (the k x k Cholesky aside), so it stands for the full basis."""
    from oracle import ghost as OG
    from oracle import operators as OO
    from oracle import panelblas as OB
    facts = host_facts()
    threads = (facts.get("blas") or {}).get("threads") or os.cpu_count()
    n1 = 8
    n2 = 128
    n3 = rows // (n1 * n2)
    N = n1 * n2 * n3
    apply = lambda X: OO.helmholtz_apply(X, (n1, n2, n3), OO.helmholtz_diagonal())
    omega = OG.matrix(np.complex128, N, k, 0xDEADBEEF)
    Y0 = OB.panelmul(apply, omega, w)

    def best_of(fn, n=repeats, warm=True):
        if warm:
            fn()
        best = None
        for _ in range(n):
            t = time.perf_counter()
            fn()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        return best

    best = best_of(lambda: OB.cholqr2(Y0, w))
    G = OB.gram(Y0, w=w)
    R, _ = OB.cholesky_upper(G, N, k)
    Q, _ = OB.cholqr2(Y0, w)
    per_op = {"panelmul_s": best_of(lambda: OB.panelmul(apply, omega, w)),
              "gram_s": best_of(lambda: OB.gram(Y0, w=w)),
              "cholesky_s": best_of(lambda: OB.cholesky_upper(G, N, k)),
              "rdiv_rows_s": best_of(lambda: OB.rdiv_rows(Y0, R, w)),
              "project_s": best_of(lambda: OB.project(Q, apply, w), n=1, warm=False)}
    return {"value": N / best, "unit": "rows/s", "cores": int(threads), "kind": "port",
            "sample": f"oracle cholqr2 (numpy + OpenBLAS/LAPACK FP64, the reference's step order) of a {N}x{k} "
                      f"ComplexF64 Helmholtz-applied ghost basis in panels of {w}: {N} of the {full_rows} rows "
                      f"({100 * N / full_rows:.1f}%), one untimed call then best of {repeats}: {best:.2f} s, "
                      f"so {N / best:.0f} rows/s; {threads} BLAS threads",
            "sample_rows": N, "sample_s": best, "host": facts,
            "per_operation": {key: round(v, 4) for key, v in per_op.items()}}


def host_facts():
    """The host the CPU baseline runs on: CPU model, core counts, memory,
    the page-lock limit of this process, and the BLAS numpy calls."""
    import platform
    import resource
    import subprocess
    facts = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
             "machine": platform.machine()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            if key.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                facts[key.strip().lower().replace(" ", "_").replace("(s)", "s")] = val.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal:"):
                facts["mem_total_GiB"] = round(int(line.split()[1]) / 2 ** 20, 1)
    except OSError:
        pass
    soft, hard = resource.getrlimit(resource.RLIMIT_MEMLOCK)
    facts["memlock_limit"] = "unlimited" if soft == resource.RLIM_INFINITY else soft
    try:
        from threadpoolctl import threadpool_info
        blas = [i for i in threadpool_info() if i.get("user_api") == "blas"]
        if blas:
            b = blas[0]
            facts["blas"] = {"library": b.get("internal_api"), "version": b.get("version"),
                             "threads": b.get("num_threads"), "architecture": b.get("architecture"),
                             "file": os.path.basename(b.get("filepath", ""))}
    except Exception:
        pass
    return facts


def pinned_probe(limit_gib=160, chunk_gib=8):
    """How much host memory this process can page-lock: fhip_host_alloc in
    chunk_gib pieces until a call fails or limit_gib is reached, then every
    piece freed. Prints one JSON line with the host's facts."""
    import ctypes
    import funicular_hip as F
    from funicular_hip import _lib
    F.hip_backend()
    lib = _lib.load()
    held, error = [], None
    t = time.perf_counter()
    while len(held) * chunk_gib < limit_gib:
        p = ctypes.c_void_p()
        st = lib.fhip_host_alloc(ctypes.byref(p), chunk_gib << 30)
        if st != 0:
            error = lib.fhip_last_error().decode(errors="replace")
            break
        held.append(p.value)
        log(f"pinned-probe: {len(held) * chunk_gib} GiB page-locked")
    dt = time.perf_counter() - t
    locked = len(held) * chunk_gib
    for p in held:
        lib.fhip_host_free(p)
    print(json.dumps({"page_locked_GiB": locked, "limit_tried_GiB": limit_gib, "chunk_GiB": chunk_gib,
                      "stopped_by": "error" if error else "limit", "error": error,
                      "seconds": round(dt, 1), "host": host_facts()}), flush=True)


def residual(backend, Y, R, G, omega):
    """||Y0 - Q R||_F / ||Y0||_F, panel by panel, with Y0 = G Omega regenerated
    from the ghost seed and Q the device-resident result."""
    N, k, w = Y.N, Y.k, Y.w
    Rd = backend.todevice(np.asfortranarray(R))
    bufs = [backend.alloc_device(np.complex128, (N, w)) for _ in range(3)]
    g = backend.alloc_device(np.complex128, (w, w))
    num = den = 0.0
    for j in range(0, k, w):
        c1 = min(j + w, k)
        b1, b2, b3 = (b.view(None, (0, c1 - j)) for b in bufs)
        gv = g.view((0, c1 - j), (0, c1 - j))
        backend.ghost_fill(b3, 0, j + 1, omega.panels[0].home.source.seed)
        G.apply(b1, b3)
        backend.rightmul_gemm(b2, Y.device.view(None, (0, c1)), Rd.view((0, c1), (j, c1)))
        backend.gram_accumulate(gv, b1, b1, 1.0, 0.0)
        den += float(np.trace(backend.to_host(gv)).real)
        backend.axpby(b2, b1, -1.0, 1.0)
        backend.gram_accumulate(gv, b2, b2, 1.0, 0.0)
        num += float(np.trace(backend.to_host(gv)).real)
    return float(np.sqrt(num / den))


def project_check(S, d):
    """S = Q^H G Q for the Helmholtz operator G = d I + (real symmetric
    stencil) and orthonormal Q satisfies S - S^H = 2 i Im(d) Q^H Q, so
    ||S - S^H - 2 i Im(d) I||_F / ||S||_F measures S at full size."""
    k = S.shape[0]
    return float(np.linalg.norm(S - S.conj().T - 2j * d.imag * np.eye(k)) / np.linalg.norm(S))


def adjoint_bench(backend):
    """The adjoint workload: Y (4,194,304 x 128 ComplexF64, ghost panels) ->
    Q R = Y -> S = Q^H G Q with G = diag(d) + the periodic tridiagonal
    (2 + 0.1i, -1), d_i = 1 + 0.5 sin(2 pi i / N); loss
    Re<dQ, Q> + Re<dS, S> for a seeded complex normal dQ and the Hermitian
    part of a seeded complex normal dS. Times the forward pass (cholqr2 +
    project, through their rrules) against the backward pass (both
    pullbacks), and checks <Ybar, E> against a central difference of the
    device forward pass along a ghost direction E."""
    import funicular_hip as F
    from funicular_hip.executor import readpanels, sweep, updatepanels, writepanels
    N, k, w = 1 << 22, 128, 32
    plan = F.ResidencyPlan(backend=backend, device_budget=200 * 1024 ** 3, host_budget=1 << 30,
                           device_resident=True)
    d = 1.0 + 0.5 * np.sin(2 * np.pi * np.arange(N) / N)
    G = F.Tridiagonal1D(N, 2 + 0.1j, -1.0, d, backend=backend)

    def ghost_copy(seed):
        src = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=seed, w=w)
        dst = F.PanelMatrix(N, k, plan=plan, w=w)
        sweep(lambda _, a, b: backend.convert_copy(b, a), readpanels(src), writepanels(dst))
        return dst

    rng = np.random.default_rng(4)
    A = rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))
    dS = (A + A.conj().T) / 2
    dQ = ghost_copy(0xDEADBEEF + 4)

    def forward(Y):
        R, pb_qr = F.rrule_cholqr2(Y)
        S, pb_pr = F.rrule_project(Y, G)
        return R, S, pb_qr, pb_pr

    def loss(Y, S):
        return float(np.real(F.gram(dQ, Y).trace()) + np.real(np.vdot(dS, S)))

    def backward(pb_qr, pb_pr):
        # Qbar goes back to the caller too, so that it is released outside
        # the timed region.
        Qbar, dbar = pb_pr(dS)
        sweep(lambda _, a, b: backend.axpby(b, a, 1.0, 1.0), readpanels(dQ), updatepanels(Qbar))
        return pb_qr(Qbar, None), dbar, Qbar

    # One untimed repetition warms the pools and the allocator, then the best
    # of three timed repetitions of each direction.
    import gc
    fwd, bwd = [], []
    R = S = pb_qr = pb_pr = Ybar = Qbar = None
    for rep in range(4):
        # The previous repetition's matrices are released before the clock starts.
        del R, S, pb_qr, pb_pr
        for pm in (Ybar, Qbar):
            if pm is not None:
                F.free(pm)
        Ybar = Qbar = None
        gc.collect()
        Y = ghost_copy(0xDEADBEEF + 3)
        backend.sync_all()
        t = time.perf_counter()
        R, S, pb_qr, pb_pr = forward(Y)
        backend.sync_all()
        t_f = time.perf_counter() - t
        t = time.perf_counter()
        Ybar, dbar, Qbar = backward(pb_qr, pb_pr)
        backend.sync_all()
        t_b = time.perf_counter() - t
        if rep > 0:
            fwd.append(t_f)
            bwd.append(t_b)
    t_fwd, t_bwd = min(fwd), min(bwd)
    log(f"adjoint: N={N} k={k}: forward {t_fwd:.3f} s, backward {t_bwd:.3f} s ({t_bwd / t_fwd:.2f}x)")

    # Central difference along E of the device forward pass.
    E = ghost_copy(0xDEADBEEF + 5)
    h = 1e-6

    def shifted(sign):
        Yh = ghost_copy(0xDEADBEEF + 3)
        sweep(lambda _, e, y: backend.axpby(y, e, sign * h, 1.0), readpanels(E), updatepanels(Yh))
        F.cholqr2(Yh)
        return loss(Yh, F.project(Yh, G))

    fd = (shifted(1.0) - shifted(-1.0)) / (2 * h)
    directional = float(np.real(F.gram(Ybar, E).trace()))
    rel = abs(fd - directional) / max(abs(fd), 1e-300)
    # And along a direction e of the diagonal parameters.
    e = np.random.default_rng(5).standard_normal(N)

    def shifted_d(sign):
        Gh = F.Tridiagonal1D(N, 2 + 0.1j, -1.0, d + sign * h * e, backend=backend)
        return float(np.real(np.vdot(dS, F.project(Y, Gh))))

    fd_d = (shifted_d(1.0) - shifted_d(-1.0)) / (2 * h)
    rel_d = abs(fd_d - float(dbar @ e)) / max(abs(fd_d), 1e-300)
    log(f"adjoint: <Ybar, E> vs central difference: rel {rel:.2e}; dS/dd along e: rel {rel_d:.2e}")
    return {"workload": f"cholqr2 + project of a {N}x{k} ComplexF64 ghost basis, G = diag(d) + tridiagonal",
            "forward_s": t_fwd, "backward_s": t_bwd, "backward_over_forward": t_bwd / t_fwd,
            "fd_rel_err_Y": rel, "fd_rel_err_d": rel_d, "fd_step": h}


def mixed_bench(backend, steps):
    """The mixed-precision workload: Y = U diag(sigma) V^H with U the Q of
    cholqr2 of an 8,388,608 x 256 ComplexF64 ghost basis (seed 0xdeadbeef),
    V the unitary Q of a seeded 256 x 256 complex normal, and sigma
    log-spaced from 1 to 1e-2 and, in a second case, to 1e-4. For each case,
    FP64-only cholqr2 against the mixed first pass, with orthogonality and
    the residual ||Y - Q R|| / ||Y|| measured on the device."""
    import funicular_hip as F
    from funicular_hip.executor import readpanels, sweep, writepanels
    from funicular_hip.panelblas import cholqr2_mixed
    N, k, w = 8388608, 256, 16
    plan = F.ResidencyPlan(backend=backend, device_budget=200 * 1024 ** 3, host_budget=1 << 30,
                           device_resident=True)
    U = F.PanelMatrix(N, k, plan=plan, w=w)
    src = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    sweep(lambda _, a, b: backend.convert_copy(b, a), readpanels(src), writepanels(U))
    F.cholqr2(U)
    rng = np.random.default_rng(0xDEADBEEF + 1)
    V = np.linalg.qr(rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k)))[0]
    Y0 = F.PanelMatrix(N, k, plan=plan, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    out = {"workload": f"cholqr2 of Y = U diag(sigma) V^H, {N}x{k} ComplexF64, panels of {w}"}

    def copy(dst, src_):
        sweep(lambda _, a, b: backend.convert_copy(b, a), readpanels(src_), writepanels(dst))

    def resid(R):
        # ||Y0 - Q R||_F / ||Y0||_F with Q in Y, through one more product.
        Z = F.PanelMatrix(N, k, plan=plan, w=w)
        F.rightmul(Z, Y, R)
        sweep(lambda _, a, b: backend.axpby(b, a, -1.0, 1.0), readpanels(Y0), F.updatepanels(Z))
        num = float(np.real(np.trace(F.gram(Z))))
        F.free(Z)
        return float(np.sqrt(num / np.real(np.trace(F.gram(Y0)))))

    from funicular_hip.executor import trace_sweeps
    for decades in (2, 4):
        sigma = 10.0 ** (-decades * np.arange(k) / (k - 1))
        F.rightmul(Y0, U, sigma[:, None] * V.conj().T)
        case = {}
        # The configuration's stated shift, 11 (N k + k (k + 1)) u32 ||Y||_2^2,
        # always applied ("fukaya"), next to the adaptive default; a shifted
        # first pass makes a third (FP64) pass run.
        for name, fn in (("fp64", F.cholqr2), ("mixed", cholqr2_mixed),
                         ("mixed_fp64_solve", lambda Z: cholqr2_mixed(Z, solve="fp64")),
                         ("mixed_fukaya", lambda Z: cholqr2_mixed(Z, shift="fukaya")),
                         ("mixed_fukaya_fp64_solve", lambda Z: cholqr2_mixed(Z, shift="fukaya", solve="fp64"))):
            copy(Y, Y0)
            with trace_sweeps() as sweeps:
                R = fn(Y)  # accuracy from the first run
            orth = float(np.linalg.norm(F.gram(Y) - np.eye(k)))
            res = resid(R)
            times = []
            for _ in range(steps):
                copy(Y, Y0)
                backend.sync_all()
                t = time.perf_counter()
                fn(Y)
                backend.sync_all()
                times.append(time.perf_counter() - t)
            case[name] = {"s": min(times), "orthogonality": orth, "residual": res, "passes": len(sweeps) // 2}
        for name in ("mixed", "mixed_fp64_solve", "mixed_fukaya", "mixed_fukaya_fp64_solve"):
            case[name]["speedup"] = case["fp64"]["s"] / case[name]["s"]
        case["speedup"] = case["mixed"]["speedup"]
        case["speedup_fp64_solve"] = case["mixed_fp64_solve"]["speedup"]
        out[f"sigma_to_1e-{decades}"] = case
        log(f"mixed workload, sigma to 1e-{decades}: {case}")
    for pm in (U, Y0, Y):
        F.free(pm)
    return out


def disk_bench(backend, scratch):
    """The disk tier under a panel sweep: Y = G Omega for a 4,194,304 x 256
    ComplexF64 basis in panels of 32 (8 panels, 17.2 GB) with a host budget
    of four panels, so every output panel goes to the scratch file and a
    read sweep (one 32 x 32 Gram per panel) brings each back, and
    cholqr2_panelwise orthonormalizes it panel pair by panel pair (the row
    sweeps of cholqr2 would need all eight panels host resident); the same
    steps with the whole basis in host memory give the cost of the tier.
    The scratch file sits on the machine's local disk, behind the page cache."""
    import funicular_hip as F
    from funicular_hip.executor import readpanels, sweep
    N, k, w = 1 << 22, 256, 32
    G = F.Tridiagonal1D(N, 2 + 0.1j, -1.0, backend=backend)
    panel = N * w * 16
    out = {"workload": f"panelmul + read sweep of a {N}x{k} ComplexF64 basis, panels of {w}",
           "panel_GB": panel / 1e9}
    for label, budget, sdir in (("disk", 4 * panel + (1 << 20), scratch), ("host", 9 * panel, None)):
        plan = F.ResidencyPlan(backend=backend, device_budget=64 * 1024 ** 3, host_budget=budget,
                               scratch_dir=sdir)
        omega = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=0xDEADBEEF, w=w)
        Y = F.PanelMatrix(N, k, plan=plan, w=w)
        C = backend.alloc_device(np.complex128, (w, w))
        backend.sync_all()
        t = time.perf_counter()
        F.panelmul(Y, G, omega)
        backend.sync_all()
        t_mul = time.perf_counter() - t
        reads0 = F.disk_reads(Y)
        t = time.perf_counter()
        sweep(lambda _, b: backend.gram_accumulate(C, b, b, 1.0, 0.0), readpanels(Y))
        backend.sync_all()
        t_read = time.perf_counter() - t
        rec = {"panelmul_s": t_mul, "read_sweep_s": t_read}
        if sdir is not None:
            rec.update({"disk_writes": F.disk_writes(Y), "disk_reads_in_read_sweep": F.disk_reads(Y) - reads0,
                        "write_GBps": F.disk_writes(Y) * panel / t_mul / 1e9,
                        "read_GBps": (F.disk_reads(Y) - reads0) * panel / t_read / 1e9})
        # cholqr2 by column panels: the row sweeps would need all 8 panels
        # host resident, which the disk case's budget of 4 does not allow.
        reads1, writes1 = F.disk_reads(Y), F.disk_writes(Y)
        t = time.perf_counter()
        F.cholqr2_panelwise(Y)
        backend.sync_all()
        t_qr = time.perf_counter() - t
        orth = float(np.linalg.norm(F.gram_panelwise(Y) - np.eye(k)))
        rec.update({"cholqr2_panelwise_s": t_qr, "cholqr2_panelwise_orthogonality": orth})
        if sdir is None:
            # The row sweeps, which this budget allows, on the same basis.
            F.panelmul(Y, G, omega)
            backend.sync_all()
            t = time.perf_counter()
            F.cholqr2(Y)
            backend.sync_all()
            rec["cholqr2_rows_s"] = time.perf_counter() - t
        if sdir is not None:
            rec.update({"cholqr2_panelwise_disk_reads": F.disk_reads(Y) - reads1,
                        "cholqr2_panelwise_disk_writes": F.disk_writes(Y) - writes1})
        out[label] = rec
        log(f"disk tier ({label}): panelmul {t_mul:.2f} s, read sweep {t_read:.2f} s, {rec}")
        F.free(Y)
        del plan, omega, Y, C
        import gc
        gc.collect()
    return out


def range_finder_bench(backend, scratch, n=160, k=256, w=32, iterations=2):
    """The disk-tier randomized range finder at a reduced size: G = the 3-D
    7-point Helmholtz stencil on an n^3 grid, Omega = GhostPanels, Y = G Omega,
    then `iterations` power steps of (Z = G^H Y, orthonormalize Z, Y = G Z,
    orthonormalize Y), every orthonormalization cholqr2_panelwise with the
    mixed first pass. Y and Z live on the scratch disk behind a host budget of
    2 (FEED_DEPTH + 2) = 8 panels (what a panel sweep over two matrices may pin
    at once) against their 16, so the row sweeps of cholqr2 could not run at
    all. Reports the time of each step, the disk traffic and ||Q^H Q - I||_F."""
    import funicular_hip as F
    from funicular_hip.plan import FEED_DEPTH
    N = n ** 3
    panel = N * w * 16
    plan = F.ResidencyPlan(backend=backend, device_budget=64 * 1024 ** 3,
                           host_budget=2 * (FEED_DEPTH + 2) * panel + (1 << 20), scratch_dir=scratch, panel_width=w)
    d = 6.0 - 0.09 * (1 + 0.05j)
    G = F.Helmholtz3D(n, n, n, d, -1.0, backend=backend)
    Gh = G.adjoint()
    omega = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    Z = F.PanelMatrix(N, k, plan=plan, w=w)
    steps = []

    def step(name, fn):
        r0 = F.disk_reads(Y) + F.disk_reads(Z)
        w0 = F.disk_writes(Y) + F.disk_writes(Z)
        backend.sync_all()
        t = time.perf_counter()
        fn()
        backend.sync_all()
        dt = time.perf_counter() - t
        reads = F.disk_reads(Y) + F.disk_reads(Z) - r0
        writes = F.disk_writes(Y) + F.disk_writes(Z) - w0
        steps.append({"step": name, "s": dt, "disk_read_GB": reads * panel / 1e9,
                      "disk_write_GB": writes * panel / 1e9})
        log(f"range finder: {name} {dt:.2f} s, disk {reads * panel / 1e9:.0f} GB read, "
            f"{writes * panel / 1e9:.0f} GB written")

    t0 = time.perf_counter()
    step("Y = G Omega", lambda: F.panelmul(Y, G, omega))
    step("cholqr2(Y)", lambda: F.cholqr2_panelwise(Y, mixed=True))
    for it in range(iterations):
        step(f"Z = G^H Y ({it + 1})", lambda: F.panelmul(Z, Gh, Y))
        step(f"cholqr2(Z) ({it + 1})", lambda: F.cholqr2_panelwise(Z, mixed=True))
        step(f"Y = G Z ({it + 1})", lambda: F.panelmul(Y, G, Z))
        step(f"cholqr2(Y) ({it + 1})", lambda: F.cholqr2_panelwise(Y, mixed=True))
    total = time.perf_counter() - t0
    orth = float(np.linalg.norm(F.gram_panelwise(Y) - np.eye(k)))
    read = sum(s["disk_read_GB"] for s in steps)
    written = sum(s["disk_write_GB"] for s in steps)
    out = {"workload": f"range finder on a {n}^3 Helmholtz grid: N={N}, k={k}, panels of {w} ({panel / 1e9:.2f} GB), "
                       f"{iterations} power iterations, mixed first pass, host budget 8 of their 16 panels",
           "total_s": total, "steps": steps, "disk_read_GB": read, "disk_write_GB": written,
           "disk_GBps": (read + written) / total, "orthogonality": orth}
    log(f"range finder: {total:.1f} s, ||Q^H Q - I||_F = {orth:.2e}")
    for m in (Y, Z):
        F.free(m)
    return out


def vendor_timings(backend, A, B, C, R, rows, k, timed):
    """rocBLAS on the same shapes, for comparison: zherk and a plain zgemm
    (A^H A, as the reference's CUDA path forms its Gram) for the Gram, ztrsm
    (right, upper, no transpose) for the solve, zgemm for the row-block
    product. Same HIP-event timing as the kernels above; TFLOP/s of zgemm_gram
    count the full 8 N k^2 flops it performs."""
    import ctypes
    lib = ctypes.CDLL("librocblas.so")
    handle = ctypes.c_void_p()
    if lib.rocblas_create_handle(ctypes.byref(handle)) != 0:
        raise RuntimeError("rocblas_create_handle failed")
    lib.rocblas_set_stream(handle, None)
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    one_r, zero_r = ctypes.c_double(1.0), ctypes.c_double(0.0)
    one_c, zero_c = (ctypes.c_double * 2)(1.0, 0.0), (ctypes.c_double * 2)(0.0, 0.0)
    FILL_UPPER, OP_N, OP_C, SIDE_R, DIAG_N = 121, 111, 113, 142, 131

    def check(st):
        if st != 0:
            raise RuntimeError(f"rocBLAS status {st}")
    lib.rocblas_zherk.argtypes = [vp, i32, i32, i32, i32, vp, vp, i32, vp, vp, i32]
    lib.rocblas_ztrsm.argtypes = [vp, i32, i32, i32, i32, i32, i32, vp, vp, i32, vp, i32]
    lib.rocblas_zgemm.argtypes = [vp, i32, i32, i32, i32, i32, vp, vp, i32, vp, i32, vp, vp, i32]
    timed("rocblas_zherk", lambda: check(lib.rocblas_zherk(handle, FILL_UPPER, OP_C, k, rows, ctypes.byref(one_r),
                                                           A.ptr, A.ld, ctypes.byref(zero_r), C.ptr, C.ld)),
          gram_flops(rows, k))
    # The reference's CUDA path forms the Gram with a plain ZGEMM (A^H B).
    timed("rocblas_zgemm_gram", lambda: check(lib.rocblas_zgemm(handle, OP_C, OP_N, k, k, rows, one_c, A.ptr, A.ld,
                                                                A.ptr, A.ld, zero_c, C.ptr, C.ld)),
          8.0 * rows * k * k)
    timed("rocblas_ztrsm", lambda: check(lib.rocblas_ztrsm(handle, SIDE_R, FILL_UPPER, OP_N, DIAG_N, rows, k, one_c,
                                                           R.ptr, R.ld, B.ptr, B.ld)), trsm_flops(rows, k))
    timed("rocblas_zgemm", lambda: check(lib.rocblas_zgemm(handle, OP_N, OP_N, rows, k, k, one_c, A.ptr, A.ld, R.ptr,
                                                           R.ld, zero_c, B.ptr, B.ld)), 8.0 * rows * k * k)
    backend.sync_all()
    lib.rocblas_destroy_handle(handle)


def kernels_only(rows, k, reps, vendor=False):
    """The per-launch shapes of the benchmark, without the pipeline around
    them: one plain copy of a rows x k ComplexF64 block (the calibration kernel
    for the HBM counters), then `reps` FP64 Grams, FP32 (mixed) Grams, FP64 and
    FP32 triangular solves, FP64 row-block products by a k x k factor, and
    the elementwise dot / norm / scale / axpby steps on it. Prints each
    kernel's average duration from HIP events (one JSON line).
    For profiling runs."""
    import funicular_hip as F
    backend = F.hip_backend()
    A = backend.alloc_device(np.complex128, (rows, k))
    B = backend.alloc_device(np.complex128, (rows, k))
    C = backend.alloc_device(np.complex128, (k, k))
    backend.ghost_fill(A, 0, 1, 7)
    backend.convert_copy(B, A)
    backend.sync_all()
    import ctypes
    from funicular_hip import _lib
    out = {"rows": rows, "k": k}

    def event():
        e = ctypes.c_void_p()
        _lib.call("fhip_event_create", ctypes.byref(e), 0)
        return e

    def timed(name, fn, flops, nbytes=None):
        # The launches go to the default stream, where the events are recorded.
        fn()  # untimed first launch
        backend.sync_all()
        e0, e1 = event(), event()
        _lib.call("fhip_event_record", e0, None)
        for _ in range(reps):
            fn()
        _lib.call("fhip_event_record", e1, None)
        _lib.call("fhip_event_sync", e1)
        ms = ctypes.c_float()
        _lib.call("fhip_event_elapsed", ctypes.byref(ms), e0, e1)
        for e in (e0, e1):
            _lib.call("fhip_event_destroy", e)
        t = ms.value / reps
        out[name] = {"ms": t, "tflops": flops / t / 1e9} if nbytes is None else {"ms": t, "gbps": nbytes / t / 1e6}

    for compute, name in ((None, "gram_fp64"), (np.complex64, "gram_fp32")):
        timed(name, lambda: backend.gram_accumulate(C, A, A, 1.0, 0.0, compute=compute), gram_flops(rows, k))
    G = backend.to_host(C)
    R = backend.todevice(np.asfortranarray(np.linalg.cholesky(G).conj().T))
    for compute, name in ((None, "trsm_fp64"), (np.complex64, "trsm_fp32")):
        timed(name, lambda: backend.rdiv_upper(B, R, compute=compute), trsm_flops(rows, k))
    timed("rightmul_fp64", lambda: backend.rightmul_gemm(B, A, R), 8.0 * rows * k * k)
    # The elementwise panel BLAS steps, HBM-bound: GB/s of algorithmic bytes
    # (each operand read once, the output written once).
    blk = rows * k * 16
    work = backend.alloc_device(np.uint8, (backend.dot_workspace_bytes(rows, k), 1))
    acc = backend.alloc_device(np.complex128, (1, 1))
    timed("dot_fp64", lambda: backend.dot_accumulate(acc, A, B, False, work), 0, 2 * blk)
    timed("norm_fp64", lambda: backend.dot_accumulate(acc, A, None, False, work), 0, blk)
    timed("scale_fp64", lambda: backend.scale(B, 0.6 + 0.8j), 0, 2 * blk)
    timed("axpby_fp64", lambda: backend.axpby(B, A, 0.5, 1.0), 0, 3 * blk)
    # The ghost fill of the block (written once, 16 B per entry).
    timed("ghost_fp64", lambda: backend.ghost_fill(B, 0, 1, 7), 0, blk)
    # The stencils on one panel of their workloads (HBM-bound; ideal bytes:
    # the panel read once and the image written once): the 216^3 Helmholtz
    # operator on 32 columns and the 2^22-row tridiagonal on 32 columns.
    for name, Gop, n_rows in (("helmholtz_fp64", F.Helmholtz3D(216, 216, 216, 6.0 - 0.09 * (1 + 0.05j), -1.0,
                                                                backend=backend), 216 ** 3),
                              ("tridiag_fp64", F.Tridiagonal1D(2 ** 22, 2 + 0.1j, -1.0, backend=backend), 2 ** 22)):
        P = backend.alloc_device(np.complex128, (n_rows, 32))
        Q = backend.alloc_device(np.complex128, (n_rows, 32))
        backend.ghost_fill(P, 0, 1, 7)
        timed(name, lambda: Gop.apply(Q, P), 0, 2 * n_rows * 32 * 16)
        backend.sync_all()
        del P, Q
    backend.sync_all()
    if vendor:
        vendor_timings(backend, A, B, C, R, rows, k, timed)
    log(f"kernels-only: {rows}x{k}, {reps} repetitions of each kernel")
    print(json.dumps(out), flush=True)


class CopyTimer:
    """HIP timing events around every staging copy a sweep issues (uploads
    on the sweep's upload stream, write-backs on its download stream), so a
    streamed operation's copies can be laid against its kernels on one time
    axis."""

    def __init__(self):
        import ctypes
        from funicular_hip import _lib
        from funicular_hip import executor
        self.ctypes = ctypes
        self._lib = _lib
        self.executor = executor
        self.records = []
        self.active = False
        self._orig_in = executor.stage_in
        self._orig_out = executor.stage_out
        timer = self

        def stage_in(t, dst, pm, step, queue, via=None):
            ev = timer._record(queue)
            r = timer._orig_in(t, dst, pm, step, queue, via=via)
            timer._done(ev, queue, "up")
            return r

        def stage_out(t, pm, src, step, queue, via=None):
            ev = timer._record(queue)
            r = timer._orig_out(t, pm, src, step, queue, via=via)
            timer._done(ev, queue, "down")
            return r
        executor.stage_in = stage_in
        executor.stage_out = stage_out

    def _event(self, stream):
        e = self.ctypes.c_void_p()
        self._lib.call("fhip_event_create", self.ctypes.byref(e), 0)
        self._lib.call("fhip_event_record", e, stream)
        return e.value

    def _record(self, queue):
        if not self.active or queue is None:
            return None
        return self._event(queue.stream)

    def _done(self, start, queue, kind):
        if start is not None:
            self.records.append((start, self._event(queue.stream), kind))


def _intervals(lib, ctypes, origin, records):
    """[(start_ms, stop_ms)] of event pairs, measured from `origin`."""
    out = []
    for start, stop in records:
        lib.call("fhip_event_sync", stop)
        a, b = ctypes.c_float(), ctypes.c_float()
        lib.call("fhip_event_elapsed", ctypes.byref(a), origin, start)
        lib.call("fhip_event_elapsed", ctypes.byref(b), origin, stop)
        out.append((a.value, b.value))
        lib.call("fhip_event_destroy", start)
        lib.call("fhip_event_destroy", stop)
    return out


def _union(iv):
    merged = []
    for a, b in sorted(iv):
        if merged and a <= merged[-1][1]:
            merged[-1][1] = max(merged[-1][1], b)
        else:
            merged.append([a, b])
    return merged


def _length(merged):
    return sum(b - a for a, b in merged)


def _overlap(m1, m2):
    total, i, j = 0.0, 0, 0
    while i < len(m1) and j < len(m2):
        lo, hi = max(m1[i][0], m2[j][0]), min(m1[i][1], m2[j][1])
        total += max(0.0, hi - lo)
        if m1[i][1] < m2[j][1]:
            i += 1
        else:
            j += 1
    return total


def overlapped(backend, kernels, copies, fn):
    """Runs fn with the kernel and copy timers on and returns (wall seconds,
    result, timeline) where the timeline gives, from the HIP events on the
    streams the work ran on, the time the kernels (Gram and solve) and the
    copies occupied, and how much of each the other covered: hidden_copy is
    the fraction of copy time that ran under kernels, hidden_compute the
    fraction of kernel time that ran under copies. Both are fractions of
    measured intervals, so they lie in [0, 1]."""
    import ctypes
    from funicular_hip import _lib
    backend.sync_all()
    origin = ctypes.c_void_p()
    _lib.call("fhip_event_create", ctypes.byref(origin), 0)
    _lib.call("fhip_event_record", origin, None)
    kernels.active = copies.active = True
    t = time.perf_counter()
    out = fn()
    backend.sync_all()
    wall = time.perf_counter() - t
    kernels.active = copies.active = False
    kiv = []
    for kind in ("gram", "trsm"):
        kiv += [(a, b) for a, b, _ in kernels.records[kind]]
    kernels.records = {"gram": [], "trsm": []}
    k_iv = _union(_intervals(_lib, ctypes, origin.value, kiv))
    up = _union(_intervals(_lib, ctypes, origin.value, [(a, b) for a, b, kind in copies.records if kind == "up"]))
    down = _union(_intervals(_lib, ctypes, origin.value, [(a, b) for a, b, kind in copies.records if kind == "down"]))
    copies.records = []
    c_iv = _union([tuple(x) for x in up + down])
    _lib.call("fhip_event_destroy", origin.value)
    t_k, t_c = _length(k_iv) / 1e3, _length(c_iv) / 1e3
    both = _overlap(k_iv, c_iv) / 1e3
    return wall, out, {"kernel_s": t_k, "copy_s": t_c, "upload_s": _length(up) / 1e3,
                       "download_s": _length(down) / 1e3, "overlap_s": both,
                       "hidden_compute": both / t_k if t_k > 0 else None,
                       "hidden_copy": both / t_c if t_c > 0 else None,
                       "link_busy_fraction": t_c / wall if wall > 0 else None}


def link_rates(nbytes=2 << 30, reps=3):
    """The host link's copy rates with page-locked host memory, through the
    library's own copy routine (fhip_memcpy2d_async, the copy engines):
    nbytes up alone, nbytes down alone, and nbytes each way at once on two
    streams. Best of `reps`, in GB/s."""
    import ctypes
    from funicular_hip import _lib
    lib = _lib.load()
    h_up, h_down, d_a, d_b = (ctypes.c_void_p() for _ in range(4))
    _lib.call("fhip_host_alloc", ctypes.byref(h_up), nbytes)
    _lib.call("fhip_host_alloc", ctypes.byref(h_down), nbytes)
    _lib.call("fhip_malloc", ctypes.byref(d_a), nbytes)
    _lib.call("fhip_malloc", ctypes.byref(d_b), nbytes)
    s1, s2 = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.call("fhip_stream_create", ctypes.byref(s1))
    _lib.call("fhip_stream_create", ctypes.byref(s2))

    def up():
        _lib.call("fhip_memcpy2d_async", d_a, nbytes, h_up, nbytes, nbytes, 1, _lib.FHIP_H2D, s1)

    def down():
        _lib.call("fhip_memcpy2d_async", h_down, nbytes, d_b, nbytes, nbytes, 1, _lib.FHIP_D2H, s2)

    def best(*fns):
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            for fn in fns:
                fn()
            _lib.call("fhip_stream_sync", s1)
            _lib.call("fhip_stream_sync", s2)
            ts.append(time.perf_counter() - t)
        return min(ts)

    up(); down()  # first touch
    _lib.call("fhip_stream_sync", s1)
    _lib.call("fhip_stream_sync", s2)
    t_up, t_down, t_both = best(up), best(down), best(up, down)
    for s in (s1, s2):
        _lib.call("fhip_stream_destroy", s)
    for d in (d_a, d_b):
        _lib.call("fhip_free", d)
    for h in (h_up, h_down):
        _lib.call("fhip_host_free", h)
    return {"bytes_each_way": nbytes, "h2d_alone_GBps": nbytes / t_up / 1e9, "d2h_alone_GBps": nbytes / t_down / 1e9,
            "both_at_once_GBps": 2 * nbytes / t_both / 1e9}


# Sweeps of one cholqr2 that only upload (Gram sweeps) and that upload and
# download (solve sweeps), in basis sizes, per schedule: the reference's two
# passes are a Gram sweep and a solve sweep each; the fused schedule folds the
# second pass's Gram into the first solve sweep.
CHOLQR2_SWEEPS = {"reference": (2, 2), "fused": (1, 2)}


STREAMED_GRIDS = (200, 180)


def streamed(args, phase_gap=0.0, grids=STREAMED_GRIDS, project_reference=False):
    """Configuration 2's own regime: FP64 host storage, the basis in
    page-locked host memory, a 64 GiB device budget, nbuffers = 2, every
    sweep streaming row blocks up (and, for the solve, back down) with the
    copies on their own streams fenced by events. The boxes page-lock at most
    64 GiB of host memory, below the 216^3 x 512 basis (82.6 GB), so the run
    takes the largest grid of `grids` whose basis page-locks under a 62 GiB
    host budget. Reports, for the reference schedule and the fused one, the
    PCIe-inclusive cholqr2 rate, the bytes and rates over the link, and the
    kernel/copy timeline (HIP events on the sweep's streams); for a Gram sweep
    also (t_copy + t_compute - t_wall) / min(t_copy, t_compute) with t_copy a
    copy-only sweep at the same nbuffers. phase_gap > 0 leaves the GPU idle
    that many seconds between phases, so a trace can be cut at the gaps."""
    import gc
    import funicular_hip as F
    from funicular_hip.executor import readpanels, rowtraversal, sweep, updatepanels
    backend = F.hip_backend()
    k, w = args.k, args.w
    # The link's own rates, before the basis takes the page-lockable memory.
    link = link_rates()
    log(f"streamed: link alone {link['h2d_alone_GBps']:.1f} GB/s up, {link['d2h_alone_GBps']:.1f} GB/s down, "
        f"{link['both_at_once_GBps']:.1f} GB/s both ways at once")
    tried = []
    plan = Y = None
    for n in grids:
        N = n ** 3
        plan = F.ResidencyPlan(backend=backend, device_budget=64 * 1024 ** 3, host_budget=62 * 1024 ** 3)
        t = time.perf_counter()
        try:
            Y = F.PanelMatrix(N, k, plan=plan, w=w)
            break
        except Exception as err:  # page-locking refused: try the next grid
            tried.append({"grid": n, "error": str(err)[:200]})
            log(f"streamed: a {n}^3 x {k} basis did not page-lock ({err}); trying a smaller grid")
            Y = plan = None
            gc.collect()
    if Y is None:
        return {"error": "no grid page-locked", "tried": tried}
    log(f"streamed: {N}x{k} host-resident basis allocated and page-locked in {time.perf_counter() - t:.1f} s")
    d = 6.0 - 0.09 * (1 + 0.05j)
    G = F.Helmholtz3D(n, n, n, d, -1.0, backend=backend)
    omega = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    c0 = F._lib.counters()
    t = time.perf_counter()
    F.panelmul(Y, G, omega)
    backend.sync_all()
    t_panelmul = time.perf_counter() - t
    c1 = F._lib.counters()
    down_pm = c1["d2h_bytes"] - c0["d2h_bytes"]
    log(f"streamed: panelmul {t_panelmul:.2f} s, {down_pm / t_panelmul / 1e9:.1f} GB/s down")
    traversal = rowtraversal(Y)

    def timed(fn):
        backend.sync_all()
        if phase_gap > 0:
            time.sleep(phase_gap)
        c0 = F._lib.counters()
        t = time.perf_counter()
        out = fn()
        backend.sync_all()
        dt = time.perf_counter() - t
        c1 = F._lib.counters()
        return dt, c1["h2d_bytes"] - c0["h2d_bytes"], c1["d2h_bytes"] - c0["d2h_bytes"], out

    noop = lambda *a: None
    t_copy1, up1, _, _ = timed(lambda: sweep(noop, readpanels(Y), nbuffers=1, traversal=traversal))
    t_copy2, up2, _, _ = timed(lambda: sweep(noop, readpanels(Y), nbuffers=2, traversal=traversal))
    # Each row block up and back down unchanged, no kernels: the link's rate
    # both ways at once in the sweeps' own copy pattern.
    t_duplex2, upd, downd, _ = timed(lambda: sweep(noop, updatepanels(Y), nbuffers=2, traversal=traversal))
    kernels = KernelTimer(backend)
    copies = CopyTimer()
    if phase_gap > 0:
        time.sleep(phase_gap)
    t_gram, _, gram_tl = overlapped(backend, kernels, copies, lambda: F.gram(Y))
    t_compute = gram_tl["kernel_s"]
    formula = (t_copy2 + t_compute - t_gram) / min(t_copy2, t_compute)
    log(f"streamed: copy-only sweep nb=1 {t_copy1:.2f} s ({up1 / t_copy1 / 1e9:.1f} GB/s), nb=2 {t_copy2:.2f} s "
        f"({up2 / t_copy2 / 1e9:.1f} GB/s), up and down {t_duplex2:.2f} s "
        f"({(upd + downd) / t_duplex2 / 1e9:.1f} GB/s); gram sweep {t_gram:.2f} s with {t_compute:.3f} s of kernels, "
        f"{gram_tl['hidden_compute']:.3f} of them under copies")
    out = {"grid": n, "N": N, "k": k, "panel_width": w, "basis_GB": N * k * 16 / 1e9,
           "budgets": "device 64 GiB, host 62 GiB page-locked (the box's limit is 64 GiB), FP64 host storage, "
                      "nbuffers 2",
           "grids_refused": tried, "link": link,
           "panelmul_s": t_panelmul, "d2h_GBps_panelmul": down_pm / t_panelmul / 1e9,
           "copy_only_sweep_nb1_s": t_copy1, "copy_only_sweep_nb2_s": t_copy2,
           "h2d_GBps_copy_only_nb1": up1 / t_copy1 / 1e9, "h2d_GBps_copy_only_nb2": up2 / t_copy2 / 1e9,
           "duplex_copy_only_sweep_nb2_s": t_duplex2,
           "duplex_GBps_copy_only_nb2": (upd + downd) / t_duplex2 / 1e9,
           "gram_sweep": dict(gram_tl, wall_s=t_gram,
                              overlap_formula=formula,
                              overlap_formula_definition="(t_copy + t_compute - t_wall) / min(t_copy, t_compute), "
                                                         "t_copy = copy-only sweep at nbuffers 2")}
    for schedule in ("reference", "fused"):
        if phase_gap > 0:
            time.sleep(phase_gap)
        c0 = F._lib.counters()
        t_chol, R, tl = overlapped(backend, kernels, copies, lambda: F.cholqr2(Y, schedule=schedule))
        c1 = F._lib.counters()
        up, down = c1["h2d_bytes"] - c0["h2d_bytes"], c1["d2h_bytes"] - c0["d2h_bytes"]
        if phase_gap > 0:  # the check's Gram sweep as a phase of its own
            time.sleep(phase_gap)
        orth = float(np.linalg.norm(F.gram(Y) - np.eye(k)))
        # What the link alone needs for this schedule's bytes: its
        # upload-only sweeps and upload-and-download sweeps, each timed as
        # the copy-only sweep of that kind (same traversal, nbuffers 2).
        n_up_only, n_duplex = CHOLQR2_SWEEPS[schedule]
        ideal = n_up_only * t_copy2 + n_duplex * t_duplex2
        out[f"cholqr2_{schedule}"] = dict(tl, s=t_chol, rows_per_s_pcie_inclusive=N / t_chol,
                                          h2d_GB=up / 1e9, d2h_GB=down / 1e9,
                                          h2d_GBps=up / t_chol / 1e9, d2h_GBps=down / t_chol / 1e9,
                                          link_GBps=(up + down) / t_chol / 1e9, orthogonality=orth,
                                          link_bound_s=ideal, fraction_of_link_bound=ideal / t_chol,
                                          link_bound_definition=f"{n_up_only} copy-only upload sweeps + "
                                          f"{n_duplex} copy-only upload-and-download sweeps, nbuffers 2")
        log(f"streamed: cholqr2 ({schedule} schedule) {t_chol:.2f} s, {N / t_chol:.0f} rows/s, "
            f"up {up / 1e9:.1f} GB, down {down / 1e9:.1f} GB, link busy {tl['link_busy_fraction']:.3f}, "
            f"kernels under copies {tl['hidden_compute']:.3f}, {ideal / t_chol:.2f} of the link-bound time "
            f"{ideal:.2f} s, ||Q^H Q - I||_F {orth:.1e}")
    # project on the orthonormal basis: by runs of held panel images (fused),
    # and on request the reference's p (p + 1)-step schedule, which uploads the
    # basis p + 1 times.
    t_proj, up_proj, _, S = timed(lambda: F.project(Y, G, schedule="fused"))
    out["project_fused"] = {"s": t_proj, "h2d_GB": up_proj / 1e9, "check": project_check(S, d)}
    log(f"streamed: project (fused, runs of held images) {t_proj:.2f} s, {up_proj / 1e9:.0f} GB up")
    if project_reference:
        t_ref, up_ref, _, S_ref = timed(lambda: F.project(Y, G, schedule="reference"))
        out["project_reference"] = {"s": t_ref, "h2d_GB": up_ref / 1e9,
                                    "relative_difference": float(np.linalg.norm(S - S_ref) / np.linalg.norm(S_ref))}
        log(f"streamed: project (reference schedule) {t_ref:.2f} s, {up_ref / 1e9:.0f} GB up")
    # Put the sweep functions back for whatever runs next.
    import funicular_hip.executor as ex
    ex.stage_in, ex.stage_out = copies._orig_in, copies._orig_out
    backend.rdiv_upper, backend.rdiv_upper_folded = kernels._orig_trsm, kernels._orig_folded
    backend.gram_accumulate = kernels._orig_gram
    F.free(Y)
    return out


def streamed_narrow(args):
    """The narrowed host tier (host_eltype ComplexF32, src/panels.jl:361-398):
    the full 216^3 x 512 basis is held in page-locked host memory as
    ComplexF32 (41.3 GB, within the 60 GiB budget), widened to ComplexF64 on
    the device as each row block arrives and narrowed on the way back, so the
    FP64 cholqr2 moves half the link bytes. Reports the PCIe-inclusive rate
    and ||Q^H Q - I||_F of the stored (ComplexF32) Q, then times project on
    that Q (runs of held images)."""
    import funicular_hip as F
    backend = F.hip_backend()
    n = args.grid
    N, k, w = n ** 3, args.k, args.w
    plan = F.ResidencyPlan(backend=backend, device_budget=64 * 1024 ** 3, host_budget=60 * 1024 ** 3,
                           host_eltype=np.complex64)
    t = time.perf_counter()
    Y = F.PanelMatrix(N, k, plan=plan, w=w)
    log(f"streamed-narrow: {N}x{k} ComplexF32 host basis allocated in {time.perf_counter() - t:.1f} s")
    d = 6.0 - 0.09 * (1 + 0.05j)
    G = F.Helmholtz3D(n, n, n, d, -1.0, backend=backend)
    omega = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    t = time.perf_counter()
    F.panelmul(Y, G, omega)
    backend.sync_all()
    t_panelmul = time.perf_counter() - t
    # The link's own rate in this process: a copy-only row sweep (nbuffers 1).
    from funicular_hip.executor import readpanels, rowtraversal, sweep, updatepanels
    c0 = F._lib.counters()
    t = time.perf_counter()
    sweep(lambda *a: None, readpanels(Y), nbuffers=1, traversal=rowtraversal(Y))
    backend.sync_all()
    t_copy = time.perf_counter() - t
    h2d_copy = (F._lib.counters()["h2d_bytes"] - c0["h2d_bytes"]) / t_copy / 1e9
    log(f"streamed-narrow: copy-only sweep {t_copy:.2f} s, {h2d_copy:.1f} GB/s up")
    c0 = F._lib.counters()
    t = time.perf_counter()
    F.cholqr2(Y, schedule="fused")
    backend.sync_all()
    t_chol = time.perf_counter() - t
    c1 = F._lib.counters()
    up, down = c1["h2d_bytes"] - c0["h2d_bytes"], c1["d2h_bytes"] - c0["d2h_bytes"]
    orth = float(np.linalg.norm(F.gram(Y) - np.eye(k)))
    log(f"streamed-narrow: cholqr2 {t_chol:.2f} s, {N / t_chol:.0f} rows/s, up {up / 1e9:.1f} GB, "
        f"down {down / 1e9:.1f} GB, ||Q^H Q - I||_F of the stored Q {orth:.2e}")
    backend.sync_all()
    g = F.panelblas.project_group(Y)
    c0 = F._lib.counters()
    t = time.perf_counter()
    S = F.project(Y, G, schedule="fused")
    backend.sync_all()
    t_proj = time.perf_counter() - t
    up_proj = F._lib.counters()["h2d_bytes"] - c0["h2d_bytes"]
    log(f"streamed-narrow: project {t_proj:.2f} s, {up_proj / 1e9:.0f} GB up, runs of {g} held images")
    return {"cholqr2_s": t_chol, "rows_per_s_pcie_inclusive": N / t_chol, "h2d_GB": up / 1e9,
            "d2h_GB": down / 1e9, "link_GBps_during_cholqr2": (up + down) / t_chol / 1e9,
            "panelmul_s": t_panelmul, "h2d_GBps_copy_only": h2d_copy, "stored_q_orthogonality": orth,
            "project_s": t_proj,
            "project_h2d_GB": up_proj / 1e9, "project_run_length": g, "project_check": project_check(S, d),
            "N": N,
            "schedule": "fused (three uploads of Y per cholqr2; project by runs of held images)",
            "budgets": "device 64 GiB, host 60 GiB page-locked, host_eltype ComplexF32, nbuffers 2"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=216, help="Helmholtz grid edge; N = grid^3")
    ap.add_argument("--k", type=int, default=512)
    ap.add_argument("--w", type=int, default=32)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint workload (N = 2^22, k = 128)")
    ap.add_argument("--no-streamed", action="store_true",
                    help="skip the host-resident (streamed over PCIe) cholqr2 of the largest grid that page-locks")
    ap.add_argument("--streamed-project-reference", action="store_true",
                    help="in the streamed run, also time project on the reference schedule (p + 1 uploads)")
    ap.add_argument("--disk", action="store_true", help="also measure the disk tier (17 GB scratch file)")
    ap.add_argument("--range-finder", action="store_true",
                    help="also run the disk-tier range finder with power iterations (34 GB of scratch files)")
    ap.add_argument("--cpu-rows", type=int, default=1 << 17)
    ap.add_argument("--streamed-narrow", action="store_true",
                    help="also run the full-size streamed cholqr2 with ComplexF32 host storage")
    ap.add_argument("--streamed-only", action="store_true",
                    help="profile mode: only the streamed run, its phases separated by 0.25 s of GPU idle so "
                         "a trace can be cut at the gaps")
    ap.add_argument("--pinned-probe", action="store_true",
                    help="only report how much host memory can be page-locked (up to 160 GiB) and the host's facts")
    ap.add_argument("--kernels-only", action="store_true",
                    help="profile mode: only the per-launch kernels at the row-block shape")
    ap.add_argument("--vendor", action="store_true",
                    help="with --kernels-only: also time rocBLAS zherk / ztrsm / zgemm on the same shapes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path shards by rows with one all-reduce of the k x k Gram; see the design notes).",
              file=sys.stderr)
        sys.exit(2)
    if args.pinned_probe:
        pinned_probe()
        return
    if args.kernels_only:
        kernels_only(args.grid ** 3 * args.w // args.k, args.k, max(args.steps, 1), vendor=args.vendor)
        return
    if args.streamed_only:
        print(json.dumps(streamed(args, phase_gap=0.25, project_reference=args.streamed_project_reference)),
              flush=True)
        return

    import funicular_hip as F
    from funicular_hip.panelblas import cholqr2_mixed

    backend = F.hip_backend()
    n = args.grid
    N, k, w = n ** 3, args.k, args.w
    extra_steps = max(1, min(args.steps, 5))
    log(f"bench: N={N} k={k} w={w} on {backend.arch}")
    plan = F.ResidencyPlan(backend=backend, device_budget=240 * 1024 ** 3, host_budget=1 << 30,
                           device_resident=True)
    d = 6.0 - 0.09 * (1 + 0.05j)
    G = F.Helmholtz3D(n, n, n, d, -1.0, backend=backend)
    omega = F.GhostPanels(F.device_randn, np.complex128, N, k, plan=plan, seed=0xDEADBEEF, w=w)
    Y = F.PanelMatrix(N, k, plan=plan, w=w)

    def regenerate():
        backend.sync_all()
        t = time.perf_counter()
        F.panelmul(Y, G, omega)
        backend.sync_all()
        return time.perf_counter() - t

    t_panelmul = regenerate()
    log(f"bench: panelmul {t_panelmul:.3f} s")
    timer = KernelTimer(backend)

    def run(mode, schedule, steps, warmup):
        fn = {"fp64": lambda Z: F.cholqr2(Z, schedule=schedule),
              "mixed": lambda Z: cholqr2_mixed(Z, schedule=schedule),
              "mixed-fp64-solve": lambda Z: cholqr2_mixed(Z, solve="fp64", schedule=schedule)}[mode]
        regenerate()
        first = fn(Y)  # from the real basis: accuracy
        backend.sync_all()
        orth = float(np.linalg.norm(F.gram(Y) - np.eye(k)))
        resid = residual(backend, Y, first, G, omega)
        for _ in range(max(warmup - 1, 0)):
            fn(Y)
        backend.sync_all()
        timer.active = True
        t = time.perf_counter()
        for _ in range(steps):
            fn(Y)
        backend.sync_all()
        dt = time.perf_counter() - t
        timer.active = False
        log(f"bench: {mode} ({schedule} schedule) {steps} x cholqr2 in {dt:.3f} s, ||Q^H Q - I||_F = {orth:.2e}, "
            f"||Y - QR||/||Y|| = {resid:.2e}")
        return dt, orth, first, timer.collect(), resid

    # The headline: the reference schedule (every pass a Gram row sweep, the
    # Gram down, a host Cholesky, R up, a solve row sweep) on a basis held in
    # HBM.
    dt64, orth64, R64, k64, res64 = run("fp64", "reference", args.steps, args.warmup)
    ms_step = dt64 / args.steps * 1e3
    value = N * args.steps / dt64
    gram64 = summarize(k64[("gram", False)], gram_flops)
    trsm64 = summarize(k64[("trsm", False)], trsm_flops)
    h = F.row_block_height(N, k, w)
    kernels = {
        "gram": {"kernel": "gram_partial_kernel + gram_reduce_kernel (FP64)", "achieved": gram64["tflops"],
                 "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": gram64["tflops"] / FP64_PEAK_TFLOPS,
                 "traffic": measured_traffic("gram_fp64") if (N, k, w) == (10077696, 512, 32) else None,
                 "avg_launch_ms": gram64["avg_ms"], "flops_per_launch": gram64["flops_per_launch"],
                 "algorithmic_bytes_per_launch": 16 * h * k, "launches": gram64["launches"],
                 "total_ms_per_step": gram64["avg_ms"] * gram64["launches"] / args.steps},
        "trsm": {"kernel": "trsm_wide_kernel (FP64; on a sweep's first block also the two fold kernels)",
                 "achieved": trsm64["tflops"], "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": trsm64["tflops"] / FP64_PEAK_TFLOPS,
                 "traffic": measured_traffic("trsm_fp64") if (N, k, w) == (10077696, 512, 32) else None,
                 "avg_launch_ms": trsm64["avg_ms"], "flops_per_launch": trsm64["flops_per_launch"],
                 "algorithmic_bytes_per_launch": 32 * h * k, "launches": trsm64["launches"],
                 "total_ms_per_step": trsm64["avg_ms"] * trsm64["launches"] / args.steps},
    }
    dominant = max(kernels, key=lambda name: kernels[name]["total_ms_per_step"])
    record = {
        "metric": METRIC, "value": value, "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"cholqr2 of a {N}x{k} ComplexF64 basis held in HBM (device-resident plan): "
                               f"{n}^3 periodic Helmholtz stencil applied to GhostPanels(seed 0xdeadbeef), "
                               f"panels of {w}, on the reference schedule (4 row sweeps of {h}-row blocks, "
                               "host Cholesky)",
                   "N": N, "k": k, "panel_width": w, "row_block": h, "schedule": "reference",
                   "tier": "device-resident (the streamed host-resident run is the 'streamed' record)",
                   "element": "ComplexF64 (FP64 arithmetic on real and imaginary parts)"},
        "roofline": dict({"bound": "mfma"}, **{key: kernels[dominant][key] for key in
                                                ("kernel", "achieved", "peak", "unit", "frac", "traffic",
                                                 "avg_launch_ms", "flops_per_launch",
                                                 "algorithmic_bytes_per_launch")}),
        "kernels": kernels,
        "errors": {"fp64_orthogonality": orth64, "fp64_residual": res64},
        "panelmul_s": t_panelmul,
    }
    # The same on the fused schedule (device Cholesky, one solve launch over
    # all rows), an opt-in.
    dtf, orthf64, _, kf, resf64 = run("fp64", "fused", extra_steps, 1)
    record["fused_schedule"] = {"rows_per_s": N * extra_steps / dtf, "ms_per_step": dtf / extra_steps * 1e3,
                                "steps": extra_steps, "orthogonality": orthf64, "residual": resf64,
                                "speedup_vs_reference_schedule": (dt64 / args.steps) / (dtf / extra_steps),
                                "trsm_tflops": summarize(kf[("trsm", False)], trsm_flops)["tflops"]}
    # project on the orthonormal basis cholqr2 left in Y: the reference's
    # p (p + 1) steps, and the fused resident route (G Q whole, one Gram sweep).
    for schedule in ("reference", "fused"):
        F.project(Y, G, schedule=schedule)
        backend.sync_all()
        t = time.perf_counter()
        S = F.project(Y, G, schedule=schedule)
        backend.sync_all()
        t_project = time.perf_counter() - t
        record[f"project_{schedule}"] = {"s": t_project, "reduction_tflops": 8.0 * N * k * k / t_project / 1e12,
                                         "antihermitian_check": project_check(S, d)}
        log(f"bench: project ({schedule}) {t_project:.3f} s, S - S^H check "
            f"{record[f'project_{schedule}']['antihermitian_check']:.2e}")
    if not args.no_mixed:
        dts, orths, Rs, ks, ress = run("mixed-fp64-solve", "reference", args.steps, args.warmup)
        dtm, orthm, Rm, km, resm = run("mixed", "reference", args.steps, args.warmup)
        regenerate()
        Rf = cholqr2_mixed(Y, shift="fukaya")
        orthfk = float(np.linalg.norm(F.gram(Y) - np.eye(k)))
        resfk = residual(backend, Y, Rf, G, omega)
        log(f"bench: mixed with the Fukaya shift: ||Q^H Q - I||_F = {orthfk:.2e}, ||Y - QR||/||Y|| = {resfk:.2e}")
        g32 = summarize(km[("gram", True)], gram_flops)
        t32 = summarize(km[("trsm", True)], trsm_flops)
        # Per-pass ratio: the Gram and solve kernels of one pass over all N rows
        # with FP64 products against the same pair with FP32 products, from the
        # live launch timings.
        pass64 = N * (gram64["ms_per_row"] + trsm64["ms_per_row"])
        pass32 = N * (g32["ms_per_row"] + t32["ms_per_row"])
        dtmf, orthmf, _, _, resmf = run("mixed", "fused", extra_steps, 1)
        dtsf, orthsf, _, _, ressf = run("mixed-fp64-solve", "fused", extra_steps, 1)
        step_f = dtf / extra_steps
        record["mixed"] = {
            "schedule": "reference (each pass a Gram sweep, host Cholesky, solve sweep), against the FP64 "
                        "reference-schedule headline",
            "rows_per_s": N * args.steps / dtm, "ms_per_step": dtm / args.steps * 1e3,
            "speedup_vs_fp64": dt64 / dtm, "orthogonality": orthm, "residual": resm,
            "speedup_at_orthogonality_bound": dt64 / dtm,
            "speedup_at_both_bounds": dt64 / dts,
            "first_pass": "FP32 Gram (FP64 chunk sums), FP32 Cholesky, FP32 solve",
            "shift": "adaptive (unshifted FP32 Cholesky unless it fails)",
            "fukaya_shift_orthogonality": orthfk, "fukaya_shift_residual": resfk,
            "gram_fp32_tflops": g32["tflops"], "trsm_fp32_tflops": t32["tflops"],
            "gram_fp32_frac_of_fp32_peak": g32["tflops"] / FP32_PEAK_TFLOPS,
            "trsm_fp32_frac_of_fp32_peak": t32["tflops"] / FP32_PEAK_TFLOPS,
            "first_pass_vs_fp64_pass": {"fp64_pass_kernels_ms": pass64, "fp32_pass_kernels_ms": pass32,
                                        "speedup": pass64 / pass32},
            "R_rel_diff_vs_fp64": float(np.linalg.norm(Rm - R64) / np.linalg.norm(R64)),
            "fp64_solve": {"rows_per_s": N * args.steps / dts, "speedup_vs_fp64": dt64 / dts,
                           "orthogonality": orths, "residual": ress},
            "fused": {"speedup_vs_fused_fp64": step_f / (dtmf / extra_steps), "orthogonality": orthmf,
                      "residual": resmf, "fp64_solve_speedup_vs_fused_fp64": step_f / (dtsf / extra_steps),
                      "fp64_solve_orthogonality": orthsf, "fp64_solve_residual": ressf}}
    # The 216^3 x 512 basis is done with: release it before the other
    # workloads allocate theirs.
    del S
    F.free(Y)
    import gc
    gc.collect()
    if not args.no_mixed:
        record["mixed_graded"] = mixed_bench(backend, args.steps)
        gc.collect()
    if not args.no_adjoint:
        record["adjoint"] = adjoint_bench(backend)
    if args.disk:
        import tempfile
        with tempfile.TemporaryDirectory(prefix="funicular-disk-") as d:
            record["disk_tier"] = disk_bench(backend, d)
    if args.range_finder:
        import tempfile
        with tempfile.TemporaryDirectory(prefix="funicular-rf-") as d:
            record["range_finder"] = range_finder_bench(backend, d)
    del Y, omega, plan
    gc.collect()
    if not args.no_streamed:
        record["streamed"] = streamed(args, project_reference=args.streamed_project_reference)
        gc.collect()
    if args.streamed_narrow:
        record["streamed_narrow"] = streamed_narrow(args)
        gc.collect()
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        record["cpu_baseline"] = cpu_baseline(k, w, args.cpu_rows)
    print(json.dumps(record), flush=True)


if __name__ == "__main__":
    main()
