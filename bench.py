"""Benchmark: LUCJ applications/s at norb=16, nelec=(8,8) (complex128) on one MI355X.

A step is one application of a spin-balanced LUCJ operator with n_reps=4
(local interaction pairs aa=(p,p+1), ab=(p,p)) to a resident 12870 x 12870
complex128 state (2.65 GB), i.e. ``UCJOpSpinBalanced._apply_unitary_``
(python/ffsim/variational/ucj_spin_balanced.py:657-696) run by the GPU
library.  Inputs are synthetic (seeded) and resident in HBM before the timed
region starts.

Printed: one JSON line with the throughput, the roofline of the dominant
kernel (the fused Givens fiber kernel; durations from HIP events recorded on
its stream during the timed region) and a CPU baseline measured on this
host's cores with the oracle's restatement of ffsim's CPU path.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def lucj_pairs(norb: int):
    pairs_aa = [(p, p + 1) for p in range(norb - 1)]
    pairs_ab = [(p, p) for p in range(norb)]
    return pairs_aa, pairs_ab


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_cpu_info() -> dict:
    """The timing host's CPU: logical CPUs, the affinity mask, the cgroup CPU
    quota (cgroup v2 cpu.max) and the model name."""
    info: dict = {"nproc": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()
        info["cgroup_cpu_quota"] = None if quota == "max" else int(quota) / int(period)
    except (OSError, ValueError):
        pass
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                info["model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    info["OMP_NUM_THREADS"] = os.environ.get("OMP_NUM_THREADS")
    return info


def usable_cpus(info: dict) -> int:
    """CPUs this process may run on at once: the affinity mask, capped by the
    cgroup quota when one is set."""
    n = info.get("affinity_cpus") or info.get("nproc") or 1
    q = info.get("cgroup_cpu_quota")
    if q:
        n = min(n, max(1, int(q)))
    return int(n)


def cpu_baseline(norb: int, nelec, op, vec_host: np.ndarray, thread_counts, repeats: int = 5,
                 single_thread_rotations: int = 12) -> dict:
    """Time a bounded sample of the same workload with the CPU oracle.

    The oracle's C restatement of ffsim's Rust kernels is built here with
    -march=native (as ffsim's .cargo/config.toml:2 builds ffsim) and run with
    ffsim's schedule: one call per Givens rotation with per-call threads over
    row pairs, one per orbital phase, a full transpose copy per beta sector
    (python/ffsim/gates/orbital_rotation.py:117-154), and a threaded
    diag-Coulomb pass.  Sample: one alpha-sector orbital rotation (the 120
    Givens rotations and 16 phase shifts), one transpose copy and one
    diag-Coulomb evolution of the norb=16 state.  ffsim's LUCJ n_reps=4
    application (ucj_spin_balanced.py:657-696) is 5 orbital rotations, i.e. 10
    sector rotations and 10 transpose copies, plus 4 diag-Coulomb passes, so
    apps/s = 1 / (10 T_rot + 10 T_transpose + 4 T_dc).  For each thread count:
    one warm-up sample, then the median of ``repeats``.  On one thread: the
    first ``single_thread_rotations`` Givens rotations scaled to the sector
    rotation by rotation count, one transpose, one diag-Coulomb pass.
    """
    os.environ["FFSIM_ORACLE_NATIVE"] = "1"
    from oracle import oracle as O

    O.build(force=True)
    O.init_cache(norb, nelec)
    da, db = math.comb(norb, nelec[0]), math.comb(norb, nelec[1])
    v = np.ascontiguousarray(vec_host.reshape(da, db))
    U = op.orbital_rotations[0]
    rots, phases = O.givens_decomposition(U)
    j_aa, j_ab = op.diag_coulomb_mats[0]
    phase_idx = [O.one_subspace_indices(norb, nelec[0], (i,)) for i in range(norb)]

    def sample(n_rot=None):
        sel = rots if n_rot is None else rots[:n_rot]
        t0 = time.perf_counter()
        for c, s_, i, j in sel:
            O._apply_adjacent(v, c, s_.conjugate(), (i, j), norb, nelec[0])
        if n_rot is None:
            for i, ph in enumerate(phases):
                O._phase_shift_in_place(v, complex(ph), phase_idx[i])
        t_rot = time.perf_counter() - t0
        if n_rot is not None:
            t_rot *= len(rots) / max(1, len(sel))
        t0 = time.perf_counter()
        vt = O._transpose_copy(v)
        t_tr = time.perf_counter() - t0
        del vt
        t0 = time.perf_counter()
        O.apply_diag_coulomb_evolution(v.reshape(-1), (j_aa, j_ab, j_aa), -1.0, norb, nelec,
                                       copy=False)
        t_dc = time.perf_counter() - t0
        return t_rot, t_tr, t_dc

    runs_by_threads = {}
    for threads in thread_counts:
        os.environ["FFSIM_ORACLE_THREADS"] = str(threads)
        sample()  # warm-up
        runs = [sample() for _ in range(repeats)]
        t_rot, t_tr, t_dc = (float(np.median([r[k] for r in runs])) for k in range(3))
        runs_by_threads[threads] = (t_rot, t_tr, t_dc, 10 * t_rot + 10 * t_tr + 4 * t_dc)
        log(f"cpu sample ({threads} threads, median of {repeats}): sector rotation {t_rot:.2f}s "
            f"transpose {t_tr:.2f}s diag-coulomb {t_dc:.2f}s")
    best = min(runs_by_threads, key=lambda k: runs_by_threads[k][3])
    t_rot, t_tr, t_dc, t_app = runs_by_threads[best]
    os.environ["FFSIM_ORACLE_THREADS"] = "1"
    r1, tr1, dc1 = sample(single_thread_rotations)
    t_app1 = 10 * r1 + 10 * tr1 + 4 * dc1
    log(f"cpu sample (1 thread): sector rotation {r1:.2f}s (from {single_thread_rotations} "
        f"rotations) transpose {tr1:.2f}s diag-coulomb {dc1:.2f}s")
    os.environ["FFSIM_ORACLE_THREADS"] = str(best)
    return {
        "value": 1.0 / t_app,
        "unit": "LUCJ applications/s",
        "cores": best,
        "kind": "port",
        "sample": (
            f"C restatement of ffsim's Rust kernels (built -march=native on this host) with "
            f"ffsim's schedule (one call per Givens rotation with per-call threads over row "
            f"pairs, one per orbital phase, transpose copies), {best} threads, one warm-up then "
            f"the median of {repeats}: one alpha-sector orbital rotation ({t_rot:.2f} s), one "
            f"transpose copy ({t_tr:.2f} s) and one diag-Coulomb evolution ({t_dc:.2f} s) of the "
            f"norb=16 (8,8) state; one n_reps=4 application = 10 sector rotations + 10 transpose "
            f"copies + 4 diag-Coulomb passes = {t_app:.1f} s (scaled from the sample by those "
            f"counts; not timed end to end)"
        ),
        "by_threads": {str(k): {"sector_rotation_s": round(v_[0], 3), "transpose_s": round(v_[1], 3),
                                "diag_coulomb_s": round(v_[2], 3),
                                "apps_per_s": round(1.0 / v_[3], 5)}
                       for k, v_ in runs_by_threads.items()},
        "single_thread": {
            "value": 1.0 / t_app1,
            "sample": (f"1 thread: {single_thread_rotations} of the {len(rots)} Givens rotations "
                       f"scaled by rotation count ({r1:.2f} s per sector rotation), one transpose "
                       f"({tr1:.2f} s), one diag-Coulomb pass ({dc1:.2f} s)"),
        },
    }


def c64_error(op, state, norb: int, nelec) -> dict:
    """One LUCJ application in complex64 and in complex128 from the same state
    (the benchmark's state, normalized); max abs difference and
    1 - |<a|b>|^2 / (<a|a><b|b>), computed on the device in FP64."""
    import torch

    a = state.to(torch.complex128, copy=True)
    a /= torch.linalg.vector_norm(a)
    b = a.to(torch.complex64)
    op.apply_in_place(a, norb, nelec)
    op.apply_in_place(b, norb, nelec)
    b = b.to(torch.complex128)
    max_abs = float((a - b).abs().max().item())
    ov = torch.vdot(a, b).item()
    na, nb = torch.vdot(a, a).real.item(), torch.vdot(b, b).real.item()
    del a, b
    torch.cuda.empty_cache()
    return {"workload": "one LUCJ n_reps=4 application from the benchmark state",
            "max_abs": max_abs, "infidelity": float(1.0 - abs(ov) ** 2 / (na * nb)),
            "c64_norm_error": float(abs(math.sqrt(nb) - 1.0))}


def copy_bandwidth(src, dst, reps: int = 10) -> float:
    """Device-to-device copy rate in GB/s (bytes read + bytes written), the
    practical HBM ceiling on this GPU for a streaming read-write kernel."""
    import torch

    dst.copy_(src)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        dst.copy_(src)
    end.record()
    end.synchronize()
    return 2.0 * src.numel() * src.element_size() * reps / (start.elapsed_time(end) * 1e6)


def measured_traffic(kernel: str, workload: str, dtype: str):
    """HBM bytes per launch of ``kernel`` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/traffic.json), or None."""
    path = ROOT / "profiles" / "traffic.json"
    if not path.exists():
        return None
    for entry in json.loads(path.read_text()):
        if (entry.get("kernel"), entry.get("workload"), entry.get("dtype")) == (kernel, workload, dtype):
            return entry
    return None


def _profiled(fn, reps: int) -> dict:
    """Run fn() reps times between profiler begin/end (after one warm-up call):
    wall ms per call and the fiber / diagonal kernels' time and algorithmic bytes."""
    import torch

    from ffsim_amd import _lib

    fn()
    torch.cuda.synchronize()
    buf = (ctypes.c_double * 6)()
    _lib.call("ffsim_amd_profile_begin")
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    _lib.call("ffsim_amd_profile_end", buf)
    fms, fn_, fb, dms, dn, db = list(buf)
    out = {"ms_per_call": round(wall * 1e3, 3)}
    if fn_:
        out["fiber_launches_per_call"] = fn_ / reps
        out["fiber_avg_ms"] = round(fms / fn_, 4)
        out["fiber_GBs"] = round(fb / (fms * 1e6), 1)
    if dn:
        out["diag_launches_per_call"] = dn / reps
        out["diag_avg_ms"] = round(dms / dn, 4)
        out["diag_GBs"] = round(db / (dms * 1e6), 1)
    return out


def lucj_rate(op, norb: int, nelec, cdtype, steps: int, warmup: int, seed: int = 0) -> dict:
    """LUCJ applications/s on a resident state of the given dtype (same seeded
    state as the headline run), with the fiber kernel's average launch."""
    import torch

    from ffsim_amd import _lib

    da, db = math.comb(norb, nelec[0]), math.comb(norb, nelec[1])
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    state = torch.randn(da * db, dtype=torch.complex128, device="cuda", generator=gen)
    state /= torch.linalg.vector_norm(state)
    state = state.to(cdtype)
    for _ in range(warmup):
        op.apply_in_place(state, norb, nelec)
    torch.cuda.synchronize()
    buf = (ctypes.c_double * 6)()
    _lib.call("ffsim_amd_profile_begin")
    t0 = time.perf_counter()
    for _ in range(steps):
        op.apply_in_place(state, norb, nelec)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.call("ffsim_amd_profile_end", buf)
    fms, fn_, fb, dms, dn, db_ = list(buf)
    del state
    torch.cuda.empty_cache()
    return {"apps_per_s": round(steps / elapsed, 3), "ms_per_app": round(1e3 * elapsed / steps, 2),
            "fiber_avg_ms": round(fms / max(fn_, 1), 4),
            "fiber_GBs": round(fb / max(fms, 1e-9) / 1e6, 1),
            "fiber_launches_per_app": fn_ / steps, "diag_launches_per_app": dn / steps}


def gates_leg(norb: int = 16, reps: int = 5, seed: int = 21) -> dict:
    """The plain drop-ins at norb=16 (8,8) on a resident state, complex128 and
    complex64: apply_orbital_rotation with a Haar 16x16 unitary (both sectors),
    apply_diag_coulomb_evolution (random symmetric J, U(-1,1), t=1.0; the
    separate diagonal kernel) and apply_num_op_sum_evolution (random 16-vector,
    U(-1,1), t=1.0).  Per call: wall ms and kernel GB/s (algorithmic bytes over
    the kernels' event time)."""
    import torch

    import ffsim_amd

    nelec = (norb // 2, norb // 2)
    rng = np.random.default_rng(seed)
    U = ffsim_amd.random.random_unitary(norb, seed=rng)
    vals = rng.uniform(-1, 1, size=norb * (norb + 1) // 2)
    J = np.zeros((norb, norb))
    J[np.triu_indices(norb)] = vals
    J = J + np.triu(J, 1).T
    lam = rng.uniform(-1, 1, size=norb)
    da, db = math.comb(norb, nelec[0]), math.comb(norb, nelec[1])
    out = {"workload": f"norb={norb} nelec={nelec} ({da}x{db}), Haar U, J U(-1,1) t=1.0, "
                       f"num-op coefficients U(-1,1) t=1.0"}
    for cdt, tag in ((torch.complex128, "c128"), (torch.complex64, "c64")):
        psi = torch.randn(da * db, dtype=torch.complex128, device="cuda")
        psi = (psi / torch.linalg.vector_norm(psi)).to(cdt)
        out[tag] = {
            "orbital_rotation": _profiled(
                lambda: ffsim_amd.apply_orbital_rotation(psi, U, norb, nelec, copy=False), reps),
            "diag_coulomb_evolution": _profiled(
                lambda: ffsim_amd.apply_diag_coulomb_evolution(psi, J, 1.0, norb, nelec,
                                                               copy=False), reps),
            "num_op_sum_evolution": _profiled(
                lambda: ffsim_amd.apply_num_op_sum_evolution(psi, lam, 1.0, norb, nelec,
                                                             copy=False), reps),
        }
        out[tag]["native_calls"] = _native_calls(psi.view(da, db), norb, nelec[0], rng, reps)
        out[tag]["one_body_contraction"] = _one_body(psi, norb, nelec, rng, reps)
        log(f"gates {tag}: {out[tag]}")
        del psi
        torch.cuda.empty_cache()
    return out


def _native_calls(vec, norb: int, nocc: int, rng, reps: int) -> dict:
    """ffsim's per-rotation calls on the resident da x db state (alpha strings on
    rows): one Givens rotation on orbitals (norb/2 - 1, norb/2) with ffsim's
    pair tables, one phase shift of the rows with orbital 0 occupied, and a
    layer of norb/2 disjoint rotations (0,1), (2,3), ... as norb/2 single calls
    and as one ffsim_apply_givens_layer call.  GB/s = bytes of the touched rows
    read and written over the wall time per call."""
    import torch

    from ffsim_amd import _cistring, native

    row_bytes = vec.shape[1] * vec.element_size()
    mid = norb // 2
    idx = _cistring.zero_one_subspace_indices(norb, nocc, (mid - 1, mid))
    half = len(idx) // 2
    s1, s2 = idx[:half], idx[half:]
    strings = _cistring.make_strings(range(norb), nocc)
    occ0 = np.flatnonzero(strings & 1).astype(np.uint64)
    theta = rng.uniform(0, 2 * np.pi, size=mid)
    phis = rng.uniform(0, 2 * np.pi, size=mid)
    layer = [(math.cos(t), math.sin(t) * complex(math.cos(p), math.sin(p)), 2 * k, 2 * k + 1)
             for k, (t, p) in enumerate(zip(theta, phis))]
    layer_tables = [_cistring.zero_one_subspace_indices(norb, nocc, (i, j))
                    for _, _, i, j in layer]

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    def singles():
        for (c, s, _, _), t in zip(layer, layer_tables):
            h = len(t) // 2
            native.apply_givens_rotation_in_place(vec, c, s, t[:h], t[h:])

    c, s = layer[0][0], layer[0][1]
    t_rot = timed(lambda: native.apply_givens_rotation_in_place(vec, c, s, s1, s2))
    t_phase = timed(lambda: native.apply_phase_shift_in_place(vec, complex(0.6, 0.8), occ0))
    t_singles = timed(singles)
    t_layer = timed(lambda: native.apply_givens_layer(vec, layer, norb, nocc))
    layer_rows = sum(len(t) for t in layer_tables)
    return {
        "givens_rotation": {"ms_per_call": round(t_rot * 1e3, 3), "rows": 2 * half,
                            "GBs": round(2 * 2 * half * row_bytes / t_rot / 1e9, 1)},
        "phase_shift": {"ms_per_call": round(t_phase * 1e3, 3), "rows": len(occ0),
                        "GBs": round(2 * len(occ0) * row_bytes / t_phase / 1e9, 1)},
        "layer_as_single_calls": {"ms": round(t_singles * 1e3, 3), "calls": len(layer),
                                  "GBs": round(2 * layer_rows * row_bytes / t_singles / 1e9, 1)},
        "layer_one_call": {"ms": round(t_layer * 1e3, 3), "rotations": len(layer)},
    }


def _one_body(psi, norb: int, nelec, rng, reps: int) -> dict:
    """contract_one_body (both sectors, out = M^ psi) on the resident state with
    a random complex M (the window passes; their rate counts psi read twice,
    out written, then read and written: 5 state sizes per call) and with a
    random Hermitian M (two orbital rotations around a number-operator pass):
    wall ms per call."""
    import torch

    from ffsim_amd.contract.one_body import contract_one_body_into

    def timed(mat):
        out = torch.empty_like(psi)
        contract_one_body_into(psi, out, mat, norb, nelec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            contract_one_body_into(psi, out, mat, norb, nelec)
        torch.cuda.synchronize()
        del out
        return (time.perf_counter() - t0) / reps

    mat = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    t = timed(mat)
    t_h = timed(mat + mat.T.conj())
    return {"ms_per_call": round(t * 1e3, 3),
            "state_pass_GBs": round(5 * psi.numel() * psi.element_size() / t / 1e9, 1),
            "hermitian_ms_per_call": round(t_h * 1e3, 3)}


def numpy_io_leg(op, norb: int, nelec, reps: int = 3, seed: int = 0) -> dict:
    """The drop-in as ffsim's API takes it: numpy state in, numpy state out
    (PCIe both ways through pinned bounce buffers), LUCJ n_reps=4, complex128."""
    import ffsim_amd

    da, db = math.comb(norb, nelec[0]), math.comb(norb, nelec[1])
    vec = ffsim_amd.random.random_state_vector(da * db, seed=seed)
    ffsim_amd.apply_unitary(vec, op, norb, nelec)
    t0 = time.perf_counter()
    for _ in range(reps):
        ffsim_amd.apply_unitary(vec, op, norb, nelec)
    elapsed = (time.perf_counter() - t0) / reps
    return {"workload": f"apply_unitary(numpy vector, LUCJ n_reps=4, norb={norb}, {nelec}), "
                        f"complex128, numpy in and out",
            "apps_per_s": round(1.0 / elapsed, 3), "ms_per_app": round(elapsed * 1e3, 1)}


def trotter_leg(norb: int = 18, n_factors: int = 10, time_step: float = 0.1, seed: int = 5) -> dict:
    """One first-order Trotter step of a synthetic double-factorized Hamiltonian
    (one-body term plus n_factors rotated diagonal Coulomb terms, J ~ N(0, 0.1),
    t = 0.1) on the norb=18 (9,9) state (48620 x 48620; 37.8 GB in complex128),
    in complex128 and in complex64 from the same initial state: time per step,
    per-kernel achieved GB/s and the complex64-vs-complex128 error on device."""
    import torch

    import ffsim_amd
    from ffsim_amd import _lib

    nelec = (norb // 2, norb // 2)
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    ham = ffsim_amd.random.random_df_hamiltonian_normal(norb, n_factors=n_factors, seed=seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    psi0 = torch.randn(dim, dtype=torch.complex128, device="cuda", generator=gen)
    psi0 /= torch.linalg.vector_norm(psi0)
    out = {"workload": f"DF Trotter step, norb={norb} nelec={nelec}, {n_factors} factors, "
                       f"t={time_step}, order 0", "dim": dim}
    results = {}
    for dtype in (torch.complex128, torch.complex64):
        tag = "c128" if dtype == torch.complex128 else "c64"
        psi = psi0.to(dtype, copy=True)  # both runs start from the same state
        if dtype == torch.complex64:
            del psi0
        torch.cuda.synchronize()
        buf = (ctypes.c_double * 6)()
        _lib.call("ffsim_amd_profile_begin")
        t0 = time.perf_counter()
        ffsim_amd.simulate_trotter_double_factorized(psi, ham, time_step, norb=norb, nelec=nelec,
                                                     copy=False)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        _lib.call("ffsim_amd_profile_end", buf)
        fms, fn, fb, dms, dn, db = list(buf)
        out[tag] = {
            "ms_per_step": round(elapsed * 1e3, 1),
            "fiber_launches": int(fn),
            "fiber_avg_ms": round(fms / max(fn, 1), 3),
            "fiber_GBs": round(fb / max(fms, 1e-9) / 1e6, 1),
            "diag_launches": int(dn),
            "diag_GBs": round(db / max(dms, 1e-9) / 1e6, 1),
        }
        results[tag] = psi
        log(f"trotter {tag}: {out[tag]}")
    psi128, psi64 = results["c128"], results["c64"].to(torch.complex128)
    out["c64_vs_c128_max_abs"] = float((psi64 - psi128).abs().max().item())
    chunk = 1 << 30  # torch.vdot takes at most 2^31 - 1 elements

    def dot(x, y):
        return sum(torch.vdot(x[i : i + chunk], y[i : i + chunk]).item()
                   for i in range(0, dim, chunk))

    overlap, n128, n64 = dot(psi128, psi64), dot(psi128, psi128).real, dot(psi64, psi64).real
    # 1 - |<a|b>|^2 / (<a|a><b|b>): the norm drift of the c64 run is reported apart
    out["c64_vs_c128_infidelity"] = float(1.0 - abs(overlap) ** 2 / (n128 * n64))
    out["c64_norm_error"] = float(abs(math.sqrt(n64) - 1.0))
    out["c128_norm_error"] = float(abs(math.sqrt(n128) - 1.0))
    del results, psi128, psi64
    torch.cuda.empty_cache()
    return out


def gradient_leg(seed: int = 12) -> dict:
    """Energy gradient of LUCJ n_reps=4 at norb=12, nelec=(6,6) (668 parameters,
    U(-pi, pi)), Hartree-Fock reference, synthetic H_DF with 8 factors: the
    adjoint sweep's wall time against central finite differences, whose cost is
    1336 energy evaluations (timed as one evaluation on the GPU and one on the
    CPU oracle, scaled)."""
    import torch

    import ffsim_amd
    from ffsim_amd.variational.gradient import UCJLayout, _apply_forward, ucj_energy_and_gradient

    norb, nelec, n_reps = 12, (6, 6), 4
    pairs = lucj_pairs(norb)
    theta = np.random.default_rng(seed).uniform(-np.pi, np.pi, 668)
    ham = ffsim_amd.random.random_df_hamiltonian_normal(norb, n_factors=8, seed=7)
    linop = ffsim_amd.linear_operator(ham, norb, nelec)
    psi0 = torch.from_numpy(ffsim_amd.hartree_fock_state(norb, nelec)).cuda()
    layout = UCJLayout.make(norb, n_reps, pairs)

    def energy():
        psi = psi0.clone()
        _apply_forward(layout, theta, psi, nelec)
        return float(torch.vdot(psi, linop.matvec_device(psi)).real.item())

    for _ in range(2):
        ucj_energy_and_gradient(theta, linop, psi0, norb=norb, nelec=nelec, n_reps=n_reps,
                                interaction_pairs=pairs)
        energy()
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        ucj_energy_and_gradient(theta, linop, psi0, norb=norb, nelec=nelec, n_reps=n_reps,
                                interaction_pairs=pairs)
    t_grad = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        energy()
    t_energy = (time.perf_counter() - t0) / reps
    return {
        "workload": "LUCJ n_reps=4 norb=12 (6,6), 668 params, HF reference, H_DF 8 factors",
        "adjoint_gradient_ms": round(t_grad * 1e3, 2),
        "gpu_energy_ms": round(t_energy * 1e3, 3),
        "gpu_central_fd_ms_estimated": round(1336 * t_energy * 1e3, 1),
        "adjoint_speedup_vs_gpu_fd": round(1336 * t_energy / t_grad, 1),
    }


def cpu_baseline_gradient(threads: int, seed: int = 12) -> dict:
    """CPU side of the gradient workload: one energy evaluation of the oracle's
    restatement of ffsim's CPU path, scaled to a central-FD gradient (1336
    evaluations, ffsim's gradient mechanism)."""
    os.environ["FFSIM_ORACLE_THREADS"] = str(threads)
    import ffsim_amd
    from oracle import oracle as O

    O.build()
    norb, nelec, n_reps = 12, (6, 6), 4
    pairs = lucj_pairs(norb)
    theta = np.random.default_rng(seed).uniform(-np.pi, np.pi, 668)
    ham = ffsim_amd.random.random_df_hamiltonian_normal(norb, n_factors=8, seed=7)
    hf = ffsim_amd.hartree_fock_state(norb, nelec)

    def h(v):
        return O.df_hamiltonian_matvec(v, ham.one_body_tensor, ham.diag_coulomb_mats,
                                       ham.orbital_rotations, ham.constant, norb, nelec, False)

    O.ucj_energy(theta, h, hf, norb, nelec, n_reps, pairs)
    t0 = time.perf_counter()
    O.ucj_energy(theta, h, hf, norb, nelec, n_reps, pairs)
    t_energy = time.perf_counter() - t0
    return {"energy_ms": round(t_energy * 1e3, 1),
            "central_fd_gradient_s_estimated": round(1336 * t_energy, 1)}


def batched_paths_leg(reps: int = 5, seed: int = 31) -> dict:
    """The batched LUCJ entry (n_reps=2) on the compound-matrix GEMM path against
    the Givens path with all states in each launch ("sweep") and state by state
    ("states"), per shape and batch size, complex128 states resident in HBM:
    milliseconds per state (wall time of the call and a synchronize, median of
    ``reps`` after a warm-up)."""
    import torch

    import ffsim_amd
    from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced

    rng = np.random.default_rng(seed)
    out = {}
    for norb in (8, 10, 12):
        nelec = (norb // 2, norb // 2)
        dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
        for batch in (1, 32):
            ops = [ffsim_amd.random.random_ucj_op_spin_balanced(
                norb, n_reps=2, interaction_pairs=lucj_pairs(norb), seed=rng) for _ in range(batch)]
            states = [torch.randn(dim, dtype=torch.complex128, device="cuda") for _ in range(batch)]
            works = [torch.empty_like(t) for t in states]
            row = {}
            for path in ("gemm", "sweep", "states"):
                os.environ["FFSIM_AMD_BATCH_PATH"] = path
                times = []
                for k in range(reps + 1):
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    UCJOpSpinBalanced.apply_batched_in_place(ops, states, norb, nelec, works)
                    torch.cuda.synchronize()
                    if k:
                        times.append(time.perf_counter() - t0)
                row[f"{path}_ms_per_state"] = round(float(np.median(times)) * 1e3 / batch, 4)
            os.environ.pop("FFSIM_AMD_BATCH_PATH", None)
            out[f"norb{norb}_batch{batch}"] = row
            del states, works
    torch.cuda.empty_cache()
    return out


def serving_leg(n_requests: int, load: float, seed: int = 99) -> dict:
    """Served mixed stream vs the unbatched plain path, same Poisson arrivals.

    Requests: LUCJ n_reps=2 (pairs aa=(p,p+1), ab=(p,p)) with random parameters,
    norb uniform in {8, 10, 12, 14}, nelec (n/2, n/2) or (n/2, n/2 - 1), complex64
    or complex128 (50/50), on a random state resident in HBM; the result stays in
    HBM.  The plain path applies one request at a time (one C-ABI call, then a
    synchronize); the server admits by HBM footprint and launches same-shape
    requests admitted together as one batch.  The arrival rate is ``load``
    times the plain path's capacity (1 / mean service time measured alone).
    """
    import torch

    import ffsim_amd
    from ffsim_amd.serving import LUCJRequest, LUCJServer

    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    reqs = []
    for _ in range(n_requests):
        norb = int(rng.choice([8, 10, 12, 14]))
        nelec = (norb // 2, norb // 2 - int(rng.integers(0, 2)))
        dtype = "c64" if rng.uniform() < 0.5 else "c128"
        op = ffsim_amd.random.random_ucj_op_spin_balanced(
            norb, n_reps=2, interaction_pairs=lucj_pairs(norb), seed=rng
        )
        dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
        st = torch.randn(dim, dtype=torch.complex128, device="cuda", generator=gen)
        st = (st / torch.linalg.vector_norm(st)).to(
            torch.complex64 if dtype == "c64" else torch.complex128)
        reqs.append(LUCJRequest(op=op, norb=norb, nelec=nelec, state=st, dtype=dtype,
                                to_host=False))

    def plain(r):
        t = r.state.clone()
        r.op.apply_in_place(t, r.norb, r.nelec)
        torch.cuda.synchronize()
        return t

    seen = set()
    for r in reqs:  # warm tables and plans of every shape, on both paths
        key = (r.norb, r.nelec, r.dtype)
        if key not in seen:
            seen.add(key)
            plain(r)
    warm = LUCJServer(hbm_budget=16 << 30, n_streams=4, max_batch=32, poll_interval=0.0)
    for r in reqs[:64]:
        warm.submit(r)
    warm.drain(timeout=600)
    service = []
    for r in reqs:
        t0 = time.perf_counter()
        plain(r)
        service.append(time.perf_counter() - t0)
    mean_service = float(np.mean(service))
    rate = load / mean_service
    offsets = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))

    # unbatched plain path: one request at a time, in arrival order
    lat_plain = []
    start = time.perf_counter()
    for r, off in zip(reqs, offsets):
        while time.perf_counter() - start < off:
            pass
        plain(r)
        lat_plain.append(time.perf_counter() - start - off)
    span_plain = time.perf_counter() - start

    budget = 16 << 30
    server = LUCJServer(hbm_budget=budget, slots_per_stream=2, n_streams=4, max_batch=32,
                        poll_interval=0.0)
    tickets = []
    start = time.perf_counter()
    k = 0
    while k < n_requests or server.step():
        now = time.perf_counter() - start
        while k < n_requests and offsets[k] <= now:
            tickets.append(server.submit(reqs[k], now=start + offsets[k]))
            k += 1
        server.step()
    span_served = time.perf_counter() - start
    torch.cuda.synchronize()
    lat_served = [t.latency for t in tickets]
    # agreement with the plain path (batched small states take the GEMM path)
    worst = {"c128": 0.0, "c64": 0.0}
    for t, r in zip(tickets[:64], reqs[:64]):
        d = float((t.result().to(torch.complex128) - plain(r).to(torch.complex128)).abs().max())
        worst[r.dtype] = max(worst[r.dtype], d)
    mismatches = int(worst["c128"] > 1e-12) + int(worst["c64"] > 2e-6)
    del tickets

    # saturation: every request queued at once, plain path one after another
    start = time.perf_counter()
    for r in reqs:
        plain(r)
    burst_plain = n_requests / (time.perf_counter() - start)
    burst = LUCJServer(hbm_budget=budget, slots_per_stream=2, n_streams=4, max_batch=32,
                       poll_interval=0.0)
    start = time.perf_counter()
    burst_tickets = [burst.submit(r) for r in reqs]
    burst.drain(timeout=600)
    torch.cuda.synchronize()
    burst_served = n_requests / (time.perf_counter() - start)
    if not all(t.done() for t in burst_tickets):
        raise RuntimeError("saturation run left requests unfinished")
    del burst_tickets
    torch.cuda.empty_cache()

    def summary(lat, span):
        lat = np.asarray(lat)
        return {
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "states_per_s": round(n_requests / span, 1),
        }

    return {
        "workload": "LUCJ n_reps=2, norb in {8,10,12,14}, nelec (n/2,n/2) or (n/2,n/2-1), "
        "c64/c128 50/50, states resident in HBM (in and out), Poisson arrivals",
        "requests": n_requests,
        "offered_load_vs_plain_capacity": load,
        "arrival_rate_per_s": round(rate, 1),
        "mean_plain_service_ms": round(mean_service * 1e3, 3),
        "served": summary(lat_served, span_served),
        "plain": summary(lat_plain, span_plain),
        "hbm_budget_bytes": budget,
        "peak_hbm_bytes": server.peak_bytes,
        "served_vs_plain_max_abs": worst,
        "served_vs_plain_out_of_tolerance": mismatches,
        "saturation_states_per_s": {"served": round(burst_served, 1),
                                    "plain": round(burst_plain, 1),
                                    "ratio": round(burst_served / burst_plain, 2)},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--norb", type=int, default=16)
    ap.add_argument("--n-reps", type=int, default=4)
    ap.add_argument("--dtype", choices=["c128", "c64"], default="c128")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-c64-check", action="store_true",
                    help="skip the c64-vs-c128 comparison (keeps profiler runs to one dtype)")
    ap.add_argument("--no-trotter", action="store_true")
    ap.add_argument("--no-gradient", action="store_true")
    ap.add_argument("--only",
                    choices=["gates", "numpy", "trotter", "gradient", "batched", "serving"],
                    default=None,
                    help="after the headline run, run this one leg and no other (for profiling)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other-dtype LUCJ rate, the plain gates and the numpy-io legs")
    ap.add_argument("--serving-requests", type=int, default=400)
    ap.add_argument("--serving-load", type=float, default=0.9)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads for the CPU baseline (0: every CPU this process may use, "
                         "and nproc too when that is more)")
    ap.add_argument(
        "--copy-probe",
        type=int,
        default=0,
        metavar="MIB",
        help="only run 3 device-to-device copies of MIB MiB each (calibrates HBM byte counters)",
    )
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py runs on one GPU only: ffsim_amd has no multi-GPU path "
            f"(asked for --gpus {args.gpus}).",
            file=sys.stderr,
        )
        return 2

    import torch

    import ffsim_amd
    from ffsim_amd import _lib

    assert torch.cuda.is_available(), "bench.py needs the GPU"
    dev = torch.device("cuda:0")
    if args.copy_probe:
        n = args.copy_probe << 20
        src = torch.ones(n, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 3 copies of {n} bytes (read {n} + write {n} per copy); "
            f"{copy_bandwidth(src, dst):.1f} GB/s")
        return 0
    norb = args.norb
    nelec = (norb // 2, norb // 2)
    da, db = math.comb(norb, nelec[0]), math.comb(norb, nelec[1])
    cdtype = torch.complex128 if args.dtype == "c128" else torch.complex64
    op = ffsim_amd.random.random_ucj_op_spin_balanced(
        norb, n_reps=args.n_reps, interaction_pairs=lucj_pairs(norb), seed=1234
    )
    log(f"workload: LUCJ n_reps={args.n_reps} norb={norb} nelec={nelec} dim={da}x{db} {args.dtype}")
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    state = torch.randn(da * db, dtype=cdtype, device=dev, generator=gen)
    state /= torch.linalg.vector_norm(state)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        op.apply_in_place(state, norb, nelec)
    torch.cuda.synchronize()

    out = (ctypes.c_double * 6)()
    _lib.call("ffsim_amd_profile_begin")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.apply_in_place(state, norb, nelec)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _lib.call("ffsim_amd_profile_end", out)
    elapsed = t1 - t0
    value = args.steps / elapsed
    fiber_ms, fiber_n, fiber_bytes, diag_ms, diag_n, diag_bytes = list(out)
    norm = float(torch.linalg.vector_norm(state).item())
    log(f"timed {args.steps} steps in {elapsed:.3f}s -> {value:.3f} app/s; |psi|={norm:.12f}")
    log(f"fiber kernel: {fiber_n:.0f} launches {fiber_ms:.2f} ms {fiber_bytes / 1e9:.1f} GB; "
        f"diag kernel: {diag_n:.0f} launches {diag_ms:.2f} ms {diag_bytes / 1e9:.1f} GB")

    avg_ms = fiber_ms / max(fiber_n, 1)
    bytes_per_launch = fiber_bytes / max(fiber_n, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": "fiber_kernel (fused Givens window pass)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "traffic_source": None,
        "avg_launch_ms": round(avg_ms, 4),
        "algorithmic_bytes_per_launch": bytes_per_launch,
        "launches_per_step": fiber_n / args.steps,
    }
    # practical ceiling: a plain device copy of 1 GiB on the same GPU, after the timed region
    probe_src = torch.ones(1 << 30, dtype=torch.uint8, device=dev)
    probe_dst = torch.empty_like(probe_src)
    copy_gbs = copy_bandwidth(probe_src, probe_dst)
    del probe_src, probe_dst
    roofline["copy_GBs"] = round(copy_gbs, 1)
    roofline["frac_of_copy"] = round(achieved / copy_gbs, 4)
    log(f"device copy: {copy_gbs:.1f} GB/s")
    tr = measured_traffic("fiber_kernel", f"lucj_norb{norb}_nreps{args.n_reps}", args.dtype)
    if tr is not None:
        roofline["traffic"] = tr["bytes_per_launch"]
        roofline["traffic_source"] = tr["source"]
    result = {
        "metric": "LUCJ applies/s + Givens HBM GB/s at norb=16 (8,8); served p50/p99 latency",
        "value": round(value, 4),
        "unit": "LUCJ applications/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded random state and LUCJ parameters)",
        "config": {
            "workload": f"LUCJ n_reps={args.n_reps} apply_unitary, norb={norb} nelec={nelec}, "
            f"state {da}x{db} {args.dtype}, interaction pairs aa=(p,p+1) ab=(p,p)",
            "norb": norb,
            "nelec": list(nelec),
            "n_reps": args.n_reps,
        },
        "roofline": roofline,
        "diag_kernel": {
            "avg_launch_ms": round(diag_ms / max(diag_n, 1), 4),
            "achieved_GBs": round(diag_bytes / max(diag_ms, 1e-9) / 1e6, 1),
        },
    }
    if not args.no_c64_check:
        result["c64_vs_c128"] = c64_error(op, state, norb, nelec)
        log(f"c64 vs c128 (one application): {result['c64_vs_c128']}")
    vec_host = None if args.no_cpu_baseline else state.to(torch.complex128).cpu().numpy()
    del state
    torch.cuda.empty_cache()
    if args.only:
        args.no_extras = args.no_trotter = args.no_gradient = args.no_serving = True
        if args.only == "gates":
            result["gates_norb16"] = gates_leg(norb)
        elif args.only == "trotter":
            result["trotter_norb18"] = trotter_leg()
        elif args.only == "gradient":
            result["gradient_norb12"] = gradient_leg()
        elif args.only == "numpy":
            result["numpy_io"] = numpy_io_leg(op, norb, nelec)
        elif args.only == "batched":
            result["batched_lucj_paths"] = batched_paths_leg()
        else:
            result["serving"] = serving_leg(args.serving_requests, args.serving_load)
    if not args.no_extras:
        other = torch.complex64 if args.dtype == "c128" else torch.complex128
        tag = "c64" if args.dtype == "c128" else "c128"
        result[f"lucj_{tag}"] = lucj_rate(op, norb, nelec, other, args.steps, args.warmup)
        log(f"lucj {tag}: {result[f'lucj_{tag}']}")
        result["gates_norb16"] = gates_leg(norb)
        result["numpy_io"] = numpy_io_leg(op, norb, nelec)
        log(f"numpy io: {result['numpy_io']}")
    if not args.no_trotter:
        log("trotter leg")
        result["trotter_norb18"] = trotter_leg()
    if not args.no_gradient:
        log("gradient leg")
        result["gradient_norb12"] = gradient_leg()
        log(f"gradient: {result['gradient_norb12']}")
    if not args.no_serving:
        log("batched paths leg")
        result["batched_lucj_paths"] = batched_paths_leg()
        log(f"batched paths: {result['batched_lucj_paths']}")
        log("serving leg")
        result["serving"] = serving_leg(args.serving_requests, args.serving_load)
        log(f"serving: {result['serving']}")
    if not args.no_cpu_baseline:
        info = host_cpu_info()
        if args.cpu_threads > 0:
            counts = [args.cpu_threads]
        else:
            counts = sorted({usable_cpus(info), int(info.get("nproc") or 1)})
        log(f"cpu host: {info}; thread counts {counts}")
        result["cpu_baseline"] = cpu_baseline(norb, nelec, op, vec_host, counts)
        result["cpu_baseline"]["host"] = info
        if not args.no_gradient:
            result["cpu_baseline"]["gradient_norb12"] = cpu_baseline_gradient(
                result["cpu_baseline"]["cores"])
        result["gpu_over_cpu"] = round(value / result["cpu_baseline"]["value"], 2)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
